// Quantum causal graph dynamics rules of the drop-in C++ API (reference:
// src/rules/qcgd.hpp): graph layout helpers, erase_create, coin and
// split_merge rules and the step / reversed_step modifiers, as device
// functors of libquids_hip.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../quids.hpp"

namespace quids::rules::qcgd {

namespace utils {
/// qcgd.hpp:11-25
inline void hash_combine(std::size_t &seed, size_t const value_64) {
    const size_t m = 0xc6a4a7935bd1e995ULL;
    seed *= m;
    seed ^= value_64 >> 47;
    seed *= m;
    seed ^= value_64;
    seed *= m;
    seed += 0xe6546b64;
}
}  // namespace utils

namespace graphs {
enum { dot_l_t = -3, dot_r_t, element_t, pair_t };
/// qcgd.hpp:35-39 (16 bytes: two int16, 4 padding bytes, the hash)
struct sub_node {
    int16_t hmlz_and_element;
    int16_t right_or_type;
    size_t hash;
};
inline uint16_t num_nodes(char const *o) {
    uint16_t n;
    std::memcpy(&n, o, 2);
    return n;
}
inline bool left(char const *o, int node) { return o[2 + node] != 0; }
inline bool right(char const *o, int node) { return o[2 + num_nodes(o) + node] != 0; }
inline uint16_t node_name_begin(char const *o, int node) {
    uint16_t v;
    std::memcpy(&v, o + 2 + 2 * num_nodes(o) + 2 * node, 2);
    return v;
}
inline sub_node node_name(char const *o, int k) {
    sub_node s;
    std::memcpy(&s, o + 4 + 4 * num_nodes(o) + 16 * k, 16);
    return s;
}
/// qcgd.hpp:114-120: left[i], right[i] <- rand() & 1 in node order
inline void randomize(char *o) {
    const uint16_t n = num_nodes(o);
    for (int i = 0; i < n; ++i) {
        o[2 + i] = std::rand() & 1;
        o[2 + n + i] = std::rand() & 1;
    }
}
/// qcgd.hpp:122-146
inline size_t hash_graph(char const *o) {
    size_t lh = 0, rh = 0, nh = 0;
    const uint16_t n = num_nodes(o);
    for (int i = 0; i < n; ++i) {
        if (left(o, i)) utils::hash_combine(lh, i);
        if (right(o, i)) utils::hash_combine(rh, i);
        utils::hash_combine(nh, node_name(o, node_name_begin(o, i)).hash);
    }
    utils::hash_combine(nh, lh);
    utils::hash_combine(nh, rh);
    return nh;
}
}  // namespace graphs

namespace utils {
/// qcgd.hpp:214-230: a cycle of `size` nodes named 0..size-1, no particles
/// (padding bytes zeroed).  The caller owns object_begin (delete[]).
inline void make_graph(char *&object_begin, char *&object_end, uint16_t size) {
    const size_t bytes = 4 + 20 * size_t(size);
    object_begin = new char[bytes]();
    object_end = object_begin + bytes;
    std::memcpy(object_begin, &size, 2);
    for (uint16_t i = 0; i <= size; ++i) std::memcpy(object_begin + 2 + 2 * size + 2 * i, &i, 2);
    for (uint16_t i = 0; i < size; ++i) {
        graphs::sub_node s{};
        s.hmlz_and_element = i == 0 ? -1 : int16_t(i + 1);
        s.right_or_type = graphs::element_t;
        s.hash = i;
        std::memcpy(object_begin + 4 + 4 * size + 16 * size_t(i), &s, 16);
    }
}
/// qcgd.hpp:232-240
inline void randomize(quids::it_t &iter) {
    for (size_t gid = 0; gid < iter.num_object; ++gid) {
        uint size;
        mag_t *mag;
        char *begin;
        iter.get_object(gid, begin, size, mag);
        graphs::randomize(begin);
    }
}
}  // namespace utils

namespace utils {
/// graphs printed by print() (qcgd.hpp:212)
inline size_t max_print_num_graphs = size_t(-1);

namespace detail {
// name tree of a node (qcgd.hpp:243-266): an element prints its index, a
// half "(x).l" / "(x).r", a pair "(x)∨(y)"
inline void print_node_name(char const *o, int k) {
    const graphs::sub_node s = graphs::node_name(o, k);
    if (s.right_or_type == graphs::element_t) {
        std::cout << std::abs(s.hmlz_and_element) - 1;
    } else if (s.right_or_type == graphs::dot_l_t || s.right_or_type == graphs::dot_r_t) {
        std::cout << "(";
        print_node_name(o, k + 1);
        std::cout << (s.right_or_type == graphs::dot_l_t ? ").l" : ").r");
    } else {
        std::cout << "(";
        print_node_name(o, k + 1);
        std::cout << ")\u2228(";
        print_node_name(o, k + s.right_or_type);
        std::cout << ")";
    }
}
}  // namespace detail

/// qcgd.hpp:242-303: the max_print_num_graphs most probable graphs, one per
/// line: the magnitude (parts below the tolerance shown as 0) and every node
/// as -|<|name|>|- with its particles.  Like the reference it leaves
/// std::cout in fixed notation with 5 digits.
inline void print(quids::it_t const &iter) {
    std::vector<size_t> gids(iter.num_object);
    std::iota(gids.begin(), gids.end(), size_t(0));
    std::vector<double> nrm(iter.num_object);
    for (size_t g = 0; g < iter.num_object; ++g) {
        uint size;
        mag_t mag;
        char const *begin;
        iter.get_object(g, begin, size, mag);
        nrm[g] = std::norm(mag);
    }
    std::stable_sort(gids.begin(), gids.end(), [&](size_t a, size_t b) { return nrm[a] > nrm[b]; });
    const size_t num_prints = std::min(iter.num_object, max_print_num_graphs);
    for (size_t i = 0; i < num_prints; ++i) {
        uint size;
        mag_t mag;
        char const *begin;
        iter.get_object(gids[i], begin, size, mag);
        const uint16_t n = graphs::num_nodes(begin);
        const double real = std::abs(mag.real()) < quids::tolerance ? 0 : mag.real();
        const double imag = std::abs(mag.imag()) < quids::tolerance ? 0 : mag.imag();
        std::cout << std::fixed << std::setprecision(5) << "\t" << real << (imag <= -1e-5 ? " - " : " + ")
                  << std::abs(imag) << "i  ";
        for (int k = 0; k < n; ++k) {
            std::cout << "-|" << (graphs::left(begin, k) ? "<" : " ") << "|";
            detail::print_node_name(begin, graphs::node_name_begin(begin, k));
            std::cout << "|" << (graphs::right(begin, k) ? ">" : " ") << "|-";
        }
        std::cout << "\n";
    }
    if (num_prints < iter.num_object) std::cout << "\t...and " << iter.num_object - num_prints << " other graphs\n";
}

/// qcgd.hpp:309-353: a JSON object of the state's statistics (total_proba,
/// num_graphs, mean and spread of the node count and of the particle
/// density, interference and deletion ratios of the last step).  As in the
/// reference, std_dev_size is sqrt(avg_squared_size) (not of the variance)
/// once the variance passes the tolerance.
inline void serialize(quids::it_t const &iter, quids::sy_it_t const &sy_it, uint indentation = 0) {
    double interference_ratio = 1, deletion_ratio = 1;
    if (sy_it.num_object > 0) {
        interference_ratio = double(sy_it.num_object_after_interferences) / double(sy_it.num_object);
        deletion_ratio = double(iter.num_object) / double(sy_it.num_object_after_interferences);
    }
    auto density = [](char const *o) {
        const double n = graphs::num_nodes(o);
        double d = 0;
        for (int i = 0; i < n; ++i) d += graphs::left(o, i) + graphs::right(o, i);
        return d / (2 * n);
    };
    const double avg_size = iter.average_value([](char const *o, char const *) { return double(graphs::num_nodes(o)); });
    const double avg_squared_size = iter.average_value([](char const *o, char const *) {
        const double n = graphs::num_nodes(o);
        return n * n;
    });
    const double avg_density = iter.average_value([&](char const *o, char const *) { return density(o); });
    const double avg_squared_density = iter.average_value([&](char const *o, char const *) {
        const double d = density(o);
        return d * d;
    });
    double std_dev_size = avg_squared_size - avg_size * avg_size;
    std_dev_size = std_dev_size < quids::tolerance ? 0 : std::sqrt(avg_squared_size);
    double std_dev_density = avg_squared_density - avg_density * avg_density;
    std_dev_density = std_dev_density < quids::tolerance ? 0 : std::sqrt(std_dev_density);
    const std::string ind(indentation, '\t');
    std::cout << "{\n";
    std::cout << ind << "\t\"total_proba\" : " << iter.total_proba << ",\n";
    std::cout << ind << "\t\"num_graphs\" : " << iter.num_object << ",\n";
    std::cout << ind << "\t\"avg_size\" : " << avg_size << ",\n";
    std::cout << ind << "\t\"std_dev_size\" : " << std_dev_size << ",\n";
    std::cout << ind << "\t\"avg_density\" : " << avg_density << ",\n";
    std::cout << ind << "\t\"std_dev_density\" : " << std_dev_density << ",\n";
    std::cout << ind << "\t\"interference_ratio\" : " << interference_ratio << ",\n";
    std::cout << ind << "\t\"deletion_ratio\" : " << deletion_ratio << "\n";
    std::cout << ind << "}";
}
}  // namespace utils

namespace detail {
/// do_ = polar(sin theta, phi), do_not = polar(cos theta, xi) and their
/// conjugates (qcgd.hpp:466-471, 541-546, 614-619)
inline void coefficients(qd_rule &d, int kind, double theta, double phi, double xi) {
    const std::complex<double> do_ = std::polar(std::sin(theta), phi), do_not = std::polar(std::cos(theta), xi);
    const std::complex<double> do_conj = std::conj(do_), do_not_conj = std::conj(do_not);
    d.kind = kind;
    d.do_re = do_.real();
    d.do_im = do_.imag();
    d.do_not_re = do_not.real();
    d.do_not_im = do_not.imag();
    d.do_conj_re = do_conj.real();
    d.do_conj_im = do_conj.imag();
    d.do_not_conj_re = do_not_conj.real();
    d.do_not_conj_im = do_not_conj.imag();
}
}  // namespace detail

/// qcgd.hpp:443-449
inline const modifier_t step(qd_modifier{QD_MOD_QCGD_STEP, 0, 0, 0});
/// qcgd.hpp:451-457
inline const modifier_t reversed_step(qd_modifier{QD_MOD_QCGD_REVERSED_STEP, 0, 0, 0});

/// qcgd.hpp:459-532
class erase_create : public device_rule_base {
public:
    erase_create(double theta, double phi = 0, double xi = 0) {
        detail::coefficients(desc_, QD_RULE_QCGD_ERASE_CREATE, theta, phi, xi);
    }
};
/// qcgd.hpp:534-605
class coin : public device_rule_base {
public:
    coin(double theta, double phi = 0, double xi = 0) { detail::coefficients(desc_, QD_RULE_QCGD_COIN, theta, phi, xi); }
};
/// qcgd.hpp:607-1036
class split_merge : public device_rule_base {
public:
    split_merge(double theta, double phi = 0, double xi = 0) {
        detail::coefficients(desc_, QD_RULE_QCGD_SPLIT_MERGE, theta, phi, xi);
    }
};

/// Command-line flags of the Quantum-graph-simulation runs
/// (qcgd.hpp:1038-1176): "<n_iter>,k=v,...|<n_nodes>,k=v,...;...|<rule>,k=v,...;..."
namespace flags {
/// per rule: (n_iter, is_rule, modifier, rule, reversed modifier, reversed rule)
typedef std::vector<std::tuple<int, bool, quids::modifier_t, quids::rule_t *, quids::modifier_t, quids::rule_t *>>
    simulator_t;

namespace detail {
// the text before the first `separator` (removed from input with it)
inline std::string strip(std::string &input, std::string const &separator) {
    const size_t end = input.find(separator);
    std::string head;
    if (end == std::string::npos) {
        head = input;
        input.clear();
    } else {
        head = input.substr(0, end);
        input.erase(0, end + separator.size());
    }
    return head;
}
// the text after `key`; the reference's substring length (first separator
// of the whole input minus the key's position, wrapping past the end when
// the key comes after that separator) is kept, the number parsers below
// stop at the first separator anyway
inline std::string value_of(std::string const &input, std::string const &key, std::string const &separator) {
    const size_t at = input.find(key);
    if (at == std::string::npos) return "";
    size_t end = input.find(separator);
    if (end == std::string::npos) end = input.size();
    return input.substr(at + key.size(), end - at);
}
inline int int_or(std::string const &input, std::string const &key, std::string const &sep, int otherwise) {
    const std::string v = value_of(input, key, sep);
    return v.empty() ? otherwise : std::atoi(v.c_str());
}
inline float float_or(std::string const &input, std::string const &key, std::string const &sep, float otherwise) {
    const std::string v = value_of(input, key, sep);
    return v.empty() ? otherwise : float(std::atof(v.c_str()));
}
}  // namespace detail

/// qcgd.hpp:1086-1117: "n_iter,seed=,reversed_n_iter=,max_print_num_graphs=,
/// tolerance=,safety_margin=,align=,simple_truncate=,
/// load_balancing_bucket_per_thread=,max_num_object=" -> (n_iter,
/// reversed_n_iter, max_num_object); seeds rand() (with the time when no
/// seed is given) and sets the globals (the tolerance passes through float,
/// as in the reference)
inline std::tuple<uint, uint, size_t> read_n_iter(const char *argv) {
    std::string arg = argv;
    const int n_iters = std::atoi(detail::strip(arg, ",").c_str());
    const std::string seed = detail::value_of(arg, "seed=", ",");
    if (!seed.empty()) std::srand(std::atoi(seed.c_str()));
    else std::srand(std::time(0));
    const int reversed_n_iters = detail::int_or(arg, "reversed_n_iter=", ",", 0);
    utils::max_print_num_graphs = detail::int_or(arg, "max_print_num_graphs=", ",", int(utils::max_print_num_graphs));
    quids::tolerance = detail::float_or(arg, "tolerance=", ",", float(quids::tolerance));
    quids::safety_margin = detail::float_or(arg, "safety_margin=", ",", quids::safety_margin);
    quids::align_byte_length = detail::int_or(arg, "align=", ",", int(quids::align_byte_length));
    quids::simple_truncation = detail::int_or(arg, "simple_truncate=", ",", quids::simple_truncation);
    quids::load_balancing_bucket_per_thread =
        detail::int_or(arg, "load_balancing_bucket_per_thread=", ",", quids::load_balancing_bucket_per_thread);
    const size_t max_num_object = detail::int_or(arg, "max_num_object=", ",", 0);
    return {uint(n_iters), uint(reversed_n_iters), max_num_object};
}

/// qcgd.hpp:1119-1137: "n_nodes,n_graphs=,real=,imag=;..." appends n_graphs
/// cycle graphs of n_nodes with magnitude (real + i imag) / sqrt(n_graphs)
/// (in float), then draws the particles of every graph of the state
inline void read_state(const char *argv, quids::it_t &state) {
    std::string args = argv, arg;
    while (!(arg = detail::strip(args, ";")).empty()) {
        const int n_node = std::atoi(detail::strip(arg, ",").c_str());
        const int n_graphs = detail::int_or(arg, "n_graphs=", ",", 1);
        const float real = detail::float_or(arg, "real=", ",", 1) / std::sqrt(float(n_graphs));
        const float imag = detail::float_or(arg, "imag=", ",", 0) / std::sqrt(float(n_graphs));
        for (int i = 0; i < n_graphs; ++i) {
            char *begin, *end;
            utils::make_graph(begin, end, uint16_t(n_node));
            state.append(begin, end, mag_t(real, imag));
            delete[] begin;
        }
    }
    utils::randomize(state);
}

/// qcgd.hpp:1139-1164: "name,theta=,phi=,xi=,n_iter=;..." with the angles in
/// units of pi (computed in float); a rule's reverse has -xi, step and
/// reversed_step reverse each other.  The caller owns the rules.
inline simulator_t read_rule(const char *argv, debug_t = [](const char *) {}) {
    std::string args = argv, arg;
    simulator_t simulator;
    while (!(arg = detail::strip(args, ";")).empty()) {
        const std::string name = detail::strip(arg, ",");
        const float theta = M_PI * detail::float_or(arg, "theta=", ",", 0.25);
        const float phi = M_PI * detail::float_or(arg, "phi=", ",", 0);
        const float xi = M_PI * detail::float_or(arg, "xi=", ",", 0);
        const int n_iter = detail::int_or(arg, "n_iter=", ",", 1);
        if (name == "split_merge")
            simulator.push_back({n_iter, true, modifier_t(), new split_merge(theta, phi, xi), modifier_t(),
                                 new split_merge(theta, phi, -xi)});
        else if (name == "erase_create")
            simulator.push_back({n_iter, true, modifier_t(), new erase_create(theta, phi, xi), modifier_t(),
                                 new erase_create(theta, phi, -xi)});
        else if (name == "coin")
            simulator.push_back(
                {n_iter, true, modifier_t(), new coin(theta, phi, xi), modifier_t(), new coin(theta, phi, -xi)});
        else if (name == "step")
            simulator.push_back({n_iter, false, step, nullptr, reversed_step, nullptr});
        else if (name == "reversed_step")
            simulator.push_back({n_iter, false, reversed_step, nullptr, step, nullptr});
    }
    return simulator;
}

/// qcgd.hpp:1166-1175: "<n_iter flags>|<state>|<rules>"
inline std::tuple<uint, uint, simulator_t, size_t> parse_simulation(const char *argv, it_t &state,
                                                                    debug_t mid_step_function = [](const char *) {}) {
    std::string args = argv;
    const auto [n_iter, reversed_n_iters, max_num_object] = read_n_iter(detail::strip(args, "|").c_str());
    read_state(detail::strip(args, "|").c_str(), state);
    auto simulator = read_rule(args.c_str(), mid_step_function);
    return {n_iter, reversed_n_iters, simulator, max_num_object};
}
}  // namespace flags

}  // namespace quids::rules::qcgd
