// Quantum causal graph dynamics rules of the drop-in C++ API (reference:
// src/rules/qcgd.hpp): graph layout helpers, erase_create, coin and
// split_merge rules and the step / reversed_step modifiers, as device
// functors of libquids_hip.
#pragma once

#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>

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

}  // namespace quids::rules::qcgd
