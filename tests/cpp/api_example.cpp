// The reference's two example programs (examples/quantum_computer_test.cpp,
// examples/qcgd_test.cpp) written against the drop-in C++ API, with a
// seeded QCGD state.  Prints one line per checkpoint:
//   <label> <num_object> <total_proba> <sum |m|^2> <hash xor>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../quids_amd/quids.hpp"
#include "../../quids_amd/rules/qcgd.hpp"
#include "../../quids_amd/rules/quantum_computer.hpp"

template <class H>
static void report(const char *label, quids::it_t const &it, H hasher) {
    double s = 0;
    unsigned long long x = 0;
    for (size_t i = 0; i < it.num_object; ++i) {
        char const *b;
        quids::uint sz;
        quids::mag_t m;
        it.get_object(i, b, sz, m);
        s += std::norm(m);
        x ^= (unsigned long long)hasher(b, b + sz);
    }
    std::printf("%s %zu %.17g %.17g %016llx\n", label, it.num_object, it.total_proba, s, x);
}

int main() {
    using namespace quids::rules;
    {
        quids::align_byte_length = 0;
        quids::rule_t *H1 = new quantum_computer::hadamard(1);
        quids::rule_t *H2 = new quantum_computer::hadamard(2);
        quids::modifier_t CNOT = quantum_computer::cnot(1, 3);
        quids::modifier_t X2 = quantum_computer::Xgate(2);
        quids::modifier_t Y0 = quantum_computer::Ygate(0);
        quids::modifier_t Z3 = quantum_computer::Zgate(3);
        quids::sy_it_t sy_it;
        quids::it_t buffer, state;
        char s1[4] = {true, true, false, false};
        char s2[5] = {false, true, true, false, true};
        state.append(s1, s1 + 4, 1 / std::sqrt(2));
        state.append(s2, s2 + 5, {0, 1 / std::sqrt(2)});
        auto hh = [&](char const *b, char const *e) { return H1->hasher(b, e); };
        quids::simulate(state, H1, buffer, sy_it, (size_t)-1);
        report("qc_h1", buffer, hh);
        quids::simulate(buffer, H2, state, sy_it, (size_t)-1);
        report("qc_h2", state, hh);
        quids::simulate(state, CNOT);
        quids::simulate(state, X2);
        quids::simulate(state, Y0);
        quids::simulate(state, Z3);
        report("qc_gates", state, hh);
        quids::simulate(state, Z3);
        quids::simulate(state, Y0);
        quids::simulate(state, X2);
        quids::simulate(state, CNOT);
        quids::simulate(state, H2, buffer, sy_it, (size_t)-1);
        quids::simulate(buffer, H1, state, sy_it, (size_t)-1);
        report("qc_reversed", state, hh);
        // the reference's Zgate squares to -I: the initial state times -1
        for (size_t i = 0; i < state.num_object; ++i) {
            char const *b;
            quids::uint sz;
            quids::mag_t m;
            state.get_object(i, b, sz, m);
            std::printf("qc_final %u %.17g %.17g\n", sz, m.real(), m.imag());
        }
    }
    {
        quids::align_byte_length = 8;
        quids::tolerance = 1e-15;
        quids::sy_it_t sy_it;
        quids::it_t state, buffer;
        char *b, *e;
        qcgd::utils::make_graph(b, e, 6);
        state.append(b, e);
        std::srand(5);
        qcgd::utils::randomize(state);
        delete[] b;
        quids::rule_t *erase_create = new qcgd::erase_create(0.3333);
        quids::rule_t *coin = new qcgd::erase_create(0.25, 0.25);
        quids::rule_t *split_merge = new qcgd::split_merge(0.25, 0.25, 0.25);
        quids::rule_t *reversed_split_merge = new qcgd::split_merge(0.25, 0.25, -0.25);
        auto hg = [](char const *b, char const *) { return qcgd::graphs::hash_graph(b); };
        report("qcgd_initial", state, hg);
        quids::simulate(state, qcgd::step);
        quids::simulate(state, coin, buffer, sy_it, (size_t)-1);
        quids::simulate(buffer, coin, state, sy_it, (size_t)-1);
        quids::simulate(state, erase_create, buffer, sy_it, (size_t)-1);
        report("qcgd_coin_ec", buffer, hg);
        quids::simulate(buffer, split_merge, state, sy_it, (size_t)-1);
        quids::simulate(state, qcgd::step);
        quids::simulate(state, split_merge, buffer, sy_it, (size_t)-1);
        report("qcgd_sm", buffer, hg);
        quids::simulate(buffer, reversed_split_merge, state, sy_it, (size_t)-1);
        quids::simulate(state, qcgd::reversed_step);
        quids::simulate(state, reversed_split_merge, buffer, sy_it, (size_t)-1);
        quids::simulate(buffer, erase_create, state, sy_it, (size_t)-1);
        quids::simulate(state, qcgd::reversed_step);
        report("qcgd_reversed", state, hg);
    }
    return 0;
}
