// ref_cli.cpp -- the reference's QCGD command-line flags, statistics and
// graph printing on one run.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference
// sources where they lie (-I <reference>/src) into oracle/_ref/ref_cli; the
// reference sources are never copied.  tests/golden/gen_golden_cli.py runs
// it to write the text fixtures tests/golden/qcgd_cli_*.txt.
//
// Usage: ref_cli "<n_iter flags>|<state>|<rules>"
// Runs flags::parse_simulation, then n_iter iterations of the rule list
// (each entry n_iter times; rules step into the other buffer) and
// reversed_n_iter iterations of the reversed list, printing
// utils::serialize after the parse and after every iteration as the
// elements of a JSON list, then utils::print of the final state.
#include "quids.hpp"
#include "rules/qcgd.hpp"

#include <iostream>
#include <utility>

int main(int argc, char **argv) {
    if (argc != 2) {
        std::cerr << "usage: ref_cli \"<n_iter flags>|<state>|<rules>\"\n";
        return 2;
    }
    namespace qcgd = quids::rules::qcgd;
    quids::it_t a, b;
    quids::sy_it_t sy;
    auto [n_iter, reversed_n_iter, simulator, max_num_object] = qcgd::flags::parse_simulation(argv[1], a);
    quids::it_t *cur = &a, *nxt = &b;
    std::cout << "[\n\t";
    qcgd::utils::serialize(*cur, sy, 1);
    for (unsigned i = 0; i < n_iter; ++i) {
        for (auto &[n, is_rule, modifier, rule, reversed_modifier, reversed_rule] : simulator)
            for (int j = 0; j < n; ++j) {
                if (is_rule) {
                    quids::simulate(*cur, rule, *nxt, sy, max_num_object);
                    std::swap(cur, nxt);
                } else {
                    quids::simulate(*cur, modifier);
                }
            }
        std::cout << ",\n\t";
        qcgd::utils::serialize(*cur, sy, 1);
    }
    for (unsigned i = 0; i < reversed_n_iter; ++i) {
        for (auto it = simulator.rbegin(); it != simulator.rend(); ++it) {
            auto &[n, is_rule, modifier, rule, reversed_modifier, reversed_rule] = *it;
            for (int j = 0; j < n; ++j) {
                if (is_rule) {
                    quids::simulate(*cur, reversed_rule, *nxt, sy, max_num_object);
                    std::swap(cur, nxt);
                } else {
                    quids::simulate(*cur, reversed_modifier);
                }
            }
        }
        std::cout << ",\n\t";
        qcgd::utils::serialize(*cur, sy, 1);
    }
    std::cout << "\n]\n";
    qcgd::utils::print(*cur);
    return 0;
}
