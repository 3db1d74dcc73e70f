// A user rule and user observables through the drop-in C++ API, the way a
// user of the reference plugs in a quids::rule subclass (README.md:112-180):
// quids::device_rule<user_hadamard> against the built-in hadamard over a
// seeded circuit (every state bit for bit), device observables against host
// observable_t lambdas, and the rule's host-side virtual methods against the
// built-in rule's.  Prints "user_rule_example ok" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../quids_amd/rules/qcgd.hpp"
#include "../../quids_amd/rules/quantum_computer.hpp"
#include "user_rules.hpp"

static int failures = 0;
#define CHECK(cond, ...)                         \
    do {                                         \
        if (!(cond)) {                           \
            std::printf("FAIL: " __VA_ARGS__);   \
            std::printf("\n");                   \
            ++failures;                          \
        }                                        \
    } while (0)

static bool same_state(quids::it_t const &a, quids::it_t const &b) {
    if (a.num_object != b.num_object || a.total_proba != b.total_proba) return false;
    for (size_t i = 0; i < a.num_object; ++i) {
        char const *pa, *pb;
        quids::uint sa, sb;
        quids::mag_t ma, mb;
        a.get_object(i, pa, sa, ma);
        b.get_object(i, pb, sb, mb);
        if (sa != sb || std::memcmp(pa, pb, sa) != 0 || ma != mb) return false;
    }
    return true;
}

int main() {
    using namespace quids::rules;
    const double coef = 1 / std::sqrt(2);
    // ---- a seeded 12-qubit circuit, built-in and user hadamards side by side
    {
        quids::align_byte_length = 8;
        const int nq = 12;
        std::vector<char> reg(nq, 0);
        quids::it_t s0, s1, u0, u1;  // (states cannot be copied: pointers swap)
        quids::it_t *a = &s0, *b = &s1, *ua = &u0, *ub = &u1;
        a->append(reg.data(), reg.data() + nq);
        ua->append(reg.data(), reg.data() + nq);
        quids::sy_it_t sy, usy;
        std::mt19937 gen(7);
        for (int step = 0; step < 16; ++step) {
            const uint32_t q = gen() % nq;
            if (step % 3 == 2) {
                const uint32_t c = (q + 1 + gen() % (nq - 1)) % nq;
                quids::simulate(*a, quantum_computer::cnot(c, q));
                quids::simulate(*ua, quantum_computer::cnot(c, q));
                continue;
            }
            quantum_computer::hadamard H(q);
            quids::device_rule<user_hadamard> UH(user_hadamard{q, coef});
            const size_t cap = step < 10 ? size_t(-1) : 300;  // truncation in the last steps
            quids::simulate(*a, &H, *b, sy, cap);
            quids::simulate(*ua, &UH, *ub, usy, cap);
            CHECK(sy.num_object == usy.num_object && sy.num_object_after_interferences == usy.num_object_after_interferences,
                  "step %d: counters differ", step);
            CHECK(same_state(*b, *ub), "step %d: states differ", step);
            std::swap(a, b);
            std::swap(ua, ub);
        }
        // average_value: device observable against the host observable_t
        const quids::device_observable<obs_ones> dev{obs_ones{}};
        const double host = ua->average_value([](char const *b, char const *e) { return obs_ones{}(b, e); });
        const double dv = ua->average_value(dev);
        CHECK(std::abs(dv - host) <= 1e-12 * std::max(1.0, std::abs(host)), "average_value %.17g vs %.17g", dv, host);
    }
    // ---- host-side virtual methods of the user rule against the built-in's
    {
        char p[6] = {1, 0, 1, 1, 0, 0};
        quantum_computer::hadamard H(2);
        quids::device_rule<user_hadamard> UH(user_hadamard{2, coef});
        quids::uint n1, m1, n2, m2;
        H.get_num_child(p, p + 6, n1, m1);
        UH.get_num_child(p, p + 6, n2, m2);
        CHECK(n1 == n2 && m1 == m2, "get_num_child");
        for (quids::uint id = 0; id < 2; ++id) {
            char c1[6], c2[6];
            quids::uint s1, s2;
            quids::mag_t g1(0.25, -0.5), g2(0.25, -0.5);
            H.populate_child(p, p + 6, c1, id, s1, g1);
            UH.populate_child(p, p + 6, c2, id, s2, g2);
            CHECK(s1 == s2 && std::memcmp(c1, c2, s1) == 0 && g1 == g2, "populate_child %u", id);
            CHECK(H.hasher(c1, c1 + s1) == UH.hasher(c2, c2 + s2), "hasher %u", id);
        }
    }
    // ---- a QCGD state: the device node-count observable against the host
    {
        quids::align_byte_length = 8;
        quids::it_t it, nx;
        quids::sy_it_t sy;
        char *gb, *ge;
        qcgd::utils::make_graph(gb, ge, 7);
        it.append(gb, ge);
        delete[] gb;
        qcgd::utils::randomize(it);
        qcgd::split_merge sm(M_PI / 4, 0);
        qcgd::erase_create ec(M_PI / 3, M_PI / 5);
        quids::simulate(it, &sm, nx, sy, size_t(-1));
        quids::simulate(nx, &ec, it, sy, size_t(-1));
        const double host = it.average_value([](char const *b, char const *e) { return obs_num_nodes{}(b, e); });
        const double dv = it.average_value(quids::device_observable<obs_num_nodes>{obs_num_nodes{}});
        CHECK(std::abs(dv - host) <= 1e-12 * std::max(1.0, std::abs(host)), "qcgd average_value %.17g vs %.17g", dv,
              host);
    }
    if (failures) return 1;
    std::printf("user_rule_example ok\n");
    return 0;
}
