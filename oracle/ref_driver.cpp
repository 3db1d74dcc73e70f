// ref_driver.cpp -- runs the reference QuIDS headers on a program file.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference
// sources where they lie (-I <reference>/src) into oracle/_ref/ref_driver; the
// reference sources are never copied.  It produces golden vectors
// (tests/golden/gen_golden.py) and is the timed CPU baseline (bench.py).
//
// Program file (little-endian):
//   char magic[4] = "QDP1"; u32 family (0 = quantum_computer, 1 = qcgd);
//   u32 n_ops; u32 align_byte_length; f64 tolerance; u32 n_obj; u32 pad;
//   n_obj x { u32 size; u32 pad; f64 re; f64 im; u8 bytes[size] }
//   n_ops x { u32 type (1 rule, 2 modifier); u32 kind (qd_rule_kind /
//             qd_modifier_kind); u32 a; u32 b; f64 x, y, z; u64 max_num_object;
//             u32 dump; u32 pad }
//     hadamard: a = bit.  QCGD rules: x, y, z = theta, phi, xi.
//     cnot: a = control, b = target.  X/Y/Z: b = bit.
// Output file: for every op with dump != 0:
//   u64 N; u64 C; u64 U; f64 total_proba; f64 seconds; u64 object_bytes;
//   N x { u32 size; u32 pad; f64 re; f64 im; u64 hash; u8 bytes[size] }
// Timing of each rule op goes to stdout as "step <i> C=<C> seconds=<s>".

#define SIMPLE_TRUNCATION
#include "quids.hpp"
#include "rules/qcgd.hpp"
#include "rules/quantum_computer.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string_view>
#include <vector>

#include "../include/quids_hip.h"

namespace {

template <class T>
T rd(std::istream &is) {
    T v;
    is.read(reinterpret_cast<char *>(&v), sizeof v);
    if (!is) throw std::runtime_error("truncated program file");
    return v;
}
template <class T>
void wr(std::ostream &os, T v) {
    os.write(reinterpret_cast<const char *>(&v), sizeof v);
}

struct Op {
    uint32_t type, kind, a, b;
    double x, y, z;
    uint64_t max_num_object;
    uint32_t dump, pad;
};

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::cerr << "usage: ref_driver program.bin out.bin\n";
        return 2;
    }
    std::ifstream is(argv[1], std::ios::binary);
    char magic[4];
    is.read(magic, 4);
    if (std::memcmp(magic, "QDP1", 4) != 0) {
        std::cerr << "bad magic\n";
        return 2;
    }
    const uint32_t family = rd<uint32_t>(is);
    const uint32_t n_ops = rd<uint32_t>(is);
    quids::align_byte_length = rd<uint32_t>(is);
    quids::tolerance = rd<double>(is);
    const uint32_t n_obj = rd<uint32_t>(is);
    rd<uint32_t>(is);

    quids::it_t state, buffer;
    quids::sy_it_t sy_it;
    for (uint32_t i = 0; i < n_obj; ++i) {
        const uint32_t size = rd<uint32_t>(is);
        rd<uint32_t>(is);
        const double re = rd<double>(is), im = rd<double>(is);
        std::vector<char> bytes(size);
        is.read(bytes.data(), size);
        state.append(bytes.data(), bytes.data() + size, quids::mag_t(re, im));
    }
    std::vector<Op> ops(n_ops);
    for (auto &op : ops) is.read(reinterpret_cast<char *>(&op), sizeof op);
    if (!is) {
        std::cerr << "truncated ops\n";
        return 2;
    }

    std::ofstream os(argv[2], std::ios::binary);
    quids::it_t *cur = &state, *nxt = &buffer;
    for (uint32_t i = 0; i < n_ops; ++i) {
        const Op &op = ops[i];
        double seconds = 0;
        uint64_t C = 0, U = 0;
        if (op.type == 1) {
            quids::rule_t *rule = nullptr;
            switch (op.kind) {
            case QD_RULE_HADAMARD: rule = new quids::rules::quantum_computer::hadamard(op.a); break;
            case QD_RULE_QCGD_ERASE_CREATE: rule = new quids::rules::qcgd::erase_create(op.x, op.y, op.z); break;
            case QD_RULE_QCGD_COIN: rule = new quids::rules::qcgd::coin(op.x, op.y, op.z); break;
            case QD_RULE_QCGD_SPLIT_MERGE: rule = new quids::rules::qcgd::split_merge(op.x, op.y, op.z); break;
            default: std::cerr << "unknown rule\n"; return 2;
            }
            auto t0 = std::chrono::steady_clock::now();
            quids::simulate(*cur, rule, *nxt, sy_it, op.max_num_object);
            auto t1 = std::chrono::steady_clock::now();
            seconds = std::chrono::duration<double>(t1 - t0).count();
            C = sy_it.num_object;
            U = sy_it.num_object_after_interferences;
            std::swap(cur, nxt);
            delete rule;
            std::printf("step %u C=%llu U=%llu N=%llu seconds=%.6f\n", i, (unsigned long long)C,
                        (unsigned long long)U, (unsigned long long)cur->num_object, seconds);
            std::fflush(stdout);
        } else {
            quids::modifier_t mod;
            switch (op.kind) {
            case QD_MOD_CNOT: mod = quids::rules::quantum_computer::cnot(op.a, op.b); break;
            case QD_MOD_XGATE: mod = quids::rules::quantum_computer::Xgate(op.b); break;
            case QD_MOD_YGATE: mod = quids::rules::quantum_computer::Ygate(op.b); break;
            case QD_MOD_ZGATE: mod = quids::rules::quantum_computer::Zgate(op.b); break;
            case QD_MOD_QCGD_STEP: mod = quids::rules::qcgd::step; break;
            case QD_MOD_QCGD_REVERSED_STEP: mod = quids::rules::qcgd::reversed_step; break;
            default: std::cerr << "unknown modifier\n"; return 2;
            }
            quids::simulate(*cur, mod);
        }
        if (op.dump) {
            const uint64_t N = cur->num_object;
            uint64_t bytes = 0;
            for (uint64_t k = 0; k < N; ++k) {
                uint32_t sz;
                quids::mag_t m;
                char const *b;
                cur->get_object(k, b, sz, m);
                bytes += sz;
            }
            wr(os, N);
            wr(os, C);
            wr(os, U);
            wr(os, cur->total_proba);
            wr(os, seconds);
            wr(os, bytes);
            for (uint64_t k = 0; k < N; ++k) {
                uint32_t sz;
                quids::mag_t m;
                char const *b;
                cur->get_object(k, b, sz, m);
                uint64_t h = family == 1 ? quids::rules::qcgd::graphs::hash_graph(b)
                                         : std::hash<std::string_view>()(std::string_view(b, sz));
                wr(os, sz);
                wr(os, uint32_t(0));
                wr(os, m.real());
                wr(os, m.imag());
                wr(os, h);
                os.write(b, sz);
            }
        }
    }
    return 0;
}
