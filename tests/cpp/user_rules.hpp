// User rules and observables of the tests, written as device functors
// against quids_amd/user_rule.hpp the way a user of the reference would
// write a quids::rule subclass (src/quids.hpp:105-146).
#pragma once
#include "../../quids_amd/user_rule.hpp"

// The reference's hadamard (src/rules/quantum_computer.hpp:31-50) restated
// as a user rule: two children, the qubit byte set or cleared, magnitude
// times -coef for the child of a set qubit that keeps it, coef otherwise.
struct user_hadamard {
    uint32_t bit;
    double coef;  // 1/std::sqrt(2), computed on the host as the reference does
    QD_HD void get_num_child(const char *pb, const char *pe, uint32_t &num_child, uint32_t &max_child_size) const {
        num_child = 2;
        max_child_size = uint32_t(pe - pb);
    }
    QD_HD void populate_child(const char *pb, const char *pe, char *child, uint32_t child_id, uint32_t &size,
                              quids::dmag &mag) const {
        mag *= pb[bit] && child_id ? -coef : coef;
        size = uint32_t(pe - pb);
        for (uint32_t i = 0; i < size; ++i) child[i] = pb[i];
        child[bit] ^= !child_id;
    }
};

// A rule with no built-in counterpart: the real rotation RY on one qubit
// byte, |0> -> c|0> + s|1>, |1> -> -s|0> + c|1>.  Child 0 keeps the byte
// (times c), child 1 flips it (times s from 0, -s from 1).  It writes its
// children through populate_child only (no populate_child_simple) and
// hashes them with its own hasher (FNV-1a), to cover those defaults.
struct user_ry {
    uint32_t bit;
    double c, s;
    QD_HD void get_num_child(const char *pb, const char *pe, uint32_t &num_child, uint32_t &max_child_size) const {
        num_child = 2;
        max_child_size = uint32_t(pe - pb);
    }
    QD_HD void populate_child(const char *pb, const char *pe, char *child, uint32_t child_id, uint32_t &size,
                              quids::dmag &mag) const {
        size = uint32_t(pe - pb);
        for (uint32_t i = 0; i < size; ++i) child[i] = pb[i];
        if (child_id == 0) {
            mag *= c;
        } else {
            mag *= pb[bit] ? -s : s;
            child[bit] = pb[bit] ? 0 : 1;
        }
    }
    QD_HD uint64_t hasher(const char *b, const char *e) const {
        uint64_t h = 1469598103934665603ull;
        for (const char *p = b; p < e; ++p) h = (h ^ uint64_t(uint8_t(*p))) * 1099511628211ull;
        return h;
    }
};

// The QCGD node count (the first 16-bit word of an object,
// src/rules/qcgd.hpp:11-60), the observable qd_average_value calls 0.
struct obs_num_nodes {
    QD_HD double operator()(const char *b, const char *) const {
        return double(uint32_t(uint8_t(b[0])) | (uint32_t(uint8_t(b[1])) << 8));
    }
};

// The number of set qubit bytes of a register.
struct obs_ones {
    QD_HD double operator()(const char *b, const char *e) const {
        double n = 0;
        for (const char *p = b; p < e; ++p) n += *p ? 1.0 : 0.0;
        return n;
    }
};
