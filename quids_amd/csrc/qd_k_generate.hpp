#pragma once
// per-parent prep and candidate generation (generic byte-level path)
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

// ===========================================================================
// phase 1: per-parent prep
// ===========================================================================
struct Parents {
    const u8 *obj;
    const u64 *begin;
    const u32 *size;
    const double2 *mag;
    u64 n;
};

// Validation bits reported through the error word.
enum : u32 { ERR_OBJECT = 1u, ERR_OVERFLOW = 2u, ERR_BIT = 4u };

// A QCGD object is well formed when its size matches its header and every
// node's name tree starts with a root the rules can read: an element, a .l /
// .r node with its inner node, or a pair with two non-empty halves.  Device
// code reads only roots, root+1 and root+right_or_type, so this bounds every
// access and every child size.
__device__ inline bool qcgd_valid(const u8 *o, u32 size) {
    if (size < 4) return false;
    const u32 n = ld16(o);
    if (n == 0 || 4 + 4 * u64(n) > size) return false;
    const u8 *nb = o + 2 + 2 * n;
    if (ld16(nb) != 0) return false;
    const u32 m = ld16(nb + 2 * n);
    if (4 + 4 * u64(n) + 16 * u64(m) != size) return false;
    Graph g(o);
    for (u32 i = 0; i < n; ++i) {
        const u32 b = ld16(nb + 2 * i), e = ld16(nb + 2 * i + 2);
        if (e <= b || e > m) return false;
        const u32 mi = e - b;
        const i16 t = g.sub_type(b);
        if (t >= 0) {
            if (t < 2 || u32(t) >= mi) return false;
        } else if (t == QCGD_DOT_L || t == QCGD_DOT_R) {
            if (mi < 2) return false;
        } else if (t != QCGD_ELEMENT) {
            return false;
        }
    }
    return true;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_prep(Parents P, const u64 *keep, qd_rule rule, u32 *nc, u64 *pre, u64 *tab,
                                              u32 *maxsz, u32 *err) {
    u32 local_max = 0;
    for (u64 p = u64(blockIdx.x) * blockDim.x + threadIdx.x; p < P.n; p += u64(gridDim.x) * blockDim.x) {
        const u8 *o = P.obj + P.begin[p];
        const u32 sz = P.size[p];
        if (keep && !((keep[p >> 6] >> (p & 63)) & 1)) {
            nc[p] = 0;
            continue;
        }
        u32 num = 0, ms = 0;
        if (KIND == QD_RULE_HADAMARD) {
            if (rule.bit >= sz) {
                atomicOr(err, ERR_BIT);
                nc[p] = 0;
                continue;
            }
            num = 2;
            ms = sz;
        } else {
            if (!qcgd_valid(o, sz)) {
                atomicOr(err, ERR_OBJECT);
                nc[p] = 0;
                continue;
            }
            Graph g(o);
            if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
                num = sm_num_child(g);
                ms = 4 * sz;
                sm_fill_table(g, tab + sm_tab_base(P.begin[p]));
                pre[p] = sm_first_hmlz(g) ? 1 : 0;
            } else {
                num = flip_num_child(g, KIND == QD_RULE_QCGD_COIN);
                ms = sz;
                pre[p] = name_fold(g);
            }
        }
        nc[p] = num;
        local_max = max(local_max, ms);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) local_max = max(local_max, (u32)__shfl_xor(int(local_max), off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) raise_max(maxsz, local_max);
}

// num_child of every parent alone (the same checks as k_prep, no tables):
// the batched step's planning pass.
template <int KIND>
__global__ __launch_bounds__(256) void k_num_child(Parents P, qd_rule rule, u32 *nc, u32 *err) {
    for (u64 p = u64(blockIdx.x) * blockDim.x + threadIdx.x; p < P.n; p += u64(gridDim.x) * blockDim.x) {
        const u8 *o = P.obj + P.begin[p];
        const u32 sz = P.size[p];
        u32 num = 0;
        if (KIND == QD_RULE_HADAMARD) {
            if (rule.bit >= sz) atomicOr(err, ERR_BIT);
            else num = 2;
        } else if (!qcgd_valid(o, sz)) {
            atomicOr(err, ERR_OBJECT);
        } else {
            Graph g(o);
            num = KIND == QD_RULE_QCGD_SPLIT_MERGE ? sm_num_child(g) : flip_num_child(g, KIND == QD_RULE_QCGD_COIN);
        }
        nc[p] = num;
    }
}

// ===========================================================================
// phase 3: generate
// ===========================================================================
__device__ inline u64 find_parent(const u64 *cb, u64 n, u64 c) {
    // largest p in [0, n) with cb[p] <= c (cb has n+1 entries, cb[n] = C > c)
    u64 lo = 0, hi = n - 1;
    while (lo < hi) {
        const u64 mid = (lo + hi + 1) >> 1;
        if (cb[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

constexpr int kHistBins = 4096;  // up to 12 bits of LDS histogram

template <int KIND>
__global__ __launch_bounds__(1024) void k_generate(Parents P, const u64 *cb, qd_rule rule, const u64 *pre,
                                                  const u64 *tab, u64 C, u64 *H0, double2 *M0, u32 b1,
                                                  u64 *hist1) {
    __shared__ u32 hist[kHistBins];
    if (b1)
        for (int i = threadIdx.x; i < (1 << b1); i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const QcgdCoef k = coef_of(rule);
    for (u64 c = u64(blockIdx.x) * blockDim.x + threadIdx.x; c < C; c += u64(gridDim.x) * blockDim.x) {
        const u64 p = find_parent(cb, P.n, c);
        const u32 j = u32(c - cb[p]);
        const double2 m0 = P.mag[p];
        cplx mag{m0.x, m0.y};
        const u8 *o = P.obj + P.begin[p];
        u64 h;
        if (KIND == QD_RULE_HADAMARD) {
            h = hadamard_child_hash(o, P.size[p], rule.bit, rule.coef, j, mag);
        } else if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
            Graph g(o);
            h = sm_child_hash(g, pre[p] != 0, tab + sm_tab_base(P.begin[p]), j, k, mag);
        } else {
            Graph g(o);
            h = flip_child_hash(g, KIND == QD_RULE_QCGD_COIN, pre[p], j, k, mag);
        }
        H0[c] = h;
        M0[c] = make_double2(mag.re, mag.im);
        if (b1) atomicAdd(&hist[h >> (64 - b1)], 1u);
    }
    if (b1) {
        __syncthreads();
        for (int i = threadIdx.x; i < (1 << b1); i += blockDim.x)
            if (hist[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist1 + i), (unsigned long long)hist[i]);
    }
}


}  // namespace qd
