#pragma once
// survivor placement, child sizes and bytes, normalisation sums, modifiers
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

// ===========================================================================
// phase 7/8: order, sizes, populate
// ===========================================================================
__global__ __launch_bounds__(256) void k_place(const u64 *UR, const double2 *UM, u64 U, const u64 *bitmap,
                                               const u64 *wprefix, u64 *orep, double2 *omag) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < U; i += u64(gridDim.x) * blockDim.x) {
        const u64 rep = UR[i];
        if (rep == kNoRep) continue;
        const u64 word = bitmap[rep >> 6];
        const u32 bit = u32(rep & 63);
        if ((word >> bit) & 1) {
            const u64 r = wprefix[rep >> 6] + __popcll(word & ((1ull << bit) - 1));
            orep[r] = rep;
            omag[r] = UM[i];
        }
    }
}

__device__ inline u32 align_pad(u32 size, u32 align) {
    if (align <= 1) return 0;
    const u32 off = align - size % align;
    return off == align ? 0 : off;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_child_size(Parents P, const u64 *cb, qd_rule rule, const u64 *orep, u64 K,
                                                    u32 align, u32 *osize, u64 *padded) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x) {
        const u64 c = orep[k];
        const u64 p = find_parent(cb, P.n, c);
        const u32 j = u32(c - cb[p]);
        const u8 *o = P.obj + P.begin[p];
        u32 s;
        if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
            Graph g(o);
            s = sm_child_size(g, sm_first_hmlz(g), j);
        } else {
            s = P.size[p];
        }
        osize[k] = s;
        padded[k] = u64(s) + align_pad(s, align);
    }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_populate(Parents P, const u64 *cb, qd_rule rule, const u64 *orep, u64 K,
                                                  const u64 *obegin, const u32 *osize, u8 *oobj) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x) {
        const u64 c = orep[k];
        const u64 p = find_parent(cb, P.n, c);
        const u32 j = u32(c - cb[p]);
        const u8 *o = P.obj + P.begin[p];
        u8 *d = oobj + obegin[k];
        u32 s;
        if (KIND == QD_RULE_HADAMARD) {
            hadamard_child_write(o, P.size[p], rule.bit, j, d);
            s = P.size[p];
        } else if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
            Graph g(o);
            s = sm_child_write(g, sm_first_hmlz(g), j, d);
        } else {
            flip_child_write(o, P.size[p], KIND == QD_RULE_QCGD_COIN, j, d);
            s = P.size[p];
        }
        const u64 e = obegin[k + 1];
        for (u64 q = obegin[k] + s; q < e; ++q) oobj[q] = 0;
        (void)osize;
    }
}

// ===========================================================================
// phase 9: canonical sum of |m|^2 and scaling
// ===========================================================================
constexpr int kSumThreads = 256;
constexpr int kSumChunk = 2048;

// One block per chunk of 2048 values: lane t sums t, t+256, ..., t+1792 in
// order, then a halving tree s[t] += s[t+stride], stride = 128 .. 1.
__global__ __launch_bounds__(kSumThreads) void k_sum_chunks(const double2 *mag, const double *vals, u64 n,
                                                            double *out) {
    __shared__ double s[kSumThreads];
    const u64 base = u64(blockIdx.x) * kSumChunk;
    double a = 0.0;
    for (int j = 0; j < kSumChunk / kSumThreads; ++j) {
        const u64 i = base + threadIdx.x + u64(kSumThreads) * j;
        if (i < n) {
            double v;
            if (mag) {
                const double2 m = mag[i];
                v = m.x * m.x + m.y * m.y;
            } else {
                v = vals[i];
            }
            a += v;
        }
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int stride = kSumThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_scale(double2 *mag, u64 n, double f) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        double2 m = mag[i];
        m.x = m.x / f;
        m.y = m.y / f;
        mag[i] = m;
    }
}

// ===========================================================================
// modifiers
// ===========================================================================
__global__ __launch_bounds__(256) void k_modifier(u8 *obj, const u64 *begin, const u32 *size, double2 *mag, u64 n,
                                                  qd_modifier m, u32 *err) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        u8 *o = obj + begin[i];
        const u32 sz = size[i];
        if (m.kind <= QD_MOD_ZGATE) {
            if (m.bit >= sz || (m.kind == QD_MOD_CNOT && m.control_bit >= sz)) {
                atomicOr(err, ERR_BIT);
                continue;
            }
        } else if (!qcgd_valid(o, sz)) {
            atomicOr(err, ERR_OBJECT);
            continue;
        }
        double2 mm = mag[i];
        cplx c{mm.x, mm.y};
        modifier_apply(m, o, sz, c);
        mag[i] = make_double2(c.re, c.im);
    }
}


}  // namespace qd
