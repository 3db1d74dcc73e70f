// qd_device.hpp -- device-side building blocks shared by the step kernels:
// block scans, the canonical |m|^2 sum, and the candidate entry format.
#pragma once

#include <hip/hip_runtime.h>

#include "qd_rules.hpp"

namespace qd {

constexpr int kWave = 64;

// Candidates after the first partition pass, as three arrays: hash, index
// in the candidate space (parents in order, then child_id:
// src/quids.hpp:666-694) and magnitude.  Generation writes only hash and
// magnitude, at the candidate's index.
struct Cands {
    u64 *h;
    u64 *c;
    double2 *m;
};

// One candidate in the partition passes: hash, candidate index, magnitude
// (32 bytes, two dwordx4 accesses, one scattered write per candidate).
struct __align__(16) Entry {
    u64 hash;
    u64 cand;
    double re, im;
};

// Raises *m to at least v.  Atomics on one address from every wavefront of
// a large grid serialise (about 11 ns each, measured), so the atomic is
// issued only when a plain read says it can raise the value; a stale read
// costs one atomic more, never a wrong result.
__device__ inline void raise_max(u32 *m, u32 v) {
    if (v > __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(m, v);
}

__device__ inline u64 wave_incl_scan(u64 v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        u64 o = __shfl_up(v, d, kWave);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive scan of one value per thread over the block; returns the
// exclusive prefix and the block total.  `tmp` holds blockDim/64 + 1 u64.
__device__ inline u64 block_excl_scan(u64 v, u64 *tmp, u64 &total) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const u64 inc = wave_incl_scan(v);
    if (lane == kWave - 1) tmp[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 s = 0;
        for (int w = 0; w < nw; ++w) {
            const u64 t = tmp[w];
            tmp[w] = s;
            s += t;
        }
        tmp[nw] = s;
    }
    __syncthreads();
    const u64 r = tmp[wid] + inc - v;
    total = tmp[nw];
    __syncthreads();
    return r;
}

}  // namespace qd
