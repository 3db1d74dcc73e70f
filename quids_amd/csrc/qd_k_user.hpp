#pragma once
// Kernels of the step with a user rule (a device functor compiled by the
// caller, include/quids_hip.h qd_user_rule): the rule's own kernels produce
// num_child, the candidates' hashes, magnitudes and sizes, and the
// survivors' bytes; these kernels glue them to the rule-independent phases
// (scan, partition + dedup, select, order).
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_k_dedup.hpp"
#include "qd_k_finalize.hpp"
#include "qd_k_generate.hpp"

namespace qd {

// Parents dropped by the pre-truncation (keep bit clear) get no children;
// *maxsz = the largest max_child_size of the parents that have some.
__global__ __launch_bounds__(256) void k_user_keep(u32 *nc, const u32 *msz, const u64 *keep, u64 N, u32 *maxsz) {
    u32 m = 0;
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += u64(gridDim.x) * blockDim.x) {
        if (keep && !((keep[i >> 6] >> (i & 63)) & 1)) nc[i] = 0;
        else if (nc[i]) m = max(m, msz[i]);
    }
    for (int d = kWave / 2; d > 0; d >>= 1) m = max(m, u32(__shfl_xor(int(m), d, kWave)));
    if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicMax(maxsz, m);
}

// hist1[d] += the number of keys whose top b1 bits are d (1 <= b1 <= 12).
__global__ __launch_bounds__(256) void k_key_hist_top(const u64 *H, u64 n, u32 b1, u64 *hist1) {
    __shared__ u32 hist[kHistBins];
    for (u32 i = threadIdx.x; i < (1u << b1); i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
        atomicAdd(&hist[H[i] >> (64 - b1)], 1u);
    __syncthreads();
    for (u32 i = threadIdx.x; i < (1u << b1); i += blockDim.x)
        if (hist[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist1 + i), (unsigned long long)hist[i]);
}

// w[i] = the modelled memory of the unique in dedup slot i (its state entry
// of per_object bytes and its padded child size, from the sizes the rule's
// generate recorded); 0 for an empty slot.
__global__ __launch_bounds__(256) void k_user_slot_weight(const u64 *UR, u64 n, const u32 *csize, u32 align,
                                                          u64 per_object, u64 *w) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        const u64 c = UR[i];
        if (c == kNoRep) {
            w[i] = 0;
            continue;
        }
        const u32 sz = csize[c];
        w[i] = per_object + sz + align_pad(sz, align);
    }
}

// Sizes and padded sizes of the survivors (candidate indices orep).
__global__ __launch_bounds__(256) void k_user_out_size(const u64 *orep, u64 K, const u32 *csize, u32 align, u32 *osize,
                                                       u64 *padded) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x) {
        const u32 s = csize[orep[k]];
        osize[k] = s;
        padded[k] = u64(s) + align_pad(s, align);
    }
}

// vals[i] = obs[i] |m_i|^2 (the observable times the probability, in the
// order of iteration::average_value, src/quids.hpp:224-226); cot[i] = 2 obs
// m_i when asked for.
__global__ __launch_bounds__(256) void k_observable_of(const double *obs, const double2 *mag, u64 N, double *vals,
                                                       double2 *cot) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += u64(gridDim.x) * blockDim.x) {
        const double v = obs[i];
        const double2 m = mag[i];
        vals[i] = v * (m.x * m.x + m.y * m.y);
        if (cot) cot[i] = make_double2(2.0 * v * m.x, 2.0 * v * m.y);
    }
}

}  // namespace qd
