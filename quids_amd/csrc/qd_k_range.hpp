#pragma once
// Hash-range passes of a step whose candidates exceed the scratch memory:
// the pool of surviving uniques carried from pass to pass.  A pass dedups
// the candidates of one range of top hash bits completely (equal hashes fall
// in the same range), so the step's uniques are the union of the passes'
// uniques, and the pool keeps the best max_num_object of them seen so far.
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_k_dedup.hpp"
#include "qd_k_fastgen.hpp"
#include "qd_k_generate.hpp"
#include "qd_k_select.hpp"

namespace qd {

constexpr int kPoolPer = 16;  // slots per thread of k_pool_append

// Appends the slots holding a unique (UR != kNoRep) to the pool: hash (the
// slot key rotated back), representative, magnitude and, with UK, the top
// 32 |m|^2 bits, in slot order within a block tile of 256 * kPoolPer slots,
// so that the uniques of one wavefront's 64 slots land side by side
// (coalesced stores).  One reservation per tile.  With UK, a unique whose
// top 24 |m|^2 bits are below t24 is left out: the pool already holds
// max_num_object keys at or above that prefix, so it can never be kept.
__global__ __launch_bounds__(256) void k_pool_append(const u64 *SK, const u64 *UR, const double2 *UM, const u32 *UK,
                                                     u32 t24, u64 nslots, u64 *Ph, u64 *Pr, double2 *Pm, u32 *Pk,
                                                     u64 *pcount) {
    constexpr int kWaves = 256 / kWave;
    constexpr int kRuns = kPoolPer * kWaves;  // runs of 64 slots in a tile, in slot order
    static_assert(kRuns == kWave, "one lane of the first wavefront per run");
    __shared__ u64 roff[kRuns];  // count, then exclusive prefix, of each run's uniques
    __shared__ u64 base;
    constexpr u64 kTile = 256 * kPoolPer;
    const u32 lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const u64 below = (1ull << lane) - 1;
    for (u64 t0 = u64(blockIdx.x) * kTile; t0 < nslots; t0 += u64(gridDim.x) * kTile) {
        u64 r[kPoolPer], m[kPoolPer];
#pragma unroll
        for (int k = 0; k < kPoolPer; ++k) {
            const u64 i = t0 + threadIdx.x + 256 * k;
            // (with UK, the 4-byte key first: below t24, the slot, empty or
            // not, is left out without reading its representative)
            r[k] = kNoRep;
            if (i < nslots && (!UK || (UK[i] >> 8) >= t24)) r[k] = UR[i];
            m[k] = __ballot(r[k] != kNoRep);
            if (lane == 0) roff[k * kWaves + wid] = __popcll(m[k]);
        }
        __syncthreads();
        if (wid == 0) {
            const u64 c = roff[lane];
            const u64 inc = wave_incl_scan(c);
            roff[lane] = inc - c;
            if (lane == kWave - 1)
                base = inc ? atomicAdd(reinterpret_cast<unsigned long long *>(pcount), (unsigned long long)inc) : 0;
        }
        __syncthreads();
        const u64 b = base;
#pragma unroll
        for (int k = 0; k < kPoolPer; ++k) {
            if (r[k] != kNoRep) {
                const u64 i = t0 + threadIdx.x + 256 * k;
                const u64 at = b + roff[k * kWaves + wid] + __popcll(m[k] & below);
                Ph[at] = rotr12(SK[i]);
                Pr[at] = r[k];
                Pm[at] = UM[i];
                if (UK) Pk[at] = UK[i];
            }
        }
        __syncthreads();  // roff and base are rewritten by the next tile
    }
}

// Keeps the pool items whose bit is set in keep, in pool order (wprefix:
// exclusive popcount prefix of keep); Pk / Qk (the 32-bit keys) may be null.
__global__ __launch_bounds__(256) void k_pool_compact(const u64 *Ph, const u64 *Pr, const double2 *Pm, const u32 *Pk,
                                                      u64 n, const u64 *keep, const u64 *wprefix, u64 *Qh, u64 *Qr,
                                                      double2 *Qm, u32 *Qk) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        const u64 w = keep[i >> 6];
        if (!((w >> (i & 63)) & 1)) continue;
        const u64 at = wprefix[i >> 6] + __popcll(w & ((1ull << (i & 63)) - 1));
        Qh[at] = Ph[i];
        Qr[at] = Pr[i];
        Qm[at] = Pm[i];
        if (Pk) Qk[at] = Pk[i];
    }
}

// Histogram of the top 12 bits of n 32-bit |m|^2 keys (the pool's).
__global__ __launch_bounds__(256) void k_ukey_hist0(const u32 *UK, u64 n, u64 *hist) {
    __shared__ u32 h[kHistBins];
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
        atomicAdd(&h[UK[i] >> 20], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
        if (h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist + i), (unsigned long long)h[i]);
}

// Marks the representatives of the pool items in the candidate bitmap.
__global__ __launch_bounds__(256) void k_mark_reps(const u64 *Pr, u64 n, u64 *bitmap) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) mark(bitmap, Pr[i]);
}

// Places every pool item at the rank of its representative among the
// marked ones (ascending candidate order).
__global__ __launch_bounds__(256) void k_place_list(const u64 *Pr, const double2 *Pm, u64 n, const u64 *bitmap,
                                                    const u64 *wprefix, u64 *orep, double2 *omag) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        const u64 rep = Pr[i];
        const u64 r = wprefix[rep >> 6] + __popcll(bitmap[rep >> 6] & ((1ull << (rep & 63)) - 1));
        orep[r] = rep;
        omag[r] = Pm[i];
    }
}

}  // namespace qd
