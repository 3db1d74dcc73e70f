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
// slot key rotated back), representative, magnitude.  One reservation per
// block tile of 256 * kPoolPer slots.
__global__ __launch_bounds__(256) void k_pool_append(const u64 *UK, const u64 *UR, const double2 *UM, u64 nslots,
                                                     u64 *Ph, u64 *Pr, double2 *Pm, u64 *pcount) {
    __shared__ u64 tmp[256 / kWave + 1];
    __shared__ u64 base;
    constexpr u64 kTile = 256 * kPoolPer;
    for (u64 t0 = u64(blockIdx.x) * kTile; t0 < nslots; t0 += u64(gridDim.x) * kTile) {
        u32 cnt = 0;
#pragma unroll
        for (int k = 0; k < kPoolPer; ++k) {
            const u64 i = t0 + threadIdx.x + 256 * k;
            cnt += (i < nslots && UR[i] != kNoRep) ? 1u : 0u;
        }
        u64 total;
        u64 at = block_excl_scan(u64(cnt), tmp, total);
        if (threadIdx.x == 0 && total)
            base = atomicAdd(reinterpret_cast<unsigned long long *>(pcount), (unsigned long long)total);
        __syncthreads();
        at += base;
#pragma unroll
        for (int k = 0; k < kPoolPer; ++k) {
            const u64 i = t0 + threadIdx.x + 256 * k;
            if (i < nslots) {
                const u64 r = UR[i];
                if (r != kNoRep) {
                    Ph[at] = rotr12(UK[i]);
                    Pr[at] = r;
                    Pm[at] = UM[i];
                    ++at;
                }
            }
        }
        __syncthreads();
    }
}

// Keeps the pool items whose bit is set in keep, in pool order (wprefix:
// exclusive popcount prefix of keep).
__global__ __launch_bounds__(256) void k_pool_compact(const u64 *Ph, const u64 *Pr, const double2 *Pm, u64 n,
                                                      const u64 *keep, const u64 *wprefix, u64 *Qh, u64 *Qr,
                                                      double2 *Qm) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        const u64 w = keep[i >> 6];
        if (!((w >> (i & 63)) & 1)) continue;
        const u64 at = wprefix[i >> 6] + __popcll(w & ((1ull << (i & 63)) - 1));
        Qh[at] = Ph[i];
        Qr[at] = Pr[i];
        Qm[at] = Pm[i];
    }
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
