#pragma once
// block scans: exclusive prefix of n loaded values
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

// ===========================================================================
// scans: exclusive prefix of n loaded values into out[0..n] (out[n] = total)
// ===========================================================================
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

struct LoadU32 {
    const u32 *p;
    __device__ u64 operator()(u64 i) const { return p[i]; }
};
struct LoadU64 {
    const u64 *p;
    __device__ u64 operator()(u64 i) const { return p[i]; }
};
struct LoadPopc {
    const u64 *p;
    __device__ u64 operator()(u64 i) const { return __popcll(p[i]); }
};

template <class Load>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(Load ld, u64 n, u64 *bsum) {
    __shared__ u64 tmp[kScanThreads / kWave + 1];
    const u64 base = u64(blockIdx.x) * kScanTile + u64(threadIdx.x) * kScanItems;
    u64 s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += ld(base + k);
    u64 total;
    block_excl_scan(s, tmp, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_bsums(u64 *bsum, u64 nb) {
    __shared__ u64 tmp[1024 / kWave + 1];
    u64 carry = 0;
    for (u64 base = 0; base < nb; base += 1024) {
        const u64 i = base + threadIdx.x;
        const u64 v = i < nb ? bsum[i] : 0;
        u64 total;
        const u64 ex = block_excl_scan(v, tmp, total);
        if (i < nb) bsum[i] = carry + ex;
        carry += total;
    }
}

template <class Load>
__global__ __launch_bounds__(kScanThreads) void k_scan_final(Load ld, u64 n, const u64 *boff, u64 *out) {
    __shared__ u64 tmp[kScanThreads / kWave + 1];
    const u64 base = u64(blockIdx.x) * kScanTile + u64(threadIdx.x) * kScanItems;
    u64 v[kScanItems];
    u64 s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? ld(base + k) : 0;
        s += v[k];
    }
    u64 total;
    u64 run = boff[blockIdx.x] + block_excl_scan(s, tmp, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
        if (base + k + 1 == n) out[n] = run;
    }
}


}  // namespace qd
