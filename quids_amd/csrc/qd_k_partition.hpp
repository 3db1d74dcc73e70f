#pragma once
// MSD radix partition of candidates on their hash
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

// ===========================================================================
// phase 4: radix partition
// ===========================================================================
constexpr int kTile1 = 4096;  // entries per workgroup in the first pass
constexpr int kLevel2Bins = 8192;  // up to 13 bits in the second pass

// First pass: from candidate order (index = candidate, hash and magnitude
// as written by generate) to 2^b1 buckets at exact offsets.
// CI, when not null, holds the candidate index of every input position
// (compacted inputs of the hash-range passes).
__global__ __launch_bounds__(256) void k_scatter1(const u64 *H0, const double2 *M0, u64 n, u32 b1, u64 *cursor,
                                                  Entry *out, const u64 *CI) {
    __shared__ u32 cnt[kHistBins];
    __shared__ u64 gbase[kHistBins];
    const int nbin = 1 << b1;
    for (int i = threadIdx.x; i < nbin; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    const u64 base = u64(blockIdx.x) * kTile1;
    constexpr int kPer = kTile1 / 256;
    u32 dr[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const u64 i = base + threadIdx.x + 256 * k;
        if (i < n) {
            const u32 d = b1 ? u32(H0[i] >> (64 - b1)) : 0u;
            const u32 r = atomicAdd(&cnt[d], 1u);
            dr[k] = (d << 16) | r;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbin; i += blockDim.x)
        if (cnt[i]) gbase[i] = atomicAdd(reinterpret_cast<unsigned long long *>(cursor + i), (unsigned long long)cnt[i]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const u64 i = base + threadIdx.x + 256 * k;
        if (i < n) {
            const double2 m = M0[i];
            Entry e;
            e.hash = H0[i];
            e.cand = CI ? CI[i] : i;
            e.re = m.x;
            e.im = m.y;
            out[gbase[dr[k] >> 16] + (dr[k] & 0xFFFF)] = e;
        }
    }
}

// Second pass, one workgroup per first-level bucket, splitting it by the
// next b2 hash bits in one read: sub-bucket d of bucket b owns cap2 slots at
// ((b << b2) + d) * cap2; workgroup-local cursors place the entries, and the
// final counts go to cnt2.  A sub-bucket that would exceed cap2 raises
// ERR_SPILL (the step then repeats this pass with exact counts).
constexpr u32 ERR_SPILL = 16u;

// Input bucket b holds off1[b+1] - off1[b] entries, read from in at
// b * in_stride (fixed-capacity first pass) or at off1[b] (in_stride 0).
// Workgroup g takes buckets g, g + gridDim.x, ... of the nbk1 (2^b1, or a
// group of them: the caller offsets in and off1 to the group's first).
__global__ __launch_bounds__(1024) void k_level2_fixed(const Entry *in, u64 in_stride, const u64 *off1, u64 nbk1,
                                                       u32 b1, u32 b2, u64 cap2, Entry *out, u64 *cnt2, u32 *err) {
    __shared__ u32 ccnt[kLevel2Bins];
    __shared__ u32 cur[kLevel2Bins];
    bool spill = false;
    for (u64 b = blockIdx.x; b < nbk1; b += gridDim.x) {
        const u64 beg = off1[b], end = off1[b + 1];
        const u64 ibase = in_stride ? b * in_stride : beg;  // entry i of [beg, end) is in[ibase + i - beg]
        const u32 nb2 = 1u << b2;
        const u32 shift = 64 - b1 - b2;
        Entry *region = out + (b << b2) * cap2;
        for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) cur[i] = 0;
        // entries of the chunk being split, and the next chunk's, loaded while
        // this one is ranked and written: four 64-bit words per entry (key,
        // candidate, magnitude bits), in plain registers
        // (lanes past the bucket's end load its first entry and drop it)
        const u64 nent = end - beg;
        if (nent == 0) {  // uniform
            for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) cnt2[(b << b2) + i] = 0;
            __syncthreads();
            continue;
        }
        const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(in + ibase);
        ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(region);
        u64 w0[4], w1[4], w2[4], w3[4], n0[4], n1[4], n2[4], n3[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const u64 i = threadIdx.x + 1024 * k;
            const u64 ii = i < nent ? i : 0;
            const ulonglong2 lo = src[2 * ii], hi = src[2 * ii + 1];
            n0[k] = lo.x;
            n1[k] = lo.y;
            n2[k] = hi.x;
            n3[k] = hi.y;
        }
        for (u64 cbase = beg; cbase < end; cbase += 4096) {
            for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) ccnt[i] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w0[k] = n0[k];
                w1[k] = n1[k];
                w2[k] = n2[k];
                w3[k] = n3[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const u64 i = cbase + 4096 - beg + threadIdx.x + 1024 * k;
                const u64 ii = i < nent ? i : 0;
                const ulonglong2 lo = src[2 * ii], hi = src[2 * ii + 1];
                n0[k] = lo.x;
                n1[k] = lo.y;
                n2[k] = hi.x;
                n3[k] = hi.y;
            }
            u32 dr[4] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const u64 i = cbase + threadIdx.x + 1024 * k;
                if (i < end) {
                    const u32 d = u32(w0[k] >> shift) & (nb2 - 1);
                    dr[k] = (d << 16) | atomicAdd(&ccnt[d], 1u);
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const u64 i = cbase + threadIdx.x + 1024 * k;
                if (i < end) {
                    const u32 d = dr[k] >> 16;
                    const u64 slot = u64(cur[d]) + (dr[k] & 0xFFFF);
                    if (slot < cap2) {
                        const u64 at = u64(d) * cap2 + slot;
                        ulonglong2 lo, hi;
                        lo.x = w0[k];
                        lo.y = w1[k];
                        hi.x = w2[k];
                        hi.y = w3[k];
                        dst[2 * at] = lo;
                        dst[2 * at + 1] = hi;
                    } else {
                        spill = true;
                    }
                }
            }
            __syncthreads();
            for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) cur[i] += ccnt[i];
        }
        __syncthreads();
        for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) cnt2[(b << b2) + i] = min<u64>(cur[i], cap2);
        __syncthreads();  // cur is cleared for the next bucket
    }
    if (spill) atomicOr(err, ERR_SPILL);
}

// The same split with exact offsets (a counting pass first): used when a
// sub-bucket spills its fixed capacity.  One workgroup per input bucket;
// positions in out and off2 are off1's minus obase (the offset of the
// group's first entry).
__global__ __launch_bounds__(1024) void k_level2_exact(const Entry *in, u64 in_stride, const u64 *off1, u64 obase,
                                                       u32 b1, u32 b2, Entry *out, u64 *off2) {
    __shared__ u32 cnt[kLevel2Bins];
    __shared__ u32 ccnt[kLevel2Bins];
    __shared__ u64 cur[kLevel2Bins];
    __shared__ u64 tmp[1024 / kWave + 1];
    const u64 b = blockIdx.x;
    const u64 beg = off1[b], end = off1[b + 1];
    const u64 ibase = in_stride ? b * in_stride : beg;
    const u32 nb2 = 1u << b2;
    const u32 shift = 64 - b1 - b2;
    for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    for (u64 i = beg + threadIdx.x; i < end; i += blockDim.x) atomicAdd(&cnt[u32(in[ibase + (i - beg)].hash >> shift) & (nb2 - 1)], 1u);
    __syncthreads();
    {
        constexpr int kPer = kLevel2Bins / 1024;
        u32 v[kPer];
        u64 s = 0;
        for (int k = 0; k < kPer; ++k) {
            const u32 d = threadIdx.x * kPer + k;
            v[k] = d < nb2 ? cnt[d] : 0;
            s += v[k];
        }
        u64 total;
        u64 run = block_excl_scan(s, tmp, total);
        for (int k = 0; k < kPer; ++k) {
            const u32 d = threadIdx.x * kPer + k;
            if (d < nb2) {
                cur[d] = beg - obase + run;
                off2[(b << b2) + d] = beg - obase + run;
            }
            run += v[k];
        }
    }
    __syncthreads();
    for (u64 cbase = beg; cbase < end; cbase += 4096) {
        for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) ccnt[i] = 0;
        __syncthreads();
        u32 dr[4];
        for (int k = 0; k < 4; ++k) {
            const u64 i = cbase + threadIdx.x + 1024 * k;
            if (i < end) {
                const u32 d = u32(in[ibase + (i - beg)].hash >> shift) & (nb2 - 1);
                dr[k] = (d << 16) | atomicAdd(&ccnt[d], 1u);
            }
        }
        __syncthreads();
        for (int k = 0; k < 4; ++k) {
            const u64 i = cbase + threadIdx.x + 1024 * k;
            if (i < end) out[cur[dr[k] >> 16] + (dr[k] & 0xFFFF)] = in[ibase + (i - beg)];
        }
        __syncthreads();
        for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) cur[i] += ccnt[i];
        __syncthreads();
    }
}


}  // namespace qd
