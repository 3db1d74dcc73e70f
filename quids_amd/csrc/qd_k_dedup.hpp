#pragma once
// per-bucket interference (dedup) in LDS
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

// ===========================================================================
// phase 5: per-bucket dedup in LDS
// ===========================================================================
// A persistent grid walks the buckets.  A bucket's entries are staged in LDS
// and bitonic-sorted on (hash, candidate index), so each hash group becomes a
// run in ascending candidate order: its first candidate is the
// representative and a segmented pairwise tree sums its magnitudes (adjacent
// pairs level by level, the oracle's canonical order).  Groups with
// |m|^2 > tolerance are written compactly at the start of the bucket's own
// slot range in the output; the rest of the range gets the sentinel
// representative ~0, so no global atomics order the output.  Buckets larger
// than the LDS capacity are listed for k_dedup_big.
constexpr int kNormBins = 2048;  // |m|^2 exponent histogram (the sign bit is 0)
constexpr u64 kNoRep = ~0ull;

__device__ inline u32 norm_top12(double nrm) { return u32(__double_as_longlong(nrm) >> 52) & 0xFFF; }

__device__ inline bool key_gt(u64 ka, u64 ca, u64 kb, u64 cb) { return ka > kb || (ka == kb && ca > cb); }

// Exclusive max-scan of one value per thread over the block.
__device__ inline u32 block_excl_max(u32 v, u64 *tmp) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
    u32 inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const u32 o = __shfl_up(inc, d, kWave);
        if (lane >= d) inc = max(inc, o);
    }
    const u32 exc_w = __shfl_up(inc, 1, kWave);
    if (lane == kWave - 1) tmp[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 m = 0;
        for (int w = 0; w < nw; ++w) {
            const u64 t = tmp[w];
            tmp[w] = m;
            m = max(m, t);
        }
    }
    __syncthreads();
    const u32 r = max(u32(tmp[wid]), lane ? exc_w : 0u);
    __syncthreads();
    return r;
}

// Per bucket: entries go into an LDS open-addressing table keyed by hash
// (home slot from the 12 hash bits below the bucket bits, linear probing),
// which names each group by its slot.  A counting sort on the slot makes each
// group contiguous; each member then finds its place in ascending candidate
// order by counting the group members with a smaller candidate index.
// perm lists the bucket in (group, candidate) order; magnitudes stay in
// place and are reached through perm.
//
// Bucket b holds woff[b+1] - woff[b] entries, read from E at b * rstride
// (fixed-capacity layout) or at woff[b] (compact layout, rstride == 0); its
// uniques go to the output slots [woff[b], woff[b+1]).
template <int T, int CAP, int TS>
__global__ __launch_bounds__(T) void k_dedup(const Entry *E, u64 rstride, const u64 *woff, u64 nbk, u32 L, double tol,
                                             u64 *UH, u64 *UR, double2 *UM, u64 *counters, u64 *nhist, u64 *biglist) {
    constexpr int PER = CAP / T;
    constexpr int SPER = TS / T;
    __shared__ u64 key[CAP];
    __shared__ u64 cand[CAP];
    __shared__ double2 mag[CAP];
    __shared__ u32 owner[TS];       // entry index + 1 of the slot's first entry, 0 = empty
    __shared__ u32 gstart[TS + 1];  // group sizes, then their exclusive prefix
    __shared__ u16 gslot[CAP];      // group slot of each entry
    __shared__ u16 perm0[CAP];      // entries grouped by slot
    __shared__ u16 perm[CAP];       // entries in (group, candidate) order
    __shared__ u32 nh[kNormBins];
    __shared__ u64 tmp[T / kWave + 1];
    __shared__ u32 maxg;
    const u32 shift = (L + 12 <= 64) ? 64 - L - 12 : 0;
    for (int i = threadIdx.x; i < kNormBins; i += T) nh[i] = 0;
    for (int i = threadIdx.x; i < TS; i += T) owner[i] = 0;
    for (int i = threadIdx.x; i <= TS; i += T) gstart[i] = 0;
    u64 uniques = 0;  // thread 0's running total

    Entry x[PER];
    u64 nbeg = 0, nnum = 0;
    auto fetch = [&](u64 bb) {
        nbeg = woff[bb];
        nnum = woff[bb + 1] - nbeg;
        const u64 rb = rstride ? bb * rstride : nbeg;
        if (nnum <= CAP) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const u32 i = threadIdx.x + T * k;
                if (i < nnum) x[k] = E[rb + i];
            }
        }
    };
    if (blockIdx.x < nbk) fetch(blockIdx.x);
    __syncthreads();
    for (u64 b = blockIdx.x; b < nbk; b += gridDim.x) {
        const u64 beg = nbeg, nn = nnum;  // uniform over the block
        if (nn > CAP) {
            if (threadIdx.x == 0) biglist[atomicAdd(reinterpret_cast<unsigned long long *>(counters + 5), 1ull)] = b;
            __syncthreads();
            if (b + gridDim.x < nbk) fetch(b + gridDim.x);
            continue;
        }
        const u32 n = u32(nn);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 i = threadIdx.x + T * k;
            if (i < n) {
                key[i] = x[k].hash;
                cand[i] = x[k].cand;
                mag[i] = make_double2(x[k].re, x[k].im);
            }
        }
        if (threadIdx.x == 0) maxg = 1;
        __syncthreads();
        if (b + gridDim.x < nbk) fetch(b + gridDim.x);
        if (n == 0) continue;  // uniform
        // insert: find or claim the group slot of every entry
        u32 rank[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 i = threadIdx.x + T * k;
            if (i < n) {
                const u64 h = key[i];
                u32 sl = u32(h >> shift) & (TS - 1);
                for (;;) {
                    const u32 o = atomicCAS(&owner[sl], 0u, i + 1);
                    if (o == 0 || key[o - 1] == h) break;
                    sl = (sl + 1) & (TS - 1);
                }
                gslot[i] = u16(sl);
                rank[k] = atomicAdd(&gstart[sl], 1u);
            }
        }
        __syncthreads();
        // group sizes -> offsets (exclusive scan over the slots)
        {
            u32 v[SPER];
            u64 sum = 0;
#pragma unroll
            for (int k = 0; k < SPER; ++k) {
                v[k] = gstart[threadIdx.x * SPER + k];
                sum += v[k];
                if (v[k] > 1) atomicMax(&maxg, v[k]);
            }
            u64 total;
            u64 run = block_excl_scan(sum, tmp, total);
#pragma unroll
            for (int k = 0; k < SPER; ++k) {
                gstart[threadIdx.x * SPER + k] = u32(run);
                run += v[k];
            }
            if (threadIdx.x == 0) gstart[TS] = u32(total);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 i = threadIdx.x + T * k;
            if (i < n) perm0[gstart[gslot[i]] + rank[k]] = u16(i);
        }
        __syncthreads();
        // place each member by the number of smaller candidates in its group
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 i = threadIdx.x + T * k;
            if (i < n) {
                const u32 sl = gslot[i];
                const u32 g0 = gstart[sl], g1 = gstart[sl + 1];
                const u64 ci = cand[i];
                u32 r = 0;
                for (u32 q = g0; q < g1; ++q) r += cand[perm0[q]] < ci;
                perm[g0 + r] = u16(i);
            }
        }
        __syncthreads();
        // segmented pairwise tree over sorted positions (values live at perm[])
        const u32 mg = maxg;
        for (u32 st = 1; st < mg; st <<= 1) {
            for (u32 p = threadIdx.x; p < n; p += T) {
                const u32 ia = perm[p];
                const u32 g0 = gstart[gslot[ia]];
                const u32 gsz = gstart[gslot[ia] + 1] - g0;
                const u32 r = p - g0;
                if ((r & (2 * st - 1)) == 0 && r + st < gsz) {
                    const u32 ib = perm[p + st];
                    const double2 a = mag[ia], c = mag[ib];
                    mag[ia] = make_double2(a.x + c.x, a.y + c.y);
                }
            }
            __syncthreads();
        }
        // compact the kept groups (their first sorted member) to the front of
        // the bucket's slot range
        u32 cnt = 0;
        u32 keepbits = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 p = threadIdx.x * PER + k;
            if (p < n) {
                const u32 ia = perm[p];
                if (gstart[gslot[ia]] == p) {
                    const double2 a = mag[ia];
                    const double nrm = a.x * a.x + a.y * a.y;
                    if (nrm > tol) {
                        keepbits |= 1u << k;
                        ++cnt;
                        atomicAdd(&nh[min(norm_top12(nrm), u32(kNormBins - 1))], 1u);
                    }
                }
            }
        }
        u64 total;
        u64 pos = beg + block_excl_scan(cnt, tmp, total);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (keepbits & (1u << k)) {
                const u32 pi = perm[threadIdx.x * PER + k];
                UH[pos] = key[pi];
                UR[pos] = cand[pi];
                UM[pos] = mag[pi];
                ++pos;
            }
        }
        for (u64 jx = beg + total + threadIdx.x; jx < beg + n; jx += T) UR[jx] = kNoRep;
        if (threadIdx.x == 0) uniques += total;
        __syncthreads();
        // clear the table: owners of the slots this bucket used, and every
        // group offset (the scan wrote all of them)
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 i = threadIdx.x + T * k;
            if (i < n) owner[gslot[i]] = 0;
        }
        for (int i = threadIdx.x; i <= TS; i += T) gstart[i] = 0;
        __syncthreads();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kNormBins; i += T)
        if (nh[i]) atomicAdd(reinterpret_cast<unsigned long long *>(nhist + i), (unsigned long long)nh[i]);
    if (threadIdx.x == 0 && uniques) atomicAdd(reinterpret_cast<unsigned long long *>(counters), uniques);
}

// The same computation for buckets beyond the LDS capacity, one workgroup
// per bucket, on a global scratch region (rare: a bucket this large needs a
// group of thousands of candidates with one hash).  Scratch of listed bucket
// q starts at soff[q] elements: keys, cands, mags (npad each) and u32 segs.
__global__ __launch_bounds__(1024) void k_dedup_big(const Entry *E, u64 rstride, const u64 *boff, const u64 *biglist,
                                                    const u64 *soff, u64 *skeyg, u64 *scandg, double2 *smagg,
                                                    u32 *segg, double tol, u64 *UH, u64 *UR, double2 *UM,
                                                    u64 *counters, u64 *nhist) {
    __shared__ u32 maxg;
    __shared__ u64 tmp[1024 / kWave + 1];
    const u64 b = biglist[blockIdx.x];
    const u64 beg = boff[b], n = boff[b + 1] - beg;
    const Entry *src = E + (rstride ? b * rstride : beg);
    u64 *skey = skeyg + soff[blockIdx.x];
    u64 *scand = scandg + soff[blockIdx.x];
    double2 *smag = smagg + soff[blockIdx.x];
    u32 *seg = segg + soff[blockIdx.x];
    u64 npad = 2;
    while (npad < n) npad <<= 1;
    for (u64 i = threadIdx.x; i < npad; i += blockDim.x) {
        if (i < n) {
            skey[i] = src[i].hash;
            scand[i] = src[i].cand;
        } else {
            skey[i] = ~0ull;
            scand[i] = ~0ull;
        }
    }
    if (threadIdx.x == 0) maxg = 1;
    __syncthreads();
    for (u64 k = 2; k <= npad; k <<= 1) {
        for (u64 j = k >> 1; j > 0; j >>= 1) {
            for (u64 t = threadIdx.x; t < npad / 2; t += blockDim.x) {
                const u64 i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                const u64 p = i | j;
                const u64 ka = skey[i], ca = scand[i], kb = skey[p], cb = scand[p];
                if (key_gt(ka, ca, kb, cb) == ((i & k) == 0)) {
                    skey[i] = kb;
                    scand[i] = cb;
                    skey[p] = ka;
                    scand[p] = ca;
                }
            }
            __syncthreads();
        }
    }
    // magnitudes: candidate c of this bucket is found by its position in the
    // input (entries carry their candidate index, not their position), so
    // re-read them by a search over the sorted (hash, cand) pairs
    for (u64 i = threadIdx.x; i < n; i += blockDim.x) {
        const u64 eh = src[i].hash, ec = src[i].cand;
        u64 lo = 0, hi = n - 1;
        while (lo < hi) {
            const u64 mid = (lo + hi) >> 1;
            if (key_gt(eh, ec, skey[mid], scand[mid])) lo = mid + 1;
            else hi = mid;
        }
        smag[lo] = make_double2(src[i].re, src[i].im);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 run = 0;
        for (u64 i = 0; i < n; ++i) {
            if (i == 0 || skey[i] != skey[i - 1]) run = u32(i);
            seg[i] = run;
            if (i + 1 - run > maxg) maxg = u32(i + 1 - run);
        }
    }
    __syncthreads();
    for (u64 st = 1; st < maxg; st <<= 1) {
        for (u64 i = threadIdx.x; i < n; i += blockDim.x) {
            const u64 r = i - seg[i];
            if ((r & (2 * st - 1)) == 0 && i + st < n && seg[i + st] == seg[i]) {
                const double2 a = smag[i], c = smag[i + st];
                smag[i] = make_double2(a.x + c.x, a.y + c.y);
            }
        }
        __syncthreads();
    }
    // serial compaction by thread 0 (this path is rare)
    if (threadIdx.x == 0) {
        u64 pos = beg;
        for (u64 i = 0; i < n; ++i) {
            if (seg[i] != i) continue;
            const double2 a = smag[i];
            const double nrm = a.x * a.x + a.y * a.y;
            if (nrm > tol) {
                UH[pos] = skey[i];
                UR[pos] = scand[i];
                UM[pos] = a;
                ++pos;
                atomicAdd(reinterpret_cast<unsigned long long *>(nhist + min(norm_top12(nrm), u32(kNormBins - 1))),
                          1ull);
            }
        }
        atomicAdd(reinterpret_cast<unsigned long long *>(counters), (unsigned long long)(pos - beg));
        for (u64 j = pos; j < beg + n; ++j) UR[j] = kNoRep;
    }
}


}  // namespace qd
