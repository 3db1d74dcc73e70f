#pragma once
// per-bucket interference (dedup) in LDS
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

// ===========================================================================
// phase 5: per-bucket dedup in LDS
// ===========================================================================
// A persistent grid walks the buckets; each workgroup stages a bucket's
// entries in LDS, groups them by hash, and writes per group the hash, the
// smallest candidate index (the representative) and the magnitudes summed
// in ascending candidate order by the pairwise tree (adjacent pairs level
// by level, the oracle's canonical order).  Groups with |m|^2 > tolerance
// are written compactly at the start of the bucket's own slot range in the
// output; the rest of the range gets the sentinel representative ~0, so no
// global atomics order the output.  k_dedup is the main kernel; it hands
// the rare buckets it cannot pack to k_dedup_sort, and buckets larger than
// the LDS capacity go to k_dedup_big.
// With UK, every slot also gets the top 32 bits of its |m|^2 bit pattern
// (0 for an empty slot): the select then reads 4 bytes per slot.
constexpr int kNormBins = 2048;  // |m|^2 exponent histogram (the sign bit is 0)
constexpr u64 kNoRep = ~0ull;

__device__ inline u32 norm_top12(double nrm) { return u32(__double_as_longlong(nrm) >> 52) & 0xFFF; }
__device__ inline u32 norm_top32(double nrm) { return u32(u64(__double_as_longlong(nrm)) >> 32); }

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

// k_dedup_sort, per bucket: entries go into an LDS open-addressing table
// keyed by hash (home slot from the 12 hash bits below the bucket bits,
// linear probing),
// which names each group by its slot.  A counting sort on the slot makes each
// group contiguous; each member then finds its place in ascending candidate
// order by counting the group members with a smaller candidate index.
// perm lists the bucket in (group, candidate) order; magnitudes stay in
// place and are reached through perm.
//
// Bucket b holds woff[b+1] - woff[b] entries, read from E at b * rstride
// (fixed-capacity layout) or at woff[b] (compact layout, rstride == 0); its
// uniques go to the output slots [woff[b], woff[b+1]).  With list, the
// buckets are list[0 .. nbk) (those k_dedup hands back); else 0 .. nbk.
template <int T, int CAP, int TS>
__global__ __launch_bounds__(T) void k_dedup_sort(const Entry *E, u64 rstride, const u64 *woff, u64 nbk, u32 L,
                                                  double tol, u64 *UH, u64 *UR, double2 *UM, u32 *UK,
                                                  u64 *counters, u64 *nhist, u64 *biglist, const u64 *list) {
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
    auto fetch = [&](u64 bq) {
        const u64 bb = list ? list[bq] : bq;
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
    for (u64 bq = blockIdx.x; bq < nbk; bq += gridDim.x) {
        const u64 b = list ? list[bq] : bq;
        const u64 beg = nbeg, nn = nnum;  // uniform over the block
        if (nn > CAP) {
            if (threadIdx.x == 0) biglist[atomicAdd(reinterpret_cast<unsigned long long *>(counters + 5), 1ull)] = b;
            __syncthreads();
            if (bq + gridDim.x < nbk) fetch(bq + gridDim.x);
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
        if (bq + gridDim.x < nbk) fetch(bq + gridDim.x);
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
        u32 nk[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 p = threadIdx.x * PER + k;
            nk[k] = 0;
            if (p < n) {
                const u32 ia = perm[p];
                if (gstart[gslot[ia]] == p) {
                    const double2 a = mag[ia];
                    const double nrm = a.x * a.x + a.y * a.y;
                    if (nrm > tol) {
                        keepbits |= 1u << k;
                        ++cnt;
                        nk[k] = norm_top32(nrm);
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
                if (UK) UK[pos] = nk[k];
                ++pos;
            }
        }
        for (u64 jx = beg + total + threadIdx.x; jx < beg + n; jx += T) {
            UR[jx] = kNoRep;
            if (UK) UK[jx] = 0;
        }
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

// A hash group of at most kGroupRegs members, in registers: sort keys
// (candidate index << 16 | entry) ordered by a sorting network (unused
// places hold ~0), then the magnitudes summed in the canonical pairwise
// order (at stride st, place r, a multiple of 2 st, takes place r + st when
// the group reaches it).
constexpr int kGroupRegs = 8;

template <int A, int B>
__device__ __attribute__((always_inline)) inline void group_ce(u64 (&c)[kGroupRegs]) {
    const u64 lo = min(c[A], c[B]), hi = max(c[A], c[B]);
    c[A] = lo;
    c[B] = hi;
}

__device__ __attribute__((always_inline)) inline void group_sort(u64 (&c)[kGroupRegs], u32 g) {
    if (g <= 4) {  // 5 comparators
        group_ce<0, 1>(c), group_ce<2, 3>(c), group_ce<0, 2>(c), group_ce<1, 3>(c), group_ce<1, 2>(c);
    } else {  // Batcher's odd-even merge network for 8 (19 comparators)
        group_ce<0, 1>(c), group_ce<2, 3>(c), group_ce<4, 5>(c), group_ce<6, 7>(c);
        group_ce<0, 2>(c), group_ce<1, 3>(c), group_ce<4, 6>(c), group_ce<5, 7>(c);
        group_ce<1, 2>(c), group_ce<5, 6>(c);
        group_ce<0, 4>(c), group_ce<1, 5>(c), group_ce<2, 6>(c), group_ce<3, 7>(c);
        group_ce<2, 4>(c), group_ce<3, 5>(c);
        group_ce<1, 2>(c), group_ce<3, 4>(c), group_ce<5, 6>(c);
    }
}

// Representative (smallest candidate index) and canonical sum of the group
// of 3 .. kGroupRegs members led by entry i, whose member list starts at j
// (candidate indices below 2^48).  The pairwise tree is evaluated as a
// binary counter over the sorted members: member q merges the partial sums
// of the levels given by the trailing one bits of q (earlier partial on the
// left), and at the end the partials of the set bits of g fold from the
// lowest level up (the higher, earlier partial on the left); this is the
// same expression tree as the level-by-level pairwise sum.
template <int CAP>
__device__ __attribute__((always_inline)) inline void group_reduce(u32 i, u32 j, const u64 *cand, const double2 *mag,
                                                                   const u16 *nxt, u64 &rc, double2 &rs) {
    u64 c[kGroupRegs];
    c[0] = (cand[i] << 16) | i;
    u32 g = 1;
#pragma unroll
    for (int q = 1; q < kGroupRegs; ++q) {
        if (j < CAP) {
            c[q] = (cand[j] << 16) | j;
            j = nxt[j];
            ++g;
        } else {
            c[q] = ~0ull;
        }
    }
    group_sort(c, g);
    double2 part[4];  // partial sum of level lv (2^lv members)
#pragma unroll
    for (int q = 0; q < kGroupRegs; ++q) {
        if (u32(q) < g) {
            double2 carry = mag[c[q] & 0xFFFF];
            int lv = 0;
#pragma unroll
            for (; lv < 3 && ((q >> lv) & 1); ++lv)
                carry = make_double2(part[lv].x + carry.x, part[lv].y + carry.y);
            part[lv] = carry;
        }
    }
    double2 acc = make_double2(0.0, 0.0);
    bool any = false;
#pragma unroll
    for (int lv = 0; lv < 4; ++lv) {
        if ((g >> lv) & 1) {
            acc = any ? make_double2(part[lv].x + acc.x, part[lv].y + acc.y) : part[lv];
            any = true;
        }
    }
    rc = c[0] >> 16;
    rs = acc;
}

// Per bucket, one LDS table slot per distinct hash (home slot from the 12
// hash bits below the bucket bits, linear probing): the entry that claims a
// slot leads the group, every entry counts itself in the slot, and every
// other entry links itself into the leader's member list.  After a barrier:
//  * a singleton is its own result;
//  * a pair is summed by its leader (a + b: addition commutes, so the
//    canonical order does not matter) with the smaller candidate index;
//  * a group of 3 .. kGroupRegs members is gathered by its leader into
//    registers, sorted by candidate index and summed pairwise in that order
//    (the reference's canonical order);
//  * members of larger groups append (slot << 52 | candidate << 12 | entry)
//    to a wide area, sorted by (group, candidate) with a bitonic sort (in
//    registers of one wavefront when it holds at most 64 entries, else
//    block-wide in LDS) and summed by a segmented pairwise tree in that
//    order.
// Groups with |m|^2 > tol (hash, smallest candidate index, sum) go to the
// front of the bucket's output range (one LDS counter), the rest of the
// range gets the sentinel representative ~0.  A bucket whose candidate
// indices do not fit the packed keys (2^40 or more) is listed in redolist
// (counted in counters[6]) for k_dedup_sort and writes nothing; one beyond
// the LDS capacity goes to biglist (counters[5]).  Layout of E, woff and
// the output as in k_dedup_sort.
constexpr int kWideCandBits = 40;

template <int T, int CAP, int TS>
__global__ __launch_bounds__(T) void k_dedup(const Entry *E, u64 rstride, const u64 *woff, u64 nbk, u32 L, double tol,
                                             u64 *UH, u64 *UR, double2 *UM, u32 *UK, u64 *counters, u64 *nhist,
                                             u64 *biglist, u64 *redolist) {
    constexpr int PER = CAP / T;
    constexpr u32 kNil = 0xFFFFFFFFu;  // empty member list (any value >= CAP ends a list)
    static_assert(CAP <= 4096 && TS <= 4096, "packed wide keys hold 12-bit slots and entries");
    __shared__ u64 key[CAP];
    __shared__ u64 cand[CAP];
    __shared__ double2 mag[CAP];
    __shared__ u64 wkey[CAP];  // wide area
    __shared__ u32 owner[TS];  // leader's entry index + 1 (0 = empty); then the group's wide-area start
    __shared__ u32 gcnt[TS];   // group size
    __shared__ u32 lnk[CAP];   // a leader's last linked member
    __shared__ u16 nxt[CAP];   // a member's next member
    __shared__ u32 nh[kNormBins];
    __shared__ u32 wcount, wmax, ocount;
    const u32 shift = (L + 12 <= 64) ? 64 - L - 12 : 0;
    for (int i = threadIdx.x; i < kNormBins; i += T) nh[i] = 0;
    for (int i = threadIdx.x; i < TS; i += T) owner[i] = gcnt[i] = 0;
    if (threadIdx.x == 0) ocount = 0;
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
                lnk[i] = kNil;
            }
        }
        if (threadIdx.x == 0) wcount = wmax = 0;
        __syncthreads();
        if (threadIdx.x == 0) ocount = 0;  // every thread has read the previous bucket's count
        if (b + gridDim.x < nbk) fetch(b + gridDim.x);
        if (n == 0) continue;  // uniform
        // insert: claim a slot (leader) or join the group that holds it
        u32 slot[PER];
        bool lead[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 i = threadIdx.x + T * k;
            lead[k] = false;
            slot[k] = 0;
            if (i < n) {
                const u64 h = key[i];
                u32 sl = u32(h >> shift) & (TS - 1);
                for (;;) {
                    const u32 o = atomicCAS(&owner[sl], 0u, i + 1);
                    if (o == 0) {
                        lead[k] = true;
                        break;
                    }
                    if (key[o - 1] == h) {
                        nxt[i] = u16(atomicExch(&lnk[o - 1], i));
                        break;
                    }
                    sl = (sl + 1) & (TS - 1);
                }
                slot[k] = sl;
                atomicAdd(&gcnt[sl], 1u);
            }
        }
        __syncthreads();
        // groups of up to kGroupRegs members in registers; members of larger
        // groups to the wide area (one LDS reservation per wavefront)
        u64 rc[PER];
        double2 rs[PER];
        bool redo = false;
        const u32 lane = threadIdx.x & (kWave - 1);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const u32 i = threadIdx.x + T * k;
            rc[k] = 0;
            rs[k] = make_double2(0.0, 0.0);
            const u32 g = i < n ? gcnt[slot[k]] : 0u;
            const bool wm = g > u32(kGroupRegs);
            const u64 wmask = __ballot(wm);
            if (wmask) {
                const int first = __ffsll((unsigned long long)wmask) - 1;
                u32 wb = 0;
                if (int(lane) == first) wb = atomicAdd(&wcount, u32(__popcll(wmask)));
                wb = __shfl(wb, first, kWave);
                if (wm) {
                    const u64 ci = cand[i];
                    if (ci >> kWideCandBits) redo = true;
                    wkey[wb + u32(__popcll(wmask & ((1ull << lane) - 1)))] = (u64(slot[k]) << 52) | (ci << 12) | i;
                    if (lead[k]) atomicMax(&wmax, g);
                }
            }
            if (lead[k] && !wm) {
                rc[k] = cand[i];
                rs[k] = mag[i];
                if (g == 2) {
                    const u32 j = lnk[i];
                    const double2 mj = mag[j];
                    rc[k] = min(rc[k], cand[j]);
                    rs[k] = make_double2(rs[k].x + mj.x, rs[k].y + mj.y);
                } else if (g > 2) {
                    if (rc[k] >> 48) redo = true;
                    else group_reduce<CAP>(i, lnk[i], cand, mag, nxt, rc[k], rs[k]);
                }
            }
        }
        if (__syncthreads_or(redo)) {
            if (threadIdx.x == 0) redolist[atomicAdd(reinterpret_cast<unsigned long long *>(counters + 6), 1ull)] = b;
#pragma unroll
            for (int k = 0; k < PER; ++k)
                if (lead[k]) owner[slot[k]] = gcnt[slot[k]] = 0;
            __syncthreads();
            continue;
        }
        const u32 W = wcount;  // uniform
        if (W && W <= u32(kWave)) {
            // one wavefront: bitonic sort in registers, then the segmented
            // pairwise tree by shuffles; no block barrier until the results
            if (threadIdx.x < kWave) {
                const u32 l = threadIdx.x;
                const bool valid = l < W;
                u64 v = valid ? wkey[l] : ~0ull;
#pragma unroll
                for (u32 kk = 2; kk <= u32(kWave); kk <<= 1) {
#pragma unroll
                    for (u32 jj = kk >> 1; jj > 0; jj >>= 1) {
                        const u64 o = __shfl_xor(v, int(jj), kWave);
                        v = (((l & jj) == 0) == ((l & kk) == 0)) ? min(v, o) : max(v, o);
                    }
                }
                const u32 sl = u32(v >> 52);
                const u32 psl = __shfl_up(sl, 1, kWave);
                u32 g0 = (l == 0 || psl != sl) ? l : 0u;  // group start lane: running max of the starts
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const u32 o = __shfl_up(g0, d, kWave);
                    if (int(l) >= d) g0 = max(g0, o);
                }
                const u32 r = l - g0;
                const u32 g = valid ? gcnt[sl] : 0u;
                const u32 ia = u32(v & 0xFFF);
                double2 m = valid ? mag[ia] : make_double2(0.0, 0.0);
                const u32 mw = wmax;
                for (u32 st = 1; st < mw; st <<= 1) {
                    const double ox = __shfl_down(m.x, int(st), kWave), oy = __shfl_down(m.y, int(st), kWave);
                    if (valid && (r & (2 * st - 1)) == 0 && r + st < g) m = make_double2(m.x + ox, m.y + oy);
                }
                if (valid) {
                    wkey[l] = v;
                    if (r == 0) {
                        owner[sl] = l;
                        mag[ia] = m;
                    }
                }
            }
            __syncthreads();
        } else if (W) {
            u32 P = 1;
            while (P < W) P <<= 1;
            for (u32 q = W + threadIdx.x; q < P; q += T) wkey[q] = ~0ull;
            __syncthreads();
            for (u32 kk = 2; kk <= P; kk <<= 1) {
                for (u32 jj = kk >> 1; jj > 0; jj >>= 1) {
                    for (u32 t = threadIdx.x; t < P / 2; t += T) {
                        const u32 lo = ((t & ~(jj - 1)) << 1) | (t & (jj - 1));
                        const u32 hi = lo | jj;
                        const u64 va = wkey[lo], vb = wkey[hi];
                        if ((va > vb) == ((lo & kk) == 0)) {
                            wkey[lo] = vb;
                            wkey[hi] = va;
                        }
                    }
                    __syncthreads();
                }
            }
            // group starts in the sorted area (the owner slots are free now)
            u32 ws[PER], wg[PER], wr[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const u32 p = threadIdx.x + T * k;
                ws[k] = 0;
                if (p < W) {
                    ws[k] = u32(wkey[p] >> 52);
                    if (p == 0 || u32(wkey[p - 1] >> 52) != ws[k]) owner[ws[k]] = p;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const u32 p = threadIdx.x + T * k;
                wg[k] = wr[k] = 0;
                if (p < W) {
                    wg[k] = gcnt[ws[k]];
                    wr[k] = p - owner[ws[k]];
                }
            }
            const u32 mw = wmax;
            for (u32 st = 1; st < mw; st <<= 1) {
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const u32 p = threadIdx.x + T * k;
                    if (p < W && (wr[k] & (2 * st - 1)) == 0 && wr[k] + st < wg[k]) {
                        const u32 ia = u32(wkey[p] & 0xFFF), ib = u32(wkey[p + st] & 0xFFF);
                        const double2 va = mag[ia], vb = mag[ib];
                        mag[ia] = make_double2(va.x + vb.x, va.y + vb.y);
                    }
                }
                __syncthreads();
            }
        }
        if (W) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                if (lead[k] && gcnt[slot[k]] > u32(kGroupRegs)) {
                    const u64 w0 = wkey[owner[slot[k]]];
                    rc[k] = (w0 >> 12) & ((1ull << kWideCandBits) - 1);
                    rs[k] = mag[u32(w0 & 0xFFF)];
                }
            }
        }
        u32 keepbits = 0;
        u32 nk[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            nk[k] = 0;
            if (lead[k]) {
                const double nrm = rs[k].x * rs[k].x + rs[k].y * rs[k].y;
                if (nrm > tol) {
                    keepbits |= 1u << k;
                    nk[k] = norm_top32(nrm);
                    atomicAdd(&nh[min(norm_top12(nrm), u32(kNormBins - 1))], 1u);
                }
            }
        }
        // output places: one LDS counter, reserved once per wavefront (the
        // order of groups within the bucket's range carries no meaning),
        // filled entry slot by entry slot so that a store instruction's
        // lanes write side by side
        u64 km[PER];
        u32 wcnt = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            km[k] = __ballot((keepbits >> k) & 1u);
            wcnt += u32(__popcll(km[k]));
        }
        u32 wbase = 0;
        if (lane == kWave - 1 && wcnt) wbase = atomicAdd(&ocount, wcnt);
        wbase = __shfl(wbase, kWave - 1, kWave);
        const u64 below = (1ull << lane) - 1;
        u32 run = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (keepbits & (1u << k)) {
                const u64 pos = beg + wbase + run + u32(__popcll(km[k] & below));
                UH[pos] = key[threadIdx.x + T * k];
                UR[pos] = rc[k];
                UM[pos] = rs[k];
                if (UK) UK[pos] = nk[k];
            }
            run += u32(__popcll(km[k]));
        }
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (lead[k]) owner[slot[k]] = gcnt[slot[k]] = 0;
        __syncthreads();
        const u32 total = ocount;  // reset after the next bucket's staging barrier
        for (u64 jx = beg + total + threadIdx.x; jx < beg + n; jx += T) {
            UR[jx] = kNoRep;
            if (UK) UK[jx] = 0;
        }
        if (threadIdx.x == 0) uniques += total;
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
                                                    u32 *UK, u64 *counters, u64 *nhist) {
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
                if (UK) UK[pos] = norm_top32(nrm);
                ++pos;
                atomicAdd(reinterpret_cast<unsigned long long *>(nhist + min(norm_top12(nrm), u32(kNormBins - 1))),
                          1ull);
            }
        }
        atomicAdd(reinterpret_cast<unsigned long long *>(counters), (unsigned long long)(pos - beg));
        for (u64 j = pos; j < beg + n; ++j) {
            UR[j] = kNoRep;
            if (UK) UK[j] = 0;
        }
    }
}


}  // namespace qd
