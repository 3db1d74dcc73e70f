#pragma once
// Batched step: many independent simulations (instances) in one launch
// sequence.  The instances' parents are consecutive in one state, so their
// candidates are consecutive too; instance i adds i * kFold to its children's
// hashes (the dedup key), so equal objects of different instances never merge
// while the merge inside an instance is the plain step's.  Truncation and
// normalisation are per instance, in the plain step's canonical orders, so
// every instance comes out exactly as its own plain simulate() call.
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_k_dedup.hpp"
#include "qd_k_finalize.hpp"
#include "qd_k_select.hpp"
#include "qd_rules.hpp"

namespace qd {

constexpr u64 kFold = 0x9E3779B97F4A7C15ull;  // odd: key = hash + instance * kFold
constexpr u32 kMaxInst = 4096;                // instances per batched call (LDS counters)
constexpr u32 kNoInst = ~0u;

// largest i < n with seg[i] <= x (seg holds n + 1 ascending offsets)
template <class T>
__device__ inline u32 find_seg(const T *seg, u32 n, u64 x) {
    u32 lo = 0, hi = n - 1;
    while (lo < hi) {
        const u32 mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// out[i] = v[idx[i]] for i < m
__global__ __launch_bounds__(256) void k_gather_u64(const u64 *v, const u64 *idx, u32 m, u64 *out) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) out[i] = v[idx[i]];
}

// out[i] = number of set bits of the bitmap below position idx[i] (the
// survivor rank of candidate idx[i]), i < m
__global__ __launch_bounds__(256) void k_rank_at(const u64 *bitmap, const u64 *wprefix, const u64 *idx, u32 m,
                                                 u64 W, u64 *out) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const u64 c = idx[i];
        const u64 w = c >> 6;
        out[i] = w >= W ? wprefix[W] : wprefix[w] + __popcll(bitmap[w] & ((1ull << (c & 63)) - 1));
    }
}

// flags[i] = 1 when instance i holds a graph of more than 64 nodes (the
// batched kernels' limit; such an instance steps on its own).
__global__ __launch_bounds__(256) void k_wide_flags(const u8 *obj, const u64 *begin, const u32 *size, u64 N,
                                                    const u64 *oseg, u32 n, u32 *flags) {
    for (u64 p = u64(blockIdx.x) * blockDim.x + threadIdx.x; p < N; p += u64(gridDim.x) * blockDim.x)
        if (size[p] >= 2 && ld16(obj + begin[p]) > 64) flags[find_seg(oseg, n, p)] = 1;
}

// Instance of a dedup slot's representative (kNoInst for empty slots).
__device__ inline u32 slot_inst(const u64 *UR, u64 s, const u64 *cseg, u32 n) {
    const u64 r = UR[s];
    return r == kNoRep ? kNoInst : find_seg(cseg, n, r);
}

// Per-instance unique counts; sinst[s] records each slot's instance
// (kNoSlotInst for an empty slot) for the kernels after it.
constexpr u16 kNoSlotInst = 0xFFFF;
static_assert(kMaxInst < kNoSlotInst, "slot instances are 16-bit");
__device__ inline u32 slot_inst16(const u16 *sinst, u64 s) {
    const u16 v = sinst[s];
    return v == kNoSlotInst ? kNoInst : u32(v);
}

__global__ __launch_bounds__(256) void k_seg_count(const u64 *UR, u64 nslots, const u64 *cseg, u32 n, u64 *ucount,
                                                   u16 *sinst) {
    __shared__ u32 cnt[kMaxInst];
    __shared__ u64 sseg[kMaxInst + 1];  // the instances' candidate offsets, searched per slot
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) cnt[i] = 0;
    for (u32 i = threadIdx.x; i <= n; i += blockDim.x) sseg[i] = cseg[i];
    __syncthreads();
    for (u64 s = u64(blockIdx.x) * blockDim.x + threadIdx.x; s < nslots; s += u64(gridDim.x) * blockDim.x) {
        const u32 inst = slot_inst(UR, s, sseg, n);
        sinst[s] = inst == kNoInst ? kNoSlotInst : u16(inst);
        if (inst != kNoInst) atomicAdd(&cnt[inst], 1u);
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < n; i += blockDim.x)
        if (cnt[i]) atomicAdd(reinterpret_cast<unsigned long long *>(ucount + i), (unsigned long long)cnt[i]);
}

// Radix-select state of one instance: the K-th largest 128-bit key
// (|m|^2 bits, ~hash) is being located 12 bits at a time; `prefix` (hi, lo)
// holds the `pos` bits found so far (lower bits zero), r how many keys with
// that prefix are still to be kept, tied how many keys carry that prefix.
// When active is 0 the instance keeps every key >= prefix (keepall: every
// key).
struct SegSel {
    u64 hi, lo, r, tied;
    u32 pos, active, keepall, err;
};

__device__ inline void seg_key(const u64 *UH, const double2 *UM, u64 slot, u32 inst, const SelKey &sk, u64 &hi,
                               u64 &lo) {
    const double2 m = UM[slot];
    const u64 h = UH[slot] - u64(inst) * kFold;  // the instance fold taken back off
    hi = sel_hi(m.x, m.y, h, sk);
    lo = ~h;
}

// Key record of a batched list: the high key word and the slot; the low
// word (~hash) is read back from the slot when a digit reaches it.
struct SegItem {
    u64 hi, slot;
};
__device__ inline u64 seg_lo(const u64 *UH, u64 slot, u32 inst) { return ~(UH[slot] - u64(inst) * kFold); }

__device__ inline bool prefix_eq(u64 hi, u64 lo, const SegSel &q) {
    if (q.pos == 0) return true;
    if (q.pos <= 64) return (hi >> (64 - q.pos)) == (q.hi >> (64 - q.pos));
    return hi == q.hi && (lo >> (128 - q.pos)) == (q.lo >> (128 - q.pos));
}

// Groups the uniques of truncated instances by instance as key records
// {hi, slot} (read once, in slot order): each block ranks its tile's slots
// per instance in LDS and reserves one range per instance.  With UK (the
// dedup's top 32 bits of each slot's |m|^2, plain keys) a record's hi holds
// only those bits (lower half zero) and the key is read in full from the
// slot once a digit reaches past them (SegLoop::part).
// A tile of 16384 slots per block of 1024 threads: one cursor reservation
// per instance per tile (fewer same-address atomics on the instances'
// cursors than tiles of 4096).
constexpr int kSegTile = 16384;
constexpr int kSegScatterThreads = 1024;
__global__ __launch_bounds__(kSegScatterThreads) void k_seg_scatter(const u16 *sinst, u64 nslots, const u64 *UH, const double2 *UM,
                                                     const u32 *UK, const u32 *trunc, u32 n, u64 *cursor,
                                                     SegItem *list, SelKey sk) {
    __shared__ u32 cnt[kMaxInst];
    __shared__ u64 base[kMaxInst];
    constexpr int kPer = kSegTile / kSegScatterThreads;
    for (u64 t0 = u64(blockIdx.x) * kSegTile; t0 < nslots; t0 += u64(gridDim.x) * kSegTile) {
        for (u32 i = threadIdx.x; i < n; i += blockDim.x) cnt[i] = 0;
        __syncthreads();
        u32 inst[kPer], rank[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const u64 s = t0 + threadIdx.x + u64(q) * kSegScatterThreads;
            inst[q] = kNoInst;
            if (s < nslots) {
                const u32 i = slot_inst16(sinst, s);
                if (i != kNoInst && trunc[i]) {
                    inst[q] = i;
                    rank[q] = atomicAdd(&cnt[i], 1u);
                }
            }
        }
        __syncthreads();
        for (u32 i = threadIdx.x; i < n; i += blockDim.x)
            if (cnt[i]) base[i] = atomicAdd(reinterpret_cast<unsigned long long *>(cursor + i), (unsigned long long)cnt[i]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if (inst[q] != kNoInst) {
                const u64 s = t0 + threadIdx.x + u64(q) * kSegScatterThreads;
                SegItem it;
                if (UK) {
                    it.hi = u64(UK[s]) << 32;
                } else {
                    u64 lo;
                    seg_key(UH, UM, s, inst[q], sk, it.hi, lo);
                }
                it.slot = s;
                list[base[inst[q]] + rank[q]] = it;
            }
        __syncthreads();
    }
}

// With plain keys (UK) and fewer than 2^32 candidates: the per-instance
// unique counts, the slots' instances and the first key list in one pass
// over the slots.  Every unique's record goes into its instance's range of
// the list, which starts at the instance's first candidate (relative to
// instance 0's; an instance has no more uniques than candidates), as 8
// bytes: the top 32 key bits over the slot.  cursor[i] starts at that range
// and ends one past its last record, so instance i holds cursor[i] minus
// its start uniques.  Records of instances that keep every unique are
// written too and never read.  Tiles of 8192 slots per 512-thread block
// (one cursor reservation per instance per tile): several blocks share a
// CU, so that one block's barriers and reservations overlap another's
// loads; each slot's instance and rank share one register.
constexpr int kSegFirstThreads = 512;
constexpr int kSegFirstTile = 8192;
static_assert(kSegFirstTile <= 65536 && kMaxInst < 65536, "instance and rank share a 32-bit word");
__global__ __launch_bounds__(kSegFirstThreads) void k_seg_scatter_first(const u64 *UR, const u32 *UK, u64 nslots,
                                                                          const u64 *cseg, u32 n, u64 *cursor,
                                                                          u16 *sinst, u64 *list8) {
    __shared__ u32 cnt[kMaxInst];
    __shared__ u32 base[kMaxInst];
    __shared__ u32 sseg[kMaxInst + 1];
    const u64 c0 = cseg[0];
    for (u32 i = threadIdx.x; i <= n; i += blockDim.x) sseg[i] = u32(cseg[i] - c0);
    constexpr int kPer = kSegFirstTile / kSegFirstThreads;
    for (u64 t0 = u64(blockIdx.x) * kSegFirstTile; t0 < nslots; t0 += u64(gridDim.x) * kSegFirstTile) {
        for (u32 i = threadIdx.x; i < n; i += blockDim.x) cnt[i] = 0;
        __syncthreads();
        u32 ir[kPer];  // instance << 16 | rank in the tile; kNoInst for an empty slot
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const u64 s = t0 + threadIdx.x + u64(q) * kSegFirstThreads;
            ir[q] = kNoInst;
            if (s < nslots) {
                const u64 r = UR[s];
                u16 si = kNoSlotInst;
                if (r != kNoRep) {
                    const u32 inst = find_seg(sseg, n, r - c0);
                    ir[q] = (inst << 16) | atomicAdd(&cnt[inst], 1u);
                    si = u16(inst);
                }
                sinst[s] = si;
            }
        }
        __syncthreads();
        for (u32 i = threadIdx.x; i < n; i += blockDim.x)
            if (cnt[i]) base[i] = u32(atomicAdd(reinterpret_cast<unsigned long long *>(cursor + i), (unsigned long long)cnt[i]));
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if (ir[q] != kNoInst) {
                const u64 s = t0 + threadIdx.x + u64(q) * kSegFirstThreads;
                list8[base[ir[q] >> 16] + (ir[q] & 0xFFFFu)] = (u64(UK[s]) << 32) | (s & 0xFFFFFFFFull);
            }
        __syncthreads();
    }
}

// Tile of a key-record list: records [begin, end) of instance inst.
struct SegTile {
    u64 begin, end;
    u32 inst, pad;
};

// The per-instance digit rounds, driven from the device: the state below
// lives in device memory, so that a fixed sequence of rounds (rounds after
// the last needed one find no active instance) needs no host round trip
// and is captured once in a hipGraph.  A round is k_seg_tiles (the tiles
// of the active instances' lists), k_seg_hist, k_seg_pick, k_seg_next (the
// next lists' offsets) and k_seg_collect.
constexpr u64 kSegTileItems = 16384;
constexpr int kSegLoopGrid = 4096;  // blocks of the tile kernels
struct SegLoop {
    SegItem *list[2];  // a round reads list[k & 1] and writes list[(k + 1) & 1]
    SegTile *tiles;    // at most (items / kSegTileItems + n) tiles
    u32 *hist;         // n rows of kHistBins counters, zero between rounds
    SegSel *sel;
    u64 *off, *cnt;    // an active instance's current list range
    u64 *cursor;       // its next list's cursor
    const u64 *UH;
    const double2 *UM;
    SelKey sk;
    u32 n, ntiles;
    u32 part;    // the first list's records hold the top 32 key bits only (UK)
    u32 narrow;  // the first list holds k_seg_scatter_first's 8-byte records
};

// Record k of a round's list; the first round reads the 8-byte records when
// SegLoop::narrow is set.
__device__ inline SegItem seg_item(const SegItem *list, u64 k, bool narrow) {
    if (narrow) {
        const u64 v = reinterpret_cast<const u64 *>(list)[k];
        return SegItem{v & ~0xFFFFFFFFull, v & 0xFFFFFFFFull};
    }
    return list[k];
}

// A record's full high key word when the current digit (bits [pos, pos +
// 12)) reaches past what it holds.
__device__ inline u64 seg_hi_at(const SegLoop *L, const SegItem &it, u32 inst, u32 pos_end) {
    if (!L->part || pos_end <= 32) return it.hi;
    u64 hi, lo;
    seg_key(L->UH, L->UM, it.slot, inst, L->sk, hi, lo);
    return hi;
}

// One block of 1024 threads: the tiles of every active instance's list.
// Per chunk of 1024 instances, the tile counts' scan in LDS; then each
// thread writes every 1024th tile of the chunk (its instance found by a
// search of the scan), so that an instance of many tiles is not written by
// one thread.
__global__ __launch_bounds__(1024) void k_seg_tiles(SegLoop *L) {
    __shared__ u64 tmp[1024 / kWave + 1];
    __shared__ u64 carry;
    __shared__ u64 sfirst[1024], soff[1024], scnt[1024];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (u32 i0 = 0; i0 < L->n; i0 += blockDim.x) {
        const u32 i = i0 + threadIdx.x;
        u64 nt = 0, off = 0, cnt = 0;
        if (i < L->n && L->sel[i].active) {
            off = L->off[i];
            cnt = L->cnt[i];
            nt = (cnt + kSegTileItems - 1) / kSegTileItems;
        }
        u64 total;
        sfirst[threadIdx.x] = block_excl_scan(nt, tmp, total);
        soff[threadIdx.x] = off;
        scnt[threadIdx.x] = cnt;
        __syncthreads();
        const u64 c0 = carry;
        for (u64 t = threadIdx.x; t < total; t += blockDim.x) {
            // the last thread j of the chunk with sfirst[j] <= t: its
            // instance has tiles (a thread with none shares its first with
            // the next thread)
            u32 lo = 0, hi = blockDim.x - 1;
            while (lo < hi) {
                const u32 mid = (lo + hi + 1) >> 1;
                if (sfirst[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            const u64 b = t - sfirst[lo];
            L->tiles[c0 + t] =
                SegTile{soff[lo] + b * kSegTileItems, soff[lo] + min(scnt[lo], (b + 1) * kSegTileItems), i0 + lo, 0};
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) L->ntiles = u32(carry);
}

// One workgroup per tile of an active instance: LDS histogram of the 12
// key bits at the instance's position, added into its global row.  Every
// record of the list carries the instance's current prefix.
__global__ __launch_bounds__(256) void k_seg_hist(const SegLoop *L, u32 parity, u32 first) {
    __shared__ u32 h[kHistBins];
    const SegItem *list = L->list[parity];
    const bool narrow = first && L->narrow;
    const u32 ntiles = L->ntiles;
    for (u32 t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const SegTile tl = L->tiles[t];
        const SegSel q = L->sel[tl.inst];
        if (!q.active) continue;  // uniform over the block
        for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = 0;
        __syncthreads();
        for (u64 k = tl.begin + threadIdx.x; k < tl.end; k += blockDim.x) {
            const SegItem it = seg_item(list, k, narrow);
            const u64 hi = seg_hi_at(L, it, tl.inst, q.pos + 12);
            const u64 lo = q.pos + 12 > 64 ? seg_lo(L->UH, it.slot, tl.inst) : 0;
            atomicAdd(&h[key_digit(hi, lo, q.pos)], 1u);
        }
        __syncthreads();
        u32 *row = L->hist + u64(tl.inst) * kHistBins;
        for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
            if (h[i]) atomicAdd(&row[i], h[i]);
        __syncthreads();
    }
}

// After a pick: carries the records that match the instance's new prefix
// (the tied ones) into the next list, at the instance's cursor.  Each
// thread holds kCollectPer records of a chunk of the tile; each store
// instruction's lanes write side by side (wavefront ballots), and the
// block makes one reservation per chunk.
constexpr int kCollectPer = 16;
__global__ __launch_bounds__(256) void k_seg_collect(const SegLoop *L, u32 parity, u32 first) {
    constexpr int kWaves = 256 / kWave;
    __shared__ u64 wsum[kWaves];
    __shared__ u64 base;
    const SegItem *list = L->list[parity];
    SegItem *out = L->list[parity ^ 1u];
    const bool narrow = first && L->narrow;
    const u32 ntiles = L->ntiles;
    const u32 lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const u64 below = (1ull << lane) - 1;
    for (u32 t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const SegTile tl = L->tiles[t];
        const SegSel q = L->sel[tl.inst];
        if (!q.active) continue;  // uniform over the block
        for (u64 k0 = tl.begin; k0 < tl.end; k0 += 256 * kCollectPer) {
            SegItem it[kCollectPer];
            u64 m[kCollectPer];
            u64 wtot = 0;
#pragma unroll
            for (int j = 0; j < kCollectPer; ++j) {
                const u64 k = k0 + threadIdx.x + 256 * j;
                bool take = false;
                it[j] = SegItem{0, 0};
                if (k < tl.end) {
                    it[j] = seg_item(list, k, narrow);
                    it[j].hi = seg_hi_at(L, it[j], tl.inst, q.pos);  // (carried in full from here on)
                    take = prefix_eq(it[j].hi, q.pos > 64 ? seg_lo(L->UH, it[j].slot, tl.inst) : 0, q);
                }
                m[j] = __ballot(take);
                wtot += u64(__popcll(m[j]));
            }
            if (lane == 0) wsum[wid] = wtot;
            __syncthreads();
            if (threadIdx.x == 0) {
                u64 acc = 0;
                for (int w = 0; w < kWaves; ++w) {
                    const u64 c = wsum[w];
                    wsum[w] = acc;
                    acc += c;
                }
                if (acc)
                    base = atomicAdd(reinterpret_cast<unsigned long long *>(L->cursor + tl.inst),
                                     (unsigned long long)acc);
            }
            __syncthreads();
            if (wtot) {  // uniform over the wavefront
                u64 pos = base + wsum[wid];
#pragma unroll
                for (int j = 0; j < kCollectPer; ++j) {
                    if ((m[j] >> lane) & 1) out[pos + u64(__popcll(m[j] & below))] = it[j];
                    pos += u64(__popcll(m[j]));
                }
            }
            __syncthreads();
        }
    }
}

// One workgroup (256 threads) per instance (a grid-stride loop over them):
// picks the digit holding the r-th largest key from the instance's
// histogram row, updates the state and clears the row for the next round.
__global__ __launch_bounds__(256) void k_seg_pick(SegLoop *L) {
    __shared__ u64 part[256];
    __shared__ u32 found_d;
    __shared__ u64 found_above, found_cnt;
    const u32 n = L->n;
    for (u32 inst = blockIdx.x; inst < n; inst += gridDim.x) {
        SegSel q = L->sel[inst];
        if (!q.active) continue;  // uniform over the block
        u32 *row = L->hist + u64(inst) * kHistBins;
        constexpr int kPer = kHistBins / 256;  // 16 bins per thread, thread 0 holds the top ones
        const u32 t = threadIdx.x;
        if (t == 0) found_cnt = ~0ull;
        u64 local = 0;
        for (int j = 0; j < kPer; ++j) local += row[kHistBins - 1 - (t * kPer + j)];
        part[t] = local;
        __syncthreads();
        // inclusive prefix over threads in order (bins descending)
        for (u32 d = 1; d < 256; d <<= 1) {
            const u64 v = t >= d ? part[t - d] : 0;
            __syncthreads();
            part[t] += v;
            __syncthreads();
        }
        const u64 before = part[t] - local;
        if (before < q.r && q.r <= before + local) {
            u64 cum = before;
            for (int j = 0; j < kPer; ++j) {
                const u32 d = kHistBins - 1 - (t * kPer + j);
                const u64 c = row[d];
                if (cum + c >= q.r) {
                    found_d = d;
                    found_above = cum;
                    found_cnt = c;
                    break;
                }
                cum += c;
            }
        }
        __syncthreads();
        for (int i = t; i < kHistBins; i += 256) row[i] = 0;
        if (t == 0) {
            if (found_cnt == ~0ull) {  // fewer keys than r: inconsistent counts
                q.active = 0;
                q.err = 1;
            } else {
                // the digit covers key bits [116 - pos, 128 - pos) (bit 127 = top of hi)
                const u64 d = found_d;
                if (q.pos < 64) {
                    const u32 sh = 64 - q.pos;
                    if (sh >= 12) q.hi |= d << (sh - 12);
                    else {
                        q.hi |= d >> (12 - sh);
                        q.lo |= d << (52 + sh);
                    }
                } else if (q.pos <= 116) {
                    q.lo |= d << (116 - q.pos);
                } else {
                    q.lo |= d >> (q.pos - 116);  // pos 120: the last 8 key bits
                }
                q.pos = min(q.pos + 12, 128u);
                q.r -= found_above;
                q.tied = found_cnt;
                if (found_cnt == q.r) q.active = 0;
                else if (q.pos >= 128) {
                    q.active = 0;
                    q.err = 1;  // equal keys: cannot happen, hashes are distinct in an instance
                }
            }
            L->sel[inst] = q;
        }
        __syncthreads();  // the shared results are reused by the next instance
    }
}

// One block of 1024 threads, after a pick: the instances still active get
// their tied records' range in the next list (in instance order), and the
// cursor of the collect.
__global__ __launch_bounds__(1024) void k_seg_next(SegLoop *L) {
    __shared__ u64 tmp[1024 / kWave + 1];
    __shared__ u64 carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (u32 i0 = 0; i0 < L->n; i0 += blockDim.x) {
        const u32 i = i0 + threadIdx.x;
        u64 c = 0;
        if (i < L->n && L->sel[i].active) c = L->sel[i].tied;
        u64 total;
        const u64 off = carry + block_excl_scan(c, tmp, total);
        if (i < L->n) {
            L->off[i] = off;
            L->cnt[i] = c;
            L->cursor[i] = off;
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
}

// Marks the survivors: every unique of an instance that is not truncated,
// and the keys >= prefix of a truncated one.  Lanes hold consecutive slots.
// With UK each instance's top 32 threshold bits sit in LDS (dynamic, n
// words; 0 for an instance that keeps everything: every key passes a zero
// threshold), so that a lane reads its instance's threshold from the block
// and the full key and threshold only when the top 32 bits tie.
__global__ __launch_bounds__(256) void k_seg_mark(const u16 *sinst, const u64 *UH, const u64 *UR, const double2 *UM,
                                                  const u32 *UK, u64 nslots, const SegSel *sel, u32 n, u64 *bitmap,
                                                  u64 *sbm, SelKey sk) {
    extern __shared__ u32 qtop[];
    if (UK)
        for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
            const SegSel q = sel[i];
            qtop[i] = q.keepall ? 0u : u32(q.hi >> 32);
        }
    __syncthreads();
    const u64 stride = u64(gridDim.x) * blockDim.x;
    for (u64 base = u64(blockIdx.x) * blockDim.x; base < nslots; base += stride) {
        const u64 s = base + threadIdx.x;
        bool keep = false;
        if (s < nslots) {
            const u32 inst = slot_inst16(sinst, s);
            if (inst != kNoInst) {
                if (UK) {
                    // the top 32 key bits decide unless they tie (plain keys)
                    const u32 top = UK[s], qt = qtop[inst];
                    if (top != qt) {
                        keep = top > qt;
                    } else {
                        const SegSel q = sel[inst];
                        if (q.keepall) keep = true;
                        else {
                            u64 hi, lo;
                            seg_key(UH, UM, s, inst, sk, hi, lo);
                            keep = hi > q.hi || (hi == q.hi && lo >= q.lo);
                        }
                    }
                } else {
                    const SegSel q = sel[inst];
                    if (q.keepall) keep = true;
                    else {
                        u64 hi, lo;
                        seg_key(UH, UM, s, inst, sk, hi, lo);
                        keep = hi > q.hi || (hi == q.hi && lo >= q.lo);
                    }
                }
            }
        }
        const u64 first = s & ~u64(kWave - 1);
        if (first < nslots) select_word(keep, s, first, UR, bitmap, sbm);
    }
}

// Segmented canonical sum, level 1: chunk c (global) of instance i covers
// survivors [kseg[i] + j * kSumChunk, ...) where j = c - chunk0[i]; sums
// |m|^2 exactly as k_sum_chunks does.
__global__ __launch_bounds__(kSumThreads) void k_seg_sum_chunks(const double2 *mag, const u64 *kseg,
                                                                const u64 *chunk0, u32 n, double *out) {
    __shared__ double s[kSumThreads];
    const u64 c = blockIdx.x;
    const u32 inst = find_seg(chunk0, n, c);
    const u64 beg = kseg[inst] + (c - chunk0[inst]) * kSumChunk, end = kseg[inst + 1];
    double a = 0.0;
    for (int j = 0; j < kSumChunk / kSumThreads; ++j) {
        const u64 i = beg + threadIdx.x + u64(kSumThreads) * j;
        if (i < end) {
            const double2 m = mag[i];
            a += m.x * m.x + m.y * m.y;
        }
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int stride = kSumThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s[0];
}

// Level 2, one block per instance over its (at most kSumChunk) chunk sums.
__global__ __launch_bounds__(kSumThreads) void k_seg_sum_final(const double *csum, const u64 *chunk0, u32 n,
                                                               double *total) {
    __shared__ double s[kSumThreads];
    const u32 inst = blockIdx.x;
    const u64 beg = chunk0[inst], end = chunk0[inst + 1];
    double a = 0.0;
    for (int j = 0; j < kSumChunk / kSumThreads; ++j) {
        const u64 i = beg + threadIdx.x + u64(kSumThreads) * j;
        if (i < end) a += csum[i];
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int stride = kSumThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[inst] = s[0];
}

// m /= sqrt(total of its instance), skipped where that factor is exactly 1.
__global__ __launch_bounds__(256) void k_seg_scale(double2 *mag, u64 K, const u64 *kseg, u32 n, const double *total) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x) {
        const u32 inst = find_seg(kseg, n, k);
        const double f = sqrt(total[inst]);
        if (f != 1.0) {
            const double2 m = mag[k];
            mag[k] = make_double2(m.x / f, m.y / f);
        }
    }
}

}  // namespace qd
