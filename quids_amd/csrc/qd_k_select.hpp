#pragma once
// top-K selection on (|m|^2, hash) and survivor marking
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

// ===========================================================================
// phase 6: top-K select on the 128-bit key (|m|^2 bits, ~hash)
// ===========================================================================
// A key under selection: (hi, lo) = (|m|^2 bits, ~hash) for uniques, whose
// `slot` is their dedup output slot; (|m|^2 bits, ~index) for parents, whose
// `slot` is their index.
struct SelItem {
    u64 hi, lo, slot;
};

__device__ inline u32 key_digit(u64 hi, u64 lo, u32 pos) {
    u64 top;
    if (pos == 0) top = hi;
    else if (pos < 64) top = (hi << pos) | (lo >> (64 - pos));
    else top = lo << (pos - 64);
    return u32(top >> 52);
}

// Appends for the lanes with pred set: one atomic per wave; returns this
// lane's slot (meaningful only when pred).
__device__ inline u64 wave_append(bool pred, u64 *counter) {
    const u64 mask = __ballot(pred);
    if (!mask) return 0;
    const int lane = threadIdx.x & (kWave - 1);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    u64 base = 0;
    if (lane == leader) base = atomicAdd(reinterpret_cast<unsigned long long *>(counter), (unsigned long long)__popcll(mask));
    base = __shfl(base, leader, kWave);
    return base + __popcll(mask & ((1ull << lane) - 1));
}

__device__ inline void mark(u64 *bitmap, u64 rep) {
    if (bitmap) atomicOr(reinterpret_cast<unsigned long long *>(bitmap + (rep >> 6)), 1ull << (rep & 63));
}

// Marks a selected item: bit `key` of the bitmap (the representative, or
// the parent index when UR is null) and, with a slot bitmap, bit `slot`.
__device__ inline void select_item(bool pred, u64 slot, const u64 *UR, u64 *bitmap, u64 *sbm) {
    if (!pred) return;
    mark(bitmap, UR ? UR[slot] : slot);
    if (sbm) mark(sbm, slot);
}

// The same for a wave whose lanes hold the 64 consecutive slots of word
// `first >> 6` (first a multiple of 64, below U): the slot word is written
// whole, so the slot bitmap needs no clearing for the words these kernels
// cover.
__device__ inline void select_word(bool pred, u64 slot, u64 first, const u64 *UR, u64 *bitmap, u64 *sbm) {
    if (pred) mark(bitmap, UR[slot]);
    const u64 m = __ballot(pred);
    if ((threadIdx.x & (kWave - 1)) == 0) sbm[first >> 6] = m;
}

// Histogram of the top 12 key bits over the slots holding a unique.
__global__ __launch_bounds__(256) void k_key_hist0(const u64 *UH, const u64 *UR, const double2 *UM, u64 U, u64 *hist,
                                                   SelKey sk) {
    __shared__ u32 h[kHistBins];
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < U; i += u64(gridDim.x) * blockDim.x) {
        if (UR[i] == kNoRep) continue;
        const double2 m = UM[i];
        atomicAdd(&h[u32(sel_hi(m.x, m.y, UH[i], sk) >> 52)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
        if (h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist + i), (unsigned long long)h[i]);
}

// First select stage over the dedup output slots: histogram of |m|^2 bits
// 51..40 among the uniques whose bits 63..52 equal d1.
__global__ __launch_bounds__(256) void k_sel_hist_u(const u64 *UH, const u64 *UR, const double2 *UM, u64 U, u32 d1,
                                                    u64 *hist, SelKey sk) {
    __shared__ u32 h[kHistBins];
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < U; i += u64(gridDim.x) * blockDim.x) {
        if (UR[i] == kNoRep) continue;
        const double2 m = UM[i];
        const u64 bits = sel_hi(m.x, m.y, UH[i], sk);
        if (u32(bits >> 52) == d1) atomicAdd(&h[u32(bits >> 40) & 0xFFF], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
        if (h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist + i), (unsigned long long)h[i]);
}

// Marks the uniques whose top 24 |m|^2 bits exceed prefix24 and collects
// the ones equal to it.
__global__ __launch_bounds__(256) void k_sel_first(const u64 *UH, const u64 *UR, const double2 *UM, u64 U,
                                                   u32 prefix24, SelItem *S, u64 *scount, u64 *bitmap,
                                                   u64 *sbm, SelKey sk) {
    const u64 stride = u64(gridDim.x) * blockDim.x;
    // uniform trip count per wave so that the appends see every lane
    for (u64 base = u64(blockIdx.x) * blockDim.x; base < U; base += stride) {
        const u64 i = base + threadIdx.x;
        bool take = false, keep = false;
        SelItem it{};
        if (i < U && UR[i] != kNoRep) {
            const double2 m = UM[i];
            const u64 bits = sel_hi(m.x, m.y, UH[i], sk);
            const u32 t = u32(bits >> 40);
            if (t > prefix24) keep = true;
            else if (t == prefix24) {
                take = true;
                it = SelItem{bits, ~UH[i], i};
            }
        }
        const u64 first = i & ~u64(kWave - 1);
        if (first < U) select_word(keep, i, first, UR, bitmap, sbm);
        const u64 at = wave_append(take, scount);
        if (take) S[at] = it;
    }
}

// The same two stages from the slots' top 32 |m|^2 bits UK (0 for an empty
// slot), written by the dedup; plain |m|^2 keys only.  The threshold's top
// 12 bits d1 must be nonzero, so that an empty slot never ties.
__global__ __launch_bounds__(256) void k_sel_hist_k(const u32 *UK, u64 U, u32 d1, u64 *hist) {
    __shared__ u32 h[kHistBins];
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    constexpr int kUnroll = 4;  // loads in flight per lane
    const u64 n4 = U / 4, stride = u64(gridDim.x) * blockDim.x;
    const uint4 *U4 = reinterpret_cast<const uint4 *>(UK);
    auto count = [&](u32 v) {
        if ((v >> 20) == d1) atomicAdd(&h[(v >> 8) & 0xFFF], 1u);
    };
    for (u64 q0 = u64(blockIdx.x) * blockDim.x + threadIdx.x; q0 < n4; q0 += kUnroll * stride) {
        uint4 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const u64 q = q0 + u * stride;
            v[u] = q < n4 ? U4[q] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            count(v[u].x);
            count(v[u].y);
            count(v[u].z);
            count(v[u].w);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < U - 4 * n4) count(UK[4 * n4 + threadIdx.x]);
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
        if (h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist + i), (unsigned long long)h[i]);
}

__global__ __launch_bounds__(256) void k_sel_first_k(const u32 *UK, const u64 *UH, const u64 *UR, const double2 *UM,
                                                     u64 U, u32 prefix24, SelItem *S, u64 *scount, u64 *bitmap,
                                                     u64 *sbm) {
    constexpr int kUnroll = 4;  // loads in flight per lane
    const u64 stride = u64(gridDim.x) * blockDim.x;
    const SelKey plain{0u, 0u, 0ull};
    // uniform trip count per wave so that the ballots see every lane; each
    // wave's lanes hold 64 consecutive slots
    for (u64 base = u64(blockIdx.x) * blockDim.x; base < U; base += kUnroll * stride) {
        u32 t[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const u64 i = base + u * stride + threadIdx.x;
            t[u] = i < U ? UK[i] >> 8 : 0u;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const u64 i = base + u * stride + threadIdx.x;
            const bool keep = i < U && t[u] > prefix24;
            const bool take = i < U && t[u] == prefix24;
            SelItem it{};
            if (take) {
                const double2 m = UM[i];
                const u64 h = UH[i];
                it = SelItem{sel_hi(m.x, m.y, h, plain), ~h, i};
            }
            const u64 first = i & ~u64(kWave - 1);
            if (first < U) select_word(keep, i, first, UR, bitmap, sbm);
            const u64 at = wave_append(take, scount);
            if (take) S[at] = it;
        }
    }
}

// ---------------------------------------------------------------------------
// The select's later rounds, driven from the device: each round histograms
// one 12-bit key digit of the tied items, picks the digit that holds the
// r-th largest key, marks the items above it and carries the items on it to
// the next round.  Counts and decisions live in SelLoop in device memory, so
// a fixed sequence of rounds (rounds after the last one return at once)
// needs no host round trip and is captured once in a hipGraph.
// ---------------------------------------------------------------------------
struct SelLoop {
    SelItem *items[2];  // the current and the next round's tied items
    const u64 *UR;      // representative of each slot (null: the slot is the key's index)
    u64 *bitmap, *sbm;
    u64 *hist;          // kHistBins counters, zero between rounds
    u64 r;              // keys still to select among the current items (1 <= r < ns)
    u64 ns;             // current items
    u64 cnt;            // items carried so far in this round
    u64 tied;           // items on the picked digit
    u32 cur;            // items[cur] holds the current items
    u32 pos;            // key bit of the current digit
    u32 d;              // the picked digit
    u32 state;          // 0 running, 1 done, 2 no convergence (equal keys)
};
constexpr int kSelLoopGrid = 1024;  // blocks of the histogram and split kernels

__global__ __launch_bounds__(1024) void k_sl_clear(SelLoop *L) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < kHistBins; i += gridDim.x * blockDim.x) L->hist[i] = 0;
}

__global__ __launch_bounds__(256) void k_sl_hist(const SelLoop *L) {
    if (L->state) return;  // uniform
    __shared__ u32 h[kHistBins];
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const SelItem *S = L->items[L->cur];
    const u64 n = L->ns;
    const u32 pos = L->pos;
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
        atomicAdd(&h[key_digit(S[i].hi, S[i].lo, pos)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
        if (h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(L->hist + i), (unsigned long long)h[i]);
}

// One block of 1024 threads: the digit d whose keys, with those of every
// higher digit, first reach r; then clears the histogram.
__global__ __launch_bounds__(1024) void k_sl_pick(SelLoop *L) {
    if (L->state) return;  // uniform
    __shared__ u64 tmp[1024 / kWave + 1];
    constexpr int kPer = kHistBins / 1024;
    const u64 r = L->r;  // read by every thread before the scan's barriers, written after them
    u64 v[kPer];
    u64 sum = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {  // digits in descending order
        v[k] = L->hist[kHistBins - 1 - (kPer * threadIdx.x + k)];
        sum += v[k];
    }
    u64 total;
    u64 above = block_excl_scan(sum, tmp, total);  // keys of higher digits than this thread's
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (above < r && above + v[k] >= r) {  // exactly one digit
            L->d = u32(kHistBins - 1 - (kPer * threadIdx.x + k));
            L->r = r - above;
            L->tied = v[k];
            L->cnt = 0;
        }
        above += v[k];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) L->hist[kHistBins - 1 - (kPer * threadIdx.x + k)] = 0;
}

// Marks the items above the picked digit (and those on it when all of them
// are taken); carries the others on it to the next round's buffer.
__global__ __launch_bounds__(256) void k_sl_split(SelLoop *L) {
    if (L->state) return;  // uniform
    const SelItem *S = L->items[L->cur];
    SelItem *out = L->items[L->cur ^ 1u];
    const u64 n = L->ns;
    const u32 pos = L->pos, d = L->d;
    const bool all = L->r == L->tied;
    const u64 stride = u64(gridDim.x) * blockDim.x;
    // uniform trip count per wave so that the appends see every lane
    for (u64 base = u64(blockIdx.x) * blockDim.x; base < n; base += stride) {
        const u64 i = base + threadIdx.x;
        bool take = false, keep = false;
        SelItem it{};
        if (i < n) {
            it = S[i];
            const u32 t = key_digit(it.hi, it.lo, pos);
            keep = t > d || (all && t == d);
            take = !all && t == d;
        }
        select_item(keep, it.slot, L->UR, L->bitmap, L->sbm);
        const u64 at = wave_append(take, &L->cnt);
        if (take) out[at] = it;
    }
}

__global__ void k_sl_advance(SelLoop *L) {
    if (threadIdx.x || blockIdx.x || L->state) return;
    if (L->r == L->tied) {
        L->state = 1;
        return;
    }
    L->ns = L->tied;
    L->cur ^= 1u;
    L->pos += 12;
    if (L->pos >= 128) L->state = 2;
}

__global__ __launch_bounds__(256) void k_sel_mark_items(const SelItem *S, u64 n, const u64 *UR, u64 *bitmap,
                                                        u64 *sbm) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
        select_item(true, S[i].slot, UR, bitmap, sbm);
}

__global__ __launch_bounds__(256) void k_mark_all(const u64 *UR, u64 U, u64 *bitmap, u64 *sbm) {
    const u64 stride = u64(gridDim.x) * blockDim.x;
    for (u64 base = u64(blockIdx.x) * blockDim.x; base < U; base += stride) {
        const u64 i = base + threadIdx.x;
        const u64 first = i & ~u64(kWave - 1);
        if (first < U) select_word(i < U && UR[i] != kNoRep, i, first, UR, bitmap, sbm);
    }
}

}  // namespace qd
