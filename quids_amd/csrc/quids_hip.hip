// quids_hip.hip -- MI355X (gfx950) implementation of the QuIDS
// rule-application step behind the C ABI of include/quids_hip.h.
//
// One simulate() call (reference: quids::simulate, src/quids.hpp:448-543)
// runs these device phases on the context's stream:
//
//  1 prep      per parent: num_child (get_num_child, quids.hpp:548-569) and a
//              per-parent hash table for the fast child hash (qd_rules.hpp)
//  2 scan      exclusive prefix of num_child -> child_begin, C candidates
//              (quids.hpp:568, 666-673)
//  3 generate  per candidate: hash + magnitude, without materialising the
//              child (quids.hpp:705-719); 32-byte entries in candidate order
//              plus a histogram of the top hash bits
//  4 partition one or two MSD radix passes on the top hash bits, LDS-ranked,
//              until every bucket fits one workgroup's LDS (the reference's
//              hash-range bucketing, quids.hpp:740-759)
//  5 dedup     per bucket, in LDS: group equal hashes, representative = the
//              smallest candidate index, magnitudes summed in ascending
//              candidate order, keep |m|^2 > tolerance (quids.hpp:785-823)
//  6 select    if U > max_num_object: radix select of the top K by
//              (|m|^2 desc, hash asc) (quids.hpp:866-900)
//  7 order     survivors placed in ascending representative order through a
//              candidate bitmap and its popcount prefix (quids.hpp:927)
//  8 finalize  child sizes, padded offsets, survivors regenerated into the
//              next state (quids.hpp:930-967)
//  9 normalize |m|^2 summed in a fixed blocked-pairwise order, then scaled
//              (quids.hpp:985-1017)
//
// Everything that depends on scheduling in the reference (representative,
// merge order, truncation ties, output order) is fixed here, so results are
// deterministic run to run.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/quids_hip.h"
#include "qd_device.hpp"
#include "qd_rules.hpp"

using namespace qd;

// ===========================================================================
// errors
// ===========================================================================
static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define QD_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? QD_ERR_OOM : QD_ERR_HIP,                       \
                        std::string(#call) + ": " + hipGetErrorString(e_));                        \
    } while (0)

#define QD_TRY(expr)                                                                               \
    do {                                                                                           \
        int rc_ = (expr);                                                                          \
        if (rc_ != QD_OK) return rc_;                                                              \
    } while (0)

// ===========================================================================
// device buffers
// ===========================================================================
struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return QD_OK;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes + bytes / 8, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            e = hipMalloc(&p, std::max<size_t>(bytes, 256));
            if (e != hipSuccess) {
                (void)hipGetLastError();
                p = nullptr;
                return fail(QD_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
            }
            want = std::max<size_t>(bytes, 256);
        }
        cap = want;
        return QD_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

struct qd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // phase timing
    bool profiling = false;
    hipEvent_t ev[2 * QD_NUM_PHASES] = {};
    bool pending[QD_NUM_PHASES] = {};
    double phase_ms[QD_NUM_PHASES] = {};
    qd_tuning tune = {768, 12, 256, 2048, 2048, {0, 0, 0}};
    u64 phase_calls[QD_NUM_PHASES] = {};
    u64 *host = nullptr;  // pinned readback area, 8192 u64
    // scratch of the step
    DBuf nc, cb, pre, tab, kept, keys;
    DBuf ebuf0, ebuf1;
    DBuf hist1, off1, cur1, off2;
    DBuf nhist, counters;
    DBuf sel0, sel1;
    DBuf bitmap, wprefix, padded, chunks, scanblk;
    DBuf biglist, bigsoff, bigkey, bigcand, bigmag, bigseg;
    std::vector<DBuf *> all() {
        return {&nc, &cb, &pre, &tab, &kept, &keys, &ebuf0, &ebuf1, &hist1, &off1, &cur1, &off2,
                &nhist, &counters, &sel0, &sel1, &bitmap, &wprefix, &padded, &chunks, &scanblk,
                &biglist, &bigsoff, &bigkey, &bigcand, &bigmag, &bigseg};
    }
};

struct qd_state {
    qd_ctx *ctx = nullptr;
    DBuf objects, begin, size, mag;
    u64 n = 0;
    u64 bytes = 0;
    double total_proba = 1;
};

struct qd_symbolic {
    qd_ctx *ctx = nullptr;
    u64 num_object = 0;
    u64 num_object_after_interferences = 0;
};

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

template <class Load>
static int run_scan(qd_ctx *ctx, Load ld, u64 n, u64 *out) {
    if (n == 0) {
        QD_HIP(hipMemsetAsync(out, 0, sizeof(u64), ctx->stream));
        return QD_OK;
    }
    const u64 nb = (n + kScanTile - 1) / kScanTile;
    QD_TRY(ctx->scanblk.ensure(nb * sizeof(u64)));
    u64 *bs = ctx->scanblk.as<u64>();
    hipLaunchKernelGGL(k_scan_reduce<Load>, dim3(nb), dim3(kScanThreads), 0, ctx->stream, ld, n, bs);
    hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(1024), 0, ctx->stream, bs, nb);
    hipLaunchKernelGGL(k_scan_final<Load>, dim3(nb), dim3(kScanThreads), 0, ctx->stream, ld, n, bs, out);
    QD_HIP(hipGetLastError());
    return QD_OK;
}

// ===========================================================================
// phase 1: per-parent prep
// ===========================================================================
struct Parents {
    const u8 *obj;
    const u64 *begin;
    const u32 *size;
    const double2 *mag;
    u64 n;
};

// Validation bits reported through the error word.
enum : u32 { ERR_OBJECT = 1u, ERR_OVERFLOW = 2u, ERR_BIT = 4u };

// A QCGD object is well formed when its size matches its header and every
// node's name tree starts with a root the rules can read: an element, a .l /
// .r node with its inner node, or a pair with two non-empty halves.  Device
// code reads only roots, root+1 and root+right_or_type, so this bounds every
// access and every child size.
__device__ inline bool qcgd_valid(const u8 *o, u32 size) {
    if (size < 4) return false;
    const u32 n = ld16(o);
    if (n == 0 || 4 + 4 * u64(n) > size) return false;
    const u8 *nb = o + 2 + 2 * n;
    if (ld16(nb) != 0) return false;
    const u32 m = ld16(nb + 2 * n);
    if (4 + 4 * u64(n) + 16 * u64(m) != size) return false;
    Graph g(o);
    for (u32 i = 0; i < n; ++i) {
        const u32 b = ld16(nb + 2 * i), e = ld16(nb + 2 * i + 2);
        if (e <= b || e > m) return false;
        const u32 mi = e - b;
        const i16 t = g.sub_type(b);
        if (t >= 0) {
            if (t < 2 || u32(t) >= mi) return false;
        } else if (t == QCGD_DOT_L || t == QCGD_DOT_R) {
            if (mi < 2) return false;
        } else if (t != QCGD_ELEMENT) {
            return false;
        }
    }
    return true;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_prep(Parents P, const u64 *keep, qd_rule rule, u32 *nc, u64 *pre, u64 *tab,
                                              u32 *maxsz, u32 *err) {
    u32 local_max = 0;
    for (u64 p = u64(blockIdx.x) * blockDim.x + threadIdx.x; p < P.n; p += u64(gridDim.x) * blockDim.x) {
        const u8 *o = P.obj + P.begin[p];
        const u32 sz = P.size[p];
        if (keep && !((keep[p >> 6] >> (p & 63)) & 1)) {
            nc[p] = 0;
            continue;
        }
        u32 num = 0, ms = 0;
        if (KIND == QD_RULE_HADAMARD) {
            if (rule.bit >= sz) {
                atomicOr(err, ERR_BIT);
                nc[p] = 0;
                continue;
            }
            num = 2;
            ms = sz;
        } else {
            if (!qcgd_valid(o, sz)) {
                atomicOr(err, ERR_OBJECT);
                nc[p] = 0;
                continue;
            }
            Graph g(o);
            if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
                num = sm_num_child(g);
                ms = 4 * sz;
                sm_fill_table(g, tab + P.begin[p] / 2);
                pre[p] = sm_first_hmlz(g) ? 1 : 0;
            } else {
                num = flip_num_child(g, KIND == QD_RULE_QCGD_COIN);
                ms = sz;
                pre[p] = name_fold(g);
            }
        }
        nc[p] = num;
        local_max = max(local_max, ms);
    }
    if (local_max) atomicMax(maxsz, local_max);
}

// ===========================================================================
// phase 3: generate
// ===========================================================================
__device__ inline u64 find_parent(const u64 *cb, u64 n, u64 c) {
    // largest p in [0, n) with cb[p] <= c (cb has n+1 entries, cb[n] = C > c)
    u64 lo = 0, hi = n - 1;
    while (lo < hi) {
        const u64 mid = (lo + hi + 1) >> 1;
        if (cb[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

constexpr int kHistBins = 4096;  // up to 12 bits of LDS histogram

template <int KIND>
__global__ __launch_bounds__(1024) void k_generate(Parents P, const u64 *cb, qd_rule rule, const u64 *pre,
                                                  const u64 *tab, u64 C, Entry *E, u32 b1, u64 *hist1) {
    __shared__ u32 hist[kHistBins];
    if (b1)
        for (int i = threadIdx.x; i < (1 << b1); i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const QcgdCoef k = coef_of(rule);
    for (u64 c = u64(blockIdx.x) * blockDim.x + threadIdx.x; c < C; c += u64(gridDim.x) * blockDim.x) {
        const u64 p = find_parent(cb, P.n, c);
        const u32 j = u32(c - cb[p]);
        const double2 m0 = P.mag[p];
        cplx mag{m0.x, m0.y};
        const u8 *o = P.obj + P.begin[p];
        u64 h;
        if (KIND == QD_RULE_HADAMARD) {
            h = hadamard_child_hash(o, P.size[p], rule.bit, rule.coef, j, mag);
        } else if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
            Graph g(o);
            h = sm_child_hash(g, pre[p] != 0, tab + P.begin[p] / 2, j, k, mag);
        } else {
            Graph g(o);
            h = flip_child_hash(g, KIND == QD_RULE_QCGD_COIN, pre[p], j, k, mag);
        }
        Entry e;
        e.hash = h;
        e.cand = c;
        e.re = mag.re;
        e.im = mag.im;
        E[c] = e;
        if (b1) atomicAdd(&hist[h >> (64 - b1)], 1u);
    }
    if (b1) {
        __syncthreads();
        for (int i = threadIdx.x; i < (1 << b1); i += blockDim.x)
            if (hist[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist1 + i), (unsigned long long)hist[i]);
    }
}

// ===========================================================================
// phase 4: radix partition
// ===========================================================================
constexpr int kTile1 = 4096;  // entries per workgroup in the first pass

__global__ __launch_bounds__(256) void k_scatter1(const Entry *in, u64 n, u32 b1, u64 *cursor, Entry *out) {
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
            const u32 d = u32(in[i].hash >> (64 - b1));
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
        if (i < n) out[gbase[dr[k] >> 16] + (dr[k] & 0xFFFF)] = in[i];
    }
}

// Second pass: one workgroup per first-level bucket splits it by the next b2
// bits; counts are workgroup-local, so placement needs no global atomics.
__global__ __launch_bounds__(1024) void k_level2(const Entry *in, const u64 *off1, u32 b1, u32 b2, Entry *out,
                                                 u64 *off2) {
    __shared__ u32 cnt[kHistBins];
    __shared__ u32 ccnt[kHistBins];
    __shared__ u64 cur[kHistBins];
    __shared__ u64 tmp[1024 / kWave + 1];
    const u64 b = blockIdx.x;
    const u64 beg = off1[b], end = off1[b + 1];
    const u32 nb2 = 1u << b2;
    const u32 shift = 64 - b1 - b2;
    for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    for (u64 i = beg + threadIdx.x; i < end; i += blockDim.x)
        atomicAdd(&cnt[u32(in[i].hash >> shift) & (nb2 - 1)], 1u);
    __syncthreads();
    // exclusive scan of cnt (nb2 <= 4096 entries, 4 per thread)
    {
        u32 v[4];
        u64 s = 0;
        for (int k = 0; k < 4; ++k) {
            const u32 d = threadIdx.x * 4 + k;
            v[k] = d < nb2 ? cnt[d] : 0;
            s += v[k];
        }
        u64 total;
        u64 run = block_excl_scan(s, tmp, total);
        for (int k = 0; k < 4; ++k) {
            const u32 d = threadIdx.x * 4 + k;
            if (d < nb2) {
                cur[d] = beg + run;
                off2[(b << b2) + d] = beg + run;
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
                const u32 d = u32(in[i].hash >> shift) & (nb2 - 1);
                dr[k] = (d << 16) | atomicAdd(&ccnt[d], 1u);
            }
        }
        __syncthreads();
        for (int k = 0; k < 4; ++k) {
            const u64 i = cbase + threadIdx.x + 1024 * k;
            if (i < end) out[cur[dr[k] >> 16] + (dr[k] & 0xFFFF)] = in[i];
        }
        __syncthreads();
        for (u32 i = threadIdx.x; i < nb2; i += blockDim.x) cur[i] += ccnt[i];
        __syncthreads();
    }
}

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
template <int T, int CAP, int TS>
__global__ __launch_bounds__(T) void k_dedup(const Entry *E, const u64 *boff, u64 nbk, u32 L, double tol, u64 *UH,
                                             u64 *UR, double2 *UM, u64 *counters, u64 *nhist, u64 *biglist) {
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
        nbeg = boff[bb];
        nnum = boff[bb + 1] - nbeg;
        if (nnum <= CAP) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const u32 i = threadIdx.x + T * k;
                if (i < nnum) x[k] = E[nbeg + i];
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
__global__ __launch_bounds__(1024) void k_dedup_big(const Entry *E, const u64 *boff, const u64 *biglist,
                                                    const u64 *soff, u64 *skeyg, u64 *scandg, double2 *smagg,
                                                    u32 *segg, double tol, u64 *UH, u64 *UR, double2 *UM,
                                                    u64 *counters, u64 *nhist) {
    __shared__ u32 maxg;
    __shared__ u64 tmp[1024 / kWave + 1];
    const u64 b = biglist[blockIdx.x];
    const u64 beg = boff[b], n = boff[b + 1] - beg;
    u64 *skey = skeyg + soff[blockIdx.x];
    u64 *scand = scandg + soff[blockIdx.x];
    double2 *smag = smagg + soff[blockIdx.x];
    u32 *seg = segg + soff[blockIdx.x];
    u64 npad = 2;
    while (npad < n) npad <<= 1;
    for (u64 i = threadIdx.x; i < npad; i += blockDim.x) {
        if (i < n) {
            skey[i] = E[beg + i].hash;
            scand[i] = E[beg + i].cand;
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
        const Entry e = E[beg + i];
        u64 lo = 0, hi = n - 1;
        while (lo < hi) {
            const u64 mid = (lo + hi) >> 1;
            if (key_gt(e.hash, e.cand, skey[mid], scand[mid])) lo = mid + 1;
            else hi = mid;
        }
        smag[lo] = make_double2(e.re, e.im);
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

// ===========================================================================
// phase 6: top-K select on the 128-bit key (|m|^2 bits, ~hash)
// ===========================================================================
struct SelItem {
    u64 hi, lo, rep;
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
    atomicOr(reinterpret_cast<unsigned long long *>(bitmap + (rep >> 6)), 1ull << (rep & 63));
}

// First select stage over the dedup output slots: histogram of |m|^2 bits
// 51..40 among the uniques whose bits 63..52 equal d1.
__global__ __launch_bounds__(256) void k_sel_hist_u(const u64 *UR, const double2 *UM, u64 U, u32 d1, u64 *hist) {
    __shared__ u32 h[kHistBins];
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < U; i += u64(gridDim.x) * blockDim.x) {
        if (UR[i] == kNoRep) continue;
        const double2 m = UM[i];
        const u64 bits = u64(__double_as_longlong(m.x * m.x + m.y * m.y));
        if (u32(bits >> 52) == d1) atomicAdd(&h[u32(bits >> 40) & 0xFFF], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
        if (h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist + i), (unsigned long long)h[i]);
}

// Marks the uniques whose top 24 |m|^2 bits exceed prefix24 and collects
// the ones equal to it.
__global__ __launch_bounds__(256) void k_sel_first(const u64 *UH, const u64 *UR, const double2 *UM, u64 U,
                                                   u32 prefix24, SelItem *S, u64 *scount, u64 *bitmap) {
    const u64 stride = u64(gridDim.x) * blockDim.x;
    // uniform trip count per wave so that wave_append sees every lane
    for (u64 base = u64(blockIdx.x) * blockDim.x; base < U; base += stride) {
        const u64 i = base + threadIdx.x;
        bool take = false;
        SelItem it{};
        if (i < U && UR[i] != kNoRep) {
            const double2 m = UM[i];
            const u64 bits = u64(__double_as_longlong(m.x * m.x + m.y * m.y));
            const u32 t = u32(bits >> 40);
            if (t > prefix24) mark(bitmap, UR[i]);
            else if (t == prefix24) {
                take = true;
                it = SelItem{bits, ~UH[i], UR[i]};
            }
        }
        const u64 at = wave_append(take, scount);
        if (take) S[at] = it;
    }
}

__global__ __launch_bounds__(256) void k_sel_hist(const SelItem *S, u64 n, u32 pos, u64 *hist) {
    __shared__ u32 h[kHistBins];
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
        atomicAdd(&h[key_digit(S[i].hi, S[i].lo, pos)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
        if (h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist + i), (unsigned long long)h[i]);
}

__global__ __launch_bounds__(256) void k_sel_split(const SelItem *S, u64 n, u32 pos, u32 d, SelItem *out,
                                                   u64 *ocount, u64 *bitmap) {
    const u64 stride = u64(gridDim.x) * blockDim.x;
    for (u64 base = u64(blockIdx.x) * blockDim.x; base < n; base += stride) {
        const u64 i = base + threadIdx.x;
        bool take = false;
        SelItem it{};
        if (i < n) {
            it = S[i];
            const u32 t = key_digit(it.hi, it.lo, pos);
            if (t > d) mark(bitmap, it.rep);
            else take = t == d;
        }
        const u64 at = wave_append(take, ocount);
        if (take) out[at] = it;
    }
}

__global__ __launch_bounds__(256) void k_sel_mark_items(const SelItem *S, u64 n, u64 *bitmap) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
        mark(bitmap, S[i].rep);
}

__global__ __launch_bounds__(256) void k_mark_all(const u64 *UR, u64 U, u64 *bitmap) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < U; i += u64(gridDim.x) * blockDim.x)
        if (UR[i] != kNoRep) mark(bitmap, UR[i]);
}

// ===========================================================================
// phase 7/8: order, sizes, populate
// ===========================================================================
__global__ __launch_bounds__(256) void k_place(const u64 *UR, const double2 *UM, u64 U, const u64 *bitmap,
                                               const u64 *wprefix, u64 *orep, double2 *omag) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < U; i += u64(gridDim.x) * blockDim.x) {
        const u64 rep = UR[i];
        if (rep == kNoRep) continue;
        const u64 word = bitmap[rep >> 6];
        const u32 bit = u32(rep & 63);
        if ((word >> bit) & 1) {
            const u64 r = wprefix[rep >> 6] + __popcll(word & ((1ull << bit) - 1));
            orep[r] = rep;
            omag[r] = UM[i];
        }
    }
}

__device__ inline u32 align_pad(u32 size, u32 align) {
    if (align <= 1) return 0;
    const u32 off = align - size % align;
    return off == align ? 0 : off;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_child_size(Parents P, const u64 *cb, qd_rule rule, const u64 *orep, u64 K,
                                                    u32 align, u32 *osize, u64 *padded) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x) {
        const u64 c = orep[k];
        const u64 p = find_parent(cb, P.n, c);
        const u32 j = u32(c - cb[p]);
        const u8 *o = P.obj + P.begin[p];
        u32 s;
        if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
            Graph g(o);
            s = sm_child_size(g, sm_first_hmlz(g), j);
        } else {
            s = P.size[p];
        }
        osize[k] = s;
        padded[k] = u64(s) + align_pad(s, align);
    }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_populate(Parents P, const u64 *cb, qd_rule rule, const u64 *orep, u64 K,
                                                  const u64 *obegin, const u32 *osize, u8 *oobj) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x) {
        const u64 c = orep[k];
        const u64 p = find_parent(cb, P.n, c);
        const u32 j = u32(c - cb[p]);
        const u8 *o = P.obj + P.begin[p];
        u8 *d = oobj + obegin[k];
        u32 s;
        if (KIND == QD_RULE_HADAMARD) {
            hadamard_child_write(o, P.size[p], rule.bit, j, d);
            s = P.size[p];
        } else if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
            Graph g(o);
            s = sm_child_write(g, sm_first_hmlz(g), j, d);
        } else {
            flip_child_write(o, P.size[p], KIND == QD_RULE_QCGD_COIN, j, d);
            s = P.size[p];
        }
        const u64 e = obegin[k + 1];
        for (u64 q = obegin[k] + s; q < e; ++q) oobj[q] = 0;
        (void)osize;
    }
}

// ===========================================================================
// phase 9: canonical sum of |m|^2 and scaling
// ===========================================================================
constexpr int kSumThreads = 256;
constexpr int kSumChunk = 2048;

// One block per chunk of 2048 values: lane t sums t, t+256, ..., t+1792 in
// order, then a halving tree s[t] += s[t+stride], stride = 128 .. 1.
__global__ __launch_bounds__(kSumThreads) void k_sum_chunks(const double2 *mag, const double *vals, u64 n,
                                                            double *out) {
    __shared__ double s[kSumThreads];
    const u64 base = u64(blockIdx.x) * kSumChunk;
    double a = 0.0;
    for (int j = 0; j < kSumChunk / kSumThreads; ++j) {
        const u64 i = base + threadIdx.x + u64(kSumThreads) * j;
        if (i < n) {
            double v;
            if (mag) {
                const double2 m = mag[i];
                v = m.x * m.x + m.y * m.y;
            } else {
                v = vals[i];
            }
            a += v;
        }
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int stride = kSumThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_scale(double2 *mag, u64 n, double f) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        double2 m = mag[i];
        m.x = m.x / f;
        m.y = m.y / f;
        mag[i] = m;
    }
}

// ===========================================================================
// modifiers
// ===========================================================================
__global__ __launch_bounds__(256) void k_modifier(u8 *obj, const u64 *begin, const u32 *size, double2 *mag, u64 n,
                                                  qd_modifier m, u32 *err) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        u8 *o = obj + begin[i];
        const u32 sz = size[i];
        if (m.kind <= QD_MOD_ZGATE) {
            if (m.bit >= sz || (m.kind == QD_MOD_CNOT && m.control_bit >= sz)) {
                atomicOr(err, ERR_BIT);
                continue;
            }
        } else if (!qcgd_valid(o, sz)) {
            atomicOr(err, ERR_OBJECT);
            continue;
        }
        double2 mm = mag[i];
        cplx c{mm.x, mm.y};
        modifier_apply(m, o, sz, c);
        mag[i] = make_double2(c.re, c.im);
    }
}

// ===========================================================================
// host orchestration
// ===========================================================================
static inline unsigned grid_for(u64 n, unsigned block = 256, unsigned cap = 65536) {
    const u64 g = (n + block - 1) / block;
    return unsigned(std::max<u64>(1, std::min<u64>(g, cap)));
}

static void prof_collect(qd_ctx *ctx);
static int sync_read(qd_ctx *ctx, const void *dev, size_t bytes, void *host) {
    QD_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    QD_HIP(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    return QD_OK;
}

static void prof_collect(qd_ctx *ctx);
static void prof_begin(qd_ctx *ctx, int ph) {
    if (!ctx->profiling) return;
    if (ctx->pending[ph]) {
        (void)hipEventSynchronize(ctx->ev[2 * ph + 1]);
        prof_collect(ctx);
    }
    (void)hipEventRecord(ctx->ev[2 * ph], ctx->stream);
}
static void prof_end(qd_ctx *ctx, int ph) {
    if (!ctx->profiling) return;
    (void)hipEventRecord(ctx->ev[2 * ph + 1], ctx->stream);
    ctx->pending[ph] = true;
}
// Folds recorded phase intervals into the totals (after a stream sync).
static void prof_collect(qd_ctx *ctx) {
    if (!ctx->profiling) return;
    for (int ph = 0; ph < QD_NUM_PHASES; ++ph) {
        if (!ctx->pending[ph]) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, ctx->ev[2 * ph], ctx->ev[2 * ph + 1]) == hipSuccess) {
            ctx->phase_ms[ph] += ms;
            ctx->phase_calls[ph] += 1;
        }
        ctx->pending[ph] = false;
    }
}

static int check_err_word(u32 w) {
    if (w & ERR_BIT) return fail(QD_ERR_INVALID, "rule or modifier bit index beyond an object's size");
    if (w & ERR_OBJECT) return fail(QD_ERR_INVALID, "malformed QCGD object (size != 4 + 4n + 16m, or n == 0)");
    if (w & ERR_OVERFLOW) return fail(QD_ERR_CAPACITY, "a dedup bucket exceeded its LDS capacity");
    return QD_OK;
}

// canonical sum of |m|^2 (mag != null) or of vals
static int canonical_sum(qd_ctx *ctx, const double2 *mag, u64 n, double *result) {
    if (n == 0) {
        *result = 0.0;
        return QD_OK;
    }
    u64 nchunk = (n + kSumChunk - 1) / kSumChunk;
    // two ping-pong halves of the chunk buffer
    u64 cap = nchunk + (nchunk + kSumChunk - 1) / kSumChunk + 8;
    QD_TRY(ctx->chunks.ensure(2 * cap * sizeof(double)));
    double *a = ctx->chunks.as<double>(), *b = a + cap;
    prof_begin(ctx, QD_PHASE_NORMALIZE);
    hipLaunchKernelGGL(k_sum_chunks, dim3(nchunk), dim3(kSumThreads), 0, ctx->stream, mag, (const double *)nullptr, n,
                       a);
    u64 m = nchunk;
    while (m > 1) {
        const u64 nc2 = (m + kSumChunk - 1) / kSumChunk;
        hipLaunchKernelGGL(k_sum_chunks, dim3(nc2), dim3(kSumThreads), 0, ctx->stream, (const double2 *)nullptr,
                           (const double *)a, m, b);
        std::swap(a, b);
        m = nc2;
    }
    QD_HIP(hipGetLastError());
    prof_end(ctx, QD_PHASE_NORMALIZE);
    QD_TRY(sync_read(ctx, a, sizeof(double), result));
    return QD_OK;
}

static int normalize_state(qd_ctx *ctx, qd_state *s) {
    double total = 0;
    QD_TRY(canonical_sum(ctx, s->mag.as<double2>(), s->n, &total));
    s->total_proba = total;
    if (s->n == 0) {
        s->total_proba = 0;
        return QD_OK;
    }
    const double f = std::sqrt(total);
    if (f != 1) {
        prof_begin(ctx, QD_PHASE_NORMALIZE);
        hipLaunchKernelGGL(k_scale, dim3(grid_for(s->n)), dim3(256), 0, ctx->stream, s->mag.as<double2>(), s->n, f);
        QD_HIP(hipGetLastError());
        prof_end(ctx, QD_PHASE_NORMALIZE);
    }
    return QD_OK;
}


__global__ __launch_bounds__(256) void k_parent_items(const double2 *mag, u64 n, SelItem *S) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        const double2 m = mag[i];
        const double nrm = m.x * m.x + m.y * m.y;
        S[i] = SelItem{u64(__double_as_longlong(nrm)), ~i, i};
    }
}

// Selects the r largest 128-bit keys among the ns items of S0, whose
// leading `pos` key bits are already known to tie; marks their rep bits in
// bitmap.  Items above the chosen digit are marked as soon as it is known;
// the tied ones are carried to the next 12-bit digit.
static int select_loop(qd_ctx *ctx, SelItem *S0, SelItem *S1, u64 ns, u64 r, u32 pos, u64 *bitmap) {
    hipStream_t s = ctx->stream;
    u64 *hist = ctx->hist1.as<u64>();
    u64 *cnt = ctx->counters.as<u64>() + 2;
    std::vector<u64> h(kHistBins);
    while (r > 0 && r < ns) {
        if (pos >= 128) return fail(QD_ERR_HIP, "select did not converge (duplicate keys)");
        QD_HIP(hipMemsetAsync(hist, 0, kHistBins * sizeof(u64), s));
        QD_HIP(hipMemsetAsync(cnt, 0, sizeof(u64), s));
        hipLaunchKernelGGL(k_sel_hist, dim3(grid_for(ns, 256, 4096)), dim3(256), 0, s, S0, ns, pos, hist);
        QD_HIP(hipGetLastError());
        QD_TRY(sync_read(ctx, hist, kHistBins * sizeof(u64), h.data()));
        u64 cum = 0;
        int d = kHistBins - 1;
        for (; d >= 0; --d) {
            if (cum + h[d] >= r) break;
            cum += h[d];
        }
        r -= cum;
        hipLaunchKernelGGL(k_sel_split, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, s, S0, ns, pos, u32(d), S1, cnt,
                           bitmap);
        QD_HIP(hipGetLastError());
        ns = h[d];
        std::swap(S0, S1);
        pos += 12;
    }
    if (r > 0 && r == ns)
        hipLaunchKernelGGL(k_sel_mark_items, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, s, S0, ns, bitmap);
    QD_HIP(hipGetLastError());
    return QD_OK;
}


// Buckets beyond the LDS capacity: one workgroup each on global scratch.
static int dedup_big_buckets(qd_ctx *ctx, const Entry *D, const u64 *boff, u64 nbig, double tol, u64 *UH, u64 *UR,
                             double2 *UM, u64 *U) {
    hipStream_t s = ctx->stream;
    std::vector<u64> list(nbig), soff(nbig + 1, 0);
    QD_TRY(sync_read(ctx, ctx->biglist.p, nbig * sizeof(u64), list.data()));
    for (u64 q = 0; q < nbig; ++q) {
        u64 be[2];
        QD_TRY(sync_read(ctx, boff + list[q], 2 * sizeof(u64), be));
        u64 npad = 2;
        while (npad < be[1] - be[0]) npad <<= 1;
        soff[q + 1] = soff[q] + npad;
    }
    const u64 tot = soff[nbig];
    QD_TRY(ctx->bigsoff.ensure((nbig + 1) * sizeof(u64)));
    QD_TRY(ctx->bigkey.ensure(tot * sizeof(u64)));
    QD_TRY(ctx->bigcand.ensure(tot * sizeof(u64)));
    QD_TRY(ctx->bigmag.ensure(tot * sizeof(double2)));
    QD_TRY(ctx->bigseg.ensure(tot * sizeof(u32)));
    QD_HIP(hipMemcpyAsync(ctx->bigsoff.p, soff.data(), (nbig + 1) * sizeof(u64), hipMemcpyHostToDevice, s));
    u64 *counters = ctx->counters.as<u64>();
    hipLaunchKernelGGL(k_dedup_big, dim3(nbig), dim3(1024), 0, s, D, boff, ctx->biglist.as<u64>(),
                       ctx->bigsoff.as<u64>(), ctx->bigkey.as<u64>(), ctx->bigcand.as<u64>(), ctx->bigmag.as<double2>(),
                       ctx->bigseg.as<u32>(), tol, UH, UR, UM, counters, ctx->nhist.as<u64>());
    QD_HIP(hipGetLastError());
    QD_TRY(sync_read(ctx, counters, sizeof(u64), U));
    return QD_OK;
}

static int state_reserve(qd_state *s, u64 n, u64 bytes) {
    QD_TRY(s->objects.ensure(std::max<u64>(bytes, 1)));
    QD_TRY(s->begin.ensure((n + 1) * sizeof(u64)));
    QD_TRY(s->size.ensure(std::max<u64>(n, 1) * sizeof(u32)));
    QD_TRY(s->mag.ensure(std::max<u64>(n, 1) * sizeof(double2)));
    return QD_OK;
}

// Dedup layout for C candidates: total bucket bits L so that buckets
// average <= 1536 entries (half the LDS capacity, so a bucket overflows only
// under a pathological hash distribution); first pass up to 12 bits, second
// pass the rest.
static void dedup_layout(u64 C, const qd_tuning &t, u32 &b1, u32 &b2) {
    u32 L = 0;
    while ((C >> L) > t.dedup_bucket_target && L < 24) ++L;
    b1 = std::min<u32>(L, t.first_pass_bits);
    if (L - b1 > 12) b1 = L - 12;
    b2 = L - b1;
}

template <int KIND>
static int simulate_kind(qd_ctx *ctx, qd_state *in, const qd_rule &rule, qd_state *out, qd_symbolic *sy, u64 max_obj,
                         const qd_params &prm, qd_step_stats &st) {
    hipStream_t s = ctx->stream;
    const u64 N = in->n;
    st.num_parent = N;
    Parents P{in->objects.as<u8>(), in->begin.as<u64>(), in->size.as<u32>(), in->mag.as<double2>(), N};

    auto finish_empty = [&]() -> int {
        QD_TRY(state_reserve(out, 0, 0));
        QD_HIP(hipMemsetAsync(out->begin.p, 0, sizeof(u64), s));
        out->n = 0;
        out->bytes = 0;
        out->total_proba = 0;
        st.total_proba = 0;
        if (sy) {
            sy->num_object = st.num_candidates;
            sy->num_object_after_interferences = st.num_unique;
        }
        QD_HIP(hipStreamSynchronize(s));
        return QD_OK;
    };
    if (N == 0) return finish_empty();

    // ---- pre-truncation of parents (quids.hpp:613-642): top max by |m|^2,
    // ties by ascending index, through the 128-bit select on (|m|^2, ~index).
    const u64 *keep = nullptr;
    u64 nkept = N;
    if (max_obj != 0 && max_obj < N) {
        if (!prm.simple_truncation)
            return fail(QD_ERR_UNSUPPORTED, "probabilistic truncation is not implemented; set simple_truncation");
        QD_TRY(ctx->sel0.ensure(N * sizeof(SelItem)));
        QD_TRY(ctx->sel1.ensure(N * sizeof(SelItem)));
        QD_TRY(ctx->kept.ensure(((N + 63) / 64) * sizeof(u64)));
        QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
        QD_TRY(ctx->hist1.ensure(kHistBins * sizeof(u64)));
        u64 *pbits = ctx->kept.as<u64>();
        QD_HIP(hipMemsetAsync(pbits, 0, ((N + 63) / 64) * sizeof(u64), s));
        hipLaunchKernelGGL(k_parent_items, dim3(grid_for(N)), dim3(256), 0, s, P.mag, N, ctx->sel0.as<SelItem>());
        QD_HIP(hipGetLastError());
        SelItem *S0 = ctx->sel0.as<SelItem>(), *S1 = ctx->sel1.as<SelItem>();
        QD_TRY(select_loop(ctx, S0, S1, N, max_obj, 0, pbits));
        keep = pbits;
        nkept = max_obj;
    }
    st.num_parent_kept = nkept;

    // ---- 1 prep
    QD_TRY(ctx->nc.ensure(N * sizeof(u32)));
    QD_TRY(ctx->cb.ensure((N + 1) * sizeof(u64)));
    QD_TRY(ctx->pre.ensure(N * sizeof(u64)));
    QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
    if (KIND == QD_RULE_QCGD_SPLIT_MERGE) QD_TRY(ctx->tab.ensure((in->bytes / 2 + 16) * sizeof(u64)));
    u64 *counters = ctx->counters.as<u64>();
    // counters: [0] unique count, [1] select count A, [2] select count B,
    // [3] low: max child size, [4] low: error word
    QD_HIP(hipMemsetAsync(counters, 0, 64 * sizeof(u64), s));
    u32 *maxsz = reinterpret_cast<u32 *>(counters + 3);
    u32 *errw = reinterpret_cast<u32 *>(counters + 4);
    prof_begin(ctx, QD_PHASE_PREP);
    hipLaunchKernelGGL(k_prep<KIND>, dim3(grid_for(N)), dim3(256), 0, s, P, keep, rule, ctx->nc.as<u32>(),
                       ctx->pre.as<u64>(), ctx->tab.as<u64>(), maxsz, errw);
    QD_HIP(hipGetLastError());

    // ---- 2 scan -> child_begin, C
    prof_end(ctx, QD_PHASE_PREP);
    prof_begin(ctx, QD_PHASE_SCAN);
    QD_TRY(run_scan(ctx, LoadU32{ctx->nc.as<u32>()}, N, ctx->cb.as<u64>()));
    prof_end(ctx, QD_PHASE_SCAN);
    u64 *hb = ctx->host;
    QD_HIP(hipMemcpyAsync(hb, ctx->cb.as<u64>() + N, sizeof(u64), hipMemcpyDeviceToHost, s));
    QD_HIP(hipMemcpyAsync(hb + 1, counters + 3, 2 * sizeof(u64), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    const u64 C = hb[0];
    QD_TRY(check_err_word(u32(hb[2])));
    st.num_candidates = C;
    if (C == 0) return finish_empty();

    // ---- 3 generate
    u32 b1, b2;
    dedup_layout(C, ctx->tune, b1, b2);
    st.dedup_levels = (b1 ? 1 : 0) + (b2 ? 1 : 0);
    QD_TRY(ctx->ebuf0.ensure(C * sizeof(Entry)));
    QD_TRY(ctx->ebuf1.ensure(C * sizeof(Entry)));
    QD_TRY(ctx->hist1.ensure(kHistBins * sizeof(u64)));
    QD_TRY(ctx->off1.ensure((kHistBins + 1) * sizeof(u64)));
    QD_TRY(ctx->cur1.ensure(kHistBins * sizeof(u64)));
    QD_TRY(ctx->nhist.ensure(kHistBins * sizeof(u64)));
    Entry *E0 = ctx->ebuf0.as<Entry>(), *E1 = ctx->ebuf1.as<Entry>();
    u64 *hist1 = ctx->hist1.as<u64>();
    QD_HIP(hipMemsetAsync(hist1, 0, kHistBins * sizeof(u64), s));
    QD_HIP(hipMemsetAsync(ctx->nhist.p, 0, kHistBins * sizeof(u64), s));
    prof_begin(ctx, QD_PHASE_GENERATE);
    {
        const unsigned gb = ctx->tune.generate_block;
        const unsigned gg = grid_for(C, gb, ctx->tune.generate_grid);
        hipLaunchKernelGGL(k_generate<KIND>, dim3(gg), dim3(gb), 0, s, P, ctx->cb.as<u64>(), rule, ctx->pre.as<u64>(),
                           ctx->tab.as<u64>(), C, E0, b1, hist1);
    }
    QD_HIP(hipGetLastError());

    prof_end(ctx, QD_PHASE_GENERATE);
    // ---- 4 partition
    const Entry *D = E0;         // entries the dedup reads
    const u64 *boff = nullptr;   // bucket offsets (nbk + 1)
    u64 nbk = 1;
    void *ubuf = ctx->ebuf1.p;   // where the uniques go (the buffer D does not read)
    if (b1 == 0) {
        QD_TRY(ctx->off1.ensure(2 * sizeof(u64)));
        u64 two[2] = {0, C};
        QD_HIP(hipMemcpyAsync(ctx->off1.p, two, sizeof two, hipMemcpyHostToDevice, s));
        QD_HIP(hipStreamSynchronize(s));
        boff = ctx->off1.as<u64>();
    } else {
        const u64 nb1 = 1ull << b1;
        prof_begin(ctx, QD_PHASE_PARTITION);
        QD_TRY(run_scan(ctx, LoadU64{hist1}, nb1, ctx->off1.as<u64>()));
        QD_HIP(hipMemcpyAsync(ctx->cur1.p, ctx->off1.p, nb1 * sizeof(u64), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_scatter1, dim3((C + kTile1 - 1) / kTile1), dim3(256), 0, s, (const Entry *)E0, C, b1,
                           ctx->cur1.as<u64>(), E1);
        QD_HIP(hipGetLastError());
        if (b2 == 0) {
            D = E1;
            boff = ctx->off1.as<u64>();
            nbk = nb1;
            ubuf = ctx->ebuf0.p;
        } else {
            const u64 nb = 1ull << (b1 + b2);
            QD_TRY(ctx->off2.ensure((nb + 1) * sizeof(u64)));
            QD_HIP(hipMemcpyAsync(ctx->off2.as<u64>() + nb, ctx->cb.as<u64>() + N, sizeof(u64),
                                  hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(k_level2, dim3(nb1), dim3(1024), 0, s, (const Entry *)E1, ctx->off1.as<u64>(), b1, b2,
                               E0, ctx->off2.as<u64>());
            QD_HIP(hipGetLastError());
            D = E0;
            boff = ctx->off2.as<u64>();
            nbk = nb;
            ubuf = ctx->ebuf1.p;
        }
    }
    if (b1) prof_end(ctx, QD_PHASE_PARTITION);
    const u32 L = b1 + b2;

    // ---- 5 dedup
    u64 *UH = static_cast<u64 *>(ubuf);
    u64 *UR = UH + C;
    double2 *UM = reinterpret_cast<double2 *>(UR + C);
    prof_begin(ctx, QD_PHASE_DEDUP);
    QD_TRY(ctx->biglist.ensure((C / 2048 + 2) * sizeof(u64)));
    {
        int ncu = 256;
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device);
        const u64 grid = std::min<u64>(nbk, u64(ncu));
        hipLaunchKernelGGL((k_dedup<1024, 2048, 4096>), dim3(grid), dim3(1024), 0, s, D, boff, nbk, L, prm.tolerance, UH,
                           UR, UM, counters, ctx->nhist.as<u64>(), ctx->biglist.as<u64>());
        QD_HIP(hipGetLastError());
    }
    prof_end(ctx, QD_PHASE_DEDUP);
    QD_HIP(hipMemcpyAsync(hb, counters, 6 * sizeof(u64), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    prof_collect(ctx);
    QD_TRY(check_err_word(u32(hb[4])));
    u64 U = hb[0];
    if (hb[5]) QD_TRY(dedup_big_buckets(ctx, D, boff, hb[5], prm.tolerance, UH, UR, UM, &U));
    st.num_unique = U;

    // memory-driven cap (max_num_object == 0, quids.hpp:459-486, 510-536):
    // keep as many survivors as the free device memory holds for the next
    // state, at the largest child size seen, with a 20% safety margin.
    if (max_obj == 0) {
        size_t freeb = 0, totalb = 0;
        QD_HIP(hipMemGetInfo(&freeb, &totalb));
        const u64 per = u64(u32(hb[3])) + 8 + sizeof(u64) + sizeof(u32) + sizeof(double2);
        max_obj = std::max<u64>(1, u64(double(freeb) * 0.8) / per);
    }

    // ---- 6 select + 7 order
    const u64 W = (C + 63) / 64;
    QD_TRY(ctx->bitmap.ensure(W * sizeof(u64)));
    QD_TRY(ctx->wprefix.ensure((W + 1) * sizeof(u64)));
    u64 *bitmap = ctx->bitmap.as<u64>();
    prof_begin(ctx, QD_PHASE_SELECT);
    QD_HIP(hipMemsetAsync(bitmap, 0, W * sizeof(u64), s));
    if (U <= max_obj) {
        if (U) hipLaunchKernelGGL(k_mark_all, dim3(grid_for(C, 256, 16384)), dim3(256), 0, s, UR, C, bitmap);
        QD_HIP(hipGetLastError());
    } else {
        if (!prm.simple_truncation)
            return fail(QD_ERR_UNSUPPORTED, "probabilistic truncation is not implemented; set simple_truncation");
        std::vector<u64> h(kHistBins);
        QD_TRY(sync_read(ctx, ctx->nhist.p, kHistBins * sizeof(u64), h.data()));
        u64 r = max_obj, cum = 0;
        int d = kHistBins - 1;
        for (; d >= 0; --d) {
            if (cum + h[d] >= r) break;
            cum += h[d];
        }
        r -= cum;
        // second 12 bits inside bin d
        const u32 d1 = u32(d);
        QD_HIP(hipMemsetAsync(hist1, 0, kHistBins * sizeof(u64), s));
        hipLaunchKernelGGL(k_sel_hist_u, dim3(grid_for(C, 256, 8192)), dim3(256), 0, s, UR, UM, C, d1, hist1);
        QD_HIP(hipGetLastError());
        QD_TRY(sync_read(ctx, hist1, kHistBins * sizeof(u64), h.data()));
        cum = 0;
        for (d = kHistBins - 1; d >= 0; --d) {
            if (cum + h[d] >= r) break;
            cum += h[d];
        }
        r -= cum;
        u64 ns = h[d];
        QD_TRY(ctx->sel0.ensure(std::max<u64>(ns, 1) * sizeof(SelItem)));
        QD_TRY(ctx->sel1.ensure(std::max<u64>(ns, 1) * sizeof(SelItem)));
        SelItem *S0 = ctx->sel0.as<SelItem>(), *S1 = ctx->sel1.as<SelItem>();
        hipLaunchKernelGGL(k_sel_first, dim3(grid_for(C, 256, 16384)), dim3(256), 0, s, UH, UR, UM, C,
                           (d1 << 12) | u32(d), S0, counters + 1, bitmap);
        QD_HIP(hipGetLastError());
        QD_TRY(select_loop(ctx, S0, S1, ns, r, 24, bitmap));
    }
    prof_end(ctx, QD_PHASE_SELECT);
    prof_begin(ctx, QD_PHASE_ORDER);
    QD_TRY(run_scan(ctx, LoadPopc{bitmap}, W, ctx->wprefix.as<u64>()));
    prof_end(ctx, QD_PHASE_ORDER);
    QD_TRY(sync_read(ctx, ctx->wprefix.as<u64>() + W, sizeof(u64), hb));
    const u64 K = hb[0];
    st.num_survivors = K;
    if (K == 0) return finish_empty();

    // ---- 8 finalize
    QD_TRY(out->begin.ensure((K + 1) * sizeof(u64)));
    QD_TRY(out->size.ensure(K * sizeof(u32)));
    QD_TRY(out->mag.ensure(K * sizeof(double2)));
    QD_TRY(ctx->padded.ensure(K * sizeof(u64)));
    QD_TRY(ctx->keys.ensure(K * sizeof(u64)));
    u64 *orep = ctx->keys.as<u64>();
    prof_begin(ctx, QD_PHASE_ORDER);
    hipLaunchKernelGGL(k_place, dim3(grid_for(C, 256, 16384)), dim3(256), 0, s, UR, UM, C, bitmap,
                       ctx->wprefix.as<u64>(), orep, out->mag.as<double2>());
    prof_end(ctx, QD_PHASE_ORDER);
    prof_begin(ctx, QD_PHASE_FINALIZE);
    hipLaunchKernelGGL(k_child_size<KIND>, dim3(grid_for(K)), dim3(256), 0, s, P, ctx->cb.as<u64>(), rule, orep, K,
                       prm.align_byte_length, out->size.as<u32>(), ctx->padded.as<u64>());
    QD_HIP(hipGetLastError());
    QD_TRY(run_scan(ctx, LoadU64{ctx->padded.as<u64>()}, K, out->begin.as<u64>()));
    prof_end(ctx, QD_PHASE_FINALIZE);
    QD_TRY(sync_read(ctx, out->begin.as<u64>() + K, sizeof(u64), hb));
    const u64 bytes = hb[0];
    QD_TRY(out->objects.ensure(std::max<u64>(bytes, 1)));
    prof_begin(ctx, QD_PHASE_FINALIZE);
    hipLaunchKernelGGL(k_populate<KIND>, dim3(grid_for(K)), dim3(256), 0, s, P, ctx->cb.as<u64>(), rule, orep, K,
                       out->begin.as<u64>(), out->size.as<u32>(), out->objects.as<u8>());
    QD_HIP(hipGetLastError());
    out->n = K;
    out->bytes = bytes;
    st.output_bytes = bytes;

    prof_end(ctx, QD_PHASE_FINALIZE);
    // ---- 9 normalize
    QD_TRY(normalize_state(ctx, out));
    prof_collect(ctx);
    st.total_proba = out->total_proba;
    if (sy) {
        sy->num_object = C;
        sy->num_object_after_interferences = U;
    }
    return QD_OK;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int qd_abi_version(void) { return QD_ABI_VERSION; }

const char *qd_status_string(int status) {
    switch (status) {
    case QD_OK: return "ok";
    case QD_ERR_INVALID: return "invalid argument";
    case QD_ERR_OOM: return "out of device memory";
    case QD_ERR_HIP: return "HIP runtime error";
    case QD_ERR_CAPACITY: return "capacity exceeded";
    case QD_ERR_UNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

const char *qd_last_error(void) { return g_last_error.c_str(); }

int qd_ctx_create(qd_ctx **out, int device) {
    if (!out) return fail(QD_ERR_INVALID, "null ctx pointer");
    *out = nullptr;
    int ndev = 0;
    QD_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(QD_ERR_INVALID, "no such device");
    QD_HIP(hipSetDevice(device));
    qd_ctx *c = new (std::nothrow) qd_ctx;
    if (!c) return fail(QD_ERR_OOM, "host allocation failed");
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(QD_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    e = hipHostMalloc(reinterpret_cast<void **>(&c->host), 8192 * sizeof(u64), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(QD_ERR_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    *out = c;
    return QD_OK;
}

int qd_ctx_destroy(qd_ctx *c) {
    if (!c) return QD_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (DBuf *b : c->all()) b->release();
    if (c->host) (void)hipHostFree(c->host);
    for (int i = 0; i < 2 * QD_NUM_PHASES; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return QD_OK;
}

int qd_ctx_synchronize(qd_ctx *c) {
    if (!c) return fail(QD_ERR_INVALID, "null ctx");
    QD_HIP(hipSetDevice(c->device));
    QD_HIP(hipStreamSynchronize(c->stream));
    return QD_OK;
}

int qd_ctx_scratch_bytes(qd_ctx *c, uint64_t *bytes) {
    if (!c || !bytes) return fail(QD_ERR_INVALID, "null argument");
    u64 t = 0;
    for (DBuf *b : c->all()) t += b->cap;
    *bytes = t;
    return QD_OK;
}

int qd_ctx_trim(qd_ctx *c) {
    if (!c) return fail(QD_ERR_INVALID, "null ctx");
    QD_HIP(hipSetDevice(c->device));
    QD_HIP(hipStreamSynchronize(c->stream));
    for (DBuf *b : c->all()) b->release();
    return QD_OK;
}

int qd_ctx_get_tuning(qd_ctx *c, qd_tuning *t) {
    if (!c || !t) return fail(QD_ERR_INVALID, "null argument");
    *t = c->tune;
    return QD_OK;
}

int qd_ctx_set_tuning(qd_ctx *c, const qd_tuning *t) {
    if (!c || !t) return fail(QD_ERR_INVALID, "null argument");
    qd_tuning n = c->tune;
    if (t->dedup_cap) {
        if (t->dedup_cap != 2048) return fail(QD_ERR_INVALID, "dedup_cap must be 2048");
        n.dedup_cap = t->dedup_cap;
    }
    if (t->dedup_bucket_target) n.dedup_bucket_target = t->dedup_bucket_target;
    if (n.dedup_bucket_target < 64 || n.dedup_bucket_target > n.dedup_cap / 2)
        return fail(QD_ERR_INVALID, "dedup_bucket_target must be in [64, dedup_cap / 2]");
    if (t->first_pass_bits) {
        if (t->first_pass_bits > 12) return fail(QD_ERR_INVALID, "first_pass_bits must be in [1, 12]");
        n.first_pass_bits = t->first_pass_bits;
    }
    if (t->generate_block) {
        const u32 g = t->generate_block;
        if (g != 64 && g != 128 && g != 256 && g != 512 && g != 1024)
            return fail(QD_ERR_INVALID, "generate_block must be 64, 128, 256, 512 or 1024");
        n.generate_block = g;
    }
    if (t->generate_grid) n.generate_grid = t->generate_grid;
    c->tune = n;
    return QD_OK;
}

int qd_ctx_set_profiling(qd_ctx *c, int enable) {
    if (!c) return fail(QD_ERR_INVALID, "null ctx");
    QD_HIP(hipSetDevice(c->device));
    if (enable && !c->ev[0])
        for (int i = 0; i < 2 * QD_NUM_PHASES; ++i) QD_HIP(hipEventCreate(&c->ev[i]));
    c->profiling = enable != 0;
    return QD_OK;
}

int qd_ctx_reset_phase_times(qd_ctx *c) {
    if (!c) return fail(QD_ERR_INVALID, "null ctx");
    QD_HIP(hipSetDevice(c->device));
    QD_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c);
    for (int i = 0; i < QD_NUM_PHASES; ++i) {
        c->phase_ms[i] = 0;
        c->phase_calls[i] = 0;
    }
    return QD_OK;
}

int qd_ctx_phase_times(qd_ctx *c, double *ms, uint64_t *calls) {
    if (!c) return fail(QD_ERR_INVALID, "null ctx");
    QD_HIP(hipSetDevice(c->device));
    QD_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c);
    for (int i = 0; i < QD_NUM_PHASES; ++i) {
        if (ms) ms[i] = c->phase_ms[i];
        if (calls) calls[i] = c->phase_calls[i];
    }
    return QD_OK;
}

int qd_state_create(qd_ctx *c, qd_state **out) {
    if (!c || !out) return fail(QD_ERR_INVALID, "null argument");
    qd_state *s = new (std::nothrow) qd_state;
    if (!s) return fail(QD_ERR_OOM, "host allocation failed");
    s->ctx = c;
    QD_HIP(hipSetDevice(c->device));
    int rc = state_reserve(s, 0, 0);
    if (rc) {
        delete s;
        return rc;
    }
    QD_HIP(hipMemsetAsync(s->begin.p, 0, sizeof(u64), c->stream));
    QD_HIP(hipStreamSynchronize(c->stream));
    *out = s;
    return QD_OK;
}

int qd_state_destroy(qd_state *s) {
    if (!s) return QD_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    s->objects.release();
    s->begin.release();
    s->size.release();
    s->mag.release();
    delete s;
    return QD_OK;
}

int qd_state_upload(qd_state *s, uint64_t n, const void *objects, const uint64_t *object_begin,
                    const uint32_t *object_size, const double *magnitude, double total_proba) {
    if (!s || (n && (!objects || !object_begin || !object_size || !magnitude)))
        return fail(QD_ERR_INVALID, "null argument");
    if (n && object_begin[0] != 0) return fail(QD_ERR_INVALID, "object_begin[0] must be 0");
    for (u64 i = 0; i < n; ++i)
        if (object_begin[i] + object_size[i] > object_begin[i + 1])
            return fail(QD_ERR_INVALID, "objects overlap: object_begin[i] + object_size[i] > object_begin[i+1]");
    qd_ctx *c = s->ctx;
    QD_HIP(hipSetDevice(c->device));
    const u64 bytes = n ? object_begin[n] : 0;
    QD_TRY(state_reserve(s, n, bytes));
    if (bytes) QD_HIP(hipMemcpyAsync(s->objects.p, objects, bytes, hipMemcpyHostToDevice, c->stream));
    if (n) {
        QD_HIP(hipMemcpyAsync(s->begin.p, object_begin, (n + 1) * sizeof(u64), hipMemcpyHostToDevice, c->stream));
        QD_HIP(hipMemcpyAsync(s->size.p, object_size, n * sizeof(u32), hipMemcpyHostToDevice, c->stream));
        QD_HIP(hipMemcpyAsync(s->mag.p, magnitude, n * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    } else {
        QD_HIP(hipMemsetAsync(s->begin.p, 0, sizeof(u64), c->stream));
    }
    QD_HIP(hipStreamSynchronize(c->stream));
    s->n = n;
    s->bytes = bytes;
    s->total_proba = total_proba;
    return QD_OK;
}

int qd_state_info(qd_state *s, uint64_t *n, uint64_t *bytes, double *total) {
    if (!s) return fail(QD_ERR_INVALID, "null state");
    if (n) *n = s->n;
    if (bytes) *bytes = s->bytes;
    if (total) *total = s->total_proba;
    return QD_OK;
}

int qd_state_download(qd_state *s, void *objects, uint64_t *object_begin, uint32_t *object_size, double *magnitude) {
    if (!s) return fail(QD_ERR_INVALID, "null state");
    qd_ctx *c = s->ctx;
    QD_HIP(hipSetDevice(c->device));
    if (objects && s->bytes) QD_HIP(hipMemcpyAsync(objects, s->objects.p, s->bytes, hipMemcpyDeviceToHost, c->stream));
    if (object_begin)
        QD_HIP(hipMemcpyAsync(object_begin, s->begin.p, (s->n + 1) * sizeof(u64), hipMemcpyDeviceToHost, c->stream));
    if (object_size && s->n)
        QD_HIP(hipMemcpyAsync(object_size, s->size.p, s->n * sizeof(u32), hipMemcpyDeviceToHost, c->stream));
    if (magnitude && s->n)
        QD_HIP(hipMemcpyAsync(magnitude, s->mag.p, s->n * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    QD_HIP(hipStreamSynchronize(c->stream));
    return QD_OK;
}

int qd_state_device_ptrs(qd_state *s, void **objects, uint64_t **object_begin, uint32_t **object_size,
                         double **magnitude) {
    if (!s) return fail(QD_ERR_INVALID, "null state");
    if (objects) *objects = s->objects.p;
    if (object_begin) *object_begin = s->begin.as<u64>();
    if (object_size) *object_size = s->size.as<u32>();
    if (magnitude) *magnitude = s->mag.as<double>();
    return QD_OK;
}

int qd_symbolic_create(qd_ctx *c, qd_symbolic **out) {
    if (!c || !out) return fail(QD_ERR_INVALID, "null argument");
    qd_symbolic *s = new (std::nothrow) qd_symbolic;
    if (!s) return fail(QD_ERR_OOM, "host allocation failed");
    s->ctx = c;
    *out = s;
    return QD_OK;
}

int qd_symbolic_destroy(qd_symbolic *s) {
    delete s;
    return QD_OK;
}

int qd_symbolic_info(qd_symbolic *s, uint64_t *num_object, uint64_t *num_after) {
    if (!s) return fail(QD_ERR_INVALID, "null symbolic iteration");
    if (num_object) *num_object = s->num_object;
    if (num_after) *num_after = s->num_object_after_interferences;
    return QD_OK;
}

int qd_simulate(qd_ctx *c, qd_state *in, const qd_rule *rule, qd_state *out, qd_symbolic *sy,
                uint64_t max_num_object, const qd_params *params, qd_step_stats *stats) {
    if (!c || !in || !rule || !out) return fail(QD_ERR_INVALID, "null argument");
    if (in == out) return fail(QD_ERR_INVALID, "input and output states must differ");
    if (in->ctx != c || out->ctx != c) return fail(QD_ERR_INVALID, "state belongs to another context");
    qd_params prm{1e-30, 8, 1};
    if (params) prm = *params;
    qd_step_stats st;
    std::memset(&st, 0, sizeof st);
    QD_HIP(hipSetDevice(c->device));
    int rc;
    switch (rule->kind) {
    case QD_RULE_HADAMARD: rc = simulate_kind<QD_RULE_HADAMARD>(c, in, *rule, out, sy, max_num_object, prm, st); break;
    case QD_RULE_QCGD_ERASE_CREATE:
        rc = simulate_kind<QD_RULE_QCGD_ERASE_CREATE>(c, in, *rule, out, sy, max_num_object, prm, st);
        break;
    case QD_RULE_QCGD_COIN: rc = simulate_kind<QD_RULE_QCGD_COIN>(c, in, *rule, out, sy, max_num_object, prm, st); break;
    case QD_RULE_QCGD_SPLIT_MERGE:
        rc = simulate_kind<QD_RULE_QCGD_SPLIT_MERGE>(c, in, *rule, out, sy, max_num_object, prm, st);
        break;
    default: return fail(QD_ERR_UNSUPPORTED, "rule kind has no device implementation");
    }
    if (stats) *stats = st;
    return rc;
}

int qd_apply_modifier(qd_ctx *c, qd_state *s, const qd_modifier *m) {
    if (!c || !s || !m) return fail(QD_ERR_INVALID, "null argument");
    if (!modifier_known(m->kind)) return fail(QD_ERR_UNSUPPORTED, "modifier kind has no device implementation");
    QD_HIP(hipSetDevice(c->device));
    if (s->n == 0) return QD_OK;
    QD_TRY(c->counters.ensure(64 * sizeof(u64)));
    u32 *errw = reinterpret_cast<u32 *>(c->counters.as<u64>() + 4);
    QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), c->stream));
    hipLaunchKernelGGL(k_modifier, dim3(grid_for(s->n)), dim3(256), 0, c->stream, s->objects.as<u8>(),
                       s->begin.as<u64>(), s->size.as<u32>(), s->mag.as<double2>(), s->n, *m, errw);
    QD_HIP(hipGetLastError());
    u32 w = 0;
    QD_TRY(sync_read(c, errw, sizeof(u32), &w));
    return check_err_word(w);
}

int qd_normalize(qd_ctx *c, qd_state *s, double *total_proba) {
    if (!c || !s) return fail(QD_ERR_INVALID, "null argument");
    QD_HIP(hipSetDevice(c->device));
    QD_TRY(normalize_state(c, s));
    if (total_proba) *total_proba = s->total_proba;
    return QD_OK;
}

int qd_rule_num_child(const qd_rule *r, const void *parent, uint32_t size, uint32_t *num_child,
                      uint32_t *max_child_size) {
    if (!r || !parent || !num_child || !max_child_size) return fail(QD_ERR_INVALID, "null argument");
    if (!rule_known(r->kind)) return fail(QD_ERR_UNSUPPORTED, "unknown rule kind");
    *num_child = rule_num_child(*r, static_cast<const u8 *>(parent), size, max_child_size);
    return QD_OK;
}

int qd_rule_populate_child(const qd_rule *r, const void *parent, uint32_t size, void *child, uint32_t capacity,
                           uint32_t child_id, uint32_t *child_size, double *mag) {
    if (!r || !parent || !child || !child_size) return fail(QD_ERR_INVALID, "null argument");
    if (!rule_known(r->kind)) return fail(QD_ERR_UNSUPPORTED, "unknown rule kind");
    const u8 *p = static_cast<const u8 *>(parent);
    const u32 need = rule_child_size(*r, p, size, child_id);
    if (need > capacity) return fail(QD_ERR_CAPACITY, "child buffer too small");
    *child_size = rule_child_write(*r, p, size, child_id, static_cast<u8 *>(child));
    if (mag) {
        const cplx m = rule_child_mag(*r, p, size, child_id, cplx{mag[0], mag[1]});
        mag[0] = m.re;
        mag[1] = m.im;
    }
    return QD_OK;
}

int qd_rule_hash(const qd_rule *r, const void *object, uint32_t size, uint64_t *hash) {
    if (!r || !object || !hash) return fail(QD_ERR_INVALID, "null argument");
    if (!rule_known(r->kind)) return fail(QD_ERR_UNSUPPORTED, "unknown rule kind");
    *hash = rule_hash_object(*r, static_cast<const u8 *>(object), size);
    return QD_OK;
}

int qd_modifier_apply_host(const qd_modifier *m, void *object, uint32_t size, double *mag) {
    if (!m || !object || !mag) return fail(QD_ERR_INVALID, "null argument");
    if (!modifier_known(m->kind)) return fail(QD_ERR_UNSUPPORTED, "unknown modifier kind");
    cplx c{mag[0], mag[1]};
    modifier_apply(*m, static_cast<u8 *>(object), size, c);
    mag[0] = c.re;
    mag[1] = c.im;
    return QD_OK;
}

}  // extern "C"
