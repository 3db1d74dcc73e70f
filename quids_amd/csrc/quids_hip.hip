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
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/quids_hip.h"
#include "qd_device.hpp"
#include "qd_rules.hpp"
#include "qd_k_scan.hpp"
#include "qd_k_generate.hpp"
#include "qd_k_fastgen.hpp"
#include "qd_k_partition.hpp"
#include "qd_k_dedup.hpp"
#include "qd_k_select.hpp"
#include "qd_k_finalize.hpp"
#include "qd_k_adjoint.hpp"
#include "qd_k_batch.hpp"
#include "qd_k_range.hpp"
#include "qd_k_user.hpp"

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
constexpr size_t kBigAlloc = 48ull << 30;

static double dbuf_ms();
static void dbuf_log(const char *what, size_t bytes, double ms, const char *name = nullptr);

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool tight = false;  // grows to the size asked (+1/8), not by 1.5x: the states' buffers
    bool view = false;   // p points into another buffer (the hash-range slab): never freed here
    const char *name = nullptr;  // for the QD_LOG_RANGE allocation log
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return QD_OK;
        if (view) {  // a view that must grow becomes a buffer of its own
            p = nullptr;
            cap = 0;
            view = false;
        }
        const size_t old = cap;
        if (p) {
            const double t0 = dbuf_ms();
            (void)hipFree(p);
            dbuf_log("hipFree", cap, dbuf_ms() - t0, name);
            p = nullptr;
            cap = 0;
        }
        // a buffer that grows takes at least 1.5x its old size, so steps of
        // varying size settle on few reallocations; but no more than
        // kBigAlloc unless asked for it: on MI355X, allocations of 64 GB and
        // more take seconds to map and to free, smaller ones about 1 ms
        size_t want = std::max<size_t>(bytes + bytes / 8, 256);
        if (old && !tight) want = std::max(want, old + old / 2);
        want = std::min(want, std::max(bytes, kBigAlloc));
        double t0 = dbuf_ms();
        hipError_t e = hipMalloc(&p, want);
        dbuf_log(e == hipSuccess ? "hipMalloc" : "hipMalloc (failed)", want, dbuf_ms() - t0, name);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            t0 = dbuf_ms();
            e = hipMalloc(&p, std::max<size_t>(bytes, 256));
            dbuf_log(e == hipSuccess ? "hipMalloc" : "hipMalloc (failed)", bytes, dbuf_ms() - t0, name);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                p = nullptr;
                size_t freeb = 0, totalb = 0;
                (void)hipMemGetInfo(&freeb, &totalb);
                return fail(QD_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed (device free " +
                                            std::to_string(freeb) + " of " + std::to_string(totalb) + " bytes)");
            }
            want = std::max<size_t>(bytes, 256);
        }
        cap = want;
        return QD_OK;
    }
    void release() {
        if (p && !view) {
            const double t0 = dbuf_ms();
            (void)hipFree(p);
            dbuf_log("hipFree", cap, dbuf_ms() - t0, name);
        }
        p = nullptr;
        cap = 0;
        view = false;
    }
    // releases, then allocates exactly `bytes`
    int alloc_exact(size_t bytes) {
        release();
        const double t0 = dbuf_ms();
        const hipError_t e = hipMalloc(&p, bytes);
        dbuf_log(e == hipSuccess ? "hipMalloc" : "hipMalloc (failed)", bytes, dbuf_ms() - t0, name);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            size_t freeb = 0, totalb = 0;
            (void)hipMemGetInfo(&freeb, &totalb);
            return fail(QD_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed (device free " +
                                        std::to_string(freeb) + " of " + std::to_string(totalb) + " bytes)");
        }
        cap = bytes;
        return QD_OK;
    }
    // points this buffer at bytes of `base` (releasing what it owned)
    void set_view(void *base, size_t bytes) {
        release();
        p = base;
        cap = bytes;
        view = true;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

static double dbuf_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// QD_LOG_RANGE: allocations and frees of 1 GB or more, with their host time.
static void dbuf_log(const char *what, size_t bytes, double ms, const char *name) {
    static const bool on = std::getenv("QD_LOG_RANGE") != nullptr;
    if (on && (bytes >= (1ull << 30) || ms > 10.0))
        std::fprintf(stderr, "[quids_hip] %s %s %.2f GB: %.1f ms\n", what, name ? name : "-", bytes / 1e9, ms);
}

// Device scratch of qd_average_value_noalloc, a file-scope array that every
// GPU the library runs on holds from module load until the process ends:
// the per-object observable values of states of up to kNoAllocMaxObjects
// objects, then the canonical sum's partial results.  Host code reaches it
// only through the address hipGetSymbolAddress returns for the device
// (qd_ctx::noalloc), never through the array's name.  Calls that overlap on
// one GPU share it and race.
constexpr u64 kNoAllocMaxObjects = 1ull << 22;
constexpr u64 kNoAllocDoubles = kNoAllocMaxObjects + 2 * ((kNoAllocMaxObjects + 2047) / 2048 + 16);
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr u64 kHostN = 8191;  // slot of the pinned area that stages a value for the device
constexpr u64 kHostSegLoop = 8000;  // first of the slots that stage the batched select loop's state
constexpr u64 kHostSelLoop = 8064;  // first of the slots that stage the select loop's state

struct qd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    qd_phase_hook hook = nullptr;  // simulate()'s phase hook and its argument
    void *hook_user = nullptr;
    int next_phase = 0;            // the next of kPhaseNames the current call reports
    double *noalloc = nullptr;  // this device's copy of `scratch` (hipGetSymbolAddress)
    // phase timing
    bool profiling = false;
    hipEvent_t ev[2 * QD_NUM_PHASES] = {};
    bool pending[QD_NUM_PHASES] = {};
    double phase_ms[QD_NUM_PHASES] = {};
    // defaults: what quids_amd.tuning picked on the 12-node QCGD workload
    qd_tuning tune = {512, 8, 256, 16384, 1024, 0, 0, 0, 0, 0};
    u64 phase_calls[QD_NUM_PHASES] = {};
    u64 *host = nullptr;  // pinned readback area, 8192 u64
    // scratch of the step
    DBuf nc, cb, pre, tab, kept, keys;
    DBuf ebuf0, ebuf1;
    DBuf hist1, off1, cur1, off2;
    DBuf nhist, counters;
    DBuf sel0, sel1;
    DBuf bitmap, wprefix, padded, chunks, scanblk;
    DBuf biglist, redolist, bigsoff, bigkey, bigcand, bigmag, bigseg;
    DBuf pdesc, wi, wb, cpar, opar, ebuf2, cnt2, slist, ukey, selloop, segloop, bsoff, bscnt;
    // the select rounds' graphs (by round count, and the batched one) and
    // the device state they read
    struct SelGraph {
        hipGraphExec_t exec = nullptr;
        const void *L = nullptr;
    } selgraph[16], seggraph;
    DBuf adjk, adjv, adjg, adjd, adjp, adjw;
    DBuf pool0, pool1, pwpre, blist2;
    DBuf mw;  // per-object memory of memory-driven truncation
    DBuf rpart[4], pk0, pk1, tags, bmat, goff;  // hash-range passes: first-pass buckets, pool select keys,
                                                // children's slice tags, per-block bucket counts, group offsets,
    DBuf rcnt, abuf;                            // region counts, a group's merged first-level buckets
    DBuf slab;  // one allocation kept across hash-range steps, carved into their big buffers (views)
    DBuf usize, uscr;  // user rules: every candidate's child size, the generate's child bytes
    // output buffers of batched chunks (objects, begin, size, mag per chunk),
    // kept between calls
    std::vector<DBuf> parts;
    DBuf bseg, boseg, brules, bsinst, bcount, bsel, btrunc, blist, bcur, btile, bhist, bkseg, bchunk, bcsum, btot;
    std::vector<DBuf *> all() {
        std::vector<DBuf *> v = all_named();
        for (DBuf &b : parts) v.push_back(&b);
        return v;
    }
    std::vector<DBuf *> all_named() {
        return {&nc, &cb, &pre, &tab, &kept, &keys, &ebuf0, &ebuf1, &hist1, &off1, &cur1, &off2,
                &nhist, &counters, &sel0, &sel1, &bitmap, &wprefix, &padded, &chunks, &scanblk,
                &biglist, &redolist, &bigsoff, &bigkey, &bigcand, &bigmag, &bigseg, &pdesc, &wi, &wb, &cpar, &opar, &ebuf2, &cnt2, &slist, &ukey, &selloop, &segloop, &bsoff, &bscnt,
                &adjk, &adjv, &adjg, &adjd, &adjp, &adjw, &bseg, &boseg, &brules, &bsinst, &bcount, &bsel, &btrunc,
                &blist, &bcur, &btile, &bhist, &bkseg, &bchunk, &bcsum, &btot, &pool0, &pool1, &pwpre, &blist2, &mw, &rpart[0], &rpart[1], &rpart[2],
                &rpart[3], &pk0, &pk1, &tags, &bmat, &goff, &rcnt, &abuf, &slab, &usize, &uscr};
    }
    void set_names() {
        static const char *const names[] = {
            "nc", "cb", "pre", "tab", "kept", "keys", "ebuf0", "ebuf1", "hist1", "off1", "cur1", "off2", "nhist",
            "counters", "sel0", "sel1", "bitmap", "wprefix", "padded", "chunks", "scanblk", "biglist", "redolist",
            "bigsoff", "bigkey", "bigcand", "bigmag", "bigseg", "pdesc", "wi", "wb", "cpar", "opar", "ebuf2", "cnt2",
            "slist", "ukey", "selloop", "segloop", "bsoff", "bscnt", "adjk", "adjv", "adjg", "adjd", "adjp", "adjw",
            "bseg", "boseg", "brules", "bsinst", "bcount", "bsel", "btrunc", "blist", "bcur", "btile", "bhist",
            "bkseg", "bchunk", "bcsum", "btot", "pool0", "pool1", "pwpre", "blist2", "mw", "rpart0", "rpart1",
            "rpart2", "rpart3", "pk0", "pk1", "tags", "bmat", "goff", "rcnt", "abuf", "slab", "usize", "uscr"};
        std::vector<DBuf *> v = all_named();
        constexpr size_t kNames = sizeof(names) / sizeof(names[0]);
        static_assert(kNames == 81, "one name per buffer");
        for (size_t i = 0; i < v.size() && i < kNames; ++i) v[i]->name = names[i];
    }
};

struct qd_state {
    qd_ctx *ctx = nullptr;
    DBuf objects{nullptr, 0, true}, begin{nullptr, 0, true}, size{nullptr, 0, true}, mag{nullptr, 0, true};
    u64 n = 0;
    u64 bytes = 0;
    double total_proba = 1;
};

struct qd_symbolic {
    qd_ctx *ctx = nullptr;
    u64 num_object = 0;
    u64 num_object_after_interferences = 0;
};

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
// host orchestration
// ===========================================================================
static inline unsigned grid_for(u64 n, unsigned block = 256, unsigned cap = 65536) {
    const u64 g = (n + block - 1) / block;
    return unsigned(std::max<u64>(1, std::min<u64>(g, cap)));
}

// The reference's mid_step_function names in the order one simulate() call
// reports them (src/quids.hpp:454-1016); every call reports all of them.
static const char *const kPhaseNames[] = {"num_child",
                                          "truncate_symbolic - prepare",
                                          "truncate_symbolic",
                                          "prepare_index",
                                          "symbolic_iteration",
                                          "compute_collisions - prepare",
                                          "compute_collisions - insert",
                                          "compute_collisions - finalize",
                                          "truncate - prepare",
                                          "truncate",
                                          "prepare_final",
                                          "final",
                                          "normalize",
                                          "end"};
enum : int {
    PH_NUM_CHILD, PH_TRUNC_SYM_PREP, PH_TRUNC_SYM, PH_PREP_INDEX, PH_SYMBOLIC, PH_COLL_PREP, PH_COLL_INSERT,
    PH_COLL_FINAL, PH_TRUNC_PREP, PH_TRUNC, PH_PREP_FINAL, PH_FINAL, PH_NORMALIZE, PH_END
};
static void prof_collect(qd_ctx *ctx);
// Reports the phase names up to `upto` not reported yet in this call, once
// the device has finished the work enqueued before (the reference calls its
// hook between phases, with the earlier phases done).
static void phase_mark(qd_ctx *ctx, int upto) {
    if (!ctx->hook || ctx->next_phase > upto) return;
    (void)hipStreamSynchronize(ctx->stream);
    prof_collect(ctx);
    for (; ctx->next_phase <= upto; ++ctx->next_phase) ctx->hook(kPhaseNames[ctx->next_phase], ctx->hook_user);
}

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
    (void)ERR_WIDE;
    (void)ERR_SPILL;
    return QD_OK;
}

// Doubles of the canonical sum's partial results for n values: two
// ping-pong halves of the first level's chunk sums.
static constexpr u64 sum_scratch_half(u64 n) {
    return (n + kSumChunk - 1) / kSumChunk + ((n + kSumChunk - 1) / kSumChunk + kSumChunk - 1) / kSumChunk + 8;
}

// canonical sum of |m|^2 (mag != null) or of vals, the partial results in
// buf (2 * sum_scratch_half(n) doubles; null: the context's buffer)
static int canonical_sum(qd_ctx *ctx, const double2 *mag, const double *vals, u64 n, double *result,
                         double *buf = nullptr) {
    if (n == 0) {
        *result = 0.0;
        return QD_OK;
    }
    u64 nchunk = (n + kSumChunk - 1) / kSumChunk;
    const u64 cap = sum_scratch_half(n);
    if (!buf) {
        QD_TRY(ctx->chunks.ensure(2 * cap * sizeof(double)));
        buf = ctx->chunks.as<double>();
    }
    double *a = buf, *b = a + cap;
    prof_begin(ctx, QD_PHASE_NORMALIZE);
    hipLaunchKernelGGL(k_sum_chunks, dim3(nchunk), dim3(kSumThreads), 0, ctx->stream, mag, vals, n, a);
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
    QD_TRY(canonical_sum(ctx, s->mag.as<double2>(), nullptr, s->n, &total));
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


// Parent keys of the pre-truncation: (|m|^2 bits, or the random key of the
// parent's hash, ~index).
__global__ __launch_bounds__(256) void k_parent_items(Parents P, qd_rule rule, SelKey sk, SelItem *S,
                                                      const u64 *phash) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += u64(gridDim.x) * blockDim.x) {
        const double2 m = P.mag[i];
        // (the object hashes from phash when given: a user rule's hasher)
        const u64 h = sk.prob ? (phash ? phash[i] : rule_hash_object(rule, P.obj + P.begin[i], P.size[i])) : 0;
        S[i] = SelItem{sel_hi(m.x, m.y, h, sk), ~i, i};
    }
}

// Runs a fixed launch sequence on the context's stream: launched one by
// one when the tuning asks for eager launches, else replayed from a graph
// captured the first time (and again when the state it reads, L, moved).
template <class Launch>
static int run_rounds(qd_ctx *ctx, qd_ctx::SelGraph &g, const void *L, Launch &&launch) {
    hipStream_t s = ctx->stream;
    if (ctx->tune.eager_select) return launch(s);
    if (g.exec && g.L != L) {
        QD_HIP(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
    }
    if (!g.exec) {
        hipGraph_t graph = nullptr;
        QD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = launch(s);
        const hipError_t e = hipStreamEndCapture(s, &graph);
        QD_TRY(rc);
        if (e != hipSuccess) return fail(QD_ERR_HIP, std::string("select graph capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g.exec = nullptr;
            return fail(QD_ERR_HIP, std::string("select graph instantiate: ") + hipGetErrorString(ei));
        }
        g.L = L;
    }
    QD_HIP(hipGraphLaunch(g.exec, s));
    return QD_OK;
}

// The batched select's digit rounds (at most 11 are needed for 128-bit
// keys from bit 0): tiles, histogram, pick, next ranges, collect.
static int launch_seg_rounds(hipStream_t s, SegLoop *L) {
    for (u32 k = 0; k < 12; ++k) {
        hipLaunchKernelGGL(k_seg_tiles, dim3(1), dim3(1024), 0, s, L);
        const u32 first = k == 0 ? 1u : 0u;
        hipLaunchKernelGGL(k_seg_hist, dim3(kSegLoopGrid), dim3(256), 0, s, (const SegLoop *)L, k & 1u, first);
        hipLaunchKernelGGL(k_seg_pick, dim3(256), dim3(256), 0, s, L);
        hipLaunchKernelGGL(k_seg_next, dim3(1), dim3(1024), 0, s, L);
        hipLaunchKernelGGL(k_seg_collect, dim3(kSegLoopGrid), dim3(256), 0, s, (const SegLoop *)L, k & 1u, first);
    }
    QD_HIP(hipGetLastError());
    return QD_OK;
}

// The rounds of the device-driven select loop: one clear, then per 12-bit
// digit a histogram, the pick, the split and the advance (fixed grids;
// sizes and decisions are read from L).
static int launch_sel_rounds(hipStream_t s, SelLoop *L, int rounds) {
    hipLaunchKernelGGL(k_sl_clear, dim3(4), dim3(1024), 0, s, L);
    for (int k = 0; k < rounds; ++k) {
        hipLaunchKernelGGL(k_sl_hist, dim3(kSelLoopGrid), dim3(256), 0, s, (const SelLoop *)L);
        hipLaunchKernelGGL(k_sl_pick, dim3(1), dim3(1024), 0, s, L);
        hipLaunchKernelGGL(k_sl_split, dim3(kSelLoopGrid), dim3(256), 0, s, L);
        hipLaunchKernelGGL(k_sl_advance, dim3(1), dim3(kWave), 0, s, L);
    }
    QD_HIP(hipGetLastError());
    return QD_OK;
}

// Selects the r largest 128-bit keys among the ns items of S0, whose
// leading `pos` key bits are already known to tie; marks their rep bits in
// bitmap (and their slots in sbm).  Items above the chosen digit are marked
// as soon as it is known; the tied ones are carried to the next 12-bit
// digit.  The rounds run from the device (SelLoop); the context replays
// them from a hipGraph captured once per round count, unless the tuning
// asks for eager launches.
static int select_loop(qd_ctx *ctx, SelItem *S0, SelItem *S1, u64 ns, u64 r, u32 pos, u64 *bitmap,
                       const u64 *UR, u64 *sbm) {
    hipStream_t s = ctx->stream;
    if (r == 0 || r > ns) return QD_OK;
    if (r == ns) {
        hipLaunchKernelGGL(k_sel_mark_items, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, s, S0, ns, UR, bitmap,
                           sbm);
        QD_HIP(hipGetLastError());
        return QD_OK;
    }
    QD_TRY(ctx->selloop.ensure(sizeof(SelLoop)));
    SelLoop *dL = ctx->selloop.as<SelLoop>();
    SelLoop h{};
    h.items[0] = S0;
    h.items[1] = S1;
    h.UR = UR;
    h.bitmap = bitmap;
    h.sbm = sbm;
    h.hist = ctx->hist1.as<u64>();
    h.r = r;
    h.ns = ns;
    h.pos = pos;
    static_assert(sizeof(SelLoop) <= 8 * (kHostN - 1 - kHostSelLoop), "staging area");
    std::memcpy(ctx->host + kHostSelLoop, &h, sizeof h);
    QD_HIP(hipMemcpyAsync(dL, ctx->host + kHostSelLoop, sizeof h, hipMemcpyHostToDevice, s));
    const int rounds = int((128 - pos + 11) / 12);
    QD_TRY(run_rounds(ctx, ctx->selgraph[rounds], dL,
                      [&](hipStream_t st) { return launch_sel_rounds(st, dL, rounds); }));
    QD_TRY(sync_read(ctx, &dL->state, sizeof(u32), ctx->host));
    if (*reinterpret_cast<const u32 *>(ctx->host) != 1)
        return fail(QD_ERR_HIP, "select did not converge (duplicate keys)");
    return QD_OK;
}

// Marks the top r slots of nslots (empty slots have UR == kNoRep) by the
// key (|m|^2 bits desc, hash asc): their bits in sbm (whole words written)
// and their representatives in bitmap.  h is the histogram of the top 12
// |m|^2 bits over the slots (it is overwritten).  UK, when not null, holds
// the top 32 key bits of every slot (0 when empty; plain |m|^2 keys): the
// first two digits then read 4 bytes per slot instead of its whole record.
// cut24, when not null, receives the top 24 key bits of the r-th largest key.
static int select_topk(qd_ctx *ctx, const u64 *UH, const u64 *UR, const double2 *UM, const u32 *UK, u64 nslots,
                       u64 r, std::vector<u64> &h, u64 *bitmap, u64 *sbm, const SelKey &sk, u32 *cut24 = nullptr) {
    hipStream_t s = ctx->stream;
    u64 *hist1 = ctx->hist1.as<u64>();
    u64 *counters = ctx->counters.as<u64>();
    u64 cum = 0;
    int d = kHistBins - 1;
    for (; d >= 0; --d) {
        if (cum + h[d] >= r) break;
        cum += h[d];
    }
    r -= cum;
    // second 12 bits inside bin d
    const u32 d1 = u32(d);
    QD_HIP(hipMemsetAsync(hist1, 0, kHistBins * sizeof(u64), s));
    QD_HIP(hipMemsetAsync(counters + 1, 0, sizeof(u64), s));
    // (an empty slot's UK of 0 ties with no nonzero d1)
    const bool compact = UK && d1 > 0 && !sk.prob;
    if (compact)
        hipLaunchKernelGGL(k_sel_hist_k, dim3(grid_for(nslots / 4 + 1, 256, 8192)), dim3(256), 0, s, UK, nslots, d1,
                           hist1);
    else
        hipLaunchKernelGGL(k_sel_hist_u, dim3(grid_for(nslots, 256, 8192)), dim3(256), 0, s, UH, UR, UM, nslots, d1,
                           hist1, sk);
    QD_HIP(hipGetLastError());
    QD_TRY(sync_read(ctx, hist1, kHistBins * sizeof(u64), h.data()));
    cum = 0;
    for (d = kHistBins - 1; d >= 0; --d) {
        if (cum + h[d] >= r) break;
        cum += h[d];
    }
    r -= cum;
    const u64 ns = h[d];
    QD_TRY(ctx->sel0.ensure(std::max<u64>(ns, 1) * sizeof(SelItem)));
    QD_TRY(ctx->sel1.ensure(std::max<u64>(ns, 1) * sizeof(SelItem)));
    SelItem *S0 = ctx->sel0.as<SelItem>(), *S1 = ctx->sel1.as<SelItem>();
    if (compact)
        hipLaunchKernelGGL(k_sel_first_k, dim3(grid_for(nslots, 256, 16384)), dim3(256), 0, s, UK, UH, UR, UM, nslots,
                           (d1 << 12) | u32(d), S0, counters + 1, bitmap, sbm);
    else
        hipLaunchKernelGGL(k_sel_first, dim3(grid_for(nslots, 256, 16384)), dim3(256), 0, s, UH, UR, UM, nslots,
                           (d1 << 12) | u32(d), S0, counters + 1, bitmap, sbm, sk);
    QD_HIP(hipGetLastError());
    if (cut24) *cut24 = (d1 << 12) | u32(d);
    return select_loop(ctx, S0, S1, ns, r, 24, bitmap, UR, sbm);
}

// Buckets beyond the LDS capacity: one workgroup each on global scratch.
static int dedup_big_buckets(qd_ctx *ctx, const Entry *D, u64 rstride, const u64 *boff, u64 nbig, double tol, u64 *UH, u64 *UR,
                             double2 *UM, u32 *UK, u64 *U) {
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
    hipLaunchKernelGGL(k_dedup_big, dim3(nbig), dim3(1024), 0, s, D, rstride, boff, ctx->biglist.as<u64>(),
                       ctx->bigsoff.as<u64>(), ctx->bigkey.as<u64>(), ctx->bigcand.as<u64>(), ctx->bigmag.as<double2>(),
                       ctx->bigseg.as<u32>(), tol, UH, UR, UM, UK, counters, ctx->nhist.as<u64>());
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
// average at most 1.25 x dedup_bucket_target entries (the target is at most
// half the LDS table capacity, so a bucket overflows only under a
// pathological hash distribution; fuller buckets cost the dedup less per
// entry); the first pass takes up to first_pass_bits, the second pass the
// rest (at most 12 bits).
static void dedup_layout(u64 C, const qd_tuning &t, u32 &b1, u32 &b2) {
    u32 L = 0;
    while ((C >> L) > t.dedup_bucket_target + t.dedup_bucket_target / 4 && L < 24) ++L;
    b1 = std::min<u32>(L, t.first_pass_bits);
    if (L - b1 > 12) b1 = L - 12;
    b2 = L - b1;
}

// Largest candidate count one pass of a step (a batched chunk or a hash
// range) may take: 2^chunk_log2,
// or what free device memory (plus the context's cached scratch) holds at
// about 150 bytes of scratch per candidate (entries, partition buffers, dedup
// output, select keys, bitmaps, growth slack), with room left for the states
// and for buffers that grow one at a time.
static u64 step_budget(qd_ctx *ctx) {
    if (ctx->tune.chunk_log2) return 1ull << ctx->tune.chunk_log2;
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
        (void)hipGetLastError();
        return 1ull << 28;
    }
    u64 cached = 0;
    for (DBuf *b : ctx->all())
        if (!b->view) cached += b->cap;
    // the largest buffer (the second partition pass, about 1.7 x 32 bytes per
    // candidate) stays within kBigAlloc
    const u64 by_alloc = u64(double(kBigAlloc) / (32.0 * 1.8));
    return std::max<u64>(std::min<u64>(u64((double(freeb) + double(cached)) * 0.45 / 150.0), by_alloc), 1ull << 20);
}

// Grows a per-candidate scratch buffer; once a step needs an eighth of what
// the pass budget needs (`full`), it takes `full` at once, so large runs map
// each big buffer a single time (large allocations and frees cost up to
// seconds on a busy device).
static int ensure_big(DBuf &b, u64 bytes, u64 full) {
    if (bytes <= b.cap && b.p) return QD_OK;
    if (full > bytes && bytes >= full / 8) bytes = full;
    return b.ensure(bytes);
}

// Host wall clock in ms, for the QD_LOG_RANGE diagnostics.
static double wall_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The hash-range passes' own buffers.  They are kept from one such step to
// the next (mapping and unmapping buffers of tens of GB costs about a
// second per step), and released when a step runs in one pass or an
// allocation does not fit beside them.
static std::vector<DBuf *> range_cache(qd_ctx *ctx) {
    return {&ctx->slab,  &ctx->rpart[0], &ctx->rpart[1], &ctx->rpart[2], &ctx->rpart[3], &ctx->tags, &ctx->abuf,
            &ctx->rcnt, &ctx->pool0,    &ctx->pool1,    &ctx->pk0,      &ctx->pk1,      &ctx->goff};
}

static void release_range_cache(qd_ctx *ctx) {
    bool any = false;
    for (DBuf *b : range_cache(ctx)) any |= b->p != nullptr;
    if (!any) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (DBuf *b : ctx->all())
        if (b->view) b->release();  // views into the slab
    for (DBuf *b : range_cache(ctx)) b->release();
}

// DBuf::ensure that, when the device is out of memory, releases the range
// passes' cached buffers and tries once more.
static int ensure_relieved(qd_ctx *ctx, DBuf &b, size_t bytes) {
    if (bytes <= b.cap && b.p) return QD_OK;
    const int rc = b.ensure(bytes);
    if (rc != QD_ERR_OOM) return rc;
    release_range_cache(ctx);
    return b.ensure(bytes);
}

// The dedup of nbk buckets (D, rstride, woff as in k_dedup) holding n
// entries, with key bits L above the buckets' LDS home slots and tables of
// dcap entries: the uniques go to the n slots UH / UR / UM (and UK), U
// counts them.  Counters 0 and 4..6 of the context must be zero.
static int dedup_phase(qd_ctx *ctx, const Entry *D, u64 rstride, const u64 *woff, u64 nbk, u32 L, double tol, u32 dcap,
                       u64 n, u64 *UH, u64 *UR, double2 *UM, u32 *UK, u64 &U) {
    hipStream_t s = ctx->stream;
    u64 *counters = ctx->counters.as<u64>();
    u64 *hb = ctx->host;
    prof_begin(ctx, QD_PHASE_DEDUP);
    QD_TRY(ctx->biglist.ensure((n / dcap + 2) * sizeof(u64)));
    QD_TRY(ctx->redolist.ensure((nbk + 1) * sizeof(u64)));
    int ncu = 256;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device);
    // persistent grids of as many blocks as stay resident (LDS bound)
    auto resident = [&](auto kern, int threads, u64 nwork) -> unsigned {
        int per = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, 0) != hipSuccess || per < 1) {
            (void)hipGetLastError();
            per = 1;
        }
        return unsigned(std::max<u64>(std::min<u64>(nwork, u64(ncu) * u64(per)), 1));
    };
    auto run_dedup = [&](auto kern, int threads) -> int {
        hipLaunchKernelGGL(kern, dim3(resident(kern, threads, nbk)), dim3(threads), 0, s, D, rstride, woff, nbk, L,
                           tol, UH, UR, UM, UK, counters, ctx->nhist.as<u64>(), ctx->biglist.as<u64>(),
                           ctx->redolist.as<u64>());
        QD_HIP(hipGetLastError());
        return QD_OK;
    };
    // buckets holding a group too large for registers: the sort-based kernel
    auto run_sort = [&](auto kern, int threads, u64 nred) -> int {
        hipLaunchKernelGGL(kern, dim3(resident(kern, threads, nred)), dim3(threads), 0, s, D, rstride, woff, nred, L,
                           tol, UH, UR, UM, UK, counters, ctx->nhist.as<u64>(), ctx->biglist.as<u64>(),
                           (const u64 *)ctx->redolist.as<u64>());
        QD_HIP(hipGetLastError());
        return QD_OK;
    };
    if (dcap == 256) QD_TRY(run_dedup(k_dedup<128, 256, 512>, 128));
    else if (dcap == 512) QD_TRY(run_dedup(k_dedup<256, 512, 1024>, 256));
    else if (dcap == 1024) QD_TRY(run_dedup(k_dedup<512, 1024, 2048>, 512));
    else QD_TRY(run_dedup(k_dedup<1024, 2048, 4096>, 1024));
    prof_end(ctx, QD_PHASE_DEDUP);
    QD_HIP(hipMemcpyAsync(hb, counters, 7 * sizeof(u64), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    prof_collect(ctx);
    if (hb[6]) {
        prof_begin(ctx, QD_PHASE_DEDUP);
        if (dcap == 256) QD_TRY(run_sort(k_dedup_sort<128, 256, 512>, 128, hb[6]));
        else if (dcap == 512) QD_TRY(run_sort(k_dedup_sort<256, 512, 1024>, 256, hb[6]));
        else if (dcap == 1024) QD_TRY(run_sort(k_dedup_sort<512, 1024, 2048>, 512, hb[6]));
        else QD_TRY(run_sort(k_dedup_sort<1024, 2048, 4096>, 1024, hb[6]));
        prof_end(ctx, QD_PHASE_DEDUP);
        QD_HIP(hipMemcpyAsync(hb, counters, 7 * sizeof(u64), hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        prof_collect(ctx);
    }
    QD_TRY(check_err_word(u32(hb[4])));
    U = hb[0];
    if (hb[5]) QD_TRY(dedup_big_buckets(ctx, D, rstride, woff, hb[5], tol, UH, UR, UM, UK, &U));
    return QD_OK;
}

// Partition (one or two radix passes on the top key bits) and dedup of n
// candidates with keys H0 and magnitudes M0 (CI: candidate index of each
// position, null when the position is the candidate index).  hist1 holds the
// histogram of the top b1 key bits.  The uniques land in the ebuf0 space
// (H0 / M0 / CI are dead by then) as n slots UH / UR / UM; U counts them.
// The context's counters must have been cleared (unique count, error word,
// oversized-bucket count).
// With in_stride != 0 the first pass is done already: E1 holds bucket d at
// [d * in_stride, ...) with its count in cur1[d] (the fused generate).
// With UK (n u32), the dedup also writes every slot's top 32 |m|^2 bits.
static int partition_dedup(qd_ctx *ctx, const u64 *H0, const double2 *M0, const u64 *CI, u64 n, u32 b1, u32 b2,
                           double tol, qd_step_stats &st, u64 *&UH, u64 *&UR, double2 *&UM, u64 &U,
                           u64 in_stride = 0, u32 *UK = nullptr) {
    hipStream_t s = ctx->stream;
    u64 *counters = ctx->counters.as<u64>();
    u32 *errw = reinterpret_cast<u32 *>(counters + 4);
    u64 *hist1 = ctx->hist1.as<u64>();
    Entry *E1 = ctx->ebuf1.as<Entry>();
    u64 *hb = ctx->host;
    const Entry *D = E1;       // candidates the dedup reads
    u64 rstride = 0;           // fixed-capacity read stride of D (0: compact)
    const u64 *woff = nullptr; // bucket output offsets (nbk + 1)
    u64 nbk = 1;
    void *ubuf = ctx->ebuf0.p; // where the uniques go (H0/M0 are dead by then)
    const u64 nb1 = 1ull << b1;
    prof_begin(ctx, QD_PHASE_PARTITION);
    if (in_stride) {
        QD_TRY(run_scan(ctx, LoadU64{ctx->cur1.as<u64>()}, nb1, ctx->off1.as<u64>()));
    } else if (b1 == 0) {
        u64 two[2] = {0, n};
        QD_HIP(hipMemcpyAsync(ctx->off1.p, two, sizeof two, hipMemcpyHostToDevice, s));
        QD_HIP(hipMemcpyAsync(ctx->cur1.p, two, sizeof(u64), hipMemcpyHostToDevice, s));
        QD_HIP(hipStreamSynchronize(s));
    } else {
        QD_TRY(run_scan(ctx, LoadU64{hist1}, nb1, ctx->off1.as<u64>()));
        QD_HIP(hipMemcpyAsync(ctx->cur1.p, ctx->off1.p, nb1 * sizeof(u64), hipMemcpyDeviceToDevice, s));
    }
    if (!in_stride) {
        hipLaunchKernelGGL(k_scatter1, dim3(unsigned((n + kTile1 - 1) / kTile1)), dim3(256), 0, s, H0, M0, n, b1,
                           ctx->cur1.as<u64>(), E1, CI);
        QD_HIP(hipGetLastError());
    }
    if (b2 == 0) {
        woff = ctx->off1.as<u64>();
        nbk = nb1;
        rstride = in_stride;
    } else {
        const u64 nb = 1ull << (b1 + b2);
        const u64 mean = (n + nb - 1) / nb;
        // mean + 6 sigma + 128 (sub-bucket sizes are near Poisson; a spill
        // takes the exact path)
        const u64 cap2 = (mean + u64(6.0 * std::sqrt(double(mean))) + 128 + 63) / 64 * 64;
        QD_TRY(ctx->off2.ensure((nb + 1) * sizeof(u64)));
        QD_TRY(ctx->cnt2.ensure(nb * sizeof(u64)));
        // fixed-capacity sub-buckets unless their padding would pass the
        // pass budget's buffer (small sub-buckets near the budget): then
        // exact offsets from a counting pass
        const u64 full2 = u64(double(step_budget(ctx)) * 1.8);
        const bool fixed = nb * cap2 <= std::max<u64>(full2, n);
        QD_TRY(ensure_big(ctx->ebuf2, std::max<u64>(fixed ? nb * cap2 : n, n) * sizeof(Entry), full2 * sizeof(Entry)));
        Entry *E2 = ctx->ebuf2.as<Entry>();
        u32 ew = ERR_SPILL;
        if (fixed) {
            // one workgroup per first-level bucket, on at most half the CUs
            // at a time: the open lines of the sub-buckets (2^b2 per
            // workgroup) then fit L2 better, and the pass runs faster than
            // on every CU
            int ncu2 = 256;
            (void)hipDeviceGetAttribute(&ncu2, hipDeviceAttributeMultiprocessorCount, ctx->device);
            const unsigned g2 = unsigned(std::min<u64>(nb1, std::max(ncu2 / 2, 1)));
            hipLaunchKernelGGL(k_level2_fixed, dim3(g2), dim3(1024), 0, s, (const Entry *)E1, in_stride,
                               ctx->off1.as<u64>(), nb1, b1, b2, cap2, E2, ctx->cnt2.as<u64>(), errw);
            QD_HIP(hipGetLastError());
            QD_TRY(run_scan(ctx, LoadU64{ctx->cnt2.as<u64>()}, nb, ctx->off2.as<u64>()));
            QD_TRY(sync_read(ctx, errw, sizeof(u32), &ew));
            rstride = cap2;
        }
        if (ew & ERR_SPILL) {  // a sub-bucket overflowed its capacity (or no fixed layout): exact offsets
            QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), s));
            ctx->host[kHostN] = n;
            QD_HIP(hipMemcpyAsync(ctx->off2.as<u64>() + nb, ctx->host + kHostN, sizeof(u64), hipMemcpyHostToDevice,
                                  s));
            hipLaunchKernelGGL(k_level2_exact, dim3(nb1), dim3(1024), 0, s, (const Entry *)E1, in_stride,
                               ctx->off1.as<u64>(), u64(0), b1, b2, E2, ctx->off2.as<u64>());
            QD_HIP(hipGetLastError());
            rstride = 0;
            if (fixed) st.reserved |= 1;  // spill fallback used
        }
        D = E2;
        woff = ctx->off2.as<u64>();
        nbk = nb;
    }
    prof_end(ctx, QD_PHASE_PARTITION);
    const u32 L = b1 + b2;

    // ---- 5 dedup
    phase_mark(ctx, PH_COLL_INSERT);
    UH = static_cast<u64 *>(ubuf);
    UR = UH + n;
    UM = reinterpret_cast<double2 *>(UR + n);
    return dedup_phase(ctx, D, rstride, woff, nbk, L, tol, ctx->tune.dedup_cap, n, UH, UR, UM, UK, U);
}

// Regenerates the K survivors (candidate indices orep, ascending; their
// magnitudes already in out->mag) into out: sizes, padded offsets, bytes.
template <int KIND>
static int populate_out(qd_ctx *ctx, const Parents &P, const qd_rule &rule, const u64 *orep, u64 K,
                        const qd_params &prm, qd_state *out, qd_step_stats &st) {
    hipStream_t s = ctx->stream;
    u64 *hb = ctx->host;
    QD_TRY(ctx->opar.ensure(std::max<u64>(K, 1) * sizeof(u64)));
    prof_begin(ctx, QD_PHASE_FINALIZE);
    hipLaunchKernelGGL(k_child_size<KIND>, dim3(grid_for(K)), dim3(256), 0, s, P, ctx->cb.as<u64>(), rule, orep, K,
                       prm.align_byte_length, out->size.as<u32>(), ctx->padded.as<u64>(), ctx->opar.as<u64>());
    QD_HIP(hipGetLastError());
    QD_TRY(run_scan(ctx, LoadU64{ctx->padded.as<u64>()}, K, out->begin.as<u64>()));
    prof_end(ctx, QD_PHASE_FINALIZE);
    QD_TRY(sync_read(ctx, out->begin.as<u64>() + K, sizeof(u64), hb));
    const u64 bytes = hb[0];
    // (after a step in hash-range passes the slab is dead here: its room is
    // the state's if the state needs it)
    QD_TRY(ensure_relieved(ctx, out->objects, std::max<u64>(bytes, 1)));
    prof_begin(ctx, QD_PHASE_FINALIZE);
    if (KIND == QD_RULE_QCGD_SPLIT_MERGE)
        hipLaunchKernelGGL(k_populate_sm, dim3(unsigned(std::min<u64>((K + kPopSegs - 1) / kPopSegs, 65536))),
                           dim3(kWave), 0, s, P,
                           ctx->cb.as<u64>(), rule, orep, (const u64 *)ctx->opar.as<u64>(), K, out->begin.as<u64>(),
                           out->size.as<u32>(), out->objects.as<u8>());
    else
        hipLaunchKernelGGL(k_populate_flip<KIND>, dim3(unsigned(std::min<u64>((K + 7) / 8, 65536))), dim3(256), 0, s,
                           P, ctx->cb.as<u64>(), rule, orep, (const u64 *)ctx->opar.as<u64>(), K,
                           out->begin.as<u64>(), out->size.as<u32>(), out->objects.as<u8>());
    QD_HIP(hipGetLastError());
    out->n = K;
    out->bytes = bytes;
    st.output_bytes = bytes;

    prof_end(ctx, QD_PHASE_FINALIZE);
    return QD_OK;
}


// ===========================================================================
// memory-driven truncation (max_num_object == 0, src/quids.hpp:459-536):
// the reference's two binary searches for how many objects (parents in
// truncation order, then uniques in truncation order) the memory keeps,
// with this library's footprints in the memory model
// (src/quids.hpp:574-589, 850-861)
// ===========================================================================
// magnitude, offset and size of an object in a state (the reference's
// ITERATION_MEMORY_SIZE, src/quids.hpp:43)
constexpr u64 kMemPerObject = sizeof(double2) + sizeof(u64) + sizeof(u32);
// step scratch per candidate in one pass (the reference's
// SYMBOLIC_ITERATION_MEMORY_SIZE plus its hash map, src/quids.hpp:44, 577):
// the 32-byte entry of the first radix pass, about 1.3 x 32 bytes of the
// second, the 32-byte dedup slot (which reuses the generate buffer's
// space), 4 bytes of select key, bitmaps, and growth slack
constexpr u64 kMemPerCandidate = 112;

// w[p] = the modelled memory of parent p: its state entry, its padded bytes
// and its candidates' scratch (iteration::get_truncated_mem_size before the
// upsize factor).
__global__ __launch_bounds__(256) void k_parent_weight(Parents P, const u32 *nc, u64 *w) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += u64(gridDim.x) * blockDim.x)
        w[i] = kMemPerObject + (P.begin[i + 1] - P.begin[i]) + u64(nc[i]) * kMemPerCandidate;
}

// w[i] = the modelled memory of the unique in dedup slot i (0 for an empty
// slot): its state entry and its child's padded size
// (symbolic_iteration::get_truncated_mem_size).
template <int KIND>
__global__ __launch_bounds__(256) void k_slot_weight(Parents P, const u64 *cb, qd_rule rule, const u64 *UR, u64 n,
                                                     u32 align, u64 *w) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        const u64 c = UR[i];
        if (c == kNoRep) {
            w[i] = 0;
            continue;
        }
        const u64 p = find_parent(cb, P.n, c);
        const u32 sz = rule_child_size(rule, P.obj + P.begin[p], P.size[p], u32(c - cb[p]));
        w[i] = kMemPerObject + sz + align_pad(sz, align);
    }
}

// *out += the sum of w[i] over i < n with bit i set in bits (every i when
// bits is null).
__global__ __launch_bounds__(256) void k_masked_sum(const u64 *w, const u64 *bits, u64 n, u64 *out) {
    u64 acc = 0;
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
        if (!bits || ((bits[i >> 6] >> (i & 63)) & 1)) acc += w[i];
    for (int d = kWave / 2; d > 0; d >>= 1) acc += __shfl_xor(acc, d, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && acc) atomicAdd(reinterpret_cast<unsigned long long *>(out), acc);
}

static u64 state_mem(const qd_state *s) { return s->objects.cap + s->begin.cap + s->size.cap + s->mag.cap; }

static u64 ctx_scratch(qd_ctx *ctx) {
    u64 t = 0;
    for (DBuf *b : ctx->all())
        if (!b->view) t += b->cap;
    return t;
}

// Free device memory as memory-driven truncation sees it (the reference
// reads the host's, src/utils/memory.hpp:25-36), at most free_mem_mib.
static int free_mem(qd_ctx *ctx, u64 &freeb) {
    size_t f = 0, total = 0;
    QD_HIP(hipMemGetInfo(&f, &total));
    freeb = f;
    if (ctx->tune.free_mem_mib) freeb = std::min<u64>(freeb, u64(ctx->tune.free_mem_mib) << 20);
    return QD_OK;
}

// target_memory of src/quids.hpp:462-469 (parent side: avail_extra = the
// symbolic iteration's memory) and 513-520 (survivor side: nonavail_extra =
// it): the next state's memory is pulled towards the previous one's by
// equalize_factor when larger; a target below zero is zero.
static u64 target_memory(u64 next_mem, u64 prev_mem, u64 avail_extra, u64 nonavail_extra, u64 freeb,
                         const qd_params &prm) {
    double next = double(next_mem);
    if (next_mem > prev_mem) next = (1 - prm.equalize_factor) * next + prm.equalize_factor * double(prev_mem);
    const double avail = std::floor(next) + double(avail_extra) + double(freeb);
    const double non_avail = double(prev_mem) + double(nonavail_extra);
    const double t = (avail + non_avail) * (1 - prm.safety_margin) - non_avail;
    return t > 0 ? u64(t) : 0;
}

// The binary search of src/quids.hpp:472-485 (and 523-535): probes mem(k),
// the modelled memory of the first k objects in truncation order, keeps
// ranks [begin, middle) when their memory (times 9/8, the reference's
// upsize policy, on the parent side) is below what is left of the target,
// and ends with the last probed count, as the reference's truncate() calls
// leave truncated_num_object.
template <class Mem>
static int mem_search(u64 n, u64 target, bool upsize, Mem &&mem, u64 &kept) {
    u64 begin = 0, end = n, mbegin = 0;
    kept = n;
    while (end > begin + 1) {
        const u64 middle = (begin + end) / 2;
        u64 mm = 0;
        QD_TRY(mem(middle, mm));
        kept = middle;
        u64 used = mm - mbegin;
        if (upsize) used += used / 8;
        if (used < target) {
            target -= used;
            begin = middle;
            mbegin = mm;
        } else {
            end = middle;
        }
    }
    return QD_OK;
}

// Parent side (src/quids.hpp:459-486): cap = the number of parents the
// memory keeps (UINT64_MAX when all fit), with ctx->nc holding every
// parent's num_child; items(S) launches the fill of the parents' select
// items (truncation keys) into S.
template <class Items>
static int mem_parent_cap_nc(qd_ctx *ctx, const Parents &P, u64 prev_mem, u64 next_mem, const qd_params &prm,
                             qd_step_stats &st, u64 &cap, Items &&items) {
    hipStream_t s = ctx->stream;
    const u64 N = P.n;
    cap = UINT64_MAX;
    QD_TRY(ctx->mw.ensure(N * sizeof(u64)));
    QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
    u64 *counters = ctx->counters.as<u64>();
    QD_HIP(hipMemsetAsync(counters + 8, 0, sizeof(u64), s));
    hipLaunchKernelGGL(k_parent_weight, dim3(grid_for(N)), dim3(256), 0, s, P, (const u32 *)ctx->nc.as<u32>(),
                       ctx->mw.as<u64>());
    hipLaunchKernelGGL(k_masked_sum, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, (const u64 *)ctx->mw.as<u64>(),
                       (const u64 *)nullptr, N, counters + 8);
    QD_HIP(hipGetLastError());
    u64 total = 0;
    QD_TRY(sync_read(ctx, counters + 8, sizeof(u64), &total));
    u64 freeb = 0;
    QD_TRY(free_mem(ctx, freeb));
    const u64 target = target_memory(next_mem, prev_mem, ctx_scratch(ctx), 0, freeb, prm);
    st.mem_target_parents = target;
    if (total + total / 8 <= target) return QD_OK;
    // top-k parents by the pre-truncation key; their memory
    QD_TRY(ctx->sel0.ensure(N * sizeof(SelItem)));
    QD_TRY(ctx->sel1.ensure(N * sizeof(SelItem)));
    QD_TRY(ctx->kept.ensure(((N + 63) / 64) * sizeof(u64)));
    QD_TRY(ctx->hist1.ensure(kHistBins * sizeof(u64)));
    u64 *pbits = ctx->kept.as<u64>();
    auto mem = [&](u64 k, u64 &m) -> int {
        QD_HIP(hipMemsetAsync(pbits, 0, ((N + 63) / 64) * sizeof(u64), s));
        items(ctx->sel0.as<SelItem>());
        QD_HIP(hipGetLastError());
        QD_TRY(select_loop(ctx, ctx->sel0.as<SelItem>(), ctx->sel1.as<SelItem>(), N, k, 0, pbits, nullptr, nullptr));
        QD_HIP(hipMemsetAsync(counters + 8, 0, sizeof(u64), s));
        hipLaunchKernelGGL(k_masked_sum, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s,
                           (const u64 *)ctx->mw.as<u64>(), (const u64 *)pbits, N, counters + 8);
        QD_HIP(hipGetLastError());
        return sync_read(ctx, counters + 8, sizeof(u64), &m);
    };
    u64 kept = N;
    QD_TRY(mem_search(N, target, true, mem, kept));
    if (kept < N) cap = kept;
    return QD_OK;
}

// The same for a built-in rule: its num_child first.
template <int KIND>
static int mem_parent_cap(qd_ctx *ctx, const Parents &P, const qd_rule &rule, u64 prev_mem, u64 next_mem,
                          const qd_params &prm, const SelKey &sk, qd_step_stats &st, u64 &cap) {
    hipStream_t s = ctx->stream;
    const u64 N = P.n;
    QD_TRY(ctx->nc.ensure(N * sizeof(u32)));
    QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
    u64 *counters = ctx->counters.as<u64>();
    QD_HIP(hipMemsetAsync(counters + 4, 0, sizeof(u64), s));
    hipLaunchKernelGGL(k_num_child<KIND>, dim3(grid_for(N)), dim3(256), 0, s, P, rule, ctx->nc.as<u32>(),
                       reinterpret_cast<u32 *>(counters + 4));
    QD_HIP(hipGetLastError());
    u64 ew = 0;
    QD_TRY(sync_read(ctx, counters + 4, sizeof(u64), &ew));
    QD_TRY(check_err_word(u32(ew)));
    return mem_parent_cap_nc(ctx, P, prev_mem, next_mem, prm, st, cap, [&](SelItem *S) {
        hipLaunchKernelGGL(k_parent_items, dim3(grid_for(N)), dim3(256), 0, s, P, rule, sk, S, (const u64 *)nullptr);
    });
}

// Survivor side (src/quids.hpp:510-536), after the dedup: max_obj = the
// number of uniques the memory keeps (UINT64_MAX when all fit).  weights()
// launches the fill of ctx->mw (C words) with every dedup slot's modelled
// memory.  A probe runs the step's own top-K select for k and sums the
// memory of the slots it marks.
template <class Weights>
static int mem_survivor_cap_w(qd_ctx *ctx, u64 prev_mem, qd_state *out, const qd_params &prm, const SelKey &sk,
                              const u64 *UH, const u64 *UR, const double2 *UM, const u32 *UK, u64 C, u64 U,
                              qd_step_stats &st, u64 &max_obj, Weights &&weights) {
    hipStream_t s = ctx->stream;
    max_obj = UINT64_MAX;
    u64 freeb = 0;
    QD_TRY(free_mem(ctx, freeb));
    const u64 target = target_memory(state_mem(out), prev_mem, 0, ctx_scratch(ctx), freeb, prm);
    st.mem_target_survivors = target;
    if (U == 0) return QD_OK;
    QD_TRY(ctx->mw.ensure(C * sizeof(u64)));
    u64 *sum = ctx->counters.as<u64>() + 12;
    auto masked = [&](const u64 *bits, u64 &m) -> int {
        QD_HIP(hipMemsetAsync(sum, 0, sizeof(u64), s));
        hipLaunchKernelGGL(k_masked_sum, dim3(grid_for(C, 256, 4096)), dim3(256), 0, s,
                           (const u64 *)ctx->mw.as<u64>(), bits, C, sum);
        QD_HIP(hipGetLastError());
        return sync_read(ctx, sum, sizeof(u64), &m);
    };
    weights();
    QD_HIP(hipGetLastError());
    u64 total = 0;
    QD_TRY(masked(nullptr, total));
    if (total <= target) return QD_OK;
    const u64 W = (C + 63) / 64;
    QD_TRY(ctx->bitmap.ensure(W * sizeof(u64)));
    QD_TRY(ctx->slist.ensure(W * sizeof(u64)));
    u64 *bitmap = ctx->bitmap.as<u64>(), *sbm = ctx->slist.as<u64>();
    auto mem = [&](u64 k, u64 &m) -> int {
        QD_HIP(hipMemsetAsync(bitmap, 0, W * sizeof(u64), s));
        std::vector<u64> h(kHistBins);
        if (sk.prob) {
            QD_HIP(hipMemsetAsync(ctx->nhist.p, 0, kHistBins * sizeof(u64), s));
            hipLaunchKernelGGL(k_key_hist0, dim3(grid_for(C, 256, 4096)), dim3(256), 0, s, UH, UR, UM, C,
                               ctx->nhist.as<u64>(), sk);
            QD_HIP(hipGetLastError());
        }
        QD_TRY(sync_read(ctx, ctx->nhist.p, kHistBins * sizeof(u64), h.data()));
        QD_TRY(select_topk(ctx, UH, UR, UM, UK, C, k, h, bitmap, sbm, sk));
        return masked(sbm, m);
    };
    u64 kept = U;
    QD_TRY(mem_search(U, target, false, mem, kept));
    if (kept < U) max_obj = kept;
    return QD_OK;
}

// The same for a built-in rule: slot weights from its child sizes.
template <int KIND>
static int mem_survivor_cap(qd_ctx *ctx, const Parents &P, const qd_rule &rule, u64 prev_mem, qd_state *out,
                            const qd_params &prm, const SelKey &sk, const u64 *UH, const u64 *UR, const double2 *UM,
                            const u32 *UK, u64 C, u64 U, qd_step_stats &st, u64 &max_obj) {
    return mem_survivor_cap_w(ctx, prev_mem, out, prm, sk, UH, UR, UM, UK, C, U, st, max_obj, [&]() {
        hipLaunchKernelGGL(k_slot_weight<KIND>, dim3(grid_for(C)), dim3(256), 0, ctx->stream, P,
                           (const u64 *)ctx->cb.as<u64>(), rule, UR, C, prm.align_byte_length, ctx->mw.as<u64>());
    });
}

// ===========================================================================
// hash-range passes (qd_k_range.hpp): steps whose candidates exceed the
// scratch memory
// ===========================================================================
// Copies host bytes into a device buffer (grown as needed).
static int upload(qd_ctx *ctx, DBuf &b, const void *src, size_t bytes) {
    QD_TRY(b.ensure(std::max<size_t>(bytes, 8)));
    if (bytes) QD_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    QD_HIP(hipStreamSynchronize(ctx->stream));  // the host source may go away
    return QD_OK;
}

// dst[i] = src[i] + add (mod 2^64) for i < n.
__global__ __launch_bounds__(256) void k_add_offset(const u64 *src, u64 n, u64 add, u64 *dst) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
        dst[i] = src[i] + add;
}

// Per-candidate bytes of a hash-range pass: the walk's 32-byte entry in
// its region (3 % capacity slack), and an eighth (one group at a time) of
// the group's merged first-level entry, of the second pass's entry at its
// sub-buckets' slack and of the group's 32-bit select keys and pool
// appends.
constexpr u32 kRangeGroups = 8;
constexpr double kRangeBytesPerCand = 32.0 * 1.03 + (32.0 + 32.0 * 2.05 + 4.0 + 36.0) / kRangeGroups;

template <int KIND>
static int range_step(qd_ctx *ctx, const Parents &P, u64 in_bytes, const qd_rule &rule, u64 C, u64 nitems,
                      u64 max_obj, const qd_params &prm, qd_state *out, qd_symbolic *sy, qd_step_stats &st,
                      const u64 *keep, bool fill_tab) {
    hipStream_t s = ctx->stream;
    const u64 N = P.n;
    const SelKey sk{prm.simple_truncation ? 0u : 1u, 0u, prm.seed};
    u64 *counters = ctx->counters.as<u64>();
    u32 *errw = reinterpret_cast<u32 *>(counters + 4);
    u64 *hb = ctx->host;
    double tl[8] = {wall_ms()};  // QD_LOG_RANGE: entry, buffers released, tags, counted, first pass, passes, end
    // what only the single-pass path and the adjoint use is not needed here:
    // give its memory to the passes (the helpers re-allocate what they use,
    // at this path's sizes).  A batched step that runs one instance here
    // holds its plan and earlier chunks' outputs in the b* buffers and
    // parts: those stay.
    for (DBuf *b : {&ctx->ebuf0, &ctx->ebuf1, &ctx->ukey, &ctx->kept, &ctx->keys, &ctx->sel0, &ctx->sel1, &ctx->padded,
                    &ctx->chunks, &ctx->biglist, &ctx->redolist, &ctx->bigsoff, &ctx->bigkey, &ctx->bigcand,
                    &ctx->bigmag, &ctx->bigseg, &ctx->opar, &ctx->adjk, &ctx->adjv, &ctx->adjg, &ctx->adjd,
                    &ctx->adjp, &ctx->adjw})
        b->release();
    if (KIND != QD_RULE_QCGD_SPLIT_MERGE) ctx->tab.release();  // split_merge's name-root table
    QD_TRY(ctx->hist1.ensure(kHistBins * sizeof(u64)));
    QD_TRY(ctx->cur1.ensure(kRangeBins * sizeof(u64)));
    QD_TRY(ctx->nhist.ensure(kHistBins * sizeof(u64)));
    u64 *hist1 = ctx->hist1.as<u64>();
    tl[1] = wall_ms();
    const unsigned gg = unsigned(std::max<u64>(
        std::min<u64>((nitems + kGenThreads - 1) / kGenThreads, ctx->tune.generate_grid), 1));
    const bool capped = max_obj < C;
    const bool ukeys = !sk.prob;  // 32-bit |m|^2 select keys (plain truncation keys only)
    const u64 W = (C + 63) / 64;
    const u64 pstride = (nitems + kRangeTile - 1) / kRangeTile * kRangeTile;  // bytes of a tag plane

    // 1. memory: one slab, kept from one hash-range step to the next (a
    // buffer of tens of GB that was freed costs about 25 ms per GB to map
    // again), carved into this step's name-root table (split_merge, when the
    // prep left it to this step), the children's tags, the pool and the
    // passes' buffers.  Its size: what the free memory and the buffers it
    // replaces hold, after room for the next state; the pass size: what it
    // holds after the table, the tags and the pool; then the tag width:
    // slices of at most about the pass size.
    u64 budget;
    u32 T = 4;
    auto tag_bits = [&](u64 b) {
        u32 t = 4;
        while (t < kMaxTagBits && double((1u << t) - 1) * double(b) < 1.02 * double(C)) ++t;
        return t;
    };
    const u64 tab_bytes = fill_tab ? (sm_tab_base(in_bytes) + kSmSlots * 64 + 16) * sizeof(u64) : 0;
    const double Kp = double(capped ? max_obj : C);
    const double pool_bytes = capped ? 3.0 * 36.0 * Kp : 36.0 * Kp;
    constexpr double kSlabSlack = 1.03;  // the passes' layout margins over budget x bytes per candidate
    std::vector<DBuf *> slabbed = {&ctx->tags, &ctx->rpart[0], &ctx->abuf, &ctx->ebuf2, &ctx->pool0,
                                   &ctx->pool1, &ctx->pk0, &ctx->pk1, &ctx->ukey};
    if (fill_tab) slabbed.push_back(&ctx->tab);
    u64 slab_want = ctx->slab.cap;
    double log_free = 0, log_reuse = 0, log_fixed = 0;  // the plan's inputs, for QD_LOG_RANGE
    if (ctx->tune.chunk_log2) {
        budget = 1ull << ctx->tune.chunk_log2;
        T = tag_bits(budget);
        const double need = double(tab_bytes) + double(T) * double(pstride) + pool_bytes +
                            kSlabSlack * kRangeBytesPerCand * double(budget) + double(64ull << 20);
        slab_want = std::max<u64>(slab_want, u64(need));
    } else {
        size_t freeb = 0, totalb = 0;
        QD_HIP(hipMemGetInfo(&freeb, &totalb));
        u64 reuse = 0;  // owned buffers the slab replaces
        for (DBuf *b : slabbed)
            if (!b->view) reuse += b->cap;
        for (DBuf &b : ctx->rpart)
            if (!b.view) reuse += b.cap;
        // the next state with its finalize scratch (Kp objects at 1.5 x the
        // parents' mean size), beyond what the output state holds already
        const double next = std::max(0.0, 1.5 * double(in_bytes) / double(std::max<u64>(N, 1)) * Kp + 60.0 * Kp -
                                              double(state_mem(out)));
        const double fixed = next + double(2ull << 30);
        const double avail = 0.95 * (double(freeb) + double(ctx->slab.cap) + double(reuse)) - fixed;
        // keep the slab while it is within 80-115 % of what would be made now
        // (the output state's buffers, counted in `next`, change size from
        // one step to the next while the states grow; `next` is an upper
        // estimate, and should the next state not fit beside the slab, the
        // finalize releases it, dead by then)
        const double cap = double(ctx->slab.cap);
        if (!(cap >= 0.8 * avail && cap <= 1.15 * avail)) slab_want = avail > 0 ? u64(avail) : 0;
        const double room = double(slab_want) - double(tab_bytes) - pool_bytes;
        auto fit = [&](u32 t) {
            const double r = room - double(t) * double(pstride);
            return r > 0 ? u64(r / (kSlabSlack * kRangeBytesPerCand)) : 0ull;
        };
        T = tag_bits(fit(5));
        budget = fit(T);
        log_free = double(freeb);
        log_reuse = double(reuse) + cap;
        log_fixed = fixed;
    }
    budget = std::max<u64>(std::min<u64>(budget, 1ull << 32), 1ull << 10);
    const u32 S = (1u << T) - 1;  // slices; tag S: no such child
    if (slab_want != ctx->slab.cap) {  // (re)make the slab: nothing in it is live
        QD_HIP(hipStreamSynchronize(s));
        for (DBuf *b : ctx->all())
            if (b->view) b->release();
        for (DBuf *b : slabbed) b->release();
        for (DBuf &b : ctx->rpart) b.release();
        QD_TRY(ctx->slab.alloc_exact(std::max<u64>(slab_want, 1ull << 20)));
    }
    u8 *slab = ctx->slab.as<u8>();
    u64 soff = 0;
    // b at the slab's next bytes, or (past its end) a buffer of its own
    auto carve = [&](DBuf &b, u64 bytes) -> int {
        bytes = (std::max<u64>(bytes, 256) + 255) / 256 * 256;
        if (soff + bytes <= ctx->slab.cap) {
            b.set_view(slab + soff, bytes);
            soff += bytes;
            return QD_OK;
        }
        if (b.view) b.release();
        return b.ensure(bytes);
    };
    if (fill_tab) {
        QD_TRY(carve(ctx->tab, tab_bytes));
        hipLaunchKernelGGL(k_prep_fast<KIND>, dim3(grid_for(N, 16, 65536)), dim3(256), 0, s, P, keep,
                           ctx->pdesc.as<PDesc>(), ctx->nc.as<u32>(), ctx->wi.as<u32>(), ctx->tab.as<u64>(),
                           reinterpret_cast<u32 *>(counters + 3), errw);
        QD_HIP(hipGetLastError());
    }
    QD_TRY(carve(ctx->tags, std::max<u64>(u64(T) * pstride, 8)));
    tl[2] = wall_ms();

    // 2. how the candidates spread over the 4096 values of the top 12 hash
    // bits, and each child's slice
    phase_mark(ctx, PH_SYMBOLIC);
    std::vector<u64> hc(kHistBins);
    QD_HIP(hipMemsetAsync(hist1, 0, kHistBins * sizeof(u64), s));
    prof_begin(ctx, QD_PHASE_GENERATE);
    hipLaunchKernelGGL(k_range_count<KIND>, dim3(gg), dim3(kGenThreads), 0, s, (const PDesc *)ctx->pdesc.as<PDesc>(),
                       P.mag, (const u64 *)ctx->wb.as<u64>(), N, (const u64 *)ctx->cpar.as<u64>(),
                       (const u64 *)ctx->tab.as<u64>(), P.begin, rule, nitems, T, ctx->tags.as<u8>(), pstride, hist1);
    QD_HIP(hipGetLastError());
    prof_end(ctx, QD_PHASE_GENERATE);
    QD_TRY(sync_read(ctx, hist1, kHistBins * sizeof(u64), hc.data()));
    tl[3] = wall_ms();

    // 3. passes: runs of whole slices within the budget; a slice beyond the
    // budget alone is split into runs of bins (its passes list all of the
    // slice's children and keep those of their bins)
    struct Pass {
        u32 rlo, rhi, slo, shi;
        u64 n;
    };
    std::vector<Pass> passes;
    std::vector<u64> sc(S, 0);
    std::vector<u32> sb(S + 1, kHistBins);  // first bin of each slice
    for (u32 d = kHistBins; d-- > 0;) {
        const u32 sl = (d * S) >> 12;
        sc[sl] += hc[d];
        sb[sl] = d;
    }
    u64 nmax = 0;
    {
        Pass cur{0, 0, 0, 0, 0};
        auto flush = [&]() {
            if (cur.n) {
                passes.push_back(cur);
                nmax = std::max(nmax, cur.n);
            }
        };
        for (u32 sl = 0; sl < S; ++sl) {
            if (sc[sl] <= budget) {
                if (cur.n && cur.n + sc[sl] > budget) {
                    flush();
                    cur = Pass{sb[sl], sb[sl], sl, sl, 0};
                }
                if (!cur.n) cur = Pass{sb[sl], sb[sl], sl, sl, 0};
                cur.n += sc[sl];
                cur.rhi = sb[sl + 1];
                cur.shi = sl + 1;
                continue;
            }
            flush();
            cur = Pass{0, 0, 0, 0, 0};
            Pass q{sb[sl], sb[sl], sl, sl + 1, 0};
            for (u32 d = sb[sl]; d < sb[sl + 1]; ++d) {
                if (hc[d] > budget)
                    return fail(QD_ERR_CAPACITY, "hash-range pass: " + std::to_string(hc[d]) +
                                                     " candidates share 12 leading hash bits, beyond the pass budget "
                                                     "of " + std::to_string(budget));
                if (q.n && q.n + hc[d] > budget) {
                    passes.push_back(q);
                    nmax = std::max(nmax, q.n);
                    q = Pass{d, d, sl, sl + 1, 0};
                }
                q.n += hc[d];
                q.rhi = d + 1;
            }
            if (q.n) {
                passes.push_back(q);
                nmax = std::max(nmax, q.n);
            }
        }
        flush();
    }
    static const bool log_plan = std::getenv("QD_LOG_RANGE") != nullptr;
    if (log_plan)
        std::fprintf(stderr,
                     "[quids_hip] hash-range step: C=%llu work items=%llu free=%.1f GB reused=%.1f GB fixed=%.1f GB "
                     "pass budget=%llu tag bits=%u passes=%zu largest=%llu\n",
                     (unsigned long long)C, (unsigned long long)nitems, log_free / 1e9, log_reuse / 1e9,
                     log_fixed / 1e9, (unsigned long long)budget, T, passes.size(), (unsigned long long)nmax);

    // 4. layout of a pass of nmax candidates: b1 first-pass bits (at most
    // 10, fused), b2 second-pass bits (at most 12), buckets of at most 1024
    // entries on average
    u32 tgt = ctx->tune.dedup_bucket_target, dcap = ctx->tune.dedup_cap;
    while ((nmax >> 22) > tgt && tgt < 1024) tgt *= 2;
    dcap = std::max(dcap, 2 * tgt);
    // (a bucket mean up to 1.25 x the target: one radix bit fewer for
    // passes just above a power of two times the target, half the buckets)
    u32 L = 0;
    while ((nmax >> L) > tgt + tgt / 4 && L < 22) ++L;
    const u32 b1 = std::min<u32>(L, 10), b2 = L - b1;
    const u64 nb1 = 1ull << b1;
    // the two-walk way's first-level buckets: all in the regions' memory
    const u32 nparts = 1, pshift = b1;
    const u64 perpart = nb1;
    const u32 ngroups = u32(std::min<u64>(nb1, kRangeGroups));
    const u64 gbk = nb1 / ngroups;  // first-level buckets per group (a group lies in one part)
    const double gmean = double(nmax) / double(ngroups);
    const u64 gcap = u64(gmean * 1.02 + 8.0 * std::sqrt(gmean)) + 4096;  // entries of a group, at most
    const u64 nb2g = gbk << b2;     // sub-buckets of a group
    const double mean2 = double(gcap) / double(std::max<u64>(nb2g, 1));
    // candidates sharing a hash fall in one sub-bucket, so the counts spread
    // more than the Poisson counts of distinct hashes: a wide margin (a
    // spill takes the exact path)
    const u64 cap2 = (u64(mean2 + 16.0 * std::sqrt(mean2)) + 256 + 63) / 64 * 64;
    const u64 pcap = std::max<u64>(capped ? 2 * max_obj + gcap : C, 1);  // trimmed to max_obj beyond 2 max_obj
    // (the walks' blocks: as many as are resident on the device at once,
    // one per wavefront slot of a SIMD on every CU, so that no block starts
    // late)
    int ncu_w = 256;
    (void)hipDeviceGetAttribute(&ncu_w, hipDeviceAttributeMultiprocessorCount, ctx->device);
    const u64 gmax = std::min<u64>(kRangeGrid, u64(std::max(ncu_w, 1)) * kRangeWalkWaves<KIND>);
    const unsigned gr = unsigned(std::max<u64>(std::min<u64>((nitems + kRangeTile - 1) / kRangeTile, gmax), 1));
    QD_TRY(ctx->bmat.ensure(u64(gr) * nb1 * sizeof(u32)));
    QD_TRY(ctx->off1.ensure((nb1 + 1) * sizeof(u64)));
    QD_TRY(ctx->cnt2.ensure(nb2g * sizeof(u64)));
    QD_TRY(ctx->off2.ensure((nb2g + 1) * sizeof(u64)));
    QD_TRY(carve(ctx->pool0, pcap * 32));
    QD_TRY(carve(ctx->pool1, (capped ? max_obj + 1 : 1) * 32));
    if (ukeys) {
        QD_TRY(carve(ctx->pk0, pcap * sizeof(u32)));
        QD_TRY(carve(ctx->pk1, (capped ? max_obj + 1 : 1) * sizeof(u32)));
    }
    QD_TRY(carve(ctx->ebuf2, std::max<u64>(nb2g * cap2, gcap) * sizeof(Entry)));
    QD_TRY(carve(ctx->ukey, gcap * sizeof(u32)));
    QD_TRY(ctx->slist.ensure(((pcap + 63) / 64) * sizeof(u64)));
    QD_TRY(ctx->pwpre.ensure(((pcap + 63) / 64 + 1) * sizeof(u64)));
    // pool: hash, representative, magnitude (and 32-bit key) per unique
    u64 *Ph = ctx->pool0.as<u64>(), *Pr = Ph + pcap;
    double2 *Pm = reinterpret_cast<double2 *>(Pr + pcap);
    const u64 qcap = capped ? max_obj + 1 : 1;
    u64 *Qh = ctx->pool1.as<u64>(), *Qr = Qh + qcap;
    double2 *Qm = reinterpret_cast<double2 *>(Qr + qcap);
    u32 *Pk = ukeys ? ctx->pk0.as<u32>() : nullptr, *Qk = ukeys ? ctx->pk1.as<u32>() : nullptr;
    u64 ns = 0, Utot = 0;
    u32 t24 = 0;  // uniques below this 24-bit |m|^2 prefix cannot enter the pool's top max_obj
    u64 *pcount = counters + 8;

    // keeps the best max_obj of the pool (ns entries) in the pool
    auto trim_pool = [&]() -> int {
        prof_begin(ctx, QD_PHASE_SELECT);
        QD_HIP(hipMemsetAsync(ctx->nhist.p, 0, kHistBins * sizeof(u64), s));
        if (ukeys)
            hipLaunchKernelGGL(k_ukey_hist0, dim3(grid_for(ns, 256, 4096)), dim3(256), 0, s, (const u32 *)Pk, ns,
                               ctx->nhist.as<u64>());
        else
            hipLaunchKernelGGL(k_key_hist0, dim3(grid_for(ns, 256, 4096)), dim3(256), 0, s, (const u64 *)Ph,
                               (const u64 *)Pr, (const double2 *)Pm, ns, ctx->nhist.as<u64>(), sk);
        QD_HIP(hipGetLastError());
        std::vector<u64> h(kHistBins);
        QD_TRY(sync_read(ctx, ctx->nhist.p, kHistBins * sizeof(u64), h.data()));
        u64 *keep = ctx->slist.as<u64>();
        u32 c24 = 0;
        QD_TRY(select_topk(ctx, Ph, Pr, Pm, Pk, ns, max_obj, h, nullptr, keep, sk, &c24));
        const u64 Wp = (ns + 63) / 64;
        QD_TRY(run_scan(ctx, LoadPopc{keep}, Wp, ctx->pwpre.as<u64>()));
        hipLaunchKernelGGL(k_pool_compact, dim3(grid_for(ns)), dim3(256), 0, s, (const u64 *)Ph, (const u64 *)Pr,
                           (const double2 *)Pm, (const u32 *)Pk, ns, (const u64 *)keep, (const u64 *)ctx->pwpre.as<u64>(),
                           Qh, Qr, Qm, Qk);
        // back to the pool's first max_obj places
        QD_HIP(hipMemcpyAsync(Ph, Qh, max_obj * sizeof(u64), hipMemcpyDeviceToDevice, s));
        QD_HIP(hipMemcpyAsync(Pr, Qr, max_obj * sizeof(u64), hipMemcpyDeviceToDevice, s));
        QD_HIP(hipMemcpyAsync(Pm, Qm, max_obj * sizeof(double2), hipMemcpyDeviceToDevice, s));
        if (ukeys) QD_HIP(hipMemcpyAsync(Pk, Qk, max_obj * sizeof(u32), hipMemcpyDeviceToDevice, s));
        prof_end(ctx, QD_PHASE_SELECT);
        ns = max_obj;
        if (ukeys && !sk.prob) t24 = std::max(t24, c24);
        return QD_OK;
    };

    std::vector<u64> c1(nb1), op(nb1 + 1), gofs(u64(ngroups) * (gbk + 1));
    std::vector<u64> psize(nparts);
    // the single walk's regions: group q's lie in part q / gpp at
    // (q % gpp) * gr * capr, block b's at b * capr within them
    const u32 gbits = ngroups >= 8 ? 3 : ngroups >= 4 ? 2 : ngroups >= 2 ? 1 : 0;
    const double rmean = gmean / double(gr);
    const u64 capr = u64(rmean * 1.03 + 8.0 * std::sqrt(rmean)) + 256;
    const u64 gbytes = u64(gr) * capr * sizeof(Entry);
    QD_TRY(ctx->rcnt.ensure(u64(ngroups) * gr * sizeof(u32)));
    QD_TRY(carve(ctx->abuf, gcap * sizeof(Entry)));
    // the regions, group after group (also the two-walk way's buckets, and
    // the order phase's bitmaps)
    QD_TRY(carve(ctx->rpart[0], std::max<u64>(u64(ngroups) * gbytes, nmax * sizeof(Entry))));
    std::vector<u32> rc(u64(ngroups) * gr);
    bool l2_exact = false;  // second pass with exact offsets (after a spill)
    tl[4] = wall_ms();
    phase_mark(ctx, PH_COLL_INSERT);
    for (const Pass &pass : passes) {
        const u64 n = pass.n;
        const RangeSel rs{ctx->tags.as<u8>(), pstride, T, pass.slo, pass.shi, pass.rlo, pass.rhi};
        // a. one walk: every child of the range into its block's region of
        // its group
        // (qd_tuning.split_generate: always the two-walk way below)
        bool single = !ctx->tune.split_generate;
        GroupBases GB{};
        if (single) {
            for (u32 g = 0; g < ngroups; ++g) GB.g[g] = ctx->rpart[0].as<Entry>() + u64(g) * gr * capr;
            QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), s));
            prof_begin(ctx, QD_PHASE_GENERATE);
            hipLaunchKernelGGL(k_range_walkwrite<KIND>, dim3(gr), dim3(kGenThreads), 0, s,
                               (const PDesc *)ctx->pdesc.as<PDesc>(), P.mag, (const u64 *)ctx->cb.as<u64>(),
                               (const u64 *)ctx->wb.as<u64>(), N, (const u64 *)ctx->cpar.as<u64>(),
                               (const u64 *)ctx->tab.as<u64>(), P.begin, rule, nitems, rs, gbits, GB, capr,
                               ctx->rcnt.as<u32>(), b1, ctx->bmat.as<u32>(), errw);
            QD_HIP(hipGetLastError());
            prof_end(ctx, QD_PHASE_GENERATE);
            u32 ew = 0;
            QD_TRY(sync_read(ctx, errw, sizeof(u32), &ew));
            single = !(ew & ERR_SPILL1);
        }
        if (single) {
            QD_TRY(sync_read(ctx, ctx->rcnt.p, rc.size() * sizeof(u32), rc.data()));
            u64 kept = 0;
            for (u32 v : rc) kept += v;
            if (kept != n) return fail(QD_ERR_HIP, "hash-range pass kept a different candidate count than it counted");
        } else {
            // a region passed its capacity: the two-walk way, counts per
            // (block, bucket), their column scan, the buckets' offsets within
            // their parts, then the writes
            QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), s));
            st.reserved |= 8;
            prof_begin(ctx, QD_PHASE_GENERATE);
            hipLaunchKernelGGL(k_range_bincount<KIND>, dim3(gr), dim3(kGenThreads), 0, s,
                               (const PDesc *)ctx->pdesc.as<PDesc>(), P.mag, (const u64 *)ctx->cb.as<u64>(),
                               (const u64 *)ctx->wb.as<u64>(), N, (const u64 *)ctx->cpar.as<u64>(),
                               (const u64 *)ctx->tab.as<u64>(), P.begin, rule, nitems, rs, b1, ctx->bmat.as<u32>());
            hipLaunchKernelGGL(k_range_colscan, dim3(unsigned(nb1)), dim3(256), 0, s, ctx->bmat.as<u32>(), u32(gr),
                               u32(nb1), u32(nb1), ctx->cur1.as<u64>());
            QD_HIP(hipGetLastError());
            prof_end(ctx, QD_PHASE_GENERATE);
            QD_TRY(sync_read(ctx, ctx->cur1.p, nb1 * sizeof(u64), c1.data()));
            u64 kept = 0;
            for (u32 q = 0; q < nparts; ++q) {
                u64 acc2 = 0;
                for (u64 d = q * perpart; d < (q + 1) * perpart; ++d) {
                    op[d] = acc2;  // bucket d's first entry within its part
                    acc2 += c1[d];
                }
                psize[q] = acc2;
                kept += acc2;
            }
            if (kept != n)
                return fail(QD_ERR_HIP, "hash-range pass kept a different candidate count than it counted");
            for (u32 g = 0; g < ngroups; ++g) {
                const u64 g0 = u64(g) * gbk;
                u64 *o = gofs.data() + u64(g) * (gbk + 1);
                for (u64 b = 0; b < gbk; ++b) o[b] = op[g0 + b];
                o[gbk] = op[g0 + gbk - 1] + c1[g0 + gbk - 1];
                if (o[gbk] - o[0] > gcap)
                    return fail(QD_ERR_CAPACITY, "hash-range pass: hash keys too skewed for the first radix pass");
            }
            for (u32 q = 0; q < nparts; ++q) QD_TRY(ctx->rpart[q].ensure(std::max<u64>(psize[q], 1) * sizeof(Entry)));
            QD_TRY(upload(ctx, ctx->off1, op.data(), nb1 * sizeof(u64)));
            QD_TRY(upload(ctx, ctx->goff, gofs.data(), gofs.size() * sizeof(u64)));
            EParts E{};
            for (u32 q = 0; q < 4; ++q) E.p[q] = q < nparts ? ctx->rpart[q].as<Entry>() : nullptr;
            E.pshift = pshift;
            prof_begin(ctx, QD_PHASE_GENERATE);
            hipLaunchKernelGGL(k_range_binwrite<KIND>, dim3(gr), dim3(kGenThreads), 0, s,
                               (const PDesc *)ctx->pdesc.as<PDesc>(), P.mag, (const u64 *)ctx->cb.as<u64>(),
                               (const u64 *)ctx->wb.as<u64>(), N, (const u64 *)ctx->cpar.as<u64>(),
                               (const u64 *)ctx->tab.as<u64>(), P.begin, rule, nitems, rs, b1,
                               (const u32 *)ctx->bmat.as<u32>(), (const u64 *)ctx->off1.as<u64>(), E);
            QD_HIP(hipGetLastError());
            prof_end(ctx, QD_PHASE_GENERATE);
        }
        st.dedup_levels = (b1 ? 1 : 0) + (b2 ? 1 : 0);
        // b. per group of first-level buckets: (single walk) its regions
        // merged into its buckets, then the second pass, dedup, pool
        for (u32 g = 0; g < ngroups; ++g) {
            const Entry *gin;  // the group's buckets: at go[b] .. go[b + 1] of gin
            const u64 *go;
            u64 obase, ng;
            Entry *ub;  // where the group's uniques go (read by then): ng * 32 bytes
            if (single) {
                ng = 0;
                for (u32 b = 0; b < gr; ++b) ng += rc[u64(g) * gr + b];
                if (ng == 0) continue;
                if (ng > gcap)
                    return fail(QD_ERR_CAPACITY, "hash-range pass: hash keys too skewed for the first radix pass");
                const u32 *rcg = ctx->rcnt.as<u32>() + u64(g) * gr;
                prof_begin(ctx, QD_PHASE_PARTITION);
                // the walk counted each block's children per first-level
                // bucket (row = block, nb1 columns): the group's columns
                u32 *bmg = ctx->bmat.as<u32>() + u64(g) * gbk;
                hipLaunchKernelGGL(k_range_colscan, dim3(unsigned(gbk)), dim3(256), 0, s, bmg, u32(gr), u32(gbk),
                                   u32(nb1), ctx->cur1.as<u64>());
                hipLaunchKernelGGL(k_excl_small, dim3(1), dim3(64), 0, s, (const u64 *)ctx->cur1.as<u64>(), u32(gbk),
                                   ctx->off1.as<u64>());
                hipLaunchKernelGGL(k_region_place, dim3(gr), dim3(kPlaceThreads), 0, s, (const Entry *)GB.g[g], capr, rcg,
                                   u32(gr),
                                   64 - b1, u32(gbk), (const u32 *)bmg, u32(nb1), (const u64 *)ctx->off1.as<u64>(),
                                   ctx->abuf.as<Entry>());
                QD_HIP(hipGetLastError());
                prof_end(ctx, QD_PHASE_PARTITION);
                gin = ctx->abuf.as<Entry>();
                go = ctx->off1.as<u64>();
                obase = 0;
                ub = GB.g[g];
            } else {
                const u64 g0 = u64(g) * gbk;
                const u64 *o = gofs.data() + u64(g) * (gbk + 1);
                ng = o[gbk] - o[0];
                if (ng == 0) continue;
                gin = ctx->rpart[g0 >> pshift].as<Entry>();  // the group's part
                go = ctx->goff.as<u64>() + u64(g) * (gbk + 1);  // its buckets' offsets within the part
                obase = o[0];
                ub = ctx->rpart[g0 >> pshift].as<Entry>() + o[0];
            }
            Entry *E2 = ctx->ebuf2.as<Entry>();
            QD_HIP(hipMemsetAsync(counters, 0, 8 * sizeof(u64), s));
            u64 rstride = 0;
            prof_begin(ctx, QD_PHASE_PARTITION);
            if (b2) {
                bool exact = l2_exact;
                if (!exact) {
                    int ncu2 = 256;
                    (void)hipDeviceGetAttribute(&ncu2, hipDeviceAttributeMultiprocessorCount, ctx->device);
                    const unsigned g2 = unsigned(std::min<u64>(gbk, std::max(ncu2 / 2, 1)));
                    hipLaunchKernelGGL(k_level2_fixed, dim3(g2), dim3(1024), 0, s, gin, u64(0), go, gbk, b1, b2, cap2,
                                       E2, ctx->cnt2.as<u64>(), errw);
                    QD_HIP(hipGetLastError());
                    QD_TRY(run_scan(ctx, LoadU64{ctx->cnt2.as<u64>()}, nb2g, ctx->off2.as<u64>()));
                    u32 ew2 = 0;
                    QD_TRY(sync_read(ctx, errw, sizeof(u32), &ew2));
                    rstride = cap2;
                    if (ew2 & ERR_SPILL) {
                        QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), s));
                        exact = true;
                        // a rule whose hashes repeat enough to spill once
                        // (split_merge) spills in most groups: exact offsets
                        // from here on
                        l2_exact = true;
                    }
                }
                if (exact) {  // a sub-bucket passed its capacity: exact offsets
                    hb[kHostN] = ng;
                    QD_HIP(hipMemcpyAsync(ctx->off2.as<u64>() + nb2g, hb + kHostN, sizeof(u64),
                                          hipMemcpyHostToDevice, s));
                    hipLaunchKernelGGL(k_level2_exact, dim3(unsigned(gbk)), dim3(1024), 0, s, gin, u64(0), go, obase, b1,
                                       b2, E2, ctx->off2.as<u64>());
                    QD_HIP(hipGetLastError());
                    rstride = 0;
                    st.reserved |= 1;
                }
            } else {
                // one level: the group's first-pass buckets are the dedup's
                QD_HIP(hipMemcpyAsync(E2, gin + obase, ng * sizeof(Entry), hipMemcpyDeviceToDevice, s));
                hipLaunchKernelGGL(k_add_offset, dim3(grid_for(gbk + 1)), dim3(256), 0, s, go, gbk + 1, u64(0) - obase,
                                   ctx->off2.as<u64>());
                QD_HIP(hipGetLastError());
            }
            prof_end(ctx, QD_PHASE_PARTITION);
            u64 *UH = reinterpret_cast<u64 *>(ub), *UR = UH + ng;
            double2 *UM = reinterpret_cast<double2 *>(UR + ng);
            u32 *UK = ukeys ? ctx->ukey.as<u32>() : nullptr;
            u64 U = 0;
            QD_TRY(dedup_phase(ctx, E2, rstride, ctx->off2.as<u64>(), b2 ? nb2g : gbk, b1 + b2, prm.tolerance, dcap,
                               ng, UH, UR, UM, UK, U));
            Utot += U;
            if (!U) continue;
            if (capped && ns + ng > pcap) QD_TRY(trim_pool());
            hb[kHostN - 1] = ns;
            QD_HIP(hipMemcpyAsync(pcount, hb + kHostN - 1, sizeof(u64), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_pool_append,
                               dim3(unsigned(std::min<u64>((ng + 256 * kPoolPer - 1) / (256 * kPoolPer), 16384))),
                               dim3(256), 0, s, (const u64 *)UH, (const u64 *)UR, (const double2 *)UM, (const u32 *)UK,
                               t24, ng, Ph, Pr, Pm, Pk, pcount);
            QD_HIP(hipGetLastError());
            QD_TRY(sync_read(ctx, pcount, sizeof(u64), &ns));
            if (capped && ns > 2 * max_obj) QD_TRY(trim_pool());
        }
    }
    tl[5] = wall_ms();
    phase_mark(ctx, PH_TRUNC);
    if (capped && ns > max_obj) QD_TRY(trim_pool());
    st.num_unique = Utot;
    const u64 K = ns;
    st.num_survivors = K;
    if (sy) {
        sy->num_object = C;
        sy->num_object_after_interferences = Utot;
    }
    if (K == 0) {
        QD_TRY(state_reserve(out, 0, 0));
        QD_HIP(hipMemsetAsync(out->begin.p, 0, sizeof(u64), s));
        QD_HIP(hipStreamSynchronize(s));
        out->n = 0;
        out->bytes = 0;
        out->total_proba = 0;
        st.total_proba = 0;
        return QD_OK;
    }
    // order: survivors in ascending representative order, through a bitmap
    // over the candidates and its popcount prefix, in the regions' memory
    // (read by now) when it holds them, else in buffers of their own
    phase_mark(ctx, PH_PREP_FINAL);
    u64 *bitmap, *wprefix;
    if (ctx->rpart[0].cap >= (2 * W + 1) * sizeof(u64)) {
        bitmap = ctx->rpart[0].as<u64>();
        wprefix = bitmap + W;
    } else {
        QD_TRY(ctx->bitmap.ensure(W * sizeof(u64)));
        QD_TRY(ctx->wprefix.ensure((W + 1) * sizeof(u64)));
        bitmap = ctx->bitmap.as<u64>();
        wprefix = ctx->wprefix.as<u64>();
    }
    prof_begin(ctx, QD_PHASE_ORDER);
    QD_HIP(hipMemsetAsync(bitmap, 0, W * sizeof(u64), s));
    hipLaunchKernelGGL(k_mark_reps, dim3(grid_for(K)), dim3(256), 0, s, (const u64 *)Pr, K, bitmap);
    QD_TRY(run_scan(ctx, LoadPopc{bitmap}, W, wprefix));
    QD_TRY(out->begin.ensure((K + 1) * sizeof(u64)));
    QD_TRY(out->size.ensure(K * sizeof(u32)));
    QD_TRY(out->mag.ensure(K * sizeof(double2)));
    QD_TRY(ctx->padded.ensure(K * sizeof(u64)));
    QD_TRY(ctx->keys.ensure(K * sizeof(u64)));
    u64 *orep = ctx->keys.as<u64>();
    hipLaunchKernelGGL(k_place_list, dim3(grid_for(K)), dim3(256), 0, s, (const u64 *)Pr, (const double2 *)Pm, K,
                       (const u64 *)bitmap, (const u64 *)wprefix, orep, out->mag.as<double2>());
    QD_HIP(hipGetLastError());
    prof_end(ctx, QD_PHASE_ORDER);
    phase_mark(ctx, PH_FINAL);
    QD_TRY(populate_out<KIND>(ctx, P, rule, orep, K, prm, out, st));
    phase_mark(ctx, PH_NORMALIZE);
    QD_TRY(normalize_state(ctx, out));
    prof_collect(ctx);
    st.total_proba = out->total_proba;
    st.reserved |= 4;  // hash-range passes
    if (log_plan) {
        tl[6] = wall_ms();
        std::fprintf(stderr,
                     "[quids_hip] hash-range step host ms: release %.1f, tags %.1f, count %.1f, pass buffers %.1f, "
                     "passes %.1f, truncate+finalize %.1f\n",
                     tl[1] - tl[0], tl[2] - tl[1], tl[3] - tl[2], tl[4] - tl[3], tl[5] - tl[4], tl[6] - tl[5]);
    }
    return QD_OK;
}

// ===========================================================================
// batched step helpers (qd_k_batch.hpp)
// ===========================================================================
// A batch: instance i owns parents [oseg[i], oseg[i+1]) of the input view;
// the step fills its candidate offsets (cseg), unique and survivor counts,
// output object offsets (kseg) and total probabilities.
struct Batch {
    u32 n = 0;
    const u64 *d_oseg = nullptr;      // device, n + 1
    const qd_rule *d_rules = nullptr; // device, n
    const u64 *caps = nullptr;        // host, n
    SelKey sk{0, 0, 0};               // truncation key (probabilistic or not)
    std::vector<u64> cseg, U, kseg;
    std::vector<double> total;
    void set_empty() {
        if (cseg.size() != n + 1) cseg.assign(n + 1, 0);
        if (U.size() != n) U.assign(n, 0);
        kseg.assign(n + 1, 0);
        total.assign(n, 0.0);
    }
};

static int batch_cseg(qd_ctx *ctx, Batch *bt) {
    QD_TRY(ctx->bseg.ensure((bt->n + 1) * sizeof(u64)));
    hipLaunchKernelGGL(k_gather_u64, dim3(grid_for(bt->n + 1)), dim3(256), 0, ctx->stream, ctx->cb.as<u64>(),
                       bt->d_oseg, bt->n + 1, ctx->bseg.as<u64>());
    QD_HIP(hipGetLastError());
    bt->cseg.resize(bt->n + 1);
    return sync_read(ctx, ctx->bseg.p, (bt->n + 1) * sizeof(u64), bt->cseg.data());
}


// Per-instance top-K over the nslots dedup output slots; marks survivors.
// The keys of truncated instances are grouped by instance once; each radix
// round histograms the current records and carries only the ones tied at
// the chosen digit into the next round.
static int batch_select(qd_ctx *ctx, Batch *bt, u64 nslots, const u64 *UH, const u64 *UR, const double2 *UM,
                        const u32 *UK, u64 *bitmap, u64 *sbm) {
    hipStream_t s = ctx->stream;
    const u32 n = bt->n;
    const u64 *cseg = ctx->bseg.as<u64>();
    QD_TRY(ctx->bsinst.ensure(std::max<u64>(nslots, 1) * sizeof(u16)));
    const u16 *sinst = ctx->bsinst.as<u16>();
    bt->U.resize(n);
    // plain keys: counts, slot instances and the first list (8-byte records,
    // in the first pass's buffer, at least 32 bytes per slot) in one pass
    const bool narrow = UK && nslots < (1ull << 32);
    if (narrow) {
        QD_TRY(ctx->bcur.ensure(n * sizeof(u64)));
        hipLaunchKernelGGL(k_add_offset, dim3(grid_for(n)), dim3(256), 0, s, cseg, u64(n), u64(0) - bt->cseg[0],
                           ctx->bcur.as<u64>());
        const unsigned sg = unsigned(std::min<u64>((nslots + kSegFirstTile - 1) / kSegFirstTile, 16384));
        hipLaunchKernelGGL(k_seg_scatter_first, dim3(std::max(sg, 1u)), dim3(kSegFirstThreads), 0, s, UR, UK,
                           nslots, cseg, n, ctx->bcur.as<u64>(), ctx->bsinst.as<u16>(), ctx->ebuf1.as<u64>());
        QD_HIP(hipGetLastError());
        QD_TRY(sync_read(ctx, ctx->bcur.p, n * sizeof(u64), bt->U.data()));
        for (u32 i = 0; i < n; ++i) bt->U[i] -= bt->cseg[i] - bt->cseg[0];
    } else {
        QD_TRY(ctx->bcount.ensure(n * sizeof(u64)));
        QD_HIP(hipMemsetAsync(ctx->bcount.p, 0, n * sizeof(u64), s));
        hipLaunchKernelGGL(k_seg_count, dim3(grid_for(nslots, 256, 8192)), dim3(256), 0, s, UR, nslots, cseg, n,
                           ctx->bcount.as<u64>(), ctx->bsinst.as<u16>());
        QD_HIP(hipGetLastError());
        QD_TRY(sync_read(ctx, ctx->bcount.p, n * sizeof(u64), bt->U.data()));
    }
    std::vector<SegSel> sel(n);
    std::vector<u32> trunc(n, 0);
    std::vector<u64> off(n, 0), cnt(n, 0);
    u64 ntrunc = 0;
    for (u32 i = 0; i < n; ++i) {
        SegSel q{};
        if (bt->U[i] > bt->caps[i]) {
            q.r = bt->caps[i];
            q.active = 1;
            trunc[i] = 1;
            off[i] = narrow ? bt->cseg[i] - bt->cseg[0] : ntrunc;
            cnt[i] = bt->U[i];
            ntrunc += bt->U[i];
        } else {
            q.keepall = 1;
        }
        sel[i] = q;
    }
    QD_TRY(upload(ctx, ctx->bsel, sel.data(), n * sizeof(SegSel)));
    SegSel *dsel = ctx->bsel.as<SegSel>();
    if (ntrunc) {
        // the key records go where the partition's entries were (dead by now):
        // the first list in the first pass's buffer (32 bytes per slot), the
        // next rounds' in the second pass's
        QD_TRY(ctx->ebuf2.ensure(ntrunc * sizeof(SegItem)));
        QD_TRY(ctx->bhist.ensure(u64(n) * kHistBins * sizeof(u32)));
        SegItem *cur = ctx->ebuf1.as<SegItem>(), *nxt = ctx->ebuf2.as<SegItem>();
        if (!narrow) {
            QD_TRY(upload(ctx, ctx->btrunc, trunc.data(), n * sizeof(u32)));
            QD_TRY(upload(ctx, ctx->bcur, off.data(), n * sizeof(u64)));
            const unsigned sg = unsigned(std::min<u64>((nslots + kSegTile - 1) / kSegTile, 16384));
            hipLaunchKernelGGL(k_seg_scatter, dim3(std::max(sg, 1u)), dim3(kSegScatterThreads), 0, s, sinst, nslots,
                               UH, UM, UK, ctx->btrunc.as<u32>(), n, ctx->bcur.as<u64>(), cur, bt->sk);
        }
        QD_HIP(hipMemsetAsync(ctx->bhist.p, 0, u64(n) * kHistBins * sizeof(u32), s));
        // the digit rounds, driven from the device (SegLoop)
        QD_TRY(upload(ctx, ctx->bsoff, off.data(), n * sizeof(u64)));
        QD_TRY(upload(ctx, ctx->bscnt, cnt.data(), n * sizeof(u64)));
        QD_TRY(ctx->btile.ensure((ntrunc / kSegTileItems + n + 1) * sizeof(SegTile)));
        QD_TRY(ctx->segloop.ensure(sizeof(SegLoop)));
        SegLoop *dL = ctx->segloop.as<SegLoop>();
        SegLoop h{};
        h.list[0] = cur;
        h.list[1] = nxt;
        h.tiles = ctx->btile.as<SegTile>();
        h.hist = ctx->bhist.as<u32>();
        h.sel = dsel;
        h.off = ctx->bsoff.as<u64>();
        h.cnt = ctx->bscnt.as<u64>();
        h.cursor = ctx->bcur.as<u64>();
        h.UH = UH;
        h.UM = UM;
        h.sk = bt->sk;
        h.part = UK ? 1u : 0u;
        h.narrow = narrow ? 1u : 0u;
        h.n = n;
        static_assert(sizeof(SegLoop) <= 8 * (kHostSelLoop - kHostSegLoop), "staging area");
        std::memcpy(ctx->host + kHostSegLoop, &h, sizeof h);
        QD_HIP(hipMemcpyAsync(dL, ctx->host + kHostSegLoop, sizeof h, hipMemcpyHostToDevice, s));
        QD_TRY(run_rounds(ctx, ctx->seggraph, dL, [&](hipStream_t st) { return launch_seg_rounds(st, dL); }));
        QD_TRY(sync_read(ctx, dsel, n * sizeof(SegSel), sel.data()));
        for (u32 i = 0; i < n; ++i)
            if (sel[i].err || sel[i].active) return fail(QD_ERR_HIP, "batched select did not converge");
    }
    hipLaunchKernelGGL(k_seg_mark, dim3(grid_for(nslots, 256, 16384)), dim3(256), UK ? n * sizeof(u32) : 0, s, sinst,
                       UH, UR, UM, UK, nslots, (const SegSel *)dsel, n, bitmap, sbm, bt->sk);
    QD_HIP(hipGetLastError());
    return QD_OK;
}

// Output object offsets of the instances: survivor ranks at their first
// candidates.
static int batch_kseg(qd_ctx *ctx, Batch *bt, const u64 *bitmap, u64 W) {
    QD_TRY(ctx->bkseg.ensure((bt->n + 1) * sizeof(u64)));
    hipLaunchKernelGGL(k_rank_at, dim3(grid_for(bt->n + 1)), dim3(256), 0, ctx->stream, bitmap,
                       ctx->wprefix.as<u64>(), ctx->bseg.as<u64>(), bt->n + 1, W, ctx->bkseg.as<u64>());
    QD_HIP(hipGetLastError());
    bt->kseg.resize(bt->n + 1);
    return sync_read(ctx, ctx->bkseg.p, (bt->n + 1) * sizeof(u64), bt->kseg.data());
}

// Per-instance canonical sums of |m|^2 and scaling, in the plain step's
// order relative to each instance's first survivor.
static int batch_normalize(qd_ctx *ctx, Batch *bt, qd_state *out) {
    hipStream_t s = ctx->stream;
    const u32 n = bt->n;
    std::vector<u64> chunk0(n + 1, 0);
    bool big = false;
    for (u32 i = 0; i < n; ++i) {
        const u64 k = bt->kseg[i + 1] - bt->kseg[i];
        const u64 c = (k + kSumChunk - 1) / kSumChunk;
        if (c > u64(kSumChunk)) big = true;
        chunk0[i + 1] = chunk0[i] + c;
    }
    const u64 M = chunk0[n];
    QD_TRY(upload(ctx, ctx->bchunk, chunk0.data(), (n + 1) * sizeof(u64)));
    QD_TRY(ctx->bcsum.ensure(std::max<u64>(M, 1) * sizeof(double)));
    QD_TRY(ctx->btot.ensure(n * sizeof(double)));
    prof_begin(ctx, QD_PHASE_NORMALIZE);
    if (M) {
        hipLaunchKernelGGL(k_seg_sum_chunks, dim3(unsigned(M)), dim3(kSumThreads), 0, s, out->mag.as<double2>(),
                           ctx->bkseg.as<u64>(), ctx->bchunk.as<u64>(), n, ctx->bcsum.as<double>());
        hipLaunchKernelGGL(k_seg_sum_final, dim3(n), dim3(kSumThreads), 0, s, ctx->bcsum.as<double>(),
                           ctx->bchunk.as<u64>(), n, ctx->btot.as<double>());
        QD_HIP(hipGetLastError());
    } else {
        QD_HIP(hipMemsetAsync(ctx->btot.p, 0, n * sizeof(double), s));
    }
    prof_end(ctx, QD_PHASE_NORMALIZE);
    bt->total.resize(n);
    QD_TRY(sync_read(ctx, ctx->btot.p, n * sizeof(double), bt->total.data()));
    if (big) {  // instances beyond kSumChunk^2 survivors: the plain multi-level sum
        for (u32 i = 0; i < n; ++i) {
            const u64 k = bt->kseg[i + 1] - bt->kseg[i];
            if ((k + kSumChunk - 1) / kSumChunk > u64(kSumChunk))
                QD_TRY(canonical_sum(ctx, out->mag.as<double2>() + bt->kseg[i], nullptr, k, &bt->total[i]));
        }
        QD_TRY(upload(ctx, ctx->btot, bt->total.data(), n * sizeof(double)));
    }
    double sum = 0;
    for (u32 i = 0; i < n; ++i) sum += bt->total[i];
    out->total_proba = sum;
    if (out->n) {
        prof_begin(ctx, QD_PHASE_NORMALIZE);
        hipLaunchKernelGGL(k_seg_scale, dim3(grid_for(out->n)), dim3(256), 0, s, out->mag.as<double2>(), out->n,
                           ctx->bkseg.as<u64>(), n, ctx->btot.as<double>());
        QD_HIP(hipGetLastError());
        prof_end(ctx, QD_PHASE_NORMALIZE);
    }
    return QD_OK;
}

// One step over the parents P (of a state whose objects span in_bytes).  With
// BATCH, the parents are the instances of bt, each with its own rule and cap.
// 6 select of a single-instance step: marks the survivors among the C dedup
// slots (U uniques) in bitmap (by candidate index) and sbm (by slot): all of
// them when U <= max_obj, else the top max_obj by the truncation key.
static int plain_select(qd_ctx *ctx, u64 C, u64 U, u64 max_obj, const SelKey &sk, const u64 *UH, const u64 *UR,
                        const double2 *UM, const u32 *UK, u64 *bitmap, u64 *sbm) {
    hipStream_t s = ctx->stream;
    if (U <= max_obj) {
        if (U) hipLaunchKernelGGL(k_mark_all, dim3(grid_for(C, 256, 16384)), dim3(256), 0, s, UR, C, bitmap, sbm);
        QD_HIP(hipGetLastError());
        return QD_OK;
    }
    std::vector<u64> h(kHistBins);
    if (sk.prob) {  // the dedup's histogram is of |m|^2; the keys are random
        QD_HIP(hipMemsetAsync(ctx->nhist.p, 0, kHistBins * sizeof(u64), s));
        hipLaunchKernelGGL(k_key_hist0, dim3(grid_for(C, 256, 4096)), dim3(256), 0, s, UH, UR, UM, C,
                           ctx->nhist.as<u64>(), sk);
        QD_HIP(hipGetLastError());
    }
    QD_TRY(sync_read(ctx, ctx->nhist.p, kHistBins * sizeof(u64), h.data()));
    return select_topk(ctx, UH, UR, UM, UK, C, max_obj, h, bitmap, sbm, sk);
}

template <int KIND, bool BATCH>
static int step_impl(qd_ctx *ctx, const Parents &P, u64 in_bytes, u64 in_mem, const qd_rule &rule, Batch *bt,
                     qd_state *out, qd_symbolic *sy, u64 max_obj, const qd_params &prm, qd_step_stats &st) {
    hipStream_t s = ctx->stream;
    const u64 N = P.n;
    st.num_parent = N;
    const SelKey sk{prm.simple_truncation ? 0u : 1u, 0u, prm.seed};

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
        if (BATCH) bt->set_empty();
        QD_HIP(hipStreamSynchronize(s));
        return QD_OK;
    };
    if (N == 0) return finish_empty();

    // ---- pre-truncation of parents (quids.hpp:613-642): top max by |m|^2,
    // ties by ascending index, through the 128-bit select on (|m|^2, ~index).
    // With max_num_object == 0 the cap comes from the memory model
    // (src/quids.hpp:459-486).
    phase_mark(ctx, PH_NUM_CHILD);
    const u64 *keep = nullptr;
    u64 nkept = N;
    u64 pcap = max_obj;
    if (!BATCH && max_obj == 0)
        QD_TRY(mem_parent_cap<KIND>(ctx, P, rule, in_mem, state_mem(out), prm, sk, st, pcap));
    phase_mark(ctx, PH_TRUNC_SYM);
    if (!BATCH && pcap != 0 && pcap < N) {
        QD_TRY(ctx->sel0.ensure(N * sizeof(SelItem)));
        QD_TRY(ctx->sel1.ensure(N * sizeof(SelItem)));
        QD_TRY(ctx->kept.ensure(((N + 63) / 64) * sizeof(u64)));
        QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
        QD_TRY(ctx->hist1.ensure(kHistBins * sizeof(u64)));
        u64 *pbits = ctx->kept.as<u64>();
        QD_HIP(hipMemsetAsync(pbits, 0, ((N + 63) / 64) * sizeof(u64), s));
        hipLaunchKernelGGL(k_parent_items, dim3(grid_for(N)), dim3(256), 0, s, P, rule, sk, ctx->sel0.as<SelItem>(),
                           (const u64 *)nullptr);
        QD_HIP(hipGetLastError());
        SelItem *S0 = ctx->sel0.as<SelItem>(), *S1 = ctx->sel1.as<SelItem>();
        QD_TRY(select_loop(ctx, S0, S1, N, pcap, 0, pbits, nullptr, nullptr));
        keep = pbits;
        nkept = pcap;
    }
    st.num_parent_kept = nkept;
    phase_mark(ctx, PH_PREP_INDEX);

    // ---- 1 prep
    const double tp0 = wall_ms();
    QD_TRY(ensure_relieved(ctx, ctx->nc, N * sizeof(u32)));
    QD_TRY(ensure_relieved(ctx, ctx->cb, (N + 1) * sizeof(u64)));
    QD_TRY(ensure_relieved(ctx, ctx->pre, N * sizeof(u64)));
    QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
    bool fast = KIND != QD_RULE_HADAMARD && !ctx->tune.generic_path;
    // split_merge's name-root table.  After a step in hash-range passes (the
    // context holds their slab) it waits until the step's path is known: a
    // range step places it in the slab, a single pass in a buffer of its
    // own, and both fill it with a second prep launch.
    const u64 tab_bytes = (sm_tab_base(in_bytes) + kSmSlots * 64 + 16) * sizeof(u64);
    bool defer_tab = KIND == QD_RULE_QCGD_SPLIT_MERGE && !BATCH && fast && ctx->slab.p != nullptr;
    if (ctx->tab.view) ctx->tab.release();  // a view into the slab is this step's to place again
    if (KIND == QD_RULE_QCGD_SPLIT_MERGE && !defer_tab) QD_TRY(ensure_relieved(ctx, ctx->tab, tab_bytes));
    u64 *counters = ctx->counters.as<u64>();
    // counters: [0] unique count, [1] select count A, [2] select count B,
    // [3] low: max child size, [4] low: error word, [5] oversized buckets
    QD_HIP(hipMemsetAsync(counters, 0, 64 * sizeof(u64), s));
    u32 *maxsz = reinterpret_cast<u32 *>(counters + 3);
    u32 *errw = reinterpret_cast<u32 *>(counters + 4);
    if (BATCH && !fast) return fail(QD_ERR_UNSUPPORTED, "the batched step needs the QCGD fast path");
    u64 *hb = ctx->host;
    u64 nitems = 0;  // work items of the fast generate
    if (fast) {
        QD_TRY(ensure_relieved(ctx, ctx->pdesc, N * sizeof(PDesc)));
        QD_TRY(ensure_relieved(ctx, ctx->wi, N * sizeof(u32)));
        QD_TRY(ensure_relieved(ctx, ctx->wb, (N + 1) * sizeof(u64)));
        static const bool log_prep = std::getenv("QD_LOG_RANGE") != nullptr;
        if (log_prep) {
            size_t fb = 0, tb = 0;
            (void)hipMemGetInfo(&fb, &tb);
            std::fprintf(stderr, "[quids_hip] step prep buffers: %.1f host ms; free %.1f GB, context scratch %.1f GB\n",
                         wall_ms() - tp0, fb / 1e9, ctx_scratch(ctx) / 1e9);
        }
        prof_begin(ctx, QD_PHASE_PREP);
        hipLaunchKernelGGL(k_prep_fast<KIND>, dim3(grid_for(N, 16, 65536)), dim3(256), 0, s, P, keep,
                           ctx->pdesc.as<PDesc>(), ctx->nc.as<u32>(), ctx->wi.as<u32>(),
                           defer_tab ? (u64 *)nullptr : ctx->tab.as<u64>(), maxsz, errw);
        QD_HIP(hipGetLastError());
        prof_end(ctx, QD_PHASE_PREP);
        prof_begin(ctx, QD_PHASE_SCAN);
        QD_TRY(run_scan(ctx, LoadU32{ctx->nc.as<u32>()}, N, ctx->cb.as<u64>()));
        QD_TRY(run_scan(ctx, LoadU32{ctx->wi.as<u32>()}, N, ctx->wb.as<u64>()));
        prof_end(ctx, QD_PHASE_SCAN);
        QD_HIP(hipMemcpyAsync(hb, ctx->cb.as<u64>() + N, sizeof(u64), hipMemcpyDeviceToHost, s));
        QD_HIP(hipMemcpyAsync(hb + 1, counters + 3, 2 * sizeof(u64), hipMemcpyDeviceToHost, s));
        QD_HIP(hipMemcpyAsync(hb + 3, ctx->wb.as<u64>() + N, sizeof(u64), hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        prof_collect(ctx);
        nitems = hb[3];
        if (nitems && !(u32(hb[2]) & ERR_WIDE)) {
            const u64 nchunk = (nitems + kGenThreads - 1) / kGenThreads;
            QD_TRY(ensure_relieved(ctx, ctx->cpar, nchunk * sizeof(u64)));
            hipLaunchKernelGGL(k_chunk_parent, dim3(grid_for(nchunk)), dim3(256), 0, s, ctx->wb.as<u64>(), N, nchunk,
                               ctx->cpar.as<u64>());
            QD_HIP(hipGetLastError());
        }
        if (BATCH && (u32(hb[2]) & ERR_WIDE))
            return fail(QD_ERR_UNSUPPORTED, "the batched step handles graphs of at most 64 nodes");
        if (u32(hb[2]) & ERR_WIDE) {  // a graph wider than 64 nodes: generic kernels
            fast = false;
            QD_HIP(hipMemsetAsync(counters + 3, 0, 2 * sizeof(u64), s));
            if (defer_tab) {  // the generic prep fills it
                release_range_cache(ctx);
                QD_TRY(ensure_relieved(ctx, ctx->tab, tab_bytes));
                defer_tab = false;
            }
        }
    }
    if (!fast) {
        prof_begin(ctx, QD_PHASE_PREP);
        hipLaunchKernelGGL(k_prep<KIND>, dim3(grid_for(N)), dim3(256), 0, s, P, keep, rule, ctx->nc.as<u32>(),
                           ctx->pre.as<u64>(), ctx->tab.as<u64>(), maxsz, errw);
        QD_HIP(hipGetLastError());
        prof_end(ctx, QD_PHASE_PREP);
        prof_begin(ctx, QD_PHASE_SCAN);
        QD_TRY(run_scan(ctx, LoadU32{ctx->nc.as<u32>()}, N, ctx->cb.as<u64>()));
        prof_end(ctx, QD_PHASE_SCAN);
        QD_HIP(hipMemcpyAsync(hb, ctx->cb.as<u64>() + N, sizeof(u64), hipMemcpyDeviceToHost, s));
        QD_HIP(hipMemcpyAsync(hb + 1, counters + 3, 2 * sizeof(u64), hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        prof_collect(ctx);
    }
    const u64 C = hb[0];
    const u32 max_child = u32(hb[1]);  // largest child size (bytes) of the kept parents
    QD_TRY(check_err_word(u32(hb[2])));
    st.num_candidates = C;
    if (BATCH) QD_TRY(batch_cseg(ctx, bt));
    if (C == 0) return finish_empty();
    if (!BATCH && fast) {
        const u64 budget = step_budget(ctx);
        if (C > budget) {
            if (max_obj == 0) {
                // memory-driven survivor cap: the uniques of a step in
                // hash-range passes are never all held at once, so the
                // survivor side's model takes every survivor at the largest
                // child size (an upper bound of the reference's sum)
                u64 freeb = 0;
                QD_TRY(free_mem(ctx, freeb));
                const u64 target = target_memory(state_mem(out), in_mem, 0, ctx_scratch(ctx), freeb, prm);
                st.mem_target_survivors = target;
                const u64 per = kMemPerObject + max_child + prm.align_byte_length;
                max_obj = std::max<u64>(1, target / per);
            }
            return range_step<KIND>(ctx, P, in_bytes, rule, C, nitems, max_obj, prm, out, sy, st, keep, defer_tab);
        }
    }
    release_range_cache(ctx);  // one pass: its buffers are sized from the free memory
    if (defer_tab) {
        QD_TRY(ensure_relieved(ctx, ctx->tab, tab_bytes));
        prof_begin(ctx, QD_PHASE_PREP);
        hipLaunchKernelGGL(k_prep_fast<KIND>, dim3(grid_for(N, 16, 65536)), dim3(256), 0, s, P, keep,
                           ctx->pdesc.as<PDesc>(), ctx->nc.as<u32>(), ctx->wi.as<u32>(), ctx->tab.as<u64>(), maxsz,
                           errw);
        QD_HIP(hipGetLastError());
        prof_end(ctx, QD_PHASE_PREP);
    }

    // ---- 3 generate: hash and magnitude at the candidate index, or (fused)
    // straight into the first pass's fixed-capacity buckets
    u32 b1, b2;
    dedup_layout(C, ctx->tune, b1, b2);
    st.dedup_levels = (b1 ? 1 : 0) + (b2 ? 1 : 0);
    bool fused = fast && (1u << b1) <= kFusedBins && !ctx->tune.split_generate;
    const u64 nbin1 = 1ull << b1;
    u64 cap1 = C;  // bucket capacity of the fused first pass: mean + 8 sigma + 4096
    if (b1) {
        const double mean = double(C) / double(nbin1);
        cap1 = std::min<u64>(C, u64(mean + 8.0 * std::sqrt(mean)) + 4096);
    }
    // compact select keys when a top-K select of plain |m|^2 keys may follow
    u32 *UK = nullptr;
    {
        const u64 full = step_budget(ctx) * 32;
        QD_TRY(ensure_big(ctx->ebuf0, C * 32, full));
        QD_TRY(ensure_big(ctx->ebuf1, std::max(C, nbin1 * cap1) * 32, full + full / 32));
        if (!sk.prob && (BATCH || max_obj == 0 || C > max_obj)) {
            QD_TRY(ensure_big(ctx->ukey, C * sizeof(u32), full / 8));
            UK = ctx->ukey.as<u32>();
        }
    }
    QD_TRY(ctx->hist1.ensure(kHistBins * sizeof(u64)));
    QD_TRY(ctx->off1.ensure((kHistBins + 1) * sizeof(u64)));
    QD_TRY(ctx->cur1.ensure(kHistBins * sizeof(u64)));
    QD_TRY(ctx->nhist.ensure(kHistBins * sizeof(u64)));
    u64 *H0 = ctx->ebuf0.as<u64>();
    double2 *M0 = reinterpret_cast<double2 *>(H0 + C);
    Entry *E1 = ctx->ebuf1.as<Entry>();
    u64 *hist1 = ctx->hist1.as<u64>();
    QD_HIP(hipMemsetAsync(hist1, 0, kHistBins * sizeof(u64), s));
    QD_HIP(hipMemsetAsync(ctx->nhist.p, 0, kHistBins * sizeof(u64), s));
    phase_mark(ctx, PH_SYMBOLIC);
    u64 in_stride = 0;
    if (fused) {
        const unsigned gg =
            unsigned(std::min<u64>((nitems + kGenThreads - 1) / kGenThreads, ctx->tune.generate_grid));
        QD_HIP(hipMemsetAsync(ctx->cur1.p, 0, nbin1 * sizeof(u64), s));
        QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), s));
        prof_begin(ctx, QD_PHASE_GENERATE);
        if (BATCH)
            hipLaunchKernelGGL((k_generate_scatter<KIND, true>), dim3(std::max(gg, 1u)), dim3(kGenThreads), 0, s,
                               ctx->pdesc.as<PDesc>(), P.mag, ctx->cb.as<u64>(), ctx->wb.as<u64>(), N, ctx->cpar.as<u64>(),
                               ctx->tab.as<u64>(), P.begin, rule, nitems, b1, cap1, ctx->cur1.as<u64>(), E1, errw,
                               bt->d_rules, bt->d_oseg, bt->n, kFold);
        else
            hipLaunchKernelGGL((k_generate_scatter<KIND, false>), dim3(std::max(gg, 1u)), dim3(kGenThreads), 0, s,
                               ctx->pdesc.as<PDesc>(), P.mag, ctx->cb.as<u64>(), ctx->wb.as<u64>(), N, ctx->cpar.as<u64>(),
                               ctx->tab.as<u64>(), P.begin, rule, nitems, b1, cap1, ctx->cur1.as<u64>(), E1, errw,
                               (const qd_rule *)nullptr, (const u64 *)nullptr, 0u, u64(0));
        QD_HIP(hipGetLastError());
        prof_end(ctx, QD_PHASE_GENERATE);
        u32 ew = 0;
        QD_TRY(sync_read(ctx, errw, sizeof(u32), &ew));
        if (ew & ERR_SPILL1) {  // a bucket passed its capacity: generate and partition again, unfused
            QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), s));
            fused = false;
            st.reserved |= 8;
        } else {
            in_stride = cap1;
        }
    }
    if (!fused) prof_begin(ctx, QD_PHASE_GENERATE);
    if (fused) {
    } else if (fast) {
        const unsigned gg =
            unsigned(std::min<u64>((nitems + kGenThreads - 1) / kGenThreads, ctx->tune.generate_grid));
        if (BATCH)
            hipLaunchKernelGGL((k_generate_fast<KIND, true>), dim3(std::max(gg, 1u)), dim3(kGenThreads), 0, s,
                               ctx->pdesc.as<PDesc>(), P.mag, ctx->cb.as<u64>(), ctx->wb.as<u64>(), N, ctx->cpar.as<u64>(),
                               ctx->tab.as<u64>(), P.begin, rule, nitems, H0, M0, b1, hist1, bt->d_rules, bt->d_oseg,
                               bt->n, kFold);
        else
            hipLaunchKernelGGL((k_generate_fast<KIND, false>), dim3(std::max(gg, 1u)), dim3(kGenThreads), 0, s,
                               ctx->pdesc.as<PDesc>(), P.mag, ctx->cb.as<u64>(), ctx->wb.as<u64>(), N, ctx->cpar.as<u64>(),
                               ctx->tab.as<u64>(), P.begin, rule, nitems, H0, M0, b1, hist1,
                               (const qd_rule *)nullptr, (const u64 *)nullptr, 0u, u64(0));
    } else {
        const unsigned gb = ctx->tune.generate_block;
        const unsigned gg = grid_for(C, gb, ctx->tune.generate_grid);
        hipLaunchKernelGGL(k_generate<KIND>, dim3(gg), dim3(gb), 0, s, P, ctx->cb.as<u64>(), rule, ctx->pre.as<u64>(),
                           ctx->tab.as<u64>(), C, H0, M0, b1, hist1);
    }
    QD_HIP(hipGetLastError());
    if (!fused) prof_end(ctx, QD_PHASE_GENERATE);

    u64 *UH, *UR;
    double2 *UM;
    u64 U;
    phase_mark(ctx, PH_COLL_PREP);
    QD_TRY(partition_dedup(ctx, H0, M0, nullptr, C, b1, b2, prm.tolerance, st, UH, UR, UM, U, in_stride, UK));
    phase_mark(ctx, PH_COLL_FINAL);
    st.num_unique = U;

    // memory-driven survivor cap (max_num_object == 0, src/quids.hpp:510-536)
    if (!BATCH && max_obj == 0) QD_TRY(mem_survivor_cap<KIND>(ctx, P, rule, in_mem, out, prm, sk, UH, UR, UM, UK, C, U,
                                                               st, max_obj));

    // ---- 6 select + 7 order
    phase_mark(ctx, PH_TRUNC);
    const u64 W = (C + 63) / 64;
    QD_TRY(ctx->bitmap.ensure(W * sizeof(u64)));
    QD_TRY(ctx->wprefix.ensure((W + 1) * sizeof(u64)));
    u64 *bitmap = ctx->bitmap.as<u64>();
    QD_TRY(ctx->slist.ensure(W * sizeof(u64)));
    u64 *sbm = ctx->slist.as<u64>();
    prof_begin(ctx, QD_PHASE_SELECT);
    QD_HIP(hipMemsetAsync(bitmap, 0, W * sizeof(u64), s));
    if (BATCH) QD_TRY(batch_select(ctx, bt, C, UH, UR, UM, UK, bitmap, sbm));
    else QD_TRY(plain_select(ctx, C, U, max_obj, sk, UH, UR, UM, UK, bitmap, sbm));
    prof_end(ctx, QD_PHASE_SELECT);
    phase_mark(ctx, PH_PREP_FINAL);
    prof_begin(ctx, QD_PHASE_ORDER);
    QD_TRY(run_scan(ctx, LoadPopc{bitmap}, W, ctx->wprefix.as<u64>()));
    prof_end(ctx, QD_PHASE_ORDER);
    QD_TRY(sync_read(ctx, ctx->wprefix.as<u64>() + W, sizeof(u64), hb));
    const u64 K = hb[0];
    st.num_survivors = K;
    if (K == 0) return finish_empty();
    if (BATCH) QD_TRY(batch_kseg(ctx, bt, bitmap, W));

    // ---- 8 finalize
    QD_TRY(out->begin.ensure((K + 1) * sizeof(u64)));
    QD_TRY(out->size.ensure(K * sizeof(u32)));
    QD_TRY(out->mag.ensure(K * sizeof(double2)));
    QD_TRY(ctx->padded.ensure(K * sizeof(u64)));
    QD_TRY(ctx->keys.ensure(K * sizeof(u64)));
    u64 *orep = ctx->keys.as<u64>();
    prof_begin(ctx, QD_PHASE_ORDER);
    hipLaunchKernelGGL(k_place, dim3(grid_for((C + 63) / 64, 256, 4096)), dim3(256), 0, s, UR, UM, (const u64 *)sbm, C, bitmap,
                       ctx->wprefix.as<u64>(), orep, out->mag.as<double2>());
    prof_end(ctx, QD_PHASE_ORDER);
    phase_mark(ctx, PH_FINAL);
    QD_TRY(populate_out<KIND>(ctx, P, rule, orep, K, prm, out, st));
    // ---- 9 normalize
    phase_mark(ctx, PH_NORMALIZE);
    if (BATCH) QD_TRY(batch_normalize(ctx, bt, out));
    else QD_TRY(normalize_state(ctx, out));
    prof_collect(ctx);
    st.total_proba = out->total_proba;
    if (sy) {
        sy->num_object = C;
        sy->num_object_after_interferences = U;
    }
    return QD_OK;
}

// The context's scratch holds nothing live between calls.  When it holds
// more than a quarter of the device while less than half is free, it is
// released, so that the coming step's buffers (sized for its own
// candidates) can be mapped; otherwise it is kept, and chained steps of
// similar size map nothing.  The range passes' cache stays: the step
// releases it once it knows it runs in one pass, or when an allocation
// needs its room.
static int relieve_memory(qd_ctx *ctx) {
    if (ctx->slab.p) return QD_OK;  // hash-range steps: their buffers are sized and kept on purpose
    size_t freeb = 0, totalb = 0;
    QD_HIP(hipMemGetInfo(&freeb, &totalb));
    if (ctx_scratch(ctx) > totalb / 4 && freeb < totalb / 2) {
        QD_HIP(hipStreamSynchronize(ctx->stream));
        const std::vector<DBuf *> keep = range_cache(ctx);
        for (DBuf *b : ctx->all())
            if (std::find(keep.begin(), keep.end(), b) == keep.end()) b->release();
    }
    return QD_OK;
}

template <int KIND>
static int simulate_kind(qd_ctx *ctx, qd_state *in, const qd_rule &rule, qd_state *out, qd_symbolic *sy, u64 max_obj,
                         const qd_params &prm, qd_step_stats &st) {
    QD_TRY(relieve_memory(ctx));
    const Parents P{in->objects.as<u8>(), in->begin.as<u64>(), in->size.as<u32>(), in->mag.as<double2>(), in->n};
    return step_impl<KIND, false>(ctx, P, in->bytes, state_mem(in), rule, nullptr, out, sy, max_obj, prm, st);
}

// ===========================================================================
// the step with a user rule (qd_user_rule, src/quids.hpp:105-146): the
// rule's callbacks launch its num_child, generate (populate_child + hasher),
// populate (populate_child_simple) and object-hash kernels; the rest is the
// single-pass path of step_impl with the same kernels, so a user rule that
// restates a built-in one gives bit-identical results
// ===========================================================================
static int user_cb(int rc, const char *what) {
    if (rc != 0)
        return fail(QD_ERR_INVALID, std::string("user rule: the ") + what + " callback returned " + std::to_string(rc));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(QD_ERR_HIP, std::string("user rule: the ") + what + " callback's launch: " + hipGetErrorString(e));
    return QD_OK;
}

// The child bytes' scratch of a user generate: slices of max_child bytes
// (rounded up to 16) for a grid of at most 2^18 threads, a multiple of 256,
// and at most about 1 GB.
static void user_scratch_shape(u64 C, u32 max_child, u64 &slice, u64 &threads) {
    slice = (u64(std::max<u32>(max_child, 1)) + 15) / 16 * 16;
    threads = std::min<u64>((C + 255) / 256 * 256, 1ull << 18);
    const u64 by_mem = std::max<u64>(256, ((1ull << 30) / slice) / 256 * 256);
    threads = std::min(threads, by_mem);
}

static int user_step(qd_ctx *ctx, qd_state *in, const qd_user_rule &ur, qd_state *out, qd_symbolic *sy, u64 max_obj,
                     const qd_params &prm, qd_step_stats &st) {
    QD_TRY(relieve_memory(ctx));
    hipStream_t s = ctx->stream;
    const Parents P{in->objects.as<u8>(), in->begin.as<u64>(), in->size.as<u32>(), in->mag.as<double2>(), in->n};
    const u64 N = P.n;
    const u64 in_mem = state_mem(in);
    st.num_parent = N;
    const SelKey sk{prm.simple_truncation ? 0u : 1u, 0u, prm.seed};
    const qd_objects_view pv{P.obj, P.begin, P.size, reinterpret_cast<const double *>(P.mag), N};
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

    // ---- num_child of every parent (max child sizes in wi), the parents'
    // hashes when the truncation keys need them
    phase_mark(ctx, PH_NUM_CHILD);
    QD_TRY(ensure_relieved(ctx, ctx->nc, N * sizeof(u32)));
    QD_TRY(ensure_relieved(ctx, ctx->wi, N * sizeof(u32)));
    QD_TRY(ensure_relieved(ctx, ctx->cb, (N + 1) * sizeof(u64)));
    QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
    prof_begin(ctx, QD_PHASE_PREP);
    QD_TRY(user_cb(ur.num_child(ur.self, &pv, ctx->nc.as<u32>(), ctx->wi.as<u32>(), (void *)s), "num_child"));
    const u64 *phash = nullptr;
    if (sk.prob) {
        QD_TRY(ensure_relieved(ctx, ctx->pre, N * sizeof(u64)));
        QD_TRY(user_cb(ur.hash_objects(ur.self, &pv, ctx->pre.as<u64>(), (void *)s), "hash_objects"));
        phash = ctx->pre.as<u64>();
    }
    prof_end(ctx, QD_PHASE_PREP);
    auto items = [&](SelItem *S) {
        hipLaunchKernelGGL(k_parent_items, dim3(grid_for(N)), dim3(256), 0, s, P, qd_rule{}, sk, S, phash);
    };

    // ---- pre-truncation of parents (quids.hpp:613-642), from the memory
    // model when max_num_object == 0 (src/quids.hpp:459-486)
    u64 pcap = max_obj;
    if (max_obj == 0) QD_TRY(mem_parent_cap_nc(ctx, P, in_mem, state_mem(out), prm, st, pcap, items));
    phase_mark(ctx, PH_TRUNC_SYM);
    const u64 *keep = nullptr;
    u64 nkept = N;
    if (pcap != 0 && pcap < N) {
        QD_TRY(ctx->sel0.ensure(N * sizeof(SelItem)));
        QD_TRY(ctx->sel1.ensure(N * sizeof(SelItem)));
        QD_TRY(ctx->kept.ensure(((N + 63) / 64) * sizeof(u64)));
        QD_TRY(ctx->hist1.ensure(kHistBins * sizeof(u64)));
        u64 *pbits = ctx->kept.as<u64>();
        QD_HIP(hipMemsetAsync(pbits, 0, ((N + 63) / 64) * sizeof(u64), s));
        items(ctx->sel0.as<SelItem>());
        QD_HIP(hipGetLastError());
        QD_TRY(select_loop(ctx, ctx->sel0.as<SelItem>(), ctx->sel1.as<SelItem>(), N, pcap, 0, pbits, nullptr,
                           nullptr));
        keep = pbits;
        nkept = pcap;
    }
    st.num_parent_kept = nkept;
    phase_mark(ctx, PH_PREP_INDEX);

    // ---- child offsets: the kept parents' num_child scanned
    u64 *counters = ctx->counters.as<u64>();
    QD_HIP(hipMemsetAsync(counters, 0, 64 * sizeof(u64), s));
    u32 *maxsz = reinterpret_cast<u32 *>(counters + 3);
    prof_begin(ctx, QD_PHASE_SCAN);
    hipLaunchKernelGGL(k_user_keep, dim3(grid_for(N)), dim3(256), 0, s, ctx->nc.as<u32>(),
                       (const u32 *)ctx->wi.as<u32>(), keep, N, maxsz);
    QD_HIP(hipGetLastError());
    QD_TRY(run_scan(ctx, LoadU32{ctx->nc.as<u32>()}, N, ctx->cb.as<u64>()));
    prof_end(ctx, QD_PHASE_SCAN);
    u64 *hb = ctx->host;
    QD_HIP(hipMemcpyAsync(hb, ctx->cb.as<u64>() + N, sizeof(u64), hipMemcpyDeviceToHost, s));
    QD_HIP(hipMemcpyAsync(hb + 1, counters + 3, sizeof(u64), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    prof_collect(ctx);
    const u64 C = hb[0];
    const u32 max_child = u32(hb[1]);
    st.num_candidates = C;
    if (C == 0) return finish_empty();
    const u64 budget = step_budget(ctx);
    if (C > budget)
        return fail(QD_ERR_CAPACITY, "user rule: " + std::to_string(C) + " candidates exceed the one-pass budget of " +
                                         std::to_string(budget) + " (user rules run in one pass)");
    release_range_cache(ctx);

    // ---- generate: the rule's populate_child + hasher per candidate
    u32 b1, b2;
    dedup_layout(C, ctx->tune, b1, b2);
    st.dedup_levels = (b1 ? 1 : 0) + (b2 ? 1 : 0);
    u32 *UK = nullptr;
    {
        const u64 full = budget * 32;
        QD_TRY(ensure_big(ctx->ebuf0, C * 32, full));
        QD_TRY(ensure_big(ctx->ebuf1, C * 32, full + full / 32));
        if (!sk.prob && (max_obj == 0 || C > max_obj)) {
            QD_TRY(ensure_big(ctx->ukey, C * sizeof(u32), full / 8));
            UK = ctx->ukey.as<u32>();
        }
    }
    QD_TRY(ensure_relieved(ctx, ctx->usize, C * sizeof(u32)));
    u64 slice = 0, threads = 0;
    user_scratch_shape(C, max_child, slice, threads);
    QD_TRY(ensure_relieved(ctx, ctx->uscr, slice * threads));
    QD_TRY(ctx->hist1.ensure(kHistBins * sizeof(u64)));
    QD_TRY(ctx->off1.ensure((kHistBins + 1) * sizeof(u64)));
    QD_TRY(ctx->cur1.ensure(kHistBins * sizeof(u64)));
    QD_TRY(ctx->nhist.ensure(kHistBins * sizeof(u64)));
    u64 *H0 = ctx->ebuf0.as<u64>();
    double2 *M0 = reinterpret_cast<double2 *>(H0 + C);
    u32 *csize = ctx->usize.as<u32>();
    QD_HIP(hipMemsetAsync(ctx->hist1.p, 0, kHistBins * sizeof(u64), s));
    QD_HIP(hipMemsetAsync(ctx->nhist.p, 0, kHistBins * sizeof(u64), s));
    phase_mark(ctx, PH_SYMBOLIC);
    prof_begin(ctx, QD_PHASE_GENERATE);
    QD_TRY(user_cb(ur.generate(ur.self, &pv, ctx->cb.as<u64>(), C, max_child, ctx->uscr.as<u8>(), slice * threads, H0,
                               reinterpret_cast<double *>(M0), csize, (void *)s),
                   "generate"));
    if (b1)
        hipLaunchKernelGGL(k_key_hist_top, dim3(grid_for(C, 256, 4096)), dim3(256), 0, s, (const u64 *)H0, C, b1,
                           ctx->hist1.as<u64>());
    QD_HIP(hipGetLastError());
    prof_end(ctx, QD_PHASE_GENERATE);

    // ---- dedup, memory-driven survivor cap, select, order
    u64 *UH, *UR;
    double2 *UM;
    u64 U;
    phase_mark(ctx, PH_COLL_PREP);
    QD_TRY(partition_dedup(ctx, H0, M0, nullptr, C, b1, b2, prm.tolerance, st, UH, UR, UM, U, 0, UK));
    phase_mark(ctx, PH_COLL_FINAL);
    st.num_unique = U;
    if (max_obj == 0)
        QD_TRY(mem_survivor_cap_w(ctx, in_mem, out, prm, sk, UH, UR, UM, UK, C, U, st, max_obj, [&]() {
            hipLaunchKernelGGL(k_user_slot_weight, dim3(grid_for(C)), dim3(256), 0, s, (const u64 *)UR, C,
                               (const u32 *)csize, prm.align_byte_length, kMemPerObject, ctx->mw.as<u64>());
        }));
    phase_mark(ctx, PH_TRUNC);
    const u64 W = (C + 63) / 64;
    QD_TRY(ctx->bitmap.ensure(W * sizeof(u64)));
    QD_TRY(ctx->wprefix.ensure((W + 1) * sizeof(u64)));
    QD_TRY(ctx->slist.ensure(W * sizeof(u64)));
    u64 *bitmap = ctx->bitmap.as<u64>(), *sbm = ctx->slist.as<u64>();
    prof_begin(ctx, QD_PHASE_SELECT);
    QD_HIP(hipMemsetAsync(bitmap, 0, W * sizeof(u64), s));
    QD_TRY(plain_select(ctx, C, U, max_obj, sk, UH, UR, UM, UK, bitmap, sbm));
    prof_end(ctx, QD_PHASE_SELECT);
    phase_mark(ctx, PH_PREP_FINAL);
    prof_begin(ctx, QD_PHASE_ORDER);
    QD_TRY(run_scan(ctx, LoadPopc{bitmap}, W, ctx->wprefix.as<u64>()));
    prof_end(ctx, QD_PHASE_ORDER);
    QD_TRY(sync_read(ctx, ctx->wprefix.as<u64>() + W, sizeof(u64), hb));
    const u64 K = hb[0];
    st.num_survivors = K;
    if (K == 0) return finish_empty();
    QD_TRY(out->begin.ensure((K + 1) * sizeof(u64)));
    QD_TRY(out->size.ensure(K * sizeof(u32)));
    QD_TRY(out->mag.ensure(K * sizeof(double2)));
    QD_TRY(ctx->padded.ensure(K * sizeof(u64)));
    QD_TRY(ctx->keys.ensure(K * sizeof(u64)));
    u64 *orep = ctx->keys.as<u64>();
    prof_begin(ctx, QD_PHASE_ORDER);
    hipLaunchKernelGGL(k_place, dim3(grid_for((C + 63) / 64, 256, 4096)), dim3(256), 0, s, UR, UM, (const u64 *)sbm, C,
                       bitmap, ctx->wprefix.as<u64>(), orep, out->mag.as<double2>());
    QD_HIP(hipGetLastError());
    prof_end(ctx, QD_PHASE_ORDER);

    // ---- final: sizes from the generate, bytes from the rule's populate
    // (the padding is zero)
    phase_mark(ctx, PH_FINAL);
    prof_begin(ctx, QD_PHASE_FINALIZE);
    hipLaunchKernelGGL(k_user_out_size, dim3(grid_for(K)), dim3(256), 0, s, (const u64 *)orep, K, (const u32 *)csize,
                       prm.align_byte_length, out->size.as<u32>(), ctx->padded.as<u64>());
    QD_HIP(hipGetLastError());
    QD_TRY(run_scan(ctx, LoadU64{ctx->padded.as<u64>()}, K, out->begin.as<u64>()));
    prof_end(ctx, QD_PHASE_FINALIZE);
    QD_TRY(sync_read(ctx, out->begin.as<u64>() + K, sizeof(u64), hb));
    const u64 bytes = hb[0];
    QD_TRY(ensure_relieved(ctx, out->objects, std::max<u64>(bytes, 1)));
    prof_begin(ctx, QD_PHASE_FINALIZE);
    QD_HIP(hipMemsetAsync(out->objects.p, 0, std::max<u64>(bytes, 1), s));
    QD_TRY(user_cb(ur.populate(ur.self, &pv, ctx->cb.as<u64>(), orep, K, out->begin.as<u64>(), out->objects.as<u8>(),
                               (void *)s),
                   "populate"));
    prof_end(ctx, QD_PHASE_FINALIZE);
    out->n = K;
    out->bytes = bytes;
    st.output_bytes = bytes;

    // ---- normalize
    phase_mark(ctx, PH_NORMALIZE);
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
// adjoint of one step (see qd_k_adjoint.hpp for the algebra)
// ===========================================================================
template <int KIND>
static int adjoint_kind(qd_ctx *ctx, qd_state *in, const qd_rule &rule, const qd_rule *dcoef, int ndir, qd_state *out,
                        const double2 *gout, double2 *gin, double *grad, u64 max_obj) {
    hipStream_t s = ctx->stream;
    const u64 N = in->n, K = out->n;
    release_range_cache(ctx);
    for (int t = 0; t < ndir; ++t) grad[t] = 0.0;
    if (N == 0) return QD_OK;
    if (max_obj < N)
        return fail(QD_ERR_UNSUPPORTED, "adjoint of a step whose parents were pre-truncated (max_num_object < N)");
    if (K == 0) {
        QD_HIP(hipMemsetAsync(gin, 0, N * sizeof(double2), s));
        QD_HIP(hipStreamSynchronize(s));
        return QD_OK;
    }
    if (K > 0xFFFFFFFFull) return fail(QD_ERR_CAPACITY, "adjoint output state beyond 2^32 objects");
    Parents P{in->objects.as<u8>(), in->begin.as<u64>(), in->size.as<u32>(), in->mag.as<double2>(), N};

    // per-parent prep (num_child, split_merge root tables / name folds)
    QD_TRY(ctx->nc.ensure(N * sizeof(u32)));
    QD_TRY(ctx->cb.ensure((N + 1) * sizeof(u64)));
    QD_TRY(ctx->pre.ensure(N * sizeof(u64)));
    QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
    if (KIND == QD_RULE_QCGD_SPLIT_MERGE) QD_TRY(ctx->tab.ensure((sm_tab_base(in->bytes) + kSmSlots * 64 + 16) * sizeof(u64)));
    u64 *counters = ctx->counters.as<u64>();
    QD_HIP(hipMemsetAsync(counters, 0, 64 * sizeof(u64), s));
    u32 *maxsz = reinterpret_cast<u32 *>(counters + 3);
    u32 *errw = reinterpret_cast<u32 *>(counters + 4);
    // QCGD graphs of at most 64 nodes: the fast walk's descriptors and work
    // items (as in the step); else, and for hadamard, the per-child kernels
    bool fast = KIND != QD_RULE_HADAMARD && !ctx->tune.generic_path;
    u64 nitems = 0;
    if (fast) {
        QD_TRY(ctx->pdesc.ensure(N * sizeof(PDesc)));
        QD_TRY(ctx->wi.ensure(N * sizeof(u32)));
        QD_TRY(ctx->wb.ensure((N + 1) * sizeof(u64)));
        hipLaunchKernelGGL(k_prep_fast<KIND>, dim3(grid_for(N, 16, 65536)), dim3(256), 0, s, P, (const u64 *)nullptr,
                           ctx->pdesc.as<PDesc>(), ctx->nc.as<u32>(), ctx->wi.as<u32>(), ctx->tab.as<u64>(), maxsz,
                           errw);
        QD_HIP(hipGetLastError());
        QD_TRY(run_scan(ctx, LoadU32{ctx->nc.as<u32>()}, N, ctx->cb.as<u64>()));
        QD_TRY(run_scan(ctx, LoadU32{ctx->wi.as<u32>()}, N, ctx->wb.as<u64>()));
        u64 h2[2];
        QD_HIP(hipMemcpyAsync(h2, ctx->wb.as<u64>() + N, sizeof(u64), hipMemcpyDeviceToHost, s));
        QD_HIP(hipMemcpyAsync(h2 + 1, counters + 4, sizeof(u64), hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        nitems = h2[0];
        if (u32(h2[1]) & ERR_WIDE) {  // a graph wider than 64 nodes
            fast = false;
            QD_HIP(hipMemsetAsync(counters + 3, 0, 2 * sizeof(u64), s));
        } else if (nitems) {
            const u64 nchunk = (nitems + kGenThreads - 1) / kGenThreads;
            QD_TRY(ensure_relieved(ctx, ctx->cpar, nchunk * sizeof(u64)));
            hipLaunchKernelGGL(k_chunk_parent, dim3(grid_for(nchunk)), dim3(256), 0, s, ctx->wb.as<u64>(), N, nchunk,
                               ctx->cpar.as<u64>());
            QD_HIP(hipGetLastError());
        }
    }
    if (!fast) {
        hipLaunchKernelGGL(k_prep<KIND>, dim3(grid_for(N)), dim3(256), 0, s, P, (const u64 *)nullptr, rule,
                           ctx->nc.as<u32>(), ctx->pre.as<u64>(), ctx->tab.as<u64>(), maxsz, errw);
        QD_HIP(hipGetLastError());
        QD_TRY(run_scan(ctx, LoadU32{ctx->nc.as<u32>()}, N, ctx->cb.as<u64>()));
    }

    // g_y from the cotangent of the normalised output
    QD_TRY(ctx->adjd.ensure(K * sizeof(double)));
    QD_TRY(ctx->adjg.ensure(K * sizeof(double2)));
    hipLaunchKernelGGL(k_adj_dot, dim3(grid_for(K)), dim3(256), 0, s, gout, out->mag.as<double2>(), K,
                       ctx->adjd.as<double>());
    QD_HIP(hipGetLastError());
    double alpha = 0;
    QD_TRY(canonical_sum(ctx, nullptr, ctx->adjd.as<double>(), K, &alpha));
    const double sn = std::sqrt(out->total_proba);
    if (!(sn > 0)) return fail(QD_ERR_INVALID, "output state has zero total probability");
    hipLaunchKernelGGL(k_adj_gy, dim3(grid_for(K)), dim3(256), 0, s, gout, out->mag.as<double2>(), K, alpha, sn,
                       ctx->adjg.as<double2>());
    QD_HIP(hipGetLastError());

    // hash -> output index table, load factor <= 1/2
    u64 T = 1024;
    while (T < 2 * K) T <<= 1;
    QD_TRY(ctx->adjk.ensure(T * sizeof(u64)));
    QD_TRY(ctx->adjv.ensure(T * sizeof(u32)));
    QD_HIP(hipMemsetAsync(ctx->adjk.p, 0xFF, T * sizeof(u64), s));
    u64 *special = counters + 8;
    hipLaunchKernelGGL(k_adj_table<KIND>, dim3(grid_for(K)), dim3(256), 0, s, out->objects.as<u8>(),
                       out->begin.as<u64>(), out->size.as<u32>(), K, rule, ctx->adjk.as<u64>(), ctx->adjv.as<u32>(),
                       T - 1, special);
    QD_HIP(hipGetLastError());
    u64 hb[16];
    QD_TRY(sync_read(ctx, counters, 16 * sizeof(u64), hb));
    QD_TRY(check_err_word(u32(hb[4])));

    // gather over all candidates
    const int nd = std::max(ndir, 0);
    QD_TRY(ctx->adjp.ensure(std::max<u64>(u64(nd) * N, 1) * sizeof(double)));
    const DualCoef dk = dual_coef_of(rule, dcoef, nd);
    if (fast) {  // one thread per work item, then per-parent sums in order
        QD_TRY(ctx->adjw.ensure(std::max<u64>(nitems, 1) * kAdjPart * sizeof(double)));
        const unsigned gg =
            unsigned(std::max<u64>(std::min<u64>((nitems + kGenThreads - 1) / kGenThreads, ctx->tune.generate_grid), 1));
        if (nitems)
            hipLaunchKernelGGL(k_adjoint_fast<KIND>, dim3(gg), dim3(kGenThreads), 0, s, ctx->pdesc.as<PDesc>(), P.mag,
                               ctx->wb.as<u64>(), N, ctx->cpar.as<u64>(), ctx->tab.as<u64>(), P.begin, dk, nd,
                               ctx->adjk.as<u64>(), ctx->adjv.as<u32>(), T - 1, hb[8], ctx->adjg.as<double2>(), nitems,
                               ctx->adjw.as<double>());
        QD_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_adjoint_reduce, dim3(grid_for(N)), dim3(256), 0, s, ctx->wb.as<u64>(), N,
                           ctx->adjw.as<double>(), nd, gin, ctx->adjp.as<double>());
    } else {  // one wavefront per parent
        hipLaunchKernelGGL(k_adjoint<KIND>, dim3(grid_for(N, 4, 65536)), dim3(256), 0, s, P, ctx->cb.as<u64>(), rule,
                           dk, nd, ctx->pre.as<u64>(), ctx->tab.as<u64>(), ctx->adjk.as<u64>(), ctx->adjv.as<u32>(),
                           T - 1, hb[8], ctx->adjg.as<double2>(), gin, ctx->adjp.as<double>());
    }
    QD_HIP(hipGetLastError());
    for (int t = 0; t < nd; ++t) QD_TRY(canonical_sum(ctx, nullptr, ctx->adjp.as<double>() + u64(t) * N, N, &grad[t]));
    QD_HIP(hipStreamSynchronize(s));
    return QD_OK;
}

// ===========================================================================
// batched step over instance chunks
// ===========================================================================

// Candidates per instance of the whole batch (prep + scan only): cseg;
// wide[i] = 1 for the instances holding graphs beyond the batched kernels'
// 64 nodes.
template <int KIND>
static int batch_plan(qd_ctx *ctx, const Parents &P, u64 in_bytes, const qd_rule &rule, Batch *bt,
                      std::vector<u32> &wide) {
    hipStream_t s = ctx->stream;
    const u64 N = P.n;
    wide.assign(bt->n, 0);
    if (N == 0) {
        bt->cseg.assign(bt->n + 1, 0);
        return QD_OK;
    }
    QD_TRY(ctx->nc.ensure(N * sizeof(u32)));
    QD_TRY(ctx->cb.ensure((N + 1) * sizeof(u64)));
    QD_TRY(ctx->counters.ensure(64 * sizeof(u64)));
    QD_TRY(ctx->btrunc.ensure(bt->n * sizeof(u32)));
    (void)in_bytes;
    u64 *counters = ctx->counters.as<u64>();
    QD_HIP(hipMemsetAsync(counters, 0, 64 * sizeof(u64), s));
    QD_HIP(hipMemsetAsync(ctx->btrunc.p, 0, bt->n * sizeof(u32), s));
    u32 *errw = reinterpret_cast<u32 *>(counters + 4);
    hipLaunchKernelGGL(k_num_child_fast<KIND>, dim3(grid_for(N, 16, 65536)), dim3(256), 0, s, P, ctx->nc.as<u32>(),
                       errw);
    hipLaunchKernelGGL(k_wide_flags, dim3(grid_for(N)), dim3(256), 0, s, P.obj, P.begin, P.size, N, bt->d_oseg, bt->n,
                       ctx->btrunc.as<u32>());
    QD_HIP(hipGetLastError());
    QD_TRY(run_scan(ctx, LoadU32{ctx->nc.as<u32>()}, N, ctx->cb.as<u64>()));
    u64 hb[2];
    QD_TRY(sync_read(ctx, counters + 3, 2 * sizeof(u64), hb));
    if (u32(hb[1]) & ERR_WIDE) {  // a graph beyond 64 nodes: every count one thread per parent
        QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), s));
        hipLaunchKernelGGL(k_num_child<KIND>, dim3(grid_for(N)), dim3(256), 0, s, P, rule, ctx->nc.as<u32>(), errw);
        QD_HIP(hipGetLastError());
        QD_TRY(run_scan(ctx, LoadU32{ctx->nc.as<u32>()}, N, ctx->cb.as<u64>()));
        QD_TRY(sync_read(ctx, counters + 3, 2 * sizeof(u64), hb));
    }
    QD_TRY(check_err_word(u32(hb[1])));
    QD_TRY(sync_read(ctx, ctx->btrunc.p, bt->n * sizeof(u32), wide.data()));
    return batch_cseg(ctx, bt);
}

template <int KIND>
static int batched_kind(qd_ctx *c, qd_state *in, const u64 *in_seg, u32 n, const qd_rule *rules, qd_state *out,
                        u64 *out_seg, const u64 *caps, const qd_params &prm, qd_step_stats *stats) {
    hipStream_t s = c->stream;
    const Parents P{in->objects.as<u8>(), in->begin.as<u64>(), in->size.as<u32>(), in->mag.as<double2>(), in->n};
    QD_TRY(relieve_memory(c));
    QD_TRY(upload(c, c->boseg, in_seg, (n + 1) * sizeof(u64)));
    QD_TRY(upload(c, c->brules, rules, n * sizeof(qd_rule)));
    Batch plan;
    plan.n = n;
    plan.d_oseg = c->boseg.as<u64>();
    std::vector<u32> wide;
    prof_begin(c, QD_PHASE_PREP);
    const int prc = batch_plan<KIND>(c, P, in->bytes, rules[0], &plan, wide);
    prof_end(c, QD_PHASE_PREP);
    QD_TRY(prc);
    // chunks of consecutive instances within the candidate budget; an
    // instance beyond the budget or with wide graphs steps alone
    const u64 budget = step_budget(c);
    std::vector<u32> cuts{0};
    u64 acc = 0;
    for (u32 i = 0; i < n; ++i) {
        const u64 ci = plan.cseg[i + 1] - plan.cseg[i];
        const bool solo = wide[i] || ci > budget;
        if (acc && (solo || acc + ci > budget)) {
            cuts.push_back(i);
            acc = 0;
        }
        acc += ci;
        if (solo && i + 1 < n) {
            cuts.push_back(i + 1);
            acc = 0;
        }
    }
    cuts.push_back(n);
    const size_t nchunk = cuts.size() - 1;

    std::vector<u64> U(n, 0), kseg(n + 1, 0);
    std::vector<double> total(n, 0.0);
    qd_step_stats st;
    std::memset(&st, 0, sizeof st);
    // chunk outputs live in buffers the context keeps between calls
    if (nchunk > 1 && c->parts.size() < 4 * nchunk) c->parts.resize(4 * nchunk);
    std::vector<qd_state> partv(nchunk > 1 ? nchunk : 0);
    for (size_t k = 0; k < partv.size(); ++k) {
        partv[k].ctx = c;
        partv[k].objects = c->parts[4 * k];
        partv[k].begin = c->parts[4 * k + 1];
        partv[k].size = c->parts[4 * k + 2];
        partv[k].mag = c->parts[4 * k + 3];
    }
    std::vector<qd_state *> parts;
    auto give_back = [&](size_t k) {  // the (possibly regrown) buffers of chunk k
        c->parts[4 * k] = partv[k].objects;
        c->parts[4 * k + 1] = partv[k].begin;
        c->parts[4 * k + 2] = partv[k].size;
        c->parts[4 * k + 3] = partv[k].mag;
    };
    auto cleanup = [&]() { parts.clear(); };
    for (size_t k = 0; k < nchunk; ++k) {
        const u32 i0 = cuts[k], i1 = cuts[k + 1];
        const u64 p0 = in_seg[i0], p1 = in_seg[i1];
        std::vector<u64> sub(i1 - i0 + 1);
        for (u32 i = i0; i <= i1; ++i) sub[i - i0] = in_seg[i] - p0;
        int rc = upload(c, c->boseg, sub.data(), sub.size() * sizeof(u64));
        if (rc) {
            cleanup();
            return rc;
        }
        Batch bt;
        bt.n = i1 - i0;
        bt.d_oseg = c->boseg.as<u64>();
        bt.d_rules = c->brules.as<qd_rule>() + i0;
        bt.caps = caps + i0;
        bt.sk = SelKey{prm.simple_truncation ? 0u : 1u, 0u, prm.seed};
        const Parents Pk{P.obj, P.begin + p0, P.size + p0, P.mag + p0, p1 - p0};
        qd_state *dst = out;
        if (nchunk > 1) {
            dst = &partv[k];
            parts.push_back(dst);
        }
        if (bt.n == 1) {
            // one instance alone: its plain step (which runs in hash-range
            // passes when it exceeds the budget by itself)
            rc = step_impl<KIND, false>(c, Pk, in->bytes, 0, rules[i0], nullptr, dst, nullptr, caps[i0], prm, st);
            bt.U.assign(1, st.num_unique);
            bt.kseg = {0, dst->n};
            bt.total.assign(1, dst->total_proba);
        } else {
            rc = step_impl<KIND, true>(c, Pk, in->bytes, 0, rules[i0], &bt, dst, nullptr, UINT64_MAX, prm, st);
        }
        if (nchunk > 1) give_back(k);
        if (rc) {
            cleanup();
            return rc;
        }
        const u64 kbase = kseg[i0];
        for (u32 i = i0; i < i1; ++i) {
            U[i] = bt.U[i - i0];
            total[i] = bt.total[i - i0];
            kseg[i + 1] = kbase + bt.kseg[i - i0 + 1];
        }
    }
    if (nchunk > 1) {  // back to back into out
        u64 K = 0, bytes = 0;
        for (qd_state *t : parts) {
            K += t->n;
            bytes += t->bytes;
        }
        int rc = state_reserve(out, K, bytes);
        if (rc) {
            cleanup();
            return rc;
        }
        u64 kb = 0, bb = 0;
        double sum = 0;
        for (qd_state *t : parts) {
            if (t->n) {
                QD_HIP(hipMemcpyAsync(out->objects.as<u8>() + bb, t->objects.p, t->bytes, hipMemcpyDeviceToDevice, s));
                QD_HIP(hipMemcpyAsync(out->size.as<u32>() + kb, t->size.p, t->n * sizeof(u32), hipMemcpyDeviceToDevice,
                                      s));
                QD_HIP(hipMemcpyAsync(out->mag.as<double2>() + kb, t->mag.p, t->n * sizeof(double2),
                                      hipMemcpyDeviceToDevice, s));
                hipLaunchKernelGGL(k_add_offset, dim3(grid_for(t->n + 1)), dim3(256), 0, s, t->begin.as<u64>(), t->n + 1,
                                   bb, out->begin.as<u64>() + kb);
                QD_HIP(hipGetLastError());
            }
            kb += t->n;
            bb += t->bytes;
            sum += t->total_proba;
        }
        if (K == 0) QD_HIP(hipMemsetAsync(out->begin.p, 0, sizeof(u64), s));
        QD_HIP(hipStreamSynchronize(s));
        out->n = K;
        out->bytes = bytes;
        out->total_proba = sum;
        cleanup();
    }
    for (u32 i = 0; i <= n; ++i) out_seg[i] = kseg[i];
    if (stats) {
        // output bytes per instance: object_begin at the instance boundaries
        std::vector<u64> ob(n + 1, 0);
        if (out->n) {
            QD_TRY(c->bcount.ensure((n + 1) * sizeof(u64)));
            QD_TRY(upload(c, c->bchunk, kseg.data(), (n + 1) * sizeof(u64)));
            hipLaunchKernelGGL(k_gather_u64, dim3(grid_for(n + 1)), dim3(256), 0, s, out->begin.as<u64>(),
                               c->bchunk.as<u64>(), n + 1, c->bcount.as<u64>());
            QD_HIP(hipGetLastError());
            QD_TRY(sync_read(c, c->bcount.p, (n + 1) * sizeof(u64), ob.data()));
        }
        for (u32 i = 0; i < n; ++i) {
            qd_step_stats &x = stats[i];
            std::memset(&x, 0, sizeof x);
            x.num_parent = x.num_parent_kept = in_seg[i + 1] - in_seg[i];
            x.num_candidates = plan.cseg[i + 1] - plan.cseg[i];
            x.num_unique = U[i];
            x.num_survivors = kseg[i + 1] - kseg[i];
            x.output_bytes = ob[i + 1] - ob[i];
            x.total_proba = total[i];
            x.dedup_levels = st.dedup_levels;
            x.reserved = st.reserved | (nchunk > 1 ? 2u : 0u);
        }
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
    c->set_names();
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
    // the device address of the no-allocation scratch on this device (the
    // device is current)
    e = hipGetSymbolAddress(reinterpret_cast<void **>(&c->noalloc), HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipHostFree(c->host);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(QD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    *out = c;
    return QD_OK;
}

int qd_ctx_destroy(qd_ctx *c) {
    if (!c) return QD_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (qd_ctx::SelGraph &g : c->selgraph)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (c->seggraph.exec) (void)hipGraphExecDestroy(c->seggraph.exec);
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

int qd_ctx_stream(qd_ctx *c, void **stream) {
    if (!c || !stream) return fail(QD_ERR_INVALID, "null argument");
    *stream = (void *)c->stream;
    return QD_OK;
}

int qd_ctx_scratch_bytes(qd_ctx *c, uint64_t *bytes) {
    if (!c || !bytes) return fail(QD_ERR_INVALID, "null argument");
    u64 t = 0;
    for (DBuf *b : c->all())
        if (!b->view) t += b->cap;
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
        if (t->dedup_cap != 256 && t->dedup_cap != 512 && t->dedup_cap != 1024 && t->dedup_cap != 2048)
            return fail(QD_ERR_INVALID, "dedup_cap must be 256, 512, 1024 or 2048");
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
    n.generic_path = t->generic_path;  // 0 = fast kernels where they apply
    if (t->chunk_log2 && (t->chunk_log2 < 10 || t->chunk_log2 > 40))
        return fail(QD_ERR_INVALID, "chunk_log2 must be 0 or in [10, 40]");
    n.chunk_log2 = t->chunk_log2;  // 0 = from free device memory
    n.split_generate = t->split_generate ? 1 : 0;
    n.eager_select = t->eager_select ? 1 : 0;
    n.free_mem_mib = t->free_mem_mib;  // 0 = the device's free memory
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

int qd_copy_probe(qd_ctx *c, uint64_t bytes, int reps, double *gbps) {
    if (!c || !gbps || bytes < 16 || reps < 1) return fail(QD_ERR_INVALID, "bad argument");
    QD_HIP(hipSetDevice(c->device));
    const u64 n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    QD_HIP(hipMalloc(&a, n * 16));
    hipError_t e = hipMalloc(&b, n * 16);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(QD_ERR_OOM, "hipMalloc of the probe buffers failed");
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemsetAsync(a, 1, n * 16, c->stream);
    float best = 0;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, c->stream);
        hipLaunchKernelGGL(k_copy_probe, dim3(grid_for(n, 256, 16384)), dim3(256), 0, c->stream,
                           static_cast<const uint4 *>(a), static_cast<uint4 *>(b), n);
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r == 0 || ms < best) best = ms;
    }
    e = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(QD_ERR_HIP, std::string("copy probe: ") + hipGetErrorString(e));
    *gbps = 2.0 * double(n * 16) / (double(best) * 1e6);
    return QD_OK;
}

int qd_state_create(qd_ctx *c, qd_state **out) {
    if (!c || !out) return fail(QD_ERR_INVALID, "null argument");
    qd_state *s = new (std::nothrow) qd_state;
    if (!s) return fail(QD_ERR_OOM, "host allocation failed");
    s->ctx = c;
    s->objects.name = "state objects";
    s->begin.name = "state begin";
    s->size.name = "state size";
    s->mag.name = "state mag";
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

int qd_state_copy(qd_state *dst, const qd_state *src) {
    if (!dst || !src) return fail(QD_ERR_INVALID, "null argument");
    if (dst == src) return QD_OK;
    if (dst->ctx != src->ctx) return fail(QD_ERR_INVALID, "states belong to different contexts");
    qd_ctx *c = dst->ctx;
    QD_HIP(hipSetDevice(c->device));
    const u64 n = src->n, bytes = src->bytes;
    QD_TRY(state_reserve(dst, n, bytes));
    hipStream_t s = c->stream;
    if (bytes) QD_HIP(hipMemcpyAsync(dst->objects.p, src->objects.p, bytes, hipMemcpyDeviceToDevice, s));
    QD_HIP(hipMemcpyAsync(dst->begin.p, src->begin.p, (n + 1) * sizeof(u64), hipMemcpyDeviceToDevice, s));
    if (n) {
        QD_HIP(hipMemcpyAsync(dst->size.p, src->size.p, n * sizeof(u32), hipMemcpyDeviceToDevice, s));
        QD_HIP(hipMemcpyAsync(dst->mag.p, src->mag.p, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    }
    QD_HIP(hipStreamSynchronize(s));
    dst->n = n;
    dst->bytes = bytes;
    dst->total_proba = src->total_proba;
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
    qd_params prm{1e-30, 8, 1, 0, 0, 0.2, 0.25};
    if (params) prm = *params;
    qd_step_stats st;
    std::memset(&st, 0, sizeof st);
    QD_HIP(hipSetDevice(c->device));
    c->next_phase = 0;
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
    if (rc == QD_OK) phase_mark(c, PH_END);
    if (stats) *stats = st;
    return rc;
}

int qd_simulate_user(qd_ctx *c, qd_state *in, const qd_user_rule *rule, qd_state *out, qd_symbolic *sy,
                     uint64_t max_num_object, const qd_params *params, qd_step_stats *stats) {
    if (!c || !in || !rule || !out) return fail(QD_ERR_INVALID, "null argument");
    if (!rule->num_child || !rule->generate || !rule->populate || !rule->hash_objects)
        return fail(QD_ERR_INVALID, "user rule: a callback is missing");
    if (in == out) return fail(QD_ERR_INVALID, "input and output states must differ");
    if (in->ctx != c || out->ctx != c) return fail(QD_ERR_INVALID, "state belongs to another context");
    qd_params prm{1e-30, 8, 1, 0, 0, 0.2, 0.25};
    if (params) prm = *params;
    qd_step_stats st;
    std::memset(&st, 0, sizeof st);
    QD_HIP(hipSetDevice(c->device));
    c->next_phase = 0;
    const int rc = user_step(c, in, *rule, out, sy, max_num_object, prm, st);
    if (rc == QD_OK) phase_mark(c, PH_END);
    if (stats) *stats = st;
    return rc;
}

int qd_ctx_set_phase_hook(qd_ctx *c, qd_phase_hook hook, void *user) {
    if (!c) return fail(QD_ERR_INVALID, "null ctx");
    c->hook = hook;
    c->hook_user = hook ? user : nullptr;
    return QD_OK;
}

int qd_apply_modifier(qd_ctx *c, qd_state *s, const qd_modifier *m) {
    if (!c || !s || !m) return fail(QD_ERR_INVALID, "null argument");
    if (!modifier_known(m->kind)) return fail(QD_ERR_UNSUPPORTED, "modifier kind has no device implementation");
    QD_HIP(hipSetDevice(c->device));
    if (s->n == 0) return QD_OK;
    QD_TRY(c->counters.ensure(64 * sizeof(u64)));
    u32 *errw = reinterpret_cast<u32 *>(c->counters.as<u64>() + 4);
    QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), c->stream));
    if (m->kind == QD_MOD_QCGD_STEP || m->kind == QD_MOD_QCGD_REVERSED_STEP)
        hipLaunchKernelGGL(k_step_modifier, dim3(grid_for(s->n, 16, 65536)), dim3(256), 0, c->stream,
                           s->objects.as<u8>(), s->begin.as<u64>(), s->size.as<u32>(), s->n, *m, errw);
    else
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

int qd_simulate_batched(qd_ctx *c, qd_state *in, const uint64_t *in_seg, uint32_t num_instances,
                        const qd_rule *rules, qd_state *out, uint64_t *out_seg, const uint64_t *max_num_object,
                        const qd_params *params, qd_step_stats *stats) {
    if (!c || !in || !in_seg || !rules || !out || !out_seg || !max_num_object)
        return fail(QD_ERR_INVALID, "null argument");
    if (in == out) return fail(QD_ERR_INVALID, "input and output states must differ");
    if (in->ctx != c || out->ctx != c) return fail(QD_ERR_INVALID, "state belongs to another context");
    const u32 n = num_instances;
    if (n == 0 || n > kMaxInst) return fail(QD_ERR_INVALID, "num_instances must be in [1, 4096]");
    if (in_seg[0] != 0 || in_seg[n] != in->n) return fail(QD_ERR_INVALID, "in_seg must run from 0 to the state size");
    const int kind = rules[0].kind;
    if (kind != QD_RULE_QCGD_ERASE_CREATE && kind != QD_RULE_QCGD_COIN && kind != QD_RULE_QCGD_SPLIT_MERGE)
        return fail(QD_ERR_UNSUPPORTED, "the batched step runs the QCGD rules");
    for (u32 i = 0; i < n; ++i) {
        if (in_seg[i + 1] < in_seg[i]) return fail(QD_ERR_INVALID, "in_seg must be ascending");
        if (rules[i].kind != kind) return fail(QD_ERR_INVALID, "all instances of a batch run the same rule kind");
        if (max_num_object[i] == 0)
            return fail(QD_ERR_UNSUPPORTED, "memory-driven truncation (max_num_object 0) is not batched");
        if (in_seg[i + 1] - in_seg[i] > max_num_object[i])
            return fail(QD_ERR_UNSUPPORTED, "batched instances cannot be pre-truncated (more parents than max)");
    }
    qd_params prm{1e-30, 8, 1, 0, 0, 0.2, 0.25};
    if (params) prm = *params;
    QD_HIP(hipSetDevice(c->device));
    c->next_phase = 0;
    switch (kind) {
    case QD_RULE_QCGD_ERASE_CREATE:
        return batched_kind<QD_RULE_QCGD_ERASE_CREATE>(c, in, in_seg, n, rules, out, out_seg, max_num_object, prm, stats);
    case QD_RULE_QCGD_COIN:
        return batched_kind<QD_RULE_QCGD_COIN>(c, in, in_seg, n, rules, out, out_seg, max_num_object, prm, stats);
    default:
        return batched_kind<QD_RULE_QCGD_SPLIT_MERGE>(c, in, in_seg, n, rules, out, out_seg, max_num_object, prm,
                                                      stats);
    }
}

int qd_simulate_adjoint(qd_ctx *c, qd_state *in, const qd_rule *rule, const qd_rule *dcoef, int ndir, qd_state *out,
                        const double *out_cot, double *in_cot, double *grad, uint64_t max_num_object) {
    if (!c || !in || !rule || !out) return fail(QD_ERR_INVALID, "null argument");
    if (ndir < 0 || ndir > kMaxDir) return fail(QD_ERR_INVALID, "ndir must be in [0, 3]");
    if (ndir > 0 && (!dcoef || !grad)) return fail(QD_ERR_INVALID, "null derivative rules or gradient");
    if (in->n && !in_cot) return fail(QD_ERR_INVALID, "null input cotangent");
    if (out->n && !out_cot) return fail(QD_ERR_INVALID, "null output cotangent");
    if (in == out) return fail(QD_ERR_INVALID, "input and output states must differ");
    if (in->ctx != c || out->ctx != c) return fail(QD_ERR_INVALID, "state belongs to another context");
    for (int t = 0; t < ndir; ++t)
        if (dcoef[t].kind != rule->kind) return fail(QD_ERR_INVALID, "derivative rule of another kind");
    QD_HIP(hipSetDevice(c->device));
    const double2 *go = reinterpret_cast<const double2 *>(out_cot);
    double2 *gi = reinterpret_cast<double2 *>(in_cot);
    double dummy[kMaxDir];
    double *gr = grad ? grad : dummy;
    switch (rule->kind) {
    case QD_RULE_HADAMARD:
        return adjoint_kind<QD_RULE_HADAMARD>(c, in, *rule, dcoef, 0, out, go, gi, gr, max_num_object);
    case QD_RULE_QCGD_ERASE_CREATE:
        return adjoint_kind<QD_RULE_QCGD_ERASE_CREATE>(c, in, *rule, dcoef, ndir, out, go, gi, gr, max_num_object);
    case QD_RULE_QCGD_COIN:
        return adjoint_kind<QD_RULE_QCGD_COIN>(c, in, *rule, dcoef, ndir, out, go, gi, gr, max_num_object);
    case QD_RULE_QCGD_SPLIT_MERGE:
        return adjoint_kind<QD_RULE_QCGD_SPLIT_MERGE>(c, in, *rule, dcoef, ndir, out, go, gi, gr, max_num_object);
    }
    return fail(QD_ERR_UNSUPPORTED, "rule kind has no device implementation");
}

int qd_modifier_adjoint(qd_ctx *c, qd_state *after, const qd_modifier *m, double *cot) {
    if (!c || !after || !m) return fail(QD_ERR_INVALID, "null argument");
    if (!modifier_known(m->kind)) return fail(QD_ERR_UNSUPPORTED, "modifier kind has no device implementation");
    if (after->ctx != c) return fail(QD_ERR_INVALID, "state belongs to another context");
    if (after->n == 0) return QD_OK;
    if (!cot) return fail(QD_ERR_INVALID, "null cotangent");
    QD_HIP(hipSetDevice(c->device));
    QD_TRY(c->counters.ensure(64 * sizeof(u64)));
    u32 *errw = reinterpret_cast<u32 *>(c->counters.as<u64>() + 4);
    QD_HIP(hipMemsetAsync(errw, 0, sizeof(u32), c->stream));
    hipLaunchKernelGGL(k_modifier_adjoint, dim3(grid_for(after->n)), dim3(256), 0, c->stream,
                       after->objects.as<u8>(), after->begin.as<u64>(), after->size.as<u32>(), after->n, *m,
                       reinterpret_cast<double2 *>(cot), errw);
    QD_HIP(hipGetLastError());
    u32 w = 0;
    QD_TRY(sync_read(c, errw, sizeof(u32), &w));
    return check_err_word(w);
}

int qd_average_value(qd_ctx *c, qd_state *s, int observable, int arg, double *value, double *cot) {
    if (!c || !s || !value) return fail(QD_ERR_INVALID, "null argument");
    if (observable < 0 || observable > 2) return fail(QD_ERR_INVALID, "unknown observable");
    if (s->ctx != c) return fail(QD_ERR_INVALID, "state belongs to another context");
    QD_HIP(hipSetDevice(c->device));
    *value = 0.0;
    if (s->n == 0) return QD_OK;
    QD_TRY(c->adjd.ensure(s->n * sizeof(double)));
    hipLaunchKernelGGL(k_observable, dim3(grid_for(s->n)), dim3(256), 0, c->stream, s->objects.as<u8>(),
                       s->begin.as<u64>(), s->size.as<u32>(), s->mag.as<double2>(), s->n, observable, arg,
                       c->adjd.as<double>(), reinterpret_cast<double2 *>(cot));
    QD_HIP(hipGetLastError());
    return canonical_sum(c, nullptr, c->adjd.as<double>(), s->n, value);
}

int qd_average_value_of(qd_ctx *c, qd_state *s, const double *obs, double *value, double *cot) {
    if (!c || !s || !value || (!obs && s->n)) return fail(QD_ERR_INVALID, "null argument");
    if (s->ctx != c) return fail(QD_ERR_INVALID, "state belongs to another context");
    QD_HIP(hipSetDevice(c->device));
    *value = 0.0;
    if (s->n == 0) return QD_OK;
    QD_TRY(c->adjd.ensure(s->n * sizeof(double)));
    hipLaunchKernelGGL(k_observable_of, dim3(grid_for(s->n)), dim3(256), 0, c->stream, obs, s->mag.as<double2>(), s->n,
                       c->adjd.as<double>(), reinterpret_cast<double2 *>(cot));
    QD_HIP(hipGetLastError());
    return canonical_sum(c, nullptr, c->adjd.as<double>(), s->n, value);
}

int qd_average_value_noalloc(qd_ctx *c, qd_state *s, int observable, int arg, double *value, double *cot) {
    if (!c || !s || !value) return fail(QD_ERR_INVALID, "null argument");
    if (observable < 0 || observable > 2) return fail(QD_ERR_INVALID, "unknown observable");
    if (s->ctx != c) return fail(QD_ERR_INVALID, "state belongs to another context");
    if (s->n > kNoAllocMaxObjects)
        return fail(QD_ERR_CAPACITY, "qd_average_value_noalloc: " + std::to_string(s->n) +
                                         " objects, beyond its static scratch of " +
                                         std::to_string(kNoAllocMaxObjects) + " (use qd_average_value)");
    static_assert(kNoAllocDoubles >= kNoAllocMaxObjects + 2 * sum_scratch_half(kNoAllocMaxObjects), "scratch size");
    QD_HIP(hipSetDevice(c->device));
    *value = 0.0;
    if (s->n == 0) return QD_OK;
    double *vals = c->noalloc;  // per-object values, then the sum's partials
    hipLaunchKernelGGL(k_observable, dim3(grid_for(s->n)), dim3(256), 0, c->stream, s->objects.as<u8>(),
                       s->begin.as<u64>(), s->size.as<u32>(), s->mag.as<double2>(), s->n, observable, arg, vals,
                       reinterpret_cast<double2 *>(cot));
    QD_HIP(hipGetLastError());
    return canonical_sum(c, nullptr, vals, s->n, value, vals + kNoAllocMaxObjects);
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
