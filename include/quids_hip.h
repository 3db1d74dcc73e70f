/*
 * quids_hip.h -- C ABI of the MI355X (gfx950) QuIDS rule-application step.
 *
 * Plain C types only: opaque handles, pointers and sizes.  Every entry point
 * returns a qd_status; 0 is success.  The C++ header quids_amd/quids.hpp
 * maps non-zero codes to std::runtime_error, as the reference throws on
 * allocation failure (src/utils/vector.hpp:124-127).
 *
 * Handles are owned by the caller and freed explicitly.  A context owns one
 * HIP stream on one device; it is not thread-safe (the reference is driven
 * from one host thread too, src/quids.hpp:448).  All device work of a call is
 * ordered on the context's stream; calls that return counters synchronise.
 */
#ifndef QUIDS_HIP_H
#define QUIDS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QD_ABI_VERSION 2

typedef enum qd_status {
    QD_OK = 0,
    QD_ERR_INVALID = 1,     /* bad argument or malformed object */
    QD_ERR_OOM = 2,         /* device allocation failed */
    QD_ERR_HIP = 3,         /* a HIP runtime call failed */
    QD_ERR_CAPACITY = 4,    /* a size limit of the library was exceeded */
    QD_ERR_UNSUPPORTED = 5  /* rule / option without a device implementation */
} qd_status;

/* Rules with a device implementation (reference: subclasses of quids::rule,
 * src/quids.hpp:105-146). */
typedef enum qd_rule_kind {
    QD_RULE_HADAMARD = 1,          /* src/rules/quantum_computer.hpp:31-50 */
    QD_RULE_QCGD_ERASE_CREATE = 2, /* src/rules/qcgd.hpp:459-532 */
    QD_RULE_QCGD_COIN = 3,         /* src/rules/qcgd.hpp:534-605 */
    QD_RULE_QCGD_SPLIT_MERGE = 4   /* src/rules/qcgd.hpp:607-1036 */
} qd_rule_kind;

/* Modifiers with a device implementation (reference: quids::modifier_t,
 * src/quids.hpp:86, applied by simulate(it, modifier), 436-438). */
typedef enum qd_modifier_kind {
    QD_MOD_CNOT = 1,               /* src/rules/quantum_computer.hpp:25-29 */
    QD_MOD_XGATE = 2,              /* src/rules/quantum_computer.hpp:52-56 */
    QD_MOD_YGATE = 3,              /* src/rules/quantum_computer.hpp:58-66 */
    QD_MOD_ZGATE = 4,              /* src/rules/quantum_computer.hpp:68-75 */
    QD_MOD_QCGD_STEP = 5,          /* src/rules/qcgd.hpp:443-449 */
    QD_MOD_QCGD_REVERSED_STEP = 6  /* src/rules/qcgd.hpp:451-457 */
} qd_modifier_kind;

/* Rule descriptor.  Coefficients are computed on the host exactly as the
 * reference constructors do: hadamard uses coef = 1/std::sqrt(2)
 * (quantum_computer.hpp:41); QCGD rules use do_ = polar(sin(theta), phi),
 * do_not = polar(cos(theta), xi) and their conjugates (qcgd.hpp:466-471,
 * 541-546, 614-619). */
typedef struct qd_rule {
    int32_t kind;        /* qd_rule_kind */
    uint32_t bit;        /* hadamard: qubit (byte) index */
    double coef;         /* hadamard: 1/sqrt(2) */
    double do_re, do_im;
    double do_not_re, do_not_im;
    double do_conj_re, do_conj_im;
    double do_not_conj_re, do_not_conj_im;
} qd_rule;

typedef struct qd_modifier {
    int32_t kind;         /* qd_modifier_kind */
    uint32_t control_bit; /* cnot: control byte index */
    uint32_t bit;         /* target byte index */
    uint32_t reserved;
} qd_modifier;

/* Globals of the reference that change results (src/quids.hpp:59-75). */
typedef struct qd_params {
    double tolerance;           /* quids::tolerance, default 1e-30 */
    uint32_t align_byte_length; /* quids::align_byte_length, default 8 */
    uint32_t simple_truncation; /* quids::simple_truncation (default 1): keep the largest |m|^2;
                                   0: keep the smallest u / |m|^2 (quids.hpp:594-608, 829-845) with
                                   u uniform in (0, 1] from Philox2x64-10 keyed by (object hash,
                                   seed) -- reproducible, unlike the reference's rand() streams */
    uint32_t reserved;
    uint64_t seed;              /* key of the probabilistic truncation's random draws */
    /* memory-driven truncation (max_num_object == 0), src/quids.hpp:64-66, 459-536 */
    double safety_margin;       /* quids::safety_margin, default 0.2 (target 80 % of the memory) */
    double equalize_factor;     /* quids::equalize_factor, default 0.25 */
} qd_params;

/* Counters of one simulate() call (reference: sy_it.num_object and
 * sy_it.num_object_after_interferences, src/quids.hpp:343-346;
 * it.num_object and it.total_proba, 151-154). */
typedef struct qd_step_stats {
    uint64_t num_parent;        /* objects of the input state */
    uint64_t num_parent_kept;   /* parents left after pre-truncation */
    uint64_t num_candidates;    /* C = sy_it.num_object */
    uint64_t num_unique;        /* U = sy_it.num_object_after_interferences */
    uint64_t num_survivors;     /* K = next.num_object */
    uint64_t output_bytes;      /* padded object bytes of the next state */
    double total_proba;         /* next.total_proba (before normalisation) */
    uint32_t dedup_levels;      /* radix partition passes used by the dedup */
    uint32_t reserved;          /* bit 0: the second pass fell back to exact offsets;
                                   bit 1: a batched step ran in several chunks;
                                   bit 2: the step ran in hash-range passes;
                                   bit 3: the fused first radix pass spilled and was redone, or
                                   a hash-range pass took two walks */
    /* memory-driven truncation only (max_num_object == 0; 0 otherwise): the
     * target_memory of the parent side and of the survivor side
     * (src/quids.hpp:469, 520) that the binary searches kept below */
    uint64_t mem_target_parents;
    uint64_t mem_target_survivors;
} qd_step_stats;

typedef struct qd_ctx qd_ctx;
typedef struct qd_state qd_state;
typedef struct qd_symbolic qd_symbolic;

/* ---- library ---------------------------------------------------------- */
int qd_abi_version(void);
const char *qd_status_string(int status);
/* Last error message recorded on this thread (never NULL). */
const char *qd_last_error(void);

/* ---- context (no reference counterpart: owns device + stream) --------- */
int qd_ctx_create(qd_ctx **ctx, int device);
int qd_ctx_destroy(qd_ctx *ctx);
int qd_ctx_synchronize(qd_ctx *ctx);
/* The context's HIP stream (a hipStream_t): work the caller enqueues on it
 * runs in order with the library's (user rules' and observables' kernels). */
int qd_ctx_stream(qd_ctx *ctx, void **stream);
/* Device-side bytes currently held by the context's scratch pool. */
int qd_ctx_scratch_bytes(qd_ctx *ctx, uint64_t *bytes);
/* Release the context's cached scratch buffers. */
int qd_ctx_trim(qd_ctx *ctx);

/* Phase timing of simulate() with HIP events on the context's stream (the
 * reference's only instrumentation is the mid_step_function hook called at
 * phase boundaries, src/quids.hpp:90, 454-990).  Times accumulate over calls
 * until reset; a phase's time spans its kernels only (no host waits). */
typedef enum qd_phase {
    QD_PHASE_PREP = 0,      /* num_child + per-parent tables */
    QD_PHASE_SCAN = 1,      /* child_begin prefix */
    QD_PHASE_GENERATE = 2,  /* child hash + magnitude */
    QD_PHASE_PARTITION = 3, /* radix passes on the hash */
    QD_PHASE_DEDUP = 4,     /* interference + tolerance */
    QD_PHASE_SELECT = 5,    /* top-K truncation */
    QD_PHASE_ORDER = 6,     /* survivor ranks */
    QD_PHASE_FINALIZE = 7,  /* sizes, offsets, survivor bytes */
    QD_PHASE_NORMALIZE = 8, /* sum |m|^2 and scale */
    QD_NUM_PHASES = 9
} qd_phase;
int qd_ctx_set_profiling(qd_ctx *ctx, int enable);

/* Phase hook of qd_simulate (the reference's mid_step_function,
 * src/quids.hpp:90, 448): called on the calling thread with the reference's
 * phase names in its order -- "num_child", "truncate_symbolic - prepare",
 * "truncate_symbolic", "prepare_index", "symbolic_iteration",
 * "compute_collisions - prepare" / "- insert" / "- finalize",
 * "truncate - prepare", "truncate", "prepare_final", "final", "normalize",
 * "end" (src/quids.hpp:454-1016) -- each once per call, after the device has
 * finished the work of the phases before it (so a hook costs a stream
 * synchronisation per phase).  NULL removes the hook. */
typedef void (*qd_phase_hook)(const char *phase, void *user);
int qd_ctx_set_phase_hook(qd_ctx *ctx, qd_phase_hook hook, void *user);
int qd_ctx_reset_phase_times(qd_ctx *ctx);
/* ms[QD_NUM_PHASES] accumulated GPU milliseconds, calls[QD_NUM_PHASES]. */
int qd_ctx_phase_times(qd_ctx *ctx, double *ms, uint64_t *calls);

/* Launch / layout parameters of the step (no reference counterpart; the
 * autotuner in quids_amd.tuning picks them).  Zero fields keep defaults. */
typedef struct qd_tuning {
    uint32_t dedup_bucket_target; /* mean candidates per final dedup bucket, 64..dedup_cap/2 (default 512) */
    uint32_t first_pass_bits;     /* max bits of the first radix pass, 1..12 (default 8) */
    uint32_t generate_block;      /* threads per generate block: 64, 128, 256, 512 or 1024 (default 256) */
    uint32_t generate_grid;       /* max generate blocks (default 16384) */
    uint32_t dedup_cap;           /* LDS bucket capacity of the dedup: 256, 512, 1024 or 2048 (default 1024; larger buckets take the global-memory path) */
    uint32_t generic_path;        /* 1: byte-level generate kernels for every rule (default 0: QCGD fast path) */
    uint32_t chunk_log2;          /* at most 2^k candidates per pass, 10..40 (default 0: as many as
                                     free device memory holds): a batched step splits its instances
                                     into chunks, and a step beyond it runs in hash-range passes */
    uint32_t split_generate;      /* 1: generate, then the first radix pass as a kernel of its own
                                     (default 0: fused when the first pass has at most 2^8 buckets);
                                     in hash-range passes, count and write in two walks per pass
                                     (default 0: one walk into per-block regions) */
    uint32_t eager_select;        /* 1: launch the top-K select's digit rounds one by one (default 0:
                                     replay them from a hipGraph the context captures once) */
    uint32_t free_mem_mib;        /* memory-driven truncation sees at most this much free device
                                     memory, in MiB (default 0: what hipMemGetInfo reports) */
} qd_tuning;
int qd_ctx_get_tuning(qd_ctx *ctx, qd_tuning *tuning);
int qd_ctx_set_tuning(qd_ctx *ctx, const qd_tuning *tuning);

/* Copies `bytes` of device memory `reps` times with a plain streaming
 * kernel (no reference counterpart): the achievable-bandwidth reference of
 * the autotuner and the calibration kernel of hardware byte counters.
 * gbps receives (read + write bytes) / time of the fastest repetition. */
int qd_copy_probe(qd_ctx *ctx, uint64_t bytes, int reps, double *gbps);

/* ---- state: quids::it_t (src/quids.hpp:149-335) ------------------------ */
int qd_state_create(qd_ctx *ctx, qd_state **state);
int qd_state_destroy(qd_state *state);
/* Replace the state's contents with N host objects.  object_begin has N+1
 * entries (byte offsets into objects, object i spans object_size[i] bytes
 * from object_begin[i]); magnitude holds N complex values as (re, im)
 * pairs.  Mirrors iteration::append (src/quids.hpp:174-188) applied N times
 * when offsets are padded to align_byte_length. */
int qd_state_upload(qd_state *state, uint64_t num_object, const void *objects,
                    const uint64_t *object_begin, const uint32_t *object_size,
                    const double *magnitude, double total_proba);
/* num_object, padded object byte length (object_begin[N]) and total_proba. */
int qd_state_info(qd_state *state, uint64_t *num_object, uint64_t *object_bytes,
                  double *total_proba);
/* Copy the state to host buffers sized from qd_state_info; any pointer may be
 * NULL to skip that array. */
int qd_state_download(qd_state *state, void *objects, uint64_t *object_begin,
                      uint32_t *object_size, double *magnitude);
/* Device-to-device copy of src into dst (same context).  The reference's
 * it_t is not copyable; this serves checkpoints of the adjoint's forward
 * pass. */
int qd_state_copy(qd_state *dst, const qd_state *src);
/* Device pointers of the state arrays (valid until the state is modified). */
int qd_state_device_ptrs(qd_state *state, void **objects, uint64_t **object_begin,
                         uint32_t **object_size, double **magnitude);

/* ---- symbolic iteration: quids::sy_it_t (src/quids.hpp:338-429) -------- */
int qd_symbolic_create(qd_ctx *ctx, qd_symbolic **sy);
int qd_symbolic_destroy(qd_symbolic *sy);
/* num_object (C) and num_object_after_interferences (U) of the last step. */
int qd_symbolic_info(qd_symbolic *sy, uint64_t *num_object,
                     uint64_t *num_object_after_interferences);

/* ---- the step: quids::simulate(it, rule, next, sy_it, max_num_object)
 *      (src/quids.hpp:448-543).  max_num_object = UINT64_MAX means no
 *      limit; 0 means memory-driven truncation from free device memory
 *      (src/quids.hpp:459-486, 510-536).  `in` and `out` must differ. */
int qd_simulate(qd_ctx *ctx, qd_state *in, const qd_rule *rule, qd_state *out,
                qd_symbolic *sy, uint64_t max_num_object, const qd_params *params,
                qd_step_stats *stats);

/* ---- user rules: the reference's plugin API, a subclass of quids::rule
 *      with get_num_child / populate_child / populate_child_simple / hasher
 *      (src/quids.hpp:105-146, README.md:112-180).  A rule the library has
 *      no kernels for is compiled by the caller as a device functor
 *      (quids_amd/user_rule.hpp turns a struct of __host__ __device__
 *      methods into this table); its callbacks enqueue the caller's kernels
 *      on `stream` (a hipStream_t) and return 0, or a nonzero code that the
 *      step reports as QD_ERR_INVALID.  The library runs everything else of
 *      the step: pre-truncation, the num_child scan, dedup, truncation,
 *      order, normalisation.  All arrays are device memory. */
typedef struct qd_objects_view {
    const uint8_t *objects; /* object bytes */
    const uint64_t *begin;  /* n + 1 padded offsets */
    const uint32_t *size;   /* n sizes */
    const double *mag;      /* n magnitudes (re, im) */
    uint64_t n;
} qd_objects_view;

typedef struct qd_user_rule {
    void *self; /* handed back to every callback */
    /* get_num_child of every parent (src/quids.hpp:557-567): num_child[i],
     * max_child_size[i] */
    int (*num_child)(void *self, const qd_objects_view *parents, uint32_t *num_child,
                     uint32_t *max_child_size, void *stream);
    /* populate_child + hasher of every candidate (src/quids.hpp:705-719):
     * candidate c is child c - child_begin[p] of the parent p with
     * child_begin[p] <= c < child_begin[p + 1] (child_begin: n + 1 offsets);
     * its magnitude starts as the parent's.  Writes hash[c], mag[2c],
     * mag[2c + 1] and size[c].  scratch holds scratch_bytes bytes for the
     * children's bytes (room for at least one child of max_child_size
     * bytes per 256 candidates, in slices of max_child_size rounded up to
     * 16 bytes). */
    int (*generate)(void *self, const qd_objects_view *parents, const uint64_t *child_begin,
                    uint64_t num_candidates, uint32_t max_child_size, uint8_t *scratch,
                    uint64_t scratch_bytes, uint64_t *hash, double *mag, uint32_t *size,
                    void *stream);
    /* populate_child_simple of the survivors (src/quids.hpp:956-967):
     * candidate cand[k] (as above) written at out_objects + out_begin[k],
     * k < n */
    int (*populate)(void *self, const qd_objects_view *parents, const uint64_t *child_begin,
                    const uint64_t *cand, uint64_t n, const uint64_t *out_begin,
                    uint8_t *out_objects, void *stream);
    /* hasher of every object of a state: the key of probabilistic
     * truncation (qd_params.simple_truncation == 0) for the parents */
    int (*hash_objects)(void *self, const qd_objects_view *objects, uint64_t *hash, void *stream);
} qd_user_rule;

/* qd_simulate with a user rule, in one pass over its candidates (a step
 * whose candidates exceed the scratch memory is rejected with
 * QD_ERR_CAPACITY).  Bit-identical to qd_simulate for a user rule that
 * restates a built-in one with the same hasher. */
int qd_simulate_user(qd_ctx *ctx, qd_state *in, const qd_user_rule *rule, qd_state *out,
                     qd_symbolic *sy, uint64_t max_num_object, const qd_params *params,
                     qd_step_stats *stats);

/* ---- batched step (no counterpart in the reference: many independent
 *      QCGD simulations advanced by one launch sequence).  `in` holds the
 *      instances back to back: instance i owns objects [in_seg[i],
 *      in_seg[i+1]) (host array, num_instances + 1 entries, in_seg[0] = 0,
 *      in_seg[num_instances] = the state size).  Instance i runs rules[i]
 *      (all of one QCGD kind, each with its own coefficients) with its own
 *      max_num_object[i] (explicit: 0 is not accepted, and an instance may not
 *      hold more objects than its cap).  The result is what num_instances
 *      separate qd_simulate calls return, bit for bit, written back to back
 *      into `out`; out_seg (host, num_instances + 1) receives the instances'
 *      object offsets and stats (host, optional, num_instances) their
 *      counters.  Equal objects of different instances never merge: the
 *      instance index is folded into the dedup key.  num_instances is 1 to
 *      4096; an instance may hold no object (it steps to an empty one); an
 *      instance holding a graph of more than 64 nodes, or more candidates
 *      than one pass takes, steps on its own inside the call, with the
 *      same results. */
int qd_simulate_batched(qd_ctx *ctx, qd_state *in, const uint64_t *in_seg, uint32_t num_instances,
                        const qd_rule *rules, qd_state *out, uint64_t *out_seg,
                        const uint64_t *max_num_object, const qd_params *params,
                        qd_step_stats *stats);

/* ---- modifier: quids::simulate(it, modifier) (src/quids.hpp:436-438,
 *      973-980) ----------------------------------------------------------- */
int qd_apply_modifier(qd_ctx *ctx, qd_state *state, const qd_modifier *mod);

/* ---- normalisation: iteration::normalize (src/quids.hpp:985-1017) ------- */
int qd_normalize(qd_ctx *ctx, qd_state *state, double *total_proba);

/* ---- expectation values: iteration::average_value (src/quids.hpp:208-234)
 *      for the QCGD observables 0 = number of nodes, 1 = particle count
 *      sum(L + R), 2 = [number of nodes == arg].  value = sum obs |m|^2 in a
 *      fixed summation order.  cot (device, optional, 2 doubles per object)
 *      receives d value / d Re m + i d value / d Im m = 2 obs m, the cotangent
 *      the adjoint starts from. */
int qd_average_value(qd_ctx *ctx, qd_state *state, int observable, int arg, double *value,
                     double *cot);
/* average_value for any observable (the reference's observable_t,
 * src/quids.hpp:83, 208-234): obs (device, one double per object) holds the
 * observable of every object, computed by the caller's own kernel
 * (quids_amd/user_rule.hpp, device_observable); value = sum obs |m|^2 in the
 * fixed summation order of qd_average_value; cot as there. */
int qd_average_value_of(qd_ctx *ctx, qd_state *state, const double *obs, double *value,
                        double *cot);
/* The same without any device allocation (no counterpart in the reference:
 * for callers that cannot allocate, such as code under HIP graph capture or
 * tight serving and training loops).  Its scratch is a static device array
 * of 32 MiB that each GPU holds until the process ends; states of more than
 * 2^22 objects are rejected with QD_ERR_CAPACITY.  Calls that overlap on one
 * GPU share that array and race.  Results are bit-identical to
 * qd_average_value. */
int qd_average_value_noalloc(qd_ctx *ctx, qd_state *state, int observable, int arg, double *value,
                             double *cot);

/* ---- adjoint of the step (no counterpart in the reference: gradients of a
 *      loss of the final state with respect to the rules' real parameters,
 *      for fitting QCGD angles).  `out` must be what qd_simulate returned for
 *      this `in` and `rule` (normalised magnitudes, total_proba before
 *      normalisation); the forward truncation is held as a fixed mask, and a
 *      step whose parents were pre-truncated (max_num_object < in's size) is
 *      rejected.  Device arrays of complex pairs (re, im):
 *        out_cot  cotangent dL/dRe m + i dL/dIm m of out's magnitudes (read);
 *        in_cot   the same for in's magnitudes (written).
 *      dcoef[i], i < ndir <= 3: a rule of the same kind whose coefficient
 *      fields hold the derivatives of rule's coefficients along real
 *      parameter i (e.g. theta, phi); grad[i] (host) receives dL/dparam_i. */
int qd_simulate_adjoint(qd_ctx *ctx, qd_state *in, const qd_rule *rule, const qd_rule *dcoef,
                        int ndir, qd_state *out, const double *out_cot, double *in_cot,
                        double *grad, uint64_t max_num_object);
/* Cotangent through simulate(it, modifier): cot (device, 2 doubles per
 * object) is multiplied in place by the conjugate of the factor the modifier
 * applied; `after` is the state after the modifier. */
int qd_modifier_adjoint(qd_ctx *ctx, qd_state *after, const qd_modifier *mod, double *cot);

/* ---- single-object rule calls on the host (the rule_t virtual methods,
 *      src/quids.hpp:115-145), run by the same functor code as the device
 *      kernels.  Used by the C++ header's rule classes. ------------------- */
int qd_rule_num_child(const qd_rule *rule, const void *parent, uint32_t parent_size,
                      uint32_t *num_child, uint32_t *max_child_size);
int qd_rule_populate_child(const qd_rule *rule, const void *parent, uint32_t parent_size,
                           void *child, uint32_t child_capacity, uint32_t child_id,
                           uint32_t *child_size, double *mag /* in: parent, out: child */);
int qd_rule_hash(const qd_rule *rule, const void *object, uint32_t size, uint64_t *hash);
int qd_modifier_apply_host(const qd_modifier *mod, void *object, uint32_t size,
                           double *mag);

#ifdef __cplusplus
}
#endif

#endif /* QUIDS_HIP_H */
