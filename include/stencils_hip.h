/*
 * stencils_hip.h — C ABI of libstencils_hip.so, the MI355X (gfx950) drop-in
 * for the `mapstencil` / `mapstencil!` hot path of Stencils.jl v0.3.6.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (STH_OK = 0) and never lets a C++ exception cross the boundary.
 * The message of the last failure on the calling thread is sth_last_error().
 *
 * Array layout follows Julia: column-major, dim 1 contiguous.  An sth_array
 * describes the LOGICAL array of a StencilArray: `size` is the logical size
 * (src/array.jl:470-473), `data` the parent's first element, `halo` the
 * number of padding cells on each side of every dim (R for Halo padding,
 * 0 for Conditional), `stride` the parent's strides in elements.  Logical
 * index i (0-based) of dim d lives at parent index i + halo.
 *
 * Status codes mirror the reference's failure modes:
 *   STH_EINVAL  — ArgumentError in the reference (size mismatch
 *                 src/gatherstencil.jl:121, radius >= axis src/array.jl:452,
 *                 kernel length src/stencils/kernel.jl:115); the Julia shim
 *                 rethrows ArgumentError with sth_last_error().
 *   STH_ENOTSUP — a combination the GPU path does not accelerate; the shim
 *                 falls back to the reference method (`invoke`).
 *   STH_ENOMEM, STH_EDEVICE — allocation / HIP runtime failures.
 */
#ifndef STENCILS_HIP_H
#define STENCILS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STH_ABI_VERSION 1

enum {
  STH_OK = 0,
  STH_EDEVICE = 5,
  STH_ENOMEM = 12,
  STH_EINVAL = 22,
  STH_ENOTSUP = 95
};

/* Element types (Julia Bool, UInt8 ... Float64; BF16 is a storage type). */
typedef enum {
  STH_BOOL = 0,
  STH_U8 = 1,
  STH_I8 = 2,
  STH_U16 = 3,
  STH_I16 = 4,
  STH_U32 = 5,
  STH_I32 = 6,
  STH_U64 = 7,
  STH_I64 = 8,
  STH_F16 = 9,
  STH_BF16 = 10,
  STH_F32 = 11,
  STH_F64 = 12
} sth_dtype;

/* Boundary conditions, src/boundary.jl:17-52. */
typedef enum {
  STH_REMOVE = 0,  /* out-of-range neighbours read `padval`            */
  STH_WRAP = 1,    /* i < 1 -> i + s ; i > s -> i - s                    */
  STH_REFLECT = 2, /* i < 1 -> 2 - i ; i > s -> 2s - i (edge not repeated) */
  STH_USE = 3      /* Halo padding only: read the halo as it is          */
} sth_boundary;

/* Padding modes, src/padding.jl:28-39. */
typedef enum {
  STH_CONDITIONAL = 0,
  STH_HALO_OUT = 1,
  STH_HALO_IN = 2
} sth_padding;

/* Stencil shape kinds (offset generators under src/stencils/). */
typedef enum {
  STH_SHAPE_POSITIONAL = 0,
  STH_SHAPE_WINDOW = 1,
  STH_SHAPE_MOORE = 2,
  STH_SHAPE_VONNEUMANN = 3,
  STH_SHAPE_CIRCLE = 4,
  STH_SHAPE_CROSS = 5,
  STH_SHAPE_ANGLEDCROSS = 6,
  STH_SHAPE_FORWARDSLASH = 7,
  STH_SHAPE_BACKSLASH = 8,
  STH_SHAPE_VERTICAL = 9,
  STH_SHAPE_HORIZONTAL = 10,
  STH_SHAPE_DIAMOND = 11,
  STH_SHAPE_ANNULUS = 12,
  STH_SHAPE_CARDINAL = 13,
  STH_SHAPE_ORDINAL = 14,
  STH_SHAPE_RECTANGLE = 15,
  STH_SHAPE_NAMED = 16
} sth_shape;

/* The fixed reduction set accelerated on the GPU. */
typedef enum {
  STH_OP_SUM = 0,           /* StaticArrays sum: left fold in offset order   */
  STH_OP_MEAN = 1,          /* sum / L (integers: see STH_OPF_INT_WRAP)      */
  STH_OP_MIN = 2,           /* minimum: left fold with Julia `min`           */
  STH_OP_MAX = 3,           /* maximum: left fold with Julia `max`           */
  STH_OP_KERNELPRODUCT = 4, /* kernelproduct, src/stencils/kernel.jl:37-43   */
  STH_OP_LIFE = 5,          /* Life-like rule on the count of live neighbours */
  STH_OP_LAPLACIAN = 6      /* c + alpha * (sum(neighbours) - L * c)          */
} sth_opkind;

/* sth_op.flags */
#define STH_OPF_ACC_F32 1u /* BF16/F16 storage: widen to f32, round once at store */
/*
 * STH_OPF_INT_WRAP: the caller states that, on its side, `sum` over a Stencil
 * of an integer element type T folds with T's own `+` (wrapping on
 * overflow), and that `mean` over it is that sum converted to Float64 and
 * divided by L.  The reference leaves both to StaticArrays (a Stencil is a
 * StaticVector, src/stencil.jl:17), which it does not pin, and Base's `sum`
 * would widen small integers instead.  Without this flag the library
 * declines (STH_ENOTSUP) `sum` over Int8..UInt32 and `mean` over every
 * integer type except Bool whenever L >= 2, i.e. wherever the two readings
 * can differ; the Julia shim sets it after checking its own StaticArrays on
 * one overflowing SVector.
 */
#define STH_OPF_INT_WRAP 2u

typedef struct sth_array {
  void* data;         /* device pointer to the parent's first element */
  int32_t dtype;      /* sth_dtype */
  int32_t ndim;       /* 1..3 */
  int64_t size[3];    /* logical size; unused dims are 1 */
  int64_t stride[3];  /* parent strides in elements */
  int32_t halo;       /* padding cells on each side of every dim */
  int32_t reserved;
} sth_array;

typedef struct sth_stencil {
  int32_t kind;            /* sth_shape (selects a build-time specialised kernel) */
  int32_t radius;          /* the reference's radius type parameter R */
  int32_t ndim;            /* stencil dims N <= array ndim; trailing dims pass through */
  int32_t length;          /* L */
  const int32_t* offsets;  /* host pointer, L * ndim ints, in fold order */
  const void* weights;     /* host pointer to L weights (Kernel) or NULL */
  int32_t weights_dtype;   /* sth_dtype of weights */
  int32_t reserved;
} sth_stencil;

typedef struct sth_op {
  int32_t kind;           /* sth_opkind */
  uint32_t flags;         /* STH_OPF_* */
  double alpha;           /* Laplacian coefficient */
  uint64_t birth_mask;    /* Life: bit s => a dead cell with s live neighbours is born */
  uint64_t survive_mask;  /* Life: bit s => a live cell with s live neighbours survives */
} sth_op;

/* Call flags. */
#define STH_ASYNC 1u     /* return without synchronising the stream */
#define STH_GRAPH 2u     /* sth_iterate: replay the steps as a captured hipGraph
                            (calls of 17 launches or more; shorter ones launch directly) */
#define STH_GENERIC 4u   /* force the runtime-offset kernel (testing / tuning) */
#define STH_NO_TUNE 16u  /* use the default kernel variant; never time variants */
#define STH_NO_FUSE 32u  /* sth_iterate: no temporal blocking (one generation per launch);
                            sth_map_multi: one pass per stencil argument */

typedef struct sth_ctx sth_ctx;

/* Library version (STH_ABI_VERSION * 10000 + minor). */
int sth_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* sth_last_error(void);

/* Context: one per host thread.  `device` is a HIP device ordinal. */
int sth_ctx_create(int device, sth_ctx** out);
int sth_ctx_destroy(sth_ctx* ctx);

/* Device memory owned by the library (the shim never frees it otherwise). */
int sth_alloc(sth_ctx* ctx, size_t bytes, void** dptr);
int sth_free(sth_ctx* ctx, void* dptr);
/* kind: 0 host->device, 1 device->host, 2 device->device. */
int sth_copy(sth_ctx* ctx, void* dst, const void* src, size_t bytes, int kind,
             void* stream);
int sth_synchronize(sth_ctx* ctx, void* stream);

/*
 * update_boundary! (src/array.jl:195-242): refresh the halo of a Halo-padded
 * array from its interior.  `radius` is the halo width R.  No-op for USE.
 */
int sth_update_boundary(sth_ctx* ctx, const sth_array* a, int32_t radius,
                        int32_t boundary, const void* padval, void* stream);

/*
 * gatherstencil!(f, dest, source) (src/gatherstencil.jl:89-103):
 * dst[I] = op(stencil(src, I)) for every logical I.  For Halo padding the
 * source halo is refreshed first, as update_boundary! does.  `padval`
 * points to one value of src->dtype (REMOVE only).
 */
int sth_map(sth_ctx* ctx, const sth_array* src, const sth_array* dst,
            const sth_stencil* st, int32_t boundary, const void* padval,
            int32_t padding, const sth_op* op, uint32_t flags, void* stream);

/*
 * One argument of a multi-array gather: a StencilArray (its parent, stencil,
 * boundary and padding, as for sth_map) or a plain array, and the fold its
 * term applies:
 *   stencil.length >= 1: op folds the argument's stencil values (sum, mean,
 *     minimum, maximum, kernelproduct, Laplacian, Life; result in the
 *     argument's element type);
 *   stencil.length == 0: the centre value, center(hood) or a plain array
 *     read at I (src/gatherstencil.jl:112-113); `op` is ignored.  A Halo
 *     StencilArray read this way still gets its halo refreshed, as the
 *     reference refreshes every StencilArray argument.
 * scaled = 1 makes the term coef * fold (coef exactly representable in the
 * element type; float types only).
 */
typedef struct sth_term {
  sth_array src;
  sth_stencil stencil;
  sth_op op;
  const void* padval;  /* one value of src.dtype (Remove only) */
  int32_t boundary;
  int32_t padding;
  double coef;
  int32_t scaled;
  int32_t reserved;
} sth_term;

#define STH_COMBINE_ADD 0 /* t_0 + t_1 + ... (left to right, each rounded) */
#define STH_COMBINE_MUL 1 /* t_0 * t_1 * ... */
#define STH_MAX_TERMS 8

/*
 * gatherstencil!(f, dest, source, args...) with several arguments
 * (src/gatherstencil.jl:89-103, kernel :105-109): dest[I] = t_0 (+|*) t_1
 * (+|*) ... with t_k the term of argument k at I, for f of the form
 * (h_0, h_1, ...) -> g_0(h_0) + c_1 * g_1(h_1) + ...  Every argument's
 * logical size equals dst's (else STH_EINVAL, the reference's
 * _checksizes, :115-121); every element type equals dst's (else
 * STH_ENOTSUP, as are term results of another type, e.g. an integer
 * mean).  Halo arguments are refreshed first (update_boundary!, :93-95).
 * Element types: Float32, Float64, UInt8, Int32, Int64 (integer `+` and
 * `*` wrap; an integer sum term needs STH_OPF_INT_WRAP as for sth_map).
 * mapstencil!(f, A::SwitchingStencilArray, B, ...) (:84-88) is this call
 * with A's source as argument 0 and A's dest buffer as dst.
 */
int sth_map_multi(sth_ctx* ctx, int32_t nterms, const sth_term* terms, const sth_array* dst,
                  int32_t combine, uint32_t flags, void* stream);

/*
 * scatterstencil!(f, op, dest, source, args...) (src/scatterstencil.jl:37-47,
 * 2-D loop :51-75, targets :78-114): every source cell I computes one value
 * per stencil offset k and folds it into dest at I + offset_k (mapped by the
 * source's boundary: Wrap wraps, Reflect reflects once, Remove and Use skip
 * targets outside), dest[J] = op(dest[J], v).  The values (`f`) come from a
 * fixed family; A is the source, B the optional second StencilArray (same
 * offsets), c the coefficient (exactly representable in the element type):
 *   STH_SCATTER_CONST          v = c                         map(_ -> c, neighbors(a))
 *   STH_SCATTER_CENTER         v = A[I]                      map(_ -> center(a), ...)
 *   STH_SCATTER_SCALED         v = A[I] * c
 *   STH_SCATTER_DOWNHILL       d = B[I] - B_k; v = d > 0 ? A[I] * c : 0
 *   STH_SCATTER_DOWNHILL_FLOW  d = B[I] - B_k; v = d > 0 ? A[I] * d * c : 0
 * (B_k: B's neighbour k of I under B's own boundary and padding).  Each dest
 * cell receives its values in the reference's order -- passes over source
 * columns j = p, p + S, ... (S = 2R + 1) for p = 1..S, rows ascending, offsets
 * in order -- so `+` rounds exactly as the reference's loop does.  2-D
 * Float32 / Float64; op STH_COMBINE_ADD, STH_COMBINE_MAX or STH_COMBINE_MIN
 * (Julia max / min).  STH_SCATTER_ZERO_DEST starts from zero instead of
 * dest's values (the SwitchingStencilArray forms clear dest first, :118-136).
 * Halo arguments are refreshed first.  A stencil whose reach exceeds its
 * radius R gets STH_ENOTSUP (the reference's passes would overlap).
 */
typedef struct sth_scatter_rule {
  int32_t kind;     /* STH_SCATTER_* */
  int32_t combine;  /* STH_COMBINE_ADD / STH_COMBINE_MAX / STH_COMBINE_MIN */
  double coef;
} sth_scatter_rule;
#define STH_SCATTER_CONST 0
#define STH_SCATTER_CENTER 1
#define STH_SCATTER_SCALED 2
#define STH_SCATTER_DOWNHILL 3
#define STH_SCATTER_DOWNHILL_FLOW 4
#define STH_COMBINE_MAX 2 /* Julia max */
#define STH_COMBINE_MIN 3 /* Julia min */
#define STH_SCATTER_ZERO_DEST 64u
int sth_scatter(sth_ctx* ctx, const sth_term* source, const sth_term* aux, const sth_array* dst,
                const sth_scatter_rule* rule, uint32_t flags, void* stream);

/*
 * `steps` calls of mapstencil!(f, A::SwitchingStencilArray)
 * (src/gatherstencil.jl:77-83 + switch, src/array.jl:610-611), ping-ponging
 * buffers a (source) and b (dest) on device.  On return *current is 0 if
 * `a` holds the newest state, 1 if `b` does (the switched wrapper's source).
 * With STH_GRAPH the steps are captured once into hipGraphs and replayed.
 * Life rules (2-D), 3-D radius-1 catalogue stencils under Wrap, Remove or
 * Reflect (arrays of 128^3 cells or more) and 2-D radius-1 catalogue stencils
 * in float types under Wrap, Remove or Reflect (2^21 cells or more, rows on
 * 16-byte boundaries)
 * advance two steps per launch in the steady
 * steps (temporal blocking; STH_NO_FUSE turns it off) and finish with single
 * steps, so both buffers end exactly as the step-by-step reference leaves them.
 */
int sth_iterate(sth_ctx* ctx, const sth_array* a, const sth_array* b,
                const sth_stencil* st, int32_t boundary, const void* padval,
                int32_t padding, const sth_op* op, int64_t steps,
                uint32_t flags, int32_t* current, void* stream);

/*
 * Autotuner.  Several kernel variants can run one call (streaming kernels with
 * different rows per wave, the LDS-tiled kernel, Life rows per wave); all of
 * them compute identical results.  sth_map / sth_iterate time the variants of
 * a (shape, element type, op, size bucket) family the first time they see it
 * -- unless STH_NO_TUNE is passed, STH_AUTOTUNE=0 is set in the environment or
 * the stream is being captured -- and reuse the fastest afterwards.
 * STH_TUNE_FILE names a file the decisions are loaded from and appended to.
 * No counterpart in the reference (KernelAbstractions picks a fixed
 * workgroup size, src/gatherstencil.jl:97-99).
 *
 * sth_autotune times every variant of this call now (replacing an earlier
 * decision), runs the call with the winner and writes the winner's name into
 * `chosen` (may be NULL).  sth_tuning_table writes "family\tvariant\n" lines;
 * `needed` receives the buffer size the table requires.
 */
int sth_autotune(sth_ctx* ctx, const sth_array* src, const sth_array* dst,
                 const sth_stencil* st, int32_t boundary, const void* padval,
                 int32_t padding, const sth_op* op, void* stream, char* chosen,
                 size_t chosen_len);
int sth_tuning_table(sth_ctx* ctx, char* buf, size_t len, size_t* needed);
/* Testing and tuning hook: while `pattern` is non-empty, every call uses the
 * first variant whose name contains it (if any), bypassing the tuner. */
int sth_set_variant(sth_ctx* ctx, const char* pattern);

/* One mapstencil call of a batch (sth_map_batch). */
typedef struct sth_job {
  sth_array src;
  sth_array dst;
  sth_stencil stencil;
  sth_op op;
  const void* padval;
  int32_t boundary;
  int32_t padding;
} sth_job;

#define STH_BATCH_SERIAL 8u /* sth_map_batch: one launch per job (the unbatched path) */

/*
 * Many independent gatherstencil!(f, dest, source) calls -- mixed sizes,
 * shapes and reductions -- in one launch per (element type, shape,
 * reduction): every job's tiles form one grid, and each workgroup finds its
 * job from its tile index.  Halo jobs get their update_boundary! refresh
 * first; jobs outside the batched catalogue, and jobs large enough to fill
 * the device alone, run as individual launches inside the call.  Each result is identical to an
 * sth_map call on that job alone.  No counterpart in the reference, which
 * runs one KernelAbstractions launch per mapstencil call
 * (src/gatherstencil.jl:97-100); callers: the Julia shim's batched
 * `mapstencil` over a vector of StencilArrays, and the serving loop.
 */
int sth_map_batch(sth_ctx* ctx, int32_t njobs, const sth_job* jobs, uint32_t flags,
                  void* stream);

/*
 * Many independent iterations with uneven step counts in one call: job i is
 * a SwitchingStencilArray whose buffers are jobs[i].src (its source) and
 * jobs[i].dst, advanced by steps[i] calls of A = mapstencil!(f_i, A)
 * (src/gatherstencil.jl:77-83, switch src/array.jl:610-611); the reference
 * counterpart is a loop over the arrays, each looping over its own steps.
 * Round r advances every array with more than r steps by one step, as one
 * sth_map_batch launch set over the arrays still running (one launch per
 * (element type, shape, reduction), large arrays on their own launches).
 * With STH_GRAPH the whole schedule is captured once as a hipGraph -- its
 * job tables in device memory the graph owns -- and replayed by later calls
 * with the same arrays, functions and step counts.  current[i] receives
 * which buffer holds array i's newest state (0: src, 1: dst).  Each result,
 * halo included, equals sth_iterate on that array alone.  STH_BATCH_SERIAL
 * runs one launch per array and step (the comparison path).
 */
int sth_iterate_batch(sth_ctx* ctx, int32_t njobs, const sth_job* jobs, const int64_t* steps,
                      uint32_t flags, int32_t* current, void* stream);

/*
 * Compile (with hipRTC, no device needed) a representative set of the
 * runtime specialisations the library builds for calls outside its
 * build-time catalogue -- streaming kernels over each storage type and
 * multi-array passes -- and return how many failed, their compiler messages
 * in `log` (NUL-terminated, at most `cap` bytes).  No counterpart in the
 * reference; for installers and tests that check the runtime compiler works
 * on a machine before any call depends on it.
 */
int sth_jit_selftest(char* log, size_t cap);

/*
 * sth_map_batch without allocating: the same calls, kernels and results as
 * sth_map_batch (each result identical to sth_map on that job alone), but
 * no device or pinned host memory is allocated, no host buffer is copied to
 * the device, and the autotuner only reuses decisions already made (it does
 * not time variants).  The batched launches' job tables live in one fixed
 * device array of STH_STATIC_BATCH_MAX_JOBS entries (819,200 bytes) per
 * device, held from library load to process exit; they are written by small
 * kernels whose arguments carry the jobs.  More than
 * STH_STATIC_BATCH_MAX_JOBS jobs is STH_EINVAL.
 * Every call on a device shares that device's array: calls that overlap on
 * one device (on different streams, or from different threads) race; calls
 * ordered on one stream do not.
 * No counterpart in the reference (one launch per mapstencil call,
 * src/gatherstencil.jl:97-100); for callers that cannot allocate device
 * memory: code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc.
 */
#define STH_STATIC_BATCH_MAX_JOBS 4096
int sth_map_batch_static(sth_ctx* ctx, int32_t njobs, const sth_job* jobs, uint32_t flags,
                         void* stream);

/*
 * Serving loop: continuous batching of iterated stencil solves.
 *
 * A request is `steps` calls of mapstencil!(f, A::SwitchingStencilArray)
 * (src/gatherstencil.jl:77-83) on its own state.  The server keeps the two
 * buffers of every admitted request resident in an arena of `hbm_budget`
 * bytes of device memory, admits requests in arrival order as soon as their
 * arrival time has passed and their buffers fit (and fewer than `max_active`
 * run), and advances every admitted request together: each step of a round
 * is one batched launch over all of them (sth_map_batch's kernels), so
 * requests that arrive while others run join at the next round.  Requests
 * of at least 2^21 cells with a two-step kernel (radius-1 catalogue stencils
 * in 2-D and 3-D, Life) run their own two-step launches
 * on their own streams instead.  Finished states (halo included) are copied to the request's
 * `result` and their buffers freed.
 * STH_SERVER_SERIAL instead serves one request at a time with one plain
 * launch per step (the unbatched path, for comparison).
 * No counterpart in the reference; callers: a serving front end (the
 * Python `Server`, a Julia task) that receives solve requests.
 */
typedef struct sth_server sth_server;

#define STH_SERVER_SERIAL 1u

typedef struct sth_server_config {
  size_t hbm_budget;         /* bytes of device memory for request buffers */
  int32_t max_active;        /* most requests admitted at once (0: no limit) */
  int32_t round_steps;       /* steps per scheduling round (0: 8) */
  uint32_t flags;            /* STH_SERVER_* */
  int32_t reserved;
} sth_server_config;

typedef struct sth_request {
  sth_array state;      /* initial state: the PARENT layout, dense (stride[1] = size[0] + 2*halo,
                           ...); data is host or device memory; copied at admission */
  sth_stencil stencil;  /* copied at submit */
  sth_op op;
  const void* padval;   /* copied at submit */
  int32_t boundary;
  int32_t padding;
  int64_t steps;
  void* result;         /* host or device memory, same layout as state: the final state */
  double arrival;       /* server-clock seconds before which the request is not admitted */
  uint64_t tag;         /* caller's identifier, returned in the completion */
} sth_request;

typedef struct sth_completion {
  uint64_t tag;
  int32_t status;       /* STH_OK or the failure of this request */
  int32_t reserved;
  double arrival;       /* server-clock seconds */
  double admitted;
  double finished;      /* when the result copy completed on the device */
  int64_t steps;
  int64_t cells;        /* logical cells of the state */
} sth_completion;

typedef struct sth_server_stats {
  int64_t submitted, completed, active, pending;
  int64_t rounds, launches;
  int64_t cell_steps;    /* completed requests' cells x steps */
  size_t hbm_in_use, hbm_peak;
  double started;        /* wall seconds since sth_server_start */
  int64_t solo_launches;   /* launches on the solo requests' own streams */
  int64_t fused_launches;  /* of those, launches that advance several steps */
} sth_server_stats;

int sth_server_create(sth_ctx* ctx, const sth_server_config* cfg, void* stream, sth_server** out);
/* Validates and queues a request (any time before or after start). */
int sth_server_submit(sth_server* srv, const sth_request* req);
/* Starts the server clock (arrival 0) and the scheduler thread. */
int sth_server_start(sth_server* srv);
/* Copies up to `max` completions not yet returned; waits up to wait_ms for one. */
int sth_server_poll(sth_server* srv, sth_completion* out, int32_t max, int32_t* n, int32_t wait_ms);
/* Waits until every submitted request has completed (or timeout_s passes: STH_EDEVICE). */
int sth_server_drain(sth_server* srv, double timeout_s);
int sth_server_stats_get(sth_server* srv, sth_server_stats* out);
/* Stops the scheduler (after draining) and frees the arena. */
int sth_server_destroy(sth_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* STENCILS_HIP_H */
