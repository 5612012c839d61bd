/*
 * libtsm: tall-skinny C = A^T B for AMD Instinct MI355X (gfx950).
 *
 * Plain C ABI: device pointers, sizes, status codes and an opaque stream
 * handle (a hipStream_t; NULL = the default stream).  No C++ or torch types
 * cross this boundary and no call throws.
 *
 * Shapes follow the reference (tsmttsm.hh:5-10, sycl_codegen_new.py:56-67):
 *   A is n x k, B is n x m, both row-major with leading dimensions
 *   lda >= k and ldb >= m (elements); C is k x m row-major, ldc >= m.
 *   Reference names: K = n, M = k, N = m.  1 <= k, m <= 64, 0 <= n < 2^62.
 *
 * Each entry point below names the reference interface it replaces.
 */
#ifndef TSM_TSM_H
#define TSM_TSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The library builds with hidden visibility; everything declared here is exported. */
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

#define TSM_VERSION_MAJOR 0
#define TSM_VERSION_MINOR 1
#define TSM_MAX_WIDTH 64

typedef void* tsm_stream_t; /* hipStream_t */

typedef enum {
  TSM_F64 = 0,  /* IEEE binary64 */
  TSM_F32 = 1,  /* IEEE binary32 */
  TSM_BF16 = 2, /* bfloat16 (upper half of a binary32) */
  /* Not values: they make every int a value of the type, so that an
   * argument outside the list (a C caller may pass any int) is rejected by
   * the checks instead of being undefined in the library's C++. */
  TSM_DTYPE_RANGE_MIN_ = -2147483647 - 1,
  TSM_DTYPE_RANGE_MAX_ = 2147483647
} tsm_dtype;

typedef enum {
  TSM_SUCCESS = 0,
  TSM_ERR_INVALID_ARGUMENT = 1, /* null pointer, n < 0, k or m outside [1,64], ld too small */
  TSM_ERR_UNSUPPORTED = 2,      /* dtype combination or kernel request not provided */
  TSM_ERR_MISALIGNED = 3,       /* pointer not aligned to its element size */
  TSM_ERR_WORKSPACE = 4,        /* caller workspace smaller than tsm_atb_workspace_size(), or
                                   none given on a stream that is being captured */
  TSM_ERR_HIP = 5,              /* a HIP runtime call failed: see tsm_last_hip_error() */
  TSM_ERR_NO_KERNEL = 6,        /* no generated kernel matches the requested config */
  TSM_STATUS_RANGE_MIN_ = -2147483647 - 1, /* not values: every int is one of the type */
  TSM_STATUS_RANGE_MAX_ = 2147483647
} tsm_status;

/* Distributions of tsm_fill_random (the synthetic inputs):
 *   TSM_DIST_UNIFORM: (w >> 11) * 2^-52 - 1, uniform on [-1, 1)
 *   TSM_DIST_POW2:    sign(bit 63) * 2^e, e = ((w >> 11) mod 41) - 20
 * where w = splitmix64_finalizer(seed + stream*0xd1b54a32d192ed03 + (idx+1)*0x9e3779b97f4a7c15)
 * and idx = row*cols + col.  FP32 and BF16 fills round the FP64 value once,
 * to nearest even. */
typedef enum {
  TSM_DIST_UNIFORM = 0,
  TSM_DIST_POW2 = 1,
  TSM_DIST_RANGE_MIN_ = -2147483647 - 1, /* not values: every int is one of the type */
  TSM_DIST_RANGE_MAX_ = 2147483647
} tsm_dist;

/* Explicit kernel choice (generator vocabulary).  Zero fields mean "tuned
 * default".  Replaces the compile-time constants baked into the generated
 * SYCL kernels: M, N, TM, TN, stride, local_size (sycl_codegen_new.py:156-160,
 * 265-267) and the max_tile_size / global_size template arguments of
 * tsmttsm4/6/9 (tsmttsm.hh:125, 230, 368). */
typedef struct {
  int tm, tn;          /* register tile: C entries owned per thread = tm x tn */
  int rows;            /* rows per LDS tile (the n-tile streamed per stage) */
  int pf;              /* tiles in flight per thread: 1 to 4 (the "leap-frog" depth);
                          direct engine: steps in flight per wave, 4, 8 or 16 (FP64/FP32)
                          or 2 or 4 (BF16) */
  int wgs_per_cu;      /* persistent workgroups per compute unit */
  int64_t grid;        /* total workgroups; overrides wgs_per_cu when > 0 (the "stride" role) */
  int load_policy;     /* 0 tuned default, 1 cached loads, 2 non-temporal (streaming) loads */
  int tile_order;      /* 0 tuned default, 1 contiguous runs of tiles per workgroup,
                          2 interleaved (tile t to workgroup t mod grid) */
  int engine;          /* 0 tuned default, 1 VALU register tiles (tm x tn), 2 matrix cores
                          (MFMA) fed through LDS tiles, 3 matrix cores fed straight from
                          global memory (BF16 through a wave-private LDS tile; pf = steps
                          in flight per wave, of 4 rows (FP64/FP32) or 32 rows (BF16);
                          inputs that are not dense run the class's staged kernel).
                          3 needs width classes that are multiples of 16; 2 also takes
                          one narrower side (BF16 4 and 8, FP32 4 and 8, FP64 2, 4 and
                          8) beside a side of 16 or more, padded to 16 in LDS, and
                          A^T A at 8. */
} tsm_kernel_config;

/* ---- the hot path ------------------------------------------------------
 * C = A^T B (accumulate == 0) or C += A^T B (accumulate != 0).
 * Replaces: the generated tsmttsm_{M}_{N}_{TM}_{TN}[_lf|_sg|_sg_lf]
 * (sycl_codegen_new.py:158-159, 263-264) and tsmttsm, tsmttsm2..9
 * (tsmttsm.hh:15-586), whose contract is C += A^T B with the caller zeroing C
 * (main.cc:67-69, 94, 119).  Unlike the reference, the result does not
 * depend on atomic ordering: the same inputs and config give the same bits.
 *
 *   in_type:  TSM_F64, TSM_F32 or TSM_BF16 (A and B share it)
 *   out_type: TSM_F64 or TSM_F32 (C).  The cross-workgroup reduction is FP64
 *             for every input type; FP32/BF16 inputs accumulate in FP32 in
 *             short row blocks folded into compensated (hi, lo) FP32 pairs.
 *   workspace: NULL to use a per-(device, stream) buffer owned by the library,
 *             or 256-byte aligned device memory of at least
 *             tsm_atb_workspace_size() bytes whose
 *             first 256 bytes are zero before its first use (the library
 *             leaves them zero after each call).  Calls sharing a workspace
 *             must be ordered on one stream.
 * Asynchronous on `stream`; no allocation when a workspace is given.  A call
 * is one kernel launch (short sums: the last workgroup adds the workgroup
 * partials) or two (grids of 256 workgroups or more, 16384 partial words or
 * more, or more than 512 outputs: a finisher kernel adds them); either form
 * can be captured in a HIP graph when a workspace is given. */
tsm_status tsm_atb(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, int m,
                   const void* A, int64_t lda, const void* B, int64_t ldb,
                   void* C, int64_t ldc, int accumulate,
                   void* workspace, size_t workspace_bytes, tsm_stream_t stream);

/* As tsm_atb with an explicit kernel configuration (NULL = tuned default).
 * Used by the autotuner and the name-compatible C++ layer (include/tsm/tsmttsm.hh). */
tsm_status tsm_atb_ex(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, int m,
                      const void* A, int64_t lda, const void* B, int64_t ldb,
                      void* C, int64_t ldc, int accumulate,
                      void* workspace, size_t workspace_bytes, tsm_stream_t stream,
                      const tsm_kernel_config* config);

/* ---- an entry point that allocates nothing --------------------------------
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  As tsm_atb with the tuned default kernel and no
 * workspace argument: the workspace is a static device array of
 * TSM_NOALLOC_SCRATCH_BYTES in the library (one per GPU, held until the
 * process ends), never a caller buffer or the per-stream pool, so the call
 * makes no allocation.  A shape whose default plan needs more
 * (tsm_atb_workspace_size() > TSM_NOALLOC_SCRATCH_BYTES; with 16 MiB no
 * default plan does on a 256-CU MI355X) returns TSM_ERR_WORKSPACE.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams, other host threads) race, so order them on one stream.
 * The first call on a device clears the array's arrival counters with a
 * kernel on `stream` and returns TSM_ERR_UNSUPPORTED if `stream` is being
 * captured; later calls may be captured. */
#define TSM_NOALLOC_SCRATCH_BYTES ((size_t)256 + ((size_t)16 << 20))
tsm_status tsm_atb_noalloc(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, int m,
                           const void* A, int64_t lda, const void* B, int64_t ldb,
                           void* C, int64_t ldc, int accumulate, tsm_stream_t stream);

/* Workspace bytes tsm_atb / tsm_atb_ex need for this shape (config may be NULL). */
size_t tsm_atb_workspace_size(tsm_dtype in_type, int64_t n, int k, int m,
                              const tsm_kernel_config* config);

/* ---- symmetric product C = A^T A (no reference counterpart) ----------------
 * The Gram matrix of a tall-skinny A (n x k, 1 <= k <= 64), reading A once
 * (half the bytes of tsm_atb(A, A)).  Same accumulation types, error bound
 * and deterministic reduction as tsm_atb; C is exactly symmetric (the
 * matrix-core classes compute the blocks on and above the diagonal and
 * mirror the others).  C is k x k (ldc >= k); accumulate, workspace and
 * stream as for tsm_atb. */
tsm_status tsm_ata(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, const void* A,
                   int64_t lda, void* C, int64_t ldc, int accumulate, void* workspace,
                   size_t workspace_bytes, tsm_stream_t stream);
/* tsm_ata with a kernel request, as tsm_atb_ex (the A^T A kernels of the
 * tuned square classes come in several rows-per-tile and pipeline depths,
 * and at widths 16, 32 and 64 also on the direct engine, engine = 3). */
tsm_status tsm_ata_ex(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, const void* A,
                      int64_t lda, void* C, int64_t ldc, int accumulate, void* workspace,
                      size_t workspace_bytes, tsm_stream_t stream,
                      const tsm_kernel_config* config);
size_t tsm_ata_workspace_size(tsm_dtype in_type, int64_t n, int k,
                              const tsm_kernel_config* config);

/* ---- differentiation (no reference counterpart) --------------------------
 * Backward of C = A^T B for an upstream gradient dC (k x m, row-major, lddc):
 *   dA = B dC^T (n x k), dB = A dC (n x m).
 * dA or dB may be NULL to skip it; both use one streaming pass over the rows.
 * dC is in out_type (TSM_F64 or TSM_F32); A, B, dA, dB are in in_type.
 * Each output is a sum of products accumulated in FP64 for FP64 inputs and in
 * FP32 otherwise, rounded once to in_type.  Dense 32- and 64-wide classes,
 * and the FP64/FP32 16-wide ones, run on matrix instructions (another
 * summation order than the narrower classes); for BF16 they split dC exactly
 * into three BF16 pieces, so every product stays exact. */
tsm_status tsm_atb_backward(tsm_dtype in_type, tsm_dtype grad_type, int64_t n, int k, int m,
                            const void* A, int64_t lda, const void* B, int64_t ldb,
                            const void* dC, int64_t lddc,
                            void* dA, int64_t ldda, void* dB, int64_t lddb,
                            tsm_stream_t stream);

/* ---- tall-skinny times small: Y = A X (no reference counterpart) ------------
 * A is n x k (in_type, row-major, lda), X is k x m (x_type: TSM_F64 or
 * TSM_F32, row-major, ldx), Y is n x m (in_type, ldy).  This is the
 * dB = A dC half of tsm_atb_backward (same kernels): each output is a sum of
 * k products in FP64 for FP64 inputs and FP32 otherwise, rounded once to
 * in_type.  Rows of 16..128 bytes with lda == k and ldy == m take the
 * LDS-staged kernel, or a matrix-core kernel for the FP64/FP32 16-wide and
 * all 32- and 64-wide classes. */
tsm_status tsm_tsmm(tsm_dtype in_type, tsm_dtype x_type, int64_t n, int k, int m,
                    const void* A, int64_t lda, const void* X, int64_t ldx, void* Y,
                    int64_t ldy, tsm_stream_t stream);

/* ---- batching (no reference counterpart) ----------------------------------
 * `count` independent problems C_i (+)= A_i^T B_i in ONE persistent launch.
 * `problems` is a HOST array; the library copies the work plan into device
 * memory (the workspace, or a plan's own buffer) before the launch.  All
 * problems share in_type/out_type; shapes may differ (1 <= k, m <= 64).
 * FP32 and BF16 problems with k or m above 32 run, inside the same launch, as
 * column blocks of at most 32 x 32 (A and B are then read once per block of
 * C's columns and rows).  The results are deterministic for a given problem
 * list.  Calls that share a workspace or a plan must be ordered on one
 * stream. */
typedef struct {
  const void* A;
  const void* B;
  void* C;
  int64_t n, lda, ldb, ldc;
  int k, m;
} tsm_problem;

size_t tsm_atb_batched_workspace_size(tsm_dtype in_type, int count, const tsm_problem* problems);

/* Uploads the plan (returning once the host array has been consumed) and
 * launches; workspace NULL = a library-owned buffer for this stream.  The
 * upload goes through a reused host buffer, so this call cannot be captured in
 * a graph (TSM_ERR_UNSUPPORTED on a capturing stream): use a plan there. */
tsm_status tsm_atb_batched(tsm_dtype in_type, tsm_dtype out_type, int count,
                           const tsm_problem* problems, int accumulate,
                           void* workspace, size_t workspace_bytes, tsm_stream_t stream);

/* A batched plan owns its device buffer: create it once (allocates and
 * uploads, synchronously), then execute it any number of times - each execute
 * is one kernel launch with no allocation, copy or memset, so it can be
 * captured in a HIP graph. */
typedef struct tsm_batched_plan_s* tsm_batched_plan;
tsm_status tsm_batched_plan_create(tsm_dtype in_type, tsm_dtype out_type, int count,
                                   const tsm_problem* problems, int accumulate,
                                   tsm_batched_plan* plan);
tsm_status tsm_batched_plan_execute(tsm_batched_plan plan, tsm_stream_t stream);
/* (Executions of one plan share its buffer: order them on one stream or with
 * events.  A graph that captured an execution uses the plan's buffer when it
 * replays, so the plan must outlive such graphs.  Destroy waits for the
 * plan's launches outside graphs, stream by stream, then frees the buffer.) */
/* Work items and launch grid of a plan (either pointer may be NULL). */
tsm_status tsm_batched_plan_info(tsm_batched_plan plan, int64_t* items, int64_t* grid);
tsm_status tsm_batched_plan_destroy(tsm_batched_plan plan);

/* ---- synthetic inputs (replaces main.cc:72-86, sycl_codegen_new.py:436-439) -- */
tsm_status tsm_fill_random(tsm_dtype type, tsm_dist dist, uint64_t seed, uint64_t stream_id,
                           void* X, int64_t rows, int64_t cols, int64_t ld,
                           tsm_stream_t stream);

/* ---- introspection ---------------------------------------------------------- */
const char* tsm_strerror(tsm_status status);
int tsm_last_hip_error(void);       /* hipError_t of the last TSM_ERR_HIP on this thread */
int tsm_version(void);              /* major * 10000 + minor * 100 */

typedef struct {
  const char* name;                 /* e.g. "tsm_atb_f64_8x8_t8x8_r256_p2" */
  int in_type;                      /* tsm_dtype */
  int kc, mc;                       /* width class: k <= kc, m <= mc */
  int tm, tn, rows, pf;              /* tm = tn = 0: matrix-core (MFMA) kernel */
  int load_policy;                  /* 1 cached, 2 non-temporal */
  int lds_bytes;
  int engine;                       /* as in tsm_kernel_config: 1, 2 or 3 */
} tsm_kernel_desc;

int tsm_num_kernels(void);
tsm_status tsm_kernel_info(int index, tsm_kernel_desc* out);
/* Kernel tsm_atb would use for this shape (NULL config = tuned default);
 * writes its index in the tsm_kernel_info table and the launch grid. */
tsm_status tsm_atb_plan(tsm_dtype in_type, int64_t n, int k, int m,
                        const tsm_kernel_config* config, int* kernel_index, int64_t* grid);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif
#endif /* TSM_TSM_H */
