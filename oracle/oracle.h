/*
 * CPU oracle for the tall-skinny product C = A^T * B.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library,
 * the Python package, bench.py's timed region) may link, load or call this
 * code.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker or as the timed CPU baseline.
 *
 * Conventions (as in the reference): A is n x k, B is n x m, both row-major
 * with leading dimensions lda >= k and ldb >= m; C is k x m row-major with
 * leading dimension ldc >= m.  Reference names: K = n (rows), M = k, N = m
 * (tsmttsm.hh:5-10, sycl_codegen_new.py:56-67).
 *
 * Which mathematical check pins each function is listed next to it; the tests
 * in tests/test_oracle.py carry those checks.
 */
#ifndef TSM_ORACLE_H
#define TSM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic inputs: counter-based SplitMix64 --------------------------
 * word(seed, stream, idx) = splitmix64_finalizer(seed + stream*S + (idx+1)*G)
 * with G = 0x9e3779b97f4a7c15 (the SplitMix64 increment) and
 * S = 0xd1b54a32d192ed03.  For stream 0 this is exactly the SplitMix64
 * sequence started from state `seed` (pinned by its published outputs).
 * Values depend only on (seed, stream, idx), never on the thread count,
 * unlike the reference's per-thread minstd_rand0 (main.cc:72-86). */
uint64_t oracle_rng_word(uint64_t seed, uint64_t stream, uint64_t idx);

/* Distributions (element idx = row*cols + col of a dense rows x cols matrix):
 *   0 = uniform[-1,1):  (w >> 11) * 2^-52 - 1            (multiple of 2^-52)
 *   1 = sign * 2^e, e uniform integer in [-20, 20]:
 *       sign from bit 63, e = ((w >> 11) % 41) - 20
 *   2 = normal(0,1) via Box-Muller from words 2*idx and 2*idx+1 (host only) */
enum { ORACLE_DIST_UNIFORM = 0, ORACLE_DIST_POW2 = 1, ORACLE_DIST_NORMAL = 2 };
double oracle_rng_value(int dist, uint64_t seed, uint64_t stream, uint64_t idx);
/* Fill a rows x cols matrix with leading dimension ld (padding untouched). */
void oracle_fill_f64(int dist, uint64_t seed, uint64_t stream,
                     int64_t rows, int64_t cols, double* out, int64_t ld);

/* ---- storage rounding (round-to-nearest-even, single rounding from FP64) */
float    oracle_round_f32(double x);
uint16_t oracle_round_bf16(double x);   /* returns the bf16 bit pattern */
double   oracle_bf16_to_f64(uint16_t h);
void     oracle_round_matrix_f32(int64_t count, const double* in, float* out);
void     oracle_round_matrix_bf16(int64_t count, const double* in, uint16_t* out);

/* ---- C = A^T B oracles -------------------------------------------------- */

/* Dot2 compensated summation (Ogita, Rump, Oishi, SIAM J. Sci. Comput. 26(6),
 * 2005, Algorithm 5.3): TwoProduct via FMA + TwoSum per term.  Rows are split
 * into 64 fixed chunks whose (sum, err) pairs are merged in chunk order with
 * TwoSum, so the result does not depend on the thread count.
 * If absC != NULL it receives |A|^T |B| (plain FP64 sums of non-negative
 * terms, relative error <= n*eps), the scale of the elementwise error bound. */
void oracle_atb_dot2(int64_t n, int k, int m,
                     const double* A, int64_t lda, const double* B, int64_t ldb,
                     double* C, int64_t ldc, double* absC);

/* Exact oracle.  Requires every A entry to be an integer multiple of 2^-qa
 * with |A * 2^qa| < 2^62 (likewise B with qb).  The products are summed
 * exactly in __int128 and rounded once (round-to-nearest-even) to FP64.
 * Returns 0 on success, -1 if an entry is not such a multiple, -2 on int128
 * overflow.  Uniform inputs are multiples of 2^-52 (qa = qb = 52, exact for
 * n <= 2^22); the sign*2^e inputs are multiples of 2^-20 (qa = qb = 20,
 * exact for n <= 2^46). */
int oracle_atb_exact(int64_t n, int k, int m,
                     const double* A, int64_t lda, int qa,
                     const double* B, int64_t ldb, int qb,
                     double* C, int64_t ldc);

/* Restatement of the reference's host oracle compute_naive (main.cc:14-27)
 * and of the generated reference_tsmttsm (sycl_codegen_new.py:392-400):
 * for each (i, j) a sequential FP64 loop over rows, generalised to k != m. */
void oracle_atb_naive(int64_t n, int k, int m,
                      const double* A, int64_t lda, const double* B, int64_t ldb,
                      double* C, int64_t ldc);

/* The reference's acceptance rule (main.cc:29-41; sycl_codegen_new.py:402-415):
 * entry e mismatches iff |ref-got| > abs_tol AND |ref-got| > rel_tol*max(|ref|,|got|).
 * Returns the index of the first mismatch, or -1 if all entries pass. */
int64_t oracle_first_mismatch(int64_t count, const double* ref, const double* got,
                              double rel_tol, double abs_tol);

/* Backward of L = <G, C>, C = A^T B: dA = B G^T (n x k), dB = A G (n x m),
 * each entry a Dot2 sum over m (resp. k) terms.  G is k x m row-major (ldg). */
void oracle_atb_backward(int64_t n, int k, int m,
                         const double* A, int64_t lda, const double* B, int64_t ldb,
                         const double* G, int64_t ldg,
                         double* dA, int64_t ldda, double* dB, int64_t lddb);

/* ---- CPU restatement of the reference kernels (ref_kernels.cc) -----------
 * Same work decomposition, per-work-item accumulators and atomic adds as the
 * reference; C += A^T B (the caller zeroes C first, as main.cc:67-69 does).
 * Work-items of a range kernel run as an OpenMP loop; an nd_range work-group
 * runs on one host thread with its items in sequence between barrier phases
 * (every barrier in these kernels lies before or after the row loop).
 * Returns 0, or -1 if the configuration is not supported (k, m not in
 * {1,2,4,8,16,32,64}, tile not dividing the width, stride constraints). */
enum {
  ORACLE_REF_GEN_BASIC = 0,  /* sycl_codegen_new.py:142-230, leap_frog=False */
  ORACLE_REF_GEN_LF    = 1,  /* sycl_codegen_new.py:142-230 + 74-135          */
  ORACLE_REF_GEN_SG    = 2,  /* sycl_codegen_new.py:233-363                   */
  ORACLE_REF_GEN_SG_LF = 3,  /* sycl_codegen_new.py:233-363 + 337-344         */
  ORACLE_REF_V1 = 11,        /* tsmttsm.hh:15-35   */
  ORACLE_REF_V2 = 12,        /* tsmttsm.hh:40-74   */
  ORACLE_REF_V3 = 13,        /* tsmttsm.hh:80-122  */
  ORACLE_REF_V4 = 14,        /* tsmttsm.hh:125-166 */
  ORACLE_REF_V5 = 15,        /* tsmttsm.hh:170-225 (first load guarded here) */
  ORACLE_REF_V6 = 16,        /* tsmttsm.hh:230-298 */
  ORACLE_REF_V7 = 17,        /* tsmttsm.hh:303-363 */
  ORACLE_REF_V8 = 18,        /* tsmttsm.hh:445-586 */
  ORACLE_REF_V9 = 19         /* tsmttsm.hh:368-439 */
};
typedef struct {
  int variant;     /* ORACLE_REF_* */
  int tm, tn;      /* register tile (generated kernels) / max_tile_size (v2-v9) */
  int64_t stride;  /* global range (work-items) */
  int local_size;  /* nd_range work-group size */
  int sg_size_gen; /* sub-group size the generator assumed (its asserts)        */
  int sg_size_rt;  /* sub-group size the runtime reports (its index math)       */
  int unroll;      /* generator --unroll                                        */
  int worker_mult; /* tsmttsm8 worker multiplier                                */
  int num_cu;      /* tsmttsm8: device compute units                            */
} oracle_ref_cfg;

int oracle_ref_run(const oracle_ref_cfg* cfg, int64_t n, int k, int m,
                   const double* A, const double* B, double* C);

/* The generator's stride rule (sycl_codegen_new.py:563-580, lcm at :31-32):
 * smallest multiple of lcm(local, num_tiles, sg*num_tiles) over all (M, TM)
 * with TM | M, scaled to about local*target_wgs work-items. */
int64_t oracle_suggest_stride(const int* Ms, int nM, const int* tiles, int nT,
                              int local_size, int sg_size, int target_wgs);
/* validate_stride (sycl_codegen_new.py:530-560): 1 if valid, 0 otherwise. */
int oracle_validate_stride(const int* Ms, int nM, const int* tiles, int nT,
                           int64_t stride, int local_size, int sg_size);

/* Timed CPU baseline: runs one reference kernel `warmup` + `runs` times on the
 * given data (C zeroed outside the timer, as main.cc:118-126) and stores the
 * per-run wall times in ns.  Returns the number of runs done (may stop early
 * once `budget_s` seconds of timed work have elapsed). */
int oracle_ref_time(const oracle_ref_cfg* cfg, int64_t n, int k, int m,
                    const double* A, const double* B, double* C,
                    int warmup, int runs, double budget_s, double* times_ns);
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
