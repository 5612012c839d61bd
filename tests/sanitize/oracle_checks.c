/* Drives every oracle entry point on small inputs in a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile):
 * any out-of-bounds access, use of uninitialised stack, signed overflow or
 * misaligned access aborts the run.  Also checks that the results agree with
 * each other (exact vs Dot2 vs naive, closed-form backward). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                       \
  do {                                         \
    if (!(cond)) {                             \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);            \
      fprintf(stderr, "\n");                   \
      ++failures;                              \
    }                                          \
  } while (0)

static double* mat(int64_t r, int64_t c, int dist, uint64_t stream) {
  double* x = malloc((size_t)(r * c > 0 ? r * c : 1) * sizeof(double));
  oracle_fill_f64(dist, 20261203, stream, r, c, x, c);
  return x;
}

static void forward(int64_t n, int k, int m, int dist) {
  double* A = mat(n, k, dist, 1);
  double* B = mat(n, m, dist, 2);
  double* Ce = malloc((size_t)k * m * sizeof(double));
  double* Cd = malloc((size_t)k * m * sizeof(double));
  double* Cn = malloc((size_t)k * m * sizeof(double));
  double* Ab = malloc((size_t)k * m * sizeof(double));
  const int q = dist == ORACLE_DIST_POW2 ? 20 : 52;
  CHECK(oracle_atb_exact(n, k, m, A, k, q, B, m, q, Ce, m) == 0, "exact n=%lld", (long long)n);
  oracle_atb_dot2(n, k, m, A, k, B, m, Cd, m, Ab);
  oracle_atb_naive(n, k, m, A, k, B, m, Cn, m);
  for (int i = 0; i < k * m; ++i) {
    CHECK(fabs(Cd[i] - Ce[i]) <= 2.3e-16 * fabs(Ce[i]), "dot2 vs exact at %d", i);
    CHECK(fabs(Cn[i] - Ce[i]) <= 4.0 * sqrt((double)(n > 0 ? n : 1)) * 2.3e-16 * Ab[i] + 1e-300,
          "naive vs exact at %d", i);
  }
  CHECK(oracle_first_mismatch(k * m, Ce, Cn, 1e-10, 1e-12) == -1, "acceptance rule");
  free(A); free(B); free(Ce); free(Cd); free(Cn); free(Ab);
}

static void reference_kernels(void) {
  const int n = 1000;
  for (int M = 1; M <= 16; M *= 2) {
    double* A = mat(n, M, 0, 3);
    double* B = mat(n, M, 0, 4);
    double* Ce = malloc((size_t)M * M * sizeof(double));
    double* C = malloc((size_t)M * M * sizeof(double));
    oracle_atb_exact(n, M, M, A, M, 52, B, M, 52, Ce, M);
    const int variants[] = {ORACLE_REF_GEN_BASIC, ORACLE_REF_GEN_LF, ORACLE_REF_GEN_SG,
                            ORACLE_REF_GEN_SG_LF, ORACLE_REF_V1, ORACLE_REF_V2, ORACLE_REF_V3,
                            ORACLE_REF_V4, ORACLE_REF_V5, ORACLE_REF_V6, ORACLE_REF_V7,
                            ORACLE_REF_V8, ORACLE_REF_V9};
    for (size_t v = 0; v < sizeof variants / sizeof variants[0]; ++v)
      for (int tile = 1; tile <= 4; tile *= 2) {
        oracle_ref_cfg c = {variants[v], tile, tile, 4096, 32, 32, 32, 1, 64, 4};
        memset(C, 0, (size_t)M * M * sizeof(double));
        const int rc = oracle_ref_run(&c, n, M, M, A, B, C);
        if (rc != 0) continue; /* configuration the variant does not support */
        CHECK(oracle_first_mismatch(M * M, Ce, C, 1e-10, 1e-12) == -1, "ref variant %d M=%d tile=%d",
              variants[v], M, tile);
      }
    free(A); free(B); free(Ce); free(C);
  }
}

static void backward(void) {
  const int64_t n = 300;
  const int k = 5, m = 3;
  double* A = mat(n, k, 0, 5);
  double* B = mat(n, m, 0, 6);
  double* G = mat(k, m, ORACLE_DIST_NORMAL, 7);
  double* dA = malloc((size_t)n * k * sizeof(double));
  double* dB = malloc((size_t)n * m * sizeof(double));
  oracle_atb_backward(n, k, m, A, k, B, m, G, m, dA, k, dB, m);
  for (int64_t r = 0; r < n; ++r)
    for (int i = 0; i < k; ++i) {
      double s = 0;
      for (int j = 0; j < m; ++j) s += B[r * m + j] * G[i * m + j];
      CHECK(fabs(s - dA[r * k + i]) <= 1e-14, "dA");
    }
  free(A); free(B); free(G); free(dA); free(dB);
}

static void strides_and_rounding(void) {
  const int Ms[] = {4, 8, 16}, tiles[] = {2, 4};
  CHECK(oracle_suggest_stride(Ms, 3, tiles, 2, 256, 32, 2048) == 524288, "default stride");
  CHECK(oracle_validate_stride(Ms, 3, tiles, 2, 524288, 256, 32) == 1, "valid stride");
  CHECK(oracle_validate_stride(Ms, 3, tiles, 2, 1000, 256, 32) == 0, "invalid stride");
  const double xs[] = {0.0, -0.0, 1.0, 1.0 + 0x1p-8, 1.0 + 0x1p-9, 3.0e38, -1.5e-40, 65504.0};
  for (size_t i = 0; i < sizeof xs / sizeof xs[0]; ++i) {
    const uint16_t h = oracle_round_bf16(xs[i]);
    /* half an ulp: 2^-8 relative, or half the smallest subnormal (2^-134) */
    CHECK(fabs(oracle_bf16_to_f64(h) - xs[i]) <= fmax(fabs(xs[i]) * 0x1p-8, 0x1p-134), "bf16 %g", xs[i]);
    CHECK((double)oracle_round_f32(xs[i]) == (double)(float)xs[i], "f32 %g", xs[i]);
  }
  double in[5] = {1.0, 2.5, -3.25, 1e-3, 7.0};
  float f[5];
  uint16_t b[5];
  oracle_round_matrix_f32(5, in, f);
  oracle_round_matrix_bf16(5, in, b);
  CHECK(f[2] == -3.25f && oracle_bf16_to_f64(b[1]) == 2.5, "matrix rounding");
}

int main(void) {
  forward(0, 4, 4, 0);
  forward(1, 1, 1, 0);
  forward(1001, 5, 3, 0);
  forward(4096, 8, 8, 0);
  forward(2048, 16, 32, ORACLE_DIST_POW2);
  forward(777, 64, 1, ORACLE_DIST_POW2);
  /* an entry off the 2^-q grid is reported, not read past */
  double A[4] = {0.5, 0x1p-60, 1.0, 2.0}, B[4] = {1.0, 1.0, 1.0, 1.0}, C[4];
  CHECK(oracle_atb_exact(2, 2, 2, A, 2, 20, B, 2, 20, C, 2) == -1, "off-grid entry");
  reference_kernels();
  backward();
  strides_and_rounding();
  double times[3];
  double* X = mat(4096, 8, 0, 9);
  double* Y = mat(4096, 8, 0, 10);
  double Cc[64];
  oracle_ref_cfg c = {ORACLE_REF_GEN_SG, 4, 4, 524288, 256, 32, 32, 1, 64, 4};
  CHECK(oracle_ref_time(&c, 4096, 8, 8, X, Y, Cc, 1, 3, 5.0, times) == 3, "ref_time");
  CHECK(oracle_num_threads() >= 1, "threads");
  free(X); free(Y);
  if (failures) {
    fprintf(stderr, "%d oracle checks failed\n", failures);
    return 1;
  }
  printf("oracle checks passed\n");
  return 0;
}
