/*
 * CPU oracle for C = A^T B: synthetic inputs, storage rounding, the Dot2
 * compensated and the exact int128 oracles, the reference's naive host loop
 * and acceptance rule, and the backward closed form.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * that TwoSum/TwoProduct stay exact; fma() is called explicitly.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SM_GAMMA  0x9e3779b97f4a7c15ULL
#define SM_STREAM 0xd1b54a32d192ed03ULL

static inline uint64_t sm_finalize(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

uint64_t oracle_rng_word(uint64_t seed, uint64_t stream, uint64_t idx) {
  return sm_finalize(seed + stream * SM_STREAM + (idx + 1) * SM_GAMMA);
}

double oracle_rng_value(int dist, uint64_t seed, uint64_t stream, uint64_t idx) {
  if (dist == ORACLE_DIST_UNIFORM) {
    uint64_t w = oracle_rng_word(seed, stream, idx);
    return ldexp((double)(w >> 11), -52) - 1.0;
  }
  if (dist == ORACLE_DIST_POW2) {
    uint64_t w = oracle_rng_word(seed, stream, idx);
    int e = (int)((w >> 11) % 41u) - 20;
    double v = ldexp(1.0, e);
    return (w >> 63) ? -v : v;
  }
  /* normal(0,1): Box-Muller on two words */
  uint64_t w1 = oracle_rng_word(seed, stream, 2 * idx);
  uint64_t w2 = oracle_rng_word(seed, stream, 2 * idx + 1);
  double u1 = ldexp((double)((w1 >> 11) + 1), -53); /* (0, 1] */
  double u2 = ldexp((double)(w2 >> 11), -53);       /* [0, 1) */
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

void oracle_fill_f64(int dist, uint64_t seed, uint64_t stream,
                     int64_t rows, int64_t cols, double* out, int64_t ld) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t c = 0; c < cols; ++c)
      out[r * ld + c] = oracle_rng_value(dist, seed, stream, (uint64_t)(r * cols + c));
}

/* ---- rounding ----------------------------------------------------------- */

float oracle_round_f32(double x) { return (float)x; /* IEEE cast: RNE */ }

uint16_t oracle_round_bf16(double x) {
  /* One rounding from FP64 to an 8-bit significand: round the FP64 pattern
   * to nearest-even at bit 45 (52 - 7 kept fraction bits), then the value is
   * exact in FP32 and its top 16 bits are the bf16 pattern.  Below 2^-126
   * (the bf16 subnormals, multiples of 2^-133) the value is rounded to the
   * nearest multiple of 2^-133 instead, ties to even.  Finite inputs. */
  uint64_t b;
  memcpy(&b, &x, 8);
  if ((b & 0x7fffffffffffffffULL) == 0) return (uint16_t)((b >> 48) & 0x8000);
  if (fabs(x) < 0x1p-126) {
    /* the scaling is exact; nearbyint rounds to nearest even (default mode) */
    const double q = nearbyint(fabs(x) * 0x1p133); /* 0 .. 128; 128 = 2^-126 */
    return (uint16_t)(((b >> 48) & 0x8000) | (uint16_t)q);
  }
  uint64_t lsb = (b >> 45) & 1u;
  b += 0x00000fffffffffffULL + lsb;
  b &= ~0x00001fffffffffffULL;
  double r;
  memcpy(&r, &b, 8);
  float f = (float)r;
  uint32_t u;
  memcpy(&u, &f, 4);
  return (uint16_t)(u >> 16);
}

double oracle_bf16_to_f64(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return (double)f;
}

void oracle_round_matrix_f32(int64_t count, const double* in, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) out[i] = (float)in[i];
}

void oracle_round_matrix_bf16(int64_t count, const double* in, uint16_t* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) out[i] = oracle_round_bf16(in[i]);
}

/* ---- Dot2 ---------------------------------------------------------------- */

static inline void two_sum(double a, double b, double* s, double* e) {
  double x = a + b;
  double z = x - a;
  *e = (a - (x - z)) + (b - z);
  *s = x;
}

#define DOT2_CHUNKS 64

void oracle_atb_dot2(int64_t n, int k, int m,
                     const double* A, int64_t lda, const double* B, int64_t ldb,
                     double* C, int64_t ldc, double* absC) {
  const int km = k * m;
  double* s = (double*)calloc((size_t)DOT2_CHUNKS * km, sizeof(double));
  double* e = (double*)calloc((size_t)DOT2_CHUNKS * km, sizeof(double));
  double* a = (double*)calloc((size_t)DOT2_CHUNKS * km, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < DOT2_CHUNKS; ++c) {
    int64_t r0 = n * c / DOT2_CHUNKS, r1 = n * (c + 1) / DOT2_CHUNKS;
    double* sc = s + (size_t)c * km;
    double* ec = e + (size_t)c * km;
    double* ac = a + (size_t)c * km;
    for (int64_t r = r0; r < r1; ++r) {
      const double* ar = A + r * lda;
      const double* br = B + r * ldb;
      for (int i = 0; i < k; ++i) {
        double ai = ar[i];
        for (int j = 0; j < m; ++j) {
          double bj = br[j];
          double p = ai * bj;
          double pe = fma(ai, bj, -p); /* TwoProduct: ai*bj = p + pe exactly */
          double t, q;
          two_sum(sc[i * m + j], p, &t, &q);
          sc[i * m + j] = t;
          ec[i * m + j] += q + pe;
          ac[i * m + j] += fabs(p);
        }
      }
    }
  }
  for (int idx = 0; idx < km; ++idx) {
    double S = 0.0, E = 0.0, AS = 0.0;
    for (int c = 0; c < DOT2_CHUNKS; ++c) {
      double t, q;
      two_sum(S, s[(size_t)c * km + idx], &t, &q);
      S = t;
      E += q + e[(size_t)c * km + idx];
      AS += a[(size_t)c * km + idx];
    }
    C[(idx / m) * ldc + idx % m] = S + E;
    if (absC) absC[(idx / m) * m + idx % m] = AS;
  }
  free(s);
  free(e);
  free(a);
}

/* ---- exact int128 ----------------------------------------------------------- */

static int to_scaled_int(double x, int q, int64_t* out) {
  double v = ldexp(x, q);
  if (!(fabs(v) < 4611686018427387904.0)) return -1; /* 2^62 */
  int64_t iv = (int64_t)v;
  if ((double)iv != v) return -1;
  *out = iv;
  return 0;
}

int oracle_atb_exact(int64_t n, int k, int m,
                     const double* A, int64_t lda, int qa,
                     const double* B, int64_t ldb, int qb,
                     double* C, int64_t ldc) {
  const int km = k * m;
  __int128* acc = (__int128*)calloc((size_t)DOT2_CHUNKS * km, sizeof(__int128));
  int status = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < DOT2_CHUNKS; ++c) {
    int64_t r0 = n * c / DOT2_CHUNKS, r1 = n * (c + 1) / DOT2_CHUNKS;
    __int128* ac = acc + (size_t)c * km;
    int64_t ai[64], bj[64];
    int bad = 0;
    for (int64_t r = r0; r < r1 && !bad; ++r) {
      for (int i = 0; i < k; ++i)
        if (to_scaled_int(A[r * lda + i], qa, &ai[i])) bad = -1;
      for (int j = 0; j < m; ++j)
        if (to_scaled_int(B[r * ldb + j], qb, &bj[j])) bad = -1;
      if (bad) break;
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < m; ++j) {
          __int128 p = (__int128)ai[i] * (__int128)bj[j];
          __int128 t;
          if (__builtin_add_overflow(ac[i * m + j], p, &t)) { bad = -2; break; }
          ac[i * m + j] = t;
        }
    }
    if (bad) {
#pragma omp critical
      status = status ? status : bad;
    }
  }
  if (status == 0) {
    for (int idx = 0; idx < km; ++idx) {
      __int128 tot = 0;
      for (int c = 0; c < DOT2_CHUNKS; ++c) {
        __int128 t;
        if (__builtin_add_overflow(tot, acc[(size_t)c * km + idx], &t)) { status = -2; break; }
        tot = t;
      }
      if (status) break;
      /* (double) of an __int128 rounds to nearest-even (libgcc __floattidf);
       * the power-of-two scaling that follows is exact. */
      C[(idx / m) * ldc + idx % m] = ldexp((double)tot, -(qa + qb));
    }
  }
  free(acc);
  return status;
}

/* ---- reference host oracle and acceptance rule ------------------------------- */

void oracle_atb_naive(int64_t n, int k, int m,
                      const double* A, int64_t lda, const double* B, int64_t ldb,
                      double* C, int64_t ldc) {
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < m; ++j) {
      double acc = 0.0;
      for (int64_t r = 0; r < n; ++r) acc += A[r * lda + i] * B[r * ldb + j];
      C[i * ldc + j] = acc;
    }
}

int64_t oracle_first_mismatch(int64_t count, const double* ref, const double* got,
                              double rel_tol, double abs_tol) {
  for (int64_t i = 0; i < count; ++i) {
    double diff = fabs(ref[i] - got[i]);
    double mag = fmax(fabs(ref[i]), fabs(got[i]));
    if (diff > abs_tol && diff > rel_tol * mag) return i;
    if (diff != diff) return i; /* NaN never passes */
  }
  return -1;
}

/* ---- backward ------------------------------------------------------------- */

static double dot2_strided(int len, const double* x, int64_t incx, const double* y, int64_t incy) {
  double s = 0.0, c = 0.0;
  for (int t = 0; t < len; ++t) {
    double a = x[t * incx], b = y[t * incy];
    double p = a * b, pe = fma(a, b, -p), u, q;
    two_sum(s, p, &u, &q);
    s = u;
    c += q + pe;
  }
  return s + c;
}

void oracle_atb_backward(int64_t n, int k, int m,
                         const double* A, int64_t lda, const double* B, int64_t ldb,
                         const double* G, int64_t ldg,
                         double* dA, int64_t ldda, double* dB, int64_t lddb) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    if (dA)
      for (int i = 0; i < k; ++i) /* dA[r,i] = sum_j B[r,j] G[i,j] */
        dA[r * ldda + i] = dot2_strided(m, B + r * ldb, 1, G + (int64_t)i * ldg, 1);
    if (dB)
      for (int j = 0; j < m; ++j) /* dB[r,j] = sum_i A[r,i] G[i,j] */
        dB[r * lddb + j] = dot2_strided(k, A + r * lda, 1, G + j, ldg);
  }
}
