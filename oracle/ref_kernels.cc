// CPU restatement of the reference's SYCL kernels for C += A^T B.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h): the timed CPU baseline and the
// subject of the work-decomposition coverage tests.  Each kernel keeps the
// reference's index math, per-work-item accumulators and atomic adds:
//   * a SYCL range kernel becomes an OpenMP loop over work-items;
//   * an nd_range work-group runs on one host thread with its work-items in
//     sequence between the barrier phases (all barriers in these kernels lie
//     before or after the row loop, so the phases are independent);
//   * a sub-group reduce_over_group sums lanes in lane order;
//   * sycl::atomic_ref<T,relaxed,device> += v becomes `omp atomic`.
// The reference uses `int` for rows and products; this restatement uses
// int64_t so that it can run the larger inputs the GPU path accepts.  For
// every input the reference itself could index (n*k < 2^31) the visited
// elements are identical.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <omp.h>

namespace {

using i64 = int64_t;

inline void atomic_add(double* p, double v) {
#pragma omp atomic
  *p += v;
}

// ---------------------------------------------------------------------------
// Generated kernels, sycl_codegen_new.py:142-363.
// A[(k) * M + m * mthreads + midx], B[(k) * N + n * nthreads + nidx] (:63-67).
// ---------------------------------------------------------------------------
template <int TM, int TN>
struct GenTile {
  int M, N, mthreads, nthreads, midx, nidx;
  const double* A;
  const double* B;
  double acc[TM][TN];

  void zero() {
    for (int a = 0; a < TM; ++a)
      for (int b = 0; b < TN; ++b) acc[a][b] = 0.0;
  }
  void load(i64 k, double* va, double* vb) const {
    for (int a = 0; a < TM; ++a) va[a] = A[k * M + a * mthreads + midx];
    for (int b = 0; b < TN; ++b) vb[b] = B[k * N + b * nthreads + nidx];
  }
  void block(const double* va, const double* vb) {
    for (int a = 0; a < TM; ++a)
      for (int b = 0; b < TN; ++b) acc[a][b] += va[a] * vb[b];
  }
  // Row loop of the basic kernel (:172-209); unroll only regroups the same
  // rows in the same order, so one loop covers both branches.
  void rows_plain(i64 K, i64 k_start, i64 k_stride, int unroll) {
    double va[TM], vb[TN];
    i64 k = k_start;
    if (unroll > 1) {
      for (; k + (i64)(unroll - 1) * k_stride < K; k += k_stride * unroll)
        for (int u = 0; u < unroll; ++u) {
          load(k + u * k_stride, va, vb);
          block(va, vb);
        }
    }
    for (; k < K; k += k_stride) {
      load(k, va, vb);
      block(va, vb);
    }
  }
  // Leap-frog body (:74-135): per-slot register prefetch, guarded next
  // loads, then a per-slot drain.
  void rows_leapfrog(i64 K, i64 k_start, i64 k_stride, int unroll) {
    double vA[8][TM] = {}, vB[8][TN] = {};
    for (int u = 0; u < unroll; ++u)
      if (k_start + u * k_stride < K) load(k_start + u * k_stride, vA[u], vB[u]);
    i64 k;
    for (k = k_start; k < K - k_stride * unroll; k += k_stride * unroll) {
      for (int u = 0; u < unroll; ++u) {
        i64 next_k = k + k_stride * (unroll + u);
        double nA[TM] = {}, nB[TN] = {};
        if (next_k < K) load(next_k, nA, nB);
        block(vA[u], vB[u]);
        std::memcpy(vA[u], nA, sizeof nA);
        std::memcpy(vB[u], nB, sizeof nB);
      }
    }
    for (int u = 0; u < unroll; ++u)
      if (k + u * k_stride < K) block(vA[u], vB[u]);
  }
  i64 cidx(int a, int b) const { return (i64)(a * mthreads + midx) * N + (b * nthreads + nidx); }
};

template <int TM, int TN>
int gen_basic(const oracle_ref_cfg& c, i64 K, int M, int N, const double* A, const double* B,
              double* C, bool lf) {
  const int mthreads = M / TM, nthreads = N / TN, nt = mthreads * nthreads;
  const i64 stride = c.stride;
  if (c.unroll < 1 || c.unroll > 8) return -1;
#pragma omp parallel for schedule(static)
  for (i64 tid = 0; tid < stride; ++tid) {
    GenTile<TM, TN> t{M, N, mthreads, nthreads, (int)((tid % nt) % mthreads),
                      (int)((tid % nt) / mthreads), A, B, {}};
    t.zero();
    const i64 k_start = tid / nt, k_stride = stride / nt;  // :166-167
    if (lf)
      t.rows_leapfrog(K, k_start, k_stride, c.unroll);
    else
      t.rows_plain(K, k_start, k_stride, c.unroll);
    for (int a = 0; a < TM; ++a)
      for (int b = 0; b < TN; ++b) atomic_add(&C[t.cidx(a, b)], t.acc[a][b]);  // :218-225
  }
  return 0;
}

template <int TM, int TN>
int gen_sg(const oracle_ref_cfg& c, i64 K, int M, int N, const double* A, const double* B,
           double* C, bool lf) {
  const int mthreads = M / TM, nthreads = N / TN, nt = mthreads * nthreads;
  const i64 stride = c.stride;
  const int L = c.local_size, sg_rt = c.sg_size_rt;
  // Generator-time asserts use the generator's sub-group size (:250-253).
  if (c.sg_size_gen <= 0 || stride % c.sg_size_gen) return -1;
  if ((stride / c.sg_size_gen) % nt) return -1;
  if (L <= 0 || stride % L || sg_rt <= 0 || L % sg_rt) return -1;
  if (c.unroll < 1 || c.unroll > 8) return -1;
  const i64 groups = stride / L;
  const int sg_per_wg = L / sg_rt;
  const i64 total_sg_rt = stride / sg_rt;  // :282
  // When the runtime sub-group size makes total_sg_rt < num_tiles the
  // reference's row loop never advances (k_stride == 0): refuse instead.
  if ((total_sg_rt / nt) * sg_rt <= 0) return -1;
#pragma omp parallel for schedule(static)
  for (i64 g = 0; g < groups; ++g) {
    for (int s = 0; s < sg_per_wg; ++s) {
      const i64 sg_gid = g * sg_per_wg + s;  // :280-281
      const int midx = (int)((sg_gid % nt) % mthreads), nidx = (int)((sg_gid % nt) / mthreads);
      const i64 k_stride = (total_sg_rt / nt) * sg_rt;  // :287
      double red[TM][TN] = {};
      for (int lane = 0; lane < sg_rt; ++lane) {
        GenTile<TM, TN> t{M, N, mthreads, nthreads, midx, nidx, A, B, {}};
        t.zero();
        const i64 k_start = (sg_gid / nt) * sg_rt + lane;  // :286
        if (lf)
          t.rows_leapfrog(K, k_start, k_stride, c.unroll);
        else
          t.rows_plain(K, k_start, k_stride, c.unroll);
        for (int a = 0; a < TM; ++a)
          for (int b = 0; b < TN; ++b) red[a][b] += t.acc[a][b];  // reduce_over_group (:353)
      }
      GenTile<TM, TN> t{M, N, mthreads, nthreads, midx, nidx, A, B, {}};
      for (int a = 0; a < TM; ++a)
        for (int b = 0; b < TN; ++b) atomic_add(&C[t.cidx(a, b)], red[a][b]);  // :354-357
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Hand-written variants, tsmttsm.hh (square: A, B are K x M, C is M x M).
// ---------------------------------------------------------------------------
int v1(const oracle_ref_cfg& c, i64 K, int M, const double* A, const double* B, double* C) {
  const i64 stride = c.stride;
#pragma omp parallel for schedule(static)
  for (i64 item = 0; item < stride; ++item) {
    double cl[64 * 64];
    std::fill(cl, cl + M * M, 0.0);
    for (i64 k = item; k < K; k += stride)  // tsmttsm.hh:22-26
      for (int a = 0; a < M; ++a)
        for (int b = 0; b < M; ++b) cl[a * M + b] += A[k * M + a] * B[k * M + b];
    for (int a = 0; a < M; ++a)
      for (int b = 0; b < M; ++b) atomic_add(&C[a * M + b], cl[a * M + b]);
  }
  return 0;
}

// Tile geometry of v2..v7, v9: contiguous (v2, v3, v5, v6) or "transposed"
// (v4, v7, v9: tsmttsm.hh:142-151, 388-389, 432).
template <int TS>
struct SqTile {
  int M, tpd, row0, col0;
  bool transposed;
  int arow(int i) const { return transposed ? row0 + i * tpd : row0 + i; }
  int bcol(int i) const { return transposed ? col0 + i * tpd : col0 + i; }
  static SqTile make(int M, i64 tile_idx, bool transposed) {
    const int tpd = M / TS;
    SqTile t{M, tpd, 0, 0, transposed};
    if (transposed) {
      t.row0 = (int)(tile_idx % tpd);
      t.col0 = (int)(tile_idx / tpd);
    } else {
      t.row0 = (int)(tile_idx / tpd) * TS;
      t.col0 = (int)(tile_idx % tpd) * TS;
    }
    return t;
  }
};

template <int TS>
void sq_rows(const SqTile<TS>& t, i64 K, i64 k_start, i64 k_stride, const double* A,
             const double* B, double (&cl)[TS][TS], bool leapfrog) {
  const int M = t.M;
  if (!leapfrog) {
    for (i64 k = k_start; k < K; k += k_stride) {
      double av[TS], bv[TS];
      for (int i = 0; i < TS; ++i) {
        av[i] = A[k * M + t.arow(i)];
        bv[i] = B[k * M + t.bcol(i)];
      }
      for (int a = 0; a < TS; ++a)
        for (int b = 0; b < TS; ++b) cl[a][b] += av[a] * bv[b];
    }
    return;
  }
  // v5 (tsmttsm.hh:193-215) and v7 (:327-353): the next row is loaded
  // before the current one is used.  v5's first load is unguarded in the
  // reference; guarding it changes no result (the loop body never runs
  // when k_start >= K).
  double cur[2][TS], nxt[2][TS];
  if (k_start < K)
    for (int i = 0; i < TS; ++i) {
      cur[0][i] = A[k_start * M + t.arow(i)];
      cur[1][i] = B[k_start * M + t.bcol(i)];
    }
  for (i64 k = k_start; k < K; k += k_stride) {
    if (k + k_stride < K)
      for (int i = 0; i < TS; ++i) {
        nxt[0][i] = A[(k + k_stride) * M + t.arow(i)];
        nxt[1][i] = B[(k + k_stride) * M + t.bcol(i)];
      }
    for (int a = 0; a < TS; ++a)
      for (int b = 0; b < TS; ++b) cl[a][b] += cur[0][a] * cur[1][b];
    std::memcpy(cur, nxt, sizeof cur);
  }
}

template <int TS>
int sq_range(const oracle_ref_cfg& c, i64 K, int M, const double* A, const double* B, double* C,
             bool transposed, bool leapfrog) {
  if ((M * M) % (TS * TS)) return -1;  // static_assert, tsmttsm.hh:43
  const int nt = (M * M) / (TS * TS);
  const i64 stride = c.stride;
  // fewer work-items than tiles: the row stride stride / nt is 0 and the
  // reference's row loop would never end (an invalid configuration)
  if (stride / nt == 0) return -1;
#pragma omp parallel for schedule(static)
  for (i64 tid = 0; tid < stride; ++tid) {
    double cl[TS][TS] = {};
    const i64 k_start = tid / nt, k_stride = stride / nt;  // tsmttsm.hh:53-54
    const auto t = SqTile<TS>::make(M, tid % nt, transposed);
    sq_rows<TS>(t, K, k_start, k_stride, A, B, cl, leapfrog);
    for (int a = 0; a < TS; ++a)
      for (int b = 0; b < TS; ++b) atomic_add(&C[(i64)t.arow(a) * M + t.bcol(b)], cl[a][b]);
  }
  return 0;
}

template <int TS>
int sq_ndrange(const oracle_ref_cfg& c, i64 K, int M, const double* A, const double* B,
               double* C, bool transposed) {
  if ((M * M) % (TS * TS)) return -1;
  const int nt = (M * M) / (TS * TS);
  const int L = c.local_size;
  if (L <= 0 || c.stride % L) return -1;
  const i64 num_groups = c.stride / L;
  // fewer work-groups than tiles: the row stride (num_groups / nt) * L is 0
  // and the reference's row loop would never end (an invalid configuration)
  if (num_groups / nt == 0) return -1;
#pragma omp parallel for schedule(static)
  for (i64 g = 0; g < num_groups; ++g) {
    const auto t = SqTile<TS>::make(M, g % nt, transposed);  // tsmttsm.hh:248-250
    double shared[TS][TS] = {};                               // C_shared, zeroed by lid 0
    for (int lid = 0; lid < L; ++lid) {
      double cl[TS][TS] = {};
      const i64 k_start = (g / nt) * L + lid, k_stride = (num_groups / nt) * L;  // :253-254
      sq_rows<TS>(t, K, k_start, k_stride, A, B, cl, false);
      for (int a = 0; a < TS; ++a)
        for (int b = 0; b < TS; ++b) shared[a][b] += cl[a][b];  // local atomics (:279-283)
    }
    for (int a = 0; a < TS; ++a)  // leader's global atomics (:288-295)
      for (int b = 0; b < TS; ++b) atomic_add(&C[(i64)t.arow(a) * M + t.bcol(b)], shared[a][b]);
  }
  return 0;
}

template <int TM>
int v8(const oracle_ref_cfg& c, i64 K, int M, const double* A, const double* B, double* C) {
  constexpr int TN = TM;
  const int ntm = (M + TM - 1) / TM, ntn = (M + TN - 1) / TN, nt = ntm * ntn;  // :451-453
  const int wg = 256;
  const i64 wg_per_tile = std::max<i64>(1, (i64)c.worker_mult * c.num_cu / wg);  // :458
  const i64 workers_per_tile = wg_per_tile * wg;
  const i64 groups = nt * wg_per_tile;
#pragma omp parallel for schedule(static)
  for (i64 g = 0; g < groups; ++g) {
    const int tile_idx = (int)(g % nt);
    const int tile_m = tile_idx / ntn, tile_n = tile_idx % ntn;  // :474-475
    int mi[TM], ni[TN];
    for (int a = 0; a < TM; ++a) mi[a] = (TM < M) ? tile_m + a * ntm : a;  // :479-495
    for (int b = 0; b < TN; ++b) ni[b] = (TM < M) ? tile_n + b * ntn : b;
    double local_c[TM * TN] = {};
    for (int lid = 0; lid < wg; ++lid) {
      const i64 worker = (g / nt) * wg + lid;  // :471
      double cl[TM][TN] = {};
      for (i64 k = worker; k < K; k += workers_per_tile)
        for (int b = 0; b < TN; ++b)
          for (int a = 0; a < TM; ++a) cl[a][b] += A[k * M + mi[a]] * B[k * M + ni[b]];
      for (int a = 0; a < TM; ++a)
        for (int b = 0; b < TN; ++b) local_c[a * TN + b] += cl[a][b];  // :565-571
    }
    for (int a = 0; a < TM; ++a)
      for (int b = 0; b < TN; ++b) atomic_add(&C[(i64)mi[a] * M + ni[b]], local_c[a * TN + b]);
  }
  return 0;
}

bool pow2_le64(int x) { return x >= 1 && x <= 64 && (x & (x - 1)) == 0; }

template <int TM, int TN>
int gen_dispatch(const oracle_ref_cfg& c, i64 n, int k, int m, const double* A, const double* B,
                 double* C) {
  switch (c.variant) {
    case ORACLE_REF_GEN_BASIC: return gen_basic<TM, TN>(c, n, k, m, A, B, C, false);
    case ORACLE_REF_GEN_LF:    return gen_basic<TM, TN>(c, n, k, m, A, B, C, true);
    case ORACLE_REF_GEN_SG:    return gen_sg<TM, TN>(c, n, k, m, A, B, C, false);
    case ORACLE_REF_GEN_SG_LF: return gen_sg<TM, TN>(c, n, k, m, A, B, C, true);
  }
  return -1;
}

template <int TS>
int sq_dispatch(const oracle_ref_cfg& c, i64 n, int M, const double* A, const double* B,
                double* C) {
  switch (c.variant) {
    case ORACLE_REF_V2:
    case ORACLE_REF_V3: return sq_range<TS>(c, n, M, A, B, C, false, false);
    case ORACLE_REF_V4: return sq_range<TS>(c, n, M, A, B, C, true, false);
    case ORACLE_REF_V5: return sq_range<TS>(c, n, M, A, B, C, false, true);
    case ORACLE_REF_V7: return sq_range<TS>(c, n, M, A, B, C, true, true);
    case ORACLE_REF_V6: return sq_ndrange<TS>(c, n, M, A, B, C, false);
    case ORACLE_REF_V9: return sq_ndrange<TS>(c, n, M, A, B, C, true);
  }
  return -1;
}

int ref_run(const oracle_ref_cfg& c, i64 n, int k, int m, const double* A, const double* B,
            double* C) {
  if (!pow2_le64(k) || !pow2_le64(m) || c.stride <= 0 || n < 0) return -1;
  if (c.variant <= ORACLE_REF_GEN_SG_LF) {
    if (c.tm < 1 || c.tn < 1 || k % c.tm || m % c.tn) return -1;
    if (c.stride % ((k / c.tm) * (m / c.tn))) return -1;  // validate_stride (:547)
#define GEN_CASE(a, b) \
  if (c.tm == a && c.tn == b) return gen_dispatch<a, b>(c, n, k, m, A, B, C);
    GEN_CASE(1, 1) GEN_CASE(2, 2) GEN_CASE(4, 4) GEN_CASE(8, 8)
    GEN_CASE(1, 2) GEN_CASE(2, 1) GEN_CASE(2, 4) GEN_CASE(4, 2) GEN_CASE(4, 8) GEN_CASE(8, 4)
#undef GEN_CASE
    return -1;
  }
  if (k != m) return -1;  // the hand-written variants are square-only
  const int M = k;
  if (c.variant == ORACLE_REF_V1) return v1(c, n, M, A, B, C);
  if (c.variant == ORACLE_REF_V8) {
    const int TM = (M * M <= 16) ? M : ((M * M <= 64) ? 2 : 4);  // tsmttsm.hh:449-450
    if (c.num_cu <= 0 || c.worker_mult <= 0) return -1;
    if (TM == 1) return v8<1>(c, n, M, A, B, C);
    if (TM == 2) return v8<2>(c, n, M, A, B, C);
    if (TM == 4) return v8<4>(c, n, M, A, B, C);
    return -1;
  }
  const int TS = std::min(c.tm, M);  // tile_size = min(max_tile_size, M)
  switch (TS) {
    case 1: return sq_dispatch<1>(c, n, M, A, B, C);
    case 2: return sq_dispatch<2>(c, n, M, A, B, C);
    case 4: return sq_dispatch<4>(c, n, M, A, B, C);
    case 8: return sq_dispatch<8>(c, n, M, A, B, C);
  }
  return -1;
}

i64 lcm64(i64 a, i64 b) { return a / std::gcd(a, b) * b; }

}  // namespace

extern "C" int oracle_ref_run(const oracle_ref_cfg* cfg, int64_t n, int k, int m,
                              const double* A, const double* B, double* C) {
  if (!cfg || !A || !B || !C) return -1;
  return ref_run(*cfg, n, k, m, A, B, C);
}

extern "C" int64_t oracle_suggest_stride(const int* Ms, int nM, const int* tiles, int nT,
                                         int local_size, int sg_size, int target_wgs) {
  i64 req = local_size;
  for (int a = 0; a < nM; ++a)
    for (int b = 0; b < nT; ++b) {
      if (Ms[a] % tiles[b]) continue;
      const i64 mt = Ms[a] / tiles[b], nt = mt * mt;
      req = lcm64(req, nt);
      req = lcm64(req, (i64)sg_size * nt);
    }
  const i64 full = lcm64(req, local_size);
  const i64 target = (i64)local_size * target_wgs;
  const i64 mult = std::max<i64>(1, (target + full - 1) / full);
  return mult * full;
}

extern "C" int oracle_validate_stride(const int* Ms, int nM, const int* tiles, int nT,
                                      int64_t stride, int local_size, int sg_size) {
  if (stride % local_size) return 0;
  int ok = 1;
  for (int a = 0; a < nM; ++a)
    for (int b = 0; b < nT; ++b) {
      if (Ms[a] % tiles[b]) continue;
      const i64 mt = Ms[a] / tiles[b], nt = mt * mt;
      if (stride % nt) ok = 0;
      if ((stride / sg_size) % nt) ok = 0;
    }
  return ok;
}

extern "C" int oracle_ref_time(const oracle_ref_cfg* cfg, int64_t n, int k, int m,
                               const double* A, const double* B, double* C, int warmup,
                               int runs, double budget_s, double* times_ns) {
  using clk = std::chrono::steady_clock;
  for (int i = 0; i < warmup; ++i) {
    std::memset(C, 0, sizeof(double) * k * m);
    if (oracle_ref_run(cfg, n, k, m, A, B, C)) return -1;
  }
  double spent = 0.0;
  int done = 0;
  for (int i = 0; i < runs; ++i) {
    std::memset(C, 0, sizeof(double) * k * m);  // zero_C outside the timer (main.cc:119)
    auto t0 = clk::now();
    oracle_ref_run(cfg, n, k, m, A, B, C);
    auto t1 = clk::now();
    times_ns[i] = std::chrono::duration<double, std::nano>(t1 - t0).count();
    spent += times_ns[i] * 1e-9;
    ++done;
    if (spent > budget_s) break;
  }
  return done;
}

extern "C" int oracle_num_threads(void) { return omp_get_max_threads(); }
