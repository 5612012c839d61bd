// Backward of C = A^T B for an upstream gradient dC (k x m):
//   dA = B dC^T  (dA[r, i] = sum_j B[r, j] dC[i, j]),  n x k
//   dB = A dC    (dB[r, j] = sum_i A[r, i] dC[i, j]),  n x m
// Both are tall-skinny times small: one streaming pass reads a row of A and
// B and writes a row of dA and dB.  dC (at most 64 x 64) is staged once per
// workgroup in LDS, zero-padded to the width class W >= max(k, m), and read
// with wave-uniform addresses (LDS broadcast).  Each thread computes whole
// rows.  Three kernels share that arithmetic:
//   staged  (dense rows of 16..64 bytes, widths up to 8): each wave moves
//           its own tiles of 64 rows (128 for 32-byte rows) between HBM and
//           its part of LDS with consecutive 16-byte loads and stores (each
//           instruction covers 1 KiB of contiguous bytes); a lane reads its
//           rows from the padded LDS tile and writes its output rows back
//           there before the tile is stored, and no workgroup barrier is
//           needed;
//   packed  (k == m == W, dense rows of at most 16 bytes; BF16 W = 8 is
//           staged): a thread moves 16 consecutive bytes, 16 / (W s) whole
//           rows, in one load and one store;
//   rowwise (any layout): a thread loads and stores its own rows directly,
//           16 bytes at a time when dense.
// Accumulation is in FP64 for FP64 inputs and FP32 otherwise (each output is
// a sum of at most 64 products, added in the same order by these three
// kernels, so they produce the same bits).  Dense rows of the 16-, 32- and
// 64-wide classes take the matrix-core kernels further down (another
// summation order; for BF16, dC split exactly into three BF16 pieces).
#include <atomic>

#include "tsm/tsm.h"
#include "tsm_internal.hpp"

namespace tsm {

struct BwdArgs {
  const void* A;
  const void* B;
  const void* dC;
  void* dA;
  void* dB;
  int64_t n, lda, ldb, lddc, ldda, lddb;
  int k, m;
  int dc_f64;   // dC is binary64 (1) or binary32 (0)
  int dense_a;  // lda == k, ldda == k, k*s % 16 == 0, 16-byte aligned A and dA
  int dense_b;  // likewise for B and dB
};

template <class S>
__device__ __forceinline__ S narrow(typename Traits<S>::Acc x);
template <>
__device__ __forceinline__ double narrow<double>(double x) { return x; }
template <>
__device__ __forceinline__ float narrow<float>(float x) { return x; }
template <>
__device__ __forceinline__ bf16 narrow<bf16>(float x) {
  // round to nearest even (finite inputs).  v_cvt_pk_bf16_f32 does the same
  // in one instruction per pair but measured 7% slower at BF16 W = 8 (staged)
  // and 1-2% faster at W = 16..64.
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf16{(uint16_t)(u >> 16)};
}

// Load `cnt` (<= W) elements of a row into registers, zero-padded to W.
template <class S, int W>
__device__ __forceinline__ void load_row(const S* row, int cnt, bool dense,
                                         typename Traits<S>::Acc (&out)[W]) {
  using Acc = typename Traits<S>::Acc;
  constexpr int ES = (int)sizeof(S);
  constexpr int PER = 16 / ES;  // elements per 16-byte chunk
  if (dense && W % PER == 0) {
#pragma unroll
    for (int c = 0; c < W / PER; ++c) {
      if (c * PER < cnt) {
        const v4u u = reinterpret_cast<const v4u*>(row)[c];
        Unpack<S>::run(u, out + c * PER);
      } else {
#pragma unroll
        for (int e = 0; e < PER; ++e) out[c * PER + e] = Acc(0);
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < W; ++c) out[c] = c < cnt ? widen(row[c]) : Acc(0);
  }
}

// out[o] = sum_t x[t] * coef(o, t) for o in [0, cnt), written to `row`.
// coef(o, t) = G[o * W + t] (dA: G = dC) or G[t * W + o] (dB: G = dC).
template <class S, int W, bool TRANSPOSED>
__device__ __forceinline__ void emit_row(S* row, int cnt, bool dense,
                                         const typename Traits<S>::Acc (&x)[W],
                                         const typename Traits<S>::Acc* G) {
  using Acc = typename Traits<S>::Acc;
  constexpr int ES = (int)sizeof(S);
  constexpr int PER = 16 / ES;
  constexpr int CH = PER < W ? PER : W;
  for (int o0 = 0; o0 < cnt; o0 += CH) {
    Acc acc[CH];
#pragma unroll
    for (int e = 0; e < CH; ++e) acc[e] = Acc(0);
#pragma unroll
    for (int t = 0; t < W; ++t)
#pragma unroll
      for (int e = 0; e < CH; ++e) {
        const int o = o0 + e;
        const Acc g = TRANSPOSED ? G[t * W + o] : G[o * W + t];
        acc[e] = fmadd(x[t], g, acc[e]);
      }
    if (dense && CH == PER && o0 + CH <= cnt) {
      S tmp[CH];
#pragma unroll
      for (int e = 0; e < CH; ++e) tmp[e] = narrow<S>(acc[e]);
      v4u u;
      __builtin_memcpy(&u, tmp, 16);
      *reinterpret_cast<v4u*>(row + o0) = u;
    } else {
#pragma unroll
      for (int e = 0; e < CH; ++e)
        if (o0 + e < cnt) row[o0 + e] = narrow<S>(acc[e]);
    }
  }
}

template <class S, int W>
__device__ __forceinline__ void bwd_body(const BwdArgs& a) {
  using Acc = typename Traits<S>::Acc;
  // dC zero-padded to W x W in LDS: G[i * W + j] = dC[i, j].
  __shared__ Acc G[W * W + 1];
  for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
    const int i = e / W, j = e % W;
    Acc v = Acc(0);
    if (i < a.k && j < a.m) {
      const int64_t off = (int64_t)i * a.lddc + j;
      v = a.dc_f64 ? (Acc) static_cast<const double*>(a.dC)[off]
                   : (Acc) static_cast<const float*>(a.dC)[off];
    }
    G[e] = v;
  }
  __syncthreads();
  const S* A = static_cast<const S*>(a.A);
  const S* B = static_cast<const S*>(a.B);
  S* dA = static_cast<S*>(a.dA);
  S* dB = static_cast<S*>(a.dB);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += stride) {
    Acc x[W];
    if (dA) {
      load_row<S, W>(B + r * a.ldb, a.m, a.dense_b != 0, x);
      emit_row<S, W, false>(dA + r * a.ldda, a.k, a.dense_a != 0, x, G);
    }
    if (dB) {
      load_row<S, W>(A + r * a.lda, a.k, a.dense_a != 0, x);
      emit_row<S, W, true>(dB + r * a.lddb, a.m, a.dense_b != 0, x, G);
    }
  }
}

// ---- staged kernel ------------------------------------------------------------------
// Each wave stages its own tiles of 64 rows in its own part of LDS, so only
// wave-level ordering is needed: the wave writes the input chunks it loaded,
// each lane reads its row, computes its output row and writes it back over
// its own row (no other lane reads that row), and the wave then stores the
// tile as consecutive 16-byte chunks.  One phase: out (rows x out_w) =
// in (rows x in_w) times the dC coefficients for the wave tile of `rows` rows
// starting at row r0; `regs` holds this lane's 16-byte chunks of the input
// tile (chunk lane + i * 64), already loaded.
template <class S, int W, bool TRANSPOSED, int RPL, int NCH>
__device__ __forceinline__ void staged_phase(const v4u (&regs)[NCH], int rows, int in_w,
                                             int out_w, S* out, int64_t r0, unsigned char* tile,
                                             const typename Traits<S>::Acc* G) {
  using Acc = typename Traits<S>::Acc;
  constexpr int ES = (int)sizeof(S);
  constexpr int PER = 16 / ES;             // elements per chunk
  constexpr int PITCH = W * ES + 16;       // padded LDS row pitch (bytes)
  const int lane = threadIdx.x & 63;
  const int in_cpr = in_w / PER, out_cpr = out_w / PER;
  const int in_chunks = rows * in_cpr;
#pragma unroll
  for (int i = 0; i < NCH; ++i) {
    const int c = lane + i * 64;
    if (c < in_chunks) {
      const int r = c / in_cpr;
      *reinterpret_cast<v4u*>(tile + r * PITCH + (c - r * in_cpr) * 16) = regs[i];
    }
  }
  wave_sync();
  // lane l owns rows l, l + 64, ... of the tile; rows past `rows` are
  // computed from stale tile contents and never stored
  Acc x[RPL][W];
#pragma unroll
  for (int j = 0; j < RPL; ++j)
#pragma unroll
    for (int c = 0; c < W / PER; ++c) {
      if (c < in_cpr) {
        const v4u u = *reinterpret_cast<const v4u*>(tile + (lane + 64 * j) * PITCH + c * 16);
        Unpack<S>::run(u, x[j] + c * PER);
      } else {
#pragma unroll
        for (int e = 0; e < PER; ++e) x[j][c * PER + e] = Acc(0);
      }
    }
#pragma unroll
  for (int oc = 0; oc < W / PER; ++oc) {
    if (oc < out_cpr) {
      Acc acc[RPL][PER];
#pragma unroll
      for (int j = 0; j < RPL; ++j)
#pragma unroll
        for (int e = 0; e < PER; ++e) acc[j][e] = Acc(0);
#pragma unroll
      for (int t = 0; t < W; ++t)
#pragma unroll
        for (int e = 0; e < PER; ++e) {
          const int o = oc * PER + e;
          const Acc g = TRANSPOSED ? G[t * W + o] : G[o * W + t];  // one read, RPL rows
#pragma unroll
          for (int j = 0; j < RPL; ++j) acc[j][e] = fmadd(x[j][t], g, acc[j][e]);
        }
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        S tmp[PER];
#pragma unroll
        for (int e = 0; e < PER; ++e) tmp[e] = narrow<S>(acc[j][e]);
        v4u u;
        __builtin_memcpy(&u, tmp, 16);
        *reinterpret_cast<v4u*>(tile + (lane + 64 * j) * PITCH + oc * 16) = u;
      }
    }
  }
  wave_sync();
  v4u* dst = reinterpret_cast<v4u*>(out + r0 * out_w);
  const int out_chunks = rows * out_cpr;
#pragma unroll
  for (int i = 0; i < NCH; ++i) {
    const int c = lane + i * 64;
    if (c < out_chunks) {
      const int r = c / out_cpr;
      const v4u u = *reinterpret_cast<const v4u*>(tile + r * PITCH + (c - r * out_cpr) * 16);
      __builtin_nontemporal_store(u, dst + c);
    }
  }
  wave_sync();  // the tile is rewritten by the next phase
}

template <class S, int NCH>
__device__ __forceinline__ void load_chunks(const S* in, int64_t r0, int rows, int w,
                                            v4u (&regs)[NCH]) {
  const v4u* src = reinterpret_cast<const v4u*>(in + r0 * w);
  const int chunks = rows * (w * (int)sizeof(S) / 16);
#pragma unroll
  for (int i = 0; i < NCH; ++i) {
    const int c = (threadIdx.x & 63) + i * 64;
    regs[i] = c < chunks ? __builtin_nontemporal_load(src + c) : v4u{0, 0, 0, 0};
  }
}

// Rows per lane of the staged kernel.  A lane that owns two rows reads each
// dC coefficient from LDS once for both; the reads are wave-uniform, and each
// still moves 64 copies through the LDS return path, which 32-byte rows with
// one row per lane load most.  Measured with two rows per lane (TB/s, dA and
// dB): 32-byte rows +5..6% (FP64 W = 4 5.38 -> 5.68, FP32 W = 8 5.36 ->
// 5.66); other widths -1..15% (more VGPRs, fewer waves).
constexpr int staged_rows_per_lane(int row_bytes) { return row_bytes == 32 ? 2 : 1; }

template <class S, int W>
__device__ __forceinline__ void bwd_staged_body(const BwdArgs& a) {
  using Acc = typename Traits<S>::Acc;
  constexpr int ES = (int)sizeof(S);
  static_assert(W * ES >= 16 && W * ES <= 128, "staged rows are 16..128 bytes");
  constexpr int PITCH = W * ES + 16;
  constexpr int RPL = staged_rows_per_lane(W * ES);
  constexpr int WT = 64 * RPL;                // rows per wave tile
  constexpr int NCH = W * ES / 16 * RPL;      // chunks per lane per input tile
  __shared__ Acc G[W * W];
  __shared__ __attribute__((aligned(16))) unsigned char tiles[kThreads / 64][WT * PITCH];
  for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
    const int i = e / W, j = e % W;
    Acc v = Acc(0);
    if (i < a.k && j < a.m) {
      const int64_t off = (int64_t)i * a.lddc + j;
      v = a.dc_f64 ? (Acc) static_cast<const double*>(a.dC)[off]
                   : (Acc) static_cast<const float*>(a.dC)[off];
    }
    G[e] = v;
  }
  __syncthreads();
  const S* A = static_cast<const S*>(a.A);
  const S* B = static_cast<const S*>(a.B);
  S* dA = static_cast<S*>(a.dA);
  S* dB = static_cast<S*>(a.dB);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned char* tile = tiles[wave];
  const int64_t ntiles = (a.n + WT - 1) / WT;
  auto rows_of = [&](int64_t t) {
    const int64_t r0 = t * WT;
    return (int)((a.n - r0) < WT ? (a.n - r0) : WT);
  };
  // Both input tiles of the wave's next tile are requested before this tile
  // is processed.  The tile stride stays an unsigned expression of gridDim.x
  // in the loop: held in a 64-bit variable it cost 18..20 VGPRs in every
  // class (FP64 W = 2: 36 -> 60) and 8% at 16-byte rows.
  v4u rb[NCH], ra[NCH];
  int64_t tt = (int64_t)blockIdx.x * (kThreads / 64) + wave;
  if (tt < ntiles) {
    if (dA) load_chunks<S, NCH>(B, tt * WT, rows_of(tt), a.m, rb);
    if (dB) load_chunks<S, NCH>(A, tt * WT, rows_of(tt), a.k, ra);
  }
  for (; tt < ntiles; tt += gridDim.x * (kThreads / 64)) {
    const int64_t r0 = tt * WT;
    const int rows = rows_of(tt);
    const int64_t nt = tt + gridDim.x * (kThreads / 64);
    v4u nb[NCH], na[NCH];
    if (nt < ntiles) {
      if (dA) load_chunks<S, NCH>(B, nt * WT, rows_of(nt), a.m, nb);
      if (dB) load_chunks<S, NCH>(A, nt * WT, rows_of(nt), a.k, na);
    }
    if (dA) staged_phase<S, W, false, RPL, NCH>(rb, rows, a.m, a.k, dA, r0, tile, G);
    if (dB) staged_phase<S, W, true, RPL, NCH>(ra, rows, a.k, a.m, dB, r0, tile, G);
#pragma unroll
    for (int i = 0; i < NCH; ++i) {
      rb[i] = nb[i];
      ra[i] = na[i];
    }
  }
}

// ---- packed kernel (square classes with rows of at most 16 bytes) --------------------
// k == m == W, W * s <= 16, dense: chunk c (16 bytes) of the input holds rows
// c * PR ... c * PR + PR - 1 (PR = 16 / (W s)), and so does chunk c of the
// output.  One thread per chunk (grid-strided): it issues the non-temporal
// 16-byte loads of both phases first, then forms each chunk's PR rows (the
// same products in the same order as emit_row) and stores them with one
// non-temporal 16-byte store; a wave covers 1 KiB of consecutive bytes per
// instruction.  Loading the dB phase's chunk before the dA phase's products
// measured +7..8.5% at FP64 W = 1 and FP32 W = 1, 2 and +2..3% for BF16
// (5.2-5.4 -> 5.5-5.7 TB/s); two or four chunks per thread and phase were
// slower (4.8-5.4).
template <class S, int W, bool TRANSPOSED>
__device__ __forceinline__ void packed_chunk(const v4u u, S* out, int64_t c,
                                             const typename Traits<S>::Acc* G) {
  using Acc = typename Traits<S>::Acc;
  constexpr int PER = 16 / (int)sizeof(S), PR = PER / W;
  static_assert(PR >= 1 && PR * W == PER, "rows of at most 16 bytes, whole rows per chunk");
  Acc x[PER];
  Unpack<S>::run(u, x);
  S y[PER];
#pragma unroll
  for (int p = 0; p < PR; ++p)
#pragma unroll
    for (int o = 0; o < W; ++o) {
      Acc acc = Acc(0);
#pragma unroll
      for (int t = 0; t < W; ++t)
        acc = fmadd(x[p * W + t], TRANSPOSED ? G[t * W + o] : G[o * W + t], acc);
      y[p * W + o] = narrow<S>(acc);
    }
  v4u v;
  __builtin_memcpy(&v, y, 16);
  __builtin_nontemporal_store(v, reinterpret_cast<v4u*>(out) + c);
}

template <class S, int W>
__device__ __forceinline__ void bwd_packed_body(const BwdArgs& a) {
  using Acc = typename Traits<S>::Acc;
  constexpr int PR = 16 / ((int)sizeof(S) * W);
  __shared__ Acc G[W * W];
  for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
    const int i = e / W, j = e % W;
    const int64_t off = (int64_t)i * a.lddc + j;
    G[e] = a.dc_f64 ? (Acc) static_cast<const double*>(a.dC)[off]
                    : (Acc) static_cast<const float*>(a.dC)[off];
  }
  __syncthreads();
  const S* A = static_cast<const S*>(a.A);
  const S* B = static_cast<const S*>(a.B);
  S* dA = static_cast<S*>(a.dA);
  S* dB = static_cast<S*>(a.dB);
  const int64_t chunks = a.n / PR;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const v4u* Bc = reinterpret_cast<const v4u*>(B);
  const v4u* Ac = reinterpret_cast<const v4u*>(A);
  for (int64_t c = first; c < chunks; c += stride) {
    const v4u ub = dA ? __builtin_nontemporal_load(Bc + c) : v4u{0, 0, 0, 0};
    const v4u ua = dB ? __builtin_nontemporal_load(Ac + c) : v4u{0, 0, 0, 0};
    if (dA) packed_chunk<S, W, false>(ub, dA, c, G);
    if (dB) packed_chunk<S, W, true>(ua, dB, c, G);
  }
  for (int64_t r = chunks * PR + first; r < a.n; r += stride) {  // the n % PR final rows
    Acc x[W];
    if (dA) {
      load_row<S, W>(B + r * W, W, false, x);
      emit_row<S, W, false>(dA + r * W, W, false, x, G);
    }
    if (dB) {
      load_row<S, W>(A + r * W, W, false, x);
      emit_row<S, W, true>(dB + r * W, W, false, x, G);
    }
  }
}

// ---- matrix-core kernel (width classes 32 and 64) ----------------------------------
// Each phase is a small GEMM per tile: out (R x W) = in (R x W) * Gp (W x W),
// with Gp[t][o] = dC[o][t] for dA and dC[t][o] for dB (dC zero-padded to W).
// Wave w owns row blocks w, w + 4, ... of 16 rows; for each it runs W / 4
// K-steps of 16x16x4 matrix instructions into W / 16 accumulator blocks:
//   FP64: v_mfma_f64_16x16x4f64 (FP64 accumulation, as the VALU kernel);
//   FP32: v_mfma_f32_16x16x4f32 (FP32 accumulation, as the VALU kernel).
// (BF16 has its own kernel on BF16 matrix instructions, below.)
// K order: in K-step st, lane group g = l / 16 takes input column
// t = g * W/4 + st, so a lane's input operands for all steps are W/4
// consecutive elements of its tile row (16-byte LDS reads).  Lane l supplies
// in[16 rb + l % 16][t] and Gp[t][16 nb + l % 16]; the Gp operands are
// built once per workgroup in LDS in lane order (one 16-byte read per lane
// covers 16 / sizeof(S) steps, consecutive lanes reading consecutive bytes).
// The result block holds row 4 (l / 16) + v (FP32) or l / 16 + 4 v (FP64) in
// register v, column l % 16.  The tile is staged in LDS with a padded pitch,
// padding columns are zeroed, the outputs overwrite the wave's own input
// rows, and the tile is stored as consecutive 16-byte chunks.  Once a phase
// has staged its input, its registers take the next tile's rows of the same
// input, so those loads are in flight through this phase and the next.
template <class S>
struct MfmaBwd;
template <>
struct MfmaBwd<double> {
  using AccV = v4d;
  static constexpr bool row_major_out = false;
};
template <>
struct MfmaBwd<float> {
  using AccV = v4f;
  static constexpr bool row_major_out = true;
};

template <class S, int W, int R>
constexpr int mfma_bwd_pitch() { return W * (int)sizeof(S) + 16; }

template <class S, int W, int R>
__device__ __forceinline__ void load_rows(const S* in, int64_t r0, int rows, int w,
                                          v4u (&regs)[R * W * (int)sizeof(S) / 16 / kThreads]) {
  constexpr int CPT = R * W * (int)sizeof(S) / 16 / kThreads;
  const v4u* src = reinterpret_cast<const v4u*>(in + r0 * w);
  const int chunks = rows * (w * (int)sizeof(S) / 16);
#pragma unroll
  for (int i = 0; i < CPT; ++i) {
    const int c = threadIdx.x + i * kThreads;
    regs[i] = c < chunks ? __builtin_nontemporal_load(src + c) : v4u{0, 0, 0, 0};
  }
}

// Stage the input tile (each thread writes the chunks it loaded), zero the
// padding columns [in_w, W) of every row, then refill `regs` (next tile).
template <class S, int W, int R, int CPT, class Refill>
__device__ __forceinline__ void stage_rows(v4u (&regs)[CPT], int rows, int in_w,
                                           unsigned char* tile, Refill refill) {
  constexpr int PITCH = mfma_bwd_pitch<S, W, R>();
  constexpr int PER = 16 / (int)sizeof(S);
  constexpr int CPR = W / PER;
  const int tid = threadIdx.x;
  const int in_cpr = in_w / PER;
  const int in_chunks = rows * in_cpr;
#pragma unroll
  for (int i = 0; i < CPT; ++i) {
    const int c = tid + i * kThreads;
    if (c < in_chunks) {
      const int r = c / in_cpr;
      *reinterpret_cast<v4u*>(tile + r * PITCH + (c - r * in_cpr) * 16) = regs[i];
    }
  }
  refill(regs);
  const int pad_cpr = CPR - in_cpr;
  for (int p = tid; p < R * pad_cpr; p += kThreads) {
    const int r = p / pad_cpr;
    *reinterpret_cast<v4u*>(tile + r * PITCH + (in_cpr + p - r * pad_cpr) * 16) =
        v4u{0, 0, 0, 0};
  }
  __syncthreads();
}

// Store the tile's `rows` output rows of out_w elements as 16-byte chunks.
template <class S, int W, int R>
__device__ __forceinline__ void store_rows(const unsigned char* tile, int rows, int out_w,
                                           S* out, int64_t r0) {
  constexpr int PITCH = mfma_bwd_pitch<S, W, R>();
  constexpr int PER = 16 / (int)sizeof(S);
  const int out_cpr = out_w / PER;
  v4u* dst = reinterpret_cast<v4u*>(out + r0 * out_w);
  const int out_chunks = rows * out_cpr;
  for (int c = threadIdx.x; c < out_chunks; c += kThreads) {
    const int r = c / out_cpr;
    const v4u u = *reinterpret_cast<const v4u*>(tile + r * PITCH + (c - r * out_cpr) * 16);
    __builtin_nontemporal_store(u, dst + c);
  }
}

template <class S, int W, int R, bool TRANSPOSED, int CPT, class Refill>
__device__ __forceinline__ void mfma_phase(v4u (&regs)[CPT], int rows, int in_w, int out_w,
                                           S* out, int64_t r0, unsigned char* tile,
                                           const S* gop, Refill refill) {
  using M = MfmaBwd<S>;
  constexpr int PITCH = mfma_bwd_pitch<S, W, R>();
  constexpr int V = 16 / (int)sizeof(S);  // steps per 16-byte operand read
  constexpr int NST = W / 4, NB = W / 16, RBW = R / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
  stage_rows<S, W, R>(regs, rows, in_w, tile, refill);
  S av[RBW][NST];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int c = 0; c < NST / V; ++c) {
      const v4u u = *reinterpret_cast<const v4u*>(
          tile + ((wave + 4 * j) * 16 + (lane & 15)) * PITCH + (g * NST + c * V) * (int)sizeof(S));
      S x[V];
      __builtin_memcpy(x, &u, 16);
#pragma unroll
      for (int v = 0; v < V; ++v) av[j][c * V + v] = x[v];
    }
  typename M::AccV acc[RBW][NB];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[j][nb] = typename M::AccV{};
  const S* po = gop + (TRANSPOSED ? W * W : 0);
#pragma unroll
  for (int c = 0; c < NST / V; ++c) {
    S gv[NB][V];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const v4u u = *reinterpret_cast<const v4u*>(po + ((nb * (NST / V) + c) * 64 + lane) * V);
      __builtin_memcpy(gv[nb], &u, 16);
    }
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int j = 0; j < RBW; ++j)
          acc[j][nb] = mfma_step(av[j][c * V + v], gv[nb][v], acc[j][nb]);
  }
  // outputs overwrite this wave's own rows of the tile
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = M::row_major_out ? 4 * g + v : g + 4 * v;
        reinterpret_cast<S*>(tile + ((wave + 4 * j) * 16 + r) * PITCH)[16 * nb + (lane & 15)] =
            acc[j][nb][v];
      }
  __syncthreads();
  store_rows<S, W, R>(tile, rows, out_w, out, r0);
  __syncthreads();  // the tile is rewritten by the next phase
}

template <class S, int W, int R>
__device__ __forceinline__ void bwd_mfma_body(const BwdArgs& a) {
  constexpr int ES = (int)sizeof(S);
  constexpr int CPT = R * W * ES / 16 / kThreads;  // input chunks per thread (dense rows)
  static_assert(W % 16 == 0 && R % 64 == 0 && CPT >= 1, "16-wide blocks, 4 waves");
  constexpr int PITCH = mfma_bwd_pitch<S, W, R>();
  constexpr int GP = W + 1;
  constexpr int V = 16 / ES, NST = W / 4, NB = W / 16;
  static_assert(R * PITCH >= W * GP * ES, "dC is staged in the tile's memory");
  __shared__ __attribute__((aligned(16))) S gop[2 * W * W];
  __shared__ __attribute__((aligned(16))) unsigned char tile[R * PITCH];
  S* Gs = reinterpret_cast<S*>(tile);
  for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
    const int i = e / W, j = e % W;
    S v = S(0);
    if (i < a.k && j < a.m) {
      const int64_t off = (int64_t)i * a.lddc + j;
      v = a.dc_f64 ? (S) static_cast<const double*>(a.dC)[off]
                   : (S) static_cast<const float*>(a.dC)[off];
    }
    Gs[i * GP + j] = v;
  }
  __syncthreads();
  // operand element (phase, nb, c, lane, v): step st = c V + v, input column
  // t = (lane / 16) NST + st, output column o = 16 nb + lane % 16; phase 0
  // (dA = B dC^T) takes dC[o][t], phase 1 (dB = A dC) dC[t][o]
  for (int idx = threadIdx.x; idx < 2 * W * W; idx += kThreads) {
    const int v = idx % V, ln = (idx / V) % 64, q = idx / (V * 64);
    const int c = q % (NST / V), nb = (q / (NST / V)) % NB, ph = q / ((NST / V) * NB);
    const int t = (ln >> 4) * NST + c * V + v, o = 16 * nb + (ln & 15);
    gop[idx] = ph ? Gs[t * GP + o] : Gs[o * GP + t];
  }
  __syncthreads();
  const S* A = static_cast<const S*>(a.A);
  const S* B = static_cast<const S*>(a.B);
  S* dA = static_cast<S*>(a.dA);
  S* dB = static_cast<S*>(a.dB);
  const int64_t ntiles = (a.n + R - 1) / R;
  auto rows_of = [&](int64_t t) {
    const int64_t r0 = t * R;
    return (int)((a.n - r0) < R ? (a.n - r0) : R);
  };
  v4u rb[CPT], ra[CPT];
  int64_t tt = blockIdx.x;
  if (tt < ntiles) {
    if (dA) load_rows<S, W, R>(B, tt * R, rows_of(tt), a.m, rb);
    if (dB) load_rows<S, W, R>(A, tt * R, rows_of(tt), a.k, ra);
  }
  // the tile stride stays an unsigned expression of gridDim.x
  for (; tt < ntiles; tt += gridDim.x) {
    const int64_t r0 = tt * R;
    const int rows = rows_of(tt);
    const int64_t nt = tt + gridDim.x;
    const bool more = nt < ntiles;
    const int nrows = more ? rows_of(nt) : 0;
    if (dA)
      mfma_phase<S, W, R, false, CPT>(rb, rows, a.m, a.k, dA, r0, tile, gop,
                                      [&](v4u (&r)[CPT]) {
                                        if (more) load_rows<S, W, R>(B, nt * R, nrows, a.m, r);
                                      });
    if (dB)
      mfma_phase<S, W, R, true, CPT>(ra, rows, a.k, a.m, dB, r0, tile, gop,
                                     [&](v4u (&r)[CPT]) {
                                       if (more) load_rows<S, W, R>(A, nt * R, nrows, a.k, r);
                                     });
  }
}

// ---- matrix-core kernel, one output column block per wave (W = 64) ----------------
// For W = 64 the four waves of a workgroup split the output columns instead
// of the rows: wave w computes columns 16 w .. 16 w + 15 of every row of the
// tile, so its dC operands for both phases (W/4 values per lane and phase:
// input column t = (l / 16) W/4 + st, output column 16 w + l % 16) stay in
// registers for the whole kernel and LDS holds only the tile.  Every wave
// reads every input row (two 16-row blocks at a time, one accumulator each,
// interleaved so that no instruction waits on the previous one's result);
// after a barrier the waves overwrite the tile with their output columns.
template <class S, int R, int CPT, class Refill>
__device__ __forceinline__ void mfma_cb_phase(v4u (&regs)[CPT], int rows, int in_w, int out_w,
                                              S* out, int64_t r0, unsigned char* tile,
                                              const S (&gv)[16], Refill refill) {
  using M = MfmaBwd<S>;
  constexpr int W = 64;
  constexpr int PITCH = mfma_bwd_pitch<S, W, R>();
  constexpr int V = 16 / (int)sizeof(S);
  constexpr int NST = W / 4, RB = R / 16;
  static_assert(RB % 2 == 0, "row blocks in pairs");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
  stage_rows<S, W, R>(regs, rows, in_w, tile, refill);
  typename M::AccV acc[RB];
#pragma unroll
  for (int rb = 0; rb < RB; rb += 2) {
    S av[2][NST];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int c = 0; c < NST / V; ++c) {
        const v4u u = *reinterpret_cast<const v4u*>(
            tile + ((rb + j) * 16 + (lane & 15)) * PITCH + (g * NST + c * V) * (int)sizeof(S));
        S x[V];
        __builtin_memcpy(x, &u, 16);
#pragma unroll
        for (int v = 0; v < V; ++v) av[j][c * V + v] = x[v];
      }
    acc[rb] = acc[rb + 1] = typename M::AccV{};
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      acc[rb] = mfma_step(av[0][st], gv[st], acc[rb]);
      acc[rb + 1] = mfma_step(av[1][st], gv[st], acc[rb + 1]);
    }
  }
  __syncthreads();  // every wave has read every input row
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = M::row_major_out ? 4 * g + v : g + 4 * v;
      reinterpret_cast<S*>(tile + (rb * 16 + r) * PITCH)[16 * wave + (lane & 15)] = acc[rb][v];
    }
  __syncthreads();
  store_rows<S, W, R>(tile, rows, out_w, out, r0);
  __syncthreads();  // the tile is rewritten by the next phase
}

template <class S, int R>
__device__ __forceinline__ void bwd_mfma_cb_body(const BwdArgs& a) {
  constexpr int W = 64;
  constexpr int ES = (int)sizeof(S);
  constexpr int CPT = R * W * ES / 16 / kThreads;
  static_assert(R % 32 == 0 && CPT >= 1, "pairs of 16-row blocks");
  constexpr int PITCH = mfma_bwd_pitch<S, W, R>();
  constexpr int GP = W + 1, NST = W / 4;
  static_assert(R * PITCH >= W * GP * ES, "dC is staged in the tile's memory");
  __shared__ __attribute__((aligned(16))) unsigned char tile[R * PITCH];
  S* Gs = reinterpret_cast<S*>(tile);
  for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
    const int i = e / W, j = e % W;
    S v = S(0);
    if (i < a.k && j < a.m) {
      const int64_t off = (int64_t)i * a.lddc + j;
      v = a.dc_f64 ? (S) static_cast<const double*>(a.dC)[off]
                   : (S) static_cast<const float*>(a.dC)[off];
    }
    Gs[i * GP + j] = v;
  }
  __syncthreads();
  // this lane's dC operands: phase 0 (dA = B dC^T) takes dC[o][t], phase 1
  // (dB = A dC) dC[t][o]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int o = 16 * wave + (lane & 15);
  S g0[NST], g1[NST];
#pragma unroll
  for (int st = 0; st < NST; ++st) {
    const int t = (lane >> 4) * NST + st;
    g0[st] = Gs[o * GP + t];
    g1[st] = Gs[t * GP + o];
  }
  __syncthreads();  // the tile memory is reused below
  const S* A = static_cast<const S*>(a.A);
  const S* B = static_cast<const S*>(a.B);
  S* dA = static_cast<S*>(a.dA);
  S* dB = static_cast<S*>(a.dB);
  const int64_t ntiles = (a.n + R - 1) / R;
  auto rows_of = [&](int64_t t) {
    const int64_t r0 = t * R;
    return (int)((a.n - r0) < R ? (a.n - r0) : R);
  };
  v4u rb[CPT], ra[CPT];
  int64_t tt = blockIdx.x;
  if (tt < ntiles) {
    if (dA) load_rows<S, W, R>(B, tt * R, rows_of(tt), a.m, rb);
    if (dB) load_rows<S, W, R>(A, tt * R, rows_of(tt), a.k, ra);
  }
  // the tile stride stays an unsigned expression of gridDim.x
  for (; tt < ntiles; tt += gridDim.x) {
    const int64_t r0 = tt * R;
    const int rows = rows_of(tt);
    const int64_t nt = tt + gridDim.x;
    const bool more = nt < ntiles;
    const int nrows = more ? rows_of(nt) : 0;
    if (dA)
      mfma_cb_phase<S, R, CPT>(rb, rows, a.m, a.k, dA, r0, tile, g0,
                                      [&](v4u (&r)[CPT]) {
                                        if (more) load_rows<S, W, R>(B, nt * R, nrows, a.m, r);
                                      });
    if (dB)
      mfma_cb_phase<S, R, CPT>(ra, rows, a.k, a.m, dB, r0, tile, g1,
                                     [&](v4u (&r)[CPT]) {
                                       if (more) load_rows<S, W, R>(A, nt * R, nrows, a.k, r);
                                     });
  }
}

// ---- BF16 matrix-core backward ----------------------------------------------------
// dC (FP32) is split exactly into three BF16 pieces, g = h + m + l with
// h = bf16(g), m = bf16(g - h), l = bf16(g - h - m) (each difference is exact
// in FP32), so every product of a BF16 input and a piece is exact and the
// sums run in the FP32 accumulation of the matrix instruction: for W = 32, 64
// v_mfma_f32_16x16x32_bf16 (K-steps of K = 32 input columns, E = 8 elements
// per lane), for W = 16 v_mfma_f32_16x16x16_bf16 (K = 16, E = 4).  A 16-row
// block times 16 output columns takes, per K-step, three instructions
// (pieces l, m, h, in that order).  Operands of lane l:
//   input  row l % 16 of the block, columns K ks + E (l / 16) + e, e < E:
//          2 E consecutive bytes of the LDS tile row;
//   piece  output column 16 nb + l % 16, input columns K ks + E (l / 16) + e:
//          built once per workgroup in LDS, in lane order.
// Output of one block: lane l, register v holds row 4 (l / 16) + v, column
// l % 16.  The instruction's FP32 rounding of its 32-product sums is biased
// toward -infinity by about 1e-8 of the sum of |products|, far below the
// BF16 rounding of the outputs.
constexpr int kBf16Pieces = 3;

typedef short v4s __attribute__((ext_vector_type(4)));
template <int K>
struct Bf16Step;
template <>
struct Bf16Step<32> {
  using Op = v8bf;
  static constexpr int E = 8;
  __device__ static v4f mma(Op a, Op b, v4f c) { return mfma_step(a, b, c); }
};
template <>
struct Bf16Step<16> {
  using Op = v4s;
  static constexpr int E = 4;
  __device__ static v4f mma(Op a, Op b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
};
template <int W>
constexpr int bf16_kstep() { return W >= 32 ? 32 : 16; }

template <int W>
constexpr int bf16_bwd_operands() { return 2 * (W / bf16_kstep<W>()) * (W / 16) * kBf16Pieces; }

template <int W, int R, bool TRANSPOSED, int CPT>
__device__ __forceinline__ void mfma_bf16_phase(
    const v4u (&regs)[CPT], int rows, int in_w, int out_w, bf16* out, int64_t r0,
    unsigned char* tile, const typename Bf16Step<bf16_kstep<W>()>::Op* ops) {
  using St = Bf16Step<bf16_kstep<W>()>;
  using Op = typename St::Op;
  constexpr int K = bf16_kstep<W>(), E = St::E;
  constexpr int PITCH = mfma_bwd_pitch<bf16, W, R>();
  constexpr int PER = 8;           // BF16 elements per chunk
  constexpr int CPR = W / PER;     // chunks per padded row
  constexpr int NB = W / 16, KS = W / K, RBW = R / 64;  // row blocks per wave
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int in_cpr = in_w / PER, out_cpr = out_w / PER;
  // stage the input tile and zero the padding columns [in_w, W) of every row
  const int in_chunks = rows * in_cpr;
#pragma unroll
  for (int i = 0; i < CPT; ++i) {
    const int c = tid + i * kThreads;
    if (c < in_chunks) {
      const int r = c / in_cpr;
      *reinterpret_cast<v4u*>(tile + r * PITCH + (c - r * in_cpr) * 16) = regs[i];
    }
  }
  const int pad_cpr = CPR - in_cpr;
  for (int p = tid; p < R * pad_cpr; p += kThreads) {
    const int r = p / pad_cpr;
    *reinterpret_cast<v4u*>(tile + r * PITCH + (in_cpr + p - r * pad_cpr) * 16) =
        v4u{0, 0, 0, 0};
  }
  __syncthreads();
  // wave w owns row blocks w, w + 4, ...; it reads and then overwrites only those rows
  Op av[RBW][KS];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      av[j][ks] = *reinterpret_cast<const Op*>(
          tile + ((wave + 4 * j) * 16 + (lane & 15)) * PITCH + (K * ks + E * (lane >> 4)) * 2);
  v4f acc[RBW][NB];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[j][nb] = v4f{};
  const Op* po = ops + (TRANSPOSED ? KS * NB * kBf16Pieces * 64 : 0);
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int p = kBf16Pieces - 1; p >= 0; --p) {
        const Op bv = po[((ks * NB + nb) * kBf16Pieces + p) * 64 + lane];
#pragma unroll
        for (int j = 0; j < RBW; ++j) acc[j][nb] = St::mma(av[j][ks], bv, acc[j][nb]);
      }
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = (wave + 4 * j) * 16 + 4 * (lane >> 4) + v;
        reinterpret_cast<bf16*>(tile + r * PITCH)[16 * nb + (lane & 15)] =
            narrow<bf16>(acc[j][nb][v]);
      }
  __syncthreads();
  v4u* dst = reinterpret_cast<v4u*>(out + r0 * out_w);
  const int out_chunks = rows * out_cpr;
  for (int c = tid; c < out_chunks; c += kThreads) {
    const int r = c / out_cpr;
    const v4u u = *reinterpret_cast<const v4u*>(tile + r * PITCH + (c - r * out_cpr) * 16);
    __builtin_nontemporal_store(u, dst + c);
  }
  __syncthreads();  // the tile is rewritten by the next phase
}

template <int W, int R>
__device__ __forceinline__ void bwd_mfma_bf16_body(const BwdArgs& a) {
  using St = Bf16Step<bf16_kstep<W>()>;
  using Op = typename St::Op;
  constexpr int K = bf16_kstep<W>(), E = St::E;
  constexpr int CPT = R * W * 2 / 16 / kThreads;  // input chunks per thread (dense rows)
  static_assert(W % K == 0 && R % 64 == 0 && CPT >= 1, "whole K-steps, 4 waves");
  constexpr int PITCH = mfma_bwd_pitch<bf16, W, R>();
  constexpr int GP = W + 1;
  constexpr int NB = W / 16, KS = W / K;
  static_assert(R * PITCH >= W * GP * 4, "dC is staged in the tile's memory");
  __shared__ Op ops[bf16_bwd_operands<W>() * 64];
  __shared__ __attribute__((aligned(16))) unsigned char tile[R * PITCH];
  float* Gs = reinterpret_cast<float*>(tile);
  for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
    const int i = e / W, j = e % W;
    float v = 0.0f;
    if (i < a.k && j < a.m) {
      const int64_t off = (int64_t)i * a.lddc + j;
      v = a.dc_f64 ? (float)static_cast<const double*>(a.dC)[off]
                   : static_cast<const float*>(a.dC)[off];
    }
    Gs[i * GP + j] = v;
  }
  __syncthreads();
  // pieces: phase 0 (dA = B dC^T) takes dC[o][t], phase 1 (dB = A dC) dC[t][o]
  for (int idx = threadIdx.x; idx < 2 * KS * NB * 64; idx += kThreads) {
    const int ln = idx & 63, q = idx >> 6;  // q = (phase * KS + ks) * NB + nb
    const int nb = q % NB, ks = (q / NB) % KS, ph = q / (NB * KS);
    uint32_t w[kBf16Pieces][E / 2];
#pragma unroll
    for (int e = 0; e < E; e += 2) {
      uint32_t h[kBf16Pieces][2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int t = K * ks + E * (ln >> 4) + e + x, o = 16 * nb + (ln & 15);
        float g = ph ? Gs[t * GP + o] : Gs[o * GP + t];
#pragma unroll
        for (int p = 0; p < kBf16Pieces; ++p) {
          const bf16 b = narrow<bf16>(g);
          h[p][x] = b.bits;
          g -= widen(b);
        }
      }
#pragma unroll
      for (int p = 0; p < kBf16Pieces; ++p) w[p][e / 2] = h[p][0] | (h[p][1] << 16);
    }
#pragma unroll
    for (int p = 0; p < kBf16Pieces; ++p) {
      Op v;
      __builtin_memcpy(&v, w[p], sizeof(Op));
      ops[(q * kBf16Pieces + p) * 64 + ln] = v;
    }
  }
  __syncthreads();
  const bf16* A = static_cast<const bf16*>(a.A);
  const bf16* B = static_cast<const bf16*>(a.B);
  bf16* dA = static_cast<bf16*>(a.dA);
  bf16* dB = static_cast<bf16*>(a.dB);
  const int64_t ntiles = (a.n + R - 1) / R;
  auto rows_of = [&](int64_t t) {
    const int64_t r0 = t * R;
    return (int)((a.n - r0) < R ? (a.n - r0) : R);
  };
  // both input tiles of the next tile are requested before this tile is
  // processed (the tile stride stays an unsigned expression of gridDim.x)
  v4u rb[CPT], ra[CPT];
  int64_t tt = blockIdx.x;
  if (tt < ntiles) {
    if (dA) load_rows<bf16, W, R>(B, tt * R, rows_of(tt), a.m, rb);
    if (dB) load_rows<bf16, W, R>(A, tt * R, rows_of(tt), a.k, ra);
  }
  for (; tt < ntiles; tt += gridDim.x) {
    const int64_t r0 = tt * R;
    const int rows = rows_of(tt);
    const int64_t nt = tt + gridDim.x;
    v4u nb_[CPT], na_[CPT];
    if (nt < ntiles) {
      if (dA) load_rows<bf16, W, R>(B, nt * R, rows_of(nt), a.m, nb_);
      if (dB) load_rows<bf16, W, R>(A, nt * R, rows_of(nt), a.k, na_);
    }
    if (dA) mfma_bf16_phase<W, R, false, CPT>(rb, rows, a.m, a.k, dA, r0, tile, ops);
    if (dB) mfma_bf16_phase<W, R, true, CPT>(ra, rows, a.k, a.m, dB, r0, tile, ops);
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      rb[i] = nb_[i];
      ra[i] = na_[i];
    }
  }
}

}  // namespace tsm

// Rows per tile: 64 for FP64 W = 32 (measured 5.61 TB/s against 5.36 with
// 128), 128 otherwise (FP32 W = 32: 5.59 against 5.33 with 64 and 5.35 with
// 256; FP32 W = 64: 5.23 against 4.57 with 256 and 5.08-5.20 with 64).  The
// 16-wide classes run here too: FP64 W = 16 with 128 rows 5.71-5.75 TB/s (64
// rows: 5.40-5.42, 256: 5.48-5.50; the staged VALU kernel 5.17-5.26), FP32
// W = 16 with 256 rows 5.73-5.77 (128: 5.39-5.40, 512: 5.45; staged
// 5.46-5.50).  Routing the 8-wide classes here (zero-padded to 16) measured
// 6-7% slower than staged.
#define TSM_BWD_MFMA(NAME, S, W, R)                                                   \
  extern "C" __global__ void __launch_bounds__(256) NAME(tsm::BwdArgs a) {          \
    tsm::bwd_mfma_body<S, W, R>(a);                                                   \
  }

TSM_BWD_MFMA(tsm_atb_bwd_mfma_f64_w16, double, 16, 128)
TSM_BWD_MFMA(tsm_atb_bwd_mfma_f32_w16, float, 16, 256)
TSM_BWD_MFMA(tsm_atb_bwd_mfma_f64_w32, double, 32, 64)
TSM_BWD_MFMA(tsm_atb_bwd_mfma_f32_w32, float, 32, 128)
TSM_BWD_MFMA(tsm_atb_bwd_mfma_f32_w64, float, 64, 128)

// FP64 W = 64: one output column block per wave, 64-row tiles (252 registers
// and 33 KiB of LDS: two workgroups per CU).  Measured 5.17-5.26 TB/s
// against 4.60-4.62 for the row-block kernel (130 KiB of LDS, one workgroup
// per CU); for FP32 W = 64 the two measured the same.
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bwd_mfma_f64_w64(tsm::BwdArgs a) {
  tsm::bwd_mfma_cb_body<double, 64>(a);
}
// Rows per tile: 512 at W = 16 (5.79-5.84 TB/s against 5.32 with 128, 5.23
// with 256, 5.21-5.31 with 1024, and 4.85-4.91 for the staged VALU kernel;
// routing W = 8 here measured 4.75-4.80 against 5.65 staged), 256 at W = 32
// (5.70 TB/s against 5.26 with 128), 128 at W = 64 (5.32 against 3.83 with
// 256).
#define TSM_BWD_MFMA_BF16(NAME, W, R)                                                 \
  extern "C" __global__ void __launch_bounds__(256) NAME(tsm::BwdArgs a) {          \
    tsm::bwd_mfma_bf16_body<W, R>(a);                                                 \
  }

TSM_BWD_MFMA_BF16(tsm_atb_bwd_mfma_bf16_w16, 16, 512)
TSM_BWD_MFMA_BF16(tsm_atb_bwd_mfma_bf16_w32, 32, 256)
TSM_BWD_MFMA_BF16(tsm_atb_bwd_mfma_bf16_w64, 64, 128)

#define TSM_BWD_STAGED(NAME, S, W)                                                    \
  extern "C" __global__ void __launch_bounds__(256) NAME(tsm::BwdArgs a) {          \
    tsm::bwd_staged_body<S, W>(a);                                                    \
  }

// Dense rows of 16..64 bytes in the classes up to 8 wide (16 and wider run on
// the matrix-core kernels).
TSM_BWD_STAGED(tsm_atb_bwd_staged_f64_w2, double, 2)
TSM_BWD_STAGED(tsm_atb_bwd_staged_f64_w4, double, 4)
TSM_BWD_STAGED(tsm_atb_bwd_staged_f64_w8, double, 8)
TSM_BWD_STAGED(tsm_atb_bwd_staged_f32_w4, float, 4)
TSM_BWD_STAGED(tsm_atb_bwd_staged_f32_w8, float, 8)
TSM_BWD_STAGED(tsm_atb_bwd_staged_bf16_w8, tsm::bf16, 8)

#define TSM_BWD_PACKED(NAME, S, W)                                                    \
  extern "C" __global__ void __launch_bounds__(256) NAME(tsm::BwdArgs a) {          \
    tsm::bwd_packed_body<S, W>(a);                                                    \
  }

// One 16-byte row per chunk for FP64 W = 2 and FP32 W = 4: 5.61-5.64 TB/s
// against 5.46-5.48 for the staged kernel; BF16 W = 8 stays staged (5.66
// against 5.17 packed).
TSM_BWD_PACKED(tsm_atb_bwd_packed_f64_w1, double, 1)
TSM_BWD_PACKED(tsm_atb_bwd_packed_f64_w2, double, 2)
TSM_BWD_PACKED(tsm_atb_bwd_packed_f32_w1, float, 1)
TSM_BWD_PACKED(tsm_atb_bwd_packed_f32_w2, float, 2)
TSM_BWD_PACKED(tsm_atb_bwd_packed_f32_w4, float, 4)
TSM_BWD_PACKED(tsm_atb_bwd_packed_bf16_w1, tsm::bf16, 1)
TSM_BWD_PACKED(tsm_atb_bwd_packed_bf16_w2, tsm::bf16, 2)
TSM_BWD_PACKED(tsm_atb_bwd_packed_bf16_w4, tsm::bf16, 4)

#define TSM_BWD_KERNEL(NAME, S, W)                                                    \
  extern "C" __global__ void __launch_bounds__(256) NAME(tsm::BwdArgs a) {          \
    tsm::bwd_body<S, W>(a);                                                           \
  }

TSM_BWD_KERNEL(tsm_atb_bwd_f64_w1, double, 1)
TSM_BWD_KERNEL(tsm_atb_bwd_f64_w2, double, 2)
TSM_BWD_KERNEL(tsm_atb_bwd_f64_w4, double, 4)
TSM_BWD_KERNEL(tsm_atb_bwd_f64_w8, double, 8)
TSM_BWD_KERNEL(tsm_atb_bwd_f64_w16, double, 16)
TSM_BWD_KERNEL(tsm_atb_bwd_f64_w32, double, 32)
TSM_BWD_KERNEL(tsm_atb_bwd_f64_w64, double, 64)
TSM_BWD_KERNEL(tsm_atb_bwd_f32_w1, float, 1)
TSM_BWD_KERNEL(tsm_atb_bwd_f32_w2, float, 2)
TSM_BWD_KERNEL(tsm_atb_bwd_f32_w4, float, 4)
TSM_BWD_KERNEL(tsm_atb_bwd_f32_w8, float, 8)
TSM_BWD_KERNEL(tsm_atb_bwd_f32_w16, float, 16)
TSM_BWD_KERNEL(tsm_atb_bwd_f32_w32, float, 32)
TSM_BWD_KERNEL(tsm_atb_bwd_f32_w64, float, 64)
TSM_BWD_KERNEL(tsm_atb_bwd_bf16_w1, tsm::bf16, 1)
TSM_BWD_KERNEL(tsm_atb_bwd_bf16_w2, tsm::bf16, 2)
TSM_BWD_KERNEL(tsm_atb_bwd_bf16_w4, tsm::bf16, 4)
TSM_BWD_KERNEL(tsm_atb_bwd_bf16_w8, tsm::bf16, 8)
TSM_BWD_KERNEL(tsm_atb_bwd_bf16_w16, tsm::bf16, 16)
TSM_BWD_KERNEL(tsm_atb_bwd_bf16_w32, tsm::bf16, 32)
TSM_BWD_KERNEL(tsm_atb_bwd_bf16_w64, tsm::bf16, 64)

namespace {
typedef void (*BwdKernel)(tsm::BwdArgs);
const BwdKernel kBwd[3][7] = {
    {tsm_atb_bwd_f64_w1, tsm_atb_bwd_f64_w2, tsm_atb_bwd_f64_w4, tsm_atb_bwd_f64_w8,
     tsm_atb_bwd_f64_w16, tsm_atb_bwd_f64_w32, tsm_atb_bwd_f64_w64},
    {tsm_atb_bwd_f32_w1, tsm_atb_bwd_f32_w2, tsm_atb_bwd_f32_w4, tsm_atb_bwd_f32_w8,
     tsm_atb_bwd_f32_w16, tsm_atb_bwd_f32_w32, tsm_atb_bwd_f32_w64},
    {tsm_atb_bwd_bf16_w1, tsm_atb_bwd_bf16_w2, tsm_atb_bwd_bf16_w4, tsm_atb_bwd_bf16_w8,
     tsm_atb_bwd_bf16_w16, tsm_atb_bwd_bf16_w32, tsm_atb_bwd_bf16_w64},
};
// Staged kernels by dtype and log2(W); null where W * s is outside 16..128 bytes.
const BwdKernel kBwdStaged[3][7] = {
    {nullptr, tsm_atb_bwd_staged_f64_w2, tsm_atb_bwd_staged_f64_w4, tsm_atb_bwd_staged_f64_w8,
     nullptr, nullptr, nullptr},
    {nullptr, nullptr, tsm_atb_bwd_staged_f32_w4, tsm_atb_bwd_staged_f32_w8, nullptr, nullptr,
     nullptr},
    {nullptr, nullptr, nullptr, tsm_atb_bwd_staged_bf16_w8, nullptr, nullptr, nullptr},
};
// Packed kernels by dtype and log2(W); null where W * s is over 16 bytes, and
// for BF16 W = 8 (staged).
const BwdKernel kBwdPacked[3][3] = {
    {tsm_atb_bwd_packed_f64_w1, tsm_atb_bwd_packed_f64_w2, nullptr},
    {tsm_atb_bwd_packed_f32_w1, tsm_atb_bwd_packed_f32_w2, tsm_atb_bwd_packed_f32_w4},
    {tsm_atb_bwd_packed_bf16_w1, tsm_atb_bwd_packed_bf16_w2, tsm_atb_bwd_packed_bf16_w4},
};
// Matrix-core kernels for the 32- and 64-wide classes (dense rows).
const BwdKernel kBwdMfma[3][2] = {
    {tsm_atb_bwd_mfma_f64_w32, tsm_atb_bwd_mfma_f64_w64},
    {tsm_atb_bwd_mfma_f32_w32, tsm_atb_bwd_mfma_f32_w64},
    {tsm_atb_bwd_mfma_bf16_w32, tsm_atb_bwd_mfma_bf16_w64},
};
int elem(int t) { return t == TSM_F64 ? 8 : t == TSM_F32 ? 4 : 2; }
}  // namespace

extern "C" tsm_status tsm_atb_backward(tsm_dtype in_type, tsm_dtype grad_type, int64_t n, int k,
                                       int m, const void* A, int64_t lda, const void* B,
                                       int64_t ldb, const void* dC, int64_t lddc, void* dA,
                                       int64_t ldda, void* dB, int64_t lddb,
                                       tsm_stream_t stream) {
  if (in_type != TSM_F64 && in_type != TSM_F32 && in_type != TSM_BF16) return TSM_ERR_UNSUPPORTED;
  if (grad_type != TSM_F64 && grad_type != TSM_F32) return TSM_ERR_UNSUPPORTED;
  if (n < 0 || k < 1 || m < 1 || k > TSM_MAX_WIDTH || m > TSM_MAX_WIDTH)
    return TSM_ERR_INVALID_ARGUMENT;
  if (lda < k || ldb < m || lddc < m) return TSM_ERR_INVALID_ARGUMENT;
  if (!dC) return TSM_ERR_INVALID_ARGUMENT;
  if (dA && ldda < k) return TSM_ERR_INVALID_ARGUMENT;
  if (dB && lddb < m) return TSM_ERR_INVALID_ARGUMENT;
  if (n > 0 && ((dA && !B) || (dB && !A))) return TSM_ERR_INVALID_ARGUMENT;
  {  // row offsets (n - 1) * ld stay far from int64 overflow, as in tsm_atb
    int64_t ld = lda > ldb ? lda : ldb;
    if (dA && ldda > ld) ld = ldda;
    if (dB && lddb > ld) ld = lddb;
    if (n > ((int64_t)1 << 62) / ld) return TSM_ERR_INVALID_ARGUMENT;
  }
  const uintptr_t es = (uintptr_t)elem(in_type);
  if ((uintptr_t)A % es || (uintptr_t)B % es || (uintptr_t)dA % es || (uintptr_t)dB % es)
    return TSM_ERR_MISALIGNED;
  if ((uintptr_t)dC % (uintptr_t)elem(grad_type)) return TSM_ERR_MISALIGNED;
  if (n == 0 || (!dA && !dB)) return TSM_SUCCESS;
  int w = 1, wi = 0;
  while (w < (k > m ? k : m)) {
    w <<= 1;
    ++wi;
  }
  tsm::BwdArgs a{};
  a.A = A;
  a.B = B;
  a.dC = dC;
  a.dA = dA;
  a.dB = dB;
  a.n = n;
  a.lda = lda;
  a.ldb = ldb;
  a.lddc = lddc;
  a.ldda = ldda;
  a.lddb = lddb;
  a.k = k;
  a.m = m;
  a.dc_f64 = grad_type == TSM_F64 ? 1 : 0;
  const auto al16 = [](const void* p) { return ((uintptr_t)p % 16) == 0; };
  a.dense_a = (lda == k && (!dA || ldda == k) && (k * (int)es) % 16 == 0 && al16(A) &&
               (!dA || al16(dA)))
                  ? 1
                  : 0;
  a.dense_b = (ldb == m && (!dB || lddb == m) && (m * (int)es) % 16 == 0 && al16(B) &&
               (!dB || al16(dB)))
                  ? 1
                  : 0;
  int cus = 0;
  const tsm_status cst = tsm_internal_device_cus(&cus);
  if (cst != TSM_SUCCESS) return cst;
  int64_t blocks = (n + 255) / 256;
  // Staged when every row that moves is dense: both dense flags cover the
  // inputs read and the outputs written (rows of k * s and m * s bytes, each
  // a multiple of 16), and the padded width fits the staged tile.
  // The staged kernel runs a persistent grid of as many workgroups as fit on
  // the device at once (its LDS tile bounds the count per CU).
  BwdKernel kern = kBwd[(int)in_type][wi];
  int64_t cap = (int64_t)cus * 8;
  const bool dense = a.dense_a && a.dense_b;
  // Square classes whose rows are at most 16 bytes (a packed kernel exists),
  // every moving matrix contiguous (leading dimension = width) and 16-byte
  // aligned.
  const bool packed = k == m && k == w && k * (int)es <= 16 && wi < 3 &&
                      kBwdPacked[(int)in_type][wi] && (!dA || (ldb == m && ldda == k && al16(B) &&
                                                               al16(dA))) &&
                      (!dB || (lda == k && lddb == m && al16(A) && al16(dB)));
  if (packed) {
    kern = kBwdPacked[(int)in_type][wi];
    const int64_t chunks = n / (16 / (k * (int)es));
    blocks = (chunks > 0 ? chunks : 1) + 255;
    blocks /= 256;
  }
  BwdKernel fast = nullptr;
  int64_t tile_rows = 256;
  if (dense && wi == 4) {
    fast = in_type == TSM_F64   ? tsm_atb_bwd_mfma_f64_w16
           : in_type == TSM_F32 ? tsm_atb_bwd_mfma_f32_w16
                                : tsm_atb_bwd_mfma_bf16_w16;
    tile_rows = 128;
  } else if (dense && wi >= 5) {
    fast = kBwdMfma[(int)in_type][wi - 5];
    tile_rows = 128;
  } else if (dense && !packed) {
    fast = kBwdStaged[(int)in_type][wi];
  }
  if (fast) {
    kern = fast;
    blocks = (n + tile_rows - 1) / tile_rows;
    // workgroups per CU, filled on first use (concurrent first calls store
    // the same value)
    static std::atomic<int> resident[3][7];
    int per_cu = resident[(int)in_type][wi].load(std::memory_order_relaxed);
    if (per_cu == 0) {
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kern),
                                                       256, 0) != hipSuccess || nb < 1)
        nb = 1;
      // The matrix-core kernels (W >= 16) run best at two FP64/FP32
      // workgroups per CU (three for BF16), below what fits: measured at 1 GiB
      // of inputs, FP32 W = 16 both gradients 372.1 -> 359.5 us and dA only
      // 179.9 -> 177.8 us, FP64 and FP32 W = 32 383.1 / 381.6 -> 369.7 /
      // 364.7 us, against four per CU.
      const int most = in_type == TSM_BF16 ? 3 : 2;
      if (wi >= 4 && nb > most) nb = most;
      per_cu = nb;
      resident[(int)in_type][wi].store(per_cu, std::memory_order_relaxed);
    }
    cap = (int64_t)cus * per_cu;
  }
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), a);
  return (tsm_status)tsm_internal_hip_status(hipGetLastError());
}

extern "C" tsm_status tsm_tsmm(tsm_dtype in_type, tsm_dtype x_type, int64_t n, int k, int m,
                               const void* A, int64_t lda, const void* X, int64_t ldx, void* Y,
                               int64_t ldy, tsm_stream_t stream) {
  if (!Y) return n == 0 ? TSM_SUCCESS : TSM_ERR_INVALID_ARGUMENT;
  // Y = A X is dB = A dC with dC = X; B is not read (ldb = m passes the checks).
  return tsm_atb_backward(in_type, x_type, n, k, m, A, lda, nullptr, m, X, ldx, nullptr, k, Y,
                          ldy, stream);
}
