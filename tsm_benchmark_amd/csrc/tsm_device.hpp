// Device building blocks of the C = A^T B kernels for gfx950 (CDNA4, wave64).
//
// One workgroup (256 threads = 4 waves) streams a contiguous chunk of rows of
// A (n x k) and B (n x m).  Rows arrive in tiles of T rows: each thread issues
// 16-byte global loads whose addresses are consecutive across the workgroup
// (a tile of a dense row-major matrix is one contiguous byte range), the
// loaded registers are written to one of two LDS stages, and the tile is
// consumed from LDS by a register-tiled outer product.  While a tile is being
// consumed the loads of the next PF tiles are in flight.
//
// Thread -> work mapping inside a tile: the k x m result (padded to the width
// class KC x MC) is cut into NTILES = (KC/TM) * (MC/TN) register tiles; thread
// t owns tile p = t % NTILES (A columns [mb*TM, mb*TM+TM), B columns
// [nb*TN, nb*TN+TN)) and the rows q, q + RG, q + 2RG, ... (q = t / NTILES,
// RG = 256 / NTILES) of every tile.
//
// Reduction (deterministic, no floating-point atomics):
//   1. wave64 recursive halving over the lanes that share a register tile
//      (each level exchanges half of the remaining values with __shfl_xor);
//   2. the four waves are added into LDS in wave order;
//   3. each workgroup writes its KC x MC FP64 partial to the workspace;
//   4. the last workgroup to arrive (an agent-scope arrival counter that the
//      last workgroup resets to zero) adds the partials in workgroup order and
//      writes C.  The summation tree depends only on the launch shape, so a
//      given shape and config gives the same bits on every run.
// FP32 and BF16 inputs accumulate in FP32 over short row blocks; each block
// sum is folded into a compensated (hi, lo) FP32 pair with TwoSum; steps 1-4
// run in FP64 for every input type.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace tsm {

constexpr int kThreads = 256;  // 4 waves of 64

// 16-byte chunk as a native vector (an SSA value, so register arrays of it
// stay in VGPRs; HIP's uint4 class is copied through memory).
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

struct bf16 {
  uint16_t bits;
};

// ---- storage / accumulator traits ------------------------------------------
template <class S>
struct Traits;
template <>
struct Traits<double> {
  using Acc = double;
  static constexpr bool compensated = false;
};
template <>
struct Traits<float> {
  using Acc = float;
  static constexpr bool compensated = true;
};
template <>
struct Traits<bf16> {
  using Acc = float;
  static constexpr bool compensated = true;
};

__device__ __forceinline__ double widen(double x) { return x; }
__device__ __forceinline__ float widen(float x) { return x; }
__device__ __forceinline__ float widen(bf16 x) { return __uint_as_float((uint32_t)x.bits << 16); }

__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Unpack one 16-byte LDS chunk into accumulator-typed values.
template <class S>
struct Unpack;
template <>
struct Unpack<double> {
  static constexpr int N = 2;
  __device__ static void run(const v4u& u, double* o) {
    o[0] = __hiloint2double((int)u.y, (int)u.x);
    o[1] = __hiloint2double((int)u.w, (int)u.z);
  }
};
template <>
struct Unpack<float> {
  static constexpr int N = 4;
  __device__ static void run(const v4u& u, float* o) {
    o[0] = __uint_as_float(u.x);
    o[1] = __uint_as_float(u.y);
    o[2] = __uint_as_float(u.z);
    o[3] = __uint_as_float(u.w);
  }
};
template <>
struct Unpack<bf16> {
  static constexpr int N = 8;
  __device__ static void run(const v4u& u, float* o) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      o[2 * i] = __uint_as_float(w[i] << 16);
      o[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
    }
  }
};

// Read CNT consecutive elements of a tile row from LDS, widened.
template <class S, int CNT>
__device__ __forceinline__ void lds_read(const S* p, typename Traits<S>::Acc* out) {
  constexpr int BYTES = CNT * (int)sizeof(S);
  if constexpr (BYTES % 16 == 0) {
    const v4u* v = reinterpret_cast<const v4u*>(p);
#pragma unroll
    for (int c = 0; c < BYTES / 16; ++c) {
      const v4u u = v[c];
      Unpack<S>::run(u, out + c * Unpack<S>::N);
    }
  } else {
#pragma unroll
    for (int c = 0; c < CNT; ++c) out[c] = widen(p[c]);
  }
}

// 16-byte global load of streamed data; NT = 1 marks it non-temporal (the
// line is not kept in the caches for reuse: A and B are read exactly once).
template <int NT>
__device__ __forceinline__ v4u stream_load(const v4u* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  return *p;
}
// Inputs live in global memory.  The batched kernel reads A and B from its
// problem table, where the compiler cannot prove that, and would emit flat
// loads (which also count against lgkmcnt and serialise the pipelined LDS
// waits); the dense tile loads take this global pointer type.
typedef __attribute__((address_space(1))) const v4u gv4u;
template <int NT>
__device__ __forceinline__ v4u stream_load(gv4u* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  return *p;
}

// TwoSum (Knuth): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float z = s - a;
  e = (a - (s - z)) + (b - z);
}

// ---- kernel arguments ---------------------------------------------------------
struct AtbArgs {
  const void* A;
  const void* B;
  void* C;
  double* partials;       // [gridDim.x][KC*MC] FP64 workgroup partials
  unsigned int* counter;  // arrival counter, zero between calls
  int64_t n, lda, ldb, ldc;
  int64_t tiles_per_wg;   // contiguous mode: consecutive tiles per workgroup
  int interleave;         // 1: tile t goes to workgroup t % gridDim.x
  int k, m;
  int accumulate;         // C += result (1) or C = result (0)
  int c_f32;              // C is binary32 (1) or binary64 (0)
  int dense;              // k == KC, m == MC, lda == k, ldb == m, 16-byte aligned A, B
  int finish;             // the last workgroup writes C (1) or a finisher kernel does (0)
};

// ---- cross-lane exchange ------------------------------------------------------------
// Value of lane (lane ^ D) of a wave64 without LDS: DPP within rows of 16
// lanes for D = 1, 2, 4, 8 and v_permlane{16,32}_swap for D = 16, 32 (lane
// mappings measured on gfx950).  `up` is (lane & D) != 0.
template <int D>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x, bool up) {
  if constexpr (D == 1) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  } else if constexpr (D == 2) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else if constexpr (D == 4) {
    const uint32_t from_up = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x12C, 0xF, 0xF, false);  // row_ror:12: lane i + 4
    const uint32_t from_dn = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x124, 0xF, 0xF, false);  // row_ror:4: lane i - 4
    return up ? from_dn : from_up;
  } else if constexpr (D == 8) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (D == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return up ? r[0] : r[1];
  } else {
    static_assert(D == 32, "xor distance within a wave64");
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return up ? r[0] : r[1];
  }
}
template <int D>
__device__ __forceinline__ double lane_xor(double v, bool up) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint64_t lo = lane_xor<D>((uint32_t)u, up), hi = lane_xor<D>((uint32_t)(u >> 32), up);
  return __longlong_as_double((long long)((hi << 32) | lo));
}
// For D = 16, 32: keep + (partner's copy of the kept half) for every lane, from
// the lower-half value lo and the upper-half value hi of this lane: one swap
// moves the partner's copy into place for both lane groups.
template <int D>
__device__ __forceinline__ double swap_add(double lo, double hi) {
  const uint64_t a = (uint64_t)__double_as_longlong(lo), b = (uint64_t)__double_as_longlong(hi);
  uint32_t d0, d1, s0, s1;
  if constexpr (D == 16) {
    const auto r0 = __builtin_amdgcn_permlane16_swap((uint32_t)a, (uint32_t)b, false, false);
    const auto r1 = __builtin_amdgcn_permlane16_swap((uint32_t)(a >> 32), (uint32_t)(b >> 32), false, false);
    d0 = r0[0]; s0 = r0[1]; d1 = r1[0]; s1 = r1[1];
  } else {
    static_assert(D == 32, "swap distance 16 or 32");
    const auto r0 = __builtin_amdgcn_permlane32_swap((uint32_t)a, (uint32_t)b, false, false);
    const auto r1 = __builtin_amdgcn_permlane32_swap((uint32_t)(a >> 32), (uint32_t)(b >> 32), false, false);
    d0 = r0[0]; s0 = r0[1]; d1 = r1[0]; s1 = r1[1];
  }
  // lanes without bit D: dst = own lo (kept), src = partner's lo;
  // lanes with bit D: dst = partner's hi, src = own hi (kept).
  const double x = __longlong_as_double((long long)(((uint64_t)d1 << 32) | d0));
  const double y = __longlong_as_double((long long)(((uint64_t)s1 << 32) | s0));
  return x + y;
}

// ---- wave64 recursive-halving reduction -------------------------------------------
// Lanes L and L ^ d (d = NTILES << l) share a register tile.  While more than
// one value is left, each level keeps half of the values (the upper half on
// the lane whose bit d is set) and adds the partner's copy of that half; the
// remaining levels are a plain butterfly.  All exchanges stay in registers.
template <int NTILES, int V>
struct WaveReduce {
  static constexpr int log2c(int x) { return x <= 1 ? 0 : 1 + log2c(x / 2); }
  static constexpr int QB = log2c(64 / NTILES);  // lane bits that index rows
  static constexpr int LV = log2c(V);
  static constexpr int H = QB < LV ? QB : LV;    // halving levels
  static constexpr int VL = V >> H;              // values each lane keeps

  template <int L>
  __device__ static __forceinline__ void halve(double (&v)[V], int lane) {
    if constexpr (L < H) {
      constexpr int d = NTILES << L;
      constexpr int half = (V >> L) >> 1;
      const bool up = (lane & d) != 0;
#pragma unroll
      for (int x = 0; x < half; ++x) {
        if constexpr (d >= 16) {
          v[x] = swap_add<d>(v[x], v[x + half]);
        } else {
          const double send = up ? v[x] : v[x + half];
          const double keep = up ? v[x + half] : v[x];
          v[x] = keep + lane_xor<d>(send, up);
        }
      }
      halve<L + 1>(v, lane);
    }
  }
  template <int L>
  __device__ static __forceinline__ void butterfly(double (&v)[V], int lane) {
    if constexpr (L < QB) {
      constexpr int d = NTILES << L;
      v[0] += lane_xor<d>(v[0], (lane & d) != 0);
      butterfly<L + 1>(v, lane);
    }
  }
  __device__ static __forceinline__ void run(double (&v)[V], int lane) {
    halve<0>(v, lane);
    butterfly<H>(v, lane);
  }
  // Index (in [0, V)) of the first value this lane holds after run().
  __device__ static __forceinline__ int base(int lane) {
    int b = 0;
#pragma unroll
    for (int l = 0; l < H; ++l)
      if (lane & (NTILES << l)) b += V >> (l + 1);
    return b;
  }
  // Among lanes holding identical copies (butterfly levels), lane 0 of the group writes.
  __device__ static __forceinline__ bool writer(int lane) {
    int mask = 0;
#pragma unroll
    for (int l = H; l < QB; ++l) mask |= NTILES << l;
    return (lane & mask) == 0;
  }
};

// ---- final reduction over workgroup partials ----------------------------------------
// Sums partials[g][o] for g in [0, G) in a fixed order for outputs o in
// [o0, o1) of the KC x MC class, then writes C.  Called by one workgroup.
// `scratch` must hold kThreads doubles of LDS.
// Stores one reduced output (class index o) into C.
__device__ __forceinline__ void store_output(const AtbArgs& a, int MC, int o, double total) {
  const int i = o / MC, j = o % MC;
  if (i < a.k && j < a.m) {
    const int64_t off = (int64_t)i * a.ldc + j;
    if (a.c_f32) {
      float* C = static_cast<float*>(a.C);
      C[off] = a.accumulate ? (float)((double)C[off] + total) : (float)total;
    } else {
      double* C = static_cast<double*>(a.C);
      C[off] = a.accumulate ? C[off] + total : total;
    }
  }
}

__device__ inline void finish_outputs(const AtbArgs& a, int G, int KCMC, int MC, int o0, int o1,
                                      double* scratch) {
  const int tid = threadIdx.x;
  const int nout = o1 - o0;
  const double* P = a.partials;
  auto store = [&](int o, double total) __attribute__((always_inline)) {
    store_output(a, MC, o, total);
  };
  // Partials g0 .. g1-1 of output o, added in g order, up to U loads in flight
  // (the partials were just written through to memory: every load is a full
  // memory round trip).
  constexpr int U = 16;
  auto sum = [&](int o, int g0, int g1) __attribute__((always_inline)) {
    double acc = 0.0;
    for (int g = g0; g < g1; g += U) {
      double v[U];
#pragma unroll
      for (int x = 0; x < U; ++x) v[x] = g + x < g1 ? P[(int64_t)(g + x) * KCMC + o] : 0.0;
#pragma unroll
      for (int x = 0; x < U; ++x) acc += v[x];
    }
    return acc;
  };
  if (nout >= kThreads) {
    for (int o = o0 + tid; o < o1; o += kThreads) store(o, sum(o, 0, G));
    return;
  }
  // nout < kThreads: SL slices of the workgroup range per output.
  const int SL = kThreads / nout;
  const int oi = tid % nout, s = tid / nout;
  if (s < SL) {
    const int g0 = (int)((int64_t)G * s / SL), g1 = (int)((int64_t)G * (s + 1) / SL);
    scratch[s * nout + oi] = sum(o0 + oi, g0, g1);
  }
  __syncthreads();
  if (tid < nout) {
    double t = 0.0;
    for (int q = 0; q < SL; ++q) t += scratch[q * nout + tid];
    store(o0 + tid, t);
  }
}

// Final reduction of G partials of a KCMC-output class by one workgroup, for
// KCMC a power of two with 2 <= KCMC <= 512.  The G x KCMC partials are read
// as 16-byte chunks c = tid + 256 * j, so every thread always meets the same two outputs
// (2 * tid mod KCMC and the next one) and all of its loads are independent.
// They are issued in batches of U = 32 guarded loads, so a thread with up to
// 32 chunks (the 256 workgroup partials of an 8 x 8 class) waits for one
// memory round trip (the partials were just written through to memory, so
// each round trip is a full memory latency).  The load registers are free
// by then: the streaming loop's registers are dead.  Each
// thread sums its chunks in j order; the 512 / KCMC threads that share a pair
// of outputs are then added in thread order.  `scratch` holds 2 * kThreads
// doubles of LDS.
template <int KCMC>
__device__ __forceinline__ void finish_small(const AtbArgs& a, int G, int MC, double* scratch) {
  static_assert(KCMC >= 2 && KCMC <= 512 && (KCMC & (KCMC - 1)) == 0, "2..512, power of two");
  const int tid = threadIdx.x;
  const v4u* P = reinterpret_cast<const v4u*>(a.partials);
  const int64_t chunks = (int64_t)G * KCMC / 2;  // 16-byte chunks
  double s0 = 0.0, s1 = 0.0;
  constexpr int U = 32;
  for (int64_t c = tid; c < chunks; c += U * kThreads) {
    v4u u[U];
#pragma unroll
    for (int x = 0; x < U; ++x) {
      const int64_t i = c + (int64_t)x * kThreads;
      u[x] = i < chunks ? P[i] : v4u{0, 0, 0, 0};
    }
#pragma unroll
    for (int x = 0; x < U; ++x) {
      if (c + (int64_t)x * kThreads < chunks) {
        s0 += __hiloint2double((int)u[x].y, (int)u[x].x);
        s1 += __hiloint2double((int)u[x].w, (int)u[x].z);
      }
    }
  }
  scratch[2 * tid] = s0;
  scratch[2 * tid + 1] = s1;
  __syncthreads();
  constexpr int PAIRS = KCMC / 2;             // threads 0..PAIRS-1 hold distinct output pairs
  constexpr int SHARE = kThreads / PAIRS;     // threads per output pair: tid, tid + PAIRS, ...
  for (int o = tid; o < KCMC; o += kThreads) {  // KCMC may exceed the 256 threads
    const int pair = o / 2, half = o % 2;
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < SHARE; ++q) t += scratch[2 * (pair + q * PAIRS) + half];
    store_output(a, MC, o, t);
  }
}

// Finisher kernel body (launched after the streaming kernel for classes of
// more than 512 outputs and for large grids, kFinishOutputs outputs per
// workgroup).  Thread t sums
// output o0 + t % 16 over slice t / 16 of the G workgroup partials, in
// workgroup order, with up to 16 independent loads in flight (each wave reads
// four partial rows of 16 consecutive outputs); the 16 slice sums are then
// added in slice order.
constexpr int kFinishOutputs = 16;
__device__ inline void finish_body(const AtbArgs& a, int G, int KC, int MC) {
  constexpr int SLICES = kThreads / kFinishOutputs, U = 16;
  __shared__ double scratch[kThreads];
  const int KCMC = KC * MC;
  const int o0 = blockIdx.x * kFinishOutputs;
  if (o0 >= KCMC) return;
  const int oi = threadIdx.x % kFinishOutputs, sl = threadIdx.x / kFinishOutputs;
  const int nout = KCMC - o0 < kFinishOutputs ? KCMC - o0 : kFinishOutputs;
  double acc = 0.0;
  if (oi < nout) {
    const int g0 = (int)((int64_t)G * sl / SLICES), g1 = (int)((int64_t)G * (sl + 1) / SLICES);
    const double* P = a.partials + o0 + oi;
    for (int g = g0; g < g1; g += U) {
      double v[U];
#pragma unroll
      for (int x = 0; x < U; ++x) v[x] = g + x < g1 ? P[(int64_t)(g + x) * KCMC] : 0.0;
#pragma unroll
      for (int x = 0; x < U; ++x) acc += v[x];
    }
  }
  scratch[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < nout) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < SLICES; ++q) t += scratch[q * kFinishOutputs + threadIdx.x];
    store_output(a, MC, o0 + threadIdx.x, t);
  }
}

// ---- MFMA consumer (width classes that are multiples of 16) -------------------------
// The k x m result is a grid of 16 x 16 blocks, each accumulated by one
// matrix instruction per K-step of rows:
//   FP64: v_mfma_f64_16x16x4f64   4 rows / step, FP64 accumulators
//   FP32: v_mfma_f32_16x16x4f32   4 rows / step, FP32 accumulators
//   BF16: v_mfma_f32_16x16x32_bf16 32 rows / step, FP32 accumulators
// Operands come from the row-major LDS tile: lane l supplies column
// 16*block + (l % 16) of the rows that the step assigns to lane group l / 16
// (the same rows for A and B, so any row order within a step is a valid
// contraction).  Output layout of one block (measured on gfx950): lane l,
// register v holds row 4*(l/16) + v (FP32, BF16) or row l/16 + 4*v (FP64),
// column l % 16.  Wave w takes the steps w, w + 4, ... of every tile, so each
// wave ends with a full KC x MC partial; FP32 accumulators are folded into
// (hi, lo) pairs every 64 (FP32) or 128 (BF16) rows per wave.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <class S>
struct MfmaTraits;
template <>
struct MfmaTraits<double> {
  using AccV = v4d;
  static constexpr int KSTEP = 4;
  static constexpr bool row_major_out = false;  // row l/16 + 4v
};
template <>
struct MfmaTraits<float> {
  using AccV = v4f;
  static constexpr int KSTEP = 4;
  static constexpr bool row_major_out = true;  // row 4*(l/16) + v
};
template <>
struct MfmaTraits<bf16> {
  using AccV = v4f;
  static constexpr int KSTEP = 32;
  static constexpr bool row_major_out = true;
};

// Operand of one 16-wide block for one K-step.
__device__ __forceinline__ double mfma_operand(const double* X, int ld, int row0, int blk,
                                               int lane) {
  return X[(row0 + (lane >> 4)) * ld + blk * 16 + (lane & 15)];
}
__device__ __forceinline__ float mfma_operand(const float* X, int ld, int row0, int blk,
                                              int lane) {
  return X[(row0 + (lane >> 4)) * ld + blk * 16 + (lane & 15)];
}
// BF16: rows row0 + 8 * (l / 16) + t, t = 0..7, of column 16 * blk + l % 16,
// with two LDS transpose reads (ds_read_b64_tr_b16).  Within each group of
// 16 lanes that instruction hands output lane 4q + e the element e of the
// 8-byte reads of lanes q, q + 4, q + 8, q + 12 (measured on gfx950), so lane
// 16g + 4j + q reads 4 columns (16 * blk + 4q ...) of row 8g + j (then of
// row 8g + 4 + j) and lane 16g + i receives column 16 * blk + i of those rows.
typedef short v4s_ __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4s_ lds_v4s_;
__device__ __forceinline__ v8bf mfma_operand(const bf16* X, int ld, int row0, int blk,
                                             int lane) {
  const int g = lane >> 4, j = (lane >> 2) & 3, q = lane & 3;
  const bf16* p0 = X + (row0 + 8 * g + j) * ld + blk * 16 + 4 * q;
  const v4s_ a = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s_*)p0);
  const v4s_ b = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s_*)(p0 + 4 * ld));
  const uint32_t w[4] = {(uint32_t)(uint16_t)a[0] | ((uint32_t)(uint16_t)a[1] << 16),
                         (uint32_t)(uint16_t)a[2] | ((uint32_t)(uint16_t)a[3] << 16),
                         (uint32_t)(uint16_t)b[0] | ((uint32_t)(uint16_t)b[1] << 16),
                         (uint32_t)(uint16_t)b[2] | ((uint32_t)(uint16_t)b[3] << 16)};
  return __builtin_bit_cast(v8bf, w);
}
// -x for eight packed BF16 values (flips the sign bits).
__device__ __forceinline__ v8bf mfma_negate(v8bf x) {
  v4u w = __builtin_bit_cast(v4u, x);
  w ^= 0x80008000u;
  return __builtin_bit_cast(v8bf, w);
}
__device__ __forceinline__ v4d mfma_step(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma_step(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma_step(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// SYM = 1 (A^T A, KC == MC): only the blocks on and above the diagonal are
// accumulated; emit() writes each off-diagonal block and its transpose, so
// the result is exactly symmetric outside the diagonal blocks.
template <class S, int KC, int MC, int T, int SYM = 0>
struct MfmaAcc {
  using Tr = MfmaTraits<S>;
  using AccV = typename Tr::AccV;
  static constexpr bool kComp = Traits<S>::compensated;
  static constexpr int IB = KC / 16, JB = MC / 16, KSTEP = Tr::KSTEP;
  static constexpr int NSTEP = T / KSTEP;
  // The BF16 instruction rounds its sum of 32 products (and the accumulator)
  // with a bias toward -infinity, about 1e-8 of the sum of |products| per
  // instruction on inputs spread over many binades.  A second accumulator
  // runs the same steps with one operand negated; its rounding bias has the
  // opposite sign in P - Q, and the result is (P - Q) / 2.
  static constexpr bool kPair = KSTEP == 32;
  static constexpr bool kNegA = IB <= JB;  // negate the operand with fewer blocks
  static_assert(!SYM || KC == MC, "A^T A is square");
  static constexpr bool used(int i, int j) { return !SYM || i <= j; }
  static_assert(KC % 16 == 0 && MC % 16 == 0, "MFMA needs widths that are multiples of 16");
  static_assert(NSTEP % 4 == 0, "each of the 4 waves takes the same number of steps");
  static_assert(!kPair || kComp, "the paired accumulators fold into (hi, lo)");
  AccV acc[IB][JB];
  AccV accn[kPair ? IB : 1][kPair ? JB : 1];
  AccV hi[kComp ? IB : 1][kComp ? JB : 1];
  AccV lo[kComp ? IB : 1][kComp ? JB : 1];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
      for (int j = 0; j < JB; ++j) {
        acc[i][j] = AccV{};
        if constexpr (kPair) accn[i][j] = AccV{};
        if constexpr (kComp) hi[i][j] = lo[i][j] = AccV{};
      }
  }
  __device__ __forceinline__ void consume(const unsigned char* a_stage,
                                         const unsigned char* b_stage, int wave, int lane) {
    const S* As = reinterpret_cast<const S*>(a_stage);
    const S* Bs = reinterpret_cast<const S*>(b_stage);
#pragma unroll
    for (int s = 0; s < NSTEP / 4; ++s) step(As, Bs, (wave + 4 * s) * KSTEP, lane);
  }
  // One K-step: rows row0 .. row0 + KSTEP - 1 of LDS tiles As (pitch KC) and
  // Bs (pitch MC).
  __device__ __forceinline__ void step(const S* As, const S* Bs, int row0, int lane) {
    {
      decltype(mfma_operand(As, KC, 0, 0, 0)) av[IB], bv[JB];
#pragma unroll
      for (int i = 0; i < IB; ++i) av[i] = mfma_operand(As, KC, row0, i, lane);
#pragma unroll
      for (int j = 0; j < JB; ++j) bv[j] = mfma_operand(Bs, MC, row0, j, lane);
#pragma unroll
      for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j)
          if (used(i, j)) acc[i][j] = mfma_step(av[i], bv[j], acc[i][j]);
      if constexpr (kPair) {
        if constexpr (kNegA) {
#pragma unroll
          for (int i = 0; i < IB; ++i) av[i] = mfma_negate(av[i]);
        } else {
#pragma unroll
          for (int j = 0; j < JB; ++j) bv[j] = mfma_negate(bv[j]);
        }
#pragma unroll
        for (int i = 0; i < IB; ++i)
#pragma unroll
          for (int j = 0; j < JB; ++j)
            if (used(i, j)) accn[i][j] = mfma_step(av[i], bv[j], accn[i][j]);
      }
    }
  }
  __device__ __forceinline__ void fold() {
    if constexpr (kComp) {
#pragma unroll
      for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            if (!used(i, j)) continue;
            float x = acc[i][j][v];
            if constexpr (kPair) {
              x -= accn[i][j][v];  // 2 x (block sum)
              accn[i][j][v] = 0.0f;
            }
            float s, e;
            two_sum(hi[i][j][v], x, s, e);
            hi[i][j][v] = s;
            lo[i][j][v] += e;
            acc[i][j][v] = 0.0f;
          }
    }
  }
  // Adds this wave's KC x MC partial into red[] (FP64, row-major KC x MC).
  __device__ __forceinline__ void emit(double* red, bool first, int lane) const {
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
      for (int j = 0; j < JB; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (!used(i, j)) continue;
          const int r = Tr::row_major_out ? 4 * (lane >> 4) + v : (lane >> 4) + 4 * v;
          const int c = lane & 15;
          const int o = (i * 16 + r) * MC + j * 16 + c;
          double x;
          if constexpr (kPair)
            x = 0.5 * ((double)hi[i][j][v] + (double)lo[i][j][v]);
          else if constexpr (kComp)
            x = (double)hi[i][j][v] + (double)lo[i][j][v];
          else
            x = acc[i][j][v];
          red[o] = first ? x : red[o] + x;
          if (SYM && i < j) {
            const int ot = (j * 16 + c) * MC + i * 16 + r;  // the transposed element
            red[ot] = first ? x : red[ot] + x;
          }
        }
  }
};

// ---- publishing workgroup partials ----------------------------------------------------
// Each XCD of gfx950 has its own L2, so a plain store is visible to the other
// XCDs only after an agent-scope release (a writeback of the whole L2).  The
// partials are instead stored with agent-scope atomic stores, which write
// through to memory; the publishing wave then only waits until its stores are
// complete before the arrival counter is incremented.  Only the last
// workgroup to arrive fences (acquire: it invalidates its L2) before reading.
//
// The ordering rests on the gfx942/gfx950 memory model of the AMDGPU backend
// (LLVM AMDGPUUsage, "Memory Model GFX942"): an agent-scope atomic store is a
// global_store with sc1 set, which writes through the issuing XCD's L2 to
// memory, and vmcnt counts it until the write is acknowledged there.  So after
// `s_waitcnt vmcnt(0)` the partial is in memory, where the counter atomic (also
// performed at memory at agent scope) cannot overtake it; the last arriver's
// acquire fence invalidates its L2 before it reads.  The relaxed store plus the
// explicit wait replace the release fence a C++ release store would emit
// (buffer_wbl2, a writeback of the whole L2).  A compiler that stopped
// lowering agent-scope stores as write-through would break this: the
// determinism and parity tests of every class exercise it.
__device__ __forceinline__ void publish(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// s_waitcnt vmcnt(0) (gfx9 encoding: vmcnt 0, expcnt 7, lgkmcnt 15): every
// vector memory operation of this wave has completed.  The signal fences
// keep the compiler from moving memory operations across the wait.
__device__ __forceinline__ void wait_vector_memory() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// ---- the streaming kernel body -----------------------------------------------------
// LDS bytes a configuration needs: two pipeline stages, the KC x MC FP64
// reduction buffer or the final-reduction scratch, whichever is largest.
template <class S, int KC, int MC, int T, int SYM = 0>
constexpr int atb_smem_bytes() {
  constexpr int stage2 = 2 * T * (KC + (SYM ? 0 : MC)) * (int)sizeof(S);
  constexpr int red = KC * MC * 8;
  constexpr int m1 = stage2 > red ? stage2 : red;
  return m1 > kThreads * 16 ? m1 : kThreads * 16;
}

// Width the consumer works on for a side of class width W: the matrix-core
// consumer (TM = TN = 0) takes 16-wide blocks, so a narrower side is staged
// into LDS rows of 16 whose remaining columns stay zero.  The register tiles
// work on the class width itself.
template <int W, int TM, int TN>
constexpr int consumer_width() {
  return (TM == 0 && TN == 0 && W < 16) ? 16 : W;
}

// Register-tile consumer (VALU): the k x m result (padded to KC x MC) is cut
// into NTILES = (KC/TM) * (MC/TN) register tiles; thread t owns tile
// p = t % NTILES (A columns [mb*TM, mb*TM+TM), B columns [nb*TN, nb*TN+TN))
// and the rows q, q + RG, q + 2RG, ... (q = t / NTILES, RG = 256 / NTILES) of
// every tile.  FP32/BF16 block sums are folded into (hi, lo) about every 32
// rows.
template <class S, int KC, int MC, int TM, int TN, int T>
struct RegTileAcc {
  using Acc = typename Traits<S>::Acc;
  static constexpr bool kComp = Traits<S>::compensated;
  static constexpr int MB = KC / TM, NB = MC / TN, NTILES = MB * NB;
  static_assert(KC % TM == 0 && MC % TN == 0, "tile must divide the width class");
  static_assert(NTILES >= 1 && NTILES <= 64 && 64 % NTILES == 0, "1..64 register tiles");
  static constexpr int RG = kThreads / NTILES;  // row lanes
  static_assert(T % RG == 0, "tile rows must be a multiple of the row lanes");
  static constexpr int RPT = T / RG;            // rows per thread per tile
  static constexpr int FOLD_TILES = RPT >= 32 ? 1 : 32 / RPT;
  int p, q, mb, nb, fold_count;
  Acc acc[TM][TN];
  Acc hi[kComp ? TM : 1][kComp ? TN : 1];
  Acc lo[kComp ? TM : 1][kComp ? TN : 1];

  __device__ __forceinline__ void zero(int tid) {
    p = tid % NTILES;
    q = tid / NTILES;
    mb = p % MB;
    nb = p / MB;
    fold_count = 0;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc[i][j] = Acc(0);
        if constexpr (kComp) hi[i][j] = lo[i][j] = Acc(0);
      }
  }
  __device__ __forceinline__ void consume(const unsigned char* a_stage,
                                         const unsigned char* b_stage) {
    const S* As = reinterpret_cast<const S*>(a_stage);
    const S* Bs = reinterpret_cast<const S*>(b_stage);
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int r = q + s * RG;
      Acc av[TM], bv[TN];
      lds_read<S, TM>(As + r * KC + mb * TM, av);
      lds_read<S, TN>(Bs + r * MC + nb * TN, bv);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = fmadd(av[i], bv[j], acc[i][j]);
    }
  }
  __device__ __forceinline__ void fold() {
    if constexpr (kComp) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          float s, e;
          two_sum(hi[i][j], acc[i][j], s, e);
          hi[i][j] = s;
          lo[i][j] += e;
          acc[i][j] = 0.0f;
        }
    }
  }
  __device__ __forceinline__ void after_tile() {
    if constexpr (kComp) {
      if (++fold_count == FOLD_TILES) {
        fold();
        fold_count = 0;
      }
    }
  }
  // Workgroup reduction into red[] (FP64 in LDS); ends with a barrier, after
  // which total(red, o) is the workgroup's sum for output o.  When four
  // KC x MC slots fit in the two stages of A's tiles (a lower bound of the
  // LDS of every kernel, A^T A included), each wave writes its slot at once
  // and total() adds the slots in wave order (one barrier); otherwise the
  // waves add into one slot in turn.  Both give ((w0 + w1) + w2) + w3.
  static constexpr bool kSlots = 4 * KC * MC * 8 <= 2 * T * KC * (int)sizeof(S);
  __device__ __forceinline__ double total(const double* red, int o) const {
    if constexpr (kSlots) {
      constexpr int K = KC * MC;
      return ((red[o] + red[K + o]) + red[2 * K + o]) + red[3 * K + o];
    } else {
      return red[o];
    }
  }
  __device__ __forceinline__ void reduce(double* red, int tid) {
    constexpr int V = TM * TN;
    double v[V];
    if constexpr (kComp) {
      fold();
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) v[i * TN + j] = (double)hi[i][j] + (double)lo[i][j];
    } else {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) v[i * TN + j] = acc[i][j];
    }
    using WR = WaveReduce<NTILES, V>;
    const int lane = tid & 63, wave = tid >> 6;
    WR::run(v, lane);
    __syncthreads();  // every wave is done with the LDS tiles
    const bool writer = WR::writer(lane);
    const int vbase = WR::base(lane);
    if constexpr (kSlots) {
      if (writer) {
#pragma unroll
        for (int x = 0; x < WR::VL; ++x) {
          const int idx = vbase + x;
          const int o = (mb * TM + idx / TN) * MC + nb * TN + idx % TN;
          red[wave * KC * MC + o] = v[x];
        }
      }
      __syncthreads();
      return;
    }
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
      if (wave == w && writer) {
#pragma unroll
        for (int x = 0; x < WR::VL; ++x) {
          const int idx = vbase + x;
          const int o = (mb * TM + idx / TN) * MC + nb * TN + idx % TN;
          red[o] = (w == 0) ? v[x] : red[o] + v[x];
        }
      }
      __syncthreads();
    }
  }
};

// The MFMA consumer with the same interface (TM = TN = 0 selects it).
// FP32/BF16 accumulators are folded into (hi, lo) every FOLD tiles: about 64
// rows per wave for FP32 (rounded products) and 128 for BF16 (exact products,
// summed in FP32 by the matrix instruction).
template <class S, int KC, int MC, int T, int SYM = 0>
struct MfmaConsumer {
  MfmaAcc<S, KC, MC, T, SYM> m;
  static constexpr int kWaveRows = T / 4;  // rows of a tile each wave consumes
  static constexpr int kFoldRows = MfmaTraits<S>::KSTEP == 32 ? 128 : 64;
  static constexpr int FOLD = kWaveRows >= kFoldRows ? 1 : kFoldRows / kWaveRows;
  int wave, lane, fold_count;
  __device__ __forceinline__ void zero(int tid) {
    wave = tid >> 6;
    lane = tid & 63;
    fold_count = 0;
    m.zero();
  }
  __device__ __forceinline__ void consume(const unsigned char* a_stage,
                                         const unsigned char* b_stage) {
    m.consume(a_stage, b_stage, wave, lane);
  }
  __device__ __forceinline__ void after_tile() {
    if (++fold_count == FOLD) {
      m.fold();
      fold_count = 0;
    }
  }
  __device__ __forceinline__ double total(const double* red, int o) const { return red[o]; }
  __device__ __forceinline__ void reduce(double* red, int) {
    m.fold();
    __syncthreads();  // every wave is done with the LDS tiles
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
      if (wave == w) m.emit(red, w == 0, lane);
      __syncthreads();
    }
  }
};

template <class S, int KC, int MC, int TM, int TN, int T, int SYM = 0>
struct ConsumerOf {
  using type = RegTileAcc<S, KC, MC, TM, TN, T>;
};
template <class S, int KC, int MC, int T, int SYM>
struct ConsumerOf<S, KC, MC, 0, 0, T, SYM> {
  using type = MfmaConsumer<S, KC, MC, T, SYM>;
};

// Streams a sequence of row tiles of one problem - the full tiles
// first + j * step (j < count) and, if owns_tail, the partial tile of
// tail_rows rows that starts at row full * T - and writes this workgroup's
// KC x MC FP64 partial sum to partial_out (global memory).  smem must hold
// atb_smem_bytes<S, KC, MC, T, SYM>() bytes.  TM = TN = 0 selects the MFMA
// consumer.  SYM = 1 computes A^T A (B is A, KC == MC): only A's tile is
// staged and the consumer reads both operands from it, so every product is
// formed and added exactly as for A^T B with B = A, from half the bytes.
template <class S, int KC, int MC, int TM, int TN, int T, int PF, int NT, int SYM = 0>
__device__ __forceinline__ void atb_partial(const AtbArgs& a, unsigned char* smem, int64_t first,
                                            int64_t step, int64_t count, bool owns_tail,
                                            int64_t full, int tail_rows, double* partial_out) {
  static_assert(!SYM || KC == MC, "A^T A is square");
  constexpr int ES = (int)sizeof(S);
  // Consumer widths: a matrix-core kernel of a side narrower than 16 (rows of
  // whole 16-byte chunks, or of 8 bytes) stages that side's rows into LDS
  // rows of KP (MP) = 16 elements, the rest zero.  The padding columns feed
  // only outputs outside the class's KC x MC block, which are dropped.
  constexpr int KP = consumer_width<KC, TM, TN>(), MP = consumer_width<MC, TM, TN>();
  constexpr bool PAD_A = KP != KC, PAD_B = MP != MC;
  static_assert(!(PAD_A || PAD_B) || ((KC * ES % 16 == 0 || KC * ES == 8) &&
                                      (MC * ES % 16 == 0 || MC * ES == 8)),
                "padded rows are whole 16-byte chunks or 8 bytes");
  constexpr int A_BYTES = T * KC * ES, B_BYTES = SYM ? 0 : T * MC * ES;  // a tile in memory
  constexpr int LA_BYTES = T * KP * ES, LB_BYTES = SYM ? 0 : T * MP * ES;  // ... in LDS
  static_assert(A_BYTES % 16 == 0 && B_BYTES % 16 == 0, "tiles are whole 16-byte chunks");
  constexpr int A_VEC = A_BYTES / 16, B_VEC = B_BYTES / 16;
  constexpr int A_PER = (A_VEC + kThreads - 1) / kThreads;
  constexpr int B_PER = (B_VEC + kThreads - 1) / kThreads;
  constexpr int STAGE = LA_BYTES + LB_BYTES;
  constexpr int B_OFF = SYM ? 0 : LA_BYTES;  // B's tile within a stage
  constexpr int KCMC = KC * MC;
  static_assert(PF >= 1 && PF <= 4, "one to four tiles in flight");
  const int tid = threadIdx.x;
  typename ConsumerOf<S, KP, MP, TM, TN, T, SYM>::type cons;
  cons.zero(tid);
  if constexpr (PAD_A || PAD_B) {
    // The padding columns are never written below: zero both stages once,
    // so that the dropped outputs are sums of zeros rather than of whatever
    // an earlier use of this LDS left there.
    v4u* z = reinterpret_cast<v4u*>(smem);
    for (int x = tid; x < 2 * STAGE / 16; x += kThreads) z[x] = v4u{0u, 0u, 0u, 0u};
    __syncthreads();
  }
  // Stores 16-byte chunk c of a tile of rows of W elements into an LDS tile
  // of rows of WP elements.  Rows of 8 bytes come two to a chunk: its halves
  // go to consecutive LDS rows.
  auto put = [](unsigned char* base, int c, const v4u& u, auto w, auto wp)
      __attribute__((always_inline)) {
    constexpr int RW = decltype(w)::value * ES, RP = decltype(wp)::value * ES;  // row bytes
    if constexpr (RW == RP) {
      reinterpret_cast<v4u*>(base)[c] = u;
    } else if constexpr (RW == 8) {
      typedef uint32_t half_chunk __attribute__((ext_vector_type(2)));
      half_chunk* p = reinterpret_cast<half_chunk*>(base);
      p[(2 * c) * (RP / 8)] = half_chunk{u.x, u.y};
      p[(2 * c + 1) * (RP / 8)] = half_chunk{u.z, u.w};
    } else {
      reinterpret_cast<v4u*>(base)[(c / (RW / 16)) * (RP / 16) + c % (RW / 16)] = u;
    }
  };
  using KW = std::integral_constant<int, KC>;
  using KWP = std::integral_constant<int, KP>;
  using MW = std::integral_constant<int, MC>;
  using MWP = std::integral_constant<int, MP>;

  const int64_t nfull = a.dense ? count : 0;

  // ---- pipelined dense tiles ---------------------------------------------------
  if (nfull > 0) {
    gv4u* gA = (gv4u*)a.A;
    gv4u* gB = (gv4u*)a.B;
    struct Regs {
      v4u a[A_PER];
      v4u b[B_PER > 0 ? B_PER : 1];
    };
    auto load = [&](Regs& r, int64_t t) __attribute__((always_inline)) {
      const int64_t tile = first + t * step;
      gv4u* ta = gA + tile * A_VEC;
      gv4u* tb = gB + tile * B_VEC;
#pragma unroll
      for (int v = 0; v < A_PER; ++v) {
        const int idx = v * kThreads + tid;
        if (A_VEC % kThreads == 0 || idx < A_VEC) r.a[v] = stream_load<NT>(ta + idx);
      }
#pragma unroll
      for (int v = 0; v < B_PER; ++v) {
        const int idx = v * kThreads + tid;
        if (B_VEC % kThreads == 0 || idx < B_VEC) r.b[v] = stream_load<NT>(tb + idx);
      }
    };
    auto store = [&](const Regs& r, unsigned char* stage) __attribute__((always_inline)) {
#pragma unroll
      for (int v = 0; v < A_PER; ++v) {
        const int idx = v * kThreads + tid;
        if (A_VEC % kThreads == 0 || idx < A_VEC) put(stage, idx, r.a[v], KW{}, KWP{});
      }
#pragma unroll
      for (int v = 0; v < B_PER; ++v) {
        const int idx = v * kThreads + tid;
        if (B_VEC % kThreads == 0 || idx < B_VEC)
          put(stage + LA_BYTES, idx, r.b[v], MW{}, MWP{});
      }
    };
    // One pipeline step on tile tt held in r: publish it to LDS, refill r with
    // tile tt + PF, consume the LDS copy.
    // The refill is unconditional (past the last tile it re-reads the last
    // tile, which is in bounds and discarded) so that the number of loads in
    // flight is the same on every path and each publish waits only for the
    // oldest tile.  (Peeling the last PF steps to skip those re-reads was
    // measured slower: the larger loop code costs more than the re-reads.)
    const int64_t last = nfull - 1;
    auto step = [&](Regs& r, int64_t tt) __attribute__((always_inline)) {
      unsigned char* stage = smem + (tt & 1) * STAGE;
      // sched_barrier keeps the phases in program order: the scheduler would
      // otherwise hoist refill loads above the publish, and since vmcnt
      // retires in issue order the publish would then wait for them too.
      __builtin_amdgcn_sched_barrier(0);
      store(r, stage);
      __syncthreads();
      load(r, tt + PF < last ? tt + PF : last);
      __builtin_amdgcn_sched_barrier(0);
      cons.consume(stage, stage + B_OFF);
      cons.after_tile();
    };
    // PF register sets, used in rotation: set p holds tiles p, p + PF, ...
    Regs r[PF];
#pragma unroll
    for (int p = 0; p < PF; ++p) load(r[p], p < last ? p : last);
    int64_t t = 0;
    for (; t + PF <= nfull; t += PF) {
#pragma unroll
      for (int p = 0; p < PF; ++p) step(r[p], t + p);
    }
#pragma unroll
    for (int p = 0; p < PF - 1; ++p)
      if (t + p < nfull) step(r[p], t + p);
  }

  // ---- element loads (strided or misaligned inputs, the partial tile) -------------
  auto elem_tile = [&](int64_t r0, int nr) __attribute__((always_inline)) {
    const S* gA = static_cast<const S*>(a.A);
    const S* gB = static_cast<const S*>(a.B);
    __syncthreads();
    S* la = reinterpret_cast<S*>(smem);
    S* lb = reinterpret_cast<S*>(smem + LA_BYTES);
    for (int e = tid; e < T * KC; e += kThreads) {
      const int r = e / KC, c = e % KC;
      S v{};
      if (r < nr && c < a.k) v = gA[(r0 + r) * a.lda + c];
      la[r * KP + c] = v;
    }
    if constexpr (!SYM) {
      for (int e = tid; e < T * MC; e += kThreads) {
        const int r = e / MC, c = e % MC;
        S v{};
        if (r < nr && c < a.m) v = gB[(r0 + r) * a.ldb + c];
        lb[r * MP + c] = v;
      }
    }
    __syncthreads();
    cons.consume(smem, smem + B_OFF);
    cons.after_tile();
  };
  if (!a.dense)
    for (int64_t j = 0; j < count; ++j) elem_tile((first + j * step) * T, T);
  if (owns_tail) elem_tile(full * T, tail_rows);

  // ---- reduction -------------------------------------------------------------------
  double* red = reinterpret_cast<double*>(smem);
  cons.reduce(red, tid);
  // the consumer's KP x MP result holds the class's KC x MC block at its top left
  for (int o = tid; o < KCMC; o += kThreads)
    publish(partial_out + o, cons.total(red, (PAD_A || PAD_B) ? (o / MC) * MP + o % MC : o));
}

// Arrival at a problem's counter after its partial is stored: returns true
// (in every thread) for the last of `expected` arrivals, which then sees all
// partials.  That workgroup resets the counter to zero after using it.
__device__ __forceinline__ bool arrive_last(unsigned int* counter, unsigned int expected,
                                            int* s_flag) {
  wait_vector_memory();  // this wave's write-through partial stores are complete
  __syncthreads();
  if (threadIdx.x == 0) *s_flag = (atomicAdd(counter, 1u) == expected - 1) ? 1 : 0;
  __syncthreads();
  const bool last = *s_flag != 0;
  // Acquire the other workgroups' partials: an agent-scope acquire fence
  // invalidates this XCD's L2 (buffer_inv sc1).  No release is needed here
  // (this workgroup's own partial went out write-through, and C is released at
  // the kernel's end), so not __threadfence(), whose seq_cst fence adds a
  // writeback of the whole L2 (buffer_wbl2): 0.5 us per call at n = 2^22.
  if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return last;
}

// One problem over the whole grid: contiguous runs of tiles_per_wg tiles per
// workgroup, or tile t to workgroup t % G (interleave); then the last
// workgroup to arrive adds the G partials in workgroup order and writes C.
template <class S, int KC, int MC, int TM, int TN, int T, int PF, int NT, int SYM = 0>
__device__ __forceinline__ void atb_body(const AtbArgs& a) {
  constexpr int SMEM = atb_smem_bytes<S, consumer_width<KC, TM, TN>(),
                                      consumer_width<MC, TM, TN>(), T, SYM>();
  __shared__ __attribute__((aligned(16))) unsigned char smem[SMEM];
  __shared__ int s_last;
  const int64_t full = a.n / T;
  const int64_t G = gridDim.x, g = blockIdx.x;
  int64_t first, step, count;
  if (a.interleave) {
    first = g;
    step = G;
    count = first < full ? (full - first + G - 1) / G : 0;
  } else {
    first = g * a.tiles_per_wg;
    step = 1;
    count = first < full ? (full - first < a.tiles_per_wg ? full - first : a.tiles_per_wg) : 0;
  }
  const int tail_rows = (int)(a.n - full * T);
  const bool owns_tail =
      tail_rows > 0 && (a.interleave ? full % G == g : full / a.tiles_per_wg == g);
  constexpr int KCMC = KC * MC;
  atb_partial<S, KC, MC, TM, TN, T, PF, NT, SYM>(a, smem, first, step, count, owns_tail, full,
                                                 tail_rows, a.partials + g * KCMC);
  if (!a.finish) return;
  if (!arrive_last(a.counter, gridDim.x, &s_last)) return;
  if constexpr (KCMC >= 2 && KCMC <= 512)
    finish_small<KCMC>(a, (int)gridDim.x, MC, reinterpret_cast<double*>(smem));
  else
    finish_outputs(a, (int)gridDim.x, KCMC, MC, 0, KCMC, reinterpret_cast<double*>(smem));
  if (threadIdx.x == 0) __hip_atomic_store(a.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- direct matrix-core engine ------------------------------------------------------
// The FP64 and FP32 matrix instructions (16x16x4) take, per lane l, one
// element of row l / 16 of a 4-row K-step; the 16 lanes of a row group supply
// the 16 columns of a block.  With A row-major those 16 lanes read the whole
// row: lane j of the group loads LB = KC * sizeof(S) / 16 bytes, which hold
// IB = KC / 16 columns, and element b of its load feeds block b.  Up to 16
// bytes a lane's columns are consecutive (j * IB ... j * IB + IB - 1).  Wider
// rows (FP64, KC = 64: 32 bytes) are read as 16-byte chunks 256 bytes apart,
// so that each load instruction covers whole 256-byte row segments: chunk x of
// lane j holds the E = 16 / sizeof(S) columns 16 E x + E j ... 16 E x + E j +
// E - 1.  (Two adjacent 16-byte chunks per lane made every instruction touch
// each cache line half: 6.1 TB/s for the loads alone at 64 x 64, against 7.0
// with whole segments.)  lane_col gives the column of block b in lane j; block
// b's output row r is C row lane_col(b, r) (likewise for B and the C
// columns).  Each wave streams its own contiguous run of 4-row steps straight
// into registers, D steps in flight: no LDS staging and no workgroup barrier
// until the final sum.
typedef uint32_t v2u __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const unsigned char gu8;
typedef __attribute__((address_space(1))) const v2u gv2u;
typedef __attribute__((address_space(1))) const uint32_t gu32;

// Byte offset of lane j's first chunk within a row, and the column that
// element b of lane j's load holds (see above).
template <class S, int LB>
__device__ __forceinline__ constexpr uint32_t lane_offset(int j) {
  return (uint32_t)(j * (LB < 16 ? LB : 16));
}
template <class S, int LB>
__device__ __forceinline__ constexpr int lane_col(int b, int j) {
  constexpr int E = 16 / (int)sizeof(S);
  if constexpr (LB < 16)
    return j * (LB / (int)sizeof(S)) + b;
  else
    return (b / E) * 16 * E + j * E + b % E;
}

// LB bytes (4, 8, 16 or 32) into w[LB / 4]: one load at p below 16 bytes,
// else 16-byte chunks at p, p + 256, ...
template <int LB, int NT>
__device__ __forceinline__ void lane_load(uint32_t (&w)[LB / 4], gu8* p) {
  static_assert(LB == 4 || LB == 8 || LB % 16 == 0, "lane load size");
  if constexpr (LB >= 16) {
#pragma unroll
    for (int x = 0; x < LB / 16; ++x) {
      const v4u v = stream_load<NT>((gv4u*)(p + 256 * x));
      w[4 * x] = v.x;
      w[4 * x + 1] = v.y;
      w[4 * x + 2] = v.z;
      w[4 * x + 3] = v.w;
    }
  } else if constexpr (LB == 8) {
    const v2u v = NT ? __builtin_nontemporal_load((gv2u*)p) : *(gv2u*)p;
    w[0] = v.x;
    w[1] = v.y;
  } else {
    w[0] = NT ? __builtin_nontemporal_load((gu32*)p) : *(gu32*)p;
  }
}
template <class S>
__device__ __forceinline__ S lane_elem(const uint32_t* w, int e);
template <>
__device__ __forceinline__ double lane_elem<double>(const uint32_t* w, int e) {
  return __hiloint2double((int)w[2 * e + 1], (int)w[2 * e]);
}
template <>
__device__ __forceinline__ float lane_elem<float>(const uint32_t* w, int e) {
  return __uint_as_float(w[e]);
}

// One wave's KC x MC partial over its steps, then the fixed-order sum of the
// four waves' partials into partial_out (red: 2 * KC * MC doubles of LDS).
// SYM = 1: B is A (C = A^T A); only blocks i <= j are accumulated and the
// others mirrored.  a.tiles_per_wg is q, the pipelined steps per wave (a
// multiple of D): wave w takes steps [w q, w q + q); the F - W q steps left
// over (F = n / 4 full steps, W waves) go one at a time to waves in turn, and
// the last wave adds the n % 4 final rows.  FP32 accumulators are folded
// into TwoSum (hi, lo) FP32 pairs every 32 steps (128 rows per wave); FP64
// appears only in the cross-wave and cross-workgroup sums.  (Folding every
// 16 steps cost 7-10% at 64 x 64: the fold's dependent chain stalls the
// matrix instructions that feed the accumulators.  At 128 rows the
// Frobenius error on 4 GiB inputs is 2.1e-7 uniform, 6.1e-8 sign * 2^e.)
// Needs a.dense.
template <class S, int KC, int MC, int D, int NT, int SYM = 0>
__device__ __forceinline__ void direct_partial(const AtbArgs& a, double* red,
                                               double* partial_out) {
  static_assert(KC % 16 == 0 && MC % 16 == 0, "matrix-core widths");
  static_assert(!SYM || KC == MC, "A^T A is square");
  constexpr int IB = KC / 16, JB = MC / 16;
  constexpr int LBA = KC * (int)sizeof(S) / 16, LBB = MC * (int)sizeof(S) / 16;
  constexpr int RA = KC * (int)sizeof(S), RB = MC * (int)sizeof(S);  // bytes per row
  constexpr bool kF32 = sizeof(S) == 4;
  constexpr int FOLD = 32;  // steps between FP32 (hi, lo) folds (D divides it)
  static_assert(!kF32 || FOLD % D == 0, "D divides the fold interval");
  using AccV = typename MfmaTraits<S>::AccV;
  auto used = [](int i, int j) constexpr { return !SYM || i <= j; };

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, j = lane & 15;
  const int64_t F = a.n / 4;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t wid = (int64_t)blockIdx.x * 4 + wave;
  const int64_t q = a.tiles_per_wg;
  const int64_t s0 = wid * q;
  gu8* A8 = (gu8*)a.A;
  gu8* B8 = (gu8*)a.B;
  const uint32_t la = (uint32_t)(g * RA) + lane_offset<S, LBA>(j);
  const uint32_t lb = (uint32_t)(g * RB) + lane_offset<S, LBB>(j);

  AccV acc[IB][JB];
  AccV hi[kF32 ? IB : 1][kF32 ? JB : 1], lo[kF32 ? IB : 1][kF32 ? JB : 1];
#pragma unroll
  for (int i = 0; i < IB; ++i)
#pragma unroll
    for (int jj = 0; jj < JB; ++jj) {
      acc[i][jj] = AccV{};
      if constexpr (kF32) hi[i][jj] = lo[i][jj] = AccV{};
    }
  uint32_t wa[D][LBA / 4], wb[SYM ? 1 : D][SYM ? 1 : LBB / 4];
  auto ld = [&](int d, int64_t step) __attribute__((always_inline)) {
    lane_load<LBA, NT>(wa[d], A8 + step * 4 * RA + la);
    if constexpr (!SYM) lane_load<LBB, NT>(wb[d], B8 + step * 4 * RB + lb);
  };
  auto consume = [&](int d) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < IB; ++i) {
      const S av = lane_elem<S>(wa[d], i);
#pragma unroll
      for (int jj = 0; jj < JB; ++jj) {
        if (!used(i, jj)) continue;
        S bv;
        if constexpr (SYM)
          bv = lane_elem<S>(wa[d], jj);
        else
          bv = lane_elem<S>(wb[d], jj);
        acc[i][jj] = mfma_step(av, bv, acc[i][jj]);
      }
    }
  };
  auto fold = [&]() __attribute__((always_inline)) {
    if constexpr (kF32) {
#pragma unroll
      for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int jj = 0; jj < JB; ++jj)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            if (!used(i, jj)) continue;
            float s, e;
            two_sum(hi[i][jj][v], acc[i][jj][v], s, e);
            hi[i][jj][v] = s;
            lo[i][jj][v] += e;
            acc[i][jj][v] = 0.0f;
          }
    }
  };

  if (q > 0) {
    const int64_t last = s0 + q - 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ld(d, s0 + d);
      __builtin_amdgcn_sched_barrier(0);
    }
    // The refill is unconditional (past the last step it re-reads that step,
    // in bounds and discarded), so every step waits only for its own loads;
    // sched_barrier keeps the refills behind the instructions that read the
    // registers they overwrite.
    int since_fold = 0;
    for (int64_t s = 0; s < q; s += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        consume(d);
        __builtin_amdgcn_sched_barrier(0);
        const int64_t nx = s0 + s + d + D;
        ld(d, nx < last ? nx : last);
        __builtin_amdgcn_sched_barrier(0);
      }
      if constexpr (kF32) {
        since_fold += D;
        if (since_fold == FOLD) {
          fold();
          since_fold = 0;
        }
      }
    }
  }
  for (int64_t st = W * q + wid; st < F; st += W) {
    ld(0, st);
    consume(0);
    fold();
  }
  if (wid == W - 1 && a.n > 4 * F) {  // the last n % 4 rows
    const int64_t r = 4 * F + g;
    if (r < a.n) {
      lane_load<LBA, NT>(wa[0], A8 + r * RA + lane_offset<S, LBA>(j));
      if constexpr (!SYM) lane_load<LBB, NT>(wb[0], B8 + r * RB + lane_offset<S, LBB>(j));
    } else {
#pragma unroll
      for (int x = 0; x < LBA / 4; ++x) wa[0][x] = 0u;
      if constexpr (!SYM)
#pragma unroll
        for (int x = 0; x < LBB / 4; ++x) wb[0][x] = 0u;
    }
    consume(0);
  }
  fold();

  // ---- fixed-order sum of the four waves: (w0 + w2) + (w1 + w3) ----------------------
  constexpr int KCMC = KC * MC;
  auto value = [&](int i, int jj, int v) __attribute__((always_inline)) -> double {
    if constexpr (kF32)
      return (double)hi[i][jj][v] + (double)lo[i][jj][v];
    else
      return acc[i][jj][v];
  };
  auto for_each = [&](auto&& f) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
      for (int jj = 0; jj < JB; ++jj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (!used(i, jj)) continue;
          const int r = MfmaTraits<S>::row_major_out ? 4 * g + v : g + 4 * v;
          const int ca = lane_col<S, LBA>(i, r), cb = lane_col<S, LBB>(jj, j);
          f(ca * MC + cb, value(i, jj, v));
          if (SYM && i < jj) f(cb * MC + ca, value(i, jj, v));
        }
  };
  if (wave >= 2) for_each([&](int o, double x) { red[(wave - 2) * KCMC + o] = x; });
  __syncthreads();
  if (wave < 2) for_each([&](int o, double x) { red[wave * KCMC + o] = x + red[wave * KCMC + o]; });
  __syncthreads();
  for (int o = threadIdx.x; o < KCMC; o += kThreads) publish(partial_out + o, red[o] + red[KCMC + o]);
}

// BF16 direct engine.  The 16x16x32 BF16 instruction takes eight rows of one
// column per lane, which row-major loads cannot supply, so each wave copies
// its own 32-row steps (the same contiguous bytes as in memory) into a
// wave-private, double-buffered LDS tile and reads the operands from it with
// the transposing reads of the staged consumer (MfmaAcc::step: paired
// accumulators against the instruction's rounding bias, (hi, lo) folds every
// 128 rows per wave).  Only the wave's own lanes touch its tile, so no
// workgroup barrier stands between the loads and the matrix instructions.
// Steps, work split and final sum as in direct_partial (steps of 32 rows).
template <int KC, int MC, int SYM>
constexpr int direct_bf16_smem() {
  constexpr int staging = 4 * 2 * 32 * (KC + (SYM ? 0 : MC)) * 2;
  constexpr int red = KC * MC * 8;
  constexpr int m1 = staging > red ? staging : red;
  return m1 > kThreads * 16 ? m1 : kThreads * 16;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int KC, int MC, int D, int NT, int SYM = 0>
__device__ __forceinline__ void direct_bf16_partial(const AtbArgs& a, unsigned char* smem,
                                                    double* partial_out) {
  constexpr int KS = 32;                        // rows per step
  constexpr int RA = KC * 2, RB = MC * 2;       // bytes per row
  constexpr int CA = KS * RA / 1024;            // 16-byte chunks per lane per step (A)
  constexpr int CB = SYM ? 0 : KS * RB / 1024;  // (B)
  static_assert(CA >= 1 && (SYM || CB >= 1), "widths of 16 or more");
  constexpr int STAGE = KS * (RA + (SYM ? 0 : RB));
  constexpr int FOLD = 4;                       // steps between (hi, lo) folds: 128 rows
  static_assert(FOLD % D == 0 && D % 2 == 0, "D even and dividing the fold interval");
  MfmaAcc<bf16, KC, MC, 4 * KS, SYM> acc;
  acc.zero();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned char* wbuf = smem + wave * 2 * STAGE;
  const int64_t F = a.n / KS;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t wid = (int64_t)blockIdx.x * 4 + wave;
  const int64_t q = a.tiles_per_wg;
  const int64_t s0 = wid * q;
  gu8* A8 = (gu8*)a.A;
  gu8* B8 = (gu8*)a.B;
  v4u ra[D][CA], rb[D][CB > 0 ? CB : 1];
  auto ld = [&](int d, int64_t st) __attribute__((always_inline)) {
    gv4u* pa = (gv4u*)(A8 + st * KS * RA);
#pragma unroll
    for (int c = 0; c < CA; ++c) ra[d][c] = stream_load<NT>(pa + lane + 64 * c);
    if constexpr (!SYM) {
      gv4u* pb = (gv4u*)(B8 + st * KS * RB);
#pragma unroll
      for (int c = 0; c < CB; ++c) rb[d][c] = stream_load<NT>(pb + lane + 64 * c);
    }
  };
  auto put = [&](int d, int buf) __attribute__((always_inline)) {
    v4u* la = reinterpret_cast<v4u*>(wbuf + buf * STAGE);
#pragma unroll
    for (int c = 0; c < CA; ++c) la[lane + 64 * c] = ra[d][c];
    if constexpr (!SYM) {
      v4u* lb = reinterpret_cast<v4u*>(wbuf + buf * STAGE + KS * RA);
#pragma unroll
      for (int c = 0; c < CB; ++c) lb[lane + 64 * c] = rb[d][c];
    }
  };
  auto compute = [&](int buf) __attribute__((always_inline)) {
    const bf16* As = reinterpret_cast<const bf16*>(wbuf + buf * STAGE);
    const bf16* Bs = SYM ? As : reinterpret_cast<const bf16*>(wbuf + buf * STAGE + KS * RA);
    acc.step(As, Bs, 0, lane);
  };
  if (q > 0) {
    const int64_t last = s0 + q - 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ld(d, s0 + d);
      __builtin_amdgcn_sched_barrier(0);
    }
    // Slot d's step goes to LDS buffer d % 2 (D is even): the buffer written
    // was last read two steps earlier, and one wave's LDS operations run in
    // order.  The refill is unconditional and clamped, as in direct_partial.
    int since_fold = 0;
    for (int64_t s = 0; s < q; s += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        put(d, d & 1);
        wave_sync();
        const int64_t nx = s0 + s + d + D;
        ld(d, nx < last ? nx : last);
        __builtin_amdgcn_sched_barrier(0);
        compute(d & 1);
        __builtin_amdgcn_sched_barrier(0);
      }
      since_fold += D;
      if (since_fold == FOLD) {
        acc.fold();
        since_fold = 0;
      }
    }
  }
  wave_sync();
  for (int64_t st = W * q + wid; st < F; st += W) {  // leftover steps, one at a time
    ld(0, st);
    put(0, 0);
    wave_sync();
    compute(0);
    acc.fold();
    wave_sync();
  }
  if (wid == W - 1 && a.n > KS * F) {  // the last n % 32 rows, zero-padded
#pragma unroll
    for (int c = 0; c < CA; ++c) {
      const int ci = lane + 64 * c;
      const int64_t row = KS * F + ci / (RA / 16);
      ra[0][c] = row < a.n ? stream_load<NT>((gv4u*)(A8 + KS * F * RA) + ci) : v4u{0, 0, 0, 0};
    }
    if constexpr (!SYM) {
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        const int ci = lane + 64 * c;
        const int64_t row = KS * F + ci / (RB / 16);
        rb[0][c] = row < a.n ? stream_load<NT>((gv4u*)(B8 + KS * F * RB) + ci) : v4u{0, 0, 0, 0};
      }
    }
    put(0, 0);
    wave_sync();
    compute(0);
  }
  acc.fold();

  // ---- the four waves' partials in wave order, then publication -----------------------
  constexpr int KCMC = KC * MC;
  double* red = reinterpret_cast<double*>(smem);
  __syncthreads();  // every wave is done with its staging tile
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (wave == w) acc.emit(red, w == 0, lane);
    __syncthreads();
  }
  for (int o = threadIdx.x; o < KCMC; o += kThreads) publish(partial_out + o, red[o]);
}

// Direct-engine kernel body: partial, then the last workgroup's final sum
// (classes up to 512 outputs) as in atb_body.
template <class S, int KC, int MC, int D, int NT, int SYM = 0>
__device__ __forceinline__ void direct_body(const AtbArgs& a) {
  constexpr int KCMC = KC * MC;
  constexpr bool kBf16 = sizeof(S) == 2;
  constexpr int SMEM_F = 2 * KCMC * 8 > kThreads * 16 ? 2 * KCMC * 8 : kThreads * 16;
  constexpr int SMEM = kBf16 ? direct_bf16_smem<KC, MC, SYM>() : SMEM_F;
  __shared__ __attribute__((aligned(16))) unsigned char smem[SMEM];
  __shared__ int s_last;
  if constexpr (kBf16)
    direct_bf16_partial<KC, MC, D, NT, SYM>(a, smem, a.partials + (int64_t)blockIdx.x * KCMC);
  else
    direct_partial<S, KC, MC, D, NT, SYM>(a, reinterpret_cast<double*>(smem),
                                          a.partials + (int64_t)blockIdx.x * KCMC);
  if (!a.finish) return;
  if (!arrive_last(a.counter, gridDim.x, &s_last)) return;
  if constexpr (KCMC >= 2 && KCMC <= 512)
    finish_small<KCMC>(a, (int)gridDim.x, MC, reinterpret_cast<double*>(smem));
  else
    finish_outputs(a, (int)gridDim.x, KCMC, MC, 0, KCMC, reinterpret_cast<double*>(smem));
  if (threadIdx.x == 0) __hip_atomic_store(a.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace tsm
