// Reference-name C++ layer over the libtsm C ABI (header only).
//
// The reference exposes its kernels as templates that take a sycl::queue and
// return a sycl::event: the hand-written tsmttsm, tsmttsm2..9 (tsmttsm.hh:15-586)
// and the generated tsmttsm_{M}_{N}_{TM}_{TN}[_lf|_sg|_sg_lf]
// (sycl_codegen_new.py:158-159, 263-264).  Here the same names take a HIP
// stream (tsm_stream_t) and return a tsm_status; the call is asynchronous on
// that stream and keeps the reference contract C += A^T B (the caller zeroes C,
// main.cc:67-69).  A and B are K x M row-major (K x N for B in the generated
// names), C is M x N row-major.
//
// The template constants that the reference bakes into each kernel (tile
// size, stride, work-group size) are tuning hints here: the requested
// register tile and depth are used when the library has that kernel, and the
// tuned default is used otherwise.  `stride` (the reference's global work-item
// count) maps to the persistent grid: stride / 256 workgroups.
#pragma once

#include <cstdint>

#include "tsm/tsm.h"

namespace tsm_compat {

template <typename T>
struct dtype_of;
template <>
struct dtype_of<double> {
  static constexpr tsm_dtype in = TSM_F64, out = TSM_F64;
};
template <>
struct dtype_of<float> {
  static constexpr tsm_dtype in = TSM_F32, out = TSM_F32;
};

// C += A^T B with a requested register tile / pipeline depth / grid.
template <typename T>
inline tsm_status hinted_call(tsm_stream_t q, int64_t K, int M, int N, const T* A, const T* B,
                              T* C, int tm, int tn, int pf, int64_t grid) {
  tsm_kernel_config cfg{tm, tn, 0, pf, 0, grid > 0 ? grid : 0, 0, 0, 0};
  tsm_status s = tsm_atb_ex(dtype_of<T>::in, dtype_of<T>::out, K, M, N, A, M, B, N, C, N, 1,
                            nullptr, 0, q, &cfg);
  if (s == TSM_ERR_NO_KERNEL) {  // hint not compiled in: tuned kernel, same grid
    tsm_kernel_config dflt{0, 0, 0, 0, 0, cfg.grid, 0, 0, 0};
    s = tsm_atb_ex(dtype_of<T>::in, dtype_of<T>::out, K, M, N, A, M, B, N, C, N, 1, nullptr, 0,
                   q, &dflt);
  }
  return s;
}

// Entry used by the generated name layer (include/tsm/tsmttsm_kernels.hh).
template <typename T>
inline tsm_status generated_call(tsm_stream_t q, int64_t K, int M, int N, const T* A, const T* B,
                                 T* C, int TM, int TN, int pf, int64_t stride, int /*unroll*/) {
  return hinted_call<T>(q, K, M, N, A, B, C, TM, TN, pf, stride / 256);
}

constexpr int tile_of(int max_tile, int M) { return max_tile < M ? max_tile : M; }

// ---- hand-written variants (square: A, B are K x M, C is M x M) -----------------
// v1 (tsmttsm.hh:15-35): full M x M accumulator per work-item.
template <typename T, int M, int stride>
inline tsm_status tsmttsm(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return hinted_call<T>(q, K, M, M, A, B, C, tile_of(8, M), tile_of(8, M), 1, stride / 256);
}
// v2, v3 (tsmttsm.hh:40-122): contiguous tiles of max_tile x max_tile.
template <typename T, int M, int stride, int max_tile_size>
inline tsm_status tsmttsm2(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  constexpr int t = tile_of(max_tile_size, M);
  return hinted_call<T>(q, K, M, M, A, B, C, t, t, 1, stride / 256);
}
template <typename T, int M, int stride, int max_tile_size>
inline tsm_status tsmttsm3(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsmttsm2<T, M, stride, max_tile_size>(q, K, A, B, C);
}
// v4 (tsmttsm.hh:125-166): "transposed" tiles (coalesced loads across items).
template <typename T, int M, int stride, int max_tile_size>
inline tsm_status tsmttsm4(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsmttsm2<T, M, stride, max_tile_size>(q, K, A, B, C);
}
// v5, v7 (tsmttsm.hh:170-225, 303-363): with register prefetch (depth 2).
template <typename T, int M, int stride, int max_tile_size>
inline tsm_status tsmttsm5(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  constexpr int t = tile_of(max_tile_size, M);
  return hinted_call<T>(q, K, M, M, A, B, C, t, t, 2, stride / 256);
}
template <typename T, int M, int stride, int max_tile_size>
inline tsm_status tsmttsm7(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsmttsm5<T, M, stride, max_tile_size>(q, K, A, B, C);
}
// v6, v9 (tsmttsm.hh:230-298, 368-439): nd_range with a work-group reduction.
template <typename T, int M, int global_size, int local_size, int max_tile_size>
inline tsm_status tsmttsm6(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  constexpr int t = tile_of(max_tile_size, M);
  return hinted_call<T>(q, K, M, M, A, B, C, t, t, 1, global_size / 256);
}
template <typename T, int M, int global_size, int local_size, int max_tile_size>
inline tsm_status tsmttsm9(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsmttsm6<T, M, global_size, local_size, max_tile_size>(q, K, A, B, C);
}
// v8 (tsmttsm.hh:445-586): adaptive tile, grid from the compute-unit count.
template <typename T, int M, int worker_mult = 64>
inline tsm_status tsmttsm8(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return hinted_call<T>(q, K, M, M, A, B, C, 0, 0, 0, 0);
}

}  // namespace tsm_compat
