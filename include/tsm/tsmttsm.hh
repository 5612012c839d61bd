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
// size, stride, work-group size, unroll) describe a SYCL work decomposition
// that has no counterpart on gfx950: every name here runs the library's tuned
// kernel and persistent grid for its width class (the same launch as
// tsm_atb), so a reference name is as fast as the library entry point.  The
// constants are kept for source compatibility and recorded by the generated
// layer (kGeneratedKernels); a caller who wants a particular kernel variant
// passes a tsm_kernel_config to tsm_atb_ex (or hinted_call below).
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

// C += A^T B with the tuned kernel and grid of the (M, N) width class.
template <typename T>
inline tsm_status tuned_call(tsm_stream_t q, int64_t K, int M, int N, const T* A, const T* B,
                             T* C) {
  return tsm_atb(dtype_of<T>::in, dtype_of<T>::out, K, M, N, A, M, B, N, C, N, 1, nullptr, 0, q);
}

// C += A^T B with a requested register tile / pipeline depth / grid (explicit
// opt-in; the reference names do not use it).
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

// Entry used by the generated name layer (include/tsm/tsmttsm_kernels.hh):
// the tuned launch; the reference's TM, TN, leap-frog depth, stride and
// unroll of the name are recorded in kGeneratedKernels, not applied.
template <typename T>
inline tsm_status generated_call(tsm_stream_t q, int64_t K, int M, int N, const T* A, const T* B,
                                 T* C) {
  return tuned_call<T>(q, K, M, N, A, B, C);
}

constexpr int tile_of(int max_tile, int M) { return max_tile < M ? max_tile : M; }

// ---- hand-written variants (square: A, B are K x M, C is M x M) -----------------
// The reference's variants differ only in how SYCL work-items share rows and
// tiles; all of them compute C += A^T B.  Each name below keeps its
// reference template parameters and runs the tuned launch.
// v1 (tsmttsm.hh:15-35): full M x M accumulator per work-item.
template <typename T, int M, int stride>
inline tsm_status tsmttsm(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  static_assert(M >= 1 && M <= TSM_MAX_WIDTH, "M in [1, 64]");
  return tuned_call<T>(q, K, M, M, A, B, C);
}
// v2, v3 (tsmttsm.hh:40-122): contiguous tiles of max_tile x max_tile.
template <typename T, int M, int stride, int max_tile_size>
inline tsm_status tsmttsm2(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  static_assert(M >= 1 && M <= TSM_MAX_WIDTH && max_tile_size >= 1, "M in [1, 64]");
  return tuned_call<T>(q, K, M, M, A, B, C);
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
// v5, v7 (tsmttsm.hh:170-225, 303-363): with register prefetch.
template <typename T, int M, int stride, int max_tile_size>
inline tsm_status tsmttsm5(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsmttsm2<T, M, stride, max_tile_size>(q, K, A, B, C);
}
template <typename T, int M, int stride, int max_tile_size>
inline tsm_status tsmttsm7(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsmttsm2<T, M, stride, max_tile_size>(q, K, A, B, C);
}
// v6, v9 (tsmttsm.hh:230-298, 368-439): nd_range with a work-group reduction.
template <typename T, int M, int global_size, int local_size, int max_tile_size>
inline tsm_status tsmttsm6(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  static_assert(local_size >= 1 && global_size % local_size == 0,
                "global size is a whole number of work-groups");
  return tsmttsm2<T, M, global_size, max_tile_size>(q, K, A, B, C);
}
template <typename T, int M, int global_size, int local_size, int max_tile_size>
inline tsm_status tsmttsm9(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsmttsm6<T, M, global_size, local_size, max_tile_size>(q, K, A, B, C);
}
// v8 (tsmttsm.hh:445-586): adaptive tile, grid from the compute-unit count.
template <typename T, int M, int worker_mult = 64>
inline tsm_status tsmttsm8(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsmttsm<T, M, 0>(q, K, A, B, C);
}

}  // namespace tsm_compat
