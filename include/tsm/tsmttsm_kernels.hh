#pragma once
// GENERATED by tsm_benchmark_amd/codegen.py --name-layer
// M=[4, 8, 16] tile-sizes=[2, 4] stride=524288 unroll=1
#include "tsm/tsmttsm.hh"

namespace tsm_compat {

template <typename T>
inline tsm_status tsmttsm_4_4_2_2(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 4, 4, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_4_4_2_2_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 4, 4, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_4_4_2_2_sg(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 4, 4, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_4_4_2_2_sg_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 4, 4, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_4_4_4_4(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 4, 4, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_4_4_4_4_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 4, 4, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_4_4_4_4_sg(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 4, 4, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_4_4_4_4_sg_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 4, 4, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_8_8_2_2(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 8, 8, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_8_8_2_2_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 8, 8, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_8_8_2_2_sg(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 8, 8, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_8_8_2_2_sg_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 8, 8, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_8_8_4_4(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 8, 8, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_8_8_4_4_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 8, 8, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_8_8_4_4_sg(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 8, 8, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_8_8_4_4_sg_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 8, 8, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_16_16_2_2(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 16, 16, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_16_16_2_2_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 16, 16, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_16_16_2_2_sg(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 16, 16, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_16_16_2_2_sg_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 16, 16, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_16_16_4_4(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 16, 16, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_16_16_4_4_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 16, 16, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_16_16_4_4_sg(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 16, 16, A, B, C);
}

template <typename T>
inline tsm_status tsmttsm_16_16_4_4_sg_lf(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {
  return tsm_compat::generated_call<T>(q, K, 16, 16, A, B, C);
}

// Every generated name with its FP64 instance, in generation order (the
// generated benchmark's list, sycl_codegen_new.py:478-487).
// The reference's decomposition of each name is recorded (tile TM x TM,
// leap-frog prefetch, sub-group reduction, stride, unroll).
struct GeneratedKernel {
  const char* name;
  int M, TM, leap_frog, sub_group;
  long long stride;
  int unroll;
  tsm_status (*fn)(tsm_stream_t, int64_t, const double*, const double*, double*);
};
inline const GeneratedKernel kGeneratedKernels[] = {
  {"tsmttsm_4_4_2_2", 4, 2, 0, 0, 524288, 1, &tsmttsm_4_4_2_2<double>},
  {"tsmttsm_4_4_2_2_lf", 4, 2, 1, 0, 524288, 1, &tsmttsm_4_4_2_2_lf<double>},
  {"tsmttsm_4_4_2_2_sg", 4, 2, 0, 1, 524288, 1, &tsmttsm_4_4_2_2_sg<double>},
  {"tsmttsm_4_4_2_2_sg_lf", 4, 2, 1, 1, 524288, 1, &tsmttsm_4_4_2_2_sg_lf<double>},
  {"tsmttsm_4_4_4_4", 4, 4, 0, 0, 524288, 1, &tsmttsm_4_4_4_4<double>},
  {"tsmttsm_4_4_4_4_lf", 4, 4, 1, 0, 524288, 1, &tsmttsm_4_4_4_4_lf<double>},
  {"tsmttsm_4_4_4_4_sg", 4, 4, 0, 1, 524288, 1, &tsmttsm_4_4_4_4_sg<double>},
  {"tsmttsm_4_4_4_4_sg_lf", 4, 4, 1, 1, 524288, 1, &tsmttsm_4_4_4_4_sg_lf<double>},
  {"tsmttsm_8_8_2_2", 8, 2, 0, 0, 524288, 1, &tsmttsm_8_8_2_2<double>},
  {"tsmttsm_8_8_2_2_lf", 8, 2, 1, 0, 524288, 1, &tsmttsm_8_8_2_2_lf<double>},
  {"tsmttsm_8_8_2_2_sg", 8, 2, 0, 1, 524288, 1, &tsmttsm_8_8_2_2_sg<double>},
  {"tsmttsm_8_8_2_2_sg_lf", 8, 2, 1, 1, 524288, 1, &tsmttsm_8_8_2_2_sg_lf<double>},
  {"tsmttsm_8_8_4_4", 8, 4, 0, 0, 524288, 1, &tsmttsm_8_8_4_4<double>},
  {"tsmttsm_8_8_4_4_lf", 8, 4, 1, 0, 524288, 1, &tsmttsm_8_8_4_4_lf<double>},
  {"tsmttsm_8_8_4_4_sg", 8, 4, 0, 1, 524288, 1, &tsmttsm_8_8_4_4_sg<double>},
  {"tsmttsm_8_8_4_4_sg_lf", 8, 4, 1, 1, 524288, 1, &tsmttsm_8_8_4_4_sg_lf<double>},
  {"tsmttsm_16_16_2_2", 16, 2, 0, 0, 524288, 1, &tsmttsm_16_16_2_2<double>},
  {"tsmttsm_16_16_2_2_lf", 16, 2, 1, 0, 524288, 1, &tsmttsm_16_16_2_2_lf<double>},
  {"tsmttsm_16_16_2_2_sg", 16, 2, 0, 1, 524288, 1, &tsmttsm_16_16_2_2_sg<double>},
  {"tsmttsm_16_16_2_2_sg_lf", 16, 2, 1, 1, 524288, 1, &tsmttsm_16_16_2_2_sg_lf<double>},
  {"tsmttsm_16_16_4_4", 16, 4, 0, 0, 524288, 1, &tsmttsm_16_16_4_4<double>},
  {"tsmttsm_16_16_4_4_lf", 16, 4, 1, 0, 524288, 1, &tsmttsm_16_16_4_4_lf<double>},
  {"tsmttsm_16_16_4_4_sg", 16, 4, 0, 1, 524288, 1, &tsmttsm_16_16_4_4_sg<double>},
  {"tsmttsm_16_16_4_4_sg_lf", 16, 4, 1, 1, 524288, 1, &tsmttsm_16_16_4_4_sg_lf<double>},
};
inline const char* const kGeneratedNames[] = {
  "tsmttsm_4_4_2_2",
  "tsmttsm_4_4_2_2_lf",
  "tsmttsm_4_4_2_2_sg",
  "tsmttsm_4_4_2_2_sg_lf",
  "tsmttsm_4_4_4_4",
  "tsmttsm_4_4_4_4_lf",
  "tsmttsm_4_4_4_4_sg",
  "tsmttsm_4_4_4_4_sg_lf",
  "tsmttsm_8_8_2_2",
  "tsmttsm_8_8_2_2_lf",
  "tsmttsm_8_8_2_2_sg",
  "tsmttsm_8_8_2_2_sg_lf",
  "tsmttsm_8_8_4_4",
  "tsmttsm_8_8_4_4_lf",
  "tsmttsm_8_8_4_4_sg",
  "tsmttsm_8_8_4_4_sg_lf",
  "tsmttsm_16_16_2_2",
  "tsmttsm_16_16_2_2_lf",
  "tsmttsm_16_16_2_2_sg",
  "tsmttsm_16_16_2_2_sg_lf",
  "tsmttsm_16_16_4_4",
  "tsmttsm_16_16_4_4_lf",
  "tsmttsm_16_16_4_4_sg",
  "tsmttsm_16_16_4_4_sg_lf",
};

}  // namespace tsm_compat
