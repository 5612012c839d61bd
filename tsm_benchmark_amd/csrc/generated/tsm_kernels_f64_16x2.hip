// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_16x2_mfma_r64_p1_nt: KC=16 MC=2 TM=0 TN=0 rows=64 pf=1 nt=1
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x2_mfma_r64_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 2, 0, 0, 64, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x2_mfma_r64_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x2_mfma_r64_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x2_mfma_r64_p2_nt: KC=16 MC=2 TM=0 TN=0 rows=64 pf=2 nt=1
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x2_mfma_r64_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 2, 0, 0, 64, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x2_mfma_r64_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x2_mfma_r64_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x2_mfma_r128_p1_nt: KC=16 MC=2 TM=0 TN=0 rows=128 pf=1 nt=1
//   matrix cores: 32 K-steps of 4 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x2_mfma_r128_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 2, 0, 0, 128, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x2_mfma_r128_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x2_mfma_r128_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x2_mfma_r128_p2_nt: KC=16 MC=2 TM=0 TN=0 rows=128 pf=2 nt=1
//   matrix cores: 32 K-steps of 4 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x2_mfma_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 2, 0, 0, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x2_mfma_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x2_mfma_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x2_mfma_r256_p1_nt: KC=16 MC=2 TM=0 TN=0 rows=256 pf=1 nt=1
//   matrix cores: 64 K-steps of 4 rows per tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x2_mfma_r256_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 2, 0, 0, 256, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x2_mfma_r256_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x2_mfma_r256_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x2_mfma_r256_p2_nt: KC=16 MC=2 TM=0 TN=0 rows=256 pf=2 nt=1
//   matrix cores: 64 K-steps of 4 rows per tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x2_mfma_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 2, 0, 0, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x2_mfma_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x2_mfma_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x2_t8x2_r128_p2_nt: KC=16 MC=2 TM=8 TN=2 rows=128 pf=2 nt=1
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x2_t8x2_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 2, 8, 2, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x2_t8x2_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x2_t8x2_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
