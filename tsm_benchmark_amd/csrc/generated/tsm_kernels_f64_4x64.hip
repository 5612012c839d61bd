// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_4x64_mfma_r16_p1_nt: KC=4 MC=64 TM=0 TN=0 rows=16 pf=1 nt=1
//   matrix cores: 4 K-steps of 4 rows per tile, LDS 20480 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x64_mfma_r16_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 64, 0, 0, 16, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x64_mfma_r16_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x64_mfma_r16_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x64_mfma_r16_p2_nt: KC=4 MC=64 TM=0 TN=0 rows=16 pf=2 nt=1
//   matrix cores: 4 K-steps of 4 rows per tile, LDS 20480 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x64_mfma_r16_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 64, 0, 0, 16, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x64_mfma_r16_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x64_mfma_r16_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x64_mfma_r32_p1_nt: KC=4 MC=64 TM=0 TN=0 rows=32 pf=1 nt=1
//   matrix cores: 8 K-steps of 4 rows per tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x64_mfma_r32_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 64, 0, 0, 32, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x64_mfma_r32_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x64_mfma_r32_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x64_mfma_r32_p2_nt: KC=4 MC=64 TM=0 TN=0 rows=32 pf=2 nt=1
//   matrix cores: 8 K-steps of 4 rows per tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x64_mfma_r32_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 64, 0, 0, 32, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x64_mfma_r32_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x64_mfma_r32_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x64_mfma_r64_p1_nt: KC=4 MC=64 TM=0 TN=0 rows=64 pf=1 nt=1
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 81920 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x64_mfma_r64_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 64, 0, 0, 64, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x64_mfma_r64_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x64_mfma_r64_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x64_mfma_r64_p2_nt: KC=4 MC=64 TM=0 TN=0 rows=64 pf=2 nt=1
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 81920 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x64_mfma_r64_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 64, 0, 0, 64, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x64_mfma_r64_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x64_mfma_r64_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x64_t4x8_r32_p2_nt: KC=4 MC=64 TM=4 TN=8 rows=32 pf=2 nt=1
//   8 register tiles x 32 row lanes, 1 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x64_t4x8_r32_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 64, 4, 8, 32, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x64_t4x8_r32_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x64_t4x8_r32_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
