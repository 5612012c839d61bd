// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_bf16_64x4_mfma_r128_p1_nt: KC=64 MC=4 TM=0 TN=0 rows=128 pf=1 nt=1
//   matrix cores: 4 K-steps of 32 rows per tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x4_mfma_r128_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 64, 4, 0, 0, 128, 1, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x4_mfma_r128_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x4_mfma_r128_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_64x4_mfma_r128_p2_nt: KC=64 MC=4 TM=0 TN=0 rows=128 pf=2 nt=1
//   matrix cores: 4 K-steps of 32 rows per tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x4_mfma_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 64, 4, 0, 0, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x4_mfma_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x4_mfma_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_64x4_mfma_r256_p1_nt: KC=64 MC=4 TM=0 TN=0 rows=256 pf=1 nt=1
//   matrix cores: 8 K-steps of 32 rows per tile, LDS 81920 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x4_mfma_r256_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 64, 4, 0, 0, 256, 1, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x4_mfma_r256_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x4_mfma_r256_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_64x4_mfma_r256_p2_nt: KC=64 MC=4 TM=0 TN=0 rows=256 pf=2 nt=1
//   matrix cores: 8 K-steps of 32 rows per tile, LDS 81920 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x4_mfma_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 64, 4, 0, 0, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x4_mfma_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x4_mfma_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_64x4_t8x4_r128_p2_nt: KC=64 MC=4 TM=8 TN=4 rows=128 pf=2 nt=1
//   8 register tiles x 32 row lanes, 4 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x4_t8x4_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 64, 4, 8, 4, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x4_t8x4_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x4_t8x4_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
