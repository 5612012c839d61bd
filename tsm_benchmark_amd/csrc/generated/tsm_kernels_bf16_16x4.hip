// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_bf16_16x4_mfma_r256_p1_nt: KC=16 MC=4 TM=0 TN=0 rows=256 pf=1 nt=1
//   matrix cores: 8 K-steps of 32 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x4_mfma_r256_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 16, 4, 0, 0, 256, 1, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x4_mfma_r256_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x4_mfma_r256_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x4_mfma_r256_p2_nt: KC=16 MC=4 TM=0 TN=0 rows=256 pf=2 nt=1
//   matrix cores: 8 K-steps of 32 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x4_mfma_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 16, 4, 0, 0, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x4_mfma_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x4_mfma_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x4_mfma_r512_p1_nt: KC=16 MC=4 TM=0 TN=0 rows=512 pf=1 nt=1
//   matrix cores: 16 K-steps of 32 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x4_mfma_r512_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 16, 4, 0, 0, 512, 1, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x4_mfma_r512_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x4_mfma_r512_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x4_mfma_r512_p2_nt: KC=16 MC=4 TM=0 TN=0 rows=512 pf=2 nt=1
//   matrix cores: 16 K-steps of 32 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x4_mfma_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 16, 4, 0, 0, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x4_mfma_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x4_mfma_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x4_mfma_r1024_p1_nt: KC=16 MC=4 TM=0 TN=0 rows=1024 pf=1 nt=1
//   matrix cores: 32 K-steps of 32 rows per tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x4_mfma_r1024_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 16, 4, 0, 0, 1024, 1, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x4_mfma_r1024_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x4_mfma_r1024_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x4_mfma_r1024_p2_nt: KC=16 MC=4 TM=0 TN=0 rows=1024 pf=2 nt=1
//   matrix cores: 32 K-steps of 32 rows per tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x4_mfma_r1024_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 16, 4, 0, 0, 1024, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x4_mfma_r1024_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x4_mfma_r1024_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x4_t8x4_r512_p2_nt: KC=16 MC=4 TM=8 TN=4 rows=512 pf=2 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x4_t8x4_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 16, 4, 8, 4, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x4_t8x4_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x4_t8x4_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
