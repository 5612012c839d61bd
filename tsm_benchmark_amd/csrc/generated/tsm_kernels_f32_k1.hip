// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f32_1x1_t1x1_r4096_p2: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=2
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_1x1_t1x1_r4096_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 1, 1, 1, 1, 4096, 2>(a);
}
hipError_t launch_tsm_atb_f32_1x1_t1x1_r4096_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_1x1_t1x1_r4096_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_1x2_t1x2_r2048_p2: KC=1 MC=2 TM=1 TN=2 rows=2048 pf=2
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_1x2_t1x2_r2048_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 1, 2, 1, 2, 2048, 2>(a);
}
hipError_t launch_tsm_atb_f32_1x2_t1x2_r2048_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_1x2_t1x2_r2048_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_1x4_t1x4_r1024_p2: KC=1 MC=4 TM=1 TN=4 rows=1024 pf=2
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_1x4_t1x4_r1024_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 1, 4, 1, 4, 1024, 2>(a);
}
hipError_t launch_tsm_atb_f32_1x4_t1x4_r1024_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_1x4_t1x4_r1024_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_1x8_t1x4_r512_p2: KC=1 MC=8 TM=1 TN=4 rows=512 pf=2
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_1x8_t1x4_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 1, 8, 1, 4, 512, 2>(a);
}
hipError_t launch_tsm_atb_f32_1x8_t1x4_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_1x8_t1x4_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_1x16_t1x4_r256_p2: KC=1 MC=16 TM=1 TN=4 rows=256 pf=2
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_1x16_t1x4_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 1, 16, 1, 4, 256, 2>(a);
}
hipError_t launch_tsm_atb_f32_1x16_t1x4_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_1x16_t1x4_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_1x32_t1x4_r128_p2: KC=1 MC=32 TM=1 TN=4 rows=128 pf=2
//   8 register tiles x 32 row lanes, 4 rows/thread/tile, LDS 33792 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_1x32_t1x4_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 1, 32, 1, 4, 128, 2>(a);
}
hipError_t launch_tsm_atb_f32_1x32_t1x4_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_1x32_t1x4_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_1x64_t1x4_r64_p2: KC=1 MC=64 TM=1 TN=4 rows=64 pf=2
//   16 register tiles x 16 row lanes, 4 rows/thread/tile, LDS 33280 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_1x64_t1x4_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 1, 64, 1, 4, 64, 2>(a);
}
hipError_t launch_tsm_atb_f32_1x64_t1x4_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_1x64_t1x4_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
