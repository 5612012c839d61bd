// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f32_64x1_t8x1_r64_p2: KC=64 MC=1 TM=8 TN=1 rows=64 pf=2
//   8 register tiles x 32 row lanes, 2 rows/thread/tile, LDS 33280 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x1_t8x1_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 1, 8, 1, 64, 2>(a);
}
hipError_t launch_tsm_atb_f32_64x1_t8x1_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x1_t8x1_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x2_t8x2_r64_p2: KC=64 MC=2 TM=8 TN=2 rows=64 pf=2
//   8 register tiles x 32 row lanes, 2 rows/thread/tile, LDS 33792 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x2_t8x2_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 2, 8, 2, 64, 2>(a);
}
hipError_t launch_tsm_atb_f32_64x2_t8x2_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x2_t8x2_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x4_t8x4_r64_p2: KC=64 MC=4 TM=8 TN=4 rows=64 pf=2
//   8 register tiles x 32 row lanes, 2 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x4_t8x4_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 4, 8, 4, 64, 2>(a);
}
hipError_t launch_tsm_atb_f32_64x4_t8x4_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x4_t8x4_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x8_t8x4_r64_p2: KC=64 MC=8 TM=8 TN=4 rows=64 pf=2
//   16 register tiles x 16 row lanes, 4 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x8_t8x4_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 8, 8, 4, 64, 2>(a);
}
hipError_t launch_tsm_atb_f32_64x8_t8x4_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x8_t8x4_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x16_t8x4_r64_p2: KC=64 MC=16 TM=8 TN=4 rows=64 pf=2
//   32 register tiles x 8 row lanes, 8 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x16_t8x4_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 16, 8, 4, 64, 2>(a);
}
hipError_t launch_tsm_atb_f32_64x16_t8x4_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x16_t8x4_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x32_t8x4_r64_p2: KC=64 MC=32 TM=8 TN=4 rows=64 pf=2
//   64 register tiles x 4 row lanes, 16 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x32_t8x4_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 32, 8, 4, 64, 2>(a);
}
hipError_t launch_tsm_atb_f32_64x32_t8x4_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x32_t8x4_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r64_p2: KC=64 MC=64 TM=8 TN=8 rows=64 pf=2
//   64 register tiles x 4 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 64, 2>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
