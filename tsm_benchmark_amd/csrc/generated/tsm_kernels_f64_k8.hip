// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_8x1_t8x1_r256_p2: KC=8 MC=1 TM=8 TN=1 rows=256 pf=2
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x1_t8x1_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 1, 8, 1, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x1_t8x1_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x1_t8x1_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x2_t8x2_r256_p2: KC=8 MC=2 TM=8 TN=2 rows=256 pf=2
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x2_t8x2_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 2, 8, 2, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x2_t8x2_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x2_t8x2_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x4_t8x4_r256_p2: KC=8 MC=4 TM=8 TN=4 rows=256 pf=2
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x4_t8x4_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 4, 8, 4, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x4_t8x4_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x4_t8x4_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t2x2_r256_p1: KC=8 MC=8 TM=2 TN=2 rows=256 pf=1
//   16 register tiles x 16 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t2x2_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 2, 2, 256, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t2x2_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t2x2_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t2x2_r256_p2: KC=8 MC=8 TM=2 TN=2 rows=256 pf=2
//   16 register tiles x 16 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t2x2_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 2, 2, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t2x2_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t2x2_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r128_p1: KC=8 MC=8 TM=4 TN=4 rows=128 pf=1
//   4 register tiles x 64 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r128_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 128, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r128_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r128_p2: KC=8 MC=8 TM=4 TN=4 rows=128 pf=2
//   4 register tiles x 64 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 128, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r256_p1: KC=8 MC=8 TM=4 TN=4 rows=256 pf=1
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 256, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r256_p2: KC=8 MC=8 TM=4 TN=4 rows=256 pf=2
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r256_p1: KC=8 MC=8 TM=8 TN=8 rows=256 pf=1
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 256, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r256_p2: KC=8 MC=8 TM=8 TN=8 rows=256 pf=2
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x16_t8x8_r128_p2: KC=8 MC=16 TM=8 TN=8 rows=128 pf=2
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x16_t8x8_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 16, 8, 8, 128, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x16_t8x8_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x16_t8x8_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x32_t8x8_r64_p2: KC=8 MC=32 TM=8 TN=8 rows=64 pf=2
//   4 register tiles x 64 row lanes, 1 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x32_t8x8_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 32, 8, 8, 64, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x32_t8x8_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x32_t8x8_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x64_t8x8_r32_p2: KC=8 MC=64 TM=8 TN=8 rows=32 pf=2
//   8 register tiles x 32 row lanes, 1 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x64_t8x8_r32_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 64, 8, 8, 32, 2>(a);
}
hipError_t launch_tsm_atb_f64_8x64_t8x8_r32_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x64_t8x8_r32_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
