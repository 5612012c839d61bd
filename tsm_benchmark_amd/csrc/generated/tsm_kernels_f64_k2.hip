// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_2x1_t2x1_r1024_p2: KC=2 MC=1 TM=2 TN=1 rows=1024 pf=2
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x1_t2x1_r1024_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 1, 2, 1, 1024, 2>(a);
}
hipError_t launch_tsm_atb_f64_2x1_t2x1_r1024_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x1_t2x1_r1024_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x2_t2x2_r512_p1: KC=2 MC=2 TM=2 TN=2 rows=512 pf=1
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x2_t2x2_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 2, 2, 2, 512, 1>(a);
}
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x2_t2x2_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x2_t2x2_r512_p2: KC=2 MC=2 TM=2 TN=2 rows=512 pf=2
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x2_t2x2_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 2, 2, 2, 512, 2>(a);
}
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x2_t2x2_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x2_t2x2_r1024_p1: KC=2 MC=2 TM=2 TN=2 rows=1024 pf=1
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x2_t2x2_r1024_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 2, 2, 2, 1024, 1>(a);
}
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x2_t2x2_r1024_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x2_t2x2_r1024_p2: KC=2 MC=2 TM=2 TN=2 rows=1024 pf=2
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x2_t2x2_r1024_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 2, 2, 2, 1024, 2>(a);
}
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x2_t2x2_r1024_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x4_t2x4_r512_p2: KC=2 MC=4 TM=2 TN=4 rows=512 pf=2
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x4_t2x4_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 4, 2, 4, 512, 2>(a);
}
hipError_t launch_tsm_atb_f64_2x4_t2x4_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x4_t2x4_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x8_t2x8_r256_p2: KC=2 MC=8 TM=2 TN=8 rows=256 pf=2
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x8_t2x8_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 8, 2, 8, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_2x8_t2x8_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x8_t2x8_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x16_t2x8_r128_p2: KC=2 MC=16 TM=2 TN=8 rows=128 pf=2
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x16_t2x8_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 16, 2, 8, 128, 2>(a);
}
hipError_t launch_tsm_atb_f64_2x16_t2x8_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x16_t2x8_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x32_t2x8_r64_p2: KC=2 MC=32 TM=2 TN=8 rows=64 pf=2
//   4 register tiles x 64 row lanes, 1 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x32_t2x8_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 32, 2, 8, 64, 2>(a);
}
hipError_t launch_tsm_atb_f64_2x32_t2x8_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x32_t2x8_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_2x64_t2x8_r32_p2: KC=2 MC=64 TM=2 TN=8 rows=32 pf=2
//   8 register tiles x 32 row lanes, 1 rows/thread/tile, LDS 33792 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_2x64_t2x8_r32_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 2, 64, 2, 8, 32, 2>(a);
}
hipError_t launch_tsm_atb_f64_2x64_t2x8_r32_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_2x64_t2x8_r32_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
