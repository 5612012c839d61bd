// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_4x1_t4x1_r512_p2: KC=4 MC=1 TM=4 TN=1 rows=512 pf=2
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x1_t4x1_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 1, 4, 1, 512, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x1_t4x1_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x1_t4x1_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x2_t4x2_r512_p2: KC=4 MC=2 TM=4 TN=2 rows=512 pf=2
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x2_t4x2_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 2, 4, 2, 512, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x2_t4x2_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x2_t4x2_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x4_t2x2_r512_p1: KC=4 MC=4 TM=2 TN=2 rows=512 pf=1
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x4_t2x2_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 4, 2, 2, 512, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x4_t2x2_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x4_t2x2_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x4_t2x2_r512_p2: KC=4 MC=4 TM=2 TN=2 rows=512 pf=2
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x4_t2x2_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 4, 2, 2, 512, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x4_t2x2_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x4_t2x2_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x4_t4x4_r256_p1: KC=4 MC=4 TM=4 TN=4 rows=256 pf=1
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x4_t4x4_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 4, 4, 4, 256, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x4_t4x4_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x4_t4x4_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x4_t4x4_r256_p2: KC=4 MC=4 TM=4 TN=4 rows=256 pf=2
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x4_t4x4_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 4, 4, 4, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x4_t4x4_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x4_t4x4_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x4_t4x4_r512_p1: KC=4 MC=4 TM=4 TN=4 rows=512 pf=1
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x4_t4x4_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 4, 4, 4, 512, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x4_t4x4_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x4_t4x4_r512_p2: KC=4 MC=4 TM=4 TN=4 rows=512 pf=2
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x4_t4x4_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 4, 4, 4, 512, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x4_t4x4_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x8_t4x8_r256_p2: KC=4 MC=8 TM=4 TN=8 rows=256 pf=2
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x8_t4x8_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 8, 4, 8, 256, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x8_t4x8_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x8_t4x8_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x16_t4x8_r128_p2: KC=4 MC=16 TM=4 TN=8 rows=128 pf=2
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x16_t4x8_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 16, 4, 8, 128, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x16_t4x8_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x16_t4x8_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x32_t4x4_r32_p1: KC=4 MC=32 TM=4 TN=4 rows=32 pf=1
//   8 register tiles x 32 row lanes, 1 rows/thread/tile, LDS 18432 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x32_t4x4_r32_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 32, 4, 4, 32, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x32_t4x4_r32_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x32_t4x4_r32_p2: KC=4 MC=32 TM=4 TN=4 rows=32 pf=2
//   8 register tiles x 32 row lanes, 1 rows/thread/tile, LDS 18432 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x32_t4x4_r32_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 32, 4, 4, 32, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x32_t4x4_r32_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x32_t4x4_r64_p1: KC=4 MC=32 TM=4 TN=4 rows=64 pf=1
//   8 register tiles x 32 row lanes, 2 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x32_t4x4_r64_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 32, 4, 4, 64, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x32_t4x4_r64_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x32_t4x4_r64_p2: KC=4 MC=32 TM=4 TN=4 rows=64 pf=2
//   8 register tiles x 32 row lanes, 2 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x32_t4x4_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 32, 4, 4, 64, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x32_t4x4_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x32_t4x8_r64_p1: KC=4 MC=32 TM=4 TN=8 rows=64 pf=1
//   4 register tiles x 64 row lanes, 1 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x32_t4x8_r64_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 32, 4, 8, 64, 1>(a);
}
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x32_t4x8_r64_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x32_t4x8_r64_p2: KC=4 MC=32 TM=4 TN=8 rows=64 pf=2
//   4 register tiles x 64 row lanes, 1 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x32_t4x8_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 32, 4, 8, 64, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x32_t4x8_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_4x64_t4x8_r32_p2: KC=4 MC=64 TM=4 TN=8 rows=32 pf=2
//   8 register tiles x 32 row lanes, 1 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_4x64_t4x8_r32_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 4, 64, 4, 8, 32, 2>(a);
}
hipError_t launch_tsm_atb_f64_4x64_t4x8_r32_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_4x64_t4x8_r32_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
