// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f32_8x8_t4x4_r256_p1: KC=8 MC=8 TM=4 TN=4 rows=256 pf=1 nt=0
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r256_p2: KC=8 MC=8 TM=4 TN=4 rows=256 pf=2 nt=0
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r256_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=256 pf=2 nt=1
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r256_p4_nt: KC=8 MC=8 TM=4 TN=4 rows=256 pf=4 nt=1
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r256_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 256, 4, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r256_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r256_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r512_p1: KC=8 MC=8 TM=4 TN=4 rows=512 pf=1 nt=0
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 512, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r512_p2: KC=8 MC=8 TM=4 TN=4 rows=512 pf=2 nt=0
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 512, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r512_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=512 pf=2 nt=1
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r512_p4_nt: KC=8 MC=8 TM=4 TN=4 rows=512 pf=4 nt=1
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r512_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 512, 4, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r512_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r1024_p1: KC=8 MC=8 TM=4 TN=4 rows=1024 pf=1 nt=0
//   4 register tiles x 64 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r1024_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 1024, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r1024_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r1024_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r1024_p2: KC=8 MC=8 TM=4 TN=4 rows=1024 pf=2 nt=0
//   4 register tiles x 64 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r1024_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 1024, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r1024_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r1024_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x4_r1024_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=1024 pf=2 nt=1
//   4 register tiles x 64 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x4_r1024_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 4, 1024, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x4_r1024_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x4_r1024_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r256_p1: KC=8 MC=8 TM=4 TN=8 rows=256 pf=1 nt=0
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r256_p2: KC=8 MC=8 TM=4 TN=8 rows=256 pf=2 nt=0
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r256_p2_nt: KC=8 MC=8 TM=4 TN=8 rows=256 pf=2 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r256_p4_nt: KC=8 MC=8 TM=4 TN=8 rows=256 pf=4 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r256_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 256, 4, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r256_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r256_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r512_p1: KC=8 MC=8 TM=4 TN=8 rows=512 pf=1 nt=0
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 512, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r512_p2: KC=8 MC=8 TM=4 TN=8 rows=512 pf=2 nt=0
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 512, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r512_p2_nt: KC=8 MC=8 TM=4 TN=8 rows=512 pf=2 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r512_p4_nt: KC=8 MC=8 TM=4 TN=8 rows=512 pf=4 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r512_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 512, 4, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r512_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r512_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r1024_p1: KC=8 MC=8 TM=4 TN=8 rows=1024 pf=1 nt=0
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r1024_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 1024, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r1024_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r1024_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r1024_p2: KC=8 MC=8 TM=4 TN=8 rows=1024 pf=2 nt=0
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r1024_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 1024, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r1024_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r1024_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t4x8_r1024_p2_nt: KC=8 MC=8 TM=4 TN=8 rows=1024 pf=2 nt=1
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t4x8_r1024_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 4, 8, 1024, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t4x8_r1024_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t4x8_r1024_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r256_p1: KC=8 MC=8 TM=8 TN=4 rows=256 pf=1 nt=0
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r256_p1_nt: KC=8 MC=8 TM=8 TN=4 rows=256 pf=1 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r256_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 256, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r256_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r256_p2: KC=8 MC=8 TM=8 TN=4 rows=256 pf=2 nt=0
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r256_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=256 pf=2 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r256_p4_nt: KC=8 MC=8 TM=8 TN=4 rows=256 pf=4 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r256_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 256, 4, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r256_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r512_p1: KC=8 MC=8 TM=8 TN=4 rows=512 pf=1 nt=0
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 512, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r512_p1_nt: KC=8 MC=8 TM=8 TN=4 rows=512 pf=1 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r512_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 512, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r512_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r512_p2: KC=8 MC=8 TM=8 TN=4 rows=512 pf=2 nt=0
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 512, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r512_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=512 pf=2 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r512_p4_nt: KC=8 MC=8 TM=8 TN=4 rows=512 pf=4 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r512_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 512, 4, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r512_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r1024_p1: KC=8 MC=8 TM=8 TN=4 rows=1024 pf=1 nt=0
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r1024_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 1024, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r1024_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r1024_p1_nt: KC=8 MC=8 TM=8 TN=4 rows=1024 pf=1 nt=1
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r1024_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 1024, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r1024_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r1024_p2: KC=8 MC=8 TM=8 TN=4 rows=1024 pf=2 nt=0
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r1024_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 1024, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r1024_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x4_r1024_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=1024 pf=2 nt=1
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x4_r1024_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 1024, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x4_r1024_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r256_p1: KC=8 MC=8 TM=8 TN=8 rows=256 pf=1 nt=0
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r256_p2: KC=8 MC=8 TM=8 TN=8 rows=256 pf=2 nt=0
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r256_p2_nt: KC=8 MC=8 TM=8 TN=8 rows=256 pf=2 nt=1
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r512_p1: KC=8 MC=8 TM=8 TN=8 rows=512 pf=1 nt=0
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 512, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r512_p2: KC=8 MC=8 TM=8 TN=8 rows=512 pf=2 nt=0
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 512, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r512_p2_nt: KC=8 MC=8 TM=8 TN=8 rows=512 pf=2 nt=1
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r1024_p1: KC=8 MC=8 TM=8 TN=8 rows=1024 pf=1 nt=0
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r1024_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 1024, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r1024_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r1024_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r1024_p2: KC=8 MC=8 TM=8 TN=8 rows=1024 pf=2 nt=0
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r1024_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 1024, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r1024_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r1024_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_8x8_t8x8_r1024_p2_nt: KC=8 MC=8 TM=8 TN=8 rows=1024 pf=2 nt=1
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_8x8_t8x8_r1024_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 8, 1024, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_8x8_t8x8_r1024_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_8x8_t8x8_r1024_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_8x8_mfma_r256_p1_nt: KC=8 MC=8 TM=0 TN=0 rows=256 pf=1 nt=1
//   matrix cores: 64 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_8x8_mfma_r256_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 0, 0, 256, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_8x8_mfma_r256_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_8x8_mfma_r256_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_8x8_mfma_r256_p2_nt: KC=8 MC=8 TM=0 TN=0 rows=256 pf=2 nt=1
//   matrix cores: 64 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_8x8_mfma_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 0, 0, 256, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_8x8_mfma_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_8x8_mfma_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_8x8_t8x4_r512_p1_nt: KC=8 MC=8 TM=8 TN=4 rows=512 pf=1 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_8x8_t8x4_r512_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 512, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_8x8_t8x4_r512_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_8x8_t8x4_r512_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_8x8_t8x4_r512_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=512 pf=2 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_8x8_t8x4_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 512, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_8x8_t8x4_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_8x8_t8x4_r1024_p1_nt: KC=8 MC=8 TM=8 TN=4 rows=1024 pf=1 nt=1
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_8x8_t8x4_r1024_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 1024, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_8x8_t8x4_r1024_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_8x8_t8x4_r1024_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_8x8_t8x4_r1024_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=1024 pf=2 nt=1
//   2 register tiles x 128 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_8x8_t8x4_r1024_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 1024, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_8x8_t8x4_r1024_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_8x8_t8x4_r1024_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_8x8_t8x4_r2048_p1_nt: KC=8 MC=8 TM=8 TN=4 rows=2048 pf=1 nt=1
//   2 register tiles x 128 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_8x8_t8x4_r2048_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 2048, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_8x8_t8x4_r2048_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_8x8_t8x4_r2048_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_8x8_t8x4_r2048_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=2048 pf=2 nt=1
//   2 register tiles x 128 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_8x8_t8x4_r2048_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 8, 8, 8, 4, 2048, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_8x8_t8x4_r2048_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_8x8_t8x4_r2048_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
