// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_8x8_t2x2_r256_p1: KC=8 MC=8 TM=2 TN=2 rows=256 pf=1 nt=0
//   16 register tiles x 16 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t2x2_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 2, 2, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t2x2_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t2x2_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t2x2_r256_p2: KC=8 MC=8 TM=2 TN=2 rows=256 pf=2 nt=0
//   16 register tiles x 16 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t2x2_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 2, 2, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t2x2_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t2x2_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r128_p1: KC=8 MC=8 TM=4 TN=4 rows=128 pf=1 nt=0
//   4 register tiles x 64 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r128_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 128, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r128_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r128_p2: KC=8 MC=8 TM=4 TN=4 rows=128 pf=2 nt=0
//   4 register tiles x 64 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 128, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r128_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=128 pf=2 nt=1
//   4 register tiles x 64 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r128_p4_nt: KC=8 MC=8 TM=4 TN=4 rows=128 pf=4 nt=1
//   4 register tiles x 64 row lanes, 2 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r128_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 128, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r128_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r256_p1: KC=8 MC=8 TM=4 TN=4 rows=256 pf=1 nt=0
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r256_p2: KC=8 MC=8 TM=4 TN=4 rows=256 pf=2 nt=0
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r256_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=256 pf=2 nt=1
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r256_p4_nt: KC=8 MC=8 TM=4 TN=4 rows=256 pf=4 nt=1
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r256_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 256, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r256_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r512_p1: KC=8 MC=8 TM=4 TN=4 rows=512 pf=1 nt=0
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 512, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r512_p2: KC=8 MC=8 TM=4 TN=4 rows=512 pf=2 nt=0
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 512, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x4_r512_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=512 pf=2 nt=1
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x4_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x4_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x4_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r128_p1: KC=8 MC=8 TM=4 TN=8 rows=128 pf=1 nt=0
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r128_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 128, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r128_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r128_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r128_p2: KC=8 MC=8 TM=4 TN=8 rows=128 pf=2 nt=0
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 128, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r128_p2_nt: KC=8 MC=8 TM=4 TN=8 rows=128 pf=2 nt=1
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r128_p4_nt: KC=8 MC=8 TM=4 TN=8 rows=128 pf=4 nt=1
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r128_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 128, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r128_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r128_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r256_p1: KC=8 MC=8 TM=4 TN=8 rows=256 pf=1 nt=0
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r256_p2: KC=8 MC=8 TM=4 TN=8 rows=256 pf=2 nt=0
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r256_p2_nt: KC=8 MC=8 TM=4 TN=8 rows=256 pf=2 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r256_p4_nt: KC=8 MC=8 TM=4 TN=8 rows=256 pf=4 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r256_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 256, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r256_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r256_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r512_p1: KC=8 MC=8 TM=4 TN=8 rows=512 pf=1 nt=0
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 512, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r512_p2: KC=8 MC=8 TM=4 TN=8 rows=512 pf=2 nt=0
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 512, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t4x8_r512_p2_nt: KC=8 MC=8 TM=4 TN=8 rows=512 pf=2 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t4x8_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 8, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t4x8_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t4x8_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r128_p1: KC=8 MC=8 TM=8 TN=4 rows=128 pf=1 nt=0
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r128_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 128, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r128_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r128_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r128_p2: KC=8 MC=8 TM=8 TN=4 rows=128 pf=2 nt=0
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 128, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r128_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=128 pf=2 nt=1
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r128_p4_nt: KC=8 MC=8 TM=8 TN=4 rows=128 pf=4 nt=1
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r128_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 128, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r128_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r128_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r256_p1: KC=8 MC=8 TM=8 TN=4 rows=256 pf=1 nt=0
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r256_p2: KC=8 MC=8 TM=8 TN=4 rows=256 pf=2 nt=0
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r256_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=256 pf=2 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r256_p4_nt: KC=8 MC=8 TM=8 TN=4 rows=256 pf=4 nt=1
//   2 register tiles x 128 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r256_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 256, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r256_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r256_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r512_p1: KC=8 MC=8 TM=8 TN=4 rows=512 pf=1 nt=0
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 512, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r512_p2: KC=8 MC=8 TM=8 TN=4 rows=512 pf=2 nt=0
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 512, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x4_r512_p2_nt: KC=8 MC=8 TM=8 TN=4 rows=512 pf=2 nt=1
//   2 register tiles x 128 row lanes, 4 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x4_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 4, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x4_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r256_p1: KC=8 MC=8 TM=8 TN=8 rows=256 pf=1 nt=0
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r256_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 256, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r256_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r256_p1_nt: KC=8 MC=8 TM=8 TN=8 rows=256 pf=1 nt=1
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r256_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 256, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r256_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r256_p2: KC=8 MC=8 TM=8 TN=8 rows=256 pf=2 nt=0
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r256_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 256, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r256_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r256_p2_nt: KC=8 MC=8 TM=8 TN=8 rows=256 pf=2 nt=1
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r512_p1: KC=8 MC=8 TM=8 TN=8 rows=512 pf=1 nt=0
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r512_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 512, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r512_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r512_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r512_p1_nt: KC=8 MC=8 TM=8 TN=8 rows=512 pf=1 nt=1
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r512_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 512, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r512_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r512_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r512_p2: KC=8 MC=8 TM=8 TN=8 rows=512 pf=2 nt=0
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r512_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 512, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r512_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r512_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_8x8_t8x8_r512_p2_nt: KC=8 MC=8 TM=8 TN=8 rows=512 pf=2 nt=1
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x8_t8x8_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 512, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x8_t8x8_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x8_t8x8_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_mfma_r128_p1_nt: KC=8 MC=8 TM=0 TN=0 rows=128 pf=1 nt=1
//   matrix cores: 32 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_mfma_r128_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 0, 0, 128, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_mfma_r128_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_mfma_r128_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_mfma_r128_p2_nt: KC=8 MC=8 TM=0 TN=0 rows=128 pf=2 nt=1
//   matrix cores: 32 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_mfma_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 0, 0, 128, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_mfma_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_mfma_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_mfma_r256_p1_nt: KC=8 MC=8 TM=0 TN=0 rows=256 pf=1 nt=1
//   matrix cores: 64 K-steps of 4 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_mfma_r256_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 0, 0, 256, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_mfma_r256_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_mfma_r256_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_mfma_r256_p2_nt: KC=8 MC=8 TM=0 TN=0 rows=256 pf=2 nt=1
//   matrix cores: 64 K-steps of 4 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_mfma_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 0, 0, 256, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_mfma_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_mfma_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t4x4_r64_p1_nt: KC=8 MC=8 TM=4 TN=4 rows=64 pf=1 nt=1
//   4 register tiles x 64 row lanes, 1 rows/thread/tile, LDS 8192 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t4x4_r64_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 64, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t4x4_r64_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t4x4_r64_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t4x4_r64_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=64 pf=2 nt=1
//   4 register tiles x 64 row lanes, 1 rows/thread/tile, LDS 8192 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t4x4_r64_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 64, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t4x4_r64_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t4x4_r64_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t4x4_r128_p1_nt: KC=8 MC=8 TM=4 TN=4 rows=128 pf=1 nt=1
//   4 register tiles x 64 row lanes, 2 rows/thread/tile, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t4x4_r128_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 128, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t4x4_r128_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t4x4_r128_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t4x4_r128_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=128 pf=2 nt=1
//   4 register tiles x 64 row lanes, 2 rows/thread/tile, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t4x4_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 128, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t4x4_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t4x4_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t4x4_r256_p1_nt: KC=8 MC=8 TM=4 TN=4 rows=256 pf=1 nt=1
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t4x4_r256_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 256, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t4x4_r256_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t4x4_r256_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t4x4_r256_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=256 pf=2 nt=1
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t4x4_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 256, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t4x4_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t4x4_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t4x4_r512_p1_nt: KC=8 MC=8 TM=4 TN=4 rows=512 pf=1 nt=1
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t4x4_r512_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 512, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t4x4_r512_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t4x4_r512_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t4x4_r512_p2_nt: KC=8 MC=8 TM=4 TN=4 rows=512 pf=2 nt=1
//   4 register tiles x 64 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t4x4_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 4, 4, 512, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t4x4_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t4x4_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_8x8_t8x8_r512_p2_nt: KC=8 MC=8 TM=8 TN=8 rows=512 pf=2 nt=1
//   1 register tiles x 256 row lanes, 2 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_8x8_t8x8_r512_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 8, 8, 8, 512, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_8x8_t8x8_r512_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_8x8_t8x8_r512_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
