// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_1x1_t1x1_r1024_p1: KC=1 MC=1 TM=1 TN=1 rows=1024 pf=1 nt=0
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r1024_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 1024, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r1024_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r1024_p1_nt: KC=1 MC=1 TM=1 TN=1 rows=1024 pf=1 nt=1
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r1024_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 1024, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r1024_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r1024_p2: KC=1 MC=1 TM=1 TN=1 rows=1024 pf=2 nt=0
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r1024_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 1024, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r1024_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r1024_p2_nt: KC=1 MC=1 TM=1 TN=1 rows=1024 pf=2 nt=1
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r1024_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 1024, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r1024_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r1024_p4_nt: KC=1 MC=1 TM=1 TN=1 rows=1024 pf=4 nt=1
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r1024_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 1024, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r1024_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r2048_p1: KC=1 MC=1 TM=1 TN=1 rows=2048 pf=1 nt=0
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r2048_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 2048, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r2048_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r2048_p1_nt: KC=1 MC=1 TM=1 TN=1 rows=2048 pf=1 nt=1
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r2048_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 2048, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r2048_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r2048_p2: KC=1 MC=1 TM=1 TN=1 rows=2048 pf=2 nt=0
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r2048_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 2048, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r2048_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r2048_p2_nt: KC=1 MC=1 TM=1 TN=1 rows=2048 pf=2 nt=1
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r2048_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 2048, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r2048_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r2048_p4_nt: KC=1 MC=1 TM=1 TN=1 rows=2048 pf=4 nt=1
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r2048_p4_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 2048, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r2048_p4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r4096_p1: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=1 nt=0
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r4096_p1(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 4096, 1, 0>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r4096_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r4096_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r4096_p1_nt: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=1 nt=1
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r4096_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 4096, 1, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r4096_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r4096_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r4096_p2: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=2 nt=0
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r4096_p2(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 4096, 2, 0>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r4096_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r4096_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_1x1_t1x1_r4096_p2_nt: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=2 nt=1
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x1_t1x1_r4096_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 4096, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x1_t1x1_r4096_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_1x1_t1x1_r2048_p1_nt: KC=1 MC=1 TM=1 TN=1 rows=2048 pf=1 nt=1
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_1x1_t1x1_r2048_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 2048, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_1x1_t1x1_r2048_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_1x1_t1x1_r2048_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_1x1_t1x1_r2048_p2_nt: KC=1 MC=1 TM=1 TN=1 rows=2048 pf=2 nt=1
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_1x1_t1x1_r2048_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 2048, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_1x1_t1x1_r2048_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_1x1_t1x1_r2048_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_1x1_t1x1_r4096_p1_nt: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=1 nt=1
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_1x1_t1x1_r4096_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 4096, 1, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_1x1_t1x1_r4096_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_1x1_t1x1_r4096_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_1x1_t1x1_r4096_p2_nt: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=2 nt=1
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_1x1_t1x1_r4096_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 1, 1, 1, 4096, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_1x1_t1x1_r4096_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
