// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f32_64x64_mfma_r32_p1: KC=64 MC=64 TM=0 TN=0 rows=32 pf=1 nt=0
//   matrix cores: 8 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r32_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 32, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r32_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r32_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r32_p1_nt: KC=64 MC=64 TM=0 TN=0 rows=32 pf=1 nt=1
//   matrix cores: 8 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r32_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 32, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r32_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r32_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r32_p2: KC=64 MC=64 TM=0 TN=0 rows=32 pf=2 nt=0
//   matrix cores: 8 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r32_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 32, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r32_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r32_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r32_p2_nt: KC=64 MC=64 TM=0 TN=0 rows=32 pf=2 nt=1
//   matrix cores: 8 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r32_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 32, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r32_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r32_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r64_p1: KC=64 MC=64 TM=0 TN=0 rows=64 pf=1 nt=0
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r64_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 64, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r64_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r64_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r64_p1_nt: KC=64 MC=64 TM=0 TN=0 rows=64 pf=1 nt=1
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r64_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 64, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r64_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r64_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r64_p2: KC=64 MC=64 TM=0 TN=0 rows=64 pf=2 nt=0
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 64, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r64_p2_nt: KC=64 MC=64 TM=0 TN=0 rows=64 pf=2 nt=1
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r64_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 64, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r64_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r64_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r128_p1: KC=64 MC=64 TM=0 TN=0 rows=128 pf=1 nt=0
//   matrix cores: 32 K-steps of 4 rows per tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r128_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 128, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r128_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r128_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r128_p1_nt: KC=64 MC=64 TM=0 TN=0 rows=128 pf=1 nt=1
//   matrix cores: 32 K-steps of 4 rows per tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r128_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 128, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r128_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r128_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r128_p2: KC=64 MC=64 TM=0 TN=0 rows=128 pf=2 nt=0
//   matrix cores: 32 K-steps of 4 rows per tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r128_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 128, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r128_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r128_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_mfma_r128_p2_nt: KC=64 MC=64 TM=0 TN=0 rows=128 pf=2 nt=1
//   matrix cores: 32 K-steps of 4 rows per tile, LDS 131072 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_mfma_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_mfma_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_mfma_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r16_p1: KC=64 MC=64 TM=8 TN=8 rows=16 pf=1 nt=0
//   64 register tiles x 4 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r16_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 16, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r16_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r16_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r16_p1_nt: KC=64 MC=64 TM=8 TN=8 rows=16 pf=1 nt=1
//   64 register tiles x 4 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r16_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 16, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r16_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r16_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r16_p2: KC=64 MC=64 TM=8 TN=8 rows=16 pf=2 nt=0
//   64 register tiles x 4 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r16_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 16, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r16_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r16_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r16_p2_nt: KC=64 MC=64 TM=8 TN=8 rows=16 pf=2 nt=1
//   64 register tiles x 4 row lanes, 4 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r16_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 16, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r16_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r16_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r32_p1: KC=64 MC=64 TM=8 TN=8 rows=32 pf=1 nt=0
//   64 register tiles x 4 row lanes, 8 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r32_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 32, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r32_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r32_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r32_p1_nt: KC=64 MC=64 TM=8 TN=8 rows=32 pf=1 nt=1
//   64 register tiles x 4 row lanes, 8 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r32_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 32, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r32_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r32_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r32_p2: KC=64 MC=64 TM=8 TN=8 rows=32 pf=2 nt=0
//   64 register tiles x 4 row lanes, 8 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r32_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 32, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r32_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r32_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r32_p2_nt: KC=64 MC=64 TM=8 TN=8 rows=32 pf=2 nt=1
//   64 register tiles x 4 row lanes, 8 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r32_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 32, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r32_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r32_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r64_p1: KC=64 MC=64 TM=8 TN=8 rows=64 pf=1 nt=0
//   64 register tiles x 4 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r64_p1(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 64, 1, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p1(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r64_p1, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r64_p1_nt: KC=64 MC=64 TM=8 TN=8 rows=64 pf=1 nt=1
//   64 register tiles x 4 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r64_p1_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 64, 1, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p1_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r64_p1_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r64_p2: KC=64 MC=64 TM=8 TN=8 rows=64 pf=2 nt=0
//   64 register tiles x 4 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 64, 2, 0>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f32_64x64_t8x8_r64_p2_nt: KC=64 MC=64 TM=8 TN=8 rows=64 pf=2 nt=1
//   64 register tiles x 4 row lanes, 16 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_64x64_t8x8_r64_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 8, 8, 64, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_64x64_t8x8_r64_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_64x64_mfma_r64_p2: KC=64 MC=64 TM=0 TN=0 rows=64 pf=2 nt=0
//   matrix cores: 16 K-steps of 4 rows per tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_64x64_mfma_r64_p2(tsm::AtbArgs a) {
  tsm::atb_body<float, 64, 64, 0, 0, 64, 2, 0, 1>(a);
}
hipError_t launch_tsm_ata_f32_64x64_mfma_r64_p2(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_64x64_mfma_r64_p2, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
