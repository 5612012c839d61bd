// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_bf16_16x16_direct_d2_nt: KC=16 MC=16, matrix cores fed from global memory, 2 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x16_direct_d2_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 16, 16, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x16_direct_d2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x16_direct_d2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x16_direct_d4_nt: KC=16 MC=16, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x16_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 16, 16, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x16_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x16_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x32_direct_d4_nt: KC=16 MC=32, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 24576 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x32_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 16, 32, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x32_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x32_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_16x64_direct_d4_nt: KC=16 MC=64, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_16x64_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 16, 64, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_16x64_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_16x64_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_32x16_direct_d4_nt: KC=32 MC=16, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 24576 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_32x16_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 32, 16, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_32x16_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_32x16_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_32x32_direct_d2_nt: KC=32 MC=32, matrix cores fed from global memory, 2 4-row steps in flight per wave, nt=1, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_32x32_direct_d2_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 32, 32, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_32x32_direct_d2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_32x32_direct_d2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_32x32_direct_d4_nt: KC=32 MC=32, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_32x32_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 32, 32, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_32x32_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_32x32_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_32x64_direct_d4_nt: KC=32 MC=64, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_32x64_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 32, 64, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_32x64_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_32x64_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_64x16_direct_d4_nt: KC=64 MC=16, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x16_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 64, 16, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x16_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x16_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_64x32_direct_d4_nt: KC=64 MC=32, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x32_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 64, 32, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x32_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x32_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_64x64_direct_d2_nt: KC=64 MC=64, matrix cores fed from global memory, 2 4-row steps in flight per wave, nt=1, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x64_direct_d2_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 64, 64, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x64_direct_d2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x64_direct_d2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_bf16_64x64_direct_d4_nt: KC=64 MC=64, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x64_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 64, 64, 4, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x64_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x64_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_bf16_16x16_direct_d2_nt: KC=16 MC=16, matrix cores fed from global memory, 2 4-row steps in flight per wave, nt=1, LDS 8192 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_bf16_16x16_direct_d2_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 16, 16, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_bf16_16x16_direct_d2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_bf16_16x16_direct_d2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_bf16_16x16_direct_d4_nt: KC=16 MC=16, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 8192 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_bf16_16x16_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 16, 16, 4, 1, 1>(a);
}
hipError_t launch_tsm_ata_bf16_16x16_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_bf16_16x16_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_bf16_32x32_direct_d2_nt: KC=32 MC=32, matrix cores fed from global memory, 2 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_bf16_32x32_direct_d2_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 32, 32, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_bf16_32x32_direct_d2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_bf16_32x32_direct_d2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_bf16_32x32_direct_d4_nt: KC=32 MC=32, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_bf16_32x32_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 32, 32, 4, 1, 1>(a);
}
hipError_t launch_tsm_ata_bf16_32x32_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_bf16_32x32_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_bf16_64x64_direct_d2_nt: KC=64 MC=64, matrix cores fed from global memory, 2 4-row steps in flight per wave, nt=1, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_bf16_64x64_direct_d2_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 64, 64, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_bf16_64x64_direct_d2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_bf16_64x64_direct_d2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_bf16_64x64_direct_d4_nt: KC=64 MC=64, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_bf16_64x64_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<tsm::bf16, 64, 64, 4, 1, 1>(a);
}
hipError_t launch_tsm_ata_bf16_64x64_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_bf16_64x64_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
