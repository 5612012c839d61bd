// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_16x16_direct_d4_nt: KC=16 MC=16, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 4096 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x16_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 16, 16, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x16_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x16_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x16_direct_d8_nt: KC=16 MC=16, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 4096 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x16_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 16, 16, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x16_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x16_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x16_direct_d16_nt: KC=16 MC=16, matrix cores fed from global memory, 16 4-row steps in flight per wave, nt=1, LDS 4096 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x16_direct_d16_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 16, 16, 16, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x16_direct_d16_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x16_direct_d16_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x32_direct_d8_nt: KC=16 MC=32, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 8192 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x32_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 16, 32, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x32_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x32_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_16x64_direct_d8_nt: KC=16 MC=64, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x64_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 16, 64, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x64_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x64_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_32x16_direct_d8_nt: KC=32 MC=16, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 8192 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_32x16_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 32, 16, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_32x16_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_32x16_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_32x32_direct_d4_nt: KC=32 MC=32, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_32x32_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 32, 32, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_32x32_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_32x32_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_32x32_direct_d8_nt: KC=32 MC=32, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_32x32_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 32, 32, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_32x32_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_32x32_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_32x32_direct_d16_nt: KC=32 MC=32, matrix cores fed from global memory, 16 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_32x32_direct_d16_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 32, 32, 16, 1>(a);
}
hipError_t launch_tsm_atb_f64_32x32_direct_d16_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_32x32_direct_d16_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_32x64_direct_d8_nt: KC=32 MC=64, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_32x64_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 32, 64, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_32x64_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_32x64_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_64x16_direct_d8_nt: KC=64 MC=16, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_64x16_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 64, 16, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_64x16_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_64x16_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_64x32_direct_d8_nt: KC=64 MC=32, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_64x32_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 64, 32, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_64x32_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_64x32_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_64x64_direct_d4_nt: KC=64 MC=64, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_64x64_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 64, 64, 4, 1>(a);
}
hipError_t launch_tsm_atb_f64_64x64_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_64x64_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_64x64_direct_d8_nt: KC=64 MC=64, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_64x64_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 64, 64, 8, 1>(a);
}
hipError_t launch_tsm_atb_f64_64x64_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_64x64_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_atb_f64_64x64_direct_d16_nt: KC=64 MC=64, matrix cores fed from global memory, 16 4-row steps in flight per wave, nt=1, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_64x64_direct_d16_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 64, 64, 16, 1>(a);
}
hipError_t launch_tsm_atb_f64_64x64_direct_d16_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_64x64_direct_d16_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_16x16_direct_d4_nt: KC=16 MC=16, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 4096 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_16x16_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 16, 16, 4, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_16x16_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_16x16_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_16x16_direct_d8_nt: KC=16 MC=16, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 4096 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_16x16_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 16, 16, 8, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_16x16_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_16x16_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_16x16_direct_d16_nt: KC=16 MC=16, matrix cores fed from global memory, 16 4-row steps in flight per wave, nt=1, LDS 4096 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_16x16_direct_d16_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 16, 16, 16, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_16x16_direct_d16_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_16x16_direct_d16_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_32x32_direct_d4_nt: KC=32 MC=32, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_32x32_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 32, 32, 4, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_32x32_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_32x32_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_32x32_direct_d8_nt: KC=32 MC=32, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_32x32_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 32, 32, 8, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_32x32_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_32x32_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_32x32_direct_d16_nt: KC=32 MC=32, matrix cores fed from global memory, 16 4-row steps in flight per wave, nt=1, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_32x32_direct_d16_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 32, 32, 16, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_32x32_direct_d16_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_32x32_direct_d16_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_64x64_direct_d4_nt: KC=64 MC=64, matrix cores fed from global memory, 4 4-row steps in flight per wave, nt=1, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_64x64_direct_d4_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 64, 64, 4, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_64x64_direct_d4_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_64x64_direct_d4_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_64x64_direct_d8_nt: KC=64 MC=64, matrix cores fed from global memory, 8 4-row steps in flight per wave, nt=1, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_64x64_direct_d8_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 64, 64, 8, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_64x64_direct_d8_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_64x64_direct_d8_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f64_64x64_direct_d16_nt: KC=64 MC=64, matrix cores fed from global memory, 16 4-row steps in flight per wave, nt=1, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f64_64x64_direct_d16_nt(tsm::AtbArgs a) {
  tsm::direct_body<double, 64, 64, 16, 1, 1>(a);
}
hipError_t launch_tsm_ata_f64_64x64_direct_d16_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f64_64x64_direct_d16_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
