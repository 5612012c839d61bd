// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f32_4x4_t4x4_r1024_p2_nt: KC=4 MC=4 TM=4 TN=4 rows=1024 pf=2 nt=1
//   1 register tiles x 256 row lanes, 4 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_4x4_t4x4_r1024_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 4, 4, 4, 4, 1024, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_4x4_t4x4_r1024_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_4x4_t4x4_r1024_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_f32_4x4_t4x4_r2048_p2_nt: KC=4 MC=4 TM=4 TN=4 rows=2048 pf=2 nt=1
//   1 register tiles x 256 row lanes, 8 rows/thread/tile, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_f32_4x4_t4x4_r2048_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 4, 4, 4, 4, 2048, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_f32_4x4_t4x4_r2048_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_f32_4x4_t4x4_r2048_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
