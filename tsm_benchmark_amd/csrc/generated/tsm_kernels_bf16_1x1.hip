// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_bf16_1x1_t1x1_r4096_p2_nt: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=2 nt=1
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 32768 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_1x1_t1x1_r4096_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 1, 1, 1, 1, 4096, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_1x1_t1x1_r4096_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// tsm_ata_bf16_1x1_t1x1_r4096_p2_nt: KC=1 MC=1 TM=1 TN=1 rows=4096 pf=2 nt=1
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 16384 B
extern "C" __global__ void __launch_bounds__(256) tsm_ata_bf16_1x1_t1x1_r4096_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 1, 1, 1, 1, 4096, 2, 1, 1>(a);
}
hipError_t launch_tsm_ata_bf16_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_ata_bf16_1x1_t1x1_r4096_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
