// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_bf16_64x2_t8x2_r128_p2_nt: KC=64 MC=2 TM=8 TN=2 rows=128 pf=2 nt=1
//   8 register tiles x 32 row lanes, 4 rows/thread/tile, LDS 33792 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_64x2_t8x2_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 64, 2, 8, 2, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_64x2_t8x2_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_64x2_t8x2_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
