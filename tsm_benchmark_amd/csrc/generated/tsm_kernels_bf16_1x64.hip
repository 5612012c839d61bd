// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_bf16_1x64_t1x4_r128_p2_nt: KC=1 MC=64 TM=1 TN=4 rows=128 pf=2 nt=1
//   16 register tiles x 16 row lanes, 8 rows/thread/tile, LDS 33280 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_1x64_t1x4_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 1, 64, 1, 4, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_1x64_t1x4_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_1x64_t1x4_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
