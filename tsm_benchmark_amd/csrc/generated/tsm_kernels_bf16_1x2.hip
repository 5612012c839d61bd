// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_bf16_1x2_t1x2_r4096_p2_nt: KC=1 MC=2 TM=1 TN=2 rows=4096 pf=2 nt=1
//   1 register tiles x 256 row lanes, 16 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_bf16_1x2_t1x2_r4096_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<tsm::bf16, 1, 2, 1, 2, 4096, 2, 1>(a);
}
hipError_t launch_tsm_atb_bf16_1x2_t1x2_r4096_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_bf16_1x2_t1x2_r4096_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
