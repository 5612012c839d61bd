// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f32_2x16_t2x4_r256_p2_nt: KC=2 MC=16 TM=2 TN=4 rows=256 pf=2 nt=1
//   4 register tiles x 64 row lanes, 4 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_2x16_t2x4_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 2, 16, 2, 4, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_2x16_t2x4_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_2x16_t2x4_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
