// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f32_2x32_t2x4_r128_p2_nt: KC=2 MC=32 TM=2 TN=4 rows=128 pf=2 nt=1
//   8 register tiles x 32 row lanes, 4 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f32_2x32_t2x4_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<float, 2, 32, 2, 4, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_f32_2x32_t2x4_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f32_2x32_t2x4_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
