// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_16x1_t8x1_r128_p2_nt: KC=16 MC=1 TM=8 TN=1 rows=128 pf=2 nt=1
//   2 register tiles x 128 row lanes, 1 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x1_t8x1_r128_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 1, 8, 1, 128, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x1_t8x1_r128_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x1_t8x1_r128_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
