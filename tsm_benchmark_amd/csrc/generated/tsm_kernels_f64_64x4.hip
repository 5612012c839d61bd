// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_64x4_t8x4_r32_p2_nt: KC=64 MC=4 TM=8 TN=4 rows=32 pf=2 nt=1
//   8 register tiles x 32 row lanes, 1 rows/thread/tile, LDS 34816 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_64x4_t8x4_r32_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 64, 4, 8, 4, 32, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_64x4_t8x4_r32_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_64x4_t8x4_r32_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
