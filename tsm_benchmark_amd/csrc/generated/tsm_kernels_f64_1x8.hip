// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_1x8_t1x8_r256_p2_nt: KC=1 MC=8 TM=1 TN=8 rows=256 pf=2 nt=1
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 36864 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_1x8_t1x8_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 1, 8, 1, 8, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_1x8_t1x8_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_1x8_t1x8_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
