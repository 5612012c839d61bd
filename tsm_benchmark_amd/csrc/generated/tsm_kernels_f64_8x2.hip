// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_8x2_t8x2_r256_p2_nt: KC=8 MC=2 TM=8 TN=2 rows=256 pf=2 nt=1
//   1 register tiles x 256 row lanes, 1 rows/thread/tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_8x2_t8x2_r256_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 8, 2, 8, 2, 256, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_8x2_t8x2_r256_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_8x2_t8x2_r256_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
