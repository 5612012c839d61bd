// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_32x16_t8x8_r64_p2_nt: KC=32 MC=16 TM=8 TN=8 rows=64 pf=2 nt=1
//   8 register tiles x 32 row lanes, 2 rows/thread/tile, LDS 49152 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_32x16_t8x8_r64_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 32, 16, 8, 8, 64, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_32x16_t8x8_r64_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_32x16_t8x8_r64_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
