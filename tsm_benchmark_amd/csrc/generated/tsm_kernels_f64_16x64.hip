// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_f64_16x64_mfma_r32_p2_nt: KC=16 MC=64 TM=0 TN=0 rows=32 pf=2 nt=1
//   matrix cores: 8 K-steps of 4 rows per tile, LDS 40960 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_f64_16x64_mfma_r32_p2_nt(tsm::AtbArgs a) {
  tsm::atb_body<double, 16, 64, 0, 0, 32, 2, 1>(a);
}
hipError_t launch_tsm_atb_f64_16x64_mfma_r32_p2_nt(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_f64_16x64_mfma_r32_p2_nt, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}
