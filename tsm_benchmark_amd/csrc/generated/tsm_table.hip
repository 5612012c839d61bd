// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

hipError_t launch_tsm_atb_bf16_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x2_t1x2_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x4_t1x4_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x8_t1x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x16_t1x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x32_t1x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x64_t1x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x1_t2x1_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x2_t2x2_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x4_t2x4_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x8_t2x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x16_t2x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x32_t2x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x64_t2x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x1_t4x1_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x2_t4x2_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x4_t4x4_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x8_t4x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x16_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x16_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x16_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x16_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x16_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x64_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x64_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x64_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x64_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x64_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x1_t8x1_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x2_t8x2_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x4_t8x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r512_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r1024_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r512_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r1024_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x8_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x16_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x16_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x16_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x16_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x16_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x64_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x64_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x64_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x64_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x64_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x1_t8x1_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x2_t8x2_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_direct_d2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x32_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x64_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x1_t8x1_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x2_t8x2_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x16_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_direct_d2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x64_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x1_t8x1_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x2_t8x2_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x4_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x4_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x4_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x4_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x4_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x8_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x8_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x8_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x8_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x8_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x16_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x32_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_direct_d2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_mfma_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_mfma_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_mfma_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_mfma_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r16_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r16_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x2_t1x2_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x4_t1x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x8_t1x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x16_t1x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x32_t1x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x64_t1x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x1_t2x1_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x2_t2x2_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x4_t2x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x8_t2x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x16_t2x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x32_t2x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x64_t2x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x1_t4x1_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x2_t4x2_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x4_t4x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x8_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x16_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x64_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x64_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x64_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x64_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x64_t4x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x1_t8x1_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x2_t8x2_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x4_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r512_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x8_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x16_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x1_t8x1_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x2_t8x2_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x4_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x4_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x4_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x4_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x4_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x8_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x8_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x8_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x8_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x8_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_mfma_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_mfma_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_mfma_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_mfma_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x32_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x1_t8x1_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x2_t8x2_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x16_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x16_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x1_t8x1_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x2_t8x2_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x16_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x32_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r16_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r16_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r4096_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r4096_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r4096_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x2_t1x2_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x4_t1x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x8_t1x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x16_t1x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x32_t1x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x64_t1x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x1_t2x1_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x4_t2x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x8_t2x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x16_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x16_t2x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_t2x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x64_mfma_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x64_mfma_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x64_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x64_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x64_t2x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x1_t4x1_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x2_t4x2_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t2x2_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t2x2_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x8_t4x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x64_mfma_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x64_mfma_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x64_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x64_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x64_t4x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x1_t8x1_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x2_t8x2_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x4_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t2x2_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t2x2_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r256_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x16_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x16_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_mfma_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_mfma_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x1_t8x1_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_t8x2_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t2x2_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t2x2_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r128_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x32_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x32_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x1_t8x1_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_t8x2_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x8_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x8_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x8_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x8_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x8_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x8_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x8_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x16_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x16_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r32_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r64_p4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x1_t8x1_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_mfma_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_mfma_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_t8x2_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_mfma_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_mfma_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_t8x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_mfma_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_mfma_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x16_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x16_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x32_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x32_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r16_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r16_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r16_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r16_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_2x2_t2x2_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_4x4_t4x4_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_mfma_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_mfma_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_t8x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_t8x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_t8x4_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_8x8_t8x4_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_direct_d2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_32x32_direct_d2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_32x32_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_32x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_32x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_32x32_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_32x32_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_32x32_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_32x32_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_64x64_direct_d2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_64x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_64x64_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_2x2_t2x2_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_4x4_t4x4_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_8x8_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_8x8_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_8x8_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_8x8_t8x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_8x8_t8x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_8x8_t8x4_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_8x8_t8x4_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_64x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_64x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_64x64_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_64x64_mfma_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_1x1_t1x1_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_1x1_t1x1_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_1x1_t1x1_r4096_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_1x1_t1x1_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_2x2_t2x2_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_2x2_t2x2_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_2x2_t2x2_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_2x2_t2x2_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_2x2_t2x2_r4096_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_2x2_t2x2_r4096_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_4x4_t4x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_4x4_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_4x4_t4x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_4x4_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_4x4_t4x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_4x4_t4x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t4x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t4x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t4x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t4x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t4x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t8x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r16_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r16_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_64x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_64x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_64x64_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_64x64_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);

static const tsm::KernelEntry kTable[] = {
  {"tsm_atb_bf16_1x1_t1x1_r4096_p2_nt", 2, 1, 1, 1, 1, 4096, 2, 1, 32768, launch_tsm_atb_bf16_1x1_t1x1_r4096_p2_nt, 1},
  {"tsm_atb_bf16_1x2_t1x2_r4096_p2_nt", 2, 1, 2, 1, 2, 4096, 2, 1, 49152, launch_tsm_atb_bf16_1x2_t1x2_r4096_p2_nt, 1},
  {"tsm_atb_bf16_1x4_t1x4_r2048_p2_nt", 2, 1, 4, 1, 4, 2048, 2, 1, 40960, launch_tsm_atb_bf16_1x4_t1x4_r2048_p2_nt, 1},
  {"tsm_atb_bf16_1x8_t1x4_r1024_p2_nt", 2, 1, 8, 1, 4, 1024, 2, 1, 36864, launch_tsm_atb_bf16_1x8_t1x4_r1024_p2_nt, 1},
  {"tsm_atb_bf16_1x16_t1x4_r512_p2_nt", 2, 1, 16, 1, 4, 512, 2, 1, 34816, launch_tsm_atb_bf16_1x16_t1x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_1x32_t1x4_r256_p2_nt", 2, 1, 32, 1, 4, 256, 2, 1, 33792, launch_tsm_atb_bf16_1x32_t1x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_1x64_t1x4_r128_p2_nt", 2, 1, 64, 1, 4, 128, 2, 1, 33280, launch_tsm_atb_bf16_1x64_t1x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_2x1_t2x1_r4096_p2_nt", 2, 2, 1, 2, 1, 4096, 2, 1, 49152, launch_tsm_atb_bf16_2x1_t2x1_r4096_p2_nt, 1},
  {"tsm_atb_bf16_2x2_t2x2_r4096_p2_nt", 2, 2, 2, 2, 2, 4096, 2, 1, 65536, launch_tsm_atb_bf16_2x2_t2x2_r4096_p2_nt, 1},
  {"tsm_atb_bf16_2x4_t2x4_r2048_p2_nt", 2, 2, 4, 2, 4, 2048, 2, 1, 49152, launch_tsm_atb_bf16_2x4_t2x4_r2048_p2_nt, 1},
  {"tsm_atb_bf16_2x8_t2x4_r1024_p2_nt", 2, 2, 8, 2, 4, 1024, 2, 1, 40960, launch_tsm_atb_bf16_2x8_t2x4_r1024_p2_nt, 1},
  {"tsm_atb_bf16_2x16_t2x4_r512_p2_nt", 2, 2, 16, 2, 4, 512, 2, 1, 36864, launch_tsm_atb_bf16_2x16_t2x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_2x32_t2x4_r256_p2_nt", 2, 2, 32, 2, 4, 256, 2, 1, 34816, launch_tsm_atb_bf16_2x32_t2x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_2x64_t2x4_r128_p2_nt", 2, 2, 64, 2, 4, 128, 2, 1, 33792, launch_tsm_atb_bf16_2x64_t2x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_4x1_t4x1_r2048_p2_nt", 2, 4, 1, 4, 1, 2048, 2, 1, 40960, launch_tsm_atb_bf16_4x1_t4x1_r2048_p2_nt, 1},
  {"tsm_atb_bf16_4x2_t4x2_r2048_p2_nt", 2, 4, 2, 4, 2, 2048, 2, 1, 49152, launch_tsm_atb_bf16_4x2_t4x2_r2048_p2_nt, 1},
  {"tsm_atb_bf16_4x4_t4x4_r2048_p2_nt", 2, 4, 4, 4, 4, 2048, 2, 1, 65536, launch_tsm_atb_bf16_4x4_t4x4_r2048_p2_nt, 1},
  {"tsm_atb_bf16_4x8_t4x4_r1024_p2_nt", 2, 4, 8, 4, 4, 1024, 2, 1, 49152, launch_tsm_atb_bf16_4x8_t4x4_r1024_p2_nt, 1},
  {"tsm_atb_bf16_4x16_mfma_r256_p1_nt", 2, 4, 16, 0, 0, 256, 1, 1, 32768, launch_tsm_atb_bf16_4x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_4x16_mfma_r256_p2_nt", 2, 4, 16, 0, 0, 256, 2, 1, 32768, launch_tsm_atb_bf16_4x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_4x16_mfma_r512_p1_nt", 2, 4, 16, 0, 0, 512, 1, 1, 65536, launch_tsm_atb_bf16_4x16_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_4x16_mfma_r512_p2_nt", 2, 4, 16, 0, 0, 512, 2, 1, 65536, launch_tsm_atb_bf16_4x16_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_4x16_mfma_r1024_p1_nt", 2, 4, 16, 0, 0, 1024, 1, 1, 131072, launch_tsm_atb_bf16_4x16_mfma_r1024_p1_nt, 2},
  {"tsm_atb_bf16_4x16_mfma_r1024_p2_nt", 2, 4, 16, 0, 0, 1024, 2, 1, 131072, launch_tsm_atb_bf16_4x16_mfma_r1024_p2_nt, 2},
  {"tsm_atb_bf16_4x16_t4x4_r512_p2_nt", 2, 4, 16, 4, 4, 512, 2, 1, 40960, launch_tsm_atb_bf16_4x16_t4x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_4x32_mfma_r128_p1_nt", 2, 4, 32, 0, 0, 128, 1, 1, 24576, launch_tsm_atb_bf16_4x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_4x32_mfma_r128_p2_nt", 2, 4, 32, 0, 0, 128, 2, 1, 24576, launch_tsm_atb_bf16_4x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_4x32_mfma_r256_p1_nt", 2, 4, 32, 0, 0, 256, 1, 1, 49152, launch_tsm_atb_bf16_4x32_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_4x32_mfma_r256_p2_nt", 2, 4, 32, 0, 0, 256, 2, 1, 49152, launch_tsm_atb_bf16_4x32_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_4x32_mfma_r512_p1_nt", 2, 4, 32, 0, 0, 512, 1, 1, 98304, launch_tsm_atb_bf16_4x32_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_4x32_mfma_r512_p2_nt", 2, 4, 32, 0, 0, 512, 2, 1, 98304, launch_tsm_atb_bf16_4x32_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_4x32_t4x4_r256_p2_nt", 2, 4, 32, 4, 4, 256, 2, 1, 36864, launch_tsm_atb_bf16_4x32_t4x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_4x64_mfma_r128_p1_nt", 2, 4, 64, 0, 0, 128, 1, 1, 40960, launch_tsm_atb_bf16_4x64_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_4x64_mfma_r128_p2_nt", 2, 4, 64, 0, 0, 128, 2, 1, 40960, launch_tsm_atb_bf16_4x64_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_4x64_mfma_r256_p1_nt", 2, 4, 64, 0, 0, 256, 1, 1, 81920, launch_tsm_atb_bf16_4x64_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_4x64_mfma_r256_p2_nt", 2, 4, 64, 0, 0, 256, 2, 1, 81920, launch_tsm_atb_bf16_4x64_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_4x64_t4x4_r128_p2_nt", 2, 4, 64, 4, 4, 128, 2, 1, 34816, launch_tsm_atb_bf16_4x64_t4x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_8x1_t8x1_r1024_p2_nt", 2, 8, 1, 8, 1, 1024, 2, 1, 36864, launch_tsm_atb_bf16_8x1_t8x1_r1024_p2_nt, 1},
  {"tsm_atb_bf16_8x2_t8x2_r1024_p2_nt", 2, 8, 2, 8, 2, 1024, 2, 1, 40960, launch_tsm_atb_bf16_8x2_t8x2_r1024_p2_nt, 1},
  {"tsm_atb_bf16_8x4_t8x4_r1024_p2_nt", 2, 8, 4, 8, 4, 1024, 2, 1, 49152, launch_tsm_atb_bf16_8x4_t8x4_r1024_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t4x4_r512_p1", 2, 8, 8, 4, 4, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t4x4_r512_p1, 1},
  {"tsm_atb_bf16_8x8_t4x4_r512_p2", 2, 8, 8, 4, 4, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t4x4_r512_p2, 1},
  {"tsm_atb_bf16_8x8_t4x4_r512_p2_nt", 2, 8, 8, 4, 4, 512, 2, 1, 32768, launch_tsm_atb_bf16_8x8_t4x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t4x4_r512_p4_nt", 2, 8, 8, 4, 4, 512, 4, 1, 32768, launch_tsm_atb_bf16_8x8_t4x4_r512_p4_nt, 1},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p1", 2, 8, 8, 4, 4, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p1, 1},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p2", 2, 8, 8, 4, 4, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p2, 1},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p2_nt", 2, 8, 8, 4, 4, 1024, 2, 1, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p4_nt", 2, 8, 8, 4, 4, 1024, 4, 1, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p4_nt, 1},
  {"tsm_atb_bf16_8x8_t4x8_r512_p1", 2, 8, 8, 4, 8, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t4x8_r512_p1, 1},
  {"tsm_atb_bf16_8x8_t4x8_r512_p2", 2, 8, 8, 4, 8, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t4x8_r512_p2, 1},
  {"tsm_atb_bf16_8x8_t4x8_r512_p2_nt", 2, 8, 8, 4, 8, 512, 2, 1, 32768, launch_tsm_atb_bf16_8x8_t4x8_r512_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t4x8_r512_p4_nt", 2, 8, 8, 4, 8, 512, 4, 1, 32768, launch_tsm_atb_bf16_8x8_t4x8_r512_p4_nt, 1},
  {"tsm_atb_bf16_8x8_t4x8_r1024_p1", 2, 8, 8, 4, 8, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t4x8_r1024_p1, 1},
  {"tsm_atb_bf16_8x8_t4x8_r1024_p2", 2, 8, 8, 4, 8, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t4x8_r1024_p2, 1},
  {"tsm_atb_bf16_8x8_t4x8_r1024_p2_nt", 2, 8, 8, 4, 8, 1024, 2, 1, 65536, launch_tsm_atb_bf16_8x8_t4x8_r1024_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t4x8_r1024_p4_nt", 2, 8, 8, 4, 8, 1024, 4, 1, 65536, launch_tsm_atb_bf16_8x8_t4x8_r1024_p4_nt, 1},
  {"tsm_atb_bf16_8x8_t4x8_r2048_p1", 2, 8, 8, 4, 8, 2048, 1, 0, 131072, launch_tsm_atb_bf16_8x8_t4x8_r2048_p1, 1},
  {"tsm_atb_bf16_8x8_t4x8_r2048_p2", 2, 8, 8, 4, 8, 2048, 2, 0, 131072, launch_tsm_atb_bf16_8x8_t4x8_r2048_p2, 1},
  {"tsm_atb_bf16_8x8_t4x8_r2048_p2_nt", 2, 8, 8, 4, 8, 2048, 2, 1, 131072, launch_tsm_atb_bf16_8x8_t4x8_r2048_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p1", 2, 8, 8, 8, 4, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p1, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p1_nt", 2, 8, 8, 8, 4, 512, 1, 1, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p1_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p2", 2, 8, 8, 8, 4, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p2, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p2_nt", 2, 8, 8, 8, 4, 512, 2, 1, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p4_nt", 2, 8, 8, 8, 4, 512, 4, 1, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p4_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p1", 2, 8, 8, 8, 4, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p1, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p1_nt", 2, 8, 8, 8, 4, 1024, 1, 1, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p1_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p2", 2, 8, 8, 8, 4, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p2, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p2_nt", 2, 8, 8, 8, 4, 1024, 2, 1, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p4_nt", 2, 8, 8, 8, 4, 1024, 4, 1, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p4_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p1", 2, 8, 8, 8, 4, 2048, 1, 0, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p1, 1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p1_nt", 2, 8, 8, 8, 4, 2048, 1, 1, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p1_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p2", 2, 8, 8, 8, 4, 2048, 2, 0, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p2, 1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p2_nt", 2, 8, 8, 8, 4, 2048, 2, 1, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x8_r512_p1", 2, 8, 8, 8, 8, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p1, 1},
  {"tsm_atb_bf16_8x8_t8x8_r512_p2", 2, 8, 8, 8, 8, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p2, 1},
  {"tsm_atb_bf16_8x8_t8x8_r512_p2_nt", 2, 8, 8, 8, 8, 512, 2, 1, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p1", 2, 8, 8, 8, 8, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p1, 1},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p2", 2, 8, 8, 8, 8, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p2, 1},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p2_nt", 2, 8, 8, 8, 8, 1024, 2, 1, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x8_r2048_p1", 2, 8, 8, 8, 8, 2048, 1, 0, 131072, launch_tsm_atb_bf16_8x8_t8x8_r2048_p1, 1},
  {"tsm_atb_bf16_8x8_t8x8_r2048_p2", 2, 8, 8, 8, 8, 2048, 2, 0, 131072, launch_tsm_atb_bf16_8x8_t8x8_r2048_p2, 1},
  {"tsm_atb_bf16_8x8_t8x8_r2048_p2_nt", 2, 8, 8, 8, 8, 2048, 2, 1, 131072, launch_tsm_atb_bf16_8x8_t8x8_r2048_p2_nt, 1},
  {"tsm_atb_bf16_8x16_mfma_r256_p1_nt", 2, 8, 16, 0, 0, 256, 1, 1, 32768, launch_tsm_atb_bf16_8x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_8x16_mfma_r256_p2_nt", 2, 8, 16, 0, 0, 256, 2, 1, 32768, launch_tsm_atb_bf16_8x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_8x16_mfma_r512_p1_nt", 2, 8, 16, 0, 0, 512, 1, 1, 65536, launch_tsm_atb_bf16_8x16_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_8x16_mfma_r512_p2_nt", 2, 8, 16, 0, 0, 512, 2, 1, 65536, launch_tsm_atb_bf16_8x16_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_8x16_mfma_r1024_p1_nt", 2, 8, 16, 0, 0, 1024, 1, 1, 131072, launch_tsm_atb_bf16_8x16_mfma_r1024_p1_nt, 2},
  {"tsm_atb_bf16_8x16_mfma_r1024_p2_nt", 2, 8, 16, 0, 0, 1024, 2, 1, 131072, launch_tsm_atb_bf16_8x16_mfma_r1024_p2_nt, 2},
  {"tsm_atb_bf16_8x16_t8x4_r512_p2_nt", 2, 8, 16, 8, 4, 512, 2, 1, 49152, launch_tsm_atb_bf16_8x16_t8x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_8x32_mfma_r128_p1_nt", 2, 8, 32, 0, 0, 128, 1, 1, 24576, launch_tsm_atb_bf16_8x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_8x32_mfma_r128_p2_nt", 2, 8, 32, 0, 0, 128, 2, 1, 24576, launch_tsm_atb_bf16_8x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_8x32_mfma_r256_p1_nt", 2, 8, 32, 0, 0, 256, 1, 1, 49152, launch_tsm_atb_bf16_8x32_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_8x32_mfma_r256_p2_nt", 2, 8, 32, 0, 0, 256, 2, 1, 49152, launch_tsm_atb_bf16_8x32_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_8x32_mfma_r512_p1_nt", 2, 8, 32, 0, 0, 512, 1, 1, 98304, launch_tsm_atb_bf16_8x32_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_8x32_mfma_r512_p2_nt", 2, 8, 32, 0, 0, 512, 2, 1, 98304, launch_tsm_atb_bf16_8x32_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_8x32_t8x4_r256_p2_nt", 2, 8, 32, 8, 4, 256, 2, 1, 40960, launch_tsm_atb_bf16_8x32_t8x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_8x64_mfma_r128_p1_nt", 2, 8, 64, 0, 0, 128, 1, 1, 40960, launch_tsm_atb_bf16_8x64_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_8x64_mfma_r128_p2_nt", 2, 8, 64, 0, 0, 128, 2, 1, 40960, launch_tsm_atb_bf16_8x64_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_8x64_mfma_r256_p1_nt", 2, 8, 64, 0, 0, 256, 1, 1, 81920, launch_tsm_atb_bf16_8x64_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_8x64_mfma_r256_p2_nt", 2, 8, 64, 0, 0, 256, 2, 1, 81920, launch_tsm_atb_bf16_8x64_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_8x64_t8x4_r128_p2_nt", 2, 8, 64, 8, 4, 128, 2, 1, 36864, launch_tsm_atb_bf16_8x64_t8x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_16x1_t8x1_r512_p2_nt", 2, 16, 1, 8, 1, 512, 2, 1, 34816, launch_tsm_atb_bf16_16x1_t8x1_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x2_t8x2_r512_p2_nt", 2, 16, 2, 8, 2, 512, 2, 1, 36864, launch_tsm_atb_bf16_16x2_t8x2_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x4_mfma_r256_p1_nt", 2, 16, 4, 0, 0, 256, 1, 1, 32768, launch_tsm_atb_bf16_16x4_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_16x4_mfma_r256_p2_nt", 2, 16, 4, 0, 0, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x4_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_16x4_mfma_r512_p1_nt", 2, 16, 4, 0, 0, 512, 1, 1, 65536, launch_tsm_atb_bf16_16x4_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_16x4_mfma_r512_p2_nt", 2, 16, 4, 0, 0, 512, 2, 1, 65536, launch_tsm_atb_bf16_16x4_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_16x4_mfma_r1024_p1_nt", 2, 16, 4, 0, 0, 1024, 1, 1, 131072, launch_tsm_atb_bf16_16x4_mfma_r1024_p1_nt, 2},
  {"tsm_atb_bf16_16x4_mfma_r1024_p2_nt", 2, 16, 4, 0, 0, 1024, 2, 1, 131072, launch_tsm_atb_bf16_16x4_mfma_r1024_p2_nt, 2},
  {"tsm_atb_bf16_16x4_t8x4_r512_p2_nt", 2, 16, 4, 8, 4, 512, 2, 1, 40960, launch_tsm_atb_bf16_16x4_t8x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x8_mfma_r256_p1_nt", 2, 16, 8, 0, 0, 256, 1, 1, 32768, launch_tsm_atb_bf16_16x8_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_16x8_mfma_r256_p2_nt", 2, 16, 8, 0, 0, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x8_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_16x8_mfma_r512_p1_nt", 2, 16, 8, 0, 0, 512, 1, 1, 65536, launch_tsm_atb_bf16_16x8_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_16x8_mfma_r512_p2_nt", 2, 16, 8, 0, 0, 512, 2, 1, 65536, launch_tsm_atb_bf16_16x8_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_16x8_mfma_r1024_p1_nt", 2, 16, 8, 0, 0, 1024, 1, 1, 131072, launch_tsm_atb_bf16_16x8_mfma_r1024_p1_nt, 2},
  {"tsm_atb_bf16_16x8_mfma_r1024_p2_nt", 2, 16, 8, 0, 0, 1024, 2, 1, 131072, launch_tsm_atb_bf16_16x8_mfma_r1024_p2_nt, 2},
  {"tsm_atb_bf16_16x8_t8x4_r512_p2_nt", 2, 16, 8, 8, 4, 512, 2, 1, 49152, launch_tsm_atb_bf16_16x8_t8x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x16_direct_d2_nt", 2, 16, 16, 0, 0, 32, 2, 1, 16384, launch_tsm_atb_bf16_16x16_direct_d2_nt, 3},
  {"tsm_atb_bf16_16x16_direct_d4_nt", 2, 16, 16, 0, 0, 32, 4, 1, 16384, launch_tsm_atb_bf16_16x16_direct_d4_nt, 3},
  {"tsm_atb_bf16_16x16_mfma_r256_p1", 2, 16, 16, 0, 0, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_mfma_r256_p1, 2},
  {"tsm_atb_bf16_16x16_mfma_r256_p1_nt", 2, 16, 16, 0, 0, 256, 1, 1, 32768, launch_tsm_atb_bf16_16x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_16x16_mfma_r256_p2", 2, 16, 16, 0, 0, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_mfma_r256_p2, 2},
  {"tsm_atb_bf16_16x16_mfma_r256_p2_nt", 2, 16, 16, 0, 0, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_16x16_mfma_r512_p1", 2, 16, 16, 0, 0, 512, 1, 0, 65536, launch_tsm_atb_bf16_16x16_mfma_r512_p1, 2},
  {"tsm_atb_bf16_16x16_mfma_r512_p1_nt", 2, 16, 16, 0, 0, 512, 1, 1, 65536, launch_tsm_atb_bf16_16x16_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_16x16_mfma_r512_p2", 2, 16, 16, 0, 0, 512, 2, 0, 65536, launch_tsm_atb_bf16_16x16_mfma_r512_p2, 2},
  {"tsm_atb_bf16_16x16_mfma_r512_p2_nt", 2, 16, 16, 0, 0, 512, 2, 1, 65536, launch_tsm_atb_bf16_16x16_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_16x16_mfma_r1024_p1", 2, 16, 16, 0, 0, 1024, 1, 0, 131072, launch_tsm_atb_bf16_16x16_mfma_r1024_p1, 2},
  {"tsm_atb_bf16_16x16_mfma_r1024_p1_nt", 2, 16, 16, 0, 0, 1024, 1, 1, 131072, launch_tsm_atb_bf16_16x16_mfma_r1024_p1_nt, 2},
  {"tsm_atb_bf16_16x16_mfma_r1024_p2", 2, 16, 16, 0, 0, 1024, 2, 0, 131072, launch_tsm_atb_bf16_16x16_mfma_r1024_p2, 2},
  {"tsm_atb_bf16_16x16_mfma_r1024_p2_nt", 2, 16, 16, 0, 0, 1024, 2, 1, 131072, launch_tsm_atb_bf16_16x16_mfma_r1024_p2_nt, 2},
  {"tsm_atb_bf16_16x16_t4x4_r128_p1", 2, 16, 16, 4, 4, 128, 1, 0, 16384, launch_tsm_atb_bf16_16x16_t4x4_r128_p1, 1},
  {"tsm_atb_bf16_16x16_t4x4_r128_p2", 2, 16, 16, 4, 4, 128, 2, 0, 16384, launch_tsm_atb_bf16_16x16_t4x4_r128_p2, 1},
  {"tsm_atb_bf16_16x16_t4x4_r128_p2_nt", 2, 16, 16, 4, 4, 128, 2, 1, 16384, launch_tsm_atb_bf16_16x16_t4x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t4x4_r128_p4_nt", 2, 16, 16, 4, 4, 128, 4, 1, 16384, launch_tsm_atb_bf16_16x16_t4x4_r128_p4_nt, 1},
  {"tsm_atb_bf16_16x16_t4x4_r256_p1", 2, 16, 16, 4, 4, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p1, 1},
  {"tsm_atb_bf16_16x16_t4x4_r256_p2", 2, 16, 16, 4, 4, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p2, 1},
  {"tsm_atb_bf16_16x16_t4x4_r256_p2_nt", 2, 16, 16, 4, 4, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t4x4_r256_p4_nt", 2, 16, 16, 4, 4, 256, 4, 1, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p4_nt, 1},
  {"tsm_atb_bf16_16x16_t4x8_r128_p1", 2, 16, 16, 4, 8, 128, 1, 0, 16384, launch_tsm_atb_bf16_16x16_t4x8_r128_p1, 1},
  {"tsm_atb_bf16_16x16_t4x8_r128_p2", 2, 16, 16, 4, 8, 128, 2, 0, 16384, launch_tsm_atb_bf16_16x16_t4x8_r128_p2, 1},
  {"tsm_atb_bf16_16x16_t4x8_r128_p2_nt", 2, 16, 16, 4, 8, 128, 2, 1, 16384, launch_tsm_atb_bf16_16x16_t4x8_r128_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t4x8_r128_p4_nt", 2, 16, 16, 4, 8, 128, 4, 1, 16384, launch_tsm_atb_bf16_16x16_t4x8_r128_p4_nt, 1},
  {"tsm_atb_bf16_16x16_t4x8_r256_p1", 2, 16, 16, 4, 8, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t4x8_r256_p1, 1},
  {"tsm_atb_bf16_16x16_t4x8_r256_p2", 2, 16, 16, 4, 8, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t4x8_r256_p2, 1},
  {"tsm_atb_bf16_16x16_t4x8_r256_p2_nt", 2, 16, 16, 4, 8, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x16_t4x8_r256_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t4x8_r256_p4_nt", 2, 16, 16, 4, 8, 256, 4, 1, 32768, launch_tsm_atb_bf16_16x16_t4x8_r256_p4_nt, 1},
  {"tsm_atb_bf16_16x16_t4x8_r512_p1", 2, 16, 16, 4, 8, 512, 1, 0, 65536, launch_tsm_atb_bf16_16x16_t4x8_r512_p1, 1},
  {"tsm_atb_bf16_16x16_t4x8_r512_p2", 2, 16, 16, 4, 8, 512, 2, 0, 65536, launch_tsm_atb_bf16_16x16_t4x8_r512_p2, 1},
  {"tsm_atb_bf16_16x16_t4x8_r512_p2_nt", 2, 16, 16, 4, 8, 512, 2, 1, 65536, launch_tsm_atb_bf16_16x16_t4x8_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p1", 2, 16, 16, 8, 4, 128, 1, 0, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p1, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p1_nt", 2, 16, 16, 8, 4, 128, 1, 1, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p1_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p2", 2, 16, 16, 8, 4, 128, 2, 0, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p2, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p2_nt", 2, 16, 16, 8, 4, 128, 2, 1, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p4_nt", 2, 16, 16, 8, 4, 128, 4, 1, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p4_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p1", 2, 16, 16, 8, 4, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p1, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p1_nt", 2, 16, 16, 8, 4, 256, 1, 1, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p1_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p2", 2, 16, 16, 8, 4, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p2, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p2_nt", 2, 16, 16, 8, 4, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p4_nt", 2, 16, 16, 8, 4, 256, 4, 1, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p4_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p1", 2, 16, 16, 8, 4, 512, 1, 0, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p1, 1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p1_nt", 2, 16, 16, 8, 4, 512, 1, 1, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p1_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p2", 2, 16, 16, 8, 4, 512, 2, 0, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p2, 1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p2_nt", 2, 16, 16, 8, 4, 512, 2, 1, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x8_r128_p1", 2, 16, 16, 8, 8, 128, 1, 0, 16384, launch_tsm_atb_bf16_16x16_t8x8_r128_p1, 1},
  {"tsm_atb_bf16_16x16_t8x8_r128_p2", 2, 16, 16, 8, 8, 128, 2, 0, 16384, launch_tsm_atb_bf16_16x16_t8x8_r128_p2, 1},
  {"tsm_atb_bf16_16x16_t8x8_r128_p2_nt", 2, 16, 16, 8, 8, 128, 2, 1, 16384, launch_tsm_atb_bf16_16x16_t8x8_r128_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x8_r256_p1", 2, 16, 16, 8, 8, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p1, 1},
  {"tsm_atb_bf16_16x16_t8x8_r256_p2", 2, 16, 16, 8, 8, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p2, 1},
  {"tsm_atb_bf16_16x16_t8x8_r256_p2_nt", 2, 16, 16, 8, 8, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x8_r512_p1", 2, 16, 16, 8, 8, 512, 1, 0, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p1, 1},
  {"tsm_atb_bf16_16x16_t8x8_r512_p2", 2, 16, 16, 8, 8, 512, 2, 0, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p2, 1},
  {"tsm_atb_bf16_16x16_t8x8_r512_p2_nt", 2, 16, 16, 8, 8, 512, 2, 1, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x32_direct_d4_nt", 2, 16, 32, 0, 0, 32, 4, 1, 24576, launch_tsm_atb_bf16_16x32_direct_d4_nt, 3},
  {"tsm_atb_bf16_16x32_mfma_r256_p2_nt", 2, 16, 32, 0, 0, 256, 2, 1, 49152, launch_tsm_atb_bf16_16x32_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_16x64_direct_d4_nt", 2, 16, 64, 0, 0, 32, 4, 1, 40960, launch_tsm_atb_bf16_16x64_direct_d4_nt, 3},
  {"tsm_atb_bf16_16x64_mfma_r128_p2_nt", 2, 16, 64, 0, 0, 128, 2, 1, 40960, launch_tsm_atb_bf16_16x64_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_32x1_t8x1_r256_p2_nt", 2, 32, 1, 8, 1, 256, 2, 1, 33792, launch_tsm_atb_bf16_32x1_t8x1_r256_p2_nt, 1},
  {"tsm_atb_bf16_32x2_t8x2_r256_p2_nt", 2, 32, 2, 8, 2, 256, 2, 1, 34816, launch_tsm_atb_bf16_32x2_t8x2_r256_p2_nt, 1},
  {"tsm_atb_bf16_32x4_mfma_r128_p1_nt", 2, 32, 4, 0, 0, 128, 1, 1, 24576, launch_tsm_atb_bf16_32x4_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_32x4_mfma_r128_p2_nt", 2, 32, 4, 0, 0, 128, 2, 1, 24576, launch_tsm_atb_bf16_32x4_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_32x4_mfma_r256_p1_nt", 2, 32, 4, 0, 0, 256, 1, 1, 49152, launch_tsm_atb_bf16_32x4_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_32x4_mfma_r256_p2_nt", 2, 32, 4, 0, 0, 256, 2, 1, 49152, launch_tsm_atb_bf16_32x4_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_32x4_mfma_r512_p1_nt", 2, 32, 4, 0, 0, 512, 1, 1, 98304, launch_tsm_atb_bf16_32x4_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_32x4_mfma_r512_p2_nt", 2, 32, 4, 0, 0, 512, 2, 1, 98304, launch_tsm_atb_bf16_32x4_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_32x4_t8x4_r256_p2_nt", 2, 32, 4, 8, 4, 256, 2, 1, 36864, launch_tsm_atb_bf16_32x4_t8x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_32x8_mfma_r128_p1_nt", 2, 32, 8, 0, 0, 128, 1, 1, 24576, launch_tsm_atb_bf16_32x8_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_32x8_mfma_r128_p2_nt", 2, 32, 8, 0, 0, 128, 2, 1, 24576, launch_tsm_atb_bf16_32x8_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_32x8_mfma_r256_p1_nt", 2, 32, 8, 0, 0, 256, 1, 1, 49152, launch_tsm_atb_bf16_32x8_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_32x8_mfma_r256_p2_nt", 2, 32, 8, 0, 0, 256, 2, 1, 49152, launch_tsm_atb_bf16_32x8_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_32x8_mfma_r512_p1_nt", 2, 32, 8, 0, 0, 512, 1, 1, 98304, launch_tsm_atb_bf16_32x8_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_32x8_mfma_r512_p2_nt", 2, 32, 8, 0, 0, 512, 2, 1, 98304, launch_tsm_atb_bf16_32x8_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_32x8_t8x4_r256_p2_nt", 2, 32, 8, 8, 4, 256, 2, 1, 40960, launch_tsm_atb_bf16_32x8_t8x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_32x16_direct_d4_nt", 2, 32, 16, 0, 0, 32, 4, 1, 24576, launch_tsm_atb_bf16_32x16_direct_d4_nt, 3},
  {"tsm_atb_bf16_32x16_mfma_r256_p2_nt", 2, 32, 16, 0, 0, 256, 2, 1, 49152, launch_tsm_atb_bf16_32x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_32x32_direct_d2_nt", 2, 32, 32, 0, 0, 32, 2, 1, 32768, launch_tsm_atb_bf16_32x32_direct_d2_nt, 3},
  {"tsm_atb_bf16_32x32_direct_d4_nt", 2, 32, 32, 0, 0, 32, 4, 1, 32768, launch_tsm_atb_bf16_32x32_direct_d4_nt, 3},
  {"tsm_atb_bf16_32x32_mfma_r128_p1", 2, 32, 32, 0, 0, 128, 1, 0, 32768, launch_tsm_atb_bf16_32x32_mfma_r128_p1, 2},
  {"tsm_atb_bf16_32x32_mfma_r128_p1_nt", 2, 32, 32, 0, 0, 128, 1, 1, 32768, launch_tsm_atb_bf16_32x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_32x32_mfma_r128_p2", 2, 32, 32, 0, 0, 128, 2, 0, 32768, launch_tsm_atb_bf16_32x32_mfma_r128_p2, 2},
  {"tsm_atb_bf16_32x32_mfma_r128_p2_nt", 2, 32, 32, 0, 0, 128, 2, 1, 32768, launch_tsm_atb_bf16_32x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_32x32_mfma_r256_p1", 2, 32, 32, 0, 0, 256, 1, 0, 65536, launch_tsm_atb_bf16_32x32_mfma_r256_p1, 2},
  {"tsm_atb_bf16_32x32_mfma_r256_p1_nt", 2, 32, 32, 0, 0, 256, 1, 1, 65536, launch_tsm_atb_bf16_32x32_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_32x32_mfma_r256_p2", 2, 32, 32, 0, 0, 256, 2, 0, 65536, launch_tsm_atb_bf16_32x32_mfma_r256_p2, 2},
  {"tsm_atb_bf16_32x32_mfma_r256_p2_nt", 2, 32, 32, 0, 0, 256, 2, 1, 65536, launch_tsm_atb_bf16_32x32_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_32x32_mfma_r512_p1", 2, 32, 32, 0, 0, 512, 1, 0, 131072, launch_tsm_atb_bf16_32x32_mfma_r512_p1, 2},
  {"tsm_atb_bf16_32x32_mfma_r512_p1_nt", 2, 32, 32, 0, 0, 512, 1, 1, 131072, launch_tsm_atb_bf16_32x32_mfma_r512_p1_nt, 2},
  {"tsm_atb_bf16_32x32_mfma_r512_p2", 2, 32, 32, 0, 0, 512, 2, 0, 131072, launch_tsm_atb_bf16_32x32_mfma_r512_p2, 2},
  {"tsm_atb_bf16_32x32_mfma_r512_p2_nt", 2, 32, 32, 0, 0, 512, 2, 1, 131072, launch_tsm_atb_bf16_32x32_mfma_r512_p2_nt, 2},
  {"tsm_atb_bf16_32x32_t4x4_r32_p1", 2, 32, 32, 4, 4, 32, 1, 0, 8192, launch_tsm_atb_bf16_32x32_t4x4_r32_p1, 1},
  {"tsm_atb_bf16_32x32_t4x4_r32_p2", 2, 32, 32, 4, 4, 32, 2, 0, 8192, launch_tsm_atb_bf16_32x32_t4x4_r32_p2, 1},
  {"tsm_atb_bf16_32x32_t4x4_r32_p2_nt", 2, 32, 32, 4, 4, 32, 2, 1, 8192, launch_tsm_atb_bf16_32x32_t4x4_r32_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t4x4_r32_p4_nt", 2, 32, 32, 4, 4, 32, 4, 1, 8192, launch_tsm_atb_bf16_32x32_t4x4_r32_p4_nt, 1},
  {"tsm_atb_bf16_32x32_t4x4_r64_p1", 2, 32, 32, 4, 4, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t4x4_r64_p1, 1},
  {"tsm_atb_bf16_32x32_t4x4_r64_p2", 2, 32, 32, 4, 4, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t4x4_r64_p2, 1},
  {"tsm_atb_bf16_32x32_t4x4_r64_p2_nt", 2, 32, 32, 4, 4, 64, 2, 1, 16384, launch_tsm_atb_bf16_32x32_t4x4_r64_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t4x4_r64_p4_nt", 2, 32, 32, 4, 4, 64, 4, 1, 16384, launch_tsm_atb_bf16_32x32_t4x4_r64_p4_nt, 1},
  {"tsm_atb_bf16_32x32_t4x8_r32_p1", 2, 32, 32, 4, 8, 32, 1, 0, 8192, launch_tsm_atb_bf16_32x32_t4x8_r32_p1, 1},
  {"tsm_atb_bf16_32x32_t4x8_r32_p2", 2, 32, 32, 4, 8, 32, 2, 0, 8192, launch_tsm_atb_bf16_32x32_t4x8_r32_p2, 1},
  {"tsm_atb_bf16_32x32_t4x8_r32_p2_nt", 2, 32, 32, 4, 8, 32, 2, 1, 8192, launch_tsm_atb_bf16_32x32_t4x8_r32_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t4x8_r32_p4_nt", 2, 32, 32, 4, 8, 32, 4, 1, 8192, launch_tsm_atb_bf16_32x32_t4x8_r32_p4_nt, 1},
  {"tsm_atb_bf16_32x32_t4x8_r64_p1", 2, 32, 32, 4, 8, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t4x8_r64_p1, 1},
  {"tsm_atb_bf16_32x32_t4x8_r64_p2", 2, 32, 32, 4, 8, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t4x8_r64_p2, 1},
  {"tsm_atb_bf16_32x32_t4x8_r64_p2_nt", 2, 32, 32, 4, 8, 64, 2, 1, 16384, launch_tsm_atb_bf16_32x32_t4x8_r64_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t4x8_r64_p4_nt", 2, 32, 32, 4, 8, 64, 4, 1, 16384, launch_tsm_atb_bf16_32x32_t4x8_r64_p4_nt, 1},
  {"tsm_atb_bf16_32x32_t4x8_r128_p1", 2, 32, 32, 4, 8, 128, 1, 0, 32768, launch_tsm_atb_bf16_32x32_t4x8_r128_p1, 1},
  {"tsm_atb_bf16_32x32_t4x8_r128_p2", 2, 32, 32, 4, 8, 128, 2, 0, 32768, launch_tsm_atb_bf16_32x32_t4x8_r128_p2, 1},
  {"tsm_atb_bf16_32x32_t4x8_r128_p2_nt", 2, 32, 32, 4, 8, 128, 2, 1, 32768, launch_tsm_atb_bf16_32x32_t4x8_r128_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p1", 2, 32, 32, 8, 4, 32, 1, 0, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p1, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p1_nt", 2, 32, 32, 8, 4, 32, 1, 1, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p1_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p2", 2, 32, 32, 8, 4, 32, 2, 0, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p2, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p2_nt", 2, 32, 32, 8, 4, 32, 2, 1, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p4_nt", 2, 32, 32, 8, 4, 32, 4, 1, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p4_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p1", 2, 32, 32, 8, 4, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p1, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p1_nt", 2, 32, 32, 8, 4, 64, 1, 1, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p1_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p2", 2, 32, 32, 8, 4, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p2, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p2_nt", 2, 32, 32, 8, 4, 64, 2, 1, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p4_nt", 2, 32, 32, 8, 4, 64, 4, 1, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p4_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p1", 2, 32, 32, 8, 4, 128, 1, 0, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p1, 1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p1_nt", 2, 32, 32, 8, 4, 128, 1, 1, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p1_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p2", 2, 32, 32, 8, 4, 128, 2, 0, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p2, 1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p2_nt", 2, 32, 32, 8, 4, 128, 2, 1, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x8_r32_p1", 2, 32, 32, 8, 8, 32, 1, 0, 8192, launch_tsm_atb_bf16_32x32_t8x8_r32_p1, 1},
  {"tsm_atb_bf16_32x32_t8x8_r32_p2", 2, 32, 32, 8, 8, 32, 2, 0, 8192, launch_tsm_atb_bf16_32x32_t8x8_r32_p2, 1},
  {"tsm_atb_bf16_32x32_t8x8_r32_p2_nt", 2, 32, 32, 8, 8, 32, 2, 1, 8192, launch_tsm_atb_bf16_32x32_t8x8_r32_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x8_r64_p1", 2, 32, 32, 8, 8, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t8x8_r64_p1, 1},
  {"tsm_atb_bf16_32x32_t8x8_r64_p2", 2, 32, 32, 8, 8, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t8x8_r64_p2, 1},
  {"tsm_atb_bf16_32x32_t8x8_r64_p2_nt", 2, 32, 32, 8, 8, 64, 2, 1, 16384, launch_tsm_atb_bf16_32x32_t8x8_r64_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x8_r128_p1", 2, 32, 32, 8, 8, 128, 1, 0, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p1, 1},
  {"tsm_atb_bf16_32x32_t8x8_r128_p2", 2, 32, 32, 8, 8, 128, 2, 0, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p2, 1},
  {"tsm_atb_bf16_32x32_t8x8_r128_p2_nt", 2, 32, 32, 8, 8, 128, 2, 1, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p2_nt, 1},
  {"tsm_atb_bf16_32x64_direct_d4_nt", 2, 32, 64, 0, 0, 32, 4, 1, 49152, launch_tsm_atb_bf16_32x64_direct_d4_nt, 3},
  {"tsm_atb_bf16_32x64_mfma_r128_p2_nt", 2, 32, 64, 0, 0, 128, 2, 1, 49152, launch_tsm_atb_bf16_32x64_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_64x1_t8x1_r128_p2_nt", 2, 64, 1, 8, 1, 128, 2, 1, 33280, launch_tsm_atb_bf16_64x1_t8x1_r128_p2_nt, 1},
  {"tsm_atb_bf16_64x2_t8x2_r128_p2_nt", 2, 64, 2, 8, 2, 128, 2, 1, 33792, launch_tsm_atb_bf16_64x2_t8x2_r128_p2_nt, 1},
  {"tsm_atb_bf16_64x4_mfma_r128_p1_nt", 2, 64, 4, 0, 0, 128, 1, 1, 40960, launch_tsm_atb_bf16_64x4_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_64x4_mfma_r128_p2_nt", 2, 64, 4, 0, 0, 128, 2, 1, 40960, launch_tsm_atb_bf16_64x4_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_64x4_mfma_r256_p1_nt", 2, 64, 4, 0, 0, 256, 1, 1, 81920, launch_tsm_atb_bf16_64x4_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_64x4_mfma_r256_p2_nt", 2, 64, 4, 0, 0, 256, 2, 1, 81920, launch_tsm_atb_bf16_64x4_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_64x4_t8x4_r128_p2_nt", 2, 64, 4, 8, 4, 128, 2, 1, 34816, launch_tsm_atb_bf16_64x4_t8x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_64x8_mfma_r128_p1_nt", 2, 64, 8, 0, 0, 128, 1, 1, 40960, launch_tsm_atb_bf16_64x8_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_64x8_mfma_r128_p2_nt", 2, 64, 8, 0, 0, 128, 2, 1, 40960, launch_tsm_atb_bf16_64x8_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_64x8_mfma_r256_p1_nt", 2, 64, 8, 0, 0, 256, 1, 1, 81920, launch_tsm_atb_bf16_64x8_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_64x8_mfma_r256_p2_nt", 2, 64, 8, 0, 0, 256, 2, 1, 81920, launch_tsm_atb_bf16_64x8_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_64x8_t8x4_r128_p2_nt", 2, 64, 8, 8, 4, 128, 2, 1, 36864, launch_tsm_atb_bf16_64x8_t8x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_64x16_direct_d4_nt", 2, 64, 16, 0, 0, 32, 4, 1, 40960, launch_tsm_atb_bf16_64x16_direct_d4_nt, 3},
  {"tsm_atb_bf16_64x16_mfma_r128_p2_nt", 2, 64, 16, 0, 0, 128, 2, 1, 40960, launch_tsm_atb_bf16_64x16_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_64x32_direct_d4_nt", 2, 64, 32, 0, 0, 32, 4, 1, 49152, launch_tsm_atb_bf16_64x32_direct_d4_nt, 3},
  {"tsm_atb_bf16_64x32_mfma_r128_p2_nt", 2, 64, 32, 0, 0, 128, 2, 1, 49152, launch_tsm_atb_bf16_64x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_64x64_direct_d2_nt", 2, 64, 64, 0, 0, 32, 2, 1, 65536, launch_tsm_atb_bf16_64x64_direct_d2_nt, 3},
  {"tsm_atb_bf16_64x64_direct_d4_nt", 2, 64, 64, 0, 0, 32, 4, 1, 65536, launch_tsm_atb_bf16_64x64_direct_d4_nt, 3},
  {"tsm_atb_bf16_64x64_mfma_r128_p1", 2, 64, 64, 0, 0, 128, 1, 0, 65536, launch_tsm_atb_bf16_64x64_mfma_r128_p1, 2},
  {"tsm_atb_bf16_64x64_mfma_r128_p1_nt", 2, 64, 64, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_bf16_64x64_mfma_r128_p1_nt, 2},
  {"tsm_atb_bf16_64x64_mfma_r128_p2", 2, 64, 64, 0, 0, 128, 2, 0, 65536, launch_tsm_atb_bf16_64x64_mfma_r128_p2, 2},
  {"tsm_atb_bf16_64x64_mfma_r128_p2_nt", 2, 64, 64, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_bf16_64x64_mfma_r128_p2_nt, 2},
  {"tsm_atb_bf16_64x64_mfma_r256_p1", 2, 64, 64, 0, 0, 256, 1, 0, 131072, launch_tsm_atb_bf16_64x64_mfma_r256_p1, 2},
  {"tsm_atb_bf16_64x64_mfma_r256_p1_nt", 2, 64, 64, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_bf16_64x64_mfma_r256_p1_nt, 2},
  {"tsm_atb_bf16_64x64_mfma_r256_p2", 2, 64, 64, 0, 0, 256, 2, 0, 131072, launch_tsm_atb_bf16_64x64_mfma_r256_p2, 2},
  {"tsm_atb_bf16_64x64_mfma_r256_p2_nt", 2, 64, 64, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_bf16_64x64_mfma_r256_p2_nt, 2},
  {"tsm_atb_bf16_64x64_t8x8_r16_p1", 2, 64, 64, 8, 8, 16, 1, 0, 32768, launch_tsm_atb_bf16_64x64_t8x8_r16_p1, 1},
  {"tsm_atb_bf16_64x64_t8x8_r16_p1_nt", 2, 64, 64, 8, 8, 16, 1, 1, 32768, launch_tsm_atb_bf16_64x64_t8x8_r16_p1_nt, 1},
  {"tsm_atb_bf16_64x64_t8x8_r16_p2", 2, 64, 64, 8, 8, 16, 2, 0, 32768, launch_tsm_atb_bf16_64x64_t8x8_r16_p2, 1},
  {"tsm_atb_bf16_64x64_t8x8_r16_p2_nt", 2, 64, 64, 8, 8, 16, 2, 1, 32768, launch_tsm_atb_bf16_64x64_t8x8_r16_p2_nt, 1},
  {"tsm_atb_bf16_64x64_t8x8_r32_p1", 2, 64, 64, 8, 8, 32, 1, 0, 32768, launch_tsm_atb_bf16_64x64_t8x8_r32_p1, 1},
  {"tsm_atb_bf16_64x64_t8x8_r32_p1_nt", 2, 64, 64, 8, 8, 32, 1, 1, 32768, launch_tsm_atb_bf16_64x64_t8x8_r32_p1_nt, 1},
  {"tsm_atb_bf16_64x64_t8x8_r32_p2", 2, 64, 64, 8, 8, 32, 2, 0, 32768, launch_tsm_atb_bf16_64x64_t8x8_r32_p2, 1},
  {"tsm_atb_bf16_64x64_t8x8_r32_p2_nt", 2, 64, 64, 8, 8, 32, 2, 1, 32768, launch_tsm_atb_bf16_64x64_t8x8_r32_p2_nt, 1},
  {"tsm_atb_bf16_64x64_t8x8_r64_p1", 2, 64, 64, 8, 8, 64, 1, 0, 32768, launch_tsm_atb_bf16_64x64_t8x8_r64_p1, 1},
  {"tsm_atb_bf16_64x64_t8x8_r64_p1_nt", 2, 64, 64, 8, 8, 64, 1, 1, 32768, launch_tsm_atb_bf16_64x64_t8x8_r64_p1_nt, 1},
  {"tsm_atb_bf16_64x64_t8x8_r64_p2", 2, 64, 64, 8, 8, 64, 2, 0, 32768, launch_tsm_atb_bf16_64x64_t8x8_r64_p2, 1},
  {"tsm_atb_bf16_64x64_t8x8_r64_p2_nt", 2, 64, 64, 8, 8, 64, 2, 1, 32768, launch_tsm_atb_bf16_64x64_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f32_1x1_t1x1_r4096_p2_nt", 1, 1, 1, 1, 1, 4096, 2, 1, 65536, launch_tsm_atb_f32_1x1_t1x1_r4096_p2_nt, 1},
  {"tsm_atb_f32_1x2_t1x2_r2048_p2_nt", 1, 1, 2, 1, 2, 2048, 2, 1, 49152, launch_tsm_atb_f32_1x2_t1x2_r2048_p2_nt, 1},
  {"tsm_atb_f32_1x4_t1x4_r1024_p2_nt", 1, 1, 4, 1, 4, 1024, 2, 1, 40960, launch_tsm_atb_f32_1x4_t1x4_r1024_p2_nt, 1},
  {"tsm_atb_f32_1x8_t1x4_r512_p2_nt", 1, 1, 8, 1, 4, 512, 2, 1, 36864, launch_tsm_atb_f32_1x8_t1x4_r512_p2_nt, 1},
  {"tsm_atb_f32_1x16_t1x4_r256_p2_nt", 1, 1, 16, 1, 4, 256, 2, 1, 34816, launch_tsm_atb_f32_1x16_t1x4_r256_p2_nt, 1},
  {"tsm_atb_f32_1x32_t1x4_r128_p2_nt", 1, 1, 32, 1, 4, 128, 2, 1, 33792, launch_tsm_atb_f32_1x32_t1x4_r128_p2_nt, 1},
  {"tsm_atb_f32_1x64_t1x4_r64_p2_nt", 1, 1, 64, 1, 4, 64, 2, 1, 33280, launch_tsm_atb_f32_1x64_t1x4_r64_p2_nt, 1},
  {"tsm_atb_f32_2x1_t2x1_r2048_p2_nt", 1, 2, 1, 2, 1, 2048, 2, 1, 49152, launch_tsm_atb_f32_2x1_t2x1_r2048_p2_nt, 1},
  {"tsm_atb_f32_2x2_t2x2_r2048_p2_nt", 1, 2, 2, 2, 2, 2048, 2, 1, 65536, launch_tsm_atb_f32_2x2_t2x2_r2048_p2_nt, 1},
  {"tsm_atb_f32_2x4_t2x4_r1024_p2_nt", 1, 2, 4, 2, 4, 1024, 2, 1, 49152, launch_tsm_atb_f32_2x4_t2x4_r1024_p2_nt, 1},
  {"tsm_atb_f32_2x8_t2x4_r512_p2_nt", 1, 2, 8, 2, 4, 512, 2, 1, 40960, launch_tsm_atb_f32_2x8_t2x4_r512_p2_nt, 1},
  {"tsm_atb_f32_2x16_t2x4_r256_p2_nt", 1, 2, 16, 2, 4, 256, 2, 1, 36864, launch_tsm_atb_f32_2x16_t2x4_r256_p2_nt, 1},
  {"tsm_atb_f32_2x32_t2x4_r128_p2_nt", 1, 2, 32, 2, 4, 128, 2, 1, 34816, launch_tsm_atb_f32_2x32_t2x4_r128_p2_nt, 1},
  {"tsm_atb_f32_2x64_t2x4_r64_p2_nt", 1, 2, 64, 2, 4, 64, 2, 1, 33792, launch_tsm_atb_f32_2x64_t2x4_r64_p2_nt, 1},
  {"tsm_atb_f32_4x1_t4x1_r1024_p2_nt", 1, 4, 1, 4, 1, 1024, 2, 1, 40960, launch_tsm_atb_f32_4x1_t4x1_r1024_p2_nt, 1},
  {"tsm_atb_f32_4x2_t4x2_r1024_p2_nt", 1, 4, 2, 4, 2, 1024, 2, 1, 49152, launch_tsm_atb_f32_4x2_t4x2_r1024_p2_nt, 1},
  {"tsm_atb_f32_4x4_t4x4_r1024_p2_nt", 1, 4, 4, 4, 4, 1024, 2, 1, 65536, launch_tsm_atb_f32_4x4_t4x4_r1024_p2_nt, 1},
  {"tsm_atb_f32_4x8_t4x4_r512_p2_nt", 1, 4, 8, 4, 4, 512, 2, 1, 49152, launch_tsm_atb_f32_4x8_t4x4_r512_p2_nt, 1},
  {"tsm_atb_f32_4x16_mfma_r128_p1_nt", 1, 4, 16, 0, 0, 128, 1, 1, 32768, launch_tsm_atb_f32_4x16_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_4x16_mfma_r128_p2_nt", 1, 4, 16, 0, 0, 128, 2, 1, 32768, launch_tsm_atb_f32_4x16_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_4x16_mfma_r256_p1_nt", 1, 4, 16, 0, 0, 256, 1, 1, 65536, launch_tsm_atb_f32_4x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_4x16_mfma_r256_p2_nt", 1, 4, 16, 0, 0, 256, 2, 1, 65536, launch_tsm_atb_f32_4x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_4x16_t4x4_r256_p2_nt", 1, 4, 16, 4, 4, 256, 2, 1, 40960, launch_tsm_atb_f32_4x16_t4x4_r256_p2_nt, 1},
  {"tsm_atb_f32_4x32_mfma_r64_p1_nt", 1, 4, 32, 0, 0, 64, 1, 1, 24576, launch_tsm_atb_f32_4x32_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_4x32_mfma_r64_p2_nt", 1, 4, 32, 0, 0, 64, 2, 1, 24576, launch_tsm_atb_f32_4x32_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_4x32_mfma_r128_p1_nt", 1, 4, 32, 0, 0, 128, 1, 1, 49152, launch_tsm_atb_f32_4x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_4x32_mfma_r128_p2_nt", 1, 4, 32, 0, 0, 128, 2, 1, 49152, launch_tsm_atb_f32_4x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_4x32_mfma_r256_p1_nt", 1, 4, 32, 0, 0, 256, 1, 1, 98304, launch_tsm_atb_f32_4x32_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_4x32_mfma_r256_p2_nt", 1, 4, 32, 0, 0, 256, 2, 1, 98304, launch_tsm_atb_f32_4x32_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_4x32_t4x4_r128_p2_nt", 1, 4, 32, 4, 4, 128, 2, 1, 36864, launch_tsm_atb_f32_4x32_t4x4_r128_p2_nt, 1},
  {"tsm_atb_f32_4x64_mfma_r32_p1_nt", 1, 4, 64, 0, 0, 32, 1, 1, 20480, launch_tsm_atb_f32_4x64_mfma_r32_p1_nt, 2},
  {"tsm_atb_f32_4x64_mfma_r32_p2_nt", 1, 4, 64, 0, 0, 32, 2, 1, 20480, launch_tsm_atb_f32_4x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f32_4x64_mfma_r64_p1_nt", 1, 4, 64, 0, 0, 64, 1, 1, 40960, launch_tsm_atb_f32_4x64_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_4x64_mfma_r64_p2_nt", 1, 4, 64, 0, 0, 64, 2, 1, 40960, launch_tsm_atb_f32_4x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_4x64_mfma_r128_p1_nt", 1, 4, 64, 0, 0, 128, 1, 1, 81920, launch_tsm_atb_f32_4x64_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_4x64_mfma_r128_p2_nt", 1, 4, 64, 0, 0, 128, 2, 1, 81920, launch_tsm_atb_f32_4x64_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_4x64_t4x4_r64_p2_nt", 1, 4, 64, 4, 4, 64, 2, 1, 34816, launch_tsm_atb_f32_4x64_t4x4_r64_p2_nt, 1},
  {"tsm_atb_f32_8x1_t8x1_r512_p2_nt", 1, 8, 1, 8, 1, 512, 2, 1, 36864, launch_tsm_atb_f32_8x1_t8x1_r512_p2_nt, 1},
  {"tsm_atb_f32_8x2_t8x2_r512_p2_nt", 1, 8, 2, 8, 2, 512, 2, 1, 40960, launch_tsm_atb_f32_8x2_t8x2_r512_p2_nt, 1},
  {"tsm_atb_f32_8x4_t8x4_r512_p2_nt", 1, 8, 4, 8, 4, 512, 2, 1, 49152, launch_tsm_atb_f32_8x4_t8x4_r512_p2_nt, 1},
  {"tsm_atb_f32_8x8_t4x4_r256_p1", 1, 8, 8, 4, 4, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t4x4_r256_p1, 1},
  {"tsm_atb_f32_8x8_t4x4_r256_p2", 1, 8, 8, 4, 4, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t4x4_r256_p2, 1},
  {"tsm_atb_f32_8x8_t4x4_r256_p2_nt", 1, 8, 8, 4, 4, 256, 2, 1, 32768, launch_tsm_atb_f32_8x8_t4x4_r256_p2_nt, 1},
  {"tsm_atb_f32_8x8_t4x4_r256_p4_nt", 1, 8, 8, 4, 4, 256, 4, 1, 32768, launch_tsm_atb_f32_8x8_t4x4_r256_p4_nt, 1},
  {"tsm_atb_f32_8x8_t4x4_r512_p1", 1, 8, 8, 4, 4, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p1, 1},
  {"tsm_atb_f32_8x8_t4x4_r512_p2", 1, 8, 8, 4, 4, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p2, 1},
  {"tsm_atb_f32_8x8_t4x4_r512_p2_nt", 1, 8, 8, 4, 4, 512, 2, 1, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p2_nt, 1},
  {"tsm_atb_f32_8x8_t4x4_r512_p4_nt", 1, 8, 8, 4, 4, 512, 4, 1, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p4_nt, 1},
  {"tsm_atb_f32_8x8_t4x4_r1024_p1", 1, 8, 8, 4, 4, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t4x4_r1024_p1, 1},
  {"tsm_atb_f32_8x8_t4x4_r1024_p2", 1, 8, 8, 4, 4, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t4x4_r1024_p2, 1},
  {"tsm_atb_f32_8x8_t4x4_r1024_p2_nt", 1, 8, 8, 4, 4, 1024, 2, 1, 131072, launch_tsm_atb_f32_8x8_t4x4_r1024_p2_nt, 1},
  {"tsm_atb_f32_8x8_t4x8_r256_p1", 1, 8, 8, 4, 8, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t4x8_r256_p1, 1},
  {"tsm_atb_f32_8x8_t4x8_r256_p2", 1, 8, 8, 4, 8, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t4x8_r256_p2, 1},
  {"tsm_atb_f32_8x8_t4x8_r256_p2_nt", 1, 8, 8, 4, 8, 256, 2, 1, 32768, launch_tsm_atb_f32_8x8_t4x8_r256_p2_nt, 1},
  {"tsm_atb_f32_8x8_t4x8_r256_p4_nt", 1, 8, 8, 4, 8, 256, 4, 1, 32768, launch_tsm_atb_f32_8x8_t4x8_r256_p4_nt, 1},
  {"tsm_atb_f32_8x8_t4x8_r512_p1", 1, 8, 8, 4, 8, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t4x8_r512_p1, 1},
  {"tsm_atb_f32_8x8_t4x8_r512_p2", 1, 8, 8, 4, 8, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t4x8_r512_p2, 1},
  {"tsm_atb_f32_8x8_t4x8_r512_p2_nt", 1, 8, 8, 4, 8, 512, 2, 1, 65536, launch_tsm_atb_f32_8x8_t4x8_r512_p2_nt, 1},
  {"tsm_atb_f32_8x8_t4x8_r512_p4_nt", 1, 8, 8, 4, 8, 512, 4, 1, 65536, launch_tsm_atb_f32_8x8_t4x8_r512_p4_nt, 1},
  {"tsm_atb_f32_8x8_t4x8_r1024_p1", 1, 8, 8, 4, 8, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t4x8_r1024_p1, 1},
  {"tsm_atb_f32_8x8_t4x8_r1024_p2", 1, 8, 8, 4, 8, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t4x8_r1024_p2, 1},
  {"tsm_atb_f32_8x8_t4x8_r1024_p2_nt", 1, 8, 8, 4, 8, 1024, 2, 1, 131072, launch_tsm_atb_f32_8x8_t4x8_r1024_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p1", 1, 8, 8, 8, 4, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p1, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p1_nt", 1, 8, 8, 8, 4, 256, 1, 1, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p1_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p2", 1, 8, 8, 8, 4, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p2, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p2_nt", 1, 8, 8, 8, 4, 256, 2, 1, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p4_nt", 1, 8, 8, 8, 4, 256, 4, 1, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p4_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p1", 1, 8, 8, 8, 4, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p1, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p1_nt", 1, 8, 8, 8, 4, 512, 1, 1, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p1_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p2", 1, 8, 8, 8, 4, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p2, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p2_nt", 1, 8, 8, 8, 4, 512, 2, 1, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p4_nt", 1, 8, 8, 8, 4, 512, 4, 1, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p4_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p1", 1, 8, 8, 8, 4, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p1, 1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p1_nt", 1, 8, 8, 8, 4, 1024, 1, 1, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p1_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p2", 1, 8, 8, 8, 4, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p2, 1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p2_nt", 1, 8, 8, 8, 4, 1024, 2, 1, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x8_r256_p1", 1, 8, 8, 8, 8, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p1, 1},
  {"tsm_atb_f32_8x8_t8x8_r256_p2", 1, 8, 8, 8, 8, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p2, 1},
  {"tsm_atb_f32_8x8_t8x8_r256_p2_nt", 1, 8, 8, 8, 8, 256, 2, 1, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x8_r512_p1", 1, 8, 8, 8, 8, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p1, 1},
  {"tsm_atb_f32_8x8_t8x8_r512_p2", 1, 8, 8, 8, 8, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p2, 1},
  {"tsm_atb_f32_8x8_t8x8_r512_p2_nt", 1, 8, 8, 8, 8, 512, 2, 1, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x8_r1024_p1", 1, 8, 8, 8, 8, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t8x8_r1024_p1, 1},
  {"tsm_atb_f32_8x8_t8x8_r1024_p2", 1, 8, 8, 8, 8, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t8x8_r1024_p2, 1},
  {"tsm_atb_f32_8x8_t8x8_r1024_p2_nt", 1, 8, 8, 8, 8, 1024, 2, 1, 131072, launch_tsm_atb_f32_8x8_t8x8_r1024_p2_nt, 1},
  {"tsm_atb_f32_8x16_mfma_r128_p1_nt", 1, 8, 16, 0, 0, 128, 1, 1, 32768, launch_tsm_atb_f32_8x16_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_8x16_mfma_r128_p2_nt", 1, 8, 16, 0, 0, 128, 2, 1, 32768, launch_tsm_atb_f32_8x16_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_8x16_mfma_r256_p1_nt", 1, 8, 16, 0, 0, 256, 1, 1, 65536, launch_tsm_atb_f32_8x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_8x16_mfma_r256_p2_nt", 1, 8, 16, 0, 0, 256, 2, 1, 65536, launch_tsm_atb_f32_8x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_8x16_t8x4_r256_p2_nt", 1, 8, 16, 8, 4, 256, 2, 1, 49152, launch_tsm_atb_f32_8x16_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f32_8x32_mfma_r64_p1_nt", 1, 8, 32, 0, 0, 64, 1, 1, 24576, launch_tsm_atb_f32_8x32_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_8x32_mfma_r64_p2_nt", 1, 8, 32, 0, 0, 64, 2, 1, 24576, launch_tsm_atb_f32_8x32_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_8x32_mfma_r128_p1_nt", 1, 8, 32, 0, 0, 128, 1, 1, 49152, launch_tsm_atb_f32_8x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_8x32_mfma_r128_p2_nt", 1, 8, 32, 0, 0, 128, 2, 1, 49152, launch_tsm_atb_f32_8x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_8x32_mfma_r256_p1_nt", 1, 8, 32, 0, 0, 256, 1, 1, 98304, launch_tsm_atb_f32_8x32_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_8x32_mfma_r256_p2_nt", 1, 8, 32, 0, 0, 256, 2, 1, 98304, launch_tsm_atb_f32_8x32_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_8x32_t8x4_r128_p2_nt", 1, 8, 32, 8, 4, 128, 2, 1, 40960, launch_tsm_atb_f32_8x32_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_8x64_mfma_r32_p1_nt", 1, 8, 64, 0, 0, 32, 1, 1, 20480, launch_tsm_atb_f32_8x64_mfma_r32_p1_nt, 2},
  {"tsm_atb_f32_8x64_mfma_r32_p2_nt", 1, 8, 64, 0, 0, 32, 2, 1, 20480, launch_tsm_atb_f32_8x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f32_8x64_mfma_r64_p1_nt", 1, 8, 64, 0, 0, 64, 1, 1, 40960, launch_tsm_atb_f32_8x64_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_8x64_mfma_r64_p2_nt", 1, 8, 64, 0, 0, 64, 2, 1, 40960, launch_tsm_atb_f32_8x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_8x64_mfma_r128_p1_nt", 1, 8, 64, 0, 0, 128, 1, 1, 81920, launch_tsm_atb_f32_8x64_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_8x64_mfma_r128_p2_nt", 1, 8, 64, 0, 0, 128, 2, 1, 81920, launch_tsm_atb_f32_8x64_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_8x64_t8x4_r64_p2_nt", 1, 8, 64, 8, 4, 64, 2, 1, 36864, launch_tsm_atb_f32_8x64_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f32_16x1_t8x1_r256_p2_nt", 1, 16, 1, 8, 1, 256, 2, 1, 34816, launch_tsm_atb_f32_16x1_t8x1_r256_p2_nt, 1},
  {"tsm_atb_f32_16x2_t8x2_r256_p2_nt", 1, 16, 2, 8, 2, 256, 2, 1, 36864, launch_tsm_atb_f32_16x2_t8x2_r256_p2_nt, 1},
  {"tsm_atb_f32_16x4_mfma_r128_p1_nt", 1, 16, 4, 0, 0, 128, 1, 1, 32768, launch_tsm_atb_f32_16x4_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_16x4_mfma_r128_p2_nt", 1, 16, 4, 0, 0, 128, 2, 1, 32768, launch_tsm_atb_f32_16x4_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_16x4_mfma_r256_p1_nt", 1, 16, 4, 0, 0, 256, 1, 1, 65536, launch_tsm_atb_f32_16x4_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_16x4_mfma_r256_p2_nt", 1, 16, 4, 0, 0, 256, 2, 1, 65536, launch_tsm_atb_f32_16x4_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_16x4_t8x4_r256_p2_nt", 1, 16, 4, 8, 4, 256, 2, 1, 40960, launch_tsm_atb_f32_16x4_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f32_16x8_mfma_r128_p1_nt", 1, 16, 8, 0, 0, 128, 1, 1, 32768, launch_tsm_atb_f32_16x8_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_16x8_mfma_r128_p2_nt", 1, 16, 8, 0, 0, 128, 2, 1, 32768, launch_tsm_atb_f32_16x8_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_16x8_mfma_r256_p1_nt", 1, 16, 8, 0, 0, 256, 1, 1, 65536, launch_tsm_atb_f32_16x8_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_16x8_mfma_r256_p2_nt", 1, 16, 8, 0, 0, 256, 2, 1, 65536, launch_tsm_atb_f32_16x8_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_16x8_t8x4_r256_p2_nt", 1, 16, 8, 8, 4, 256, 2, 1, 49152, launch_tsm_atb_f32_16x8_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f32_16x16_direct_d4_nt", 1, 16, 16, 0, 0, 4, 4, 1, 4096, launch_tsm_atb_f32_16x16_direct_d4_nt, 3},
  {"tsm_atb_f32_16x16_direct_d8_nt", 1, 16, 16, 0, 0, 4, 8, 1, 4096, launch_tsm_atb_f32_16x16_direct_d8_nt, 3},
  {"tsm_atb_f32_16x16_direct_d16_nt", 1, 16, 16, 0, 0, 4, 16, 1, 4096, launch_tsm_atb_f32_16x16_direct_d16_nt, 3},
  {"tsm_atb_f32_16x16_mfma_r128_p1", 1, 16, 16, 0, 0, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_mfma_r128_p1, 2},
  {"tsm_atb_f32_16x16_mfma_r128_p1_nt", 1, 16, 16, 0, 0, 128, 1, 1, 32768, launch_tsm_atb_f32_16x16_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_16x16_mfma_r128_p2", 1, 16, 16, 0, 0, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_mfma_r128_p2, 2},
  {"tsm_atb_f32_16x16_mfma_r128_p2_nt", 1, 16, 16, 0, 0, 128, 2, 1, 32768, launch_tsm_atb_f32_16x16_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_16x16_mfma_r256_p1", 1, 16, 16, 0, 0, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_mfma_r256_p1, 2},
  {"tsm_atb_f32_16x16_mfma_r256_p1_nt", 1, 16, 16, 0, 0, 256, 1, 1, 65536, launch_tsm_atb_f32_16x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_16x16_mfma_r256_p2", 1, 16, 16, 0, 0, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_mfma_r256_p2, 2},
  {"tsm_atb_f32_16x16_mfma_r256_p2_nt", 1, 16, 16, 0, 0, 256, 2, 1, 65536, launch_tsm_atb_f32_16x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_16x16_t4x4_r128_p1", 1, 16, 16, 4, 4, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t4x4_r128_p1, 1},
  {"tsm_atb_f32_16x16_t4x4_r128_p2", 1, 16, 16, 4, 4, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t4x4_r128_p2, 1},
  {"tsm_atb_f32_16x16_t4x4_r128_p2_nt", 1, 16, 16, 4, 4, 128, 2, 1, 32768, launch_tsm_atb_f32_16x16_t4x4_r128_p2_nt, 1},
  {"tsm_atb_f32_16x16_t4x4_r128_p4_nt", 1, 16, 16, 4, 4, 128, 4, 1, 32768, launch_tsm_atb_f32_16x16_t4x4_r128_p4_nt, 1},
  {"tsm_atb_f32_16x16_t4x4_r256_p1", 1, 16, 16, 4, 4, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p1, 1},
  {"tsm_atb_f32_16x16_t4x4_r256_p2", 1, 16, 16, 4, 4, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p2, 1},
  {"tsm_atb_f32_16x16_t4x4_r256_p2_nt", 1, 16, 16, 4, 4, 256, 2, 1, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p2_nt, 1},
  {"tsm_atb_f32_16x16_t4x4_r256_p4_nt", 1, 16, 16, 4, 4, 256, 4, 1, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p4_nt, 1},
  {"tsm_atb_f32_16x16_t4x8_r128_p1", 1, 16, 16, 4, 8, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t4x8_r128_p1, 1},
  {"tsm_atb_f32_16x16_t4x8_r128_p2", 1, 16, 16, 4, 8, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t4x8_r128_p2, 1},
  {"tsm_atb_f32_16x16_t4x8_r128_p2_nt", 1, 16, 16, 4, 8, 128, 2, 1, 32768, launch_tsm_atb_f32_16x16_t4x8_r128_p2_nt, 1},
  {"tsm_atb_f32_16x16_t4x8_r128_p4_nt", 1, 16, 16, 4, 8, 128, 4, 1, 32768, launch_tsm_atb_f32_16x16_t4x8_r128_p4_nt, 1},
  {"tsm_atb_f32_16x16_t4x8_r256_p1", 1, 16, 16, 4, 8, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t4x8_r256_p1, 1},
  {"tsm_atb_f32_16x16_t4x8_r256_p2", 1, 16, 16, 4, 8, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t4x8_r256_p2, 1},
  {"tsm_atb_f32_16x16_t4x8_r256_p2_nt", 1, 16, 16, 4, 8, 256, 2, 1, 65536, launch_tsm_atb_f32_16x16_t4x8_r256_p2_nt, 1},
  {"tsm_atb_f32_16x16_t4x8_r256_p4_nt", 1, 16, 16, 4, 8, 256, 4, 1, 65536, launch_tsm_atb_f32_16x16_t4x8_r256_p4_nt, 1},
  {"tsm_atb_f32_16x16_t4x8_r512_p1", 1, 16, 16, 4, 8, 512, 1, 0, 131072, launch_tsm_atb_f32_16x16_t4x8_r512_p1, 1},
  {"tsm_atb_f32_16x16_t4x8_r512_p2", 1, 16, 16, 4, 8, 512, 2, 0, 131072, launch_tsm_atb_f32_16x16_t4x8_r512_p2, 1},
  {"tsm_atb_f32_16x16_t4x8_r512_p2_nt", 1, 16, 16, 4, 8, 512, 2, 1, 131072, launch_tsm_atb_f32_16x16_t4x8_r512_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p1", 1, 16, 16, 8, 4, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p1, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p1_nt", 1, 16, 16, 8, 4, 128, 1, 1, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p1_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p2", 1, 16, 16, 8, 4, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p2, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p2_nt", 1, 16, 16, 8, 4, 128, 2, 1, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p4_nt", 1, 16, 16, 8, 4, 128, 4, 1, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p4_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p1", 1, 16, 16, 8, 4, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p1, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p1_nt", 1, 16, 16, 8, 4, 256, 1, 1, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p1_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p2", 1, 16, 16, 8, 4, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p2, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p2_nt", 1, 16, 16, 8, 4, 256, 2, 1, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p4_nt", 1, 16, 16, 8, 4, 256, 4, 1, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p4_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r512_p1", 1, 16, 16, 8, 4, 512, 1, 0, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p1, 1},
  {"tsm_atb_f32_16x16_t8x4_r512_p1_nt", 1, 16, 16, 8, 4, 512, 1, 1, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p1_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r512_p2", 1, 16, 16, 8, 4, 512, 2, 0, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p2, 1},
  {"tsm_atb_f32_16x16_t8x4_r512_p2_nt", 1, 16, 16, 8, 4, 512, 2, 1, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x8_r128_p1", 1, 16, 16, 8, 8, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p1, 1},
  {"tsm_atb_f32_16x16_t8x8_r128_p2", 1, 16, 16, 8, 8, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p2, 1},
  {"tsm_atb_f32_16x16_t8x8_r128_p2_nt", 1, 16, 16, 8, 8, 128, 2, 1, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x8_r256_p1", 1, 16, 16, 8, 8, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p1, 1},
  {"tsm_atb_f32_16x16_t8x8_r256_p2", 1, 16, 16, 8, 8, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p2, 1},
  {"tsm_atb_f32_16x16_t8x8_r256_p2_nt", 1, 16, 16, 8, 8, 256, 2, 1, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x8_r512_p1", 1, 16, 16, 8, 8, 512, 1, 0, 131072, launch_tsm_atb_f32_16x16_t8x8_r512_p1, 1},
  {"tsm_atb_f32_16x16_t8x8_r512_p2", 1, 16, 16, 8, 8, 512, 2, 0, 131072, launch_tsm_atb_f32_16x16_t8x8_r512_p2, 1},
  {"tsm_atb_f32_16x16_t8x8_r512_p2_nt", 1, 16, 16, 8, 8, 512, 2, 1, 131072, launch_tsm_atb_f32_16x16_t8x8_r512_p2_nt, 1},
  {"tsm_atb_f32_16x32_direct_d8_nt", 1, 16, 32, 0, 0, 4, 8, 1, 8192, launch_tsm_atb_f32_16x32_direct_d8_nt, 3},
  {"tsm_atb_f32_16x32_t8x4_r128_p2_nt", 1, 16, 32, 8, 4, 128, 2, 1, 49152, launch_tsm_atb_f32_16x32_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_16x64_direct_d8_nt", 1, 16, 64, 0, 0, 4, 8, 1, 16384, launch_tsm_atb_f32_16x64_direct_d8_nt, 3},
  {"tsm_atb_f32_16x64_mfma_r64_p2_nt", 1, 16, 64, 0, 0, 64, 2, 1, 40960, launch_tsm_atb_f32_16x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_32x1_t8x1_r128_p2_nt", 1, 32, 1, 8, 1, 128, 2, 1, 33792, launch_tsm_atb_f32_32x1_t8x1_r128_p2_nt, 1},
  {"tsm_atb_f32_32x2_t8x2_r128_p2_nt", 1, 32, 2, 8, 2, 128, 2, 1, 34816, launch_tsm_atb_f32_32x2_t8x2_r128_p2_nt, 1},
  {"tsm_atb_f32_32x4_mfma_r64_p1_nt", 1, 32, 4, 0, 0, 64, 1, 1, 24576, launch_tsm_atb_f32_32x4_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_32x4_mfma_r64_p2_nt", 1, 32, 4, 0, 0, 64, 2, 1, 24576, launch_tsm_atb_f32_32x4_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_32x4_mfma_r128_p1_nt", 1, 32, 4, 0, 0, 128, 1, 1, 49152, launch_tsm_atb_f32_32x4_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_32x4_mfma_r128_p2_nt", 1, 32, 4, 0, 0, 128, 2, 1, 49152, launch_tsm_atb_f32_32x4_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_32x4_mfma_r256_p1_nt", 1, 32, 4, 0, 0, 256, 1, 1, 98304, launch_tsm_atb_f32_32x4_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_32x4_mfma_r256_p2_nt", 1, 32, 4, 0, 0, 256, 2, 1, 98304, launch_tsm_atb_f32_32x4_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_32x4_t8x4_r128_p2_nt", 1, 32, 4, 8, 4, 128, 2, 1, 36864, launch_tsm_atb_f32_32x4_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_32x8_mfma_r64_p1_nt", 1, 32, 8, 0, 0, 64, 1, 1, 24576, launch_tsm_atb_f32_32x8_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_32x8_mfma_r64_p2_nt", 1, 32, 8, 0, 0, 64, 2, 1, 24576, launch_tsm_atb_f32_32x8_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_32x8_mfma_r128_p1_nt", 1, 32, 8, 0, 0, 128, 1, 1, 49152, launch_tsm_atb_f32_32x8_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_32x8_mfma_r128_p2_nt", 1, 32, 8, 0, 0, 128, 2, 1, 49152, launch_tsm_atb_f32_32x8_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_32x8_mfma_r256_p1_nt", 1, 32, 8, 0, 0, 256, 1, 1, 98304, launch_tsm_atb_f32_32x8_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_32x8_mfma_r256_p2_nt", 1, 32, 8, 0, 0, 256, 2, 1, 98304, launch_tsm_atb_f32_32x8_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_32x8_t8x4_r128_p2_nt", 1, 32, 8, 8, 4, 128, 2, 1, 40960, launch_tsm_atb_f32_32x8_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_32x16_direct_d8_nt", 1, 32, 16, 0, 0, 4, 8, 1, 8192, launch_tsm_atb_f32_32x16_direct_d8_nt, 3},
  {"tsm_atb_f32_32x16_t8x4_r128_p2_nt", 1, 32, 16, 8, 4, 128, 2, 1, 49152, launch_tsm_atb_f32_32x16_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_32x32_direct_d4_nt", 1, 32, 32, 0, 0, 4, 4, 1, 16384, launch_tsm_atb_f32_32x32_direct_d4_nt, 3},
  {"tsm_atb_f32_32x32_direct_d8_nt", 1, 32, 32, 0, 0, 4, 8, 1, 16384, launch_tsm_atb_f32_32x32_direct_d8_nt, 3},
  {"tsm_atb_f32_32x32_direct_d16_nt", 1, 32, 32, 0, 0, 4, 16, 1, 16384, launch_tsm_atb_f32_32x32_direct_d16_nt, 3},
  {"tsm_atb_f32_32x32_mfma_r64_p1", 1, 32, 32, 0, 0, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_mfma_r64_p1, 2},
  {"tsm_atb_f32_32x32_mfma_r64_p1_nt", 1, 32, 32, 0, 0, 64, 1, 1, 32768, launch_tsm_atb_f32_32x32_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_32x32_mfma_r64_p2", 1, 32, 32, 0, 0, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_mfma_r64_p2, 2},
  {"tsm_atb_f32_32x32_mfma_r64_p2_nt", 1, 32, 32, 0, 0, 64, 2, 1, 32768, launch_tsm_atb_f32_32x32_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_32x32_mfma_r128_p1", 1, 32, 32, 0, 0, 128, 1, 0, 65536, launch_tsm_atb_f32_32x32_mfma_r128_p1, 2},
  {"tsm_atb_f32_32x32_mfma_r128_p1_nt", 1, 32, 32, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_f32_32x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_32x32_mfma_r128_p2", 1, 32, 32, 0, 0, 128, 2, 0, 65536, launch_tsm_atb_f32_32x32_mfma_r128_p2, 2},
  {"tsm_atb_f32_32x32_mfma_r128_p2_nt", 1, 32, 32, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_f32_32x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_32x32_mfma_r256_p1", 1, 32, 32, 0, 0, 256, 1, 0, 131072, launch_tsm_atb_f32_32x32_mfma_r256_p1, 2},
  {"tsm_atb_f32_32x32_mfma_r256_p1_nt", 1, 32, 32, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_f32_32x32_mfma_r256_p1_nt, 2},
  {"tsm_atb_f32_32x32_mfma_r256_p2", 1, 32, 32, 0, 0, 256, 2, 0, 131072, launch_tsm_atb_f32_32x32_mfma_r256_p2, 2},
  {"tsm_atb_f32_32x32_mfma_r256_p2_nt", 1, 32, 32, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_f32_32x32_mfma_r256_p2_nt, 2},
  {"tsm_atb_f32_32x32_t4x4_r32_p1", 1, 32, 32, 4, 4, 32, 1, 0, 16384, launch_tsm_atb_f32_32x32_t4x4_r32_p1, 1},
  {"tsm_atb_f32_32x32_t4x4_r32_p2", 1, 32, 32, 4, 4, 32, 2, 0, 16384, launch_tsm_atb_f32_32x32_t4x4_r32_p2, 1},
  {"tsm_atb_f32_32x32_t4x4_r32_p2_nt", 1, 32, 32, 4, 4, 32, 2, 1, 16384, launch_tsm_atb_f32_32x32_t4x4_r32_p2_nt, 1},
  {"tsm_atb_f32_32x32_t4x4_r32_p4_nt", 1, 32, 32, 4, 4, 32, 4, 1, 16384, launch_tsm_atb_f32_32x32_t4x4_r32_p4_nt, 1},
  {"tsm_atb_f32_32x32_t4x4_r64_p1", 1, 32, 32, 4, 4, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p1, 1},
  {"tsm_atb_f32_32x32_t4x4_r64_p2", 1, 32, 32, 4, 4, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p2, 1},
  {"tsm_atb_f32_32x32_t4x4_r64_p2_nt", 1, 32, 32, 4, 4, 64, 2, 1, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p2_nt, 1},
  {"tsm_atb_f32_32x32_t4x4_r64_p4_nt", 1, 32, 32, 4, 4, 64, 4, 1, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p4_nt, 1},
  {"tsm_atb_f32_32x32_t4x8_r32_p1", 1, 32, 32, 4, 8, 32, 1, 0, 16384, launch_tsm_atb_f32_32x32_t4x8_r32_p1, 1},
  {"tsm_atb_f32_32x32_t4x8_r32_p2", 1, 32, 32, 4, 8, 32, 2, 0, 16384, launch_tsm_atb_f32_32x32_t4x8_r32_p2, 1},
  {"tsm_atb_f32_32x32_t4x8_r32_p2_nt", 1, 32, 32, 4, 8, 32, 2, 1, 16384, launch_tsm_atb_f32_32x32_t4x8_r32_p2_nt, 1},
  {"tsm_atb_f32_32x32_t4x8_r32_p4_nt", 1, 32, 32, 4, 8, 32, 4, 1, 16384, launch_tsm_atb_f32_32x32_t4x8_r32_p4_nt, 1},
  {"tsm_atb_f32_32x32_t4x8_r64_p1", 1, 32, 32, 4, 8, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t4x8_r64_p1, 1},
  {"tsm_atb_f32_32x32_t4x8_r64_p2", 1, 32, 32, 4, 8, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t4x8_r64_p2, 1},
  {"tsm_atb_f32_32x32_t4x8_r64_p2_nt", 1, 32, 32, 4, 8, 64, 2, 1, 32768, launch_tsm_atb_f32_32x32_t4x8_r64_p2_nt, 1},
  {"tsm_atb_f32_32x32_t4x8_r64_p4_nt", 1, 32, 32, 4, 8, 64, 4, 1, 32768, launch_tsm_atb_f32_32x32_t4x8_r64_p4_nt, 1},
  {"tsm_atb_f32_32x32_t4x8_r128_p1", 1, 32, 32, 4, 8, 128, 1, 0, 65536, launch_tsm_atb_f32_32x32_t4x8_r128_p1, 1},
  {"tsm_atb_f32_32x32_t4x8_r128_p2", 1, 32, 32, 4, 8, 128, 2, 0, 65536, launch_tsm_atb_f32_32x32_t4x8_r128_p2, 1},
  {"tsm_atb_f32_32x32_t4x8_r128_p2_nt", 1, 32, 32, 4, 8, 128, 2, 1, 65536, launch_tsm_atb_f32_32x32_t4x8_r128_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p1", 1, 32, 32, 8, 4, 32, 1, 0, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p1, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p1_nt", 1, 32, 32, 8, 4, 32, 1, 1, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p1_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p2", 1, 32, 32, 8, 4, 32, 2, 0, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p2, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p2_nt", 1, 32, 32, 8, 4, 32, 2, 1, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p4_nt", 1, 32, 32, 8, 4, 32, 4, 1, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p4_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p1", 1, 32, 32, 8, 4, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p1, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p1_nt", 1, 32, 32, 8, 4, 64, 1, 1, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p1_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p2", 1, 32, 32, 8, 4, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p2, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p2_nt", 1, 32, 32, 8, 4, 64, 2, 1, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p4_nt", 1, 32, 32, 8, 4, 64, 4, 1, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p4_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r128_p1", 1, 32, 32, 8, 4, 128, 1, 0, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p1, 1},
  {"tsm_atb_f32_32x32_t8x4_r128_p1_nt", 1, 32, 32, 8, 4, 128, 1, 1, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p1_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r128_p2", 1, 32, 32, 8, 4, 128, 2, 0, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p2, 1},
  {"tsm_atb_f32_32x32_t8x4_r128_p2_nt", 1, 32, 32, 8, 4, 128, 2, 1, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x8_r32_p1", 1, 32, 32, 8, 8, 32, 1, 0, 16384, launch_tsm_atb_f32_32x32_t8x8_r32_p1, 1},
  {"tsm_atb_f32_32x32_t8x8_r32_p2", 1, 32, 32, 8, 8, 32, 2, 0, 16384, launch_tsm_atb_f32_32x32_t8x8_r32_p2, 1},
  {"tsm_atb_f32_32x32_t8x8_r32_p2_nt", 1, 32, 32, 8, 8, 32, 2, 1, 16384, launch_tsm_atb_f32_32x32_t8x8_r32_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x8_r64_p1", 1, 32, 32, 8, 8, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p1, 1},
  {"tsm_atb_f32_32x32_t8x8_r64_p2", 1, 32, 32, 8, 8, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p2, 1},
  {"tsm_atb_f32_32x32_t8x8_r64_p2_nt", 1, 32, 32, 8, 8, 64, 2, 1, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x8_r128_p1", 1, 32, 32, 8, 8, 128, 1, 0, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p1, 1},
  {"tsm_atb_f32_32x32_t8x8_r128_p2", 1, 32, 32, 8, 8, 128, 2, 0, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p2, 1},
  {"tsm_atb_f32_32x32_t8x8_r128_p2_nt", 1, 32, 32, 8, 8, 128, 2, 1, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p2_nt, 1},
  {"tsm_atb_f32_32x64_direct_d8_nt", 1, 32, 64, 0, 0, 4, 8, 1, 32768, launch_tsm_atb_f32_32x64_direct_d8_nt, 3},
  {"tsm_atb_f32_32x64_mfma_r64_p2_nt", 1, 32, 64, 0, 0, 64, 2, 1, 49152, launch_tsm_atb_f32_32x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_64x1_t8x1_r64_p2_nt", 1, 64, 1, 8, 1, 64, 2, 1, 33280, launch_tsm_atb_f32_64x1_t8x1_r64_p2_nt, 1},
  {"tsm_atb_f32_64x2_t8x2_r64_p2_nt", 1, 64, 2, 8, 2, 64, 2, 1, 33792, launch_tsm_atb_f32_64x2_t8x2_r64_p2_nt, 1},
  {"tsm_atb_f32_64x4_mfma_r32_p1_nt", 1, 64, 4, 0, 0, 32, 1, 1, 20480, launch_tsm_atb_f32_64x4_mfma_r32_p1_nt, 2},
  {"tsm_atb_f32_64x4_mfma_r32_p2_nt", 1, 64, 4, 0, 0, 32, 2, 1, 20480, launch_tsm_atb_f32_64x4_mfma_r32_p2_nt, 2},
  {"tsm_atb_f32_64x4_mfma_r64_p1_nt", 1, 64, 4, 0, 0, 64, 1, 1, 40960, launch_tsm_atb_f32_64x4_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_64x4_mfma_r64_p2_nt", 1, 64, 4, 0, 0, 64, 2, 1, 40960, launch_tsm_atb_f32_64x4_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_64x4_mfma_r128_p1_nt", 1, 64, 4, 0, 0, 128, 1, 1, 81920, launch_tsm_atb_f32_64x4_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_64x4_mfma_r128_p2_nt", 1, 64, 4, 0, 0, 128, 2, 1, 81920, launch_tsm_atb_f32_64x4_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_64x4_t8x4_r64_p2_nt", 1, 64, 4, 8, 4, 64, 2, 1, 34816, launch_tsm_atb_f32_64x4_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f32_64x8_mfma_r32_p1_nt", 1, 64, 8, 0, 0, 32, 1, 1, 20480, launch_tsm_atb_f32_64x8_mfma_r32_p1_nt, 2},
  {"tsm_atb_f32_64x8_mfma_r32_p2_nt", 1, 64, 8, 0, 0, 32, 2, 1, 20480, launch_tsm_atb_f32_64x8_mfma_r32_p2_nt, 2},
  {"tsm_atb_f32_64x8_mfma_r64_p1_nt", 1, 64, 8, 0, 0, 64, 1, 1, 40960, launch_tsm_atb_f32_64x8_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_64x8_mfma_r64_p2_nt", 1, 64, 8, 0, 0, 64, 2, 1, 40960, launch_tsm_atb_f32_64x8_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_64x8_mfma_r128_p1_nt", 1, 64, 8, 0, 0, 128, 1, 1, 81920, launch_tsm_atb_f32_64x8_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_64x8_mfma_r128_p2_nt", 1, 64, 8, 0, 0, 128, 2, 1, 81920, launch_tsm_atb_f32_64x8_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_64x8_t8x4_r64_p2_nt", 1, 64, 8, 8, 4, 64, 2, 1, 36864, launch_tsm_atb_f32_64x8_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f32_64x16_direct_d8_nt", 1, 64, 16, 0, 0, 4, 8, 1, 16384, launch_tsm_atb_f32_64x16_direct_d8_nt, 3},
  {"tsm_atb_f32_64x16_mfma_r64_p2_nt", 1, 64, 16, 0, 0, 64, 2, 1, 40960, launch_tsm_atb_f32_64x16_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_64x32_direct_d8_nt", 1, 64, 32, 0, 0, 4, 8, 1, 32768, launch_tsm_atb_f32_64x32_direct_d8_nt, 3},
  {"tsm_atb_f32_64x32_mfma_r64_p2_nt", 1, 64, 32, 0, 0, 64, 2, 1, 49152, launch_tsm_atb_f32_64x32_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_64x64_direct_d4_nt", 1, 64, 64, 0, 0, 4, 4, 1, 65536, launch_tsm_atb_f32_64x64_direct_d4_nt, 3},
  {"tsm_atb_f32_64x64_direct_d8_nt", 1, 64, 64, 0, 0, 4, 8, 1, 65536, launch_tsm_atb_f32_64x64_direct_d8_nt, 3},
  {"tsm_atb_f32_64x64_direct_d16_nt", 1, 64, 64, 0, 0, 4, 16, 1, 65536, launch_tsm_atb_f32_64x64_direct_d16_nt, 3},
  {"tsm_atb_f32_64x64_mfma_r32_p1", 1, 64, 64, 0, 0, 32, 1, 0, 32768, launch_tsm_atb_f32_64x64_mfma_r32_p1, 2},
  {"tsm_atb_f32_64x64_mfma_r32_p1_nt", 1, 64, 64, 0, 0, 32, 1, 1, 32768, launch_tsm_atb_f32_64x64_mfma_r32_p1_nt, 2},
  {"tsm_atb_f32_64x64_mfma_r32_p2", 1, 64, 64, 0, 0, 32, 2, 0, 32768, launch_tsm_atb_f32_64x64_mfma_r32_p2, 2},
  {"tsm_atb_f32_64x64_mfma_r32_p2_nt", 1, 64, 64, 0, 0, 32, 2, 1, 32768, launch_tsm_atb_f32_64x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f32_64x64_mfma_r64_p1", 1, 64, 64, 0, 0, 64, 1, 0, 65536, launch_tsm_atb_f32_64x64_mfma_r64_p1, 2},
  {"tsm_atb_f32_64x64_mfma_r64_p1_nt", 1, 64, 64, 0, 0, 64, 1, 1, 65536, launch_tsm_atb_f32_64x64_mfma_r64_p1_nt, 2},
  {"tsm_atb_f32_64x64_mfma_r64_p2", 1, 64, 64, 0, 0, 64, 2, 0, 65536, launch_tsm_atb_f32_64x64_mfma_r64_p2, 2},
  {"tsm_atb_f32_64x64_mfma_r64_p2_nt", 1, 64, 64, 0, 0, 64, 2, 1, 65536, launch_tsm_atb_f32_64x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f32_64x64_mfma_r128_p1", 1, 64, 64, 0, 0, 128, 1, 0, 131072, launch_tsm_atb_f32_64x64_mfma_r128_p1, 2},
  {"tsm_atb_f32_64x64_mfma_r128_p1_nt", 1, 64, 64, 0, 0, 128, 1, 1, 131072, launch_tsm_atb_f32_64x64_mfma_r128_p1_nt, 2},
  {"tsm_atb_f32_64x64_mfma_r128_p2", 1, 64, 64, 0, 0, 128, 2, 0, 131072, launch_tsm_atb_f32_64x64_mfma_r128_p2, 2},
  {"tsm_atb_f32_64x64_mfma_r128_p2_nt", 1, 64, 64, 0, 0, 128, 2, 1, 131072, launch_tsm_atb_f32_64x64_mfma_r128_p2_nt, 2},
  {"tsm_atb_f32_64x64_t8x8_r16_p1", 1, 64, 64, 8, 8, 16, 1, 0, 32768, launch_tsm_atb_f32_64x64_t8x8_r16_p1, 1},
  {"tsm_atb_f32_64x64_t8x8_r16_p1_nt", 1, 64, 64, 8, 8, 16, 1, 1, 32768, launch_tsm_atb_f32_64x64_t8x8_r16_p1_nt, 1},
  {"tsm_atb_f32_64x64_t8x8_r16_p2", 1, 64, 64, 8, 8, 16, 2, 0, 32768, launch_tsm_atb_f32_64x64_t8x8_r16_p2, 1},
  {"tsm_atb_f32_64x64_t8x8_r16_p2_nt", 1, 64, 64, 8, 8, 16, 2, 1, 32768, launch_tsm_atb_f32_64x64_t8x8_r16_p2_nt, 1},
  {"tsm_atb_f32_64x64_t8x8_r32_p1", 1, 64, 64, 8, 8, 32, 1, 0, 32768, launch_tsm_atb_f32_64x64_t8x8_r32_p1, 1},
  {"tsm_atb_f32_64x64_t8x8_r32_p1_nt", 1, 64, 64, 8, 8, 32, 1, 1, 32768, launch_tsm_atb_f32_64x64_t8x8_r32_p1_nt, 1},
  {"tsm_atb_f32_64x64_t8x8_r32_p2", 1, 64, 64, 8, 8, 32, 2, 0, 32768, launch_tsm_atb_f32_64x64_t8x8_r32_p2, 1},
  {"tsm_atb_f32_64x64_t8x8_r32_p2_nt", 1, 64, 64, 8, 8, 32, 2, 1, 32768, launch_tsm_atb_f32_64x64_t8x8_r32_p2_nt, 1},
  {"tsm_atb_f32_64x64_t8x8_r64_p1", 1, 64, 64, 8, 8, 64, 1, 0, 65536, launch_tsm_atb_f32_64x64_t8x8_r64_p1, 1},
  {"tsm_atb_f32_64x64_t8x8_r64_p1_nt", 1, 64, 64, 8, 8, 64, 1, 1, 65536, launch_tsm_atb_f32_64x64_t8x8_r64_p1_nt, 1},
  {"tsm_atb_f32_64x64_t8x8_r64_p2", 1, 64, 64, 8, 8, 64, 2, 0, 65536, launch_tsm_atb_f32_64x64_t8x8_r64_p2, 1},
  {"tsm_atb_f32_64x64_t8x8_r64_p2_nt", 1, 64, 64, 8, 8, 64, 2, 1, 65536, launch_tsm_atb_f32_64x64_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r1024_p1", 0, 1, 1, 1, 1, 1024, 1, 0, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p1, 1},
  {"tsm_atb_f64_1x1_t1x1_r1024_p1_nt", 0, 1, 1, 1, 1, 1024, 1, 1, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p1_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r1024_p2", 0, 1, 1, 1, 1, 1024, 2, 0, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p2, 1},
  {"tsm_atb_f64_1x1_t1x1_r1024_p2_nt", 0, 1, 1, 1, 1, 1024, 2, 1, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p2_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r1024_p4_nt", 0, 1, 1, 1, 1, 1024, 4, 1, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p4_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p1", 0, 1, 1, 1, 1, 2048, 1, 0, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p1, 1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p1_nt", 0, 1, 1, 1, 1, 2048, 1, 1, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p1_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p2", 0, 1, 1, 1, 1, 2048, 2, 0, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p2, 1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p2_nt", 0, 1, 1, 1, 1, 2048, 2, 1, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p2_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p4_nt", 0, 1, 1, 1, 1, 2048, 4, 1, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p4_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r4096_p1", 0, 1, 1, 1, 1, 4096, 1, 0, 131072, launch_tsm_atb_f64_1x1_t1x1_r4096_p1, 1},
  {"tsm_atb_f64_1x1_t1x1_r4096_p1_nt", 0, 1, 1, 1, 1, 4096, 1, 1, 131072, launch_tsm_atb_f64_1x1_t1x1_r4096_p1_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r4096_p2", 0, 1, 1, 1, 1, 4096, 2, 0, 131072, launch_tsm_atb_f64_1x1_t1x1_r4096_p2, 1},
  {"tsm_atb_f64_1x1_t1x1_r4096_p2_nt", 0, 1, 1, 1, 1, 4096, 2, 1, 131072, launch_tsm_atb_f64_1x1_t1x1_r4096_p2_nt, 1},
  {"tsm_atb_f64_1x2_t1x2_r1024_p2_nt", 0, 1, 2, 1, 2, 1024, 2, 1, 49152, launch_tsm_atb_f64_1x2_t1x2_r1024_p2_nt, 1},
  {"tsm_atb_f64_1x4_t1x4_r512_p2_nt", 0, 1, 4, 1, 4, 512, 2, 1, 40960, launch_tsm_atb_f64_1x4_t1x4_r512_p2_nt, 1},
  {"tsm_atb_f64_1x8_t1x8_r256_p2_nt", 0, 1, 8, 1, 8, 256, 2, 1, 36864, launch_tsm_atb_f64_1x8_t1x8_r256_p2_nt, 1},
  {"tsm_atb_f64_1x16_t1x8_r128_p2_nt", 0, 1, 16, 1, 8, 128, 2, 1, 34816, launch_tsm_atb_f64_1x16_t1x8_r128_p2_nt, 1},
  {"tsm_atb_f64_1x32_t1x8_r64_p2_nt", 0, 1, 32, 1, 8, 64, 2, 1, 33792, launch_tsm_atb_f64_1x32_t1x8_r64_p2_nt, 1},
  {"tsm_atb_f64_1x64_t1x8_r32_p2_nt", 0, 1, 64, 1, 8, 32, 2, 1, 33280, launch_tsm_atb_f64_1x64_t1x8_r32_p2_nt, 1},
  {"tsm_atb_f64_2x1_t2x1_r1024_p2_nt", 0, 2, 1, 2, 1, 1024, 2, 1, 49152, launch_tsm_atb_f64_2x1_t2x1_r1024_p2_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r512_p1", 0, 2, 2, 2, 2, 512, 1, 0, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p1, 1},
  {"tsm_atb_f64_2x2_t2x2_r512_p1_nt", 0, 2, 2, 2, 2, 512, 1, 1, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p1_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r512_p2", 0, 2, 2, 2, 2, 512, 2, 0, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p2, 1},
  {"tsm_atb_f64_2x2_t2x2_r512_p2_nt", 0, 2, 2, 2, 2, 512, 2, 1, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p2_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r512_p4_nt", 0, 2, 2, 2, 2, 512, 4, 1, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p4_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p1", 0, 2, 2, 2, 2, 1024, 1, 0, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p1, 1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p1_nt", 0, 2, 2, 2, 2, 1024, 1, 1, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p1_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p2", 0, 2, 2, 2, 2, 1024, 2, 0, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p2, 1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p2_nt", 0, 2, 2, 2, 2, 1024, 2, 1, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p2_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p4_nt", 0, 2, 2, 2, 2, 1024, 4, 1, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p4_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r2048_p1", 0, 2, 2, 2, 2, 2048, 1, 0, 131072, launch_tsm_atb_f64_2x2_t2x2_r2048_p1, 1},
  {"tsm_atb_f64_2x2_t2x2_r2048_p1_nt", 0, 2, 2, 2, 2, 2048, 1, 1, 131072, launch_tsm_atb_f64_2x2_t2x2_r2048_p1_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r2048_p2", 0, 2, 2, 2, 2, 2048, 2, 0, 131072, launch_tsm_atb_f64_2x2_t2x2_r2048_p2, 1},
  {"tsm_atb_f64_2x2_t2x2_r2048_p2_nt", 0, 2, 2, 2, 2, 2048, 2, 1, 131072, launch_tsm_atb_f64_2x2_t2x2_r2048_p2_nt, 1},
  {"tsm_atb_f64_2x4_t2x4_r512_p2_nt", 0, 2, 4, 2, 4, 512, 2, 1, 49152, launch_tsm_atb_f64_2x4_t2x4_r512_p2_nt, 1},
  {"tsm_atb_f64_2x8_t2x8_r256_p2_nt", 0, 2, 8, 2, 8, 256, 2, 1, 40960, launch_tsm_atb_f64_2x8_t2x8_r256_p2_nt, 1},
  {"tsm_atb_f64_2x16_mfma_r64_p1_nt", 0, 2, 16, 0, 0, 64, 1, 1, 32768, launch_tsm_atb_f64_2x16_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_2x16_mfma_r64_p2_nt", 0, 2, 16, 0, 0, 64, 2, 1, 32768, launch_tsm_atb_f64_2x16_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_2x16_mfma_r128_p1_nt", 0, 2, 16, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_f64_2x16_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_2x16_mfma_r128_p2_nt", 0, 2, 16, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_f64_2x16_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_2x16_mfma_r256_p1_nt", 0, 2, 16, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_f64_2x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_f64_2x16_mfma_r256_p2_nt", 0, 2, 16, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_f64_2x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_f64_2x16_t2x8_r128_p2_nt", 0, 2, 16, 2, 8, 128, 2, 1, 36864, launch_tsm_atb_f64_2x16_t2x8_r128_p2_nt, 1},
  {"tsm_atb_f64_2x32_mfma_r32_p1_nt", 0, 2, 32, 0, 0, 32, 1, 1, 24576, launch_tsm_atb_f64_2x32_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_2x32_mfma_r32_p2_nt", 0, 2, 32, 0, 0, 32, 2, 1, 24576, launch_tsm_atb_f64_2x32_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_2x32_mfma_r64_p1_nt", 0, 2, 32, 0, 0, 64, 1, 1, 49152, launch_tsm_atb_f64_2x32_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_2x32_mfma_r64_p2_nt", 0, 2, 32, 0, 0, 64, 2, 1, 49152, launch_tsm_atb_f64_2x32_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_2x32_mfma_r128_p1_nt", 0, 2, 32, 0, 0, 128, 1, 1, 98304, launch_tsm_atb_f64_2x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_2x32_mfma_r128_p2_nt", 0, 2, 32, 0, 0, 128, 2, 1, 98304, launch_tsm_atb_f64_2x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_2x32_t2x8_r64_p2_nt", 0, 2, 32, 2, 8, 64, 2, 1, 34816, launch_tsm_atb_f64_2x32_t2x8_r64_p2_nt, 1},
  {"tsm_atb_f64_2x64_mfma_r16_p1_nt", 0, 2, 64, 0, 0, 16, 1, 1, 20480, launch_tsm_atb_f64_2x64_mfma_r16_p1_nt, 2},
  {"tsm_atb_f64_2x64_mfma_r16_p2_nt", 0, 2, 64, 0, 0, 16, 2, 1, 20480, launch_tsm_atb_f64_2x64_mfma_r16_p2_nt, 2},
  {"tsm_atb_f64_2x64_mfma_r32_p1_nt", 0, 2, 64, 0, 0, 32, 1, 1, 40960, launch_tsm_atb_f64_2x64_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_2x64_mfma_r32_p2_nt", 0, 2, 64, 0, 0, 32, 2, 1, 40960, launch_tsm_atb_f64_2x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_2x64_mfma_r64_p1_nt", 0, 2, 64, 0, 0, 64, 1, 1, 81920, launch_tsm_atb_f64_2x64_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_2x64_mfma_r64_p2_nt", 0, 2, 64, 0, 0, 64, 2, 1, 81920, launch_tsm_atb_f64_2x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_2x64_t2x8_r32_p2_nt", 0, 2, 64, 2, 8, 32, 2, 1, 33792, launch_tsm_atb_f64_2x64_t2x8_r32_p2_nt, 1},
  {"tsm_atb_f64_4x1_t4x1_r512_p2_nt", 0, 4, 1, 4, 1, 512, 2, 1, 40960, launch_tsm_atb_f64_4x1_t4x1_r512_p2_nt, 1},
  {"tsm_atb_f64_4x2_t4x2_r512_p2_nt", 0, 4, 2, 4, 2, 512, 2, 1, 49152, launch_tsm_atb_f64_4x2_t4x2_r512_p2_nt, 1},
  {"tsm_atb_f64_4x4_t2x2_r512_p1", 0, 4, 4, 2, 2, 512, 1, 0, 65536, launch_tsm_atb_f64_4x4_t2x2_r512_p1, 1},
  {"tsm_atb_f64_4x4_t2x2_r512_p2", 0, 4, 4, 2, 2, 512, 2, 0, 65536, launch_tsm_atb_f64_4x4_t2x2_r512_p2, 1},
  {"tsm_atb_f64_4x4_t4x4_r256_p1", 0, 4, 4, 4, 4, 256, 1, 0, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p1, 1},
  {"tsm_atb_f64_4x4_t4x4_r256_p1_nt", 0, 4, 4, 4, 4, 256, 1, 1, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p1_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r256_p2", 0, 4, 4, 4, 4, 256, 2, 0, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p2, 1},
  {"tsm_atb_f64_4x4_t4x4_r256_p2_nt", 0, 4, 4, 4, 4, 256, 2, 1, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p2_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r256_p4_nt", 0, 4, 4, 4, 4, 256, 4, 1, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p4_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r512_p1", 0, 4, 4, 4, 4, 512, 1, 0, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p1, 1},
  {"tsm_atb_f64_4x4_t4x4_r512_p1_nt", 0, 4, 4, 4, 4, 512, 1, 1, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p1_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r512_p2", 0, 4, 4, 4, 4, 512, 2, 0, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p2, 1},
  {"tsm_atb_f64_4x4_t4x4_r512_p2_nt", 0, 4, 4, 4, 4, 512, 2, 1, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p2_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r512_p4_nt", 0, 4, 4, 4, 4, 512, 4, 1, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p4_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p1", 0, 4, 4, 4, 4, 1024, 1, 0, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p1, 1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p1_nt", 0, 4, 4, 4, 4, 1024, 1, 1, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p1_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p2", 0, 4, 4, 4, 4, 1024, 2, 0, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p2, 1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p2_nt", 0, 4, 4, 4, 4, 1024, 2, 1, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p2_nt, 1},
  {"tsm_atb_f64_4x8_t4x8_r256_p2_nt", 0, 4, 8, 4, 8, 256, 2, 1, 49152, launch_tsm_atb_f64_4x8_t4x8_r256_p2_nt, 1},
  {"tsm_atb_f64_4x16_mfma_r64_p1_nt", 0, 4, 16, 0, 0, 64, 1, 1, 32768, launch_tsm_atb_f64_4x16_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_4x16_mfma_r64_p2_nt", 0, 4, 16, 0, 0, 64, 2, 1, 32768, launch_tsm_atb_f64_4x16_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_4x16_mfma_r128_p1_nt", 0, 4, 16, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_f64_4x16_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_4x16_mfma_r128_p2_nt", 0, 4, 16, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_f64_4x16_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_4x16_mfma_r256_p1_nt", 0, 4, 16, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_f64_4x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_f64_4x16_mfma_r256_p2_nt", 0, 4, 16, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_f64_4x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_f64_4x16_t4x8_r128_p2_nt", 0, 4, 16, 4, 8, 128, 2, 1, 40960, launch_tsm_atb_f64_4x16_t4x8_r128_p2_nt, 1},
  {"tsm_atb_f64_4x32_mfma_r32_p1_nt", 0, 4, 32, 0, 0, 32, 1, 1, 24576, launch_tsm_atb_f64_4x32_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_4x32_mfma_r32_p2_nt", 0, 4, 32, 0, 0, 32, 2, 1, 24576, launch_tsm_atb_f64_4x32_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_4x32_mfma_r64_p1_nt", 0, 4, 32, 0, 0, 64, 1, 1, 49152, launch_tsm_atb_f64_4x32_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_4x32_mfma_r64_p2_nt", 0, 4, 32, 0, 0, 64, 2, 1, 49152, launch_tsm_atb_f64_4x32_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_4x32_mfma_r128_p1_nt", 0, 4, 32, 0, 0, 128, 1, 1, 98304, launch_tsm_atb_f64_4x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_4x32_mfma_r128_p2_nt", 0, 4, 32, 0, 0, 128, 2, 1, 98304, launch_tsm_atb_f64_4x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_4x32_t4x4_r32_p1", 0, 4, 32, 4, 4, 32, 1, 0, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p1, 1},
  {"tsm_atb_f64_4x32_t4x4_r32_p2", 0, 4, 32, 4, 4, 32, 2, 0, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p2, 1},
  {"tsm_atb_f64_4x32_t4x4_r32_p2_nt", 0, 4, 32, 4, 4, 32, 2, 1, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p2_nt, 1},
  {"tsm_atb_f64_4x32_t4x4_r32_p4_nt", 0, 4, 32, 4, 4, 32, 4, 1, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p4_nt, 1},
  {"tsm_atb_f64_4x32_t4x4_r64_p1", 0, 4, 32, 4, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p1, 1},
  {"tsm_atb_f64_4x32_t4x4_r64_p2", 0, 4, 32, 4, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p2, 1},
  {"tsm_atb_f64_4x32_t4x4_r64_p2_nt", 0, 4, 32, 4, 4, 64, 2, 1, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p2_nt, 1},
  {"tsm_atb_f64_4x32_t4x4_r64_p4_nt", 0, 4, 32, 4, 4, 64, 4, 1, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p4_nt, 1},
  {"tsm_atb_f64_4x32_t4x4_r128_p1", 0, 4, 32, 4, 4, 128, 1, 0, 73728, launch_tsm_atb_f64_4x32_t4x4_r128_p1, 1},
  {"tsm_atb_f64_4x32_t4x4_r128_p2", 0, 4, 32, 4, 4, 128, 2, 0, 73728, launch_tsm_atb_f64_4x32_t4x4_r128_p2, 1},
  {"tsm_atb_f64_4x32_t4x4_r128_p2_nt", 0, 4, 32, 4, 4, 128, 2, 1, 73728, launch_tsm_atb_f64_4x32_t4x4_r128_p2_nt, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p1", 0, 4, 32, 4, 8, 64, 1, 0, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p1, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p1_nt", 0, 4, 32, 4, 8, 64, 1, 1, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p1_nt, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p2", 0, 4, 32, 4, 8, 64, 2, 0, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p2, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p2_nt", 0, 4, 32, 4, 8, 64, 2, 1, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p2_nt, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p4_nt", 0, 4, 32, 4, 8, 64, 4, 1, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p4_nt, 1},
  {"tsm_atb_f64_4x32_t4x8_r128_p1", 0, 4, 32, 4, 8, 128, 1, 0, 73728, launch_tsm_atb_f64_4x32_t4x8_r128_p1, 1},
  {"tsm_atb_f64_4x32_t4x8_r128_p1_nt", 0, 4, 32, 4, 8, 128, 1, 1, 73728, launch_tsm_atb_f64_4x32_t4x8_r128_p1_nt, 1},
  {"tsm_atb_f64_4x32_t4x8_r128_p2", 0, 4, 32, 4, 8, 128, 2, 0, 73728, launch_tsm_atb_f64_4x32_t4x8_r128_p2, 1},
  {"tsm_atb_f64_4x32_t4x8_r128_p2_nt", 0, 4, 32, 4, 8, 128, 2, 1, 73728, launch_tsm_atb_f64_4x32_t4x8_r128_p2_nt, 1},
  {"tsm_atb_f64_4x64_mfma_r16_p1_nt", 0, 4, 64, 0, 0, 16, 1, 1, 20480, launch_tsm_atb_f64_4x64_mfma_r16_p1_nt, 2},
  {"tsm_atb_f64_4x64_mfma_r16_p2_nt", 0, 4, 64, 0, 0, 16, 2, 1, 20480, launch_tsm_atb_f64_4x64_mfma_r16_p2_nt, 2},
  {"tsm_atb_f64_4x64_mfma_r32_p1_nt", 0, 4, 64, 0, 0, 32, 1, 1, 40960, launch_tsm_atb_f64_4x64_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_4x64_mfma_r32_p2_nt", 0, 4, 64, 0, 0, 32, 2, 1, 40960, launch_tsm_atb_f64_4x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_4x64_mfma_r64_p1_nt", 0, 4, 64, 0, 0, 64, 1, 1, 81920, launch_tsm_atb_f64_4x64_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_4x64_mfma_r64_p2_nt", 0, 4, 64, 0, 0, 64, 2, 1, 81920, launch_tsm_atb_f64_4x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_4x64_t4x8_r32_p2_nt", 0, 4, 64, 4, 8, 32, 2, 1, 34816, launch_tsm_atb_f64_4x64_t4x8_r32_p2_nt, 1},
  {"tsm_atb_f64_8x1_t8x1_r256_p2_nt", 0, 8, 1, 8, 1, 256, 2, 1, 36864, launch_tsm_atb_f64_8x1_t8x1_r256_p2_nt, 1},
  {"tsm_atb_f64_8x2_t8x2_r256_p2_nt", 0, 8, 2, 8, 2, 256, 2, 1, 40960, launch_tsm_atb_f64_8x2_t8x2_r256_p2_nt, 1},
  {"tsm_atb_f64_8x4_t8x4_r256_p2_nt", 0, 8, 4, 8, 4, 256, 2, 1, 49152, launch_tsm_atb_f64_8x4_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f64_8x8_t2x2_r256_p1", 0, 8, 8, 2, 2, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t2x2_r256_p1, 1},
  {"tsm_atb_f64_8x8_t2x2_r256_p2", 0, 8, 8, 2, 2, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t2x2_r256_p2, 1},
  {"tsm_atb_f64_8x8_t4x4_r128_p1", 0, 8, 8, 4, 4, 128, 1, 0, 32768, launch_tsm_atb_f64_8x8_t4x4_r128_p1, 1},
  {"tsm_atb_f64_8x8_t4x4_r128_p2", 0, 8, 8, 4, 4, 128, 2, 0, 32768, launch_tsm_atb_f64_8x8_t4x4_r128_p2, 1},
  {"tsm_atb_f64_8x8_t4x4_r128_p2_nt", 0, 8, 8, 4, 4, 128, 2, 1, 32768, launch_tsm_atb_f64_8x8_t4x4_r128_p2_nt, 1},
  {"tsm_atb_f64_8x8_t4x4_r128_p4_nt", 0, 8, 8, 4, 4, 128, 4, 1, 32768, launch_tsm_atb_f64_8x8_t4x4_r128_p4_nt, 1},
  {"tsm_atb_f64_8x8_t4x4_r256_p1", 0, 8, 8, 4, 4, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p1, 1},
  {"tsm_atb_f64_8x8_t4x4_r256_p2", 0, 8, 8, 4, 4, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p2, 1},
  {"tsm_atb_f64_8x8_t4x4_r256_p2_nt", 0, 8, 8, 4, 4, 256, 2, 1, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p2_nt, 1},
  {"tsm_atb_f64_8x8_t4x4_r256_p4_nt", 0, 8, 8, 4, 4, 256, 4, 1, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p4_nt, 1},
  {"tsm_atb_f64_8x8_t4x4_r512_p1", 0, 8, 8, 4, 4, 512, 1, 0, 131072, launch_tsm_atb_f64_8x8_t4x4_r512_p1, 1},
  {"tsm_atb_f64_8x8_t4x4_r512_p2", 0, 8, 8, 4, 4, 512, 2, 0, 131072, launch_tsm_atb_f64_8x8_t4x4_r512_p2, 1},
  {"tsm_atb_f64_8x8_t4x4_r512_p2_nt", 0, 8, 8, 4, 4, 512, 2, 1, 131072, launch_tsm_atb_f64_8x8_t4x4_r512_p2_nt, 1},
  {"tsm_atb_f64_8x8_t4x8_r128_p1", 0, 8, 8, 4, 8, 128, 1, 0, 32768, launch_tsm_atb_f64_8x8_t4x8_r128_p1, 1},
  {"tsm_atb_f64_8x8_t4x8_r128_p2", 0, 8, 8, 4, 8, 128, 2, 0, 32768, launch_tsm_atb_f64_8x8_t4x8_r128_p2, 1},
  {"tsm_atb_f64_8x8_t4x8_r128_p2_nt", 0, 8, 8, 4, 8, 128, 2, 1, 32768, launch_tsm_atb_f64_8x8_t4x8_r128_p2_nt, 1},
  {"tsm_atb_f64_8x8_t4x8_r128_p4_nt", 0, 8, 8, 4, 8, 128, 4, 1, 32768, launch_tsm_atb_f64_8x8_t4x8_r128_p4_nt, 1},
  {"tsm_atb_f64_8x8_t4x8_r256_p1", 0, 8, 8, 4, 8, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t4x8_r256_p1, 1},
  {"tsm_atb_f64_8x8_t4x8_r256_p2", 0, 8, 8, 4, 8, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t4x8_r256_p2, 1},
  {"tsm_atb_f64_8x8_t4x8_r256_p2_nt", 0, 8, 8, 4, 8, 256, 2, 1, 65536, launch_tsm_atb_f64_8x8_t4x8_r256_p2_nt, 1},
  {"tsm_atb_f64_8x8_t4x8_r256_p4_nt", 0, 8, 8, 4, 8, 256, 4, 1, 65536, launch_tsm_atb_f64_8x8_t4x8_r256_p4_nt, 1},
  {"tsm_atb_f64_8x8_t4x8_r512_p1", 0, 8, 8, 4, 8, 512, 1, 0, 131072, launch_tsm_atb_f64_8x8_t4x8_r512_p1, 1},
  {"tsm_atb_f64_8x8_t4x8_r512_p2", 0, 8, 8, 4, 8, 512, 2, 0, 131072, launch_tsm_atb_f64_8x8_t4x8_r512_p2, 1},
  {"tsm_atb_f64_8x8_t4x8_r512_p2_nt", 0, 8, 8, 4, 8, 512, 2, 1, 131072, launch_tsm_atb_f64_8x8_t4x8_r512_p2_nt, 1},
  {"tsm_atb_f64_8x8_t8x4_r128_p1", 0, 8, 8, 8, 4, 128, 1, 0, 32768, launch_tsm_atb_f64_8x8_t8x4_r128_p1, 1},
  {"tsm_atb_f64_8x8_t8x4_r128_p2", 0, 8, 8, 8, 4, 128, 2, 0, 32768, launch_tsm_atb_f64_8x8_t8x4_r128_p2, 1},
  {"tsm_atb_f64_8x8_t8x4_r128_p2_nt", 0, 8, 8, 8, 4, 128, 2, 1, 32768, launch_tsm_atb_f64_8x8_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f64_8x8_t8x4_r128_p4_nt", 0, 8, 8, 8, 4, 128, 4, 1, 32768, launch_tsm_atb_f64_8x8_t8x4_r128_p4_nt, 1},
  {"tsm_atb_f64_8x8_t8x4_r256_p1", 0, 8, 8, 8, 4, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t8x4_r256_p1, 1},
  {"tsm_atb_f64_8x8_t8x4_r256_p2", 0, 8, 8, 8, 4, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t8x4_r256_p2, 1},
  {"tsm_atb_f64_8x8_t8x4_r256_p2_nt", 0, 8, 8, 8, 4, 256, 2, 1, 65536, launch_tsm_atb_f64_8x8_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f64_8x8_t8x4_r256_p4_nt", 0, 8, 8, 8, 4, 256, 4, 1, 65536, launch_tsm_atb_f64_8x8_t8x4_r256_p4_nt, 1},
  {"tsm_atb_f64_8x8_t8x4_r512_p1", 0, 8, 8, 8, 4, 512, 1, 0, 131072, launch_tsm_atb_f64_8x8_t8x4_r512_p1, 1},
  {"tsm_atb_f64_8x8_t8x4_r512_p2", 0, 8, 8, 8, 4, 512, 2, 0, 131072, launch_tsm_atb_f64_8x8_t8x4_r512_p2, 1},
  {"tsm_atb_f64_8x8_t8x4_r512_p2_nt", 0, 8, 8, 8, 4, 512, 2, 1, 131072, launch_tsm_atb_f64_8x8_t8x4_r512_p2_nt, 1},
  {"tsm_atb_f64_8x8_t8x8_r256_p1", 0, 8, 8, 8, 8, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p1, 1},
  {"tsm_atb_f64_8x8_t8x8_r256_p1_nt", 0, 8, 8, 8, 8, 256, 1, 1, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p1_nt, 1},
  {"tsm_atb_f64_8x8_t8x8_r256_p2", 0, 8, 8, 8, 8, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p2, 1},
  {"tsm_atb_f64_8x8_t8x8_r256_p2_nt", 0, 8, 8, 8, 8, 256, 2, 1, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p2_nt, 1},
  {"tsm_atb_f64_8x8_t8x8_r512_p1", 0, 8, 8, 8, 8, 512, 1, 0, 131072, launch_tsm_atb_f64_8x8_t8x8_r512_p1, 1},
  {"tsm_atb_f64_8x8_t8x8_r512_p1_nt", 0, 8, 8, 8, 8, 512, 1, 1, 131072, launch_tsm_atb_f64_8x8_t8x8_r512_p1_nt, 1},
  {"tsm_atb_f64_8x8_t8x8_r512_p2", 0, 8, 8, 8, 8, 512, 2, 0, 131072, launch_tsm_atb_f64_8x8_t8x8_r512_p2, 1},
  {"tsm_atb_f64_8x8_t8x8_r512_p2_nt", 0, 8, 8, 8, 8, 512, 2, 1, 131072, launch_tsm_atb_f64_8x8_t8x8_r512_p2_nt, 1},
  {"tsm_atb_f64_8x16_mfma_r64_p1_nt", 0, 8, 16, 0, 0, 64, 1, 1, 32768, launch_tsm_atb_f64_8x16_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_8x16_mfma_r64_p2_nt", 0, 8, 16, 0, 0, 64, 2, 1, 32768, launch_tsm_atb_f64_8x16_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_8x16_mfma_r128_p1_nt", 0, 8, 16, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_f64_8x16_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_8x16_mfma_r128_p2_nt", 0, 8, 16, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_f64_8x16_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_8x16_mfma_r256_p1_nt", 0, 8, 16, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_f64_8x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_f64_8x16_mfma_r256_p2_nt", 0, 8, 16, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_f64_8x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_f64_8x16_t8x8_r128_p2_nt", 0, 8, 16, 8, 8, 128, 2, 1, 49152, launch_tsm_atb_f64_8x16_t8x8_r128_p2_nt, 1},
  {"tsm_atb_f64_8x32_mfma_r32_p1_nt", 0, 8, 32, 0, 0, 32, 1, 1, 24576, launch_tsm_atb_f64_8x32_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_8x32_mfma_r32_p2_nt", 0, 8, 32, 0, 0, 32, 2, 1, 24576, launch_tsm_atb_f64_8x32_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_8x32_mfma_r64_p1_nt", 0, 8, 32, 0, 0, 64, 1, 1, 49152, launch_tsm_atb_f64_8x32_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_8x32_mfma_r64_p2_nt", 0, 8, 32, 0, 0, 64, 2, 1, 49152, launch_tsm_atb_f64_8x32_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_8x32_mfma_r128_p1_nt", 0, 8, 32, 0, 0, 128, 1, 1, 98304, launch_tsm_atb_f64_8x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_8x32_mfma_r128_p2_nt", 0, 8, 32, 0, 0, 128, 2, 1, 98304, launch_tsm_atb_f64_8x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_8x32_t8x8_r64_p2_nt", 0, 8, 32, 8, 8, 64, 2, 1, 40960, launch_tsm_atb_f64_8x32_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f64_8x64_mfma_r16_p1_nt", 0, 8, 64, 0, 0, 16, 1, 1, 20480, launch_tsm_atb_f64_8x64_mfma_r16_p1_nt, 2},
  {"tsm_atb_f64_8x64_mfma_r16_p2_nt", 0, 8, 64, 0, 0, 16, 2, 1, 20480, launch_tsm_atb_f64_8x64_mfma_r16_p2_nt, 2},
  {"tsm_atb_f64_8x64_mfma_r32_p1_nt", 0, 8, 64, 0, 0, 32, 1, 1, 40960, launch_tsm_atb_f64_8x64_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_8x64_mfma_r32_p2_nt", 0, 8, 64, 0, 0, 32, 2, 1, 40960, launch_tsm_atb_f64_8x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_8x64_mfma_r64_p1_nt", 0, 8, 64, 0, 0, 64, 1, 1, 81920, launch_tsm_atb_f64_8x64_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_8x64_mfma_r64_p2_nt", 0, 8, 64, 0, 0, 64, 2, 1, 81920, launch_tsm_atb_f64_8x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_8x64_t8x8_r32_p2_nt", 0, 8, 64, 8, 8, 32, 2, 1, 36864, launch_tsm_atb_f64_8x64_t8x8_r32_p2_nt, 1},
  {"tsm_atb_f64_16x1_t8x1_r128_p2_nt", 0, 16, 1, 8, 1, 128, 2, 1, 34816, launch_tsm_atb_f64_16x1_t8x1_r128_p2_nt, 1},
  {"tsm_atb_f64_16x2_mfma_r64_p1_nt", 0, 16, 2, 0, 0, 64, 1, 1, 32768, launch_tsm_atb_f64_16x2_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_16x2_mfma_r64_p2_nt", 0, 16, 2, 0, 0, 64, 2, 1, 32768, launch_tsm_atb_f64_16x2_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_16x2_mfma_r128_p1_nt", 0, 16, 2, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_f64_16x2_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_16x2_mfma_r128_p2_nt", 0, 16, 2, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_f64_16x2_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_16x2_mfma_r256_p1_nt", 0, 16, 2, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_f64_16x2_mfma_r256_p1_nt, 2},
  {"tsm_atb_f64_16x2_mfma_r256_p2_nt", 0, 16, 2, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_f64_16x2_mfma_r256_p2_nt, 2},
  {"tsm_atb_f64_16x2_t8x2_r128_p2_nt", 0, 16, 2, 8, 2, 128, 2, 1, 36864, launch_tsm_atb_f64_16x2_t8x2_r128_p2_nt, 1},
  {"tsm_atb_f64_16x4_mfma_r64_p1_nt", 0, 16, 4, 0, 0, 64, 1, 1, 32768, launch_tsm_atb_f64_16x4_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_16x4_mfma_r64_p2_nt", 0, 16, 4, 0, 0, 64, 2, 1, 32768, launch_tsm_atb_f64_16x4_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_16x4_mfma_r128_p1_nt", 0, 16, 4, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_f64_16x4_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_16x4_mfma_r128_p2_nt", 0, 16, 4, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_f64_16x4_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_16x4_mfma_r256_p1_nt", 0, 16, 4, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_f64_16x4_mfma_r256_p1_nt, 2},
  {"tsm_atb_f64_16x4_mfma_r256_p2_nt", 0, 16, 4, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_f64_16x4_mfma_r256_p2_nt, 2},
  {"tsm_atb_f64_16x4_t8x4_r128_p2_nt", 0, 16, 4, 8, 4, 128, 2, 1, 40960, launch_tsm_atb_f64_16x4_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f64_16x8_mfma_r64_p1_nt", 0, 16, 8, 0, 0, 64, 1, 1, 32768, launch_tsm_atb_f64_16x8_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_16x8_mfma_r64_p2_nt", 0, 16, 8, 0, 0, 64, 2, 1, 32768, launch_tsm_atb_f64_16x8_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_16x8_mfma_r128_p1_nt", 0, 16, 8, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_f64_16x8_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_16x8_mfma_r128_p2_nt", 0, 16, 8, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_f64_16x8_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_16x8_mfma_r256_p1_nt", 0, 16, 8, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_f64_16x8_mfma_r256_p1_nt, 2},
  {"tsm_atb_f64_16x8_mfma_r256_p2_nt", 0, 16, 8, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_f64_16x8_mfma_r256_p2_nt, 2},
  {"tsm_atb_f64_16x8_t8x8_r128_p2_nt", 0, 16, 8, 8, 8, 128, 2, 1, 49152, launch_tsm_atb_f64_16x8_t8x8_r128_p2_nt, 1},
  {"tsm_atb_f64_16x16_direct_d4_nt", 0, 16, 16, 0, 0, 4, 4, 1, 4096, launch_tsm_atb_f64_16x16_direct_d4_nt, 3},
  {"tsm_atb_f64_16x16_direct_d8_nt", 0, 16, 16, 0, 0, 4, 8, 1, 4096, launch_tsm_atb_f64_16x16_direct_d8_nt, 3},
  {"tsm_atb_f64_16x16_direct_d16_nt", 0, 16, 16, 0, 0, 4, 16, 1, 4096, launch_tsm_atb_f64_16x16_direct_d16_nt, 3},
  {"tsm_atb_f64_16x16_mfma_r64_p1", 0, 16, 16, 0, 0, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_mfma_r64_p1, 2},
  {"tsm_atb_f64_16x16_mfma_r64_p1_nt", 0, 16, 16, 0, 0, 64, 1, 1, 32768, launch_tsm_atb_f64_16x16_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_16x16_mfma_r64_p2", 0, 16, 16, 0, 0, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_mfma_r64_p2, 2},
  {"tsm_atb_f64_16x16_mfma_r64_p2_nt", 0, 16, 16, 0, 0, 64, 2, 1, 32768, launch_tsm_atb_f64_16x16_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_16x16_mfma_r128_p1", 0, 16, 16, 0, 0, 128, 1, 0, 65536, launch_tsm_atb_f64_16x16_mfma_r128_p1, 2},
  {"tsm_atb_f64_16x16_mfma_r128_p1_nt", 0, 16, 16, 0, 0, 128, 1, 1, 65536, launch_tsm_atb_f64_16x16_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_16x16_mfma_r128_p2", 0, 16, 16, 0, 0, 128, 2, 0, 65536, launch_tsm_atb_f64_16x16_mfma_r128_p2, 2},
  {"tsm_atb_f64_16x16_mfma_r128_p2_nt", 0, 16, 16, 0, 0, 128, 2, 1, 65536, launch_tsm_atb_f64_16x16_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_16x16_mfma_r256_p1", 0, 16, 16, 0, 0, 256, 1, 0, 131072, launch_tsm_atb_f64_16x16_mfma_r256_p1, 2},
  {"tsm_atb_f64_16x16_mfma_r256_p1_nt", 0, 16, 16, 0, 0, 256, 1, 1, 131072, launch_tsm_atb_f64_16x16_mfma_r256_p1_nt, 2},
  {"tsm_atb_f64_16x16_mfma_r256_p2", 0, 16, 16, 0, 0, 256, 2, 0, 131072, launch_tsm_atb_f64_16x16_mfma_r256_p2, 2},
  {"tsm_atb_f64_16x16_mfma_r256_p2_nt", 0, 16, 16, 0, 0, 256, 2, 1, 131072, launch_tsm_atb_f64_16x16_mfma_r256_p2_nt, 2},
  {"tsm_atb_f64_16x16_t2x2_r64_p1", 0, 16, 16, 2, 2, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_t2x2_r64_p1, 1},
  {"tsm_atb_f64_16x16_t2x2_r64_p2", 0, 16, 16, 2, 2, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_t2x2_r64_p2, 1},
  {"tsm_atb_f64_16x16_t4x4_r64_p1", 0, 16, 16, 4, 4, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_t4x4_r64_p1, 1},
  {"tsm_atb_f64_16x16_t4x4_r64_p2", 0, 16, 16, 4, 4, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_t4x4_r64_p2, 1},
  {"tsm_atb_f64_16x16_t4x4_r64_p2_nt", 0, 16, 16, 4, 4, 64, 2, 1, 32768, launch_tsm_atb_f64_16x16_t4x4_r64_p2_nt, 1},
  {"tsm_atb_f64_16x16_t4x4_r64_p4_nt", 0, 16, 16, 4, 4, 64, 4, 1, 32768, launch_tsm_atb_f64_16x16_t4x4_r64_p4_nt, 1},
  {"tsm_atb_f64_16x16_t4x4_r128_p1", 0, 16, 16, 4, 4, 128, 1, 0, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p1, 1},
  {"tsm_atb_f64_16x16_t4x4_r128_p2", 0, 16, 16, 4, 4, 128, 2, 0, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p2, 1},
  {"tsm_atb_f64_16x16_t4x4_r128_p2_nt", 0, 16, 16, 4, 4, 128, 2, 1, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p2_nt, 1},
  {"tsm_atb_f64_16x16_t4x4_r128_p4_nt", 0, 16, 16, 4, 4, 128, 4, 1, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p4_nt, 1},
  {"tsm_atb_f64_16x16_t4x4_r256_p1", 0, 16, 16, 4, 4, 256, 1, 0, 131072, launch_tsm_atb_f64_16x16_t4x4_r256_p1, 1},
  {"tsm_atb_f64_16x16_t4x4_r256_p2", 0, 16, 16, 4, 4, 256, 2, 0, 131072, launch_tsm_atb_f64_16x16_t4x4_r256_p2, 1},
  {"tsm_atb_f64_16x16_t4x4_r256_p2_nt", 0, 16, 16, 4, 4, 256, 2, 1, 131072, launch_tsm_atb_f64_16x16_t4x4_r256_p2_nt, 1},
  {"tsm_atb_f64_16x16_t4x8_r64_p1", 0, 16, 16, 4, 8, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_t4x8_r64_p1, 1},
  {"tsm_atb_f64_16x16_t4x8_r64_p2", 0, 16, 16, 4, 8, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_t4x8_r64_p2, 1},
  {"tsm_atb_f64_16x16_t4x8_r64_p2_nt", 0, 16, 16, 4, 8, 64, 2, 1, 32768, launch_tsm_atb_f64_16x16_t4x8_r64_p2_nt, 1},
  {"tsm_atb_f64_16x16_t4x8_r64_p4_nt", 0, 16, 16, 4, 8, 64, 4, 1, 32768, launch_tsm_atb_f64_16x16_t4x8_r64_p4_nt, 1},
  {"tsm_atb_f64_16x16_t4x8_r128_p1", 0, 16, 16, 4, 8, 128, 1, 0, 65536, launch_tsm_atb_f64_16x16_t4x8_r128_p1, 1},
  {"tsm_atb_f64_16x16_t4x8_r128_p2", 0, 16, 16, 4, 8, 128, 2, 0, 65536, launch_tsm_atb_f64_16x16_t4x8_r128_p2, 1},
  {"tsm_atb_f64_16x16_t4x8_r128_p2_nt", 0, 16, 16, 4, 8, 128, 2, 1, 65536, launch_tsm_atb_f64_16x16_t4x8_r128_p2_nt, 1},
  {"tsm_atb_f64_16x16_t4x8_r128_p4_nt", 0, 16, 16, 4, 8, 128, 4, 1, 65536, launch_tsm_atb_f64_16x16_t4x8_r128_p4_nt, 1},
  {"tsm_atb_f64_16x16_t4x8_r256_p1", 0, 16, 16, 4, 8, 256, 1, 0, 131072, launch_tsm_atb_f64_16x16_t4x8_r256_p1, 1},
  {"tsm_atb_f64_16x16_t4x8_r256_p2", 0, 16, 16, 4, 8, 256, 2, 0, 131072, launch_tsm_atb_f64_16x16_t4x8_r256_p2, 1},
  {"tsm_atb_f64_16x16_t4x8_r256_p2_nt", 0, 16, 16, 4, 8, 256, 2, 1, 131072, launch_tsm_atb_f64_16x16_t4x8_r256_p2_nt, 1},
  {"tsm_atb_f64_16x16_t8x4_r64_p1", 0, 16, 16, 8, 4, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_t8x4_r64_p1, 1},
  {"tsm_atb_f64_16x16_t8x4_r64_p2", 0, 16, 16, 8, 4, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_t8x4_r64_p2, 1},
  {"tsm_atb_f64_16x16_t8x4_r64_p2_nt", 0, 16, 16, 8, 4, 64, 2, 1, 32768, launch_tsm_atb_f64_16x16_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f64_16x16_t8x4_r64_p4_nt", 0, 16, 16, 8, 4, 64, 4, 1, 32768, launch_tsm_atb_f64_16x16_t8x4_r64_p4_nt, 1},
  {"tsm_atb_f64_16x16_t8x4_r128_p1", 0, 16, 16, 8, 4, 128, 1, 0, 65536, launch_tsm_atb_f64_16x16_t8x4_r128_p1, 1},
  {"tsm_atb_f64_16x16_t8x4_r128_p2", 0, 16, 16, 8, 4, 128, 2, 0, 65536, launch_tsm_atb_f64_16x16_t8x4_r128_p2, 1},
  {"tsm_atb_f64_16x16_t8x4_r128_p2_nt", 0, 16, 16, 8, 4, 128, 2, 1, 65536, launch_tsm_atb_f64_16x16_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f64_16x16_t8x4_r128_p4_nt", 0, 16, 16, 8, 4, 128, 4, 1, 65536, launch_tsm_atb_f64_16x16_t8x4_r128_p4_nt, 1},
  {"tsm_atb_f64_16x16_t8x4_r256_p1", 0, 16, 16, 8, 4, 256, 1, 0, 131072, launch_tsm_atb_f64_16x16_t8x4_r256_p1, 1},
  {"tsm_atb_f64_16x16_t8x4_r256_p2", 0, 16, 16, 8, 4, 256, 2, 0, 131072, launch_tsm_atb_f64_16x16_t8x4_r256_p2, 1},
  {"tsm_atb_f64_16x16_t8x4_r256_p2_nt", 0, 16, 16, 8, 4, 256, 2, 1, 131072, launch_tsm_atb_f64_16x16_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f64_16x16_t8x8_r64_p1", 0, 16, 16, 8, 8, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p1, 1},
  {"tsm_atb_f64_16x16_t8x8_r64_p1_nt", 0, 16, 16, 8, 8, 64, 1, 1, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p1_nt, 1},
  {"tsm_atb_f64_16x16_t8x8_r64_p2", 0, 16, 16, 8, 8, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p2, 1},
  {"tsm_atb_f64_16x16_t8x8_r64_p2_nt", 0, 16, 16, 8, 8, 64, 2, 1, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f64_16x16_t8x8_r128_p1", 0, 16, 16, 8, 8, 128, 1, 0, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p1, 1},
  {"tsm_atb_f64_16x16_t8x8_r128_p1_nt", 0, 16, 16, 8, 8, 128, 1, 1, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p1_nt, 1},
  {"tsm_atb_f64_16x16_t8x8_r128_p2", 0, 16, 16, 8, 8, 128, 2, 0, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p2, 1},
  {"tsm_atb_f64_16x16_t8x8_r128_p2_nt", 0, 16, 16, 8, 8, 128, 2, 1, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p2_nt, 1},
  {"tsm_atb_f64_16x16_t8x8_r256_p1", 0, 16, 16, 8, 8, 256, 1, 0, 131072, launch_tsm_atb_f64_16x16_t8x8_r256_p1, 1},
  {"tsm_atb_f64_16x16_t8x8_r256_p1_nt", 0, 16, 16, 8, 8, 256, 1, 1, 131072, launch_tsm_atb_f64_16x16_t8x8_r256_p1_nt, 1},
  {"tsm_atb_f64_16x16_t8x8_r256_p2", 0, 16, 16, 8, 8, 256, 2, 0, 131072, launch_tsm_atb_f64_16x16_t8x8_r256_p2, 1},
  {"tsm_atb_f64_16x16_t8x8_r256_p2_nt", 0, 16, 16, 8, 8, 256, 2, 1, 131072, launch_tsm_atb_f64_16x16_t8x8_r256_p2_nt, 1},
  {"tsm_atb_f64_16x32_direct_d8_nt", 0, 16, 32, 0, 0, 4, 8, 1, 8192, launch_tsm_atb_f64_16x32_direct_d8_nt, 3},
  {"tsm_atb_f64_16x32_t8x8_r64_p2_nt", 0, 16, 32, 8, 8, 64, 2, 1, 49152, launch_tsm_atb_f64_16x32_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f64_16x64_direct_d8_nt", 0, 16, 64, 0, 0, 4, 8, 1, 16384, launch_tsm_atb_f64_16x64_direct_d8_nt, 3},
  {"tsm_atb_f64_16x64_mfma_r32_p2_nt", 0, 16, 64, 0, 0, 32, 2, 1, 40960, launch_tsm_atb_f64_16x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_32x1_t8x1_r64_p2_nt", 0, 32, 1, 8, 1, 64, 2, 1, 33792, launch_tsm_atb_f64_32x1_t8x1_r64_p2_nt, 1},
  {"tsm_atb_f64_32x2_mfma_r32_p1_nt", 0, 32, 2, 0, 0, 32, 1, 1, 24576, launch_tsm_atb_f64_32x2_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_32x2_mfma_r32_p2_nt", 0, 32, 2, 0, 0, 32, 2, 1, 24576, launch_tsm_atb_f64_32x2_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_32x2_mfma_r64_p1_nt", 0, 32, 2, 0, 0, 64, 1, 1, 49152, launch_tsm_atb_f64_32x2_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_32x2_mfma_r64_p2_nt", 0, 32, 2, 0, 0, 64, 2, 1, 49152, launch_tsm_atb_f64_32x2_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_32x2_mfma_r128_p1_nt", 0, 32, 2, 0, 0, 128, 1, 1, 98304, launch_tsm_atb_f64_32x2_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_32x2_mfma_r128_p2_nt", 0, 32, 2, 0, 0, 128, 2, 1, 98304, launch_tsm_atb_f64_32x2_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_32x2_t8x2_r64_p2_nt", 0, 32, 2, 8, 2, 64, 2, 1, 34816, launch_tsm_atb_f64_32x2_t8x2_r64_p2_nt, 1},
  {"tsm_atb_f64_32x4_mfma_r32_p1_nt", 0, 32, 4, 0, 0, 32, 1, 1, 24576, launch_tsm_atb_f64_32x4_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_32x4_mfma_r32_p2_nt", 0, 32, 4, 0, 0, 32, 2, 1, 24576, launch_tsm_atb_f64_32x4_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_32x4_mfma_r64_p1_nt", 0, 32, 4, 0, 0, 64, 1, 1, 49152, launch_tsm_atb_f64_32x4_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_32x4_mfma_r64_p2_nt", 0, 32, 4, 0, 0, 64, 2, 1, 49152, launch_tsm_atb_f64_32x4_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_32x4_mfma_r128_p1_nt", 0, 32, 4, 0, 0, 128, 1, 1, 98304, launch_tsm_atb_f64_32x4_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_32x4_mfma_r128_p2_nt", 0, 32, 4, 0, 0, 128, 2, 1, 98304, launch_tsm_atb_f64_32x4_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_32x4_t4x4_r32_p1", 0, 32, 4, 4, 4, 32, 1, 0, 18432, launch_tsm_atb_f64_32x4_t4x4_r32_p1, 1},
  {"tsm_atb_f64_32x4_t4x4_r32_p2", 0, 32, 4, 4, 4, 32, 2, 0, 18432, launch_tsm_atb_f64_32x4_t4x4_r32_p2, 1},
  {"tsm_atb_f64_32x4_t4x4_r32_p2_nt", 0, 32, 4, 4, 4, 32, 2, 1, 18432, launch_tsm_atb_f64_32x4_t4x4_r32_p2_nt, 1},
  {"tsm_atb_f64_32x4_t4x4_r32_p4_nt", 0, 32, 4, 4, 4, 32, 4, 1, 18432, launch_tsm_atb_f64_32x4_t4x4_r32_p4_nt, 1},
  {"tsm_atb_f64_32x4_t4x4_r64_p1", 0, 32, 4, 4, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p1, 1},
  {"tsm_atb_f64_32x4_t4x4_r64_p2", 0, 32, 4, 4, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p2, 1},
  {"tsm_atb_f64_32x4_t4x4_r64_p2_nt", 0, 32, 4, 4, 4, 64, 2, 1, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p2_nt, 1},
  {"tsm_atb_f64_32x4_t4x4_r64_p4_nt", 0, 32, 4, 4, 4, 64, 4, 1, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p4_nt, 1},
  {"tsm_atb_f64_32x4_t4x4_r128_p1", 0, 32, 4, 4, 4, 128, 1, 0, 73728, launch_tsm_atb_f64_32x4_t4x4_r128_p1, 1},
  {"tsm_atb_f64_32x4_t4x4_r128_p2", 0, 32, 4, 4, 4, 128, 2, 0, 73728, launch_tsm_atb_f64_32x4_t4x4_r128_p2, 1},
  {"tsm_atb_f64_32x4_t4x4_r128_p2_nt", 0, 32, 4, 4, 4, 128, 2, 1, 73728, launch_tsm_atb_f64_32x4_t4x4_r128_p2_nt, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p1", 0, 32, 4, 8, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p1, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p1_nt", 0, 32, 4, 8, 4, 64, 1, 1, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p1_nt, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p2", 0, 32, 4, 8, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p2, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p2_nt", 0, 32, 4, 8, 4, 64, 2, 1, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p4_nt", 0, 32, 4, 8, 4, 64, 4, 1, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p4_nt, 1},
  {"tsm_atb_f64_32x4_t8x4_r128_p1", 0, 32, 4, 8, 4, 128, 1, 0, 73728, launch_tsm_atb_f64_32x4_t8x4_r128_p1, 1},
  {"tsm_atb_f64_32x4_t8x4_r128_p1_nt", 0, 32, 4, 8, 4, 128, 1, 1, 73728, launch_tsm_atb_f64_32x4_t8x4_r128_p1_nt, 1},
  {"tsm_atb_f64_32x4_t8x4_r128_p2", 0, 32, 4, 8, 4, 128, 2, 0, 73728, launch_tsm_atb_f64_32x4_t8x4_r128_p2, 1},
  {"tsm_atb_f64_32x4_t8x4_r128_p2_nt", 0, 32, 4, 8, 4, 128, 2, 1, 73728, launch_tsm_atb_f64_32x4_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f64_32x8_mfma_r32_p1_nt", 0, 32, 8, 0, 0, 32, 1, 1, 24576, launch_tsm_atb_f64_32x8_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_32x8_mfma_r32_p2_nt", 0, 32, 8, 0, 0, 32, 2, 1, 24576, launch_tsm_atb_f64_32x8_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_32x8_mfma_r64_p1_nt", 0, 32, 8, 0, 0, 64, 1, 1, 49152, launch_tsm_atb_f64_32x8_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_32x8_mfma_r64_p2_nt", 0, 32, 8, 0, 0, 64, 2, 1, 49152, launch_tsm_atb_f64_32x8_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_32x8_mfma_r128_p1_nt", 0, 32, 8, 0, 0, 128, 1, 1, 98304, launch_tsm_atb_f64_32x8_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_32x8_mfma_r128_p2_nt", 0, 32, 8, 0, 0, 128, 2, 1, 98304, launch_tsm_atb_f64_32x8_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_32x8_t8x8_r64_p2_nt", 0, 32, 8, 8, 8, 64, 2, 1, 40960, launch_tsm_atb_f64_32x8_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f64_32x16_direct_d8_nt", 0, 32, 16, 0, 0, 4, 8, 1, 8192, launch_tsm_atb_f64_32x16_direct_d8_nt, 3},
  {"tsm_atb_f64_32x16_t8x8_r64_p2_nt", 0, 32, 16, 8, 8, 64, 2, 1, 49152, launch_tsm_atb_f64_32x16_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f64_32x32_direct_d4_nt", 0, 32, 32, 0, 0, 4, 4, 1, 16384, launch_tsm_atb_f64_32x32_direct_d4_nt, 3},
  {"tsm_atb_f64_32x32_direct_d8_nt", 0, 32, 32, 0, 0, 4, 8, 1, 16384, launch_tsm_atb_f64_32x32_direct_d8_nt, 3},
  {"tsm_atb_f64_32x32_direct_d16_nt", 0, 32, 32, 0, 0, 4, 16, 1, 16384, launch_tsm_atb_f64_32x32_direct_d16_nt, 3},
  {"tsm_atb_f64_32x32_mfma_r32_p1", 0, 32, 32, 0, 0, 32, 1, 0, 32768, launch_tsm_atb_f64_32x32_mfma_r32_p1, 2},
  {"tsm_atb_f64_32x32_mfma_r32_p1_nt", 0, 32, 32, 0, 0, 32, 1, 1, 32768, launch_tsm_atb_f64_32x32_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_32x32_mfma_r32_p2", 0, 32, 32, 0, 0, 32, 2, 0, 32768, launch_tsm_atb_f64_32x32_mfma_r32_p2, 2},
  {"tsm_atb_f64_32x32_mfma_r32_p2_nt", 0, 32, 32, 0, 0, 32, 2, 1, 32768, launch_tsm_atb_f64_32x32_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_32x32_mfma_r64_p1", 0, 32, 32, 0, 0, 64, 1, 0, 65536, launch_tsm_atb_f64_32x32_mfma_r64_p1, 2},
  {"tsm_atb_f64_32x32_mfma_r64_p1_nt", 0, 32, 32, 0, 0, 64, 1, 1, 65536, launch_tsm_atb_f64_32x32_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_32x32_mfma_r64_p2", 0, 32, 32, 0, 0, 64, 2, 0, 65536, launch_tsm_atb_f64_32x32_mfma_r64_p2, 2},
  {"tsm_atb_f64_32x32_mfma_r64_p2_nt", 0, 32, 32, 0, 0, 64, 2, 1, 65536, launch_tsm_atb_f64_32x32_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_32x32_mfma_r128_p1", 0, 32, 32, 0, 0, 128, 1, 0, 131072, launch_tsm_atb_f64_32x32_mfma_r128_p1, 2},
  {"tsm_atb_f64_32x32_mfma_r128_p1_nt", 0, 32, 32, 0, 0, 128, 1, 1, 131072, launch_tsm_atb_f64_32x32_mfma_r128_p1_nt, 2},
  {"tsm_atb_f64_32x32_mfma_r128_p2", 0, 32, 32, 0, 0, 128, 2, 0, 131072, launch_tsm_atb_f64_32x32_mfma_r128_p2, 2},
  {"tsm_atb_f64_32x32_mfma_r128_p2_nt", 0, 32, 32, 0, 0, 128, 2, 1, 131072, launch_tsm_atb_f64_32x32_mfma_r128_p2_nt, 2},
  {"tsm_atb_f64_32x32_t4x4_r32_p1", 0, 32, 32, 4, 4, 32, 1, 0, 32768, launch_tsm_atb_f64_32x32_t4x4_r32_p1, 1},
  {"tsm_atb_f64_32x32_t4x4_r32_p2", 0, 32, 32, 4, 4, 32, 2, 0, 32768, launch_tsm_atb_f64_32x32_t4x4_r32_p2, 1},
  {"tsm_atb_f64_32x32_t4x4_r32_p2_nt", 0, 32, 32, 4, 4, 32, 2, 1, 32768, launch_tsm_atb_f64_32x32_t4x4_r32_p2_nt, 1},
  {"tsm_atb_f64_32x32_t4x4_r32_p4_nt", 0, 32, 32, 4, 4, 32, 4, 1, 32768, launch_tsm_atb_f64_32x32_t4x4_r32_p4_nt, 1},
  {"tsm_atb_f64_32x32_t4x4_r64_p1", 0, 32, 32, 4, 4, 64, 1, 0, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p1, 1},
  {"tsm_atb_f64_32x32_t4x4_r64_p2", 0, 32, 32, 4, 4, 64, 2, 0, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p2, 1},
  {"tsm_atb_f64_32x32_t4x4_r64_p2_nt", 0, 32, 32, 4, 4, 64, 2, 1, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p2_nt, 1},
  {"tsm_atb_f64_32x32_t4x4_r64_p4_nt", 0, 32, 32, 4, 4, 64, 4, 1, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p4_nt, 1},
  {"tsm_atb_f64_32x32_t4x8_r32_p1", 0, 32, 32, 4, 8, 32, 1, 0, 32768, launch_tsm_atb_f64_32x32_t4x8_r32_p1, 1},
  {"tsm_atb_f64_32x32_t4x8_r32_p2", 0, 32, 32, 4, 8, 32, 2, 0, 32768, launch_tsm_atb_f64_32x32_t4x8_r32_p2, 1},
  {"tsm_atb_f64_32x32_t4x8_r32_p2_nt", 0, 32, 32, 4, 8, 32, 2, 1, 32768, launch_tsm_atb_f64_32x32_t4x8_r32_p2_nt, 1},
  {"tsm_atb_f64_32x32_t4x8_r32_p4_nt", 0, 32, 32, 4, 8, 32, 4, 1, 32768, launch_tsm_atb_f64_32x32_t4x8_r32_p4_nt, 1},
  {"tsm_atb_f64_32x32_t4x8_r64_p1", 0, 32, 32, 4, 8, 64, 1, 0, 65536, launch_tsm_atb_f64_32x32_t4x8_r64_p1, 1},
  {"tsm_atb_f64_32x32_t4x8_r64_p2", 0, 32, 32, 4, 8, 64, 2, 0, 65536, launch_tsm_atb_f64_32x32_t4x8_r64_p2, 1},
  {"tsm_atb_f64_32x32_t4x8_r64_p2_nt", 0, 32, 32, 4, 8, 64, 2, 1, 65536, launch_tsm_atb_f64_32x32_t4x8_r64_p2_nt, 1},
  {"tsm_atb_f64_32x32_t4x8_r64_p4_nt", 0, 32, 32, 4, 8, 64, 4, 1, 65536, launch_tsm_atb_f64_32x32_t4x8_r64_p4_nt, 1},
  {"tsm_atb_f64_32x32_t4x8_r128_p1", 0, 32, 32, 4, 8, 128, 1, 0, 131072, launch_tsm_atb_f64_32x32_t4x8_r128_p1, 1},
  {"tsm_atb_f64_32x32_t4x8_r128_p2", 0, 32, 32, 4, 8, 128, 2, 0, 131072, launch_tsm_atb_f64_32x32_t4x8_r128_p2, 1},
  {"tsm_atb_f64_32x32_t4x8_r128_p2_nt", 0, 32, 32, 4, 8, 128, 2, 1, 131072, launch_tsm_atb_f64_32x32_t4x8_r128_p2_nt, 1},
  {"tsm_atb_f64_32x32_t8x4_r32_p1", 0, 32, 32, 8, 4, 32, 1, 0, 32768, launch_tsm_atb_f64_32x32_t8x4_r32_p1, 1},
  {"tsm_atb_f64_32x32_t8x4_r32_p2", 0, 32, 32, 8, 4, 32, 2, 0, 32768, launch_tsm_atb_f64_32x32_t8x4_r32_p2, 1},
  {"tsm_atb_f64_32x32_t8x4_r32_p2_nt", 0, 32, 32, 8, 4, 32, 2, 1, 32768, launch_tsm_atb_f64_32x32_t8x4_r32_p2_nt, 1},
  {"tsm_atb_f64_32x32_t8x4_r32_p4_nt", 0, 32, 32, 8, 4, 32, 4, 1, 32768, launch_tsm_atb_f64_32x32_t8x4_r32_p4_nt, 1},
  {"tsm_atb_f64_32x32_t8x4_r64_p1", 0, 32, 32, 8, 4, 64, 1, 0, 65536, launch_tsm_atb_f64_32x32_t8x4_r64_p1, 1},
  {"tsm_atb_f64_32x32_t8x4_r64_p2", 0, 32, 32, 8, 4, 64, 2, 0, 65536, launch_tsm_atb_f64_32x32_t8x4_r64_p2, 1},
  {"tsm_atb_f64_32x32_t8x4_r64_p2_nt", 0, 32, 32, 8, 4, 64, 2, 1, 65536, launch_tsm_atb_f64_32x32_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f64_32x32_t8x4_r64_p4_nt", 0, 32, 32, 8, 4, 64, 4, 1, 65536, launch_tsm_atb_f64_32x32_t8x4_r64_p4_nt, 1},
  {"tsm_atb_f64_32x32_t8x4_r128_p1", 0, 32, 32, 8, 4, 128, 1, 0, 131072, launch_tsm_atb_f64_32x32_t8x4_r128_p1, 1},
  {"tsm_atb_f64_32x32_t8x4_r128_p2", 0, 32, 32, 8, 4, 128, 2, 0, 131072, launch_tsm_atb_f64_32x32_t8x4_r128_p2, 1},
  {"tsm_atb_f64_32x32_t8x4_r128_p2_nt", 0, 32, 32, 8, 4, 128, 2, 1, 131072, launch_tsm_atb_f64_32x32_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f64_32x32_t8x8_r32_p1", 0, 32, 32, 8, 8, 32, 1, 0, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p1, 1},
  {"tsm_atb_f64_32x32_t8x8_r32_p1_nt", 0, 32, 32, 8, 8, 32, 1, 1, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p1_nt, 1},
  {"tsm_atb_f64_32x32_t8x8_r32_p2", 0, 32, 32, 8, 8, 32, 2, 0, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p2, 1},
  {"tsm_atb_f64_32x32_t8x8_r32_p2_nt", 0, 32, 32, 8, 8, 32, 2, 1, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p2_nt, 1},
  {"tsm_atb_f64_32x32_t8x8_r64_p1", 0, 32, 32, 8, 8, 64, 1, 0, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p1, 1},
  {"tsm_atb_f64_32x32_t8x8_r64_p1_nt", 0, 32, 32, 8, 8, 64, 1, 1, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p1_nt, 1},
  {"tsm_atb_f64_32x32_t8x8_r64_p2", 0, 32, 32, 8, 8, 64, 2, 0, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p2, 1},
  {"tsm_atb_f64_32x32_t8x8_r64_p2_nt", 0, 32, 32, 8, 8, 64, 2, 1, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f64_32x32_t8x8_r128_p1", 0, 32, 32, 8, 8, 128, 1, 0, 131072, launch_tsm_atb_f64_32x32_t8x8_r128_p1, 1},
  {"tsm_atb_f64_32x32_t8x8_r128_p1_nt", 0, 32, 32, 8, 8, 128, 1, 1, 131072, launch_tsm_atb_f64_32x32_t8x8_r128_p1_nt, 1},
  {"tsm_atb_f64_32x32_t8x8_r128_p2", 0, 32, 32, 8, 8, 128, 2, 0, 131072, launch_tsm_atb_f64_32x32_t8x8_r128_p2, 1},
  {"tsm_atb_f64_32x32_t8x8_r128_p2_nt", 0, 32, 32, 8, 8, 128, 2, 1, 131072, launch_tsm_atb_f64_32x32_t8x8_r128_p2_nt, 1},
  {"tsm_atb_f64_32x64_direct_d8_nt", 0, 32, 64, 0, 0, 4, 8, 1, 32768, launch_tsm_atb_f64_32x64_direct_d8_nt, 3},
  {"tsm_atb_f64_32x64_mfma_r32_p2_nt", 0, 32, 64, 0, 0, 32, 2, 1, 49152, launch_tsm_atb_f64_32x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_64x1_t8x1_r32_p2_nt", 0, 64, 1, 8, 1, 32, 2, 1, 33280, launch_tsm_atb_f64_64x1_t8x1_r32_p2_nt, 1},
  {"tsm_atb_f64_64x2_mfma_r16_p1_nt", 0, 64, 2, 0, 0, 16, 1, 1, 20480, launch_tsm_atb_f64_64x2_mfma_r16_p1_nt, 2},
  {"tsm_atb_f64_64x2_mfma_r16_p2_nt", 0, 64, 2, 0, 0, 16, 2, 1, 20480, launch_tsm_atb_f64_64x2_mfma_r16_p2_nt, 2},
  {"tsm_atb_f64_64x2_mfma_r32_p1_nt", 0, 64, 2, 0, 0, 32, 1, 1, 40960, launch_tsm_atb_f64_64x2_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_64x2_mfma_r32_p2_nt", 0, 64, 2, 0, 0, 32, 2, 1, 40960, launch_tsm_atb_f64_64x2_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_64x2_mfma_r64_p1_nt", 0, 64, 2, 0, 0, 64, 1, 1, 81920, launch_tsm_atb_f64_64x2_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_64x2_mfma_r64_p2_nt", 0, 64, 2, 0, 0, 64, 2, 1, 81920, launch_tsm_atb_f64_64x2_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_64x2_t8x2_r32_p2_nt", 0, 64, 2, 8, 2, 32, 2, 1, 33792, launch_tsm_atb_f64_64x2_t8x2_r32_p2_nt, 1},
  {"tsm_atb_f64_64x4_mfma_r16_p1_nt", 0, 64, 4, 0, 0, 16, 1, 1, 20480, launch_tsm_atb_f64_64x4_mfma_r16_p1_nt, 2},
  {"tsm_atb_f64_64x4_mfma_r16_p2_nt", 0, 64, 4, 0, 0, 16, 2, 1, 20480, launch_tsm_atb_f64_64x4_mfma_r16_p2_nt, 2},
  {"tsm_atb_f64_64x4_mfma_r32_p1_nt", 0, 64, 4, 0, 0, 32, 1, 1, 40960, launch_tsm_atb_f64_64x4_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_64x4_mfma_r32_p2_nt", 0, 64, 4, 0, 0, 32, 2, 1, 40960, launch_tsm_atb_f64_64x4_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_64x4_mfma_r64_p1_nt", 0, 64, 4, 0, 0, 64, 1, 1, 81920, launch_tsm_atb_f64_64x4_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_64x4_mfma_r64_p2_nt", 0, 64, 4, 0, 0, 64, 2, 1, 81920, launch_tsm_atb_f64_64x4_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_64x4_t8x4_r32_p2_nt", 0, 64, 4, 8, 4, 32, 2, 1, 34816, launch_tsm_atb_f64_64x4_t8x4_r32_p2_nt, 1},
  {"tsm_atb_f64_64x8_mfma_r16_p1_nt", 0, 64, 8, 0, 0, 16, 1, 1, 20480, launch_tsm_atb_f64_64x8_mfma_r16_p1_nt, 2},
  {"tsm_atb_f64_64x8_mfma_r16_p2_nt", 0, 64, 8, 0, 0, 16, 2, 1, 20480, launch_tsm_atb_f64_64x8_mfma_r16_p2_nt, 2},
  {"tsm_atb_f64_64x8_mfma_r32_p1_nt", 0, 64, 8, 0, 0, 32, 1, 1, 40960, launch_tsm_atb_f64_64x8_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_64x8_mfma_r32_p2_nt", 0, 64, 8, 0, 0, 32, 2, 1, 40960, launch_tsm_atb_f64_64x8_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_64x8_mfma_r64_p1_nt", 0, 64, 8, 0, 0, 64, 1, 1, 81920, launch_tsm_atb_f64_64x8_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_64x8_mfma_r64_p2_nt", 0, 64, 8, 0, 0, 64, 2, 1, 81920, launch_tsm_atb_f64_64x8_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_64x8_t8x8_r32_p2_nt", 0, 64, 8, 8, 8, 32, 2, 1, 36864, launch_tsm_atb_f64_64x8_t8x8_r32_p2_nt, 1},
  {"tsm_atb_f64_64x16_direct_d8_nt", 0, 64, 16, 0, 0, 4, 8, 1, 16384, launch_tsm_atb_f64_64x16_direct_d8_nt, 3},
  {"tsm_atb_f64_64x16_mfma_r32_p2_nt", 0, 64, 16, 0, 0, 32, 2, 1, 40960, launch_tsm_atb_f64_64x16_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_64x32_direct_d8_nt", 0, 64, 32, 0, 0, 4, 8, 1, 32768, launch_tsm_atb_f64_64x32_direct_d8_nt, 3},
  {"tsm_atb_f64_64x32_mfma_r32_p2_nt", 0, 64, 32, 0, 0, 32, 2, 1, 49152, launch_tsm_atb_f64_64x32_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_64x64_direct_d4_nt", 0, 64, 64, 0, 0, 4, 4, 1, 65536, launch_tsm_atb_f64_64x64_direct_d4_nt, 3},
  {"tsm_atb_f64_64x64_direct_d8_nt", 0, 64, 64, 0, 0, 4, 8, 1, 65536, launch_tsm_atb_f64_64x64_direct_d8_nt, 3},
  {"tsm_atb_f64_64x64_direct_d16_nt", 0, 64, 64, 0, 0, 4, 16, 1, 65536, launch_tsm_atb_f64_64x64_direct_d16_nt, 3},
  {"tsm_atb_f64_64x64_mfma_r16_p1", 0, 64, 64, 0, 0, 16, 1, 0, 32768, launch_tsm_atb_f64_64x64_mfma_r16_p1, 2},
  {"tsm_atb_f64_64x64_mfma_r16_p1_nt", 0, 64, 64, 0, 0, 16, 1, 1, 32768, launch_tsm_atb_f64_64x64_mfma_r16_p1_nt, 2},
  {"tsm_atb_f64_64x64_mfma_r16_p2", 0, 64, 64, 0, 0, 16, 2, 0, 32768, launch_tsm_atb_f64_64x64_mfma_r16_p2, 2},
  {"tsm_atb_f64_64x64_mfma_r16_p2_nt", 0, 64, 64, 0, 0, 16, 2, 1, 32768, launch_tsm_atb_f64_64x64_mfma_r16_p2_nt, 2},
  {"tsm_atb_f64_64x64_mfma_r32_p1", 0, 64, 64, 0, 0, 32, 1, 0, 65536, launch_tsm_atb_f64_64x64_mfma_r32_p1, 2},
  {"tsm_atb_f64_64x64_mfma_r32_p1_nt", 0, 64, 64, 0, 0, 32, 1, 1, 65536, launch_tsm_atb_f64_64x64_mfma_r32_p1_nt, 2},
  {"tsm_atb_f64_64x64_mfma_r32_p2", 0, 64, 64, 0, 0, 32, 2, 0, 65536, launch_tsm_atb_f64_64x64_mfma_r32_p2, 2},
  {"tsm_atb_f64_64x64_mfma_r32_p2_nt", 0, 64, 64, 0, 0, 32, 2, 1, 65536, launch_tsm_atb_f64_64x64_mfma_r32_p2_nt, 2},
  {"tsm_atb_f64_64x64_mfma_r64_p1", 0, 64, 64, 0, 0, 64, 1, 0, 131072, launch_tsm_atb_f64_64x64_mfma_r64_p1, 2},
  {"tsm_atb_f64_64x64_mfma_r64_p1_nt", 0, 64, 64, 0, 0, 64, 1, 1, 131072, launch_tsm_atb_f64_64x64_mfma_r64_p1_nt, 2},
  {"tsm_atb_f64_64x64_mfma_r64_p2", 0, 64, 64, 0, 0, 64, 2, 0, 131072, launch_tsm_atb_f64_64x64_mfma_r64_p2, 2},
  {"tsm_atb_f64_64x64_mfma_r64_p2_nt", 0, 64, 64, 0, 0, 64, 2, 1, 131072, launch_tsm_atb_f64_64x64_mfma_r64_p2_nt, 2},
  {"tsm_atb_f64_64x64_t8x8_r16_p1", 0, 64, 64, 8, 8, 16, 1, 0, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p1, 1},
  {"tsm_atb_f64_64x64_t8x8_r16_p1_nt", 0, 64, 64, 8, 8, 16, 1, 1, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p1_nt, 1},
  {"tsm_atb_f64_64x64_t8x8_r16_p2", 0, 64, 64, 8, 8, 16, 2, 0, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p2, 1},
  {"tsm_atb_f64_64x64_t8x8_r16_p2_nt", 0, 64, 64, 8, 8, 16, 2, 1, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p2_nt, 1},
  {"tsm_atb_f64_64x64_t8x8_r32_p1", 0, 64, 64, 8, 8, 32, 1, 0, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p1, 1},
  {"tsm_atb_f64_64x64_t8x8_r32_p1_nt", 0, 64, 64, 8, 8, 32, 1, 1, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p1_nt, 1},
  {"tsm_atb_f64_64x64_t8x8_r32_p2", 0, 64, 64, 8, 8, 32, 2, 0, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p2, 1},
  {"tsm_atb_f64_64x64_t8x8_r32_p2_nt", 0, 64, 64, 8, 8, 32, 2, 1, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p2_nt, 1},
  {"tsm_atb_f64_64x64_t8x8_r64_p1", 0, 64, 64, 8, 8, 64, 1, 0, 131072, launch_tsm_atb_f64_64x64_t8x8_r64_p1, 1},
  {"tsm_atb_f64_64x64_t8x8_r64_p1_nt", 0, 64, 64, 8, 8, 64, 1, 1, 131072, launch_tsm_atb_f64_64x64_t8x8_r64_p1_nt, 1},
  {"tsm_atb_f64_64x64_t8x8_r64_p2", 0, 64, 64, 8, 8, 64, 2, 0, 131072, launch_tsm_atb_f64_64x64_t8x8_r64_p2, 1},
  {"tsm_atb_f64_64x64_t8x8_r64_p2_nt", 0, 64, 64, 8, 8, 64, 2, 1, 131072, launch_tsm_atb_f64_64x64_t8x8_r64_p2_nt, 1},
};

static const tsm::KernelEntry kSym[] = {
  {"tsm_ata_bf16_1x1_t1x1_r4096_p2_nt", 2, 1, 1, 1, 1, 4096, 2, 1, 16384, launch_tsm_ata_bf16_1x1_t1x1_r4096_p2_nt, 1},
  {"tsm_ata_bf16_2x2_t2x2_r4096_p2_nt", 2, 2, 2, 2, 2, 4096, 2, 1, 32768, launch_tsm_ata_bf16_2x2_t2x2_r4096_p2_nt, 1},
  {"tsm_ata_bf16_4x4_t4x4_r4096_p2_nt", 2, 4, 4, 4, 4, 4096, 2, 1, 65536, launch_tsm_ata_bf16_4x4_t4x4_r4096_p2_nt, 1},
  {"tsm_ata_bf16_8x8_mfma_r512_p1_nt", 2, 8, 8, 0, 0, 512, 1, 1, 32768, launch_tsm_ata_bf16_8x8_mfma_r512_p1_nt, 2},
  {"tsm_ata_bf16_8x8_mfma_r512_p2_nt", 2, 8, 8, 0, 0, 512, 2, 1, 32768, launch_tsm_ata_bf16_8x8_mfma_r512_p2_nt, 2},
  {"tsm_ata_bf16_8x8_mfma_r1024_p1_nt", 2, 8, 8, 0, 0, 1024, 1, 1, 65536, launch_tsm_ata_bf16_8x8_mfma_r1024_p1_nt, 2},
  {"tsm_ata_bf16_8x8_mfma_r1024_p2_nt", 2, 8, 8, 0, 0, 1024, 2, 1, 65536, launch_tsm_ata_bf16_8x8_mfma_r1024_p2_nt, 2},
  {"tsm_ata_bf16_8x8_mfma_r2048_p1_nt", 2, 8, 8, 0, 0, 2048, 1, 1, 131072, launch_tsm_ata_bf16_8x8_mfma_r2048_p1_nt, 2},
  {"tsm_ata_bf16_8x8_mfma_r2048_p2_nt", 2, 8, 8, 0, 0, 2048, 2, 1, 131072, launch_tsm_ata_bf16_8x8_mfma_r2048_p2_nt, 2},
  {"tsm_ata_bf16_8x8_t8x4_r512_p1_nt", 2, 8, 8, 8, 4, 512, 1, 1, 16384, launch_tsm_ata_bf16_8x8_t8x4_r512_p1_nt, 1},
  {"tsm_ata_bf16_8x8_t8x4_r512_p2_nt", 2, 8, 8, 8, 4, 512, 2, 1, 16384, launch_tsm_ata_bf16_8x8_t8x4_r512_p2_nt, 1},
  {"tsm_ata_bf16_8x8_t8x4_r1024_p1_nt", 2, 8, 8, 8, 4, 1024, 1, 1, 32768, launch_tsm_ata_bf16_8x8_t8x4_r1024_p1_nt, 1},
  {"tsm_ata_bf16_8x8_t8x4_r1024_p2_nt", 2, 8, 8, 8, 4, 1024, 2, 1, 32768, launch_tsm_ata_bf16_8x8_t8x4_r1024_p2_nt, 1},
  {"tsm_ata_bf16_8x8_t8x4_r2048_p1_nt", 2, 8, 8, 8, 4, 2048, 1, 1, 65536, launch_tsm_ata_bf16_8x8_t8x4_r2048_p1_nt, 1},
  {"tsm_ata_bf16_8x8_t8x4_r2048_p2_nt", 2, 8, 8, 8, 4, 2048, 2, 1, 65536, launch_tsm_ata_bf16_8x8_t8x4_r2048_p2_nt, 1},
  {"tsm_ata_bf16_16x16_direct_d2_nt", 2, 16, 16, 0, 0, 32, 2, 1, 8192, launch_tsm_ata_bf16_16x16_direct_d2_nt, 3},
  {"tsm_ata_bf16_16x16_direct_d4_nt", 2, 16, 16, 0, 0, 32, 4, 1, 8192, launch_tsm_ata_bf16_16x16_direct_d4_nt, 3},
  {"tsm_ata_bf16_16x16_mfma_r512_p1_nt", 2, 16, 16, 0, 0, 512, 1, 1, 32768, launch_tsm_ata_bf16_16x16_mfma_r512_p1_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r512_p2_nt", 2, 16, 16, 0, 0, 512, 2, 1, 32768, launch_tsm_ata_bf16_16x16_mfma_r512_p2_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r1024_p1_nt", 2, 16, 16, 0, 0, 1024, 1, 1, 65536, launch_tsm_ata_bf16_16x16_mfma_r1024_p1_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r1024_p2_nt", 2, 16, 16, 0, 0, 1024, 2, 1, 65536, launch_tsm_ata_bf16_16x16_mfma_r1024_p2_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r2048_p1_nt", 2, 16, 16, 0, 0, 2048, 1, 1, 131072, launch_tsm_ata_bf16_16x16_mfma_r2048_p1_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r2048_p2_nt", 2, 16, 16, 0, 0, 2048, 2, 1, 131072, launch_tsm_ata_bf16_16x16_mfma_r2048_p2_nt, 2},
  {"tsm_ata_bf16_32x32_direct_d2_nt", 2, 32, 32, 0, 0, 32, 2, 1, 16384, launch_tsm_ata_bf16_32x32_direct_d2_nt, 3},
  {"tsm_ata_bf16_32x32_direct_d4_nt", 2, 32, 32, 0, 0, 32, 4, 1, 16384, launch_tsm_ata_bf16_32x32_direct_d4_nt, 3},
  {"tsm_ata_bf16_32x32_mfma_r256_p1_nt", 2, 32, 32, 0, 0, 256, 1, 1, 32768, launch_tsm_ata_bf16_32x32_mfma_r256_p1_nt, 2},
  {"tsm_ata_bf16_32x32_mfma_r256_p2_nt", 2, 32, 32, 0, 0, 256, 2, 1, 32768, launch_tsm_ata_bf16_32x32_mfma_r256_p2_nt, 2},
  {"tsm_ata_bf16_32x32_mfma_r512_p1_nt", 2, 32, 32, 0, 0, 512, 1, 1, 65536, launch_tsm_ata_bf16_32x32_mfma_r512_p1_nt, 2},
  {"tsm_ata_bf16_32x32_mfma_r512_p2_nt", 2, 32, 32, 0, 0, 512, 2, 1, 65536, launch_tsm_ata_bf16_32x32_mfma_r512_p2_nt, 2},
  {"tsm_ata_bf16_32x32_mfma_r1024_p1_nt", 2, 32, 32, 0, 0, 1024, 1, 1, 131072, launch_tsm_ata_bf16_32x32_mfma_r1024_p1_nt, 2},
  {"tsm_ata_bf16_32x32_mfma_r1024_p2_nt", 2, 32, 32, 0, 0, 1024, 2, 1, 131072, launch_tsm_ata_bf16_32x32_mfma_r1024_p2_nt, 2},
  {"tsm_ata_bf16_64x64_direct_d2_nt", 2, 64, 64, 0, 0, 32, 2, 1, 32768, launch_tsm_ata_bf16_64x64_direct_d2_nt, 3},
  {"tsm_ata_bf16_64x64_direct_d4_nt", 2, 64, 64, 0, 0, 32, 4, 1, 32768, launch_tsm_ata_bf16_64x64_direct_d4_nt, 3},
  {"tsm_ata_bf16_64x64_mfma_r256_p2_nt", 2, 64, 64, 0, 0, 256, 2, 1, 65536, launch_tsm_ata_bf16_64x64_mfma_r256_p2_nt, 2},
  {"tsm_ata_f32_1x1_t1x1_r4096_p2_nt", 1, 1, 1, 1, 1, 4096, 2, 1, 32768, launch_tsm_ata_f32_1x1_t1x1_r4096_p2_nt, 1},
  {"tsm_ata_f32_2x2_t2x2_r4096_p2_nt", 1, 2, 2, 2, 2, 4096, 2, 1, 65536, launch_tsm_ata_f32_2x2_t2x2_r4096_p2_nt, 1},
  {"tsm_ata_f32_4x4_t4x4_r2048_p2_nt", 1, 4, 4, 4, 4, 2048, 2, 1, 65536, launch_tsm_ata_f32_4x4_t4x4_r2048_p2_nt, 1},
  {"tsm_ata_f32_8x8_mfma_r256_p1_nt", 1, 8, 8, 0, 0, 256, 1, 1, 32768, launch_tsm_ata_f32_8x8_mfma_r256_p1_nt, 2},
  {"tsm_ata_f32_8x8_mfma_r256_p2_nt", 1, 8, 8, 0, 0, 256, 2, 1, 32768, launch_tsm_ata_f32_8x8_mfma_r256_p2_nt, 2},
  {"tsm_ata_f32_8x8_t8x4_r512_p1_nt", 1, 8, 8, 8, 4, 512, 1, 1, 32768, launch_tsm_ata_f32_8x8_t8x4_r512_p1_nt, 1},
  {"tsm_ata_f32_8x8_t8x4_r512_p2_nt", 1, 8, 8, 8, 4, 512, 2, 1, 32768, launch_tsm_ata_f32_8x8_t8x4_r512_p2_nt, 1},
  {"tsm_ata_f32_8x8_t8x4_r1024_p1_nt", 1, 8, 8, 8, 4, 1024, 1, 1, 65536, launch_tsm_ata_f32_8x8_t8x4_r1024_p1_nt, 1},
  {"tsm_ata_f32_8x8_t8x4_r1024_p2_nt", 1, 8, 8, 8, 4, 1024, 2, 1, 65536, launch_tsm_ata_f32_8x8_t8x4_r1024_p2_nt, 1},
  {"tsm_ata_f32_8x8_t8x4_r2048_p1_nt", 1, 8, 8, 8, 4, 2048, 1, 1, 131072, launch_tsm_ata_f32_8x8_t8x4_r2048_p1_nt, 1},
  {"tsm_ata_f32_8x8_t8x4_r2048_p2_nt", 1, 8, 8, 8, 4, 2048, 2, 1, 131072, launch_tsm_ata_f32_8x8_t8x4_r2048_p2_nt, 1},
  {"tsm_ata_f32_16x16_direct_d4_nt", 1, 16, 16, 0, 0, 4, 4, 1, 4096, launch_tsm_ata_f32_16x16_direct_d4_nt, 3},
  {"tsm_ata_f32_16x16_direct_d8_nt", 1, 16, 16, 0, 0, 4, 8, 1, 4096, launch_tsm_ata_f32_16x16_direct_d8_nt, 3},
  {"tsm_ata_f32_16x16_direct_d16_nt", 1, 16, 16, 0, 0, 4, 16, 1, 4096, launch_tsm_ata_f32_16x16_direct_d16_nt, 3},
  {"tsm_ata_f32_16x16_mfma_r64_p1_nt", 1, 16, 16, 0, 0, 64, 1, 1, 8192, launch_tsm_ata_f32_16x16_mfma_r64_p1_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r64_p2_nt", 1, 16, 16, 0, 0, 64, 2, 1, 8192, launch_tsm_ata_f32_16x16_mfma_r64_p2_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r128_p1_nt", 1, 16, 16, 0, 0, 128, 1, 1, 16384, launch_tsm_ata_f32_16x16_mfma_r128_p1_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r128_p2_nt", 1, 16, 16, 0, 0, 128, 2, 1, 16384, launch_tsm_ata_f32_16x16_mfma_r128_p2_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r256_p1_nt", 1, 16, 16, 0, 0, 256, 1, 1, 32768, launch_tsm_ata_f32_16x16_mfma_r256_p1_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r256_p2_nt", 1, 16, 16, 0, 0, 256, 2, 1, 32768, launch_tsm_ata_f32_16x16_mfma_r256_p2_nt, 2},
  {"tsm_ata_f32_32x32_direct_d4_nt", 1, 32, 32, 0, 0, 4, 4, 1, 16384, launch_tsm_ata_f32_32x32_direct_d4_nt, 3},
  {"tsm_ata_f32_32x32_direct_d8_nt", 1, 32, 32, 0, 0, 4, 8, 1, 16384, launch_tsm_ata_f32_32x32_direct_d8_nt, 3},
  {"tsm_ata_f32_32x32_direct_d16_nt", 1, 32, 32, 0, 0, 4, 16, 1, 16384, launch_tsm_ata_f32_32x32_direct_d16_nt, 3},
  {"tsm_ata_f32_32x32_mfma_r64_p1_nt", 1, 32, 32, 0, 0, 64, 1, 1, 16384, launch_tsm_ata_f32_32x32_mfma_r64_p1_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r64_p2_nt", 1, 32, 32, 0, 0, 64, 2, 1, 16384, launch_tsm_ata_f32_32x32_mfma_r64_p2_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r128_p1_nt", 1, 32, 32, 0, 0, 128, 1, 1, 32768, launch_tsm_ata_f32_32x32_mfma_r128_p1_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r128_p2_nt", 1, 32, 32, 0, 0, 128, 2, 1, 32768, launch_tsm_ata_f32_32x32_mfma_r128_p2_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r256_p1_nt", 1, 32, 32, 0, 0, 256, 1, 1, 65536, launch_tsm_ata_f32_32x32_mfma_r256_p1_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r256_p2_nt", 1, 32, 32, 0, 0, 256, 2, 1, 65536, launch_tsm_ata_f32_32x32_mfma_r256_p2_nt, 2},
  {"tsm_ata_f32_64x64_direct_d4_nt", 1, 64, 64, 0, 0, 4, 4, 1, 65536, launch_tsm_ata_f32_64x64_direct_d4_nt, 3},
  {"tsm_ata_f32_64x64_direct_d8_nt", 1, 64, 64, 0, 0, 4, 8, 1, 65536, launch_tsm_ata_f32_64x64_direct_d8_nt, 3},
  {"tsm_ata_f32_64x64_direct_d16_nt", 1, 64, 64, 0, 0, 4, 16, 1, 65536, launch_tsm_ata_f32_64x64_direct_d16_nt, 3},
  {"tsm_ata_f32_64x64_mfma_r64_p2", 1, 64, 64, 0, 0, 64, 2, 0, 32768, launch_tsm_ata_f32_64x64_mfma_r64_p2, 2},
  {"tsm_ata_f64_1x1_t1x1_r2048_p1_nt", 0, 1, 1, 1, 1, 2048, 1, 1, 32768, launch_tsm_ata_f64_1x1_t1x1_r2048_p1_nt, 1},
  {"tsm_ata_f64_1x1_t1x1_r2048_p2_nt", 0, 1, 1, 1, 1, 2048, 2, 1, 32768, launch_tsm_ata_f64_1x1_t1x1_r2048_p2_nt, 1},
  {"tsm_ata_f64_1x1_t1x1_r4096_p1_nt", 0, 1, 1, 1, 1, 4096, 1, 1, 65536, launch_tsm_ata_f64_1x1_t1x1_r4096_p1_nt, 1},
  {"tsm_ata_f64_1x1_t1x1_r4096_p2_nt", 0, 1, 1, 1, 1, 4096, 2, 1, 65536, launch_tsm_ata_f64_1x1_t1x1_r4096_p2_nt, 1},
  {"tsm_ata_f64_2x2_t2x2_r1024_p1_nt", 0, 2, 2, 2, 2, 1024, 1, 1, 32768, launch_tsm_ata_f64_2x2_t2x2_r1024_p1_nt, 1},
  {"tsm_ata_f64_2x2_t2x2_r1024_p2_nt", 0, 2, 2, 2, 2, 1024, 2, 1, 32768, launch_tsm_ata_f64_2x2_t2x2_r1024_p2_nt, 1},
  {"tsm_ata_f64_2x2_t2x2_r2048_p1_nt", 0, 2, 2, 2, 2, 2048, 1, 1, 65536, launch_tsm_ata_f64_2x2_t2x2_r2048_p1_nt, 1},
  {"tsm_ata_f64_2x2_t2x2_r2048_p2_nt", 0, 2, 2, 2, 2, 2048, 2, 1, 65536, launch_tsm_ata_f64_2x2_t2x2_r2048_p2_nt, 1},
  {"tsm_ata_f64_2x2_t2x2_r4096_p1_nt", 0, 2, 2, 2, 2, 4096, 1, 1, 131072, launch_tsm_ata_f64_2x2_t2x2_r4096_p1_nt, 1},
  {"tsm_ata_f64_2x2_t2x2_r4096_p2_nt", 0, 2, 2, 2, 2, 4096, 2, 1, 131072, launch_tsm_ata_f64_2x2_t2x2_r4096_p2_nt, 1},
  {"tsm_ata_f64_4x4_t4x4_r256_p1_nt", 0, 4, 4, 4, 4, 256, 1, 1, 16384, launch_tsm_ata_f64_4x4_t4x4_r256_p1_nt, 1},
  {"tsm_ata_f64_4x4_t4x4_r256_p2_nt", 0, 4, 4, 4, 4, 256, 2, 1, 16384, launch_tsm_ata_f64_4x4_t4x4_r256_p2_nt, 1},
  {"tsm_ata_f64_4x4_t4x4_r512_p1_nt", 0, 4, 4, 4, 4, 512, 1, 1, 32768, launch_tsm_ata_f64_4x4_t4x4_r512_p1_nt, 1},
  {"tsm_ata_f64_4x4_t4x4_r512_p2_nt", 0, 4, 4, 4, 4, 512, 2, 1, 32768, launch_tsm_ata_f64_4x4_t4x4_r512_p2_nt, 1},
  {"tsm_ata_f64_4x4_t4x4_r1024_p1_nt", 0, 4, 4, 4, 4, 1024, 1, 1, 65536, launch_tsm_ata_f64_4x4_t4x4_r1024_p1_nt, 1},
  {"tsm_ata_f64_4x4_t4x4_r1024_p2_nt", 0, 4, 4, 4, 4, 1024, 2, 1, 65536, launch_tsm_ata_f64_4x4_t4x4_r1024_p2_nt, 1},
  {"tsm_ata_f64_8x8_mfma_r128_p1_nt", 0, 8, 8, 0, 0, 128, 1, 1, 32768, launch_tsm_ata_f64_8x8_mfma_r128_p1_nt, 2},
  {"tsm_ata_f64_8x8_mfma_r128_p2_nt", 0, 8, 8, 0, 0, 128, 2, 1, 32768, launch_tsm_ata_f64_8x8_mfma_r128_p2_nt, 2},
  {"tsm_ata_f64_8x8_mfma_r256_p1_nt", 0, 8, 8, 0, 0, 256, 1, 1, 65536, launch_tsm_ata_f64_8x8_mfma_r256_p1_nt, 2},
  {"tsm_ata_f64_8x8_mfma_r256_p2_nt", 0, 8, 8, 0, 0, 256, 2, 1, 65536, launch_tsm_ata_f64_8x8_mfma_r256_p2_nt, 2},
  {"tsm_ata_f64_8x8_t4x4_r64_p1_nt", 0, 8, 8, 4, 4, 64, 1, 1, 8192, launch_tsm_ata_f64_8x8_t4x4_r64_p1_nt, 1},
  {"tsm_ata_f64_8x8_t4x4_r64_p2_nt", 0, 8, 8, 4, 4, 64, 2, 1, 8192, launch_tsm_ata_f64_8x8_t4x4_r64_p2_nt, 1},
  {"tsm_ata_f64_8x8_t4x4_r128_p1_nt", 0, 8, 8, 4, 4, 128, 1, 1, 16384, launch_tsm_ata_f64_8x8_t4x4_r128_p1_nt, 1},
  {"tsm_ata_f64_8x8_t4x4_r128_p2_nt", 0, 8, 8, 4, 4, 128, 2, 1, 16384, launch_tsm_ata_f64_8x8_t4x4_r128_p2_nt, 1},
  {"tsm_ata_f64_8x8_t4x4_r256_p1_nt", 0, 8, 8, 4, 4, 256, 1, 1, 32768, launch_tsm_ata_f64_8x8_t4x4_r256_p1_nt, 1},
  {"tsm_ata_f64_8x8_t4x4_r256_p2_nt", 0, 8, 8, 4, 4, 256, 2, 1, 32768, launch_tsm_ata_f64_8x8_t4x4_r256_p2_nt, 1},
  {"tsm_ata_f64_8x8_t4x4_r512_p1_nt", 0, 8, 8, 4, 4, 512, 1, 1, 65536, launch_tsm_ata_f64_8x8_t4x4_r512_p1_nt, 1},
  {"tsm_ata_f64_8x8_t4x4_r512_p2_nt", 0, 8, 8, 4, 4, 512, 2, 1, 65536, launch_tsm_ata_f64_8x8_t4x4_r512_p2_nt, 1},
  {"tsm_ata_f64_8x8_t8x8_r512_p2_nt", 0, 8, 8, 8, 8, 512, 2, 1, 65536, launch_tsm_ata_f64_8x8_t8x8_r512_p2_nt, 1},
  {"tsm_ata_f64_16x16_direct_d4_nt", 0, 16, 16, 0, 0, 4, 4, 1, 4096, launch_tsm_ata_f64_16x16_direct_d4_nt, 3},
  {"tsm_ata_f64_16x16_direct_d8_nt", 0, 16, 16, 0, 0, 4, 8, 1, 4096, launch_tsm_ata_f64_16x16_direct_d8_nt, 3},
  {"tsm_ata_f64_16x16_direct_d16_nt", 0, 16, 16, 0, 0, 4, 16, 1, 4096, launch_tsm_ata_f64_16x16_direct_d16_nt, 3},
  {"tsm_ata_f64_16x16_mfma_r32_p1_nt", 0, 16, 16, 0, 0, 32, 1, 1, 8192, launch_tsm_ata_f64_16x16_mfma_r32_p1_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r32_p2_nt", 0, 16, 16, 0, 0, 32, 2, 1, 8192, launch_tsm_ata_f64_16x16_mfma_r32_p2_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r64_p1_nt", 0, 16, 16, 0, 0, 64, 1, 1, 16384, launch_tsm_ata_f64_16x16_mfma_r64_p1_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r64_p2_nt", 0, 16, 16, 0, 0, 64, 2, 1, 16384, launch_tsm_ata_f64_16x16_mfma_r64_p2_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r128_p1_nt", 0, 16, 16, 0, 0, 128, 1, 1, 32768, launch_tsm_ata_f64_16x16_mfma_r128_p1_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r128_p2_nt", 0, 16, 16, 0, 0, 128, 2, 1, 32768, launch_tsm_ata_f64_16x16_mfma_r128_p2_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r256_p1_nt", 0, 16, 16, 0, 0, 256, 1, 1, 65536, launch_tsm_ata_f64_16x16_mfma_r256_p1_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r256_p2_nt", 0, 16, 16, 0, 0, 256, 2, 1, 65536, launch_tsm_ata_f64_16x16_mfma_r256_p2_nt, 2},
  {"tsm_ata_f64_32x32_direct_d4_nt", 0, 32, 32, 0, 0, 4, 4, 1, 16384, launch_tsm_ata_f64_32x32_direct_d4_nt, 3},
  {"tsm_ata_f64_32x32_direct_d8_nt", 0, 32, 32, 0, 0, 4, 8, 1, 16384, launch_tsm_ata_f64_32x32_direct_d8_nt, 3},
  {"tsm_ata_f64_32x32_direct_d16_nt", 0, 32, 32, 0, 0, 4, 16, 1, 16384, launch_tsm_ata_f64_32x32_direct_d16_nt, 3},
  {"tsm_ata_f64_32x32_mfma_r16_p1_nt", 0, 32, 32, 0, 0, 16, 1, 1, 8192, launch_tsm_ata_f64_32x32_mfma_r16_p1_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r16_p2_nt", 0, 32, 32, 0, 0, 16, 2, 1, 8192, launch_tsm_ata_f64_32x32_mfma_r16_p2_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r32_p1_nt", 0, 32, 32, 0, 0, 32, 1, 1, 16384, launch_tsm_ata_f64_32x32_mfma_r32_p1_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r32_p2_nt", 0, 32, 32, 0, 0, 32, 2, 1, 16384, launch_tsm_ata_f64_32x32_mfma_r32_p2_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r64_p1_nt", 0, 32, 32, 0, 0, 64, 1, 1, 32768, launch_tsm_ata_f64_32x32_mfma_r64_p1_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r64_p2_nt", 0, 32, 32, 0, 0, 64, 2, 1, 32768, launch_tsm_ata_f64_32x32_mfma_r64_p2_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r128_p1_nt", 0, 32, 32, 0, 0, 128, 1, 1, 65536, launch_tsm_ata_f64_32x32_mfma_r128_p1_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r128_p2_nt", 0, 32, 32, 0, 0, 128, 2, 1, 65536, launch_tsm_ata_f64_32x32_mfma_r128_p2_nt, 2},
  {"tsm_ata_f64_64x64_direct_d4_nt", 0, 64, 64, 0, 0, 4, 4, 1, 65536, launch_tsm_ata_f64_64x64_direct_d4_nt, 3},
  {"tsm_ata_f64_64x64_direct_d8_nt", 0, 64, 64, 0, 0, 4, 8, 1, 65536, launch_tsm_ata_f64_64x64_direct_d8_nt, 3},
  {"tsm_ata_f64_64x64_direct_d16_nt", 0, 64, 64, 0, 0, 4, 16, 1, 65536, launch_tsm_ata_f64_64x64_direct_d16_nt, 3},
  {"tsm_ata_f64_64x64_mfma_r32_p2_nt", 0, 64, 64, 0, 0, 32, 2, 1, 32768, launch_tsm_ata_f64_64x64_mfma_r32_p2_nt, 2},
};

static const tsm::TunedEntry kSymTuned[] = {
  {2, 1, 1, 0, 1, 1, 0, 1},  // tsm_ata_bf16_1x1_t1x1_r4096_p2_nt
  {2, 2, 2, 1, 1, 0, 1, 0},  // tsm_ata_bf16_2x2_t2x2_r4096_p2_nt
  {2, 4, 4, 2, 1, 1, 2, 1},  // tsm_ata_bf16_4x4_t4x4_r4096_p2_nt
  {2, 8, 8, 6, 1, 1, 6, 1},  // tsm_ata_bf16_8x8_mfma_r1024_p2_nt
  {2, 16, 16, 18, 1, 1, 18, 1},  // tsm_ata_bf16_16x16_mfma_r512_p2_nt
  {2, 32, 32, 26, 1, 1, 26, 1},  // tsm_ata_bf16_32x32_mfma_r256_p2_nt
  {2, 64, 64, 32, 1, 0, 33, 1},  // tsm_ata_bf16_64x64_direct_d4_nt (else tsm_ata_bf16_64x64_mfma_r256_p2_nt)
  {1, 1, 1, 34, 1, 1, 34, 1},  // tsm_ata_f32_1x1_t1x1_r4096_p2_nt
  {1, 2, 2, 35, 1, 1, 35, 1},  // tsm_ata_f32_2x2_t2x2_r4096_p2_nt
  {1, 4, 4, 36, 1, 1, 36, 1},  // tsm_ata_f32_4x4_t4x4_r2048_p2_nt
  {1, 8, 8, 41, 2, 0, 41, 0},  // tsm_ata_f32_8x8_t8x4_r1024_p1_nt
  {1, 16, 16, 52, 2, 0, 52, 0},  // tsm_ata_f32_16x16_mfma_r256_p1_nt
  {1, 32, 32, 56, 3, 0, 60, 1},  // tsm_ata_f32_32x32_direct_d16_nt (else tsm_ata_f32_32x32_mfma_r128_p2_nt)
  {1, 64, 64, 65, 2, 0, 66, 1},  // tsm_ata_f32_64x64_direct_d16_nt (else tsm_ata_f32_64x64_mfma_r64_p2)
  {0, 1, 1, 67, 2, 1, 67, 1},  // tsm_ata_f64_1x1_t1x1_r2048_p1_nt
  {0, 2, 2, 72, 1, 1, 72, 1},  // tsm_ata_f64_2x2_t2x2_r1024_p2_nt
  {0, 4, 4, 80, 1, 1, 80, 1},  // tsm_ata_f64_4x4_t4x4_r512_p2_nt
  {0, 8, 8, 95, 1, 0, 95, 0},  // tsm_ata_f64_8x8_t8x8_r512_p2_nt
  {0, 16, 16, 104, 1, 1, 104, 1},  // tsm_ata_f64_16x16_mfma_r128_p2_nt
  {0, 32, 32, 108, 1, 0, 112, 1},  // tsm_ata_f64_32x32_direct_d8_nt (else tsm_ata_f64_32x32_mfma_r32_p1_nt)
  {0, 64, 64, 120, 2, 0, 121, 0},  // tsm_ata_f64_64x64_direct_d16_nt (else tsm_ata_f64_64x64_mfma_r32_p2_nt)
};

static const tsm::TunedEntry kTuned[] = {
  {2, 1, 1, 0, 1, 1, 0, 1},  // tsm_atb_bf16_1x1_t1x1_r4096_p2_nt
  {2, 1, 2, 1, 1, 1, 1, 1},  // tsm_atb_bf16_1x2_t1x2_r4096_p2_nt
  {2, 1, 4, 2, 1, 1, 2, 1},  // tsm_atb_bf16_1x4_t1x4_r2048_p2_nt
  {2, 1, 8, 3, 2, 1, 3, 1},  // tsm_atb_bf16_1x8_t1x4_r1024_p2_nt
  {2, 1, 16, 4, 1, 1, 4, 1},  // tsm_atb_bf16_1x16_t1x4_r512_p2_nt
  {2, 1, 32, 5, 2, 1, 5, 1},  // tsm_atb_bf16_1x32_t1x4_r256_p2_nt
  {2, 1, 64, 6, 2, 1, 6, 1},  // tsm_atb_bf16_1x64_t1x4_r128_p2_nt
  {2, 2, 1, 7, 1, 0, 7, 0},  // tsm_atb_bf16_2x1_t2x1_r4096_p2_nt
  {2, 2, 2, 8, 1, 0, 8, 0},  // tsm_atb_bf16_2x2_t2x2_r4096_p2_nt
  {2, 2, 4, 9, 1, 1, 9, 1},  // tsm_atb_bf16_2x4_t2x4_r2048_p2_nt
  {2, 2, 8, 10, 1, 1, 10, 1},  // tsm_atb_bf16_2x8_t2x4_r1024_p2_nt
  {2, 2, 16, 11, 2, 1, 11, 1},  // tsm_atb_bf16_2x16_t2x4_r512_p2_nt
  {2, 2, 32, 12, 2, 1, 12, 1},  // tsm_atb_bf16_2x32_t2x4_r256_p2_nt
  {2, 2, 64, 13, 2, 1, 13, 1},  // tsm_atb_bf16_2x64_t2x4_r128_p2_nt
  {2, 4, 1, 14, 1, 1, 14, 1},  // tsm_atb_bf16_4x1_t4x1_r2048_p2_nt
  {2, 4, 2, 15, 1, 1, 15, 1},  // tsm_atb_bf16_4x2_t4x2_r2048_p2_nt
  {2, 4, 4, 16, 1, 1, 16, 1},  // tsm_atb_bf16_4x4_t4x4_r2048_p2_nt
  {2, 4, 8, 17, 1, 1, 17, 1},  // tsm_atb_bf16_4x8_t4x4_r1024_p2_nt
  {2, 4, 16, 19, 2, 1, 19, 1},  // tsm_atb_bf16_4x16_mfma_r256_p2_nt
  {2, 4, 32, 28, 1, 1, 28, 1},  // tsm_atb_bf16_4x32_mfma_r256_p2_nt
  {2, 4, 64, 32, 2, 1, 32, 1},  // tsm_atb_bf16_4x64_mfma_r128_p1_nt
  {2, 8, 1, 37, 2, 1, 37, 1},  // tsm_atb_bf16_8x1_t8x1_r1024_p2_nt
  {2, 8, 2, 38, 1, 1, 38, 1},  // tsm_atb_bf16_8x2_t8x2_r1024_p2_nt
  {2, 8, 4, 39, 1, 1, 39, 1},  // tsm_atb_bf16_8x4_t8x4_r1024_p2_nt
  {2, 8, 8, 65, 2, 1, 65, 1},  // tsm_atb_bf16_8x8_t8x4_r1024_p1_nt
  {2, 8, 16, 83, 1, 1, 83, 1},  // tsm_atb_bf16_8x16_mfma_r256_p2_nt
  {2, 8, 32, 90, 2, 1, 90, 1},  // tsm_atb_bf16_8x32_mfma_r128_p2_nt
  {2, 8, 64, 96, 2, 1, 96, 1},  // tsm_atb_bf16_8x64_mfma_r128_p1_nt
  {2, 16, 1, 101, 2, 1, 101, 1},  // tsm_atb_bf16_16x1_t8x1_r512_p2_nt
  {2, 16, 2, 102, 2, 1, 102, 1},  // tsm_atb_bf16_16x2_t8x2_r512_p2_nt
  {2, 16, 4, 109, 1, 0, 109, 0},  // tsm_atb_bf16_16x4_t8x4_r512_p2_nt
  {2, 16, 8, 111, 1, 1, 111, 1},  // tsm_atb_bf16_16x8_mfma_r256_p2_nt
  {2, 16, 16, 128, 1, 1, 128, 1},  // tsm_atb_bf16_16x16_mfma_r1024_p1_nt
  {2, 16, 32, 174, 1, 1, 174, 1},  // tsm_atb_bf16_16x32_mfma_r256_p2_nt
  {2, 16, 64, 176, 1, 1, 176, 1},  // tsm_atb_bf16_16x64_mfma_r128_p2_nt
  {2, 32, 1, 177, 2, 1, 177, 1},  // tsm_atb_bf16_32x1_t8x1_r256_p2_nt
  {2, 32, 2, 178, 2, 1, 178, 1},  // tsm_atb_bf16_32x2_t8x2_r256_p2_nt
  {2, 32, 4, 185, 2, 1, 185, 1},  // tsm_atb_bf16_32x4_t8x4_r256_p2_nt
  {2, 32, 8, 189, 1, 1, 189, 1},  // tsm_atb_bf16_32x8_mfma_r256_p2_nt
  {2, 32, 16, 194, 1, 1, 194, 1},  // tsm_atb_bf16_32x16_mfma_r256_p2_nt
  {2, 32, 32, 206, 1, 1, 206, 1},  // tsm_atb_bf16_32x32_mfma_r512_p1_nt
  {2, 32, 64, 252, 1, 1, 252, 1},  // tsm_atb_bf16_32x64_mfma_r128_p2_nt
  {2, 64, 1, 253, 2, 1, 253, 1},  // tsm_atb_bf16_64x1_t8x1_r128_p2_nt
  {2, 64, 2, 254, 2, 1, 254, 1},  // tsm_atb_bf16_64x2_t8x2_r128_p2_nt
  {2, 64, 4, 255, 2, 1, 255, 1},  // tsm_atb_bf16_64x4_mfma_r128_p1_nt
  {2, 64, 8, 260, 2, 1, 260, 1},  // tsm_atb_bf16_64x8_mfma_r128_p1_nt
  {2, 64, 16, 266, 1, 1, 266, 1},  // tsm_atb_bf16_64x16_mfma_r128_p2_nt
  {2, 64, 32, 268, 1, 1, 268, 1},  // tsm_atb_bf16_64x32_mfma_r128_p2_nt
  {2, 64, 64, 270, 1, 0, 278, 1},  // tsm_atb_bf16_64x64_direct_d4_nt (else tsm_atb_bf16_64x64_mfma_r256_p2_nt)
  {1, 1, 1, 291, 1, 1, 291, 1},  // tsm_atb_f32_1x1_t1x1_r4096_p2_nt
  {1, 1, 2, 292, 1, 1, 292, 1},  // tsm_atb_f32_1x2_t1x2_r2048_p2_nt
  {1, 1, 4, 293, 1, 1, 293, 1},  // tsm_atb_f32_1x4_t1x4_r1024_p2_nt
  {1, 1, 8, 294, 1, 1, 294, 1},  // tsm_atb_f32_1x8_t1x4_r512_p2_nt
  {1, 1, 16, 295, 1, 1, 295, 1},  // tsm_atb_f32_1x16_t1x4_r256_p2_nt
  {1, 1, 32, 296, 1, 1, 296, 1},  // tsm_atb_f32_1x32_t1x4_r128_p2_nt
  {1, 1, 64, 297, 1, 1, 297, 1},  // tsm_atb_f32_1x64_t1x4_r64_p2_nt
  {1, 2, 1, 298, 1, 1, 298, 1},  // tsm_atb_f32_2x1_t2x1_r2048_p2_nt
  {1, 2, 2, 299, 1, 1, 299, 1},  // tsm_atb_f32_2x2_t2x2_r2048_p2_nt
  {1, 2, 4, 300, 1, 1, 300, 1},  // tsm_atb_f32_2x4_t2x4_r1024_p2_nt
  {1, 2, 8, 301, 1, 1, 301, 1},  // tsm_atb_f32_2x8_t2x4_r512_p2_nt
  {1, 2, 16, 302, 1, 1, 302, 1},  // tsm_atb_f32_2x16_t2x4_r256_p2_nt
  {1, 2, 32, 303, 1, 1, 303, 1},  // tsm_atb_f32_2x32_t2x4_r128_p2_nt
  {1, 2, 64, 304, 1, 1, 304, 1},  // tsm_atb_f32_2x64_t2x4_r64_p2_nt
  {1, 4, 1, 305, 1, 1, 305, 1},  // tsm_atb_f32_4x1_t4x1_r1024_p2_nt
  {1, 4, 2, 306, 1, 1, 306, 1},  // tsm_atb_f32_4x2_t4x2_r1024_p2_nt
  {1, 4, 4, 307, 1, 1, 307, 1},  // tsm_atb_f32_4x4_t4x4_r1024_p2_nt
  {1, 4, 8, 308, 1, 1, 308, 1},  // tsm_atb_f32_4x8_t4x4_r512_p2_nt
  {1, 4, 16, 313, 1, 1, 313, 1},  // tsm_atb_f32_4x16_t4x4_r256_p2_nt
  {1, 4, 32, 320, 1, 1, 320, 1},  // tsm_atb_f32_4x32_t4x4_r128_p2_nt
  {1, 4, 64, 327, 1, 1, 327, 1},  // tsm_atb_f32_4x64_t4x4_r64_p2_nt
  {1, 8, 1, 328, 2, 1, 328, 1},  // tsm_atb_f32_8x1_t8x1_r512_p2_nt
  {1, 8, 2, 329, 1, 1, 329, 1},  // tsm_atb_f32_8x2_t8x2_r512_p2_nt
  {1, 8, 4, 330, 1, 1, 330, 1},  // tsm_atb_f32_8x4_t8x4_r512_p2_nt
  {1, 8, 8, 366, 1, 1, 366, 1},  // tsm_atb_f32_8x8_t8x4_r1024_p2_nt
  {1, 8, 16, 380, 1, 1, 380, 1},  // tsm_atb_f32_8x16_t8x4_r256_p2_nt
  {1, 8, 32, 382, 2, 1, 382, 1},  // tsm_atb_f32_8x32_mfma_r64_p2_nt
  {1, 8, 64, 394, 2, 1, 394, 1},  // tsm_atb_f32_8x64_t8x4_r64_p2_nt
  {1, 16, 1, 395, 2, 1, 395, 1},  // tsm_atb_f32_16x1_t8x1_r256_p2_nt
  {1, 16, 2, 396, 2, 1, 396, 1},  // tsm_atb_f32_16x2_t8x2_r256_p2_nt
  {1, 16, 4, 401, 1, 0, 401, 0},  // tsm_atb_f32_16x4_t8x4_r256_p2_nt
  {1, 16, 8, 406, 1, 1, 406, 1},  // tsm_atb_f32_16x8_t8x4_r256_p2_nt
  {1, 16, 16, 413, 1, 1, 413, 1},  // tsm_atb_f32_16x16_mfma_r128_p2_nt
  {1, 16, 32, 460, 1, 0, 461, 0},  // tsm_atb_f32_16x32_direct_d8_nt (else tsm_atb_f32_16x32_t8x4_r128_p2_nt)
  {1, 16, 64, 463, 1, 1, 463, 1},  // tsm_atb_f32_16x64_mfma_r64_p2_nt
  {1, 32, 1, 464, 2, 1, 464, 1},  // tsm_atb_f32_32x1_t8x1_r128_p2_nt
  {1, 32, 2, 465, 2, 1, 465, 1},  // tsm_atb_f32_32x2_t8x2_r128_p2_nt
  {1, 32, 4, 472, 2, 1, 472, 1},  // tsm_atb_f32_32x4_t8x4_r128_p2_nt
  {1, 32, 8, 473, 3, 1, 473, 1},  // tsm_atb_f32_32x8_mfma_r64_p1_nt
  {1, 32, 16, 480, 1, 0, 481, 0},  // tsm_atb_f32_32x16_direct_d8_nt (else tsm_atb_f32_32x16_t8x4_r128_p2_nt)
  {1, 32, 32, 492, 1, 1, 492, 1},  // tsm_atb_f32_32x32_mfma_r128_p2_nt
  {1, 32, 64, 539, 1, 0, 540, 0},  // tsm_atb_f32_32x64_direct_d8_nt (else tsm_atb_f32_32x64_mfma_r64_p2_nt)
  {1, 64, 1, 541, 2, 1, 541, 1},  // tsm_atb_f32_64x1_t8x1_r64_p2_nt
  {1, 64, 2, 542, 2, 1, 542, 1},  // tsm_atb_f32_64x2_t8x2_r64_p2_nt
  {1, 64, 4, 549, 2, 1, 549, 1},  // tsm_atb_f32_64x4_t8x4_r64_p2_nt
  {1, 64, 8, 556, 2, 1, 556, 1},  // tsm_atb_f32_64x8_t8x4_r64_p2_nt
  {1, 64, 16, 558, 1, 1, 558, 1},  // tsm_atb_f32_64x16_mfma_r64_p2_nt
  {1, 64, 32, 559, 1, 0, 560, 1},  // tsm_atb_f32_64x32_direct_d8_nt (else tsm_atb_f32_64x32_mfma_r64_p2_nt)
  {1, 64, 64, 562, 1, 0, 570, 1},  // tsm_atb_f32_64x64_direct_d8_nt (else tsm_atb_f32_64x64_mfma_r64_p2)
  {0, 1, 1, 599, 1, 1, 599, 1},  // tsm_atb_f64_1x1_t1x1_r4096_p1_nt
  {0, 1, 2, 602, 1, 1, 602, 1},  // tsm_atb_f64_1x2_t1x2_r1024_p2_nt
  {0, 1, 4, 603, 1, 1, 603, 1},  // tsm_atb_f64_1x4_t1x4_r512_p2_nt
  {0, 1, 8, 604, 1, 1, 604, 1},  // tsm_atb_f64_1x8_t1x8_r256_p2_nt
  {0, 1, 16, 605, 1, 1, 605, 1},  // tsm_atb_f64_1x16_t1x8_r128_p2_nt
  {0, 1, 32, 606, 1, 1, 606, 1},  // tsm_atb_f64_1x32_t1x8_r64_p2_nt
  {0, 1, 64, 607, 1, 1, 607, 1},  // tsm_atb_f64_1x64_t1x8_r32_p2_nt
  {0, 2, 1, 608, 1, 1, 608, 1},  // tsm_atb_f64_2x1_t2x1_r1024_p2_nt
  {0, 2, 2, 620, 1, 1, 620, 1},  // tsm_atb_f64_2x2_t2x2_r2048_p1_nt
  {0, 2, 4, 623, 1, 1, 623, 1},  // tsm_atb_f64_2x4_t2x4_r512_p2_nt
  {0, 2, 8, 624, 1, 1, 624, 1},  // tsm_atb_f64_2x8_t2x8_r256_p2_nt
  {0, 2, 16, 631, 1, 1, 631, 1},  // tsm_atb_f64_2x16_t2x8_r128_p2_nt
  {0, 2, 32, 638, 1, 1, 638, 1},  // tsm_atb_f64_2x32_t2x8_r64_p2_nt
  {0, 2, 64, 645, 1, 1, 645, 1},  // tsm_atb_f64_2x64_t2x8_r32_p2_nt
  {0, 4, 1, 646, 1, 1, 646, 1},  // tsm_atb_f64_4x1_t4x1_r512_p2_nt
  {0, 4, 2, 647, 1, 1, 647, 1},  // tsm_atb_f64_4x2_t4x2_r512_p2_nt
  {0, 4, 4, 651, 2, 1, 651, 1},  // tsm_atb_f64_4x4_t4x4_r256_p1_nt
  {0, 4, 8, 664, 1, 1, 664, 1},  // tsm_atb_f64_4x8_t4x8_r256_p2_nt
  {0, 4, 16, 671, 1, 1, 671, 1},  // tsm_atb_f64_4x16_t4x8_r128_p2_nt
  {0, 4, 32, 690, 2, 1, 690, 1},  // tsm_atb_f64_4x32_t4x8_r64_p1_nt
  {0, 4, 64, 700, 2, 1, 700, 1},  // tsm_atb_f64_4x64_mfma_r32_p1_nt
  {0, 8, 1, 705, 2, 1, 705, 1},  // tsm_atb_f64_8x1_t8x1_r256_p2_nt
  {0, 8, 2, 706, 1, 1, 706, 1},  // tsm_atb_f64_8x2_t8x2_r256_p2_nt
  {0, 8, 4, 707, 1, 1, 707, 1},  // tsm_atb_f64_8x4_t8x4_r256_p2_nt
  {0, 8, 8, 712, 1, 1, 712, 1},  // tsm_atb_f64_8x8_t4x4_r128_p2_nt
  {0, 8, 16, 757, 1, 1, 757, 1},  // tsm_atb_f64_8x16_t8x8_r128_p2_nt
  {0, 8, 32, 764, 1, 1, 764, 1},  // tsm_atb_f64_8x32_t8x8_r64_p2_nt
  {0, 8, 64, 767, 2, 1, 767, 1},  // tsm_atb_f64_8x64_mfma_r32_p1_nt
  {0, 16, 1, 772, 2, 1, 772, 1},  // tsm_atb_f64_16x1_t8x1_r128_p2_nt
  {0, 16, 2, 779, 2, 1, 779, 1},  // tsm_atb_f64_16x2_t8x2_r128_p2_nt
  {0, 16, 4, 786, 1, 1, 786, 1},  // tsm_atb_f64_16x4_t8x4_r128_p2_nt
  {0, 16, 8, 793, 1, 1, 793, 1},  // tsm_atb_f64_16x8_t8x8_r128_p2_nt
  {0, 16, 16, 800, 1, 1, 800, 1},  // tsm_atb_f64_16x16_mfma_r64_p2_nt
  {0, 16, 32, 856, 1, 0, 857, 0},  // tsm_atb_f64_16x32_direct_d8_nt (else tsm_atb_f64_16x32_t8x8_r64_p2_nt)
  {0, 16, 64, 859, 1, 1, 859, 1},  // tsm_atb_f64_16x64_mfma_r32_p2_nt
  {0, 32, 1, 860, 2, 1, 860, 1},  // tsm_atb_f64_32x1_t8x1_r64_p2_nt
  {0, 32, 2, 867, 2, 1, 867, 1},  // tsm_atb_f64_32x2_t8x2_r64_p2_nt
  {0, 32, 4, 877, 1, 1, 877, 1},  // tsm_atb_f64_32x4_t4x4_r32_p4_nt
  {0, 32, 8, 900, 1, 1, 900, 1},  // tsm_atb_f64_32x8_t8x8_r64_p2_nt
  {0, 32, 16, 901, 1, 0, 902, 0},  // tsm_atb_f64_32x16_direct_d8_nt (else tsm_atb_f64_32x16_t8x8_r64_p2_nt)
  {0, 32, 32, 907, 2, 1, 907, 1},  // tsm_atb_f64_32x32_mfma_r32_p1_nt
  {0, 32, 64, 960, 1, 0, 961, 0},  // tsm_atb_f64_32x64_direct_d8_nt (else tsm_atb_f64_32x64_mfma_r32_p2_nt)
  {0, 64, 1, 962, 2, 1, 962, 1},  // tsm_atb_f64_64x1_t8x1_r32_p2_nt
  {0, 64, 2, 969, 2, 1, 969, 1},  // tsm_atb_f64_64x2_t8x2_r32_p2_nt
  {0, 64, 4, 976, 2, 1, 976, 1},  // tsm_atb_f64_64x4_t8x4_r32_p2_nt
  {0, 64, 8, 979, 2, 1, 979, 1},  // tsm_atb_f64_64x8_mfma_r32_p1_nt
  {0, 64, 16, 985, 1, 1, 985, 1},  // tsm_atb_f64_64x16_mfma_r32_p2_nt
  {0, 64, 32, 986, 1, 0, 987, 0},  // tsm_atb_f64_64x32_direct_d8_nt (else tsm_atb_f64_64x32_mfma_r32_p2_nt)
  {0, 64, 64, 989, 1, 0, 998, 0},  // tsm_atb_f64_64x64_direct_d8_nt (else tsm_atb_f64_64x64_mfma_r32_p2_nt)
};

hipError_t launch_tsm_atb_batched_f64(const tsm::BatchArgs&, unsigned, hipStream_t);
const tsm::BatchClass* classes_tsm_atb_batched_f64(int* count);
int occupancy_tsm_atb_batched_f64();
hipError_t launch_tsm_atb_batched_f32(const tsm::BatchArgs&, unsigned, hipStream_t);
const tsm::BatchClass* classes_tsm_atb_batched_f32(int* count);
int occupancy_tsm_atb_batched_f32();
hipError_t launch_tsm_atb_batched_bf16(const tsm::BatchArgs&, unsigned, hipStream_t);
const tsm::BatchClass* classes_tsm_atb_batched_bf16(int* count);
int occupancy_tsm_atb_batched_bf16();
static const tsm::BatchedKernel* batched_all() {
  static tsm::BatchedKernel all[3];
  static bool init = [] {
    all[0].name = "tsm_atb_batched_f64"; all[0].dtype = 0;
    all[0].classes = classes_tsm_atb_batched_f64(&all[0].n_classes);
    all[0].lds_bytes = 81920; all[0].launch = launch_tsm_atb_batched_f64;
    all[0].occupancy = occupancy_tsm_atb_batched_f64;
    all[1].name = "tsm_atb_batched_f32"; all[1].dtype = 1;
    all[1].classes = classes_tsm_atb_batched_f32(&all[1].n_classes);
    all[1].lds_bytes = 81920; all[1].launch = launch_tsm_atb_batched_f32;
    all[1].occupancy = occupancy_tsm_atb_batched_f32;
    all[2].name = "tsm_atb_batched_bf16"; all[2].dtype = 2;
    all[2].classes = classes_tsm_atb_batched_bf16(&all[2].n_classes);
    all[2].lds_bytes = 65536; all[2].launch = launch_tsm_atb_batched_bf16;
    all[2].occupancy = occupancy_tsm_atb_batched_bf16;
    return true;
  }();
  (void)init;
  return all;
}
extern "C" const tsm::BatchedKernel* tsm_gen_batched_table(int* count) {
  *count = 3;
  return batched_all();
}
extern "C" const tsm::KernelEntry* tsm_gen_kernel_table(int* count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
extern "C" const tsm::KernelEntry* tsm_gen_sym_table(int* count) {
  *count = (int)(sizeof(kSym) / sizeof(kSym[0]));
  return kSym;
}
extern "C" const tsm::TunedEntry* tsm_gen_sym_tuned_table(int* count) {
  *count = (int)(sizeof(kSymTuned) / sizeof(kSymTuned[0]));
  return kSymTuned;
}
extern "C" const tsm::TunedEntry* tsm_gen_tuned_table(int* count) {
  *count = (int)(sizeof(kTuned) / sizeof(kTuned[0]));
  return kTuned;
}
