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
hipError_t launch_tsm_atb_bf16_4x16_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_4x16_t4x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_t4x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_2x16_t2x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_t2x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_4x16_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_16x2_t8x2_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_32x2_t8x2_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_64x2_t8x2_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
  {"tsm_atb_bf16_4x16_t4x4_r512_p2_nt", 2, 4, 16, 4, 4, 512, 2, 1, 40960, launch_tsm_atb_bf16_4x16_t4x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_4x32_t4x4_r256_p2_nt", 2, 4, 32, 4, 4, 256, 2, 1, 36864, launch_tsm_atb_bf16_4x32_t4x4_r256_p2_nt, 1},
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
  {"tsm_atb_f32_4x16_t4x4_r256_p2_nt", 1, 4, 16, 4, 4, 256, 2, 1, 40960, launch_tsm_atb_f32_4x16_t4x4_r256_p2_nt, 1},
  {"tsm_atb_f32_4x32_t4x4_r128_p2_nt", 1, 4, 32, 4, 4, 128, 2, 1, 36864, launch_tsm_atb_f32_4x32_t4x4_r128_p2_nt, 1},
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
  {"tsm_atb_f64_2x16_t2x8_r128_p2_nt", 0, 2, 16, 2, 8, 128, 2, 1, 36864, launch_tsm_atb_f64_2x16_t2x8_r128_p2_nt, 1},
  {"tsm_atb_f64_2x32_t2x8_r64_p2_nt", 0, 2, 32, 2, 8, 64, 2, 1, 34816, launch_tsm_atb_f64_2x32_t2x8_r64_p2_nt, 1},
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
  {"tsm_atb_f64_4x16_t4x8_r128_p2_nt", 0, 4, 16, 4, 8, 128, 2, 1, 40960, launch_tsm_atb_f64_4x16_t4x8_r128_p2_nt, 1},
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
  {"tsm_atb_f64_16x2_t8x2_r128_p2_nt", 0, 16, 2, 8, 2, 128, 2, 1, 36864, launch_tsm_atb_f64_16x2_t8x2_r128_p2_nt, 1},
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
  {"tsm_atb_f64_32x2_t8x2_r64_p2_nt", 0, 32, 2, 8, 2, 64, 2, 1, 34816, launch_tsm_atb_f64_32x2_t8x2_r64_p2_nt, 1},
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
  {"tsm_atb_f64_64x2_t8x2_r32_p2_nt", 0, 64, 2, 8, 2, 32, 2, 1, 33792, launch_tsm_atb_f64_64x2_t8x2_r32_p2_nt, 1},
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
  {2, 4, 16, 18, 2, 0, 18, 0},  // tsm_atb_bf16_4x16_t4x4_r512_p2_nt
  {2, 4, 32, 19, 2, 1, 19, 1},  // tsm_atb_bf16_4x32_t4x4_r256_p2_nt
  {2, 4, 64, 20, 2, 1, 20, 1},  // tsm_atb_bf16_4x64_t4x4_r128_p2_nt
  {2, 8, 1, 21, 2, 1, 21, 1},  // tsm_atb_bf16_8x1_t8x1_r1024_p2_nt
  {2, 8, 2, 22, 1, 1, 22, 1},  // tsm_atb_bf16_8x2_t8x2_r1024_p2_nt
  {2, 8, 4, 23, 1, 1, 23, 1},  // tsm_atb_bf16_8x4_t8x4_r1024_p2_nt
  {2, 8, 8, 49, 2, 1, 49, 1},  // tsm_atb_bf16_8x8_t8x4_r1024_p1_nt
  {2, 8, 16, 67, 1, 1, 67, 1},  // tsm_atb_bf16_8x16_mfma_r256_p2_nt
  {2, 8, 32, 74, 2, 1, 74, 1},  // tsm_atb_bf16_8x32_mfma_r128_p2_nt
  {2, 8, 64, 80, 2, 1, 80, 1},  // tsm_atb_bf16_8x64_mfma_r128_p1_nt
  {2, 16, 1, 85, 2, 1, 85, 1},  // tsm_atb_bf16_16x1_t8x1_r512_p2_nt
  {2, 16, 2, 86, 2, 1, 86, 1},  // tsm_atb_bf16_16x2_t8x2_r512_p2_nt
  {2, 16, 4, 87, 1, 0, 87, 0},  // tsm_atb_bf16_16x4_t8x4_r512_p2_nt
  {2, 16, 8, 89, 1, 1, 89, 1},  // tsm_atb_bf16_16x8_mfma_r256_p2_nt
  {2, 16, 16, 106, 1, 1, 106, 1},  // tsm_atb_bf16_16x16_mfma_r1024_p1_nt
  {2, 16, 32, 152, 1, 1, 152, 1},  // tsm_atb_bf16_16x32_mfma_r256_p2_nt
  {2, 16, 64, 154, 1, 1, 154, 1},  // tsm_atb_bf16_16x64_mfma_r128_p2_nt
  {2, 32, 1, 155, 2, 1, 155, 1},  // tsm_atb_bf16_32x1_t8x1_r256_p2_nt
  {2, 32, 2, 156, 2, 1, 156, 1},  // tsm_atb_bf16_32x2_t8x2_r256_p2_nt
  {2, 32, 4, 157, 2, 1, 157, 1},  // tsm_atb_bf16_32x4_t8x4_r256_p2_nt
  {2, 32, 8, 161, 1, 1, 161, 1},  // tsm_atb_bf16_32x8_mfma_r256_p2_nt
  {2, 32, 16, 166, 1, 1, 166, 1},  // tsm_atb_bf16_32x16_mfma_r256_p2_nt
  {2, 32, 32, 178, 1, 1, 178, 1},  // tsm_atb_bf16_32x32_mfma_r512_p1_nt
  {2, 32, 64, 224, 1, 1, 224, 1},  // tsm_atb_bf16_32x64_mfma_r128_p2_nt
  {2, 64, 1, 225, 2, 1, 225, 1},  // tsm_atb_bf16_64x1_t8x1_r128_p2_nt
  {2, 64, 2, 226, 2, 1, 226, 1},  // tsm_atb_bf16_64x2_t8x2_r128_p2_nt
  {2, 64, 4, 227, 2, 1, 227, 1},  // tsm_atb_bf16_64x4_t8x4_r128_p2_nt
  {2, 64, 8, 228, 2, 1, 228, 1},  // tsm_atb_bf16_64x8_mfma_r128_p1_nt
  {2, 64, 16, 234, 1, 1, 234, 1},  // tsm_atb_bf16_64x16_mfma_r128_p2_nt
  {2, 64, 32, 236, 1, 1, 236, 1},  // tsm_atb_bf16_64x32_mfma_r128_p2_nt
  {2, 64, 64, 238, 1, 0, 246, 1},  // tsm_atb_bf16_64x64_direct_d4_nt (else tsm_atb_bf16_64x64_mfma_r256_p2_nt)
  {1, 1, 1, 259, 1, 1, 259, 1},  // tsm_atb_f32_1x1_t1x1_r4096_p2_nt
  {1, 1, 2, 260, 1, 1, 260, 1},  // tsm_atb_f32_1x2_t1x2_r2048_p2_nt
  {1, 1, 4, 261, 1, 1, 261, 1},  // tsm_atb_f32_1x4_t1x4_r1024_p2_nt
  {1, 1, 8, 262, 1, 1, 262, 1},  // tsm_atb_f32_1x8_t1x4_r512_p2_nt
  {1, 1, 16, 263, 1, 1, 263, 1},  // tsm_atb_f32_1x16_t1x4_r256_p2_nt
  {1, 1, 32, 264, 1, 1, 264, 1},  // tsm_atb_f32_1x32_t1x4_r128_p2_nt
  {1, 1, 64, 265, 1, 1, 265, 1},  // tsm_atb_f32_1x64_t1x4_r64_p2_nt
  {1, 2, 1, 266, 1, 1, 266, 1},  // tsm_atb_f32_2x1_t2x1_r2048_p2_nt
  {1, 2, 2, 267, 1, 1, 267, 1},  // tsm_atb_f32_2x2_t2x2_r2048_p2_nt
  {1, 2, 4, 268, 1, 1, 268, 1},  // tsm_atb_f32_2x4_t2x4_r1024_p2_nt
  {1, 2, 8, 269, 1, 1, 269, 1},  // tsm_atb_f32_2x8_t2x4_r512_p2_nt
  {1, 2, 16, 270, 1, 1, 270, 1},  // tsm_atb_f32_2x16_t2x4_r256_p2_nt
  {1, 2, 32, 271, 1, 1, 271, 1},  // tsm_atb_f32_2x32_t2x4_r128_p2_nt
  {1, 2, 64, 272, 1, 1, 272, 1},  // tsm_atb_f32_2x64_t2x4_r64_p2_nt
  {1, 4, 1, 273, 1, 1, 273, 1},  // tsm_atb_f32_4x1_t4x1_r1024_p2_nt
  {1, 4, 2, 274, 1, 1, 274, 1},  // tsm_atb_f32_4x2_t4x2_r1024_p2_nt
  {1, 4, 4, 275, 1, 1, 275, 1},  // tsm_atb_f32_4x4_t4x4_r1024_p2_nt
  {1, 4, 8, 276, 1, 1, 276, 1},  // tsm_atb_f32_4x8_t4x4_r512_p2_nt
  {1, 4, 16, 277, 1, 1, 277, 1},  // tsm_atb_f32_4x16_t4x4_r256_p2_nt
  {1, 4, 32, 278, 1, 1, 278, 1},  // tsm_atb_f32_4x32_t4x4_r128_p2_nt
  {1, 4, 64, 279, 1, 1, 279, 1},  // tsm_atb_f32_4x64_t4x4_r64_p2_nt
  {1, 8, 1, 280, 2, 1, 280, 1},  // tsm_atb_f32_8x1_t8x1_r512_p2_nt
  {1, 8, 2, 281, 1, 1, 281, 1},  // tsm_atb_f32_8x2_t8x2_r512_p2_nt
  {1, 8, 4, 282, 1, 1, 282, 1},  // tsm_atb_f32_8x4_t8x4_r512_p2_nt
  {1, 8, 8, 318, 1, 1, 318, 1},  // tsm_atb_f32_8x8_t8x4_r1024_p2_nt
  {1, 8, 16, 332, 1, 1, 332, 1},  // tsm_atb_f32_8x16_t8x4_r256_p2_nt
  {1, 8, 32, 334, 2, 1, 334, 1},  // tsm_atb_f32_8x32_mfma_r64_p2_nt
  {1, 8, 64, 346, 2, 1, 346, 1},  // tsm_atb_f32_8x64_t8x4_r64_p2_nt
  {1, 16, 1, 347, 2, 1, 347, 1},  // tsm_atb_f32_16x1_t8x1_r256_p2_nt
  {1, 16, 2, 348, 2, 1, 348, 1},  // tsm_atb_f32_16x2_t8x2_r256_p2_nt
  {1, 16, 4, 349, 1, 0, 349, 0},  // tsm_atb_f32_16x4_t8x4_r256_p2_nt
  {1, 16, 8, 354, 1, 1, 354, 1},  // tsm_atb_f32_16x8_t8x4_r256_p2_nt
  {1, 16, 16, 361, 1, 1, 361, 1},  // tsm_atb_f32_16x16_mfma_r128_p2_nt
  {1, 16, 32, 408, 1, 0, 409, 0},  // tsm_atb_f32_16x32_direct_d8_nt (else tsm_atb_f32_16x32_t8x4_r128_p2_nt)
  {1, 16, 64, 411, 1, 1, 411, 1},  // tsm_atb_f32_16x64_mfma_r64_p2_nt
  {1, 32, 1, 412, 2, 1, 412, 1},  // tsm_atb_f32_32x1_t8x1_r128_p2_nt
  {1, 32, 2, 413, 2, 1, 413, 1},  // tsm_atb_f32_32x2_t8x2_r128_p2_nt
  {1, 32, 4, 414, 2, 1, 414, 1},  // tsm_atb_f32_32x4_t8x4_r128_p2_nt
  {1, 32, 8, 415, 3, 1, 415, 1},  // tsm_atb_f32_32x8_mfma_r64_p1_nt
  {1, 32, 16, 422, 1, 0, 423, 0},  // tsm_atb_f32_32x16_direct_d8_nt (else tsm_atb_f32_32x16_t8x4_r128_p2_nt)
  {1, 32, 32, 434, 1, 1, 434, 1},  // tsm_atb_f32_32x32_mfma_r128_p2_nt
  {1, 32, 64, 481, 1, 0, 482, 0},  // tsm_atb_f32_32x64_direct_d8_nt (else tsm_atb_f32_32x64_mfma_r64_p2_nt)
  {1, 64, 1, 483, 2, 1, 483, 1},  // tsm_atb_f32_64x1_t8x1_r64_p2_nt
  {1, 64, 2, 484, 2, 1, 484, 1},  // tsm_atb_f32_64x2_t8x2_r64_p2_nt
  {1, 64, 4, 485, 2, 1, 485, 1},  // tsm_atb_f32_64x4_t8x4_r64_p2_nt
  {1, 64, 8, 492, 2, 1, 492, 1},  // tsm_atb_f32_64x8_t8x4_r64_p2_nt
  {1, 64, 16, 494, 1, 1, 494, 1},  // tsm_atb_f32_64x16_mfma_r64_p2_nt
  {1, 64, 32, 495, 1, 0, 496, 1},  // tsm_atb_f32_64x32_direct_d8_nt (else tsm_atb_f32_64x32_mfma_r64_p2_nt)
  {1, 64, 64, 498, 1, 0, 506, 1},  // tsm_atb_f32_64x64_direct_d8_nt (else tsm_atb_f32_64x64_mfma_r64_p2)
  {0, 1, 1, 535, 1, 1, 535, 1},  // tsm_atb_f64_1x1_t1x1_r4096_p1_nt
  {0, 1, 2, 538, 1, 1, 538, 1},  // tsm_atb_f64_1x2_t1x2_r1024_p2_nt
  {0, 1, 4, 539, 1, 1, 539, 1},  // tsm_atb_f64_1x4_t1x4_r512_p2_nt
  {0, 1, 8, 540, 1, 1, 540, 1},  // tsm_atb_f64_1x8_t1x8_r256_p2_nt
  {0, 1, 16, 541, 1, 1, 541, 1},  // tsm_atb_f64_1x16_t1x8_r128_p2_nt
  {0, 1, 32, 542, 1, 1, 542, 1},  // tsm_atb_f64_1x32_t1x8_r64_p2_nt
  {0, 1, 64, 543, 1, 1, 543, 1},  // tsm_atb_f64_1x64_t1x8_r32_p2_nt
  {0, 2, 1, 544, 1, 1, 544, 1},  // tsm_atb_f64_2x1_t2x1_r1024_p2_nt
  {0, 2, 2, 556, 1, 1, 556, 1},  // tsm_atb_f64_2x2_t2x2_r2048_p1_nt
  {0, 2, 4, 559, 1, 1, 559, 1},  // tsm_atb_f64_2x4_t2x4_r512_p2_nt
  {0, 2, 8, 560, 1, 1, 560, 1},  // tsm_atb_f64_2x8_t2x8_r256_p2_nt
  {0, 2, 16, 561, 1, 1, 561, 1},  // tsm_atb_f64_2x16_t2x8_r128_p2_nt
  {0, 2, 32, 562, 1, 1, 562, 1},  // tsm_atb_f64_2x32_t2x8_r64_p2_nt
  {0, 2, 64, 563, 1, 1, 563, 1},  // tsm_atb_f64_2x64_t2x8_r32_p2_nt
  {0, 4, 1, 564, 1, 1, 564, 1},  // tsm_atb_f64_4x1_t4x1_r512_p2_nt
  {0, 4, 2, 565, 1, 1, 565, 1},  // tsm_atb_f64_4x2_t4x2_r512_p2_nt
  {0, 4, 4, 569, 2, 1, 569, 1},  // tsm_atb_f64_4x4_t4x4_r256_p1_nt
  {0, 4, 8, 582, 1, 1, 582, 1},  // tsm_atb_f64_4x8_t4x8_r256_p2_nt
  {0, 4, 16, 583, 1, 1, 583, 1},  // tsm_atb_f64_4x16_t4x8_r128_p2_nt
  {0, 4, 32, 596, 2, 1, 596, 1},  // tsm_atb_f64_4x32_t4x8_r64_p1_nt
  {0, 4, 64, 604, 1, 1, 604, 1},  // tsm_atb_f64_4x64_t4x8_r32_p2_nt
  {0, 8, 1, 605, 2, 1, 605, 1},  // tsm_atb_f64_8x1_t8x1_r256_p2_nt
  {0, 8, 2, 606, 1, 1, 606, 1},  // tsm_atb_f64_8x2_t8x2_r256_p2_nt
  {0, 8, 4, 607, 1, 1, 607, 1},  // tsm_atb_f64_8x4_t8x4_r256_p2_nt
  {0, 8, 8, 612, 1, 1, 612, 1},  // tsm_atb_f64_8x8_t4x4_r128_p2_nt
  {0, 8, 16, 657, 1, 1, 657, 1},  // tsm_atb_f64_8x16_t8x8_r128_p2_nt
  {0, 8, 32, 664, 1, 1, 664, 1},  // tsm_atb_f64_8x32_t8x8_r64_p2_nt
  {0, 8, 64, 667, 2, 1, 667, 1},  // tsm_atb_f64_8x64_mfma_r32_p1_nt
  {0, 16, 1, 672, 2, 1, 672, 1},  // tsm_atb_f64_16x1_t8x1_r128_p2_nt
  {0, 16, 2, 673, 2, 1, 673, 1},  // tsm_atb_f64_16x2_t8x2_r128_p2_nt
  {0, 16, 4, 674, 1, 1, 674, 1},  // tsm_atb_f64_16x4_t8x4_r128_p2_nt
  {0, 16, 8, 681, 1, 1, 681, 1},  // tsm_atb_f64_16x8_t8x8_r128_p2_nt
  {0, 16, 16, 688, 1, 1, 688, 1},  // tsm_atb_f64_16x16_mfma_r64_p2_nt
  {0, 16, 32, 744, 1, 0, 745, 0},  // tsm_atb_f64_16x32_direct_d8_nt (else tsm_atb_f64_16x32_t8x8_r64_p2_nt)
  {0, 16, 64, 747, 1, 1, 747, 1},  // tsm_atb_f64_16x64_mfma_r32_p2_nt
  {0, 32, 1, 748, 2, 1, 748, 1},  // tsm_atb_f64_32x1_t8x1_r64_p2_nt
  {0, 32, 2, 749, 2, 1, 749, 1},  // tsm_atb_f64_32x2_t8x2_r64_p2_nt
  {0, 32, 4, 753, 1, 1, 753, 1},  // tsm_atb_f64_32x4_t4x4_r32_p4_nt
  {0, 32, 8, 776, 1, 1, 776, 1},  // tsm_atb_f64_32x8_t8x8_r64_p2_nt
  {0, 32, 16, 777, 1, 0, 778, 0},  // tsm_atb_f64_32x16_direct_d8_nt (else tsm_atb_f64_32x16_t8x8_r64_p2_nt)
  {0, 32, 32, 783, 2, 1, 783, 1},  // tsm_atb_f64_32x32_mfma_r32_p1_nt
  {0, 32, 64, 836, 1, 0, 837, 0},  // tsm_atb_f64_32x64_direct_d8_nt (else tsm_atb_f64_32x64_mfma_r32_p2_nt)
  {0, 64, 1, 838, 2, 1, 838, 1},  // tsm_atb_f64_64x1_t8x1_r32_p2_nt
  {0, 64, 2, 839, 2, 1, 839, 1},  // tsm_atb_f64_64x2_t8x2_r32_p2_nt
  {0, 64, 4, 840, 2, 1, 840, 1},  // tsm_atb_f64_64x4_t8x4_r32_p2_nt
  {0, 64, 8, 843, 2, 1, 843, 1},  // tsm_atb_f64_64x8_mfma_r32_p1_nt
  {0, 64, 16, 849, 1, 1, 849, 1},  // tsm_atb_f64_64x16_mfma_r32_p2_nt
  {0, 64, 32, 850, 1, 0, 851, 0},  // tsm_atb_f64_64x32_direct_d8_nt (else tsm_atb_f64_64x32_mfma_r32_p2_nt)
  {0, 64, 64, 853, 1, 0, 862, 0},  // tsm_atb_f64_64x64_direct_d8_nt (else tsm_atb_f64_64x64_mfma_r32_p2_nt)
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
