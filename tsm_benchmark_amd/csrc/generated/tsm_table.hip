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
hipError_t launch_tsm_atb_bf16_8x16_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x64_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x1_t8x1_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x2_t8x2_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_8x16_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x1_t8x1_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x2_t8x2_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x4_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_8x16_t8x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_t8x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_t8x8_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x1_t8x1_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_t8x2_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
  {"tsm_atb_bf16_8x16_t8x4_r512_p2_nt", 2, 8, 16, 8, 4, 512, 2, 1, 49152, launch_tsm_atb_bf16_8x16_t8x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_8x32_t8x4_r256_p2_nt", 2, 8, 32, 8, 4, 256, 2, 1, 40960, launch_tsm_atb_bf16_8x32_t8x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_8x64_t8x4_r128_p2_nt", 2, 8, 64, 8, 4, 128, 2, 1, 36864, launch_tsm_atb_bf16_8x64_t8x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_16x1_t8x1_r512_p2_nt", 2, 16, 1, 8, 1, 512, 2, 1, 34816, launch_tsm_atb_bf16_16x1_t8x1_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x2_t8x2_r512_p2_nt", 2, 16, 2, 8, 2, 512, 2, 1, 36864, launch_tsm_atb_bf16_16x2_t8x2_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x4_t8x4_r512_p2_nt", 2, 16, 4, 8, 4, 512, 2, 1, 40960, launch_tsm_atb_bf16_16x4_t8x4_r512_p2_nt, 1},
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
  {"tsm_atb_f32_8x16_t8x4_r256_p2_nt", 1, 8, 16, 8, 4, 256, 2, 1, 49152, launch_tsm_atb_f32_8x16_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f32_8x32_t8x4_r128_p2_nt", 1, 8, 32, 8, 4, 128, 2, 1, 40960, launch_tsm_atb_f32_8x32_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_8x64_t8x4_r64_p2_nt", 1, 8, 64, 8, 4, 64, 2, 1, 36864, launch_tsm_atb_f32_8x64_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f32_16x1_t8x1_r256_p2_nt", 1, 16, 1, 8, 1, 256, 2, 1, 34816, launch_tsm_atb_f32_16x1_t8x1_r256_p2_nt, 1},
  {"tsm_atb_f32_16x2_t8x2_r256_p2_nt", 1, 16, 2, 8, 2, 256, 2, 1, 36864, launch_tsm_atb_f32_16x2_t8x2_r256_p2_nt, 1},
  {"tsm_atb_f32_16x4_t8x4_r256_p2_nt", 1, 16, 4, 8, 4, 256, 2, 1, 40960, launch_tsm_atb_f32_16x4_t8x4_r256_p2_nt, 1},
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
  {"tsm_atb_f64_8x16_t8x8_r128_p2_nt", 0, 8, 16, 8, 8, 128, 2, 1, 49152, launch_tsm_atb_f64_8x16_t8x8_r128_p2_nt, 1},
  {"tsm_atb_f64_8x32_t8x8_r64_p2_nt", 0, 8, 32, 8, 8, 64, 2, 1, 40960, launch_tsm_atb_f64_8x32_t8x8_r64_p2_nt, 1},
  {"tsm_atb_f64_8x64_t8x8_r32_p2_nt", 0, 8, 64, 8, 8, 32, 2, 1, 36864, launch_tsm_atb_f64_8x64_t8x8_r32_p2_nt, 1},
  {"tsm_atb_f64_16x1_t8x1_r128_p2_nt", 0, 16, 1, 8, 1, 128, 2, 1, 34816, launch_tsm_atb_f64_16x1_t8x1_r128_p2_nt, 1},
  {"tsm_atb_f64_16x2_t8x2_r128_p2_nt", 0, 16, 2, 8, 2, 128, 2, 1, 36864, launch_tsm_atb_f64_16x2_t8x2_r128_p2_nt, 1},
  {"tsm_atb_f64_16x4_t8x4_r128_p2_nt", 0, 16, 4, 8, 4, 128, 2, 1, 40960, launch_tsm_atb_f64_16x4_t8x4_r128_p2_nt, 1},
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
  {2, 8, 8, 5, 6, 1, 5, 1},  // tsm_ata_bf16_8x8_t8x4_r1024_p1_nt
  {2, 16, 16, 12, 1, 1, 12, 1},  // tsm_ata_bf16_16x16_mfma_r512_p2_nt
  {2, 32, 32, 20, 1, 1, 20, 1},  // tsm_ata_bf16_32x32_mfma_r256_p2_nt
  {2, 64, 64, 26, 1, 0, 27, 1},  // tsm_ata_bf16_64x64_direct_d4_nt (else tsm_ata_bf16_64x64_mfma_r256_p2_nt)
  {1, 1, 1, 28, 1, 1, 28, 1},  // tsm_ata_f32_1x1_t1x1_r4096_p2_nt
  {1, 2, 2, 29, 1, 1, 29, 1},  // tsm_ata_f32_2x2_t2x2_r4096_p2_nt
  {1, 4, 4, 30, 1, 1, 30, 1},  // tsm_ata_f32_4x4_t4x4_r2048_p2_nt
  {1, 8, 8, 33, 2, 0, 33, 0},  // tsm_ata_f32_8x8_t8x4_r1024_p1_nt
  {1, 16, 16, 44, 2, 0, 44, 0},  // tsm_ata_f32_16x16_mfma_r256_p1_nt
  {1, 32, 32, 48, 3, 0, 52, 1},  // tsm_ata_f32_32x32_direct_d16_nt (else tsm_ata_f32_32x32_mfma_r128_p2_nt)
  {1, 64, 64, 57, 2, 0, 58, 1},  // tsm_ata_f32_64x64_direct_d16_nt (else tsm_ata_f32_64x64_mfma_r64_p2)
  {0, 1, 1, 59, 2, 1, 59, 1},  // tsm_ata_f64_1x1_t1x1_r2048_p1_nt
  {0, 2, 2, 64, 1, 1, 64, 1},  // tsm_ata_f64_2x2_t2x2_r1024_p2_nt
  {0, 4, 4, 72, 1, 1, 72, 1},  // tsm_ata_f64_4x4_t4x4_r512_p2_nt
  {0, 8, 8, 83, 1, 0, 83, 0},  // tsm_ata_f64_8x8_t8x8_r512_p2_nt
  {0, 16, 16, 92, 1, 1, 92, 1},  // tsm_ata_f64_16x16_mfma_r128_p2_nt
  {0, 32, 32, 96, 1, 0, 100, 1},  // tsm_ata_f64_32x32_direct_d8_nt (else tsm_ata_f64_32x32_mfma_r32_p1_nt)
  {0, 64, 64, 108, 2, 0, 109, 0},  // tsm_ata_f64_64x64_direct_d16_nt (else tsm_ata_f64_64x64_mfma_r32_p2_nt)
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
  {2, 8, 16, 66, 2, 0, 66, 0},  // tsm_atb_bf16_8x16_t8x4_r512_p2_nt
  {2, 8, 32, 67, 2, 0, 67, 0},  // tsm_atb_bf16_8x32_t8x4_r256_p2_nt
  {2, 8, 64, 68, 2, 0, 68, 0},  // tsm_atb_bf16_8x64_t8x4_r128_p2_nt
  {2, 16, 1, 69, 2, 1, 69, 1},  // tsm_atb_bf16_16x1_t8x1_r512_p2_nt
  {2, 16, 2, 70, 2, 1, 70, 1},  // tsm_atb_bf16_16x2_t8x2_r512_p2_nt
  {2, 16, 4, 71, 1, 0, 71, 0},  // tsm_atb_bf16_16x4_t8x4_r512_p2_nt
  {2, 16, 8, 72, 2, 1, 72, 1},  // tsm_atb_bf16_16x8_t8x4_r512_p2_nt
  {2, 16, 16, 84, 1, 1, 84, 1},  // tsm_atb_bf16_16x16_mfma_r1024_p1_nt
  {2, 16, 32, 130, 1, 1, 130, 1},  // tsm_atb_bf16_16x32_mfma_r256_p2_nt
  {2, 16, 64, 132, 1, 1, 132, 1},  // tsm_atb_bf16_16x64_mfma_r128_p2_nt
  {2, 32, 1, 133, 2, 1, 133, 1},  // tsm_atb_bf16_32x1_t8x1_r256_p2_nt
  {2, 32, 2, 134, 2, 1, 134, 1},  // tsm_atb_bf16_32x2_t8x2_r256_p2_nt
  {2, 32, 4, 135, 2, 1, 135, 1},  // tsm_atb_bf16_32x4_t8x4_r256_p2_nt
  {2, 32, 8, 136, 2, 0, 136, 0},  // tsm_atb_bf16_32x8_t8x4_r256_p2_nt
  {2, 32, 16, 138, 1, 1, 138, 1},  // tsm_atb_bf16_32x16_mfma_r256_p2_nt
  {2, 32, 32, 150, 1, 1, 150, 1},  // tsm_atb_bf16_32x32_mfma_r512_p1_nt
  {2, 32, 64, 196, 1, 1, 196, 1},  // tsm_atb_bf16_32x64_mfma_r128_p2_nt
  {2, 64, 1, 197, 2, 1, 197, 1},  // tsm_atb_bf16_64x1_t8x1_r128_p2_nt
  {2, 64, 2, 198, 2, 1, 198, 1},  // tsm_atb_bf16_64x2_t8x2_r128_p2_nt
  {2, 64, 4, 199, 2, 1, 199, 1},  // tsm_atb_bf16_64x4_t8x4_r128_p2_nt
  {2, 64, 8, 200, 2, 1, 200, 1},  // tsm_atb_bf16_64x8_t8x4_r128_p2_nt
  {2, 64, 16, 202, 1, 1, 202, 1},  // tsm_atb_bf16_64x16_mfma_r128_p2_nt
  {2, 64, 32, 204, 1, 1, 204, 1},  // tsm_atb_bf16_64x32_mfma_r128_p2_nt
  {2, 64, 64, 206, 1, 0, 214, 1},  // tsm_atb_bf16_64x64_direct_d4_nt (else tsm_atb_bf16_64x64_mfma_r256_p2_nt)
  {1, 1, 1, 227, 1, 1, 227, 1},  // tsm_atb_f32_1x1_t1x1_r4096_p2_nt
  {1, 1, 2, 228, 1, 1, 228, 1},  // tsm_atb_f32_1x2_t1x2_r2048_p2_nt
  {1, 1, 4, 229, 1, 1, 229, 1},  // tsm_atb_f32_1x4_t1x4_r1024_p2_nt
  {1, 1, 8, 230, 1, 1, 230, 1},  // tsm_atb_f32_1x8_t1x4_r512_p2_nt
  {1, 1, 16, 231, 1, 1, 231, 1},  // tsm_atb_f32_1x16_t1x4_r256_p2_nt
  {1, 1, 32, 232, 1, 1, 232, 1},  // tsm_atb_f32_1x32_t1x4_r128_p2_nt
  {1, 1, 64, 233, 1, 1, 233, 1},  // tsm_atb_f32_1x64_t1x4_r64_p2_nt
  {1, 2, 1, 234, 1, 1, 234, 1},  // tsm_atb_f32_2x1_t2x1_r2048_p2_nt
  {1, 2, 2, 235, 1, 1, 235, 1},  // tsm_atb_f32_2x2_t2x2_r2048_p2_nt
  {1, 2, 4, 236, 1, 1, 236, 1},  // tsm_atb_f32_2x4_t2x4_r1024_p2_nt
  {1, 2, 8, 237, 1, 1, 237, 1},  // tsm_atb_f32_2x8_t2x4_r512_p2_nt
  {1, 2, 16, 238, 1, 1, 238, 1},  // tsm_atb_f32_2x16_t2x4_r256_p2_nt
  {1, 2, 32, 239, 1, 1, 239, 1},  // tsm_atb_f32_2x32_t2x4_r128_p2_nt
  {1, 2, 64, 240, 1, 1, 240, 1},  // tsm_atb_f32_2x64_t2x4_r64_p2_nt
  {1, 4, 1, 241, 1, 1, 241, 1},  // tsm_atb_f32_4x1_t4x1_r1024_p2_nt
  {1, 4, 2, 242, 1, 1, 242, 1},  // tsm_atb_f32_4x2_t4x2_r1024_p2_nt
  {1, 4, 4, 243, 1, 1, 243, 1},  // tsm_atb_f32_4x4_t4x4_r1024_p2_nt
  {1, 4, 8, 244, 1, 1, 244, 1},  // tsm_atb_f32_4x8_t4x4_r512_p2_nt
  {1, 4, 16, 245, 1, 1, 245, 1},  // tsm_atb_f32_4x16_t4x4_r256_p2_nt
  {1, 4, 32, 246, 1, 1, 246, 1},  // tsm_atb_f32_4x32_t4x4_r128_p2_nt
  {1, 4, 64, 247, 1, 1, 247, 1},  // tsm_atb_f32_4x64_t4x4_r64_p2_nt
  {1, 8, 1, 248, 2, 1, 248, 1},  // tsm_atb_f32_8x1_t8x1_r512_p2_nt
  {1, 8, 2, 249, 1, 1, 249, 1},  // tsm_atb_f32_8x2_t8x2_r512_p2_nt
  {1, 8, 4, 250, 1, 1, 250, 1},  // tsm_atb_f32_8x4_t8x4_r512_p2_nt
  {1, 8, 8, 286, 1, 1, 286, 1},  // tsm_atb_f32_8x8_t8x4_r1024_p2_nt
  {1, 8, 16, 296, 1, 1, 296, 1},  // tsm_atb_f32_8x16_t8x4_r256_p2_nt
  {1, 8, 32, 297, 1, 1, 297, 1},  // tsm_atb_f32_8x32_t8x4_r128_p2_nt
  {1, 8, 64, 298, 2, 1, 298, 1},  // tsm_atb_f32_8x64_t8x4_r64_p2_nt
  {1, 16, 1, 299, 2, 1, 299, 1},  // tsm_atb_f32_16x1_t8x1_r256_p2_nt
  {1, 16, 2, 300, 2, 1, 300, 1},  // tsm_atb_f32_16x2_t8x2_r256_p2_nt
  {1, 16, 4, 301, 1, 0, 301, 0},  // tsm_atb_f32_16x4_t8x4_r256_p2_nt
  {1, 16, 8, 302, 1, 1, 302, 1},  // tsm_atb_f32_16x8_t8x4_r256_p2_nt
  {1, 16, 16, 309, 1, 1, 309, 1},  // tsm_atb_f32_16x16_mfma_r128_p2_nt
  {1, 16, 32, 356, 1, 0, 357, 0},  // tsm_atb_f32_16x32_direct_d8_nt (else tsm_atb_f32_16x32_t8x4_r128_p2_nt)
  {1, 16, 64, 359, 1, 1, 359, 1},  // tsm_atb_f32_16x64_mfma_r64_p2_nt
  {1, 32, 1, 360, 2, 1, 360, 1},  // tsm_atb_f32_32x1_t8x1_r128_p2_nt
  {1, 32, 2, 361, 2, 1, 361, 1},  // tsm_atb_f32_32x2_t8x2_r128_p2_nt
  {1, 32, 4, 362, 2, 1, 362, 1},  // tsm_atb_f32_32x4_t8x4_r128_p2_nt
  {1, 32, 8, 363, 1, 0, 363, 0},  // tsm_atb_f32_32x8_t8x4_r128_p2_nt
  {1, 32, 16, 364, 1, 0, 365, 0},  // tsm_atb_f32_32x16_direct_d8_nt (else tsm_atb_f32_32x16_t8x4_r128_p2_nt)
  {1, 32, 32, 376, 1, 1, 376, 1},  // tsm_atb_f32_32x32_mfma_r128_p2_nt
  {1, 32, 64, 423, 1, 0, 424, 0},  // tsm_atb_f32_32x64_direct_d8_nt (else tsm_atb_f32_32x64_mfma_r64_p2_nt)
  {1, 64, 1, 425, 2, 1, 425, 1},  // tsm_atb_f32_64x1_t8x1_r64_p2_nt
  {1, 64, 2, 426, 2, 1, 426, 1},  // tsm_atb_f32_64x2_t8x2_r64_p2_nt
  {1, 64, 4, 427, 2, 1, 427, 1},  // tsm_atb_f32_64x4_t8x4_r64_p2_nt
  {1, 64, 8, 428, 2, 1, 428, 1},  // tsm_atb_f32_64x8_t8x4_r64_p2_nt
  {1, 64, 16, 430, 1, 1, 430, 1},  // tsm_atb_f32_64x16_mfma_r64_p2_nt
  {1, 64, 32, 431, 1, 0, 432, 1},  // tsm_atb_f32_64x32_direct_d8_nt (else tsm_atb_f32_64x32_mfma_r64_p2_nt)
  {1, 64, 64, 434, 1, 0, 442, 1},  // tsm_atb_f32_64x64_direct_d8_nt (else tsm_atb_f32_64x64_mfma_r64_p2)
  {0, 1, 1, 471, 1, 1, 471, 1},  // tsm_atb_f64_1x1_t1x1_r4096_p1_nt
  {0, 1, 2, 474, 1, 1, 474, 1},  // tsm_atb_f64_1x2_t1x2_r1024_p2_nt
  {0, 1, 4, 475, 1, 1, 475, 1},  // tsm_atb_f64_1x4_t1x4_r512_p2_nt
  {0, 1, 8, 476, 1, 1, 476, 1},  // tsm_atb_f64_1x8_t1x8_r256_p2_nt
  {0, 1, 16, 477, 1, 1, 477, 1},  // tsm_atb_f64_1x16_t1x8_r128_p2_nt
  {0, 1, 32, 478, 1, 1, 478, 1},  // tsm_atb_f64_1x32_t1x8_r64_p2_nt
  {0, 1, 64, 479, 1, 1, 479, 1},  // tsm_atb_f64_1x64_t1x8_r32_p2_nt
  {0, 2, 1, 480, 1, 1, 480, 1},  // tsm_atb_f64_2x1_t2x1_r1024_p2_nt
  {0, 2, 2, 492, 1, 1, 492, 1},  // tsm_atb_f64_2x2_t2x2_r2048_p1_nt
  {0, 2, 4, 495, 1, 1, 495, 1},  // tsm_atb_f64_2x4_t2x4_r512_p2_nt
  {0, 2, 8, 496, 1, 1, 496, 1},  // tsm_atb_f64_2x8_t2x8_r256_p2_nt
  {0, 2, 16, 497, 1, 1, 497, 1},  // tsm_atb_f64_2x16_t2x8_r128_p2_nt
  {0, 2, 32, 498, 1, 1, 498, 1},  // tsm_atb_f64_2x32_t2x8_r64_p2_nt
  {0, 2, 64, 499, 1, 1, 499, 1},  // tsm_atb_f64_2x64_t2x8_r32_p2_nt
  {0, 4, 1, 500, 1, 1, 500, 1},  // tsm_atb_f64_4x1_t4x1_r512_p2_nt
  {0, 4, 2, 501, 1, 1, 501, 1},  // tsm_atb_f64_4x2_t4x2_r512_p2_nt
  {0, 4, 4, 505, 2, 1, 505, 1},  // tsm_atb_f64_4x4_t4x4_r256_p1_nt
  {0, 4, 8, 518, 1, 1, 518, 1},  // tsm_atb_f64_4x8_t4x8_r256_p2_nt
  {0, 4, 16, 519, 1, 1, 519, 1},  // tsm_atb_f64_4x16_t4x8_r128_p2_nt
  {0, 4, 32, 532, 2, 1, 532, 1},  // tsm_atb_f64_4x32_t4x8_r64_p1_nt
  {0, 4, 64, 540, 1, 1, 540, 1},  // tsm_atb_f64_4x64_t4x8_r32_p2_nt
  {0, 8, 1, 541, 2, 1, 541, 1},  // tsm_atb_f64_8x1_t8x1_r256_p2_nt
  {0, 8, 2, 542, 1, 1, 542, 1},  // tsm_atb_f64_8x2_t8x2_r256_p2_nt
  {0, 8, 4, 543, 1, 1, 543, 1},  // tsm_atb_f64_8x4_t8x4_r256_p2_nt
  {0, 8, 8, 548, 1, 1, 548, 1},  // tsm_atb_f64_8x8_t4x4_r128_p2_nt
  {0, 8, 16, 587, 1, 1, 587, 1},  // tsm_atb_f64_8x16_t8x8_r128_p2_nt
  {0, 8, 32, 588, 1, 1, 588, 1},  // tsm_atb_f64_8x32_t8x8_r64_p2_nt
  {0, 8, 64, 589, 2, 1, 589, 1},  // tsm_atb_f64_8x64_t8x8_r32_p2_nt
  {0, 16, 1, 590, 2, 1, 590, 1},  // tsm_atb_f64_16x1_t8x1_r128_p2_nt
  {0, 16, 2, 591, 2, 1, 591, 1},  // tsm_atb_f64_16x2_t8x2_r128_p2_nt
  {0, 16, 4, 592, 1, 1, 592, 1},  // tsm_atb_f64_16x4_t8x4_r128_p2_nt
  {0, 16, 8, 593, 1, 1, 593, 1},  // tsm_atb_f64_16x8_t8x8_r128_p2_nt
  {0, 16, 16, 600, 1, 1, 600, 1},  // tsm_atb_f64_16x16_mfma_r64_p2_nt
  {0, 16, 32, 656, 1, 0, 657, 0},  // tsm_atb_f64_16x32_direct_d8_nt (else tsm_atb_f64_16x32_t8x8_r64_p2_nt)
  {0, 16, 64, 659, 1, 1, 659, 1},  // tsm_atb_f64_16x64_mfma_r32_p2_nt
  {0, 32, 1, 660, 2, 1, 660, 1},  // tsm_atb_f64_32x1_t8x1_r64_p2_nt
  {0, 32, 2, 661, 2, 1, 661, 1},  // tsm_atb_f64_32x2_t8x2_r64_p2_nt
  {0, 32, 4, 665, 1, 1, 665, 1},  // tsm_atb_f64_32x4_t4x4_r32_p4_nt
  {0, 32, 8, 682, 1, 1, 682, 1},  // tsm_atb_f64_32x8_t8x8_r64_p2_nt
  {0, 32, 16, 683, 1, 0, 684, 0},  // tsm_atb_f64_32x16_direct_d8_nt (else tsm_atb_f64_32x16_t8x8_r64_p2_nt)
  {0, 32, 32, 689, 2, 1, 689, 1},  // tsm_atb_f64_32x32_mfma_r32_p1_nt
  {0, 32, 64, 742, 1, 0, 743, 0},  // tsm_atb_f64_32x64_direct_d8_nt (else tsm_atb_f64_32x64_mfma_r32_p2_nt)
  {0, 64, 1, 744, 2, 1, 744, 1},  // tsm_atb_f64_64x1_t8x1_r32_p2_nt
  {0, 64, 2, 745, 2, 1, 745, 1},  // tsm_atb_f64_64x2_t8x2_r32_p2_nt
  {0, 64, 4, 746, 2, 1, 746, 1},  // tsm_atb_f64_64x4_t8x4_r32_p2_nt
  {0, 64, 8, 747, 2, 1, 747, 1},  // tsm_atb_f64_64x8_t8x8_r32_p2_nt
  {0, 64, 16, 749, 1, 1, 749, 1},  // tsm_atb_f64_64x16_mfma_r32_p2_nt
  {0, 64, 32, 750, 1, 0, 751, 0},  // tsm_atb_f64_64x32_direct_d8_nt (else tsm_atb_f64_64x32_mfma_r32_p2_nt)
  {0, 64, 64, 753, 1, 0, 762, 0},  // tsm_atb_f64_64x64_direct_d8_nt (else tsm_atb_f64_64x64_mfma_r32_p2_nt)
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
