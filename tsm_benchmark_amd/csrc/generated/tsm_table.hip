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
  {1, 4, 16, 281, 1, 1, 281, 1},  // tsm_atb_f32_4x16_t4x4_r256_p2_nt
  {1, 4, 32, 288, 1, 1, 288, 1},  // tsm_atb_f32_4x32_t4x4_r128_p2_nt
  {1, 4, 64, 295, 1, 1, 295, 1},  // tsm_atb_f32_4x64_t4x4_r64_p2_nt
  {1, 8, 1, 296, 2, 1, 296, 1},  // tsm_atb_f32_8x1_t8x1_r512_p2_nt
  {1, 8, 2, 297, 1, 1, 297, 1},  // tsm_atb_f32_8x2_t8x2_r512_p2_nt
  {1, 8, 4, 298, 1, 1, 298, 1},  // tsm_atb_f32_8x4_t8x4_r512_p2_nt
  {1, 8, 8, 334, 1, 1, 334, 1},  // tsm_atb_f32_8x8_t8x4_r1024_p2_nt
  {1, 8, 16, 348, 1, 1, 348, 1},  // tsm_atb_f32_8x16_t8x4_r256_p2_nt
  {1, 8, 32, 350, 2, 1, 350, 1},  // tsm_atb_f32_8x32_mfma_r64_p2_nt
  {1, 8, 64, 362, 2, 1, 362, 1},  // tsm_atb_f32_8x64_t8x4_r64_p2_nt
  {1, 16, 1, 363, 2, 1, 363, 1},  // tsm_atb_f32_16x1_t8x1_r256_p2_nt
  {1, 16, 2, 364, 2, 1, 364, 1},  // tsm_atb_f32_16x2_t8x2_r256_p2_nt
  {1, 16, 4, 369, 1, 0, 369, 0},  // tsm_atb_f32_16x4_t8x4_r256_p2_nt
  {1, 16, 8, 374, 1, 1, 374, 1},  // tsm_atb_f32_16x8_t8x4_r256_p2_nt
  {1, 16, 16, 381, 1, 1, 381, 1},  // tsm_atb_f32_16x16_mfma_r128_p2_nt
  {1, 16, 32, 428, 1, 0, 429, 0},  // tsm_atb_f32_16x32_direct_d8_nt (else tsm_atb_f32_16x32_t8x4_r128_p2_nt)
  {1, 16, 64, 431, 1, 1, 431, 1},  // tsm_atb_f32_16x64_mfma_r64_p2_nt
  {1, 32, 1, 432, 2, 1, 432, 1},  // tsm_atb_f32_32x1_t8x1_r128_p2_nt
  {1, 32, 2, 433, 2, 1, 433, 1},  // tsm_atb_f32_32x2_t8x2_r128_p2_nt
  {1, 32, 4, 440, 2, 1, 440, 1},  // tsm_atb_f32_32x4_t8x4_r128_p2_nt
  {1, 32, 8, 441, 3, 1, 441, 1},  // tsm_atb_f32_32x8_mfma_r64_p1_nt
  {1, 32, 16, 448, 1, 0, 449, 0},  // tsm_atb_f32_32x16_direct_d8_nt (else tsm_atb_f32_32x16_t8x4_r128_p2_nt)
  {1, 32, 32, 460, 1, 1, 460, 1},  // tsm_atb_f32_32x32_mfma_r128_p2_nt
  {1, 32, 64, 507, 1, 0, 508, 0},  // tsm_atb_f32_32x64_direct_d8_nt (else tsm_atb_f32_32x64_mfma_r64_p2_nt)
  {1, 64, 1, 509, 2, 1, 509, 1},  // tsm_atb_f32_64x1_t8x1_r64_p2_nt
  {1, 64, 2, 510, 2, 1, 510, 1},  // tsm_atb_f32_64x2_t8x2_r64_p2_nt
  {1, 64, 4, 517, 2, 1, 517, 1},  // tsm_atb_f32_64x4_t8x4_r64_p2_nt
  {1, 64, 8, 524, 2, 1, 524, 1},  // tsm_atb_f32_64x8_t8x4_r64_p2_nt
  {1, 64, 16, 526, 1, 1, 526, 1},  // tsm_atb_f32_64x16_mfma_r64_p2_nt
  {1, 64, 32, 527, 1, 0, 528, 1},  // tsm_atb_f32_64x32_direct_d8_nt (else tsm_atb_f32_64x32_mfma_r64_p2_nt)
  {1, 64, 64, 530, 1, 0, 538, 1},  // tsm_atb_f32_64x64_direct_d8_nt (else tsm_atb_f32_64x64_mfma_r64_p2)
  {0, 1, 1, 567, 1, 1, 567, 1},  // tsm_atb_f64_1x1_t1x1_r4096_p1_nt
  {0, 1, 2, 570, 1, 1, 570, 1},  // tsm_atb_f64_1x2_t1x2_r1024_p2_nt
  {0, 1, 4, 571, 1, 1, 571, 1},  // tsm_atb_f64_1x4_t1x4_r512_p2_nt
  {0, 1, 8, 572, 1, 1, 572, 1},  // tsm_atb_f64_1x8_t1x8_r256_p2_nt
  {0, 1, 16, 573, 1, 1, 573, 1},  // tsm_atb_f64_1x16_t1x8_r128_p2_nt
  {0, 1, 32, 574, 1, 1, 574, 1},  // tsm_atb_f64_1x32_t1x8_r64_p2_nt
  {0, 1, 64, 575, 1, 1, 575, 1},  // tsm_atb_f64_1x64_t1x8_r32_p2_nt
  {0, 2, 1, 576, 1, 1, 576, 1},  // tsm_atb_f64_2x1_t2x1_r1024_p2_nt
  {0, 2, 2, 588, 1, 1, 588, 1},  // tsm_atb_f64_2x2_t2x2_r2048_p1_nt
  {0, 2, 4, 591, 1, 1, 591, 1},  // tsm_atb_f64_2x4_t2x4_r512_p2_nt
  {0, 2, 8, 592, 1, 1, 592, 1},  // tsm_atb_f64_2x8_t2x8_r256_p2_nt
  {0, 2, 16, 599, 1, 1, 599, 1},  // tsm_atb_f64_2x16_t2x8_r128_p2_nt
  {0, 2, 32, 606, 1, 1, 606, 1},  // tsm_atb_f64_2x32_t2x8_r64_p2_nt
  {0, 2, 64, 613, 1, 1, 613, 1},  // tsm_atb_f64_2x64_t2x8_r32_p2_nt
  {0, 4, 1, 614, 1, 1, 614, 1},  // tsm_atb_f64_4x1_t4x1_r512_p2_nt
  {0, 4, 2, 615, 1, 1, 615, 1},  // tsm_atb_f64_4x2_t4x2_r512_p2_nt
  {0, 4, 4, 619, 2, 1, 619, 1},  // tsm_atb_f64_4x4_t4x4_r256_p1_nt
  {0, 4, 8, 632, 1, 1, 632, 1},  // tsm_atb_f64_4x8_t4x8_r256_p2_nt
  {0, 4, 16, 639, 1, 1, 639, 1},  // tsm_atb_f64_4x16_t4x8_r128_p2_nt
  {0, 4, 32, 658, 2, 1, 658, 1},  // tsm_atb_f64_4x32_t4x8_r64_p1_nt
  {0, 4, 64, 668, 2, 1, 668, 1},  // tsm_atb_f64_4x64_mfma_r32_p1_nt
  {0, 8, 1, 673, 2, 1, 673, 1},  // tsm_atb_f64_8x1_t8x1_r256_p2_nt
  {0, 8, 2, 674, 1, 1, 674, 1},  // tsm_atb_f64_8x2_t8x2_r256_p2_nt
  {0, 8, 4, 675, 1, 1, 675, 1},  // tsm_atb_f64_8x4_t8x4_r256_p2_nt
  {0, 8, 8, 680, 1, 1, 680, 1},  // tsm_atb_f64_8x8_t4x4_r128_p2_nt
  {0, 8, 16, 725, 1, 1, 725, 1},  // tsm_atb_f64_8x16_t8x8_r128_p2_nt
  {0, 8, 32, 732, 1, 1, 732, 1},  // tsm_atb_f64_8x32_t8x8_r64_p2_nt
  {0, 8, 64, 735, 2, 1, 735, 1},  // tsm_atb_f64_8x64_mfma_r32_p1_nt
  {0, 16, 1, 740, 2, 1, 740, 1},  // tsm_atb_f64_16x1_t8x1_r128_p2_nt
  {0, 16, 2, 747, 2, 1, 747, 1},  // tsm_atb_f64_16x2_t8x2_r128_p2_nt
  {0, 16, 4, 754, 1, 1, 754, 1},  // tsm_atb_f64_16x4_t8x4_r128_p2_nt
  {0, 16, 8, 761, 1, 1, 761, 1},  // tsm_atb_f64_16x8_t8x8_r128_p2_nt
  {0, 16, 16, 768, 1, 1, 768, 1},  // tsm_atb_f64_16x16_mfma_r64_p2_nt
  {0, 16, 32, 824, 1, 0, 825, 0},  // tsm_atb_f64_16x32_direct_d8_nt (else tsm_atb_f64_16x32_t8x8_r64_p2_nt)
  {0, 16, 64, 827, 1, 1, 827, 1},  // tsm_atb_f64_16x64_mfma_r32_p2_nt
  {0, 32, 1, 828, 2, 1, 828, 1},  // tsm_atb_f64_32x1_t8x1_r64_p2_nt
  {0, 32, 2, 835, 2, 1, 835, 1},  // tsm_atb_f64_32x2_t8x2_r64_p2_nt
  {0, 32, 4, 845, 1, 1, 845, 1},  // tsm_atb_f64_32x4_t4x4_r32_p4_nt
  {0, 32, 8, 868, 1, 1, 868, 1},  // tsm_atb_f64_32x8_t8x8_r64_p2_nt
  {0, 32, 16, 869, 1, 0, 870, 0},  // tsm_atb_f64_32x16_direct_d8_nt (else tsm_atb_f64_32x16_t8x8_r64_p2_nt)
  {0, 32, 32, 875, 2, 1, 875, 1},  // tsm_atb_f64_32x32_mfma_r32_p1_nt
  {0, 32, 64, 928, 1, 0, 929, 0},  // tsm_atb_f64_32x64_direct_d8_nt (else tsm_atb_f64_32x64_mfma_r32_p2_nt)
  {0, 64, 1, 930, 2, 1, 930, 1},  // tsm_atb_f64_64x1_t8x1_r32_p2_nt
  {0, 64, 2, 937, 2, 1, 937, 1},  // tsm_atb_f64_64x2_t8x2_r32_p2_nt
  {0, 64, 4, 944, 2, 1, 944, 1},  // tsm_atb_f64_64x4_t8x4_r32_p2_nt
  {0, 64, 8, 947, 2, 1, 947, 1},  // tsm_atb_f64_64x8_mfma_r32_p1_nt
  {0, 64, 16, 953, 1, 1, 953, 1},  // tsm_atb_f64_64x16_mfma_r32_p2_nt
  {0, 64, 32, 954, 1, 0, 955, 0},  // tsm_atb_f64_64x32_direct_d8_nt (else tsm_atb_f64_64x32_mfma_r32_p2_nt)
  {0, 64, 64, 957, 1, 0, 966, 0},  // tsm_atb_f64_64x64_direct_d8_nt (else tsm_atb_f64_64x64_mfma_r32_p2_nt)
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
