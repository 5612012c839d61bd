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
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t4x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t4x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_16x16_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x8_t4x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_t4x8_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_16x16_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_32x4_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_ata_bf16_16x16_mfma_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r2048_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_bf16_16x16_mfma_r2048_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_ata_f32_16x16_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_16x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_32x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_64x64_direct_d4_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_64x64_direct_d8_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_64x64_direct_d16_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f32_64x64_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_ata_f64_8x8_t8x8_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t8x8_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t8x8_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t8x8_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t8x8_r1024_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_8x8_t8x8_r1024_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r32_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r32_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_16x16_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_ata_f64_32x32_mfma_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
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
  {"tsm_atb_bf16_8x8_t4x4_r1024_p1", 2, 8, 8, 4, 4, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p1, 1},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p2", 2, 8, 8, 4, 4, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p2, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p1", 2, 8, 8, 8, 4, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p1, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p1_nt", 2, 8, 8, 8, 4, 512, 1, 1, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p1_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p2", 2, 8, 8, 8, 4, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p2, 1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p2_nt", 2, 8, 8, 8, 4, 512, 2, 1, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p1", 2, 8, 8, 8, 4, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p1, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p1_nt", 2, 8, 8, 8, 4, 1024, 1, 1, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p1_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p2", 2, 8, 8, 8, 4, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p2, 1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p2_nt", 2, 8, 8, 8, 4, 1024, 2, 1, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p1", 2, 8, 8, 8, 4, 2048, 1, 0, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p1, 1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p1_nt", 2, 8, 8, 8, 4, 2048, 1, 1, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p1_nt, 1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p2", 2, 8, 8, 8, 4, 2048, 2, 0, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p2, 1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p2_nt", 2, 8, 8, 8, 4, 2048, 2, 1, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p2_nt, 1},
  {"tsm_atb_bf16_8x8_t8x8_r512_p1", 2, 8, 8, 8, 8, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p1, 1},
  {"tsm_atb_bf16_8x8_t8x8_r512_p2", 2, 8, 8, 8, 8, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p2, 1},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p1", 2, 8, 8, 8, 8, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p1, 1},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p2", 2, 8, 8, 8, 8, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p2, 1},
  {"tsm_atb_bf16_8x8_t8x8_r2048_p1", 2, 8, 8, 8, 8, 2048, 1, 0, 131072, launch_tsm_atb_bf16_8x8_t8x8_r2048_p1, 1},
  {"tsm_atb_bf16_8x8_t8x8_r2048_p2", 2, 8, 8, 8, 8, 2048, 2, 0, 131072, launch_tsm_atb_bf16_8x8_t8x8_r2048_p2, 1},
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
  {"tsm_atb_bf16_16x16_t4x4_r256_p1", 2, 16, 16, 4, 4, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p1, 1},
  {"tsm_atb_bf16_16x16_t4x4_r256_p2", 2, 16, 16, 4, 4, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p2, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p1", 2, 16, 16, 8, 4, 128, 1, 0, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p1, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p1_nt", 2, 16, 16, 8, 4, 128, 1, 1, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p1_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p2", 2, 16, 16, 8, 4, 128, 2, 0, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p2, 1},
  {"tsm_atb_bf16_16x16_t8x4_r128_p2_nt", 2, 16, 16, 8, 4, 128, 2, 1, 16384, launch_tsm_atb_bf16_16x16_t8x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p1", 2, 16, 16, 8, 4, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p1, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p1_nt", 2, 16, 16, 8, 4, 256, 1, 1, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p1_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p2", 2, 16, 16, 8, 4, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p2, 1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p2_nt", 2, 16, 16, 8, 4, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p1", 2, 16, 16, 8, 4, 512, 1, 0, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p1, 1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p1_nt", 2, 16, 16, 8, 4, 512, 1, 1, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p1_nt, 1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p2", 2, 16, 16, 8, 4, 512, 2, 0, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p2, 1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p2_nt", 2, 16, 16, 8, 4, 512, 2, 1, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p2_nt, 1},
  {"tsm_atb_bf16_16x16_t8x8_r128_p1", 2, 16, 16, 8, 8, 128, 1, 0, 16384, launch_tsm_atb_bf16_16x16_t8x8_r128_p1, 1},
  {"tsm_atb_bf16_16x16_t8x8_r128_p2", 2, 16, 16, 8, 8, 128, 2, 0, 16384, launch_tsm_atb_bf16_16x16_t8x8_r128_p2, 1},
  {"tsm_atb_bf16_16x16_t8x8_r256_p1", 2, 16, 16, 8, 8, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p1, 1},
  {"tsm_atb_bf16_16x16_t8x8_r256_p2", 2, 16, 16, 8, 8, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p2, 1},
  {"tsm_atb_bf16_16x16_t8x8_r512_p1", 2, 16, 16, 8, 8, 512, 1, 0, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p1, 1},
  {"tsm_atb_bf16_16x16_t8x8_r512_p2", 2, 16, 16, 8, 8, 512, 2, 0, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p2, 1},
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
  {"tsm_atb_bf16_32x32_t4x4_r64_p1", 2, 32, 32, 4, 4, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t4x4_r64_p1, 1},
  {"tsm_atb_bf16_32x32_t4x4_r64_p2", 2, 32, 32, 4, 4, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t4x4_r64_p2, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p1", 2, 32, 32, 8, 4, 32, 1, 0, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p1, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p1_nt", 2, 32, 32, 8, 4, 32, 1, 1, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p1_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p2", 2, 32, 32, 8, 4, 32, 2, 0, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p2, 1},
  {"tsm_atb_bf16_32x32_t8x4_r32_p2_nt", 2, 32, 32, 8, 4, 32, 2, 1, 8192, launch_tsm_atb_bf16_32x32_t8x4_r32_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p1", 2, 32, 32, 8, 4, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p1, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p1_nt", 2, 32, 32, 8, 4, 64, 1, 1, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p1_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p2", 2, 32, 32, 8, 4, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p2, 1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p2_nt", 2, 32, 32, 8, 4, 64, 2, 1, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p1", 2, 32, 32, 8, 4, 128, 1, 0, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p1, 1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p1_nt", 2, 32, 32, 8, 4, 128, 1, 1, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p1_nt, 1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p2", 2, 32, 32, 8, 4, 128, 2, 0, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p2, 1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p2_nt", 2, 32, 32, 8, 4, 128, 2, 1, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p2_nt, 1},
  {"tsm_atb_bf16_32x32_t8x8_r32_p1", 2, 32, 32, 8, 8, 32, 1, 0, 8192, launch_tsm_atb_bf16_32x32_t8x8_r32_p1, 1},
  {"tsm_atb_bf16_32x32_t8x8_r32_p2", 2, 32, 32, 8, 8, 32, 2, 0, 8192, launch_tsm_atb_bf16_32x32_t8x8_r32_p2, 1},
  {"tsm_atb_bf16_32x32_t8x8_r64_p1", 2, 32, 32, 8, 8, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t8x8_r64_p1, 1},
  {"tsm_atb_bf16_32x32_t8x8_r64_p2", 2, 32, 32, 8, 8, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t8x8_r64_p2, 1},
  {"tsm_atb_bf16_32x32_t8x8_r128_p1", 2, 32, 32, 8, 8, 128, 1, 0, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p1, 1},
  {"tsm_atb_bf16_32x32_t8x8_r128_p2", 2, 32, 32, 8, 8, 128, 2, 0, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p2, 1},
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
  {"tsm_atb_f32_8x8_t4x4_r512_p1", 1, 8, 8, 4, 4, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p1, 1},
  {"tsm_atb_f32_8x8_t4x4_r512_p2", 1, 8, 8, 4, 4, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p2, 1},
  {"tsm_atb_f32_8x8_t4x4_r1024_p1", 1, 8, 8, 4, 4, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t4x4_r1024_p1, 1},
  {"tsm_atb_f32_8x8_t4x4_r1024_p2", 1, 8, 8, 4, 4, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t4x4_r1024_p2, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p1", 1, 8, 8, 8, 4, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p1, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p1_nt", 1, 8, 8, 8, 4, 256, 1, 1, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p1_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p2", 1, 8, 8, 8, 4, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p2, 1},
  {"tsm_atb_f32_8x8_t8x4_r256_p2_nt", 1, 8, 8, 8, 4, 256, 2, 1, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p1", 1, 8, 8, 8, 4, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p1, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p1_nt", 1, 8, 8, 8, 4, 512, 1, 1, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p1_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p2", 1, 8, 8, 8, 4, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p2, 1},
  {"tsm_atb_f32_8x8_t8x4_r512_p2_nt", 1, 8, 8, 8, 4, 512, 2, 1, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p1", 1, 8, 8, 8, 4, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p1, 1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p1_nt", 1, 8, 8, 8, 4, 1024, 1, 1, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p1_nt, 1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p2", 1, 8, 8, 8, 4, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p2, 1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p2_nt", 1, 8, 8, 8, 4, 1024, 2, 1, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p2_nt, 1},
  {"tsm_atb_f32_8x8_t8x8_r256_p1", 1, 8, 8, 8, 8, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p1, 1},
  {"tsm_atb_f32_8x8_t8x8_r256_p2", 1, 8, 8, 8, 8, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p2, 1},
  {"tsm_atb_f32_8x8_t8x8_r512_p1", 1, 8, 8, 8, 8, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p1, 1},
  {"tsm_atb_f32_8x8_t8x8_r512_p2", 1, 8, 8, 8, 8, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p2, 1},
  {"tsm_atb_f32_8x8_t8x8_r1024_p1", 1, 8, 8, 8, 8, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t8x8_r1024_p1, 1},
  {"tsm_atb_f32_8x8_t8x8_r1024_p2", 1, 8, 8, 8, 8, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t8x8_r1024_p2, 1},
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
  {"tsm_atb_f32_16x16_t4x4_r256_p1", 1, 16, 16, 4, 4, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p1, 1},
  {"tsm_atb_f32_16x16_t4x4_r256_p2", 1, 16, 16, 4, 4, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p2, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p1", 1, 16, 16, 8, 4, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p1, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p1_nt", 1, 16, 16, 8, 4, 128, 1, 1, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p1_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p2", 1, 16, 16, 8, 4, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p2, 1},
  {"tsm_atb_f32_16x16_t8x4_r128_p2_nt", 1, 16, 16, 8, 4, 128, 2, 1, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p1", 1, 16, 16, 8, 4, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p1, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p1_nt", 1, 16, 16, 8, 4, 256, 1, 1, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p1_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p2", 1, 16, 16, 8, 4, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p2, 1},
  {"tsm_atb_f32_16x16_t8x4_r256_p2_nt", 1, 16, 16, 8, 4, 256, 2, 1, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r512_p1", 1, 16, 16, 8, 4, 512, 1, 0, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p1, 1},
  {"tsm_atb_f32_16x16_t8x4_r512_p1_nt", 1, 16, 16, 8, 4, 512, 1, 1, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p1_nt, 1},
  {"tsm_atb_f32_16x16_t8x4_r512_p2", 1, 16, 16, 8, 4, 512, 2, 0, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p2, 1},
  {"tsm_atb_f32_16x16_t8x4_r512_p2_nt", 1, 16, 16, 8, 4, 512, 2, 1, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p2_nt, 1},
  {"tsm_atb_f32_16x16_t8x8_r128_p1", 1, 16, 16, 8, 8, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p1, 1},
  {"tsm_atb_f32_16x16_t8x8_r128_p2", 1, 16, 16, 8, 8, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p2, 1},
  {"tsm_atb_f32_16x16_t8x8_r256_p1", 1, 16, 16, 8, 8, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p1, 1},
  {"tsm_atb_f32_16x16_t8x8_r256_p2", 1, 16, 16, 8, 8, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p2, 1},
  {"tsm_atb_f32_16x16_t8x8_r512_p1", 1, 16, 16, 8, 8, 512, 1, 0, 131072, launch_tsm_atb_f32_16x16_t8x8_r512_p1, 1},
  {"tsm_atb_f32_16x16_t8x8_r512_p2", 1, 16, 16, 8, 8, 512, 2, 0, 131072, launch_tsm_atb_f32_16x16_t8x8_r512_p2, 1},
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
  {"tsm_atb_f32_32x32_t4x4_r64_p1", 1, 32, 32, 4, 4, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p1, 1},
  {"tsm_atb_f32_32x32_t4x4_r64_p2", 1, 32, 32, 4, 4, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p2, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p1", 1, 32, 32, 8, 4, 32, 1, 0, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p1, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p1_nt", 1, 32, 32, 8, 4, 32, 1, 1, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p1_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p2", 1, 32, 32, 8, 4, 32, 2, 0, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p2, 1},
  {"tsm_atb_f32_32x32_t8x4_r32_p2_nt", 1, 32, 32, 8, 4, 32, 2, 1, 16384, launch_tsm_atb_f32_32x32_t8x4_r32_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p1", 1, 32, 32, 8, 4, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p1, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p1_nt", 1, 32, 32, 8, 4, 64, 1, 1, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p1_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p2", 1, 32, 32, 8, 4, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p2, 1},
  {"tsm_atb_f32_32x32_t8x4_r64_p2_nt", 1, 32, 32, 8, 4, 64, 2, 1, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r128_p1", 1, 32, 32, 8, 4, 128, 1, 0, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p1, 1},
  {"tsm_atb_f32_32x32_t8x4_r128_p1_nt", 1, 32, 32, 8, 4, 128, 1, 1, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p1_nt, 1},
  {"tsm_atb_f32_32x32_t8x4_r128_p2", 1, 32, 32, 8, 4, 128, 2, 0, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p2, 1},
  {"tsm_atb_f32_32x32_t8x4_r128_p2_nt", 1, 32, 32, 8, 4, 128, 2, 1, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p2_nt, 1},
  {"tsm_atb_f32_32x32_t8x8_r32_p1", 1, 32, 32, 8, 8, 32, 1, 0, 16384, launch_tsm_atb_f32_32x32_t8x8_r32_p1, 1},
  {"tsm_atb_f32_32x32_t8x8_r32_p2", 1, 32, 32, 8, 8, 32, 2, 0, 16384, launch_tsm_atb_f32_32x32_t8x8_r32_p2, 1},
  {"tsm_atb_f32_32x32_t8x8_r64_p1", 1, 32, 32, 8, 8, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p1, 1},
  {"tsm_atb_f32_32x32_t8x8_r64_p2", 1, 32, 32, 8, 8, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p2, 1},
  {"tsm_atb_f32_32x32_t8x8_r128_p1", 1, 32, 32, 8, 8, 128, 1, 0, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p1, 1},
  {"tsm_atb_f32_32x32_t8x8_r128_p2", 1, 32, 32, 8, 8, 128, 2, 0, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p2, 1},
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
  {"tsm_atb_f64_1x1_t1x1_r2048_p1", 0, 1, 1, 1, 1, 2048, 1, 0, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p1, 1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p1_nt", 0, 1, 1, 1, 1, 2048, 1, 1, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p1_nt, 1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p2", 0, 1, 1, 1, 1, 2048, 2, 0, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p2, 1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p2_nt", 0, 1, 1, 1, 1, 2048, 2, 1, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p2_nt, 1},
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
  {"tsm_atb_f64_2x2_t2x2_r1024_p1", 0, 2, 2, 2, 2, 1024, 1, 0, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p1, 1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p1_nt", 0, 2, 2, 2, 2, 1024, 1, 1, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p1_nt, 1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p2", 0, 2, 2, 2, 2, 1024, 2, 0, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p2, 1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p2_nt", 0, 2, 2, 2, 2, 1024, 2, 1, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p2_nt, 1},
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
  {"tsm_atb_f64_4x4_t4x4_r512_p1", 0, 4, 4, 4, 4, 512, 1, 0, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p1, 1},
  {"tsm_atb_f64_4x4_t4x4_r512_p1_nt", 0, 4, 4, 4, 4, 512, 1, 1, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p1_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r512_p2", 0, 4, 4, 4, 4, 512, 2, 0, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p2, 1},
  {"tsm_atb_f64_4x4_t4x4_r512_p2_nt", 0, 4, 4, 4, 4, 512, 2, 1, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p2_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p1", 0, 4, 4, 4, 4, 1024, 1, 0, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p1, 1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p1_nt", 0, 4, 4, 4, 4, 1024, 1, 1, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p1_nt, 1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p2", 0, 4, 4, 4, 4, 1024, 2, 0, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p2, 1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p2_nt", 0, 4, 4, 4, 4, 1024, 2, 1, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p2_nt, 1},
  {"tsm_atb_f64_4x8_t4x8_r256_p2_nt", 0, 4, 8, 4, 8, 256, 2, 1, 49152, launch_tsm_atb_f64_4x8_t4x8_r256_p2_nt, 1},
  {"tsm_atb_f64_4x16_t4x8_r128_p2_nt", 0, 4, 16, 4, 8, 128, 2, 1, 40960, launch_tsm_atb_f64_4x16_t4x8_r128_p2_nt, 1},
  {"tsm_atb_f64_4x32_t4x4_r32_p1", 0, 4, 32, 4, 4, 32, 1, 0, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p1, 1},
  {"tsm_atb_f64_4x32_t4x4_r32_p2", 0, 4, 32, 4, 4, 32, 2, 0, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p2, 1},
  {"tsm_atb_f64_4x32_t4x4_r64_p1", 0, 4, 32, 4, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p1, 1},
  {"tsm_atb_f64_4x32_t4x4_r64_p2", 0, 4, 32, 4, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p2, 1},
  {"tsm_atb_f64_4x32_t4x4_r128_p1", 0, 4, 32, 4, 4, 128, 1, 0, 73728, launch_tsm_atb_f64_4x32_t4x4_r128_p1, 1},
  {"tsm_atb_f64_4x32_t4x4_r128_p2", 0, 4, 32, 4, 4, 128, 2, 0, 73728, launch_tsm_atb_f64_4x32_t4x4_r128_p2, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p1", 0, 4, 32, 4, 8, 64, 1, 0, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p1, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p1_nt", 0, 4, 32, 4, 8, 64, 1, 1, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p1_nt, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p2", 0, 4, 32, 4, 8, 64, 2, 0, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p2, 1},
  {"tsm_atb_f64_4x32_t4x8_r64_p2_nt", 0, 4, 32, 4, 8, 64, 2, 1, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p2_nt, 1},
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
  {"tsm_atb_f64_8x8_t4x4_r256_p1", 0, 8, 8, 4, 4, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p1, 1},
  {"tsm_atb_f64_8x8_t4x4_r256_p2", 0, 8, 8, 4, 4, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p2, 1},
  {"tsm_atb_f64_8x8_t4x4_r512_p1", 0, 8, 8, 4, 4, 512, 1, 0, 131072, launch_tsm_atb_f64_8x8_t4x4_r512_p1, 1},
  {"tsm_atb_f64_8x8_t4x4_r512_p2", 0, 8, 8, 4, 4, 512, 2, 0, 131072, launch_tsm_atb_f64_8x8_t4x4_r512_p2, 1},
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
  {"tsm_atb_f64_16x16_t4x4_r128_p1", 0, 16, 16, 4, 4, 128, 1, 0, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p1, 1},
  {"tsm_atb_f64_16x16_t4x4_r128_p2", 0, 16, 16, 4, 4, 128, 2, 0, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p2, 1},
  {"tsm_atb_f64_16x16_t4x4_r256_p1", 0, 16, 16, 4, 4, 256, 1, 0, 131072, launch_tsm_atb_f64_16x16_t4x4_r256_p1, 1},
  {"tsm_atb_f64_16x16_t4x4_r256_p2", 0, 16, 16, 4, 4, 256, 2, 0, 131072, launch_tsm_atb_f64_16x16_t4x4_r256_p2, 1},
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
  {"tsm_atb_f64_32x4_t4x4_r64_p1", 0, 32, 4, 4, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p1, 1},
  {"tsm_atb_f64_32x4_t4x4_r64_p2", 0, 32, 4, 4, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p2, 1},
  {"tsm_atb_f64_32x4_t4x4_r128_p1", 0, 32, 4, 4, 4, 128, 1, 0, 73728, launch_tsm_atb_f64_32x4_t4x4_r128_p1, 1},
  {"tsm_atb_f64_32x4_t4x4_r128_p2", 0, 32, 4, 4, 4, 128, 2, 0, 73728, launch_tsm_atb_f64_32x4_t4x4_r128_p2, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p1", 0, 32, 4, 8, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p1, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p1_nt", 0, 32, 4, 8, 4, 64, 1, 1, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p1_nt, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p2", 0, 32, 4, 8, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p2, 1},
  {"tsm_atb_f64_32x4_t8x4_r64_p2_nt", 0, 32, 4, 8, 4, 64, 2, 1, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p2_nt, 1},
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
  {"tsm_atb_f64_32x32_t4x4_r64_p1", 0, 32, 32, 4, 4, 64, 1, 0, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p1, 1},
  {"tsm_atb_f64_32x32_t4x4_r64_p2", 0, 32, 32, 4, 4, 64, 2, 0, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p2, 1},
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
  {"tsm_ata_bf16_16x16_mfma_r512_p1_nt", 2, 16, 16, 0, 0, 512, 1, 1, 32768, launch_tsm_ata_bf16_16x16_mfma_r512_p1_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r512_p2_nt", 2, 16, 16, 0, 0, 512, 2, 1, 32768, launch_tsm_ata_bf16_16x16_mfma_r512_p2_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r1024_p1_nt", 2, 16, 16, 0, 0, 1024, 1, 1, 65536, launch_tsm_ata_bf16_16x16_mfma_r1024_p1_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r1024_p2_nt", 2, 16, 16, 0, 0, 1024, 2, 1, 65536, launch_tsm_ata_bf16_16x16_mfma_r1024_p2_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r2048_p1_nt", 2, 16, 16, 0, 0, 2048, 1, 1, 131072, launch_tsm_ata_bf16_16x16_mfma_r2048_p1_nt, 2},
  {"tsm_ata_bf16_16x16_mfma_r2048_p2_nt", 2, 16, 16, 0, 0, 2048, 2, 1, 131072, launch_tsm_ata_bf16_16x16_mfma_r2048_p2_nt, 2},
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
  {"tsm_ata_f32_16x16_mfma_r64_p1_nt", 1, 16, 16, 0, 0, 64, 1, 1, 8192, launch_tsm_ata_f32_16x16_mfma_r64_p1_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r64_p2_nt", 1, 16, 16, 0, 0, 64, 2, 1, 8192, launch_tsm_ata_f32_16x16_mfma_r64_p2_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r128_p1_nt", 1, 16, 16, 0, 0, 128, 1, 1, 16384, launch_tsm_ata_f32_16x16_mfma_r128_p1_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r128_p2_nt", 1, 16, 16, 0, 0, 128, 2, 1, 16384, launch_tsm_ata_f32_16x16_mfma_r128_p2_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r256_p1_nt", 1, 16, 16, 0, 0, 256, 1, 1, 32768, launch_tsm_ata_f32_16x16_mfma_r256_p1_nt, 2},
  {"tsm_ata_f32_16x16_mfma_r256_p2_nt", 1, 16, 16, 0, 0, 256, 2, 1, 32768, launch_tsm_ata_f32_16x16_mfma_r256_p2_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r64_p1_nt", 1, 32, 32, 0, 0, 64, 1, 1, 16384, launch_tsm_ata_f32_32x32_mfma_r64_p1_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r64_p2_nt", 1, 32, 32, 0, 0, 64, 2, 1, 16384, launch_tsm_ata_f32_32x32_mfma_r64_p2_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r128_p1_nt", 1, 32, 32, 0, 0, 128, 1, 1, 32768, launch_tsm_ata_f32_32x32_mfma_r128_p1_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r128_p2_nt", 1, 32, 32, 0, 0, 128, 2, 1, 32768, launch_tsm_ata_f32_32x32_mfma_r128_p2_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r256_p1_nt", 1, 32, 32, 0, 0, 256, 1, 1, 65536, launch_tsm_ata_f32_32x32_mfma_r256_p1_nt, 2},
  {"tsm_ata_f32_32x32_mfma_r256_p2_nt", 1, 32, 32, 0, 0, 256, 2, 1, 65536, launch_tsm_ata_f32_32x32_mfma_r256_p2_nt, 2},
  {"tsm_ata_f32_64x64_direct_d4_nt", 1, 64, 64, 0, 0, 4, 4, 1, 65536, launch_tsm_ata_f32_64x64_direct_d4_nt, 3},
  {"tsm_ata_f32_64x64_direct_d8_nt", 1, 64, 64, 0, 0, 4, 8, 1, 65536, launch_tsm_ata_f32_64x64_direct_d8_nt, 3},
  {"tsm_ata_f32_64x64_direct_d16_nt", 1, 64, 64, 0, 0, 4, 16, 1, 65536, launch_tsm_ata_f32_64x64_direct_d16_nt, 3},
  {"tsm_ata_f32_64x64_mfma_r64_p2_nt", 1, 64, 64, 0, 0, 64, 2, 1, 32768, launch_tsm_ata_f32_64x64_mfma_r64_p2_nt, 2},
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
  {"tsm_ata_f64_8x8_t8x8_r256_p1_nt", 0, 8, 8, 8, 8, 256, 1, 1, 32768, launch_tsm_ata_f64_8x8_t8x8_r256_p1_nt, 1},
  {"tsm_ata_f64_8x8_t8x8_r256_p2_nt", 0, 8, 8, 8, 8, 256, 2, 1, 32768, launch_tsm_ata_f64_8x8_t8x8_r256_p2_nt, 1},
  {"tsm_ata_f64_8x8_t8x8_r512_p1_nt", 0, 8, 8, 8, 8, 512, 1, 1, 65536, launch_tsm_ata_f64_8x8_t8x8_r512_p1_nt, 1},
  {"tsm_ata_f64_8x8_t8x8_r512_p2_nt", 0, 8, 8, 8, 8, 512, 2, 1, 65536, launch_tsm_ata_f64_8x8_t8x8_r512_p2_nt, 1},
  {"tsm_ata_f64_8x8_t8x8_r1024_p1_nt", 0, 8, 8, 8, 8, 1024, 1, 1, 131072, launch_tsm_ata_f64_8x8_t8x8_r1024_p1_nt, 1},
  {"tsm_ata_f64_8x8_t8x8_r1024_p2_nt", 0, 8, 8, 8, 8, 1024, 2, 1, 131072, launch_tsm_ata_f64_8x8_t8x8_r1024_p2_nt, 1},
  {"tsm_ata_f64_16x16_mfma_r32_p1_nt", 0, 16, 16, 0, 0, 32, 1, 1, 8192, launch_tsm_ata_f64_16x16_mfma_r32_p1_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r32_p2_nt", 0, 16, 16, 0, 0, 32, 2, 1, 8192, launch_tsm_ata_f64_16x16_mfma_r32_p2_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r64_p1_nt", 0, 16, 16, 0, 0, 64, 1, 1, 16384, launch_tsm_ata_f64_16x16_mfma_r64_p1_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r64_p2_nt", 0, 16, 16, 0, 0, 64, 2, 1, 16384, launch_tsm_ata_f64_16x16_mfma_r64_p2_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r128_p1_nt", 0, 16, 16, 0, 0, 128, 1, 1, 32768, launch_tsm_ata_f64_16x16_mfma_r128_p1_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r128_p2_nt", 0, 16, 16, 0, 0, 128, 2, 1, 32768, launch_tsm_ata_f64_16x16_mfma_r128_p2_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r256_p1_nt", 0, 16, 16, 0, 0, 256, 1, 1, 65536, launch_tsm_ata_f64_16x16_mfma_r256_p1_nt, 2},
  {"tsm_ata_f64_16x16_mfma_r256_p2_nt", 0, 16, 16, 0, 0, 256, 2, 1, 65536, launch_tsm_ata_f64_16x16_mfma_r256_p2_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r64_p1_nt", 0, 32, 32, 0, 0, 64, 1, 1, 32768, launch_tsm_ata_f64_32x32_mfma_r64_p1_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r64_p2_nt", 0, 32, 32, 0, 0, 64, 2, 1, 32768, launch_tsm_ata_f64_32x32_mfma_r64_p2_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r128_p1_nt", 0, 32, 32, 0, 0, 128, 1, 1, 65536, launch_tsm_ata_f64_32x32_mfma_r128_p1_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r128_p2_nt", 0, 32, 32, 0, 0, 128, 2, 1, 65536, launch_tsm_ata_f64_32x32_mfma_r128_p2_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r256_p1_nt", 0, 32, 32, 0, 0, 256, 1, 1, 131072, launch_tsm_ata_f64_32x32_mfma_r256_p1_nt, 2},
  {"tsm_ata_f64_32x32_mfma_r256_p2_nt", 0, 32, 32, 0, 0, 256, 2, 1, 131072, launch_tsm_ata_f64_32x32_mfma_r256_p2_nt, 2},
  {"tsm_ata_f64_64x64_direct_d4_nt", 0, 64, 64, 0, 0, 4, 4, 1, 65536, launch_tsm_ata_f64_64x64_direct_d4_nt, 3},
  {"tsm_ata_f64_64x64_direct_d8_nt", 0, 64, 64, 0, 0, 4, 8, 1, 65536, launch_tsm_ata_f64_64x64_direct_d8_nt, 3},
  {"tsm_ata_f64_64x64_direct_d16_nt", 0, 64, 64, 0, 0, 4, 16, 1, 65536, launch_tsm_ata_f64_64x64_direct_d16_nt, 3},
  {"tsm_ata_f64_64x64_mfma_r32_p2_nt", 0, 64, 64, 0, 0, 32, 2, 1, 32768, launch_tsm_ata_f64_64x64_mfma_r32_p2_nt, 2},
};

static const tsm::TunedEntry kSymTuned[] = {
  {2, 1, 1, 0, 1, 1, 0, 1},  // tsm_ata_bf16_1x1_t1x1_r4096_p2_nt
  {2, 2, 2, 1, 1, 0, 1, 0},  // tsm_ata_bf16_2x2_t2x2_r4096_p2_nt
  {2, 4, 4, 2, 1, 1, 2, 1},  // tsm_ata_bf16_4x4_t4x4_r4096_p2_nt
  {2, 8, 8, 6, 6, 1, 6, 1},  // tsm_ata_bf16_8x8_t8x4_r1024_p2_nt
  {2, 16, 16, 10, 1, 1, 10, 1},  // tsm_ata_bf16_16x16_mfma_r512_p2_nt
  {2, 32, 32, 16, 1, 1, 16, 1},  // tsm_ata_bf16_32x32_mfma_r256_p2_nt
  {2, 64, 64, 22, 1, 0, 23, 1},  // tsm_ata_bf16_64x64_direct_d4_nt (else tsm_ata_bf16_64x64_mfma_r256_p2_nt)
  {1, 1, 1, 24, 1, 1, 24, 1},  // tsm_ata_f32_1x1_t1x1_r4096_p2_nt
  {1, 2, 2, 25, 1, 1, 25, 1},  // tsm_ata_f32_2x2_t2x2_r4096_p2_nt
  {1, 4, 4, 26, 1, 1, 26, 1},  // tsm_ata_f32_4x4_t4x4_r2048_p2_nt
  {1, 8, 8, 29, 2, 0, 29, 0},  // tsm_ata_f32_8x8_t8x4_r1024_p1_nt
  {1, 16, 16, 38, 1, 0, 38, 0},  // tsm_ata_f32_16x16_mfma_r256_p2_nt
  {1, 32, 32, 42, 3, 0, 42, 0},  // tsm_ata_f32_32x32_mfma_r128_p2_nt
  {1, 64, 64, 47, 2, 0, 48, 1},  // tsm_ata_f32_64x64_direct_d16_nt (else tsm_ata_f32_64x64_mfma_r64_p2_nt)
  {0, 1, 1, 49, 2, 1, 49, 1},  // tsm_ata_f64_1x1_t1x1_r2048_p1_nt
  {0, 2, 2, 54, 1, 1, 54, 1},  // tsm_ata_f64_2x2_t2x2_r1024_p2_nt
  {0, 4, 4, 62, 1, 1, 62, 1},  // tsm_ata_f64_4x4_t4x4_r512_p2_nt
  {0, 8, 8, 68, 1, 0, 68, 0},  // tsm_ata_f64_8x8_t8x8_r512_p2_nt
  {0, 16, 16, 78, 1, 0, 78, 0},  // tsm_ata_f64_16x16_mfma_r256_p2_nt
  {0, 32, 32, 81, 2, 0, 81, 0},  // tsm_ata_f64_32x32_mfma_r128_p1_nt
  {0, 64, 64, 87, 2, 0, 88, 0},  // tsm_ata_f64_64x64_direct_d16_nt (else tsm_ata_f64_64x64_mfma_r32_p2_nt)
};

static const tsm::TunedEntry kTuned[] = {
  {2, 1, 1, 0, 1, 1, 0, 1},  // tsm_atb_bf16_1x1_t1x1_r4096_p2_nt
  {2, 1, 2, 1, 1, 1, 1, 1},  // tsm_atb_bf16_1x2_t1x2_r4096_p2_nt
  {2, 1, 4, 2, 1, 1, 2, 1},  // tsm_atb_bf16_1x4_t1x4_r2048_p2_nt
  {2, 1, 8, 3, 2, 1, 3, 1},  // tsm_atb_bf16_1x8_t1x4_r1024_p2_nt
  {2, 1, 16, 4, 1, 1, 4, 1},  // tsm_atb_bf16_1x16_t1x4_r512_p2_nt
  {2, 1, 32, 5, 1, 1, 5, 1},  // tsm_atb_bf16_1x32_t1x4_r256_p2_nt
  {2, 1, 64, 6, 1, 1, 6, 1},  // tsm_atb_bf16_1x64_t1x4_r128_p2_nt
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
  {2, 8, 8, 33, 2, 1, 33, 1},  // tsm_atb_bf16_8x8_t8x4_r1024_p1_nt
  {2, 8, 16, 46, 2, 0, 46, 0},  // tsm_atb_bf16_8x16_t8x4_r512_p2_nt
  {2, 8, 32, 47, 2, 0, 47, 0},  // tsm_atb_bf16_8x32_t8x4_r256_p2_nt
  {2, 8, 64, 48, 2, 0, 48, 0},  // tsm_atb_bf16_8x64_t8x4_r128_p2_nt
  {2, 16, 1, 49, 2, 1, 49, 1},  // tsm_atb_bf16_16x1_t8x1_r512_p2_nt
  {2, 16, 2, 50, 2, 1, 50, 1},  // tsm_atb_bf16_16x2_t8x2_r512_p2_nt
  {2, 16, 4, 51, 1, 0, 51, 0},  // tsm_atb_bf16_16x4_t8x4_r512_p2_nt
  {2, 16, 8, 52, 2, 1, 52, 1},  // tsm_atb_bf16_16x8_t8x4_r512_p2_nt
  {2, 16, 16, 64, 1, 1, 64, 1},  // tsm_atb_bf16_16x16_mfma_r1024_p1_nt
  {2, 16, 32, 90, 1, 1, 90, 1},  // tsm_atb_bf16_16x32_mfma_r256_p2_nt
  {2, 16, 64, 92, 1, 1, 92, 1},  // tsm_atb_bf16_16x64_mfma_r128_p2_nt
  {2, 32, 1, 93, 2, 1, 93, 1},  // tsm_atb_bf16_32x1_t8x1_r256_p2_nt
  {2, 32, 2, 94, 2, 1, 94, 1},  // tsm_atb_bf16_32x2_t8x2_r256_p2_nt
  {2, 32, 4, 95, 2, 1, 95, 1},  // tsm_atb_bf16_32x4_t8x4_r256_p2_nt
  {2, 32, 8, 96, 2, 0, 96, 0},  // tsm_atb_bf16_32x8_t8x4_r256_p2_nt
  {2, 32, 16, 98, 1, 1, 98, 1},  // tsm_atb_bf16_32x16_mfma_r256_p2_nt
  {2, 32, 32, 110, 1, 1, 110, 1},  // tsm_atb_bf16_32x32_mfma_r512_p1_nt
  {2, 32, 64, 136, 1, 1, 136, 1},  // tsm_atb_bf16_32x64_mfma_r128_p2_nt
  {2, 64, 1, 137, 2, 1, 137, 1},  // tsm_atb_bf16_64x1_t8x1_r128_p2_nt
  {2, 64, 2, 138, 2, 1, 138, 1},  // tsm_atb_bf16_64x2_t8x2_r128_p2_nt
  {2, 64, 4, 139, 2, 1, 139, 1},  // tsm_atb_bf16_64x4_t8x4_r128_p2_nt
  {2, 64, 8, 140, 2, 1, 140, 1},  // tsm_atb_bf16_64x8_t8x4_r128_p2_nt
  {2, 64, 16, 142, 1, 1, 142, 1},  // tsm_atb_bf16_64x16_mfma_r128_p2_nt
  {2, 64, 32, 144, 2, 1, 144, 1},  // tsm_atb_bf16_64x32_mfma_r128_p2_nt
  {2, 64, 64, 146, 1, 0, 154, 1},  // tsm_atb_bf16_64x64_direct_d4_nt (else tsm_atb_bf16_64x64_mfma_r256_p2_nt)
  {1, 1, 1, 167, 1, 1, 167, 1},  // tsm_atb_f32_1x1_t1x1_r4096_p2_nt
  {1, 1, 2, 168, 1, 1, 168, 1},  // tsm_atb_f32_1x2_t1x2_r2048_p2_nt
  {1, 1, 4, 169, 1, 1, 169, 1},  // tsm_atb_f32_1x4_t1x4_r1024_p2_nt
  {1, 1, 8, 170, 1, 1, 170, 1},  // tsm_atb_f32_1x8_t1x4_r512_p2_nt
  {1, 1, 16, 171, 1, 1, 171, 1},  // tsm_atb_f32_1x16_t1x4_r256_p2_nt
  {1, 1, 32, 172, 1, 1, 172, 1},  // tsm_atb_f32_1x32_t1x4_r128_p2_nt
  {1, 1, 64, 173, 1, 1, 173, 1},  // tsm_atb_f32_1x64_t1x4_r64_p2_nt
  {1, 2, 1, 174, 1, 1, 174, 1},  // tsm_atb_f32_2x1_t2x1_r2048_p2_nt
  {1, 2, 2, 175, 1, 1, 175, 1},  // tsm_atb_f32_2x2_t2x2_r2048_p2_nt
  {1, 2, 4, 176, 1, 1, 176, 1},  // tsm_atb_f32_2x4_t2x4_r1024_p2_nt
  {1, 2, 8, 177, 1, 1, 177, 1},  // tsm_atb_f32_2x8_t2x4_r512_p2_nt
  {1, 2, 16, 178, 1, 1, 178, 1},  // tsm_atb_f32_2x16_t2x4_r256_p2_nt
  {1, 2, 32, 179, 1, 1, 179, 1},  // tsm_atb_f32_2x32_t2x4_r128_p2_nt
  {1, 2, 64, 180, 1, 1, 180, 1},  // tsm_atb_f32_2x64_t2x4_r64_p2_nt
  {1, 4, 1, 181, 1, 1, 181, 1},  // tsm_atb_f32_4x1_t4x1_r1024_p2_nt
  {1, 4, 2, 182, 1, 1, 182, 1},  // tsm_atb_f32_4x2_t4x2_r1024_p2_nt
  {1, 4, 4, 183, 1, 1, 183, 1},  // tsm_atb_f32_4x4_t4x4_r1024_p2_nt
  {1, 4, 8, 184, 1, 1, 184, 1},  // tsm_atb_f32_4x8_t4x4_r512_p2_nt
  {1, 4, 16, 185, 1, 1, 185, 1},  // tsm_atb_f32_4x16_t4x4_r256_p2_nt
  {1, 4, 32, 186, 1, 1, 186, 1},  // tsm_atb_f32_4x32_t4x4_r128_p2_nt
  {1, 4, 64, 187, 1, 1, 187, 1},  // tsm_atb_f32_4x64_t4x4_r64_p2_nt
  {1, 8, 1, 188, 2, 1, 188, 1},  // tsm_atb_f32_8x1_t8x1_r512_p2_nt
  {1, 8, 2, 189, 1, 1, 189, 1},  // tsm_atb_f32_8x2_t8x2_r512_p2_nt
  {1, 8, 4, 190, 1, 1, 190, 1},  // tsm_atb_f32_8x4_t8x4_r512_p2_nt
  {1, 8, 8, 208, 1, 1, 208, 1},  // tsm_atb_f32_8x8_t8x4_r1024_p2_nt
  {1, 8, 16, 215, 1, 1, 215, 1},  // tsm_atb_f32_8x16_t8x4_r256_p2_nt
  {1, 8, 32, 216, 1, 1, 216, 1},  // tsm_atb_f32_8x32_t8x4_r128_p2_nt
  {1, 8, 64, 217, 2, 1, 217, 1},  // tsm_atb_f32_8x64_t8x4_r64_p2_nt
  {1, 16, 1, 218, 2, 1, 218, 1},  // tsm_atb_f32_16x1_t8x1_r256_p2_nt
  {1, 16, 2, 219, 2, 1, 219, 1},  // tsm_atb_f32_16x2_t8x2_r256_p2_nt
  {1, 16, 4, 220, 1, 0, 220, 0},  // tsm_atb_f32_16x4_t8x4_r256_p2_nt
  {1, 16, 8, 221, 1, 1, 221, 1},  // tsm_atb_f32_16x8_t8x4_r256_p2_nt
  {1, 16, 16, 228, 1, 1, 228, 1},  // tsm_atb_f32_16x16_mfma_r128_p2_nt
  {1, 16, 32, 255, 1, 0, 256, 0},  // tsm_atb_f32_16x32_direct_d8_nt (else tsm_atb_f32_16x32_t8x4_r128_p2_nt)
  {1, 16, 64, 258, 1, 1, 258, 1},  // tsm_atb_f32_16x64_mfma_r64_p2_nt
  {1, 32, 1, 259, 2, 1, 259, 1},  // tsm_atb_f32_32x1_t8x1_r128_p2_nt
  {1, 32, 2, 260, 2, 1, 260, 1},  // tsm_atb_f32_32x2_t8x2_r128_p2_nt
  {1, 32, 4, 261, 2, 1, 261, 1},  // tsm_atb_f32_32x4_t8x4_r128_p2_nt
  {1, 32, 8, 262, 1, 0, 262, 0},  // tsm_atb_f32_32x8_t8x4_r128_p2_nt
  {1, 32, 16, 263, 1, 0, 264, 0},  // tsm_atb_f32_32x16_direct_d8_nt (else tsm_atb_f32_32x16_t8x4_r128_p2_nt)
  {1, 32, 32, 275, 1, 1, 275, 1},  // tsm_atb_f32_32x32_mfma_r128_p2_nt
  {1, 32, 64, 303, 2, 1, 303, 1},  // tsm_atb_f32_32x64_mfma_r64_p2_nt
  {1, 64, 1, 304, 2, 1, 304, 1},  // tsm_atb_f32_64x1_t8x1_r64_p2_nt
  {1, 64, 2, 305, 2, 1, 305, 1},  // tsm_atb_f32_64x2_t8x2_r64_p2_nt
  {1, 64, 4, 306, 2, 1, 306, 1},  // tsm_atb_f32_64x4_t8x4_r64_p2_nt
  {1, 64, 8, 307, 2, 1, 307, 1},  // tsm_atb_f32_64x8_t8x4_r64_p2_nt
  {1, 64, 16, 309, 1, 1, 309, 1},  // tsm_atb_f32_64x16_mfma_r64_p2_nt
  {1, 64, 32, 310, 1, 0, 311, 1},  // tsm_atb_f32_64x32_direct_d8_nt (else tsm_atb_f32_64x32_mfma_r64_p2_nt)
  {1, 64, 64, 314, 1, 0, 322, 1},  // tsm_atb_f32_64x64_direct_d16_nt (else tsm_atb_f32_64x64_mfma_r64_p2_nt)
  {0, 1, 1, 348, 1, 1, 348, 1},  // tsm_atb_f64_1x1_t1x1_r4096_p1_nt
  {0, 1, 2, 351, 1, 1, 351, 1},  // tsm_atb_f64_1x2_t1x2_r1024_p2_nt
  {0, 1, 4, 352, 1, 1, 352, 1},  // tsm_atb_f64_1x4_t1x4_r512_p2_nt
  {0, 1, 8, 353, 1, 1, 353, 1},  // tsm_atb_f64_1x8_t1x8_r256_p2_nt
  {0, 1, 16, 354, 1, 1, 354, 1},  // tsm_atb_f64_1x16_t1x8_r128_p2_nt
  {0, 1, 32, 355, 1, 1, 355, 1},  // tsm_atb_f64_1x32_t1x8_r64_p2_nt
  {0, 1, 64, 356, 1, 1, 356, 1},  // tsm_atb_f64_1x64_t1x8_r32_p2_nt
  {0, 2, 1, 357, 1, 1, 357, 1},  // tsm_atb_f64_2x1_t2x1_r1024_p2_nt
  {0, 2, 2, 367, 1, 1, 367, 1},  // tsm_atb_f64_2x2_t2x2_r2048_p1_nt
  {0, 2, 4, 370, 1, 1, 370, 1},  // tsm_atb_f64_2x4_t2x4_r512_p2_nt
  {0, 2, 8, 371, 1, 1, 371, 1},  // tsm_atb_f64_2x8_t2x8_r256_p2_nt
  {0, 2, 16, 372, 1, 1, 372, 1},  // tsm_atb_f64_2x16_t2x8_r128_p2_nt
  {0, 2, 32, 373, 1, 1, 373, 1},  // tsm_atb_f64_2x32_t2x8_r64_p2_nt
  {0, 2, 64, 374, 1, 1, 374, 1},  // tsm_atb_f64_2x64_t2x8_r32_p2_nt
  {0, 4, 1, 375, 1, 1, 375, 1},  // tsm_atb_f64_4x1_t4x1_r512_p2_nt
  {0, 4, 2, 376, 1, 1, 376, 1},  // tsm_atb_f64_4x2_t4x2_r512_p2_nt
  {0, 4, 4, 380, 2, 1, 380, 1},  // tsm_atb_f64_4x4_t4x4_r256_p1_nt
  {0, 4, 8, 391, 1, 1, 391, 1},  // tsm_atb_f64_4x8_t4x8_r256_p2_nt
  {0, 4, 16, 392, 1, 1, 392, 1},  // tsm_atb_f64_4x16_t4x8_r128_p2_nt
  {0, 4, 32, 400, 2, 1, 400, 1},  // tsm_atb_f64_4x32_t4x8_r64_p1_nt
  {0, 4, 64, 407, 1, 1, 407, 1},  // tsm_atb_f64_4x64_t4x8_r32_p2_nt
  {0, 8, 1, 408, 2, 1, 408, 1},  // tsm_atb_f64_8x1_t8x1_r256_p2_nt
  {0, 8, 2, 409, 1, 1, 409, 1},  // tsm_atb_f64_8x2_t8x2_r256_p2_nt
  {0, 8, 4, 410, 1, 1, 410, 1},  // tsm_atb_f64_8x4_t8x4_r256_p2_nt
  {0, 8, 8, 422, 1, 1, 422, 1},  // tsm_atb_f64_8x8_t8x8_r256_p2_nt
  {0, 8, 16, 427, 1, 1, 427, 1},  // tsm_atb_f64_8x16_t8x8_r128_p2_nt
  {0, 8, 32, 428, 1, 1, 428, 1},  // tsm_atb_f64_8x32_t8x8_r64_p2_nt
  {0, 8, 64, 429, 2, 1, 429, 1},  // tsm_atb_f64_8x64_t8x8_r32_p2_nt
  {0, 16, 1, 430, 2, 1, 430, 1},  // tsm_atb_f64_16x1_t8x1_r128_p2_nt
  {0, 16, 2, 431, 2, 1, 431, 1},  // tsm_atb_f64_16x2_t8x2_r128_p2_nt
  {0, 16, 4, 432, 1, 1, 432, 1},  // tsm_atb_f64_16x4_t8x4_r128_p2_nt
  {0, 16, 8, 433, 1, 1, 433, 1},  // tsm_atb_f64_16x8_t8x8_r128_p2_nt
  {0, 16, 16, 440, 1, 1, 440, 1},  // tsm_atb_f64_16x16_mfma_r64_p2_nt
  {0, 16, 32, 469, 1, 0, 470, 0},  // tsm_atb_f64_16x32_direct_d8_nt (else tsm_atb_f64_16x32_t8x8_r64_p2_nt)
  {0, 16, 64, 472, 1, 1, 472, 1},  // tsm_atb_f64_16x64_mfma_r32_p2_nt
  {0, 32, 1, 473, 2, 1, 473, 1},  // tsm_atb_f64_32x1_t8x1_r64_p2_nt
  {0, 32, 2, 474, 2, 1, 474, 1},  // tsm_atb_f64_32x2_t8x2_r64_p2_nt
  {0, 32, 4, 482, 2, 1, 482, 1},  // tsm_atb_f64_32x4_t8x4_r64_p1_nt
  {0, 32, 8, 489, 1, 1, 489, 1},  // tsm_atb_f64_32x8_t8x8_r64_p2_nt
  {0, 32, 16, 490, 1, 0, 491, 0},  // tsm_atb_f64_32x16_direct_d8_nt (else tsm_atb_f64_32x16_t8x8_r64_p2_nt)
  {0, 32, 32, 506, 1, 0, 506, 0},  // tsm_atb_f64_32x32_mfma_r128_p2_nt
  {0, 32, 64, 524, 2, 0, 524, 0},  // tsm_atb_f64_32x64_mfma_r32_p2_nt
  {0, 64, 1, 525, 2, 1, 525, 1},  // tsm_atb_f64_64x1_t8x1_r32_p2_nt
  {0, 64, 2, 526, 2, 1, 526, 1},  // tsm_atb_f64_64x2_t8x2_r32_p2_nt
  {0, 64, 4, 527, 2, 1, 527, 1},  // tsm_atb_f64_64x4_t8x4_r32_p2_nt
  {0, 64, 8, 528, 2, 1, 528, 1},  // tsm_atb_f64_64x8_t8x8_r32_p2_nt
  {0, 64, 16, 530, 1, 1, 530, 1},  // tsm_atb_f64_64x16_mfma_r32_p2_nt
  {0, 64, 32, 532, 1, 0, 532, 0},  // tsm_atb_f64_64x32_mfma_r32_p2_nt
  {0, 64, 64, 535, 2, 0, 543, 0},  // tsm_atb_f64_64x64_direct_d16_nt (else tsm_atb_f64_64x64_mfma_r32_p2_nt)
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
