// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

hipError_t launch_tsm_atb_bf16_1x1_t1x1_r4096_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x2_t1x2_r4096_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x4_t1x4_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x8_t1x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x16_t1x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x32_t1x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_1x64_t1x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x1_t2x1_r4096_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x2_t2x2_r4096_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x4_t2x4_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x8_t2x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x16_t2x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x32_t2x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_2x64_t2x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x1_t4x1_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x2_t4x2_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x4_t4x4_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x8_t4x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x16_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x32_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_4x64_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x1_t8x1_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x2_t8x2_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x4_t8x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_bf16_8x16_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x64_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x1_t8x1_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x2_t8x2_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x32_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x64_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x1_t8x1_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x2_t8x2_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x64_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x1_t8x1_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x2_t8x2_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x4_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x8_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x1_t1x1_r4096_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x2_t1x2_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x4_t1x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x8_t1x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x16_t1x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x32_t1x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_1x64_t1x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x1_t2x1_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x2_t2x2_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x4_t2x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x8_t2x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x16_t2x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x32_t2x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_2x64_t2x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x1_t4x1_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x2_t4x2_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x4_t4x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x8_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x32_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_4x64_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x1_t8x1_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x2_t8x2_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x4_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_8x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x64_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x1_t8x1_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x2_t8x2_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x4_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x8_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_16x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x64_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x1_t8x1_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x2_t8x2_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p1_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p2_nt(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x64_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x1_t8x1_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x2_t8x2_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x16_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_1x2_t1x2_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x4_t1x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x8_t1x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x16_t1x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x32_t1x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x64_t1x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x1_t2x1_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_2x4_t2x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x8_t2x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x16_t2x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x32_t2x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x64_t2x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x1_t4x1_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x2_t4x2_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_4x8_t4x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_4x64_t4x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x1_t8x1_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x2_t8x2_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x4_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_8x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x1_t8x1_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_t8x2_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_16x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x1_t8x1_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_t8x2_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_32x8_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x16_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_32x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x1_t8x1_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_t8x2_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_t8x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x16_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x32_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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

static const tsm::KernelEntry kTable[] = {
  {"tsm_atb_bf16_1x1_t1x1_r4096_p2", 2, 1, 1, 1, 1, 4096, 2, 0, 32768, launch_tsm_atb_bf16_1x1_t1x1_r4096_p2},
  {"tsm_atb_bf16_1x2_t1x2_r4096_p2", 2, 1, 2, 1, 2, 4096, 2, 0, 49152, launch_tsm_atb_bf16_1x2_t1x2_r4096_p2},
  {"tsm_atb_bf16_1x4_t1x4_r2048_p2", 2, 1, 4, 1, 4, 2048, 2, 0, 40960, launch_tsm_atb_bf16_1x4_t1x4_r2048_p2},
  {"tsm_atb_bf16_1x8_t1x4_r1024_p2", 2, 1, 8, 1, 4, 1024, 2, 0, 36864, launch_tsm_atb_bf16_1x8_t1x4_r1024_p2},
  {"tsm_atb_bf16_1x16_t1x4_r512_p2", 2, 1, 16, 1, 4, 512, 2, 0, 34816, launch_tsm_atb_bf16_1x16_t1x4_r512_p2},
  {"tsm_atb_bf16_1x32_t1x4_r256_p2", 2, 1, 32, 1, 4, 256, 2, 0, 33792, launch_tsm_atb_bf16_1x32_t1x4_r256_p2},
  {"tsm_atb_bf16_1x64_t1x4_r128_p2", 2, 1, 64, 1, 4, 128, 2, 0, 33280, launch_tsm_atb_bf16_1x64_t1x4_r128_p2},
  {"tsm_atb_bf16_2x1_t2x1_r4096_p2", 2, 2, 1, 2, 1, 4096, 2, 0, 49152, launch_tsm_atb_bf16_2x1_t2x1_r4096_p2},
  {"tsm_atb_bf16_2x2_t2x2_r4096_p2", 2, 2, 2, 2, 2, 4096, 2, 0, 65536, launch_tsm_atb_bf16_2x2_t2x2_r4096_p2},
  {"tsm_atb_bf16_2x4_t2x4_r2048_p2", 2, 2, 4, 2, 4, 2048, 2, 0, 49152, launch_tsm_atb_bf16_2x4_t2x4_r2048_p2},
  {"tsm_atb_bf16_2x8_t2x4_r1024_p2", 2, 2, 8, 2, 4, 1024, 2, 0, 40960, launch_tsm_atb_bf16_2x8_t2x4_r1024_p2},
  {"tsm_atb_bf16_2x16_t2x4_r512_p2", 2, 2, 16, 2, 4, 512, 2, 0, 36864, launch_tsm_atb_bf16_2x16_t2x4_r512_p2},
  {"tsm_atb_bf16_2x32_t2x4_r256_p2", 2, 2, 32, 2, 4, 256, 2, 0, 34816, launch_tsm_atb_bf16_2x32_t2x4_r256_p2},
  {"tsm_atb_bf16_2x64_t2x4_r128_p2", 2, 2, 64, 2, 4, 128, 2, 0, 33792, launch_tsm_atb_bf16_2x64_t2x4_r128_p2},
  {"tsm_atb_bf16_4x1_t4x1_r2048_p2", 2, 4, 1, 4, 1, 2048, 2, 0, 40960, launch_tsm_atb_bf16_4x1_t4x1_r2048_p2},
  {"tsm_atb_bf16_4x2_t4x2_r2048_p2", 2, 4, 2, 4, 2, 2048, 2, 0, 49152, launch_tsm_atb_bf16_4x2_t4x2_r2048_p2},
  {"tsm_atb_bf16_4x4_t4x4_r2048_p2", 2, 4, 4, 4, 4, 2048, 2, 0, 65536, launch_tsm_atb_bf16_4x4_t4x4_r2048_p2},
  {"tsm_atb_bf16_4x8_t4x4_r1024_p2", 2, 4, 8, 4, 4, 1024, 2, 0, 49152, launch_tsm_atb_bf16_4x8_t4x4_r1024_p2},
  {"tsm_atb_bf16_4x16_t4x4_r512_p2", 2, 4, 16, 4, 4, 512, 2, 0, 40960, launch_tsm_atb_bf16_4x16_t4x4_r512_p2},
  {"tsm_atb_bf16_4x32_t4x4_r256_p2", 2, 4, 32, 4, 4, 256, 2, 0, 36864, launch_tsm_atb_bf16_4x32_t4x4_r256_p2},
  {"tsm_atb_bf16_4x64_t4x4_r128_p2", 2, 4, 64, 4, 4, 128, 2, 0, 34816, launch_tsm_atb_bf16_4x64_t4x4_r128_p2},
  {"tsm_atb_bf16_8x1_t8x1_r1024_p2", 2, 8, 1, 8, 1, 1024, 2, 0, 36864, launch_tsm_atb_bf16_8x1_t8x1_r1024_p2},
  {"tsm_atb_bf16_8x2_t8x2_r1024_p2", 2, 8, 2, 8, 2, 1024, 2, 0, 40960, launch_tsm_atb_bf16_8x2_t8x2_r1024_p2},
  {"tsm_atb_bf16_8x4_t8x4_r1024_p2", 2, 8, 4, 8, 4, 1024, 2, 0, 49152, launch_tsm_atb_bf16_8x4_t8x4_r1024_p2},
  {"tsm_atb_bf16_8x8_t4x4_r512_p1", 2, 8, 8, 4, 4, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t4x4_r512_p1},
  {"tsm_atb_bf16_8x8_t4x4_r512_p2", 2, 8, 8, 4, 4, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t4x4_r512_p2},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p1", 2, 8, 8, 4, 4, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p1},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p2", 2, 8, 8, 4, 4, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p2},
  {"tsm_atb_bf16_8x8_t8x4_r512_p1", 2, 8, 8, 8, 4, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p1_nt", 2, 8, 8, 8, 4, 512, 1, 1, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p1_nt},
  {"tsm_atb_bf16_8x8_t8x4_r512_p2", 2, 8, 8, 8, 4, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p2},
  {"tsm_atb_bf16_8x8_t8x4_r512_p2_nt", 2, 8, 8, 8, 4, 512, 2, 1, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p2_nt},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p1", 2, 8, 8, 8, 4, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p1_nt", 2, 8, 8, 8, 4, 1024, 1, 1, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p1_nt},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p2", 2, 8, 8, 8, 4, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p2},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p2_nt", 2, 8, 8, 8, 4, 1024, 2, 1, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p2_nt},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p1", 2, 8, 8, 8, 4, 2048, 1, 0, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p1},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p1_nt", 2, 8, 8, 8, 4, 2048, 1, 1, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p1_nt},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p2", 2, 8, 8, 8, 4, 2048, 2, 0, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p2},
  {"tsm_atb_bf16_8x8_t8x4_r2048_p2_nt", 2, 8, 8, 8, 4, 2048, 2, 1, 131072, launch_tsm_atb_bf16_8x8_t8x4_r2048_p2_nt},
  {"tsm_atb_bf16_8x8_t8x8_r512_p1", 2, 8, 8, 8, 8, 512, 1, 0, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p1},
  {"tsm_atb_bf16_8x8_t8x8_r512_p2", 2, 8, 8, 8, 8, 512, 2, 0, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p2},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p1", 2, 8, 8, 8, 8, 1024, 1, 0, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p1},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p2", 2, 8, 8, 8, 8, 1024, 2, 0, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p2},
  {"tsm_atb_bf16_8x8_t8x8_r2048_p1", 2, 8, 8, 8, 8, 2048, 1, 0, 131072, launch_tsm_atb_bf16_8x8_t8x8_r2048_p1},
  {"tsm_atb_bf16_8x8_t8x8_r2048_p2", 2, 8, 8, 8, 8, 2048, 2, 0, 131072, launch_tsm_atb_bf16_8x8_t8x8_r2048_p2},
  {"tsm_atb_bf16_8x16_t8x4_r512_p2", 2, 8, 16, 8, 4, 512, 2, 0, 49152, launch_tsm_atb_bf16_8x16_t8x4_r512_p2},
  {"tsm_atb_bf16_8x32_t8x4_r256_p2", 2, 8, 32, 8, 4, 256, 2, 0, 40960, launch_tsm_atb_bf16_8x32_t8x4_r256_p2},
  {"tsm_atb_bf16_8x64_t8x4_r128_p2", 2, 8, 64, 8, 4, 128, 2, 0, 36864, launch_tsm_atb_bf16_8x64_t8x4_r128_p2},
  {"tsm_atb_bf16_16x1_t8x1_r512_p2", 2, 16, 1, 8, 1, 512, 2, 0, 34816, launch_tsm_atb_bf16_16x1_t8x1_r512_p2},
  {"tsm_atb_bf16_16x2_t8x2_r512_p2", 2, 16, 2, 8, 2, 512, 2, 0, 36864, launch_tsm_atb_bf16_16x2_t8x2_r512_p2},
  {"tsm_atb_bf16_16x4_t8x4_r512_p2", 2, 16, 4, 8, 4, 512, 2, 0, 40960, launch_tsm_atb_bf16_16x4_t8x4_r512_p2},
  {"tsm_atb_bf16_16x8_t8x4_r512_p2", 2, 16, 8, 8, 4, 512, 2, 0, 49152, launch_tsm_atb_bf16_16x8_t8x4_r512_p2},
  {"tsm_atb_bf16_16x16_t4x4_r256_p1", 2, 16, 16, 4, 4, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p1},
  {"tsm_atb_bf16_16x16_t4x4_r256_p2", 2, 16, 16, 4, 4, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p2},
  {"tsm_atb_bf16_16x16_t8x4_r256_p1", 2, 16, 16, 8, 4, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p1_nt", 2, 16, 16, 8, 4, 256, 1, 1, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p1_nt},
  {"tsm_atb_bf16_16x16_t8x4_r256_p2", 2, 16, 16, 8, 4, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p2},
  {"tsm_atb_bf16_16x16_t8x4_r256_p2_nt", 2, 16, 16, 8, 4, 256, 2, 1, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p2_nt},
  {"tsm_atb_bf16_16x16_t8x4_r512_p1", 2, 16, 16, 8, 4, 512, 1, 0, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p1_nt", 2, 16, 16, 8, 4, 512, 1, 1, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p1_nt},
  {"tsm_atb_bf16_16x16_t8x4_r512_p2", 2, 16, 16, 8, 4, 512, 2, 0, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p2},
  {"tsm_atb_bf16_16x16_t8x4_r512_p2_nt", 2, 16, 16, 8, 4, 512, 2, 1, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p2_nt},
  {"tsm_atb_bf16_16x16_t8x8_r256_p1", 2, 16, 16, 8, 8, 256, 1, 0, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p1},
  {"tsm_atb_bf16_16x16_t8x8_r256_p2", 2, 16, 16, 8, 8, 256, 2, 0, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p2},
  {"tsm_atb_bf16_16x16_t8x8_r512_p1", 2, 16, 16, 8, 8, 512, 1, 0, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p1},
  {"tsm_atb_bf16_16x16_t8x8_r512_p2", 2, 16, 16, 8, 8, 512, 2, 0, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p2},
  {"tsm_atb_bf16_16x16_t8x8_r1024_p1", 2, 16, 16, 8, 8, 1024, 1, 0, 131072, launch_tsm_atb_bf16_16x16_t8x8_r1024_p1},
  {"tsm_atb_bf16_16x16_t8x8_r1024_p2", 2, 16, 16, 8, 8, 1024, 2, 0, 131072, launch_tsm_atb_bf16_16x16_t8x8_r1024_p2},
  {"tsm_atb_bf16_16x32_t8x4_r256_p2", 2, 16, 32, 8, 4, 256, 2, 0, 49152, launch_tsm_atb_bf16_16x32_t8x4_r256_p2},
  {"tsm_atb_bf16_16x64_t8x4_r128_p2", 2, 16, 64, 8, 4, 128, 2, 0, 40960, launch_tsm_atb_bf16_16x64_t8x4_r128_p2},
  {"tsm_atb_bf16_32x1_t8x1_r256_p2", 2, 32, 1, 8, 1, 256, 2, 0, 33792, launch_tsm_atb_bf16_32x1_t8x1_r256_p2},
  {"tsm_atb_bf16_32x2_t8x2_r256_p2", 2, 32, 2, 8, 2, 256, 2, 0, 34816, launch_tsm_atb_bf16_32x2_t8x2_r256_p2},
  {"tsm_atb_bf16_32x4_t8x4_r256_p2", 2, 32, 4, 8, 4, 256, 2, 0, 36864, launch_tsm_atb_bf16_32x4_t8x4_r256_p2},
  {"tsm_atb_bf16_32x8_t8x4_r256_p2", 2, 32, 8, 8, 4, 256, 2, 0, 40960, launch_tsm_atb_bf16_32x8_t8x4_r256_p2},
  {"tsm_atb_bf16_32x16_t8x4_r256_p2", 2, 32, 16, 8, 4, 256, 2, 0, 49152, launch_tsm_atb_bf16_32x16_t8x4_r256_p2},
  {"tsm_atb_bf16_32x32_t4x4_r64_p1", 2, 32, 32, 4, 4, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t4x4_r64_p1},
  {"tsm_atb_bf16_32x32_t4x4_r64_p2", 2, 32, 32, 4, 4, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t4x4_r64_p2},
  {"tsm_atb_bf16_32x32_t8x4_r64_p1", 2, 32, 32, 8, 4, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p1},
  {"tsm_atb_bf16_32x32_t8x4_r64_p1_nt", 2, 32, 32, 8, 4, 64, 1, 1, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p1_nt},
  {"tsm_atb_bf16_32x32_t8x4_r64_p2", 2, 32, 32, 8, 4, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p2},
  {"tsm_atb_bf16_32x32_t8x4_r64_p2_nt", 2, 32, 32, 8, 4, 64, 2, 1, 16384, launch_tsm_atb_bf16_32x32_t8x4_r64_p2_nt},
  {"tsm_atb_bf16_32x32_t8x4_r128_p1", 2, 32, 32, 8, 4, 128, 1, 0, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p1_nt", 2, 32, 32, 8, 4, 128, 1, 1, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p1_nt},
  {"tsm_atb_bf16_32x32_t8x4_r128_p2", 2, 32, 32, 8, 4, 128, 2, 0, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p2},
  {"tsm_atb_bf16_32x32_t8x4_r128_p2_nt", 2, 32, 32, 8, 4, 128, 2, 1, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p2_nt},
  {"tsm_atb_bf16_32x32_t8x8_r64_p1", 2, 32, 32, 8, 8, 64, 1, 0, 16384, launch_tsm_atb_bf16_32x32_t8x8_r64_p1},
  {"tsm_atb_bf16_32x32_t8x8_r64_p2", 2, 32, 32, 8, 8, 64, 2, 0, 16384, launch_tsm_atb_bf16_32x32_t8x8_r64_p2},
  {"tsm_atb_bf16_32x32_t8x8_r128_p1", 2, 32, 32, 8, 8, 128, 1, 0, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p1},
  {"tsm_atb_bf16_32x32_t8x8_r128_p2", 2, 32, 32, 8, 8, 128, 2, 0, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p2},
  {"tsm_atb_bf16_32x32_t8x8_r256_p1", 2, 32, 32, 8, 8, 256, 1, 0, 65536, launch_tsm_atb_bf16_32x32_t8x8_r256_p1},
  {"tsm_atb_bf16_32x32_t8x8_r256_p2", 2, 32, 32, 8, 8, 256, 2, 0, 65536, launch_tsm_atb_bf16_32x32_t8x8_r256_p2},
  {"tsm_atb_bf16_32x64_t8x4_r64_p2", 2, 32, 64, 8, 4, 64, 2, 0, 24576, launch_tsm_atb_bf16_32x64_t8x4_r64_p2},
  {"tsm_atb_bf16_64x1_t8x1_r128_p2", 2, 64, 1, 8, 1, 128, 2, 0, 33280, launch_tsm_atb_bf16_64x1_t8x1_r128_p2},
  {"tsm_atb_bf16_64x2_t8x2_r128_p2", 2, 64, 2, 8, 2, 128, 2, 0, 33792, launch_tsm_atb_bf16_64x2_t8x2_r128_p2},
  {"tsm_atb_bf16_64x4_t8x4_r128_p2", 2, 64, 4, 8, 4, 128, 2, 0, 34816, launch_tsm_atb_bf16_64x4_t8x4_r128_p2},
  {"tsm_atb_bf16_64x8_t8x4_r128_p2", 2, 64, 8, 8, 4, 128, 2, 0, 36864, launch_tsm_atb_bf16_64x8_t8x4_r128_p2},
  {"tsm_atb_bf16_64x16_t8x4_r128_p2", 2, 64, 16, 8, 4, 128, 2, 0, 40960, launch_tsm_atb_bf16_64x16_t8x4_r128_p2},
  {"tsm_atb_bf16_64x32_t8x4_r64_p2", 2, 64, 32, 8, 4, 64, 2, 0, 24576, launch_tsm_atb_bf16_64x32_t8x4_r64_p2},
  {"tsm_atb_bf16_64x64_t8x8_r64_p2", 2, 64, 64, 8, 8, 64, 2, 0, 32768, launch_tsm_atb_bf16_64x64_t8x8_r64_p2},
  {"tsm_atb_f32_1x1_t1x1_r4096_p2", 1, 1, 1, 1, 1, 4096, 2, 0, 65536, launch_tsm_atb_f32_1x1_t1x1_r4096_p2},
  {"tsm_atb_f32_1x2_t1x2_r2048_p2", 1, 1, 2, 1, 2, 2048, 2, 0, 49152, launch_tsm_atb_f32_1x2_t1x2_r2048_p2},
  {"tsm_atb_f32_1x4_t1x4_r1024_p2", 1, 1, 4, 1, 4, 1024, 2, 0, 40960, launch_tsm_atb_f32_1x4_t1x4_r1024_p2},
  {"tsm_atb_f32_1x8_t1x4_r512_p2", 1, 1, 8, 1, 4, 512, 2, 0, 36864, launch_tsm_atb_f32_1x8_t1x4_r512_p2},
  {"tsm_atb_f32_1x16_t1x4_r256_p2", 1, 1, 16, 1, 4, 256, 2, 0, 34816, launch_tsm_atb_f32_1x16_t1x4_r256_p2},
  {"tsm_atb_f32_1x32_t1x4_r128_p2", 1, 1, 32, 1, 4, 128, 2, 0, 33792, launch_tsm_atb_f32_1x32_t1x4_r128_p2},
  {"tsm_atb_f32_1x64_t1x4_r64_p2", 1, 1, 64, 1, 4, 64, 2, 0, 33280, launch_tsm_atb_f32_1x64_t1x4_r64_p2},
  {"tsm_atb_f32_2x1_t2x1_r2048_p2", 1, 2, 1, 2, 1, 2048, 2, 0, 49152, launch_tsm_atb_f32_2x1_t2x1_r2048_p2},
  {"tsm_atb_f32_2x2_t2x2_r2048_p2", 1, 2, 2, 2, 2, 2048, 2, 0, 65536, launch_tsm_atb_f32_2x2_t2x2_r2048_p2},
  {"tsm_atb_f32_2x4_t2x4_r1024_p2", 1, 2, 4, 2, 4, 1024, 2, 0, 49152, launch_tsm_atb_f32_2x4_t2x4_r1024_p2},
  {"tsm_atb_f32_2x8_t2x4_r512_p2", 1, 2, 8, 2, 4, 512, 2, 0, 40960, launch_tsm_atb_f32_2x8_t2x4_r512_p2},
  {"tsm_atb_f32_2x16_t2x4_r256_p2", 1, 2, 16, 2, 4, 256, 2, 0, 36864, launch_tsm_atb_f32_2x16_t2x4_r256_p2},
  {"tsm_atb_f32_2x32_t2x4_r128_p2", 1, 2, 32, 2, 4, 128, 2, 0, 34816, launch_tsm_atb_f32_2x32_t2x4_r128_p2},
  {"tsm_atb_f32_2x64_t2x4_r64_p2", 1, 2, 64, 2, 4, 64, 2, 0, 33792, launch_tsm_atb_f32_2x64_t2x4_r64_p2},
  {"tsm_atb_f32_4x1_t4x1_r1024_p2", 1, 4, 1, 4, 1, 1024, 2, 0, 40960, launch_tsm_atb_f32_4x1_t4x1_r1024_p2},
  {"tsm_atb_f32_4x2_t4x2_r1024_p2", 1, 4, 2, 4, 2, 1024, 2, 0, 49152, launch_tsm_atb_f32_4x2_t4x2_r1024_p2},
  {"tsm_atb_f32_4x4_t4x4_r1024_p2", 1, 4, 4, 4, 4, 1024, 2, 0, 65536, launch_tsm_atb_f32_4x4_t4x4_r1024_p2},
  {"tsm_atb_f32_4x8_t4x4_r512_p2", 1, 4, 8, 4, 4, 512, 2, 0, 49152, launch_tsm_atb_f32_4x8_t4x4_r512_p2},
  {"tsm_atb_f32_4x16_t4x4_r256_p2", 1, 4, 16, 4, 4, 256, 2, 0, 40960, launch_tsm_atb_f32_4x16_t4x4_r256_p2},
  {"tsm_atb_f32_4x32_t4x4_r128_p2", 1, 4, 32, 4, 4, 128, 2, 0, 36864, launch_tsm_atb_f32_4x32_t4x4_r128_p2},
  {"tsm_atb_f32_4x64_t4x4_r64_p2", 1, 4, 64, 4, 4, 64, 2, 0, 34816, launch_tsm_atb_f32_4x64_t4x4_r64_p2},
  {"tsm_atb_f32_8x1_t8x1_r512_p2", 1, 8, 1, 8, 1, 512, 2, 0, 36864, launch_tsm_atb_f32_8x1_t8x1_r512_p2},
  {"tsm_atb_f32_8x2_t8x2_r512_p2", 1, 8, 2, 8, 2, 512, 2, 0, 40960, launch_tsm_atb_f32_8x2_t8x2_r512_p2},
  {"tsm_atb_f32_8x4_t8x4_r512_p2", 1, 8, 4, 8, 4, 512, 2, 0, 49152, launch_tsm_atb_f32_8x4_t8x4_r512_p2},
  {"tsm_atb_f32_8x8_t4x4_r256_p1", 1, 8, 8, 4, 4, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t4x4_r256_p1},
  {"tsm_atb_f32_8x8_t4x4_r256_p2", 1, 8, 8, 4, 4, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t4x4_r256_p2},
  {"tsm_atb_f32_8x8_t4x4_r512_p1", 1, 8, 8, 4, 4, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p1},
  {"tsm_atb_f32_8x8_t4x4_r512_p2", 1, 8, 8, 4, 4, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p2},
  {"tsm_atb_f32_8x8_t4x4_r1024_p1", 1, 8, 8, 4, 4, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t4x4_r1024_p1},
  {"tsm_atb_f32_8x8_t4x4_r1024_p2", 1, 8, 8, 4, 4, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t4x4_r1024_p2},
  {"tsm_atb_f32_8x8_t8x4_r256_p1", 1, 8, 8, 8, 4, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p1},
  {"tsm_atb_f32_8x8_t8x4_r256_p1_nt", 1, 8, 8, 8, 4, 256, 1, 1, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p1_nt},
  {"tsm_atb_f32_8x8_t8x4_r256_p2", 1, 8, 8, 8, 4, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p2},
  {"tsm_atb_f32_8x8_t8x4_r256_p2_nt", 1, 8, 8, 8, 4, 256, 2, 1, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p2_nt},
  {"tsm_atb_f32_8x8_t8x4_r512_p1", 1, 8, 8, 8, 4, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p1},
  {"tsm_atb_f32_8x8_t8x4_r512_p1_nt", 1, 8, 8, 8, 4, 512, 1, 1, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p1_nt},
  {"tsm_atb_f32_8x8_t8x4_r512_p2", 1, 8, 8, 8, 4, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p2},
  {"tsm_atb_f32_8x8_t8x4_r512_p2_nt", 1, 8, 8, 8, 4, 512, 2, 1, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p2_nt},
  {"tsm_atb_f32_8x8_t8x4_r1024_p1", 1, 8, 8, 8, 4, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p1},
  {"tsm_atb_f32_8x8_t8x4_r1024_p1_nt", 1, 8, 8, 8, 4, 1024, 1, 1, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p1_nt},
  {"tsm_atb_f32_8x8_t8x4_r1024_p2", 1, 8, 8, 8, 4, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p2},
  {"tsm_atb_f32_8x8_t8x4_r1024_p2_nt", 1, 8, 8, 8, 4, 1024, 2, 1, 131072, launch_tsm_atb_f32_8x8_t8x4_r1024_p2_nt},
  {"tsm_atb_f32_8x8_t8x8_r256_p1", 1, 8, 8, 8, 8, 256, 1, 0, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p1},
  {"tsm_atb_f32_8x8_t8x8_r256_p2", 1, 8, 8, 8, 8, 256, 2, 0, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p2},
  {"tsm_atb_f32_8x8_t8x8_r512_p1", 1, 8, 8, 8, 8, 512, 1, 0, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p1},
  {"tsm_atb_f32_8x8_t8x8_r512_p2", 1, 8, 8, 8, 8, 512, 2, 0, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p2},
  {"tsm_atb_f32_8x8_t8x8_r1024_p1", 1, 8, 8, 8, 8, 1024, 1, 0, 131072, launch_tsm_atb_f32_8x8_t8x8_r1024_p1},
  {"tsm_atb_f32_8x8_t8x8_r1024_p2", 1, 8, 8, 8, 8, 1024, 2, 0, 131072, launch_tsm_atb_f32_8x8_t8x8_r1024_p2},
  {"tsm_atb_f32_8x16_t8x4_r256_p2", 1, 8, 16, 8, 4, 256, 2, 0, 49152, launch_tsm_atb_f32_8x16_t8x4_r256_p2},
  {"tsm_atb_f32_8x32_t8x4_r128_p2", 1, 8, 32, 8, 4, 128, 2, 0, 40960, launch_tsm_atb_f32_8x32_t8x4_r128_p2},
  {"tsm_atb_f32_8x64_t8x4_r64_p2", 1, 8, 64, 8, 4, 64, 2, 0, 36864, launch_tsm_atb_f32_8x64_t8x4_r64_p2},
  {"tsm_atb_f32_16x1_t8x1_r256_p2", 1, 16, 1, 8, 1, 256, 2, 0, 34816, launch_tsm_atb_f32_16x1_t8x1_r256_p2},
  {"tsm_atb_f32_16x2_t8x2_r256_p2", 1, 16, 2, 8, 2, 256, 2, 0, 36864, launch_tsm_atb_f32_16x2_t8x2_r256_p2},
  {"tsm_atb_f32_16x4_t8x4_r256_p2", 1, 16, 4, 8, 4, 256, 2, 0, 40960, launch_tsm_atb_f32_16x4_t8x4_r256_p2},
  {"tsm_atb_f32_16x8_t8x4_r256_p2", 1, 16, 8, 8, 4, 256, 2, 0, 49152, launch_tsm_atb_f32_16x8_t8x4_r256_p2},
  {"tsm_atb_f32_16x16_t4x4_r128_p1", 1, 16, 16, 4, 4, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t4x4_r128_p1},
  {"tsm_atb_f32_16x16_t4x4_r128_p2", 1, 16, 16, 4, 4, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t4x4_r128_p2},
  {"tsm_atb_f32_16x16_t4x4_r256_p1", 1, 16, 16, 4, 4, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p1},
  {"tsm_atb_f32_16x16_t4x4_r256_p2", 1, 16, 16, 4, 4, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p2},
  {"tsm_atb_f32_16x16_t8x4_r128_p1", 1, 16, 16, 8, 4, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p1},
  {"tsm_atb_f32_16x16_t8x4_r128_p1_nt", 1, 16, 16, 8, 4, 128, 1, 1, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p1_nt},
  {"tsm_atb_f32_16x16_t8x4_r128_p2", 1, 16, 16, 8, 4, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p2},
  {"tsm_atb_f32_16x16_t8x4_r128_p2_nt", 1, 16, 16, 8, 4, 128, 2, 1, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p2_nt},
  {"tsm_atb_f32_16x16_t8x4_r256_p1", 1, 16, 16, 8, 4, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p1},
  {"tsm_atb_f32_16x16_t8x4_r256_p1_nt", 1, 16, 16, 8, 4, 256, 1, 1, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p1_nt},
  {"tsm_atb_f32_16x16_t8x4_r256_p2", 1, 16, 16, 8, 4, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p2},
  {"tsm_atb_f32_16x16_t8x4_r256_p2_nt", 1, 16, 16, 8, 4, 256, 2, 1, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p2_nt},
  {"tsm_atb_f32_16x16_t8x4_r512_p1", 1, 16, 16, 8, 4, 512, 1, 0, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p1},
  {"tsm_atb_f32_16x16_t8x4_r512_p1_nt", 1, 16, 16, 8, 4, 512, 1, 1, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p1_nt},
  {"tsm_atb_f32_16x16_t8x4_r512_p2", 1, 16, 16, 8, 4, 512, 2, 0, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p2},
  {"tsm_atb_f32_16x16_t8x4_r512_p2_nt", 1, 16, 16, 8, 4, 512, 2, 1, 131072, launch_tsm_atb_f32_16x16_t8x4_r512_p2_nt},
  {"tsm_atb_f32_16x16_t8x8_r128_p1", 1, 16, 16, 8, 8, 128, 1, 0, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p1},
  {"tsm_atb_f32_16x16_t8x8_r128_p2", 1, 16, 16, 8, 8, 128, 2, 0, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p2},
  {"tsm_atb_f32_16x16_t8x8_r256_p1", 1, 16, 16, 8, 8, 256, 1, 0, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p1},
  {"tsm_atb_f32_16x16_t8x8_r256_p2", 1, 16, 16, 8, 8, 256, 2, 0, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p2},
  {"tsm_atb_f32_16x16_t8x8_r512_p1", 1, 16, 16, 8, 8, 512, 1, 0, 131072, launch_tsm_atb_f32_16x16_t8x8_r512_p1},
  {"tsm_atb_f32_16x16_t8x8_r512_p2", 1, 16, 16, 8, 8, 512, 2, 0, 131072, launch_tsm_atb_f32_16x16_t8x8_r512_p2},
  {"tsm_atb_f32_16x32_t8x4_r128_p2", 1, 16, 32, 8, 4, 128, 2, 0, 49152, launch_tsm_atb_f32_16x32_t8x4_r128_p2},
  {"tsm_atb_f32_16x64_t8x4_r64_p2", 1, 16, 64, 8, 4, 64, 2, 0, 40960, launch_tsm_atb_f32_16x64_t8x4_r64_p2},
  {"tsm_atb_f32_32x1_t8x1_r128_p2", 1, 32, 1, 8, 1, 128, 2, 0, 33792, launch_tsm_atb_f32_32x1_t8x1_r128_p2},
  {"tsm_atb_f32_32x2_t8x2_r128_p2", 1, 32, 2, 8, 2, 128, 2, 0, 34816, launch_tsm_atb_f32_32x2_t8x2_r128_p2},
  {"tsm_atb_f32_32x4_t8x4_r128_p2", 1, 32, 4, 8, 4, 128, 2, 0, 36864, launch_tsm_atb_f32_32x4_t8x4_r128_p2},
  {"tsm_atb_f32_32x8_t8x4_r128_p2", 1, 32, 8, 8, 4, 128, 2, 0, 40960, launch_tsm_atb_f32_32x8_t8x4_r128_p2},
  {"tsm_atb_f32_32x16_t8x4_r128_p2", 1, 32, 16, 8, 4, 128, 2, 0, 49152, launch_tsm_atb_f32_32x16_t8x4_r128_p2},
  {"tsm_atb_f32_32x32_t4x4_r64_p1", 1, 32, 32, 4, 4, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p1},
  {"tsm_atb_f32_32x32_t4x4_r64_p2", 1, 32, 32, 4, 4, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p2},
  {"tsm_atb_f32_32x32_t8x4_r64_p1", 1, 32, 32, 8, 4, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p1},
  {"tsm_atb_f32_32x32_t8x4_r64_p1_nt", 1, 32, 32, 8, 4, 64, 1, 1, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p1_nt},
  {"tsm_atb_f32_32x32_t8x4_r64_p2", 1, 32, 32, 8, 4, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p2},
  {"tsm_atb_f32_32x32_t8x4_r64_p2_nt", 1, 32, 32, 8, 4, 64, 2, 1, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p2_nt},
  {"tsm_atb_f32_32x32_t8x4_r128_p1", 1, 32, 32, 8, 4, 128, 1, 0, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p1},
  {"tsm_atb_f32_32x32_t8x4_r128_p1_nt", 1, 32, 32, 8, 4, 128, 1, 1, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p1_nt},
  {"tsm_atb_f32_32x32_t8x4_r128_p2", 1, 32, 32, 8, 4, 128, 2, 0, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p2},
  {"tsm_atb_f32_32x32_t8x4_r128_p2_nt", 1, 32, 32, 8, 4, 128, 2, 1, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p2_nt},
  {"tsm_atb_f32_32x32_t8x8_r64_p1", 1, 32, 32, 8, 8, 64, 1, 0, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p1},
  {"tsm_atb_f32_32x32_t8x8_r64_p2", 1, 32, 32, 8, 8, 64, 2, 0, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p2},
  {"tsm_atb_f32_32x32_t8x8_r128_p1", 1, 32, 32, 8, 8, 128, 1, 0, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p1},
  {"tsm_atb_f32_32x32_t8x8_r128_p2", 1, 32, 32, 8, 8, 128, 2, 0, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p2},
  {"tsm_atb_f32_32x32_t8x8_r256_p1", 1, 32, 32, 8, 8, 256, 1, 0, 131072, launch_tsm_atb_f32_32x32_t8x8_r256_p1},
  {"tsm_atb_f32_32x32_t8x8_r256_p2", 1, 32, 32, 8, 8, 256, 2, 0, 131072, launch_tsm_atb_f32_32x32_t8x8_r256_p2},
  {"tsm_atb_f32_32x64_t8x4_r64_p2", 1, 32, 64, 8, 4, 64, 2, 0, 49152, launch_tsm_atb_f32_32x64_t8x4_r64_p2},
  {"tsm_atb_f32_64x1_t8x1_r64_p2", 1, 64, 1, 8, 1, 64, 2, 0, 33280, launch_tsm_atb_f32_64x1_t8x1_r64_p2},
  {"tsm_atb_f32_64x2_t8x2_r64_p2", 1, 64, 2, 8, 2, 64, 2, 0, 33792, launch_tsm_atb_f32_64x2_t8x2_r64_p2},
  {"tsm_atb_f32_64x4_t8x4_r64_p2", 1, 64, 4, 8, 4, 64, 2, 0, 34816, launch_tsm_atb_f32_64x4_t8x4_r64_p2},
  {"tsm_atb_f32_64x8_t8x4_r64_p2", 1, 64, 8, 8, 4, 64, 2, 0, 36864, launch_tsm_atb_f32_64x8_t8x4_r64_p2},
  {"tsm_atb_f32_64x16_t8x4_r64_p2", 1, 64, 16, 8, 4, 64, 2, 0, 40960, launch_tsm_atb_f32_64x16_t8x4_r64_p2},
  {"tsm_atb_f32_64x32_t8x4_r64_p2", 1, 64, 32, 8, 4, 64, 2, 0, 49152, launch_tsm_atb_f32_64x32_t8x4_r64_p2},
  {"tsm_atb_f32_64x64_t8x8_r64_p2", 1, 64, 64, 8, 8, 64, 2, 0, 65536, launch_tsm_atb_f32_64x64_t8x8_r64_p2},
  {"tsm_atb_f64_1x1_t1x1_r1024_p1", 0, 1, 1, 1, 1, 1024, 1, 0, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p1},
  {"tsm_atb_f64_1x1_t1x1_r1024_p1_nt", 0, 1, 1, 1, 1, 1024, 1, 1, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p1_nt},
  {"tsm_atb_f64_1x1_t1x1_r1024_p2", 0, 1, 1, 1, 1, 1024, 2, 0, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p2},
  {"tsm_atb_f64_1x1_t1x1_r1024_p2_nt", 0, 1, 1, 1, 1, 1024, 2, 1, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p2_nt},
  {"tsm_atb_f64_1x1_t1x1_r2048_p1", 0, 1, 1, 1, 1, 2048, 1, 0, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p1_nt", 0, 1, 1, 1, 1, 2048, 1, 1, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p1_nt},
  {"tsm_atb_f64_1x1_t1x1_r2048_p2", 0, 1, 1, 1, 1, 2048, 2, 0, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p2},
  {"tsm_atb_f64_1x1_t1x1_r2048_p2_nt", 0, 1, 1, 1, 1, 2048, 2, 1, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p2_nt},
  {"tsm_atb_f64_1x1_t1x1_r4096_p1", 0, 1, 1, 1, 1, 4096, 1, 0, 131072, launch_tsm_atb_f64_1x1_t1x1_r4096_p1},
  {"tsm_atb_f64_1x1_t1x1_r4096_p1_nt", 0, 1, 1, 1, 1, 4096, 1, 1, 131072, launch_tsm_atb_f64_1x1_t1x1_r4096_p1_nt},
  {"tsm_atb_f64_1x1_t1x1_r4096_p2", 0, 1, 1, 1, 1, 4096, 2, 0, 131072, launch_tsm_atb_f64_1x1_t1x1_r4096_p2},
  {"tsm_atb_f64_1x1_t1x1_r4096_p2_nt", 0, 1, 1, 1, 1, 4096, 2, 1, 131072, launch_tsm_atb_f64_1x1_t1x1_r4096_p2_nt},
  {"tsm_atb_f64_1x2_t1x2_r1024_p2", 0, 1, 2, 1, 2, 1024, 2, 0, 49152, launch_tsm_atb_f64_1x2_t1x2_r1024_p2},
  {"tsm_atb_f64_1x4_t1x4_r512_p2", 0, 1, 4, 1, 4, 512, 2, 0, 40960, launch_tsm_atb_f64_1x4_t1x4_r512_p2},
  {"tsm_atb_f64_1x8_t1x8_r256_p2", 0, 1, 8, 1, 8, 256, 2, 0, 36864, launch_tsm_atb_f64_1x8_t1x8_r256_p2},
  {"tsm_atb_f64_1x16_t1x8_r128_p2", 0, 1, 16, 1, 8, 128, 2, 0, 34816, launch_tsm_atb_f64_1x16_t1x8_r128_p2},
  {"tsm_atb_f64_1x32_t1x8_r64_p2", 0, 1, 32, 1, 8, 64, 2, 0, 33792, launch_tsm_atb_f64_1x32_t1x8_r64_p2},
  {"tsm_atb_f64_1x64_t1x8_r32_p2", 0, 1, 64, 1, 8, 32, 2, 0, 33280, launch_tsm_atb_f64_1x64_t1x8_r32_p2},
  {"tsm_atb_f64_2x1_t2x1_r1024_p2", 0, 2, 1, 2, 1, 1024, 2, 0, 49152, launch_tsm_atb_f64_2x1_t2x1_r1024_p2},
  {"tsm_atb_f64_2x2_t2x2_r512_p1", 0, 2, 2, 2, 2, 512, 1, 0, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p1},
  {"tsm_atb_f64_2x2_t2x2_r512_p1_nt", 0, 2, 2, 2, 2, 512, 1, 1, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p1_nt},
  {"tsm_atb_f64_2x2_t2x2_r512_p2", 0, 2, 2, 2, 2, 512, 2, 0, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p2},
  {"tsm_atb_f64_2x2_t2x2_r512_p2_nt", 0, 2, 2, 2, 2, 512, 2, 1, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p2_nt},
  {"tsm_atb_f64_2x2_t2x2_r1024_p1", 0, 2, 2, 2, 2, 1024, 1, 0, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p1_nt", 0, 2, 2, 2, 2, 1024, 1, 1, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p1_nt},
  {"tsm_atb_f64_2x2_t2x2_r1024_p2", 0, 2, 2, 2, 2, 1024, 2, 0, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p2},
  {"tsm_atb_f64_2x2_t2x2_r1024_p2_nt", 0, 2, 2, 2, 2, 1024, 2, 1, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p2_nt},
  {"tsm_atb_f64_2x2_t2x2_r2048_p1", 0, 2, 2, 2, 2, 2048, 1, 0, 131072, launch_tsm_atb_f64_2x2_t2x2_r2048_p1},
  {"tsm_atb_f64_2x2_t2x2_r2048_p1_nt", 0, 2, 2, 2, 2, 2048, 1, 1, 131072, launch_tsm_atb_f64_2x2_t2x2_r2048_p1_nt},
  {"tsm_atb_f64_2x2_t2x2_r2048_p2", 0, 2, 2, 2, 2, 2048, 2, 0, 131072, launch_tsm_atb_f64_2x2_t2x2_r2048_p2},
  {"tsm_atb_f64_2x2_t2x2_r2048_p2_nt", 0, 2, 2, 2, 2, 2048, 2, 1, 131072, launch_tsm_atb_f64_2x2_t2x2_r2048_p2_nt},
  {"tsm_atb_f64_2x4_t2x4_r512_p2", 0, 2, 4, 2, 4, 512, 2, 0, 49152, launch_tsm_atb_f64_2x4_t2x4_r512_p2},
  {"tsm_atb_f64_2x8_t2x8_r256_p2", 0, 2, 8, 2, 8, 256, 2, 0, 40960, launch_tsm_atb_f64_2x8_t2x8_r256_p2},
  {"tsm_atb_f64_2x16_t2x8_r128_p2", 0, 2, 16, 2, 8, 128, 2, 0, 36864, launch_tsm_atb_f64_2x16_t2x8_r128_p2},
  {"tsm_atb_f64_2x32_t2x8_r64_p2", 0, 2, 32, 2, 8, 64, 2, 0, 34816, launch_tsm_atb_f64_2x32_t2x8_r64_p2},
  {"tsm_atb_f64_2x64_t2x8_r32_p2", 0, 2, 64, 2, 8, 32, 2, 0, 33792, launch_tsm_atb_f64_2x64_t2x8_r32_p2},
  {"tsm_atb_f64_4x1_t4x1_r512_p2", 0, 4, 1, 4, 1, 512, 2, 0, 40960, launch_tsm_atb_f64_4x1_t4x1_r512_p2},
  {"tsm_atb_f64_4x2_t4x2_r512_p2", 0, 4, 2, 4, 2, 512, 2, 0, 49152, launch_tsm_atb_f64_4x2_t4x2_r512_p2},
  {"tsm_atb_f64_4x4_t2x2_r512_p1", 0, 4, 4, 2, 2, 512, 1, 0, 65536, launch_tsm_atb_f64_4x4_t2x2_r512_p1},
  {"tsm_atb_f64_4x4_t2x2_r512_p2", 0, 4, 4, 2, 2, 512, 2, 0, 65536, launch_tsm_atb_f64_4x4_t2x2_r512_p2},
  {"tsm_atb_f64_4x4_t4x4_r256_p1", 0, 4, 4, 4, 4, 256, 1, 0, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p1},
  {"tsm_atb_f64_4x4_t4x4_r256_p1_nt", 0, 4, 4, 4, 4, 256, 1, 1, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p1_nt},
  {"tsm_atb_f64_4x4_t4x4_r256_p2", 0, 4, 4, 4, 4, 256, 2, 0, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p2},
  {"tsm_atb_f64_4x4_t4x4_r256_p2_nt", 0, 4, 4, 4, 4, 256, 2, 1, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p2_nt},
  {"tsm_atb_f64_4x4_t4x4_r512_p1", 0, 4, 4, 4, 4, 512, 1, 0, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p1},
  {"tsm_atb_f64_4x4_t4x4_r512_p1_nt", 0, 4, 4, 4, 4, 512, 1, 1, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p1_nt},
  {"tsm_atb_f64_4x4_t4x4_r512_p2", 0, 4, 4, 4, 4, 512, 2, 0, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p2},
  {"tsm_atb_f64_4x4_t4x4_r512_p2_nt", 0, 4, 4, 4, 4, 512, 2, 1, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p2_nt},
  {"tsm_atb_f64_4x4_t4x4_r1024_p1", 0, 4, 4, 4, 4, 1024, 1, 0, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p1},
  {"tsm_atb_f64_4x4_t4x4_r1024_p1_nt", 0, 4, 4, 4, 4, 1024, 1, 1, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p1_nt},
  {"tsm_atb_f64_4x4_t4x4_r1024_p2", 0, 4, 4, 4, 4, 1024, 2, 0, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p2},
  {"tsm_atb_f64_4x4_t4x4_r1024_p2_nt", 0, 4, 4, 4, 4, 1024, 2, 1, 131072, launch_tsm_atb_f64_4x4_t4x4_r1024_p2_nt},
  {"tsm_atb_f64_4x8_t4x8_r256_p2", 0, 4, 8, 4, 8, 256, 2, 0, 49152, launch_tsm_atb_f64_4x8_t4x8_r256_p2},
  {"tsm_atb_f64_4x16_t4x8_r128_p2", 0, 4, 16, 4, 8, 128, 2, 0, 40960, launch_tsm_atb_f64_4x16_t4x8_r128_p2},
  {"tsm_atb_f64_4x32_t4x4_r32_p1", 0, 4, 32, 4, 4, 32, 1, 0, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p1},
  {"tsm_atb_f64_4x32_t4x4_r32_p2", 0, 4, 32, 4, 4, 32, 2, 0, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p2},
  {"tsm_atb_f64_4x32_t4x4_r64_p1", 0, 4, 32, 4, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p1},
  {"tsm_atb_f64_4x32_t4x4_r64_p2", 0, 4, 32, 4, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p2},
  {"tsm_atb_f64_4x32_t4x4_r128_p1", 0, 4, 32, 4, 4, 128, 1, 0, 73728, launch_tsm_atb_f64_4x32_t4x4_r128_p1},
  {"tsm_atb_f64_4x32_t4x4_r128_p2", 0, 4, 32, 4, 4, 128, 2, 0, 73728, launch_tsm_atb_f64_4x32_t4x4_r128_p2},
  {"tsm_atb_f64_4x32_t4x8_r64_p1", 0, 4, 32, 4, 8, 64, 1, 0, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p1},
  {"tsm_atb_f64_4x32_t4x8_r64_p1_nt", 0, 4, 32, 4, 8, 64, 1, 1, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p1_nt},
  {"tsm_atb_f64_4x32_t4x8_r64_p2", 0, 4, 32, 4, 8, 64, 2, 0, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p2},
  {"tsm_atb_f64_4x32_t4x8_r64_p2_nt", 0, 4, 32, 4, 8, 64, 2, 1, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p2_nt},
  {"tsm_atb_f64_4x32_t4x8_r128_p1", 0, 4, 32, 4, 8, 128, 1, 0, 73728, launch_tsm_atb_f64_4x32_t4x8_r128_p1},
  {"tsm_atb_f64_4x32_t4x8_r128_p1_nt", 0, 4, 32, 4, 8, 128, 1, 1, 73728, launch_tsm_atb_f64_4x32_t4x8_r128_p1_nt},
  {"tsm_atb_f64_4x32_t4x8_r128_p2", 0, 4, 32, 4, 8, 128, 2, 0, 73728, launch_tsm_atb_f64_4x32_t4x8_r128_p2},
  {"tsm_atb_f64_4x32_t4x8_r128_p2_nt", 0, 4, 32, 4, 8, 128, 2, 1, 73728, launch_tsm_atb_f64_4x32_t4x8_r128_p2_nt},
  {"tsm_atb_f64_4x64_t4x8_r32_p2", 0, 4, 64, 4, 8, 32, 2, 0, 34816, launch_tsm_atb_f64_4x64_t4x8_r32_p2},
  {"tsm_atb_f64_8x1_t8x1_r256_p2", 0, 8, 1, 8, 1, 256, 2, 0, 36864, launch_tsm_atb_f64_8x1_t8x1_r256_p2},
  {"tsm_atb_f64_8x2_t8x2_r256_p2", 0, 8, 2, 8, 2, 256, 2, 0, 40960, launch_tsm_atb_f64_8x2_t8x2_r256_p2},
  {"tsm_atb_f64_8x4_t8x4_r256_p2", 0, 8, 4, 8, 4, 256, 2, 0, 49152, launch_tsm_atb_f64_8x4_t8x4_r256_p2},
  {"tsm_atb_f64_8x8_t2x2_r256_p1", 0, 8, 8, 2, 2, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t2x2_r256_p1},
  {"tsm_atb_f64_8x8_t2x2_r256_p2", 0, 8, 8, 2, 2, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t2x2_r256_p2},
  {"tsm_atb_f64_8x8_t4x4_r128_p1", 0, 8, 8, 4, 4, 128, 1, 0, 32768, launch_tsm_atb_f64_8x8_t4x4_r128_p1},
  {"tsm_atb_f64_8x8_t4x4_r128_p2", 0, 8, 8, 4, 4, 128, 2, 0, 32768, launch_tsm_atb_f64_8x8_t4x4_r128_p2},
  {"tsm_atb_f64_8x8_t4x4_r256_p1", 0, 8, 8, 4, 4, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p1},
  {"tsm_atb_f64_8x8_t4x4_r256_p2", 0, 8, 8, 4, 4, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p2},
  {"tsm_atb_f64_8x8_t4x4_r512_p1", 0, 8, 8, 4, 4, 512, 1, 0, 131072, launch_tsm_atb_f64_8x8_t4x4_r512_p1},
  {"tsm_atb_f64_8x8_t4x4_r512_p2", 0, 8, 8, 4, 4, 512, 2, 0, 131072, launch_tsm_atb_f64_8x8_t4x4_r512_p2},
  {"tsm_atb_f64_8x8_t8x8_r256_p1", 0, 8, 8, 8, 8, 256, 1, 0, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p1},
  {"tsm_atb_f64_8x8_t8x8_r256_p1_nt", 0, 8, 8, 8, 8, 256, 1, 1, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p1_nt},
  {"tsm_atb_f64_8x8_t8x8_r256_p2", 0, 8, 8, 8, 8, 256, 2, 0, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p2},
  {"tsm_atb_f64_8x8_t8x8_r256_p2_nt", 0, 8, 8, 8, 8, 256, 2, 1, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p2_nt},
  {"tsm_atb_f64_8x8_t8x8_r512_p1", 0, 8, 8, 8, 8, 512, 1, 0, 131072, launch_tsm_atb_f64_8x8_t8x8_r512_p1},
  {"tsm_atb_f64_8x8_t8x8_r512_p1_nt", 0, 8, 8, 8, 8, 512, 1, 1, 131072, launch_tsm_atb_f64_8x8_t8x8_r512_p1_nt},
  {"tsm_atb_f64_8x8_t8x8_r512_p2", 0, 8, 8, 8, 8, 512, 2, 0, 131072, launch_tsm_atb_f64_8x8_t8x8_r512_p2},
  {"tsm_atb_f64_8x8_t8x8_r512_p2_nt", 0, 8, 8, 8, 8, 512, 2, 1, 131072, launch_tsm_atb_f64_8x8_t8x8_r512_p2_nt},
  {"tsm_atb_f64_8x16_t8x8_r128_p2", 0, 8, 16, 8, 8, 128, 2, 0, 49152, launch_tsm_atb_f64_8x16_t8x8_r128_p2},
  {"tsm_atb_f64_8x32_t8x8_r64_p2", 0, 8, 32, 8, 8, 64, 2, 0, 40960, launch_tsm_atb_f64_8x32_t8x8_r64_p2},
  {"tsm_atb_f64_8x64_t8x8_r32_p2", 0, 8, 64, 8, 8, 32, 2, 0, 36864, launch_tsm_atb_f64_8x64_t8x8_r32_p2},
  {"tsm_atb_f64_16x1_t8x1_r128_p2", 0, 16, 1, 8, 1, 128, 2, 0, 34816, launch_tsm_atb_f64_16x1_t8x1_r128_p2},
  {"tsm_atb_f64_16x2_t8x2_r128_p2", 0, 16, 2, 8, 2, 128, 2, 0, 36864, launch_tsm_atb_f64_16x2_t8x2_r128_p2},
  {"tsm_atb_f64_16x4_t8x4_r128_p2", 0, 16, 4, 8, 4, 128, 2, 0, 40960, launch_tsm_atb_f64_16x4_t8x4_r128_p2},
  {"tsm_atb_f64_16x8_t8x8_r128_p2", 0, 16, 8, 8, 8, 128, 2, 0, 49152, launch_tsm_atb_f64_16x8_t8x8_r128_p2},
  {"tsm_atb_f64_16x16_t2x2_r64_p1", 0, 16, 16, 2, 2, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_t2x2_r64_p1},
  {"tsm_atb_f64_16x16_t2x2_r64_p2", 0, 16, 16, 2, 2, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_t2x2_r64_p2},
  {"tsm_atb_f64_16x16_t4x4_r64_p1", 0, 16, 16, 4, 4, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_t4x4_r64_p1},
  {"tsm_atb_f64_16x16_t4x4_r64_p2", 0, 16, 16, 4, 4, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_t4x4_r64_p2},
  {"tsm_atb_f64_16x16_t4x4_r128_p1", 0, 16, 16, 4, 4, 128, 1, 0, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p1},
  {"tsm_atb_f64_16x16_t4x4_r128_p2", 0, 16, 16, 4, 4, 128, 2, 0, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p2},
  {"tsm_atb_f64_16x16_t4x4_r256_p1", 0, 16, 16, 4, 4, 256, 1, 0, 131072, launch_tsm_atb_f64_16x16_t4x4_r256_p1},
  {"tsm_atb_f64_16x16_t4x4_r256_p2", 0, 16, 16, 4, 4, 256, 2, 0, 131072, launch_tsm_atb_f64_16x16_t4x4_r256_p2},
  {"tsm_atb_f64_16x16_t8x8_r64_p1", 0, 16, 16, 8, 8, 64, 1, 0, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p1},
  {"tsm_atb_f64_16x16_t8x8_r64_p1_nt", 0, 16, 16, 8, 8, 64, 1, 1, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p1_nt},
  {"tsm_atb_f64_16x16_t8x8_r64_p2", 0, 16, 16, 8, 8, 64, 2, 0, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p2},
  {"tsm_atb_f64_16x16_t8x8_r64_p2_nt", 0, 16, 16, 8, 8, 64, 2, 1, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p2_nt},
  {"tsm_atb_f64_16x16_t8x8_r128_p1", 0, 16, 16, 8, 8, 128, 1, 0, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p1},
  {"tsm_atb_f64_16x16_t8x8_r128_p1_nt", 0, 16, 16, 8, 8, 128, 1, 1, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p1_nt},
  {"tsm_atb_f64_16x16_t8x8_r128_p2", 0, 16, 16, 8, 8, 128, 2, 0, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p2},
  {"tsm_atb_f64_16x16_t8x8_r128_p2_nt", 0, 16, 16, 8, 8, 128, 2, 1, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p2_nt},
  {"tsm_atb_f64_16x16_t8x8_r256_p1", 0, 16, 16, 8, 8, 256, 1, 0, 131072, launch_tsm_atb_f64_16x16_t8x8_r256_p1},
  {"tsm_atb_f64_16x16_t8x8_r256_p1_nt", 0, 16, 16, 8, 8, 256, 1, 1, 131072, launch_tsm_atb_f64_16x16_t8x8_r256_p1_nt},
  {"tsm_atb_f64_16x16_t8x8_r256_p2", 0, 16, 16, 8, 8, 256, 2, 0, 131072, launch_tsm_atb_f64_16x16_t8x8_r256_p2},
  {"tsm_atb_f64_16x16_t8x8_r256_p2_nt", 0, 16, 16, 8, 8, 256, 2, 1, 131072, launch_tsm_atb_f64_16x16_t8x8_r256_p2_nt},
  {"tsm_atb_f64_16x32_t8x8_r64_p2", 0, 16, 32, 8, 8, 64, 2, 0, 49152, launch_tsm_atb_f64_16x32_t8x8_r64_p2},
  {"tsm_atb_f64_16x64_t8x8_r32_p2", 0, 16, 64, 8, 8, 32, 2, 0, 40960, launch_tsm_atb_f64_16x64_t8x8_r32_p2},
  {"tsm_atb_f64_32x1_t8x1_r64_p2", 0, 32, 1, 8, 1, 64, 2, 0, 33792, launch_tsm_atb_f64_32x1_t8x1_r64_p2},
  {"tsm_atb_f64_32x2_t8x2_r64_p2", 0, 32, 2, 8, 2, 64, 2, 0, 34816, launch_tsm_atb_f64_32x2_t8x2_r64_p2},
  {"tsm_atb_f64_32x4_t4x4_r32_p1", 0, 32, 4, 4, 4, 32, 1, 0, 18432, launch_tsm_atb_f64_32x4_t4x4_r32_p1},
  {"tsm_atb_f64_32x4_t4x4_r32_p2", 0, 32, 4, 4, 4, 32, 2, 0, 18432, launch_tsm_atb_f64_32x4_t4x4_r32_p2},
  {"tsm_atb_f64_32x4_t4x4_r64_p1", 0, 32, 4, 4, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p1},
  {"tsm_atb_f64_32x4_t4x4_r64_p2", 0, 32, 4, 4, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p2},
  {"tsm_atb_f64_32x4_t4x4_r128_p1", 0, 32, 4, 4, 4, 128, 1, 0, 73728, launch_tsm_atb_f64_32x4_t4x4_r128_p1},
  {"tsm_atb_f64_32x4_t4x4_r128_p2", 0, 32, 4, 4, 4, 128, 2, 0, 73728, launch_tsm_atb_f64_32x4_t4x4_r128_p2},
  {"tsm_atb_f64_32x4_t8x4_r64_p1", 0, 32, 4, 8, 4, 64, 1, 0, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p1},
  {"tsm_atb_f64_32x4_t8x4_r64_p1_nt", 0, 32, 4, 8, 4, 64, 1, 1, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p1_nt},
  {"tsm_atb_f64_32x4_t8x4_r64_p2", 0, 32, 4, 8, 4, 64, 2, 0, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p2},
  {"tsm_atb_f64_32x4_t8x4_r64_p2_nt", 0, 32, 4, 8, 4, 64, 2, 1, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p2_nt},
  {"tsm_atb_f64_32x4_t8x4_r128_p1", 0, 32, 4, 8, 4, 128, 1, 0, 73728, launch_tsm_atb_f64_32x4_t8x4_r128_p1},
  {"tsm_atb_f64_32x4_t8x4_r128_p1_nt", 0, 32, 4, 8, 4, 128, 1, 1, 73728, launch_tsm_atb_f64_32x4_t8x4_r128_p1_nt},
  {"tsm_atb_f64_32x4_t8x4_r128_p2", 0, 32, 4, 8, 4, 128, 2, 0, 73728, launch_tsm_atb_f64_32x4_t8x4_r128_p2},
  {"tsm_atb_f64_32x4_t8x4_r128_p2_nt", 0, 32, 4, 8, 4, 128, 2, 1, 73728, launch_tsm_atb_f64_32x4_t8x4_r128_p2_nt},
  {"tsm_atb_f64_32x8_t8x8_r64_p2", 0, 32, 8, 8, 8, 64, 2, 0, 40960, launch_tsm_atb_f64_32x8_t8x8_r64_p2},
  {"tsm_atb_f64_32x16_t8x8_r64_p2", 0, 32, 16, 8, 8, 64, 2, 0, 49152, launch_tsm_atb_f64_32x16_t8x8_r64_p2},
  {"tsm_atb_f64_32x32_t4x4_r32_p1", 0, 32, 32, 4, 4, 32, 1, 0, 32768, launch_tsm_atb_f64_32x32_t4x4_r32_p1},
  {"tsm_atb_f64_32x32_t4x4_r32_p2", 0, 32, 32, 4, 4, 32, 2, 0, 32768, launch_tsm_atb_f64_32x32_t4x4_r32_p2},
  {"tsm_atb_f64_32x32_t4x4_r64_p1", 0, 32, 32, 4, 4, 64, 1, 0, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p1},
  {"tsm_atb_f64_32x32_t4x4_r64_p2", 0, 32, 32, 4, 4, 64, 2, 0, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p2},
  {"tsm_atb_f64_32x32_t8x8_r32_p1", 0, 32, 32, 8, 8, 32, 1, 0, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p1},
  {"tsm_atb_f64_32x32_t8x8_r32_p1_nt", 0, 32, 32, 8, 8, 32, 1, 1, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p1_nt},
  {"tsm_atb_f64_32x32_t8x8_r32_p2", 0, 32, 32, 8, 8, 32, 2, 0, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p2},
  {"tsm_atb_f64_32x32_t8x8_r32_p2_nt", 0, 32, 32, 8, 8, 32, 2, 1, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p2_nt},
  {"tsm_atb_f64_32x32_t8x8_r64_p1", 0, 32, 32, 8, 8, 64, 1, 0, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p1},
  {"tsm_atb_f64_32x32_t8x8_r64_p1_nt", 0, 32, 32, 8, 8, 64, 1, 1, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p1_nt},
  {"tsm_atb_f64_32x32_t8x8_r64_p2", 0, 32, 32, 8, 8, 64, 2, 0, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p2},
  {"tsm_atb_f64_32x32_t8x8_r64_p2_nt", 0, 32, 32, 8, 8, 64, 2, 1, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p2_nt},
  {"tsm_atb_f64_32x32_t8x8_r128_p1", 0, 32, 32, 8, 8, 128, 1, 0, 131072, launch_tsm_atb_f64_32x32_t8x8_r128_p1},
  {"tsm_atb_f64_32x32_t8x8_r128_p1_nt", 0, 32, 32, 8, 8, 128, 1, 1, 131072, launch_tsm_atb_f64_32x32_t8x8_r128_p1_nt},
  {"tsm_atb_f64_32x32_t8x8_r128_p2", 0, 32, 32, 8, 8, 128, 2, 0, 131072, launch_tsm_atb_f64_32x32_t8x8_r128_p2},
  {"tsm_atb_f64_32x32_t8x8_r128_p2_nt", 0, 32, 32, 8, 8, 128, 2, 1, 131072, launch_tsm_atb_f64_32x32_t8x8_r128_p2_nt},
  {"tsm_atb_f64_32x64_t8x8_r32_p2", 0, 32, 64, 8, 8, 32, 2, 0, 49152, launch_tsm_atb_f64_32x64_t8x8_r32_p2},
  {"tsm_atb_f64_64x1_t8x1_r32_p2", 0, 64, 1, 8, 1, 32, 2, 0, 33280, launch_tsm_atb_f64_64x1_t8x1_r32_p2},
  {"tsm_atb_f64_64x2_t8x2_r32_p2", 0, 64, 2, 8, 2, 32, 2, 0, 33792, launch_tsm_atb_f64_64x2_t8x2_r32_p2},
  {"tsm_atb_f64_64x4_t8x4_r32_p2", 0, 64, 4, 8, 4, 32, 2, 0, 34816, launch_tsm_atb_f64_64x4_t8x4_r32_p2},
  {"tsm_atb_f64_64x8_t8x8_r32_p2", 0, 64, 8, 8, 8, 32, 2, 0, 36864, launch_tsm_atb_f64_64x8_t8x8_r32_p2},
  {"tsm_atb_f64_64x16_t8x8_r32_p2", 0, 64, 16, 8, 8, 32, 2, 0, 40960, launch_tsm_atb_f64_64x16_t8x8_r32_p2},
  {"tsm_atb_f64_64x32_t8x8_r32_p2", 0, 64, 32, 8, 8, 32, 2, 0, 49152, launch_tsm_atb_f64_64x32_t8x8_r32_p2},
  {"tsm_atb_f64_64x64_t8x8_r16_p1", 0, 64, 64, 8, 8, 16, 1, 0, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p1},
  {"tsm_atb_f64_64x64_t8x8_r16_p1_nt", 0, 64, 64, 8, 8, 16, 1, 1, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p1_nt},
  {"tsm_atb_f64_64x64_t8x8_r16_p2", 0, 64, 64, 8, 8, 16, 2, 0, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p2},
  {"tsm_atb_f64_64x64_t8x8_r16_p2_nt", 0, 64, 64, 8, 8, 16, 2, 1, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p2_nt},
  {"tsm_atb_f64_64x64_t8x8_r32_p1", 0, 64, 64, 8, 8, 32, 1, 0, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p1},
  {"tsm_atb_f64_64x64_t8x8_r32_p1_nt", 0, 64, 64, 8, 8, 32, 1, 1, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p1_nt},
  {"tsm_atb_f64_64x64_t8x8_r32_p2", 0, 64, 64, 8, 8, 32, 2, 0, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p2},
  {"tsm_atb_f64_64x64_t8x8_r32_p2_nt", 0, 64, 64, 8, 8, 32, 2, 1, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p2_nt},
  {"tsm_atb_f64_64x64_t8x8_r64_p1", 0, 64, 64, 8, 8, 64, 1, 0, 131072, launch_tsm_atb_f64_64x64_t8x8_r64_p1},
  {"tsm_atb_f64_64x64_t8x8_r64_p1_nt", 0, 64, 64, 8, 8, 64, 1, 1, 131072, launch_tsm_atb_f64_64x64_t8x8_r64_p1_nt},
  {"tsm_atb_f64_64x64_t8x8_r64_p2", 0, 64, 64, 8, 8, 64, 2, 0, 131072, launch_tsm_atb_f64_64x64_t8x8_r64_p2},
  {"tsm_atb_f64_64x64_t8x8_r64_p2_nt", 0, 64, 64, 8, 8, 64, 2, 1, 131072, launch_tsm_atb_f64_64x64_t8x8_r64_p2_nt},
};

static const tsm::TunedEntry kTuned[] = {
  {2, 1, 1, 0, 2, 0},  // tsm_atb_bf16_1x1_t1x1_r4096_p2
  {2, 1, 2, 1, 2, 0},  // tsm_atb_bf16_1x2_t1x2_r4096_p2
  {2, 1, 4, 2, 2, 0},  // tsm_atb_bf16_1x4_t1x4_r2048_p2
  {2, 1, 8, 3, 2, 0},  // tsm_atb_bf16_1x8_t1x4_r1024_p2
  {2, 1, 16, 4, 2, 0},  // tsm_atb_bf16_1x16_t1x4_r512_p2
  {2, 1, 32, 5, 2, 0},  // tsm_atb_bf16_1x32_t1x4_r256_p2
  {2, 1, 64, 6, 2, 0},  // tsm_atb_bf16_1x64_t1x4_r128_p2
  {2, 2, 1, 7, 2, 0},  // tsm_atb_bf16_2x1_t2x1_r4096_p2
  {2, 2, 2, 8, 2, 0},  // tsm_atb_bf16_2x2_t2x2_r4096_p2
  {2, 2, 4, 9, 2, 0},  // tsm_atb_bf16_2x4_t2x4_r2048_p2
  {2, 2, 8, 10, 2, 0},  // tsm_atb_bf16_2x8_t2x4_r1024_p2
  {2, 2, 16, 11, 2, 0},  // tsm_atb_bf16_2x16_t2x4_r512_p2
  {2, 2, 32, 12, 2, 0},  // tsm_atb_bf16_2x32_t2x4_r256_p2
  {2, 2, 64, 13, 2, 0},  // tsm_atb_bf16_2x64_t2x4_r128_p2
  {2, 4, 1, 14, 2, 0},  // tsm_atb_bf16_4x1_t4x1_r2048_p2
  {2, 4, 2, 15, 2, 0},  // tsm_atb_bf16_4x2_t4x2_r2048_p2
  {2, 4, 4, 16, 2, 0},  // tsm_atb_bf16_4x4_t4x4_r2048_p2
  {2, 4, 8, 17, 2, 0},  // tsm_atb_bf16_4x8_t4x4_r1024_p2
  {2, 4, 16, 18, 2, 0},  // tsm_atb_bf16_4x16_t4x4_r512_p2
  {2, 4, 32, 19, 2, 0},  // tsm_atb_bf16_4x32_t4x4_r256_p2
  {2, 4, 64, 20, 2, 0},  // tsm_atb_bf16_4x64_t4x4_r128_p2
  {2, 8, 1, 21, 2, 0},  // tsm_atb_bf16_8x1_t8x1_r1024_p2
  {2, 8, 2, 22, 2, 0},  // tsm_atb_bf16_8x2_t8x2_r1024_p2
  {2, 8, 4, 23, 2, 0},  // tsm_atb_bf16_8x4_t8x4_r1024_p2
  {2, 8, 8, 43, 1, 1},  // tsm_atb_bf16_8x8_t8x8_r1024_p2
  {2, 8, 16, 46, 2, 0},  // tsm_atb_bf16_8x16_t8x4_r512_p2
  {2, 8, 32, 47, 2, 0},  // tsm_atb_bf16_8x32_t8x4_r256_p2
  {2, 8, 64, 48, 2, 0},  // tsm_atb_bf16_8x64_t8x4_r128_p2
  {2, 16, 1, 49, 2, 0},  // tsm_atb_bf16_16x1_t8x1_r512_p2
  {2, 16, 2, 50, 2, 0},  // tsm_atb_bf16_16x2_t8x2_r512_p2
  {2, 16, 4, 51, 2, 0},  // tsm_atb_bf16_16x4_t8x4_r512_p2
  {2, 16, 8, 52, 2, 0},  // tsm_atb_bf16_16x8_t8x4_r512_p2
  {2, 16, 16, 58, 2, 0},  // tsm_atb_bf16_16x16_t8x4_r256_p2_nt
  {2, 16, 32, 69, 2, 0},  // tsm_atb_bf16_16x32_t8x4_r256_p2
  {2, 16, 64, 70, 2, 0},  // tsm_atb_bf16_16x64_t8x4_r128_p2
  {2, 32, 1, 71, 2, 0},  // tsm_atb_bf16_32x1_t8x1_r256_p2
  {2, 32, 2, 72, 2, 0},  // tsm_atb_bf16_32x2_t8x2_r256_p2
  {2, 32, 4, 73, 2, 0},  // tsm_atb_bf16_32x4_t8x4_r256_p2
  {2, 32, 8, 74, 2, 0},  // tsm_atb_bf16_32x8_t8x4_r256_p2
  {2, 32, 16, 75, 2, 0},  // tsm_atb_bf16_32x16_t8x4_r256_p2
  {2, 32, 32, 88, 2, 1},  // tsm_atb_bf16_32x32_t8x8_r128_p1
  {2, 32, 64, 92, 2, 0},  // tsm_atb_bf16_32x64_t8x4_r64_p2
  {2, 64, 1, 93, 2, 0},  // tsm_atb_bf16_64x1_t8x1_r128_p2
  {2, 64, 2, 94, 2, 0},  // tsm_atb_bf16_64x2_t8x2_r128_p2
  {2, 64, 4, 95, 2, 0},  // tsm_atb_bf16_64x4_t8x4_r128_p2
  {2, 64, 8, 96, 2, 0},  // tsm_atb_bf16_64x8_t8x4_r128_p2
  {2, 64, 16, 97, 2, 0},  // tsm_atb_bf16_64x16_t8x4_r128_p2
  {2, 64, 32, 98, 2, 0},  // tsm_atb_bf16_64x32_t8x4_r64_p2
  {2, 64, 64, 99, 2, 0},  // tsm_atb_bf16_64x64_t8x8_r64_p2
  {1, 1, 1, 100, 2, 0},  // tsm_atb_f32_1x1_t1x1_r4096_p2
  {1, 1, 2, 101, 2, 0},  // tsm_atb_f32_1x2_t1x2_r2048_p2
  {1, 1, 4, 102, 2, 0},  // tsm_atb_f32_1x4_t1x4_r1024_p2
  {1, 1, 8, 103, 2, 0},  // tsm_atb_f32_1x8_t1x4_r512_p2
  {1, 1, 16, 104, 2, 0},  // tsm_atb_f32_1x16_t1x4_r256_p2
  {1, 1, 32, 105, 2, 0},  // tsm_atb_f32_1x32_t1x4_r128_p2
  {1, 1, 64, 106, 2, 0},  // tsm_atb_f32_1x64_t1x4_r64_p2
  {1, 2, 1, 107, 2, 0},  // tsm_atb_f32_2x1_t2x1_r2048_p2
  {1, 2, 2, 108, 2, 0},  // tsm_atb_f32_2x2_t2x2_r2048_p2
  {1, 2, 4, 109, 2, 0},  // tsm_atb_f32_2x4_t2x4_r1024_p2
  {1, 2, 8, 110, 2, 0},  // tsm_atb_f32_2x8_t2x4_r512_p2
  {1, 2, 16, 111, 2, 0},  // tsm_atb_f32_2x16_t2x4_r256_p2
  {1, 2, 32, 112, 2, 0},  // tsm_atb_f32_2x32_t2x4_r128_p2
  {1, 2, 64, 113, 2, 0},  // tsm_atb_f32_2x64_t2x4_r64_p2
  {1, 4, 1, 114, 2, 0},  // tsm_atb_f32_4x1_t4x1_r1024_p2
  {1, 4, 2, 115, 2, 0},  // tsm_atb_f32_4x2_t4x2_r1024_p2
  {1, 4, 4, 116, 2, 0},  // tsm_atb_f32_4x4_t4x4_r1024_p2
  {1, 4, 8, 117, 2, 0},  // tsm_atb_f32_4x8_t4x4_r512_p2
  {1, 4, 16, 118, 2, 0},  // tsm_atb_f32_4x16_t4x4_r256_p2
  {1, 4, 32, 119, 2, 0},  // tsm_atb_f32_4x32_t4x4_r128_p2
  {1, 4, 64, 120, 2, 0},  // tsm_atb_f32_4x64_t4x4_r64_p2
  {1, 8, 1, 121, 2, 0},  // tsm_atb_f32_8x1_t8x1_r512_p2
  {1, 8, 2, 122, 2, 0},  // tsm_atb_f32_8x2_t8x2_r512_p2
  {1, 8, 4, 123, 2, 0},  // tsm_atb_f32_8x4_t8x4_r512_p2
  {1, 8, 8, 141, 1, 1},  // tsm_atb_f32_8x8_t8x4_r1024_p2_nt
  {1, 8, 16, 148, 2, 0},  // tsm_atb_f32_8x16_t8x4_r256_p2
  {1, 8, 32, 149, 2, 0},  // tsm_atb_f32_8x32_t8x4_r128_p2
  {1, 8, 64, 150, 2, 0},  // tsm_atb_f32_8x64_t8x4_r64_p2
  {1, 16, 1, 151, 2, 0},  // tsm_atb_f32_16x1_t8x1_r256_p2
  {1, 16, 2, 152, 2, 0},  // tsm_atb_f32_16x2_t8x2_r256_p2
  {1, 16, 4, 153, 2, 0},  // tsm_atb_f32_16x4_t8x4_r256_p2
  {1, 16, 8, 154, 2, 0},  // tsm_atb_f32_16x8_t8x4_r256_p2
  {1, 16, 16, 166, 1, 0},  // tsm_atb_f32_16x16_t8x4_r256_p2_nt
  {1, 16, 32, 177, 2, 0},  // tsm_atb_f32_16x32_t8x4_r128_p2
  {1, 16, 64, 178, 2, 0},  // tsm_atb_f32_16x64_t8x4_r64_p2
  {1, 32, 1, 179, 2, 0},  // tsm_atb_f32_32x1_t8x1_r128_p2
  {1, 32, 2, 180, 2, 0},  // tsm_atb_f32_32x2_t8x2_r128_p2
  {1, 32, 4, 181, 2, 0},  // tsm_atb_f32_32x4_t8x4_r128_p2
  {1, 32, 8, 182, 2, 0},  // tsm_atb_f32_32x8_t8x4_r128_p2
  {1, 32, 16, 183, 2, 0},  // tsm_atb_f32_32x16_t8x4_r128_p2
  {1, 32, 32, 191, 2, 1},  // tsm_atb_f32_32x32_t8x4_r128_p1_nt
  {1, 32, 64, 200, 2, 0},  // tsm_atb_f32_32x64_t8x4_r64_p2
  {1, 64, 1, 201, 2, 0},  // tsm_atb_f32_64x1_t8x1_r64_p2
  {1, 64, 2, 202, 2, 0},  // tsm_atb_f32_64x2_t8x2_r64_p2
  {1, 64, 4, 203, 2, 0},  // tsm_atb_f32_64x4_t8x4_r64_p2
  {1, 64, 8, 204, 2, 0},  // tsm_atb_f32_64x8_t8x4_r64_p2
  {1, 64, 16, 205, 2, 0},  // tsm_atb_f32_64x16_t8x4_r64_p2
  {1, 64, 32, 206, 2, 0},  // tsm_atb_f32_64x32_t8x4_r64_p2
  {1, 64, 64, 207, 2, 0},  // tsm_atb_f32_64x64_t8x8_r64_p2
  {0, 1, 1, 211, 1, 1},  // tsm_atb_f64_1x1_t1x1_r1024_p2_nt
  {0, 1, 2, 220, 2, 0},  // tsm_atb_f64_1x2_t1x2_r1024_p2
  {0, 1, 4, 221, 2, 0},  // tsm_atb_f64_1x4_t1x4_r512_p2
  {0, 1, 8, 222, 2, 0},  // tsm_atb_f64_1x8_t1x8_r256_p2
  {0, 1, 16, 223, 2, 0},  // tsm_atb_f64_1x16_t1x8_r128_p2
  {0, 1, 32, 224, 2, 0},  // tsm_atb_f64_1x32_t1x8_r64_p2
  {0, 1, 64, 225, 2, 0},  // tsm_atb_f64_1x64_t1x8_r32_p2
  {0, 2, 1, 226, 2, 0},  // tsm_atb_f64_2x1_t2x1_r1024_p2
  {0, 2, 2, 236, 1, 1},  // tsm_atb_f64_2x2_t2x2_r2048_p1_nt
  {0, 2, 4, 239, 2, 0},  // tsm_atb_f64_2x4_t2x4_r512_p2
  {0, 2, 8, 240, 2, 0},  // tsm_atb_f64_2x8_t2x8_r256_p2
  {0, 2, 16, 241, 2, 0},  // tsm_atb_f64_2x16_t2x8_r128_p2
  {0, 2, 32, 242, 2, 0},  // tsm_atb_f64_2x32_t2x8_r64_p2
  {0, 2, 64, 243, 2, 0},  // tsm_atb_f64_2x64_t2x8_r32_p2
  {0, 4, 1, 244, 2, 0},  // tsm_atb_f64_4x1_t4x1_r512_p2
  {0, 4, 2, 245, 2, 0},  // tsm_atb_f64_4x2_t4x2_r512_p2
  {0, 4, 4, 255, 1, 1},  // tsm_atb_f64_4x4_t4x4_r512_p2_nt
  {0, 4, 8, 260, 2, 0},  // tsm_atb_f64_4x8_t4x8_r256_p2
  {0, 4, 16, 261, 2, 0},  // tsm_atb_f64_4x16_t4x8_r128_p2
  {0, 4, 32, 271, 1, 1},  // tsm_atb_f64_4x32_t4x8_r64_p2_nt
  {0, 4, 64, 276, 2, 0},  // tsm_atb_f64_4x64_t4x8_r32_p2
  {0, 8, 1, 277, 2, 0},  // tsm_atb_f64_8x1_t8x1_r256_p2
  {0, 8, 2, 278, 2, 0},  // tsm_atb_f64_8x2_t8x2_r256_p2
  {0, 8, 4, 279, 2, 0},  // tsm_atb_f64_8x4_t8x4_r256_p2
  {0, 8, 8, 291, 1, 1},  // tsm_atb_f64_8x8_t8x8_r256_p2_nt
  {0, 8, 16, 296, 2, 0},  // tsm_atb_f64_8x16_t8x8_r128_p2
  {0, 8, 32, 297, 2, 0},  // tsm_atb_f64_8x32_t8x8_r64_p2
  {0, 8, 64, 298, 2, 0},  // tsm_atb_f64_8x64_t8x8_r32_p2
  {0, 16, 1, 299, 2, 0},  // tsm_atb_f64_16x1_t8x1_r128_p2
  {0, 16, 2, 300, 2, 0},  // tsm_atb_f64_16x2_t8x2_r128_p2
  {0, 16, 4, 301, 2, 0},  // tsm_atb_f64_16x4_t8x4_r128_p2
  {0, 16, 8, 302, 2, 0},  // tsm_atb_f64_16x8_t8x8_r128_p2
  {0, 16, 16, 318, 1, 1},  // tsm_atb_f64_16x16_t8x8_r128_p2_nt
  {0, 16, 32, 323, 2, 0},  // tsm_atb_f64_16x32_t8x8_r64_p2
  {0, 16, 64, 324, 2, 0},  // tsm_atb_f64_16x64_t8x8_r32_p2
  {0, 32, 1, 325, 2, 0},  // tsm_atb_f64_32x1_t8x1_r64_p2
  {0, 32, 2, 326, 2, 0},  // tsm_atb_f64_32x2_t8x2_r64_p2
  {0, 32, 4, 340, 1, 1},  // tsm_atb_f64_32x4_t8x4_r128_p2_nt
  {0, 32, 8, 341, 2, 0},  // tsm_atb_f64_32x8_t8x8_r64_p2
  {0, 32, 16, 342, 2, 0},  // tsm_atb_f64_32x16_t8x8_r64_p2
  {0, 32, 32, 352, 1, 1},  // tsm_atb_f64_32x32_t8x8_r64_p1_nt
  {0, 32, 64, 359, 2, 0},  // tsm_atb_f64_32x64_t8x8_r32_p2
  {0, 64, 1, 360, 2, 0},  // tsm_atb_f64_64x1_t8x1_r32_p2
  {0, 64, 2, 361, 2, 0},  // tsm_atb_f64_64x2_t8x2_r32_p2
  {0, 64, 4, 362, 2, 0},  // tsm_atb_f64_64x4_t8x4_r32_p2
  {0, 64, 8, 363, 2, 0},  // tsm_atb_f64_64x8_t8x8_r32_p2
  {0, 64, 16, 364, 2, 0},  // tsm_atb_f64_64x16_t8x8_r32_p2
  {0, 64, 32, 365, 2, 0},  // tsm_atb_f64_64x32_t8x8_r32_p2
  {0, 64, 64, 369, 2, 1},  // tsm_atb_f64_64x64_t8x8_r16_p2_nt
};

hipError_t launch_tsm_atb_batched_f64(const tsm::BatchArgs&, unsigned, hipStream_t);
const tsm::BatchClass* classes_tsm_atb_batched_f64(int* count);
hipError_t launch_tsm_atb_batched_f32(const tsm::BatchArgs&, unsigned, hipStream_t);
const tsm::BatchClass* classes_tsm_atb_batched_f32(int* count);
hipError_t launch_tsm_atb_batched_bf16(const tsm::BatchArgs&, unsigned, hipStream_t);
const tsm::BatchClass* classes_tsm_atb_batched_bf16(int* count);
static const tsm::BatchedKernel* batched_all() {
  static tsm::BatchedKernel all[3];
  static bool init = [] {
    all[0].name = "tsm_atb_batched_f64"; all[0].dtype = 0;
    all[0].classes = classes_tsm_atb_batched_f64(&all[0].n_classes);
    all[0].lds_bytes = 65536; all[0].launch = launch_tsm_atb_batched_f64;
    all[1].name = "tsm_atb_batched_f32"; all[1].dtype = 1;
    all[1].classes = classes_tsm_atb_batched_f32(&all[1].n_classes);
    all[1].lds_bytes = 65536; all[1].launch = launch_tsm_atb_batched_f32;
    all[2].name = "tsm_atb_batched_bf16"; all[2].dtype = 2;
    all[2].classes = classes_tsm_atb_batched_bf16(&all[2].n_classes);
    all[2].lds_bytes = 65536; all[2].launch = launch_tsm_atb_batched_bf16;
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
extern "C" const tsm::TunedEntry* tsm_gen_tuned_table(int* count) {
  *count = (int)(sizeof(kTuned) / sizeof(kTuned[0]));
  return kTuned;
}
