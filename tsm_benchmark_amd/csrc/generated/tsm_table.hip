// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

hipError_t launch_tsm_atb_bf16_1x1_t1x1_r8192_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x4_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x8_t8x8_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x16_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x32_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_8x64_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x1_t8x1_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x2_t8x2_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x4_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x16_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x32_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_16x64_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x1_t8x1_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x2_t8x2_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x4_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x8_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x32_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_32x64_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x1_t8x1_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x2_t8x2_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x4_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x8_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_bf16_64x64_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_8x8_t8x8_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f32_16x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x16_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_16x64_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x1_t8x1_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x2_t8x2_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x4_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x8_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x16_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x32_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_32x64_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x1_t8x1_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x2_t8x2_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x4_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x8_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x16_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x32_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f32_64x64_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x1_t1x1_r2048_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x2_t1x2_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x4_t1x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x8_t1x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x16_t1x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x32_t1x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_1x64_t1x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x1_t2x1_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_2x2_t2x2_r1024_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_4x4_t4x4_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x4_t4x4_r512_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x8_t4x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x16_t4x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_4x32_t4x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
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
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x8_t8x8_r256_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_8x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x1_t8x1_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x2_t8x2_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x4_t8x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x8_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t2x2_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t2x2_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t4x4_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r128_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x16_t8x8_r128_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_16x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x1_t8x1_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x2_t8x2_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x4_t8x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x8_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x16_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t4x4_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r64_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x32_t8x8_r64_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_32x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x1_t8x1_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x2_t8x2_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x4_t8x4_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x8_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x16_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x32_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r16_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r16_p2(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r32_p1(const tsm::AtbArgs&, unsigned, hipStream_t);
hipError_t launch_tsm_atb_f64_64x64_t8x8_r32_p2(const tsm::AtbArgs&, unsigned, hipStream_t);

static const tsm::KernelEntry kTable[] = {
  {"tsm_atb_bf16_1x1_t1x1_r8192_p2", 2, 1, 1, 1, 1, 8192, 2, 65536, launch_tsm_atb_bf16_1x1_t1x1_r8192_p2},
  {"tsm_atb_bf16_1x2_t1x2_r4096_p2", 2, 1, 2, 1, 2, 4096, 2, 49152, launch_tsm_atb_bf16_1x2_t1x2_r4096_p2},
  {"tsm_atb_bf16_1x4_t1x4_r2048_p2", 2, 1, 4, 1, 4, 2048, 2, 40960, launch_tsm_atb_bf16_1x4_t1x4_r2048_p2},
  {"tsm_atb_bf16_1x8_t1x4_r1024_p2", 2, 1, 8, 1, 4, 1024, 2, 36864, launch_tsm_atb_bf16_1x8_t1x4_r1024_p2},
  {"tsm_atb_bf16_1x16_t1x4_r512_p2", 2, 1, 16, 1, 4, 512, 2, 34816, launch_tsm_atb_bf16_1x16_t1x4_r512_p2},
  {"tsm_atb_bf16_1x32_t1x4_r256_p2", 2, 1, 32, 1, 4, 256, 2, 33792, launch_tsm_atb_bf16_1x32_t1x4_r256_p2},
  {"tsm_atb_bf16_1x64_t1x4_r128_p2", 2, 1, 64, 1, 4, 128, 2, 33280, launch_tsm_atb_bf16_1x64_t1x4_r128_p2},
  {"tsm_atb_bf16_2x1_t2x1_r4096_p2", 2, 2, 1, 2, 1, 4096, 2, 49152, launch_tsm_atb_bf16_2x1_t2x1_r4096_p2},
  {"tsm_atb_bf16_2x2_t2x2_r4096_p2", 2, 2, 2, 2, 2, 4096, 2, 65536, launch_tsm_atb_bf16_2x2_t2x2_r4096_p2},
  {"tsm_atb_bf16_2x4_t2x4_r2048_p2", 2, 2, 4, 2, 4, 2048, 2, 49152, launch_tsm_atb_bf16_2x4_t2x4_r2048_p2},
  {"tsm_atb_bf16_2x8_t2x4_r1024_p2", 2, 2, 8, 2, 4, 1024, 2, 40960, launch_tsm_atb_bf16_2x8_t2x4_r1024_p2},
  {"tsm_atb_bf16_2x16_t2x4_r512_p2", 2, 2, 16, 2, 4, 512, 2, 36864, launch_tsm_atb_bf16_2x16_t2x4_r512_p2},
  {"tsm_atb_bf16_2x32_t2x4_r256_p2", 2, 2, 32, 2, 4, 256, 2, 34816, launch_tsm_atb_bf16_2x32_t2x4_r256_p2},
  {"tsm_atb_bf16_2x64_t2x4_r128_p2", 2, 2, 64, 2, 4, 128, 2, 33792, launch_tsm_atb_bf16_2x64_t2x4_r128_p2},
  {"tsm_atb_bf16_4x1_t4x1_r2048_p2", 2, 4, 1, 4, 1, 2048, 2, 40960, launch_tsm_atb_bf16_4x1_t4x1_r2048_p2},
  {"tsm_atb_bf16_4x2_t4x2_r2048_p2", 2, 4, 2, 4, 2, 2048, 2, 49152, launch_tsm_atb_bf16_4x2_t4x2_r2048_p2},
  {"tsm_atb_bf16_4x4_t4x4_r2048_p2", 2, 4, 4, 4, 4, 2048, 2, 65536, launch_tsm_atb_bf16_4x4_t4x4_r2048_p2},
  {"tsm_atb_bf16_4x8_t4x4_r1024_p2", 2, 4, 8, 4, 4, 1024, 2, 49152, launch_tsm_atb_bf16_4x8_t4x4_r1024_p2},
  {"tsm_atb_bf16_4x16_t4x4_r512_p2", 2, 4, 16, 4, 4, 512, 2, 40960, launch_tsm_atb_bf16_4x16_t4x4_r512_p2},
  {"tsm_atb_bf16_4x32_t4x4_r256_p2", 2, 4, 32, 4, 4, 256, 2, 36864, launch_tsm_atb_bf16_4x32_t4x4_r256_p2},
  {"tsm_atb_bf16_4x64_t4x4_r128_p2", 2, 4, 64, 4, 4, 128, 2, 34816, launch_tsm_atb_bf16_4x64_t4x4_r128_p2},
  {"tsm_atb_bf16_8x1_t8x1_r1024_p2", 2, 8, 1, 8, 1, 1024, 2, 36864, launch_tsm_atb_bf16_8x1_t8x1_r1024_p2},
  {"tsm_atb_bf16_8x2_t8x2_r1024_p2", 2, 8, 2, 8, 2, 1024, 2, 40960, launch_tsm_atb_bf16_8x2_t8x2_r1024_p2},
  {"tsm_atb_bf16_8x4_t8x4_r1024_p2", 2, 8, 4, 8, 4, 1024, 2, 49152, launch_tsm_atb_bf16_8x4_t8x4_r1024_p2},
  {"tsm_atb_bf16_8x8_t4x4_r512_p1", 2, 8, 8, 4, 4, 512, 1, 32768, launch_tsm_atb_bf16_8x8_t4x4_r512_p1},
  {"tsm_atb_bf16_8x8_t4x4_r512_p2", 2, 8, 8, 4, 4, 512, 2, 32768, launch_tsm_atb_bf16_8x8_t4x4_r512_p2},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p1", 2, 8, 8, 4, 4, 1024, 1, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p1},
  {"tsm_atb_bf16_8x8_t4x4_r1024_p2", 2, 8, 8, 4, 4, 1024, 2, 65536, launch_tsm_atb_bf16_8x8_t4x4_r1024_p2},
  {"tsm_atb_bf16_8x8_t8x4_r512_p1", 2, 8, 8, 8, 4, 512, 1, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p1},
  {"tsm_atb_bf16_8x8_t8x4_r512_p2", 2, 8, 8, 8, 4, 512, 2, 32768, launch_tsm_atb_bf16_8x8_t8x4_r512_p2},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p1", 2, 8, 8, 8, 4, 1024, 1, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p1},
  {"tsm_atb_bf16_8x8_t8x4_r1024_p2", 2, 8, 8, 8, 4, 1024, 2, 65536, launch_tsm_atb_bf16_8x8_t8x4_r1024_p2},
  {"tsm_atb_bf16_8x8_t8x8_r512_p1", 2, 8, 8, 8, 8, 512, 1, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p1},
  {"tsm_atb_bf16_8x8_t8x8_r512_p2", 2, 8, 8, 8, 8, 512, 2, 32768, launch_tsm_atb_bf16_8x8_t8x8_r512_p2},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p1", 2, 8, 8, 8, 8, 1024, 1, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p1},
  {"tsm_atb_bf16_8x8_t8x8_r1024_p2", 2, 8, 8, 8, 8, 1024, 2, 65536, launch_tsm_atb_bf16_8x8_t8x8_r1024_p2},
  {"tsm_atb_bf16_8x16_t8x4_r512_p2", 2, 8, 16, 8, 4, 512, 2, 49152, launch_tsm_atb_bf16_8x16_t8x4_r512_p2},
  {"tsm_atb_bf16_8x32_t8x4_r256_p2", 2, 8, 32, 8, 4, 256, 2, 40960, launch_tsm_atb_bf16_8x32_t8x4_r256_p2},
  {"tsm_atb_bf16_8x64_t8x4_r128_p2", 2, 8, 64, 8, 4, 128, 2, 36864, launch_tsm_atb_bf16_8x64_t8x4_r128_p2},
  {"tsm_atb_bf16_16x1_t8x1_r512_p2", 2, 16, 1, 8, 1, 512, 2, 34816, launch_tsm_atb_bf16_16x1_t8x1_r512_p2},
  {"tsm_atb_bf16_16x2_t8x2_r512_p2", 2, 16, 2, 8, 2, 512, 2, 36864, launch_tsm_atb_bf16_16x2_t8x2_r512_p2},
  {"tsm_atb_bf16_16x4_t8x4_r512_p2", 2, 16, 4, 8, 4, 512, 2, 40960, launch_tsm_atb_bf16_16x4_t8x4_r512_p2},
  {"tsm_atb_bf16_16x8_t8x4_r512_p2", 2, 16, 8, 8, 4, 512, 2, 49152, launch_tsm_atb_bf16_16x8_t8x4_r512_p2},
  {"tsm_atb_bf16_16x16_t4x4_r256_p1", 2, 16, 16, 4, 4, 256, 1, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p1},
  {"tsm_atb_bf16_16x16_t4x4_r256_p2", 2, 16, 16, 4, 4, 256, 2, 32768, launch_tsm_atb_bf16_16x16_t4x4_r256_p2},
  {"tsm_atb_bf16_16x16_t4x4_r512_p1", 2, 16, 16, 4, 4, 512, 1, 65536, launch_tsm_atb_bf16_16x16_t4x4_r512_p1},
  {"tsm_atb_bf16_16x16_t4x4_r512_p2", 2, 16, 16, 4, 4, 512, 2, 65536, launch_tsm_atb_bf16_16x16_t4x4_r512_p2},
  {"tsm_atb_bf16_16x16_t8x4_r256_p1", 2, 16, 16, 8, 4, 256, 1, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p1},
  {"tsm_atb_bf16_16x16_t8x4_r256_p2", 2, 16, 16, 8, 4, 256, 2, 32768, launch_tsm_atb_bf16_16x16_t8x4_r256_p2},
  {"tsm_atb_bf16_16x16_t8x4_r512_p1", 2, 16, 16, 8, 4, 512, 1, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p1},
  {"tsm_atb_bf16_16x16_t8x4_r512_p2", 2, 16, 16, 8, 4, 512, 2, 65536, launch_tsm_atb_bf16_16x16_t8x4_r512_p2},
  {"tsm_atb_bf16_16x16_t8x8_r256_p1", 2, 16, 16, 8, 8, 256, 1, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p1},
  {"tsm_atb_bf16_16x16_t8x8_r256_p2", 2, 16, 16, 8, 8, 256, 2, 32768, launch_tsm_atb_bf16_16x16_t8x8_r256_p2},
  {"tsm_atb_bf16_16x16_t8x8_r512_p1", 2, 16, 16, 8, 8, 512, 1, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p1},
  {"tsm_atb_bf16_16x16_t8x8_r512_p2", 2, 16, 16, 8, 8, 512, 2, 65536, launch_tsm_atb_bf16_16x16_t8x8_r512_p2},
  {"tsm_atb_bf16_16x32_t8x4_r256_p2", 2, 16, 32, 8, 4, 256, 2, 49152, launch_tsm_atb_bf16_16x32_t8x4_r256_p2},
  {"tsm_atb_bf16_16x64_t8x4_r128_p2", 2, 16, 64, 8, 4, 128, 2, 40960, launch_tsm_atb_bf16_16x64_t8x4_r128_p2},
  {"tsm_atb_bf16_32x1_t8x1_r256_p2", 2, 32, 1, 8, 1, 256, 2, 33792, launch_tsm_atb_bf16_32x1_t8x1_r256_p2},
  {"tsm_atb_bf16_32x2_t8x2_r256_p2", 2, 32, 2, 8, 2, 256, 2, 34816, launch_tsm_atb_bf16_32x2_t8x2_r256_p2},
  {"tsm_atb_bf16_32x4_t8x4_r256_p2", 2, 32, 4, 8, 4, 256, 2, 36864, launch_tsm_atb_bf16_32x4_t8x4_r256_p2},
  {"tsm_atb_bf16_32x8_t8x4_r256_p2", 2, 32, 8, 8, 4, 256, 2, 40960, launch_tsm_atb_bf16_32x8_t8x4_r256_p2},
  {"tsm_atb_bf16_32x16_t8x4_r256_p2", 2, 32, 16, 8, 4, 256, 2, 49152, launch_tsm_atb_bf16_32x16_t8x4_r256_p2},
  {"tsm_atb_bf16_32x32_t4x4_r128_p1", 2, 32, 32, 4, 4, 128, 1, 32768, launch_tsm_atb_bf16_32x32_t4x4_r128_p1},
  {"tsm_atb_bf16_32x32_t4x4_r128_p2", 2, 32, 32, 4, 4, 128, 2, 32768, launch_tsm_atb_bf16_32x32_t4x4_r128_p2},
  {"tsm_atb_bf16_32x32_t4x4_r256_p1", 2, 32, 32, 4, 4, 256, 1, 65536, launch_tsm_atb_bf16_32x32_t4x4_r256_p1},
  {"tsm_atb_bf16_32x32_t4x4_r256_p2", 2, 32, 32, 4, 4, 256, 2, 65536, launch_tsm_atb_bf16_32x32_t4x4_r256_p2},
  {"tsm_atb_bf16_32x32_t8x4_r128_p1", 2, 32, 32, 8, 4, 128, 1, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p1},
  {"tsm_atb_bf16_32x32_t8x4_r128_p2", 2, 32, 32, 8, 4, 128, 2, 32768, launch_tsm_atb_bf16_32x32_t8x4_r128_p2},
  {"tsm_atb_bf16_32x32_t8x4_r256_p1", 2, 32, 32, 8, 4, 256, 1, 65536, launch_tsm_atb_bf16_32x32_t8x4_r256_p1},
  {"tsm_atb_bf16_32x32_t8x4_r256_p2", 2, 32, 32, 8, 4, 256, 2, 65536, launch_tsm_atb_bf16_32x32_t8x4_r256_p2},
  {"tsm_atb_bf16_32x32_t8x8_r128_p1", 2, 32, 32, 8, 8, 128, 1, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p1},
  {"tsm_atb_bf16_32x32_t8x8_r128_p2", 2, 32, 32, 8, 8, 128, 2, 32768, launch_tsm_atb_bf16_32x32_t8x8_r128_p2},
  {"tsm_atb_bf16_32x32_t8x8_r256_p1", 2, 32, 32, 8, 8, 256, 1, 65536, launch_tsm_atb_bf16_32x32_t8x8_r256_p1},
  {"tsm_atb_bf16_32x32_t8x8_r256_p2", 2, 32, 32, 8, 8, 256, 2, 65536, launch_tsm_atb_bf16_32x32_t8x8_r256_p2},
  {"tsm_atb_bf16_32x64_t8x4_r128_p2", 2, 32, 64, 8, 4, 128, 2, 49152, launch_tsm_atb_bf16_32x64_t8x4_r128_p2},
  {"tsm_atb_bf16_64x1_t8x1_r128_p2", 2, 64, 1, 8, 1, 128, 2, 33280, launch_tsm_atb_bf16_64x1_t8x1_r128_p2},
  {"tsm_atb_bf16_64x2_t8x2_r128_p2", 2, 64, 2, 8, 2, 128, 2, 33792, launch_tsm_atb_bf16_64x2_t8x2_r128_p2},
  {"tsm_atb_bf16_64x4_t8x4_r128_p2", 2, 64, 4, 8, 4, 128, 2, 34816, launch_tsm_atb_bf16_64x4_t8x4_r128_p2},
  {"tsm_atb_bf16_64x8_t8x4_r128_p2", 2, 64, 8, 8, 4, 128, 2, 36864, launch_tsm_atb_bf16_64x8_t8x4_r128_p2},
  {"tsm_atb_bf16_64x16_t8x4_r128_p2", 2, 64, 16, 8, 4, 128, 2, 40960, launch_tsm_atb_bf16_64x16_t8x4_r128_p2},
  {"tsm_atb_bf16_64x32_t8x4_r128_p2", 2, 64, 32, 8, 4, 128, 2, 49152, launch_tsm_atb_bf16_64x32_t8x4_r128_p2},
  {"tsm_atb_bf16_64x64_t8x8_r128_p2", 2, 64, 64, 8, 8, 128, 2, 65536, launch_tsm_atb_bf16_64x64_t8x8_r128_p2},
  {"tsm_atb_f32_1x1_t1x1_r4096_p2", 1, 1, 1, 1, 1, 4096, 2, 65536, launch_tsm_atb_f32_1x1_t1x1_r4096_p2},
  {"tsm_atb_f32_1x2_t1x2_r2048_p2", 1, 1, 2, 1, 2, 2048, 2, 49152, launch_tsm_atb_f32_1x2_t1x2_r2048_p2},
  {"tsm_atb_f32_1x4_t1x4_r1024_p2", 1, 1, 4, 1, 4, 1024, 2, 40960, launch_tsm_atb_f32_1x4_t1x4_r1024_p2},
  {"tsm_atb_f32_1x8_t1x4_r512_p2", 1, 1, 8, 1, 4, 512, 2, 36864, launch_tsm_atb_f32_1x8_t1x4_r512_p2},
  {"tsm_atb_f32_1x16_t1x4_r256_p2", 1, 1, 16, 1, 4, 256, 2, 34816, launch_tsm_atb_f32_1x16_t1x4_r256_p2},
  {"tsm_atb_f32_1x32_t1x4_r128_p2", 1, 1, 32, 1, 4, 128, 2, 33792, launch_tsm_atb_f32_1x32_t1x4_r128_p2},
  {"tsm_atb_f32_1x64_t1x4_r64_p2", 1, 1, 64, 1, 4, 64, 2, 33280, launch_tsm_atb_f32_1x64_t1x4_r64_p2},
  {"tsm_atb_f32_2x1_t2x1_r2048_p2", 1, 2, 1, 2, 1, 2048, 2, 49152, launch_tsm_atb_f32_2x1_t2x1_r2048_p2},
  {"tsm_atb_f32_2x2_t2x2_r2048_p2", 1, 2, 2, 2, 2, 2048, 2, 65536, launch_tsm_atb_f32_2x2_t2x2_r2048_p2},
  {"tsm_atb_f32_2x4_t2x4_r1024_p2", 1, 2, 4, 2, 4, 1024, 2, 49152, launch_tsm_atb_f32_2x4_t2x4_r1024_p2},
  {"tsm_atb_f32_2x8_t2x4_r512_p2", 1, 2, 8, 2, 4, 512, 2, 40960, launch_tsm_atb_f32_2x8_t2x4_r512_p2},
  {"tsm_atb_f32_2x16_t2x4_r256_p2", 1, 2, 16, 2, 4, 256, 2, 36864, launch_tsm_atb_f32_2x16_t2x4_r256_p2},
  {"tsm_atb_f32_2x32_t2x4_r128_p2", 1, 2, 32, 2, 4, 128, 2, 34816, launch_tsm_atb_f32_2x32_t2x4_r128_p2},
  {"tsm_atb_f32_2x64_t2x4_r64_p2", 1, 2, 64, 2, 4, 64, 2, 33792, launch_tsm_atb_f32_2x64_t2x4_r64_p2},
  {"tsm_atb_f32_4x1_t4x1_r1024_p2", 1, 4, 1, 4, 1, 1024, 2, 40960, launch_tsm_atb_f32_4x1_t4x1_r1024_p2},
  {"tsm_atb_f32_4x2_t4x2_r1024_p2", 1, 4, 2, 4, 2, 1024, 2, 49152, launch_tsm_atb_f32_4x2_t4x2_r1024_p2},
  {"tsm_atb_f32_4x4_t4x4_r1024_p2", 1, 4, 4, 4, 4, 1024, 2, 65536, launch_tsm_atb_f32_4x4_t4x4_r1024_p2},
  {"tsm_atb_f32_4x8_t4x4_r512_p2", 1, 4, 8, 4, 4, 512, 2, 49152, launch_tsm_atb_f32_4x8_t4x4_r512_p2},
  {"tsm_atb_f32_4x16_t4x4_r256_p2", 1, 4, 16, 4, 4, 256, 2, 40960, launch_tsm_atb_f32_4x16_t4x4_r256_p2},
  {"tsm_atb_f32_4x32_t4x4_r128_p2", 1, 4, 32, 4, 4, 128, 2, 36864, launch_tsm_atb_f32_4x32_t4x4_r128_p2},
  {"tsm_atb_f32_4x64_t4x4_r64_p2", 1, 4, 64, 4, 4, 64, 2, 34816, launch_tsm_atb_f32_4x64_t4x4_r64_p2},
  {"tsm_atb_f32_8x1_t8x1_r512_p2", 1, 8, 1, 8, 1, 512, 2, 36864, launch_tsm_atb_f32_8x1_t8x1_r512_p2},
  {"tsm_atb_f32_8x2_t8x2_r512_p2", 1, 8, 2, 8, 2, 512, 2, 40960, launch_tsm_atb_f32_8x2_t8x2_r512_p2},
  {"tsm_atb_f32_8x4_t8x4_r512_p2", 1, 8, 4, 8, 4, 512, 2, 49152, launch_tsm_atb_f32_8x4_t8x4_r512_p2},
  {"tsm_atb_f32_8x8_t4x4_r256_p1", 1, 8, 8, 4, 4, 256, 1, 32768, launch_tsm_atb_f32_8x8_t4x4_r256_p1},
  {"tsm_atb_f32_8x8_t4x4_r256_p2", 1, 8, 8, 4, 4, 256, 2, 32768, launch_tsm_atb_f32_8x8_t4x4_r256_p2},
  {"tsm_atb_f32_8x8_t4x4_r512_p1", 1, 8, 8, 4, 4, 512, 1, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p1},
  {"tsm_atb_f32_8x8_t4x4_r512_p2", 1, 8, 8, 4, 4, 512, 2, 65536, launch_tsm_atb_f32_8x8_t4x4_r512_p2},
  {"tsm_atb_f32_8x8_t8x4_r256_p1", 1, 8, 8, 8, 4, 256, 1, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p1},
  {"tsm_atb_f32_8x8_t8x4_r256_p2", 1, 8, 8, 8, 4, 256, 2, 32768, launch_tsm_atb_f32_8x8_t8x4_r256_p2},
  {"tsm_atb_f32_8x8_t8x4_r512_p1", 1, 8, 8, 8, 4, 512, 1, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p1},
  {"tsm_atb_f32_8x8_t8x4_r512_p2", 1, 8, 8, 8, 4, 512, 2, 65536, launch_tsm_atb_f32_8x8_t8x4_r512_p2},
  {"tsm_atb_f32_8x8_t8x8_r256_p1", 1, 8, 8, 8, 8, 256, 1, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p1},
  {"tsm_atb_f32_8x8_t8x8_r256_p2", 1, 8, 8, 8, 8, 256, 2, 32768, launch_tsm_atb_f32_8x8_t8x8_r256_p2},
  {"tsm_atb_f32_8x8_t8x8_r512_p1", 1, 8, 8, 8, 8, 512, 1, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p1},
  {"tsm_atb_f32_8x8_t8x8_r512_p2", 1, 8, 8, 8, 8, 512, 2, 65536, launch_tsm_atb_f32_8x8_t8x8_r512_p2},
  {"tsm_atb_f32_8x16_t8x4_r256_p2", 1, 8, 16, 8, 4, 256, 2, 49152, launch_tsm_atb_f32_8x16_t8x4_r256_p2},
  {"tsm_atb_f32_8x32_t8x4_r128_p2", 1, 8, 32, 8, 4, 128, 2, 40960, launch_tsm_atb_f32_8x32_t8x4_r128_p2},
  {"tsm_atb_f32_8x64_t8x4_r64_p2", 1, 8, 64, 8, 4, 64, 2, 36864, launch_tsm_atb_f32_8x64_t8x4_r64_p2},
  {"tsm_atb_f32_16x1_t8x1_r256_p2", 1, 16, 1, 8, 1, 256, 2, 34816, launch_tsm_atb_f32_16x1_t8x1_r256_p2},
  {"tsm_atb_f32_16x2_t8x2_r256_p2", 1, 16, 2, 8, 2, 256, 2, 36864, launch_tsm_atb_f32_16x2_t8x2_r256_p2},
  {"tsm_atb_f32_16x4_t8x4_r256_p2", 1, 16, 4, 8, 4, 256, 2, 40960, launch_tsm_atb_f32_16x4_t8x4_r256_p2},
  {"tsm_atb_f32_16x8_t8x4_r256_p2", 1, 16, 8, 8, 4, 256, 2, 49152, launch_tsm_atb_f32_16x8_t8x4_r256_p2},
  {"tsm_atb_f32_16x16_t4x4_r128_p1", 1, 16, 16, 4, 4, 128, 1, 32768, launch_tsm_atb_f32_16x16_t4x4_r128_p1},
  {"tsm_atb_f32_16x16_t4x4_r128_p2", 1, 16, 16, 4, 4, 128, 2, 32768, launch_tsm_atb_f32_16x16_t4x4_r128_p2},
  {"tsm_atb_f32_16x16_t4x4_r256_p1", 1, 16, 16, 4, 4, 256, 1, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p1},
  {"tsm_atb_f32_16x16_t4x4_r256_p2", 1, 16, 16, 4, 4, 256, 2, 65536, launch_tsm_atb_f32_16x16_t4x4_r256_p2},
  {"tsm_atb_f32_16x16_t8x4_r128_p1", 1, 16, 16, 8, 4, 128, 1, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p1},
  {"tsm_atb_f32_16x16_t8x4_r128_p2", 1, 16, 16, 8, 4, 128, 2, 32768, launch_tsm_atb_f32_16x16_t8x4_r128_p2},
  {"tsm_atb_f32_16x16_t8x4_r256_p1", 1, 16, 16, 8, 4, 256, 1, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p1},
  {"tsm_atb_f32_16x16_t8x4_r256_p2", 1, 16, 16, 8, 4, 256, 2, 65536, launch_tsm_atb_f32_16x16_t8x4_r256_p2},
  {"tsm_atb_f32_16x16_t8x8_r128_p1", 1, 16, 16, 8, 8, 128, 1, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p1},
  {"tsm_atb_f32_16x16_t8x8_r128_p2", 1, 16, 16, 8, 8, 128, 2, 32768, launch_tsm_atb_f32_16x16_t8x8_r128_p2},
  {"tsm_atb_f32_16x16_t8x8_r256_p1", 1, 16, 16, 8, 8, 256, 1, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p1},
  {"tsm_atb_f32_16x16_t8x8_r256_p2", 1, 16, 16, 8, 8, 256, 2, 65536, launch_tsm_atb_f32_16x16_t8x8_r256_p2},
  {"tsm_atb_f32_16x32_t8x4_r128_p2", 1, 16, 32, 8, 4, 128, 2, 49152, launch_tsm_atb_f32_16x32_t8x4_r128_p2},
  {"tsm_atb_f32_16x64_t8x4_r64_p2", 1, 16, 64, 8, 4, 64, 2, 40960, launch_tsm_atb_f32_16x64_t8x4_r64_p2},
  {"tsm_atb_f32_32x1_t8x1_r128_p2", 1, 32, 1, 8, 1, 128, 2, 33792, launch_tsm_atb_f32_32x1_t8x1_r128_p2},
  {"tsm_atb_f32_32x2_t8x2_r128_p2", 1, 32, 2, 8, 2, 128, 2, 34816, launch_tsm_atb_f32_32x2_t8x2_r128_p2},
  {"tsm_atb_f32_32x4_t8x4_r128_p2", 1, 32, 4, 8, 4, 128, 2, 36864, launch_tsm_atb_f32_32x4_t8x4_r128_p2},
  {"tsm_atb_f32_32x8_t8x4_r128_p2", 1, 32, 8, 8, 4, 128, 2, 40960, launch_tsm_atb_f32_32x8_t8x4_r128_p2},
  {"tsm_atb_f32_32x16_t8x4_r128_p2", 1, 32, 16, 8, 4, 128, 2, 49152, launch_tsm_atb_f32_32x16_t8x4_r128_p2},
  {"tsm_atb_f32_32x32_t4x4_r64_p1", 1, 32, 32, 4, 4, 64, 1, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p1},
  {"tsm_atb_f32_32x32_t4x4_r64_p2", 1, 32, 32, 4, 4, 64, 2, 32768, launch_tsm_atb_f32_32x32_t4x4_r64_p2},
  {"tsm_atb_f32_32x32_t4x4_r128_p1", 1, 32, 32, 4, 4, 128, 1, 65536, launch_tsm_atb_f32_32x32_t4x4_r128_p1},
  {"tsm_atb_f32_32x32_t4x4_r128_p2", 1, 32, 32, 4, 4, 128, 2, 65536, launch_tsm_atb_f32_32x32_t4x4_r128_p2},
  {"tsm_atb_f32_32x32_t8x4_r64_p1", 1, 32, 32, 8, 4, 64, 1, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p1},
  {"tsm_atb_f32_32x32_t8x4_r64_p2", 1, 32, 32, 8, 4, 64, 2, 32768, launch_tsm_atb_f32_32x32_t8x4_r64_p2},
  {"tsm_atb_f32_32x32_t8x4_r128_p1", 1, 32, 32, 8, 4, 128, 1, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p1},
  {"tsm_atb_f32_32x32_t8x4_r128_p2", 1, 32, 32, 8, 4, 128, 2, 65536, launch_tsm_atb_f32_32x32_t8x4_r128_p2},
  {"tsm_atb_f32_32x32_t8x8_r64_p1", 1, 32, 32, 8, 8, 64, 1, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p1},
  {"tsm_atb_f32_32x32_t8x8_r64_p2", 1, 32, 32, 8, 8, 64, 2, 32768, launch_tsm_atb_f32_32x32_t8x8_r64_p2},
  {"tsm_atb_f32_32x32_t8x8_r128_p1", 1, 32, 32, 8, 8, 128, 1, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p1},
  {"tsm_atb_f32_32x32_t8x8_r128_p2", 1, 32, 32, 8, 8, 128, 2, 65536, launch_tsm_atb_f32_32x32_t8x8_r128_p2},
  {"tsm_atb_f32_32x64_t8x4_r64_p2", 1, 32, 64, 8, 4, 64, 2, 49152, launch_tsm_atb_f32_32x64_t8x4_r64_p2},
  {"tsm_atb_f32_64x1_t8x1_r64_p2", 1, 64, 1, 8, 1, 64, 2, 33280, launch_tsm_atb_f32_64x1_t8x1_r64_p2},
  {"tsm_atb_f32_64x2_t8x2_r64_p2", 1, 64, 2, 8, 2, 64, 2, 33792, launch_tsm_atb_f32_64x2_t8x2_r64_p2},
  {"tsm_atb_f32_64x4_t8x4_r64_p2", 1, 64, 4, 8, 4, 64, 2, 34816, launch_tsm_atb_f32_64x4_t8x4_r64_p2},
  {"tsm_atb_f32_64x8_t8x4_r64_p2", 1, 64, 8, 8, 4, 64, 2, 36864, launch_tsm_atb_f32_64x8_t8x4_r64_p2},
  {"tsm_atb_f32_64x16_t8x4_r64_p2", 1, 64, 16, 8, 4, 64, 2, 40960, launch_tsm_atb_f32_64x16_t8x4_r64_p2},
  {"tsm_atb_f32_64x32_t8x4_r64_p2", 1, 64, 32, 8, 4, 64, 2, 49152, launch_tsm_atb_f32_64x32_t8x4_r64_p2},
  {"tsm_atb_f32_64x64_t8x8_r64_p2", 1, 64, 64, 8, 8, 64, 2, 65536, launch_tsm_atb_f32_64x64_t8x8_r64_p2},
  {"tsm_atb_f64_1x1_t1x1_r1024_p1", 0, 1, 1, 1, 1, 1024, 1, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p1},
  {"tsm_atb_f64_1x1_t1x1_r1024_p2", 0, 1, 1, 1, 1, 1024, 2, 32768, launch_tsm_atb_f64_1x1_t1x1_r1024_p2},
  {"tsm_atb_f64_1x1_t1x1_r2048_p1", 0, 1, 1, 1, 1, 2048, 1, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p1},
  {"tsm_atb_f64_1x1_t1x1_r2048_p2", 0, 1, 1, 1, 1, 2048, 2, 65536, launch_tsm_atb_f64_1x1_t1x1_r2048_p2},
  {"tsm_atb_f64_1x2_t1x2_r1024_p2", 0, 1, 2, 1, 2, 1024, 2, 49152, launch_tsm_atb_f64_1x2_t1x2_r1024_p2},
  {"tsm_atb_f64_1x4_t1x4_r512_p2", 0, 1, 4, 1, 4, 512, 2, 40960, launch_tsm_atb_f64_1x4_t1x4_r512_p2},
  {"tsm_atb_f64_1x8_t1x8_r256_p2", 0, 1, 8, 1, 8, 256, 2, 36864, launch_tsm_atb_f64_1x8_t1x8_r256_p2},
  {"tsm_atb_f64_1x16_t1x8_r128_p2", 0, 1, 16, 1, 8, 128, 2, 34816, launch_tsm_atb_f64_1x16_t1x8_r128_p2},
  {"tsm_atb_f64_1x32_t1x8_r64_p2", 0, 1, 32, 1, 8, 64, 2, 33792, launch_tsm_atb_f64_1x32_t1x8_r64_p2},
  {"tsm_atb_f64_1x64_t1x8_r32_p2", 0, 1, 64, 1, 8, 32, 2, 33280, launch_tsm_atb_f64_1x64_t1x8_r32_p2},
  {"tsm_atb_f64_2x1_t2x1_r1024_p2", 0, 2, 1, 2, 1, 1024, 2, 49152, launch_tsm_atb_f64_2x1_t2x1_r1024_p2},
  {"tsm_atb_f64_2x2_t2x2_r512_p1", 0, 2, 2, 2, 2, 512, 1, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p1},
  {"tsm_atb_f64_2x2_t2x2_r512_p2", 0, 2, 2, 2, 2, 512, 2, 32768, launch_tsm_atb_f64_2x2_t2x2_r512_p2},
  {"tsm_atb_f64_2x2_t2x2_r1024_p1", 0, 2, 2, 2, 2, 1024, 1, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p1},
  {"tsm_atb_f64_2x2_t2x2_r1024_p2", 0, 2, 2, 2, 2, 1024, 2, 65536, launch_tsm_atb_f64_2x2_t2x2_r1024_p2},
  {"tsm_atb_f64_2x4_t2x4_r512_p2", 0, 2, 4, 2, 4, 512, 2, 49152, launch_tsm_atb_f64_2x4_t2x4_r512_p2},
  {"tsm_atb_f64_2x8_t2x8_r256_p2", 0, 2, 8, 2, 8, 256, 2, 40960, launch_tsm_atb_f64_2x8_t2x8_r256_p2},
  {"tsm_atb_f64_2x16_t2x8_r128_p2", 0, 2, 16, 2, 8, 128, 2, 36864, launch_tsm_atb_f64_2x16_t2x8_r128_p2},
  {"tsm_atb_f64_2x32_t2x8_r64_p2", 0, 2, 32, 2, 8, 64, 2, 34816, launch_tsm_atb_f64_2x32_t2x8_r64_p2},
  {"tsm_atb_f64_2x64_t2x8_r32_p2", 0, 2, 64, 2, 8, 32, 2, 33792, launch_tsm_atb_f64_2x64_t2x8_r32_p2},
  {"tsm_atb_f64_4x1_t4x1_r512_p2", 0, 4, 1, 4, 1, 512, 2, 40960, launch_tsm_atb_f64_4x1_t4x1_r512_p2},
  {"tsm_atb_f64_4x2_t4x2_r512_p2", 0, 4, 2, 4, 2, 512, 2, 49152, launch_tsm_atb_f64_4x2_t4x2_r512_p2},
  {"tsm_atb_f64_4x4_t2x2_r512_p1", 0, 4, 4, 2, 2, 512, 1, 65536, launch_tsm_atb_f64_4x4_t2x2_r512_p1},
  {"tsm_atb_f64_4x4_t2x2_r512_p2", 0, 4, 4, 2, 2, 512, 2, 65536, launch_tsm_atb_f64_4x4_t2x2_r512_p2},
  {"tsm_atb_f64_4x4_t4x4_r256_p1", 0, 4, 4, 4, 4, 256, 1, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p1},
  {"tsm_atb_f64_4x4_t4x4_r256_p2", 0, 4, 4, 4, 4, 256, 2, 32768, launch_tsm_atb_f64_4x4_t4x4_r256_p2},
  {"tsm_atb_f64_4x4_t4x4_r512_p1", 0, 4, 4, 4, 4, 512, 1, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p1},
  {"tsm_atb_f64_4x4_t4x4_r512_p2", 0, 4, 4, 4, 4, 512, 2, 65536, launch_tsm_atb_f64_4x4_t4x4_r512_p2},
  {"tsm_atb_f64_4x8_t4x8_r256_p2", 0, 4, 8, 4, 8, 256, 2, 49152, launch_tsm_atb_f64_4x8_t4x8_r256_p2},
  {"tsm_atb_f64_4x16_t4x8_r128_p2", 0, 4, 16, 4, 8, 128, 2, 40960, launch_tsm_atb_f64_4x16_t4x8_r128_p2},
  {"tsm_atb_f64_4x32_t4x4_r32_p1", 0, 4, 32, 4, 4, 32, 1, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p1},
  {"tsm_atb_f64_4x32_t4x4_r32_p2", 0, 4, 32, 4, 4, 32, 2, 18432, launch_tsm_atb_f64_4x32_t4x4_r32_p2},
  {"tsm_atb_f64_4x32_t4x4_r64_p1", 0, 4, 32, 4, 4, 64, 1, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p1},
  {"tsm_atb_f64_4x32_t4x4_r64_p2", 0, 4, 32, 4, 4, 64, 2, 36864, launch_tsm_atb_f64_4x32_t4x4_r64_p2},
  {"tsm_atb_f64_4x32_t4x8_r64_p1", 0, 4, 32, 4, 8, 64, 1, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p1},
  {"tsm_atb_f64_4x32_t4x8_r64_p2", 0, 4, 32, 4, 8, 64, 2, 36864, launch_tsm_atb_f64_4x32_t4x8_r64_p2},
  {"tsm_atb_f64_4x64_t4x8_r32_p2", 0, 4, 64, 4, 8, 32, 2, 34816, launch_tsm_atb_f64_4x64_t4x8_r32_p2},
  {"tsm_atb_f64_8x1_t8x1_r256_p2", 0, 8, 1, 8, 1, 256, 2, 36864, launch_tsm_atb_f64_8x1_t8x1_r256_p2},
  {"tsm_atb_f64_8x2_t8x2_r256_p2", 0, 8, 2, 8, 2, 256, 2, 40960, launch_tsm_atb_f64_8x2_t8x2_r256_p2},
  {"tsm_atb_f64_8x4_t8x4_r256_p2", 0, 8, 4, 8, 4, 256, 2, 49152, launch_tsm_atb_f64_8x4_t8x4_r256_p2},
  {"tsm_atb_f64_8x8_t2x2_r256_p1", 0, 8, 8, 2, 2, 256, 1, 65536, launch_tsm_atb_f64_8x8_t2x2_r256_p1},
  {"tsm_atb_f64_8x8_t2x2_r256_p2", 0, 8, 8, 2, 2, 256, 2, 65536, launch_tsm_atb_f64_8x8_t2x2_r256_p2},
  {"tsm_atb_f64_8x8_t4x4_r128_p1", 0, 8, 8, 4, 4, 128, 1, 32768, launch_tsm_atb_f64_8x8_t4x4_r128_p1},
  {"tsm_atb_f64_8x8_t4x4_r128_p2", 0, 8, 8, 4, 4, 128, 2, 32768, launch_tsm_atb_f64_8x8_t4x4_r128_p2},
  {"tsm_atb_f64_8x8_t4x4_r256_p1", 0, 8, 8, 4, 4, 256, 1, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p1},
  {"tsm_atb_f64_8x8_t4x4_r256_p2", 0, 8, 8, 4, 4, 256, 2, 65536, launch_tsm_atb_f64_8x8_t4x4_r256_p2},
  {"tsm_atb_f64_8x8_t8x8_r256_p1", 0, 8, 8, 8, 8, 256, 1, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p1},
  {"tsm_atb_f64_8x8_t8x8_r256_p2", 0, 8, 8, 8, 8, 256, 2, 65536, launch_tsm_atb_f64_8x8_t8x8_r256_p2},
  {"tsm_atb_f64_8x16_t8x8_r128_p2", 0, 8, 16, 8, 8, 128, 2, 49152, launch_tsm_atb_f64_8x16_t8x8_r128_p2},
  {"tsm_atb_f64_8x32_t8x8_r64_p2", 0, 8, 32, 8, 8, 64, 2, 40960, launch_tsm_atb_f64_8x32_t8x8_r64_p2},
  {"tsm_atb_f64_8x64_t8x8_r32_p2", 0, 8, 64, 8, 8, 32, 2, 36864, launch_tsm_atb_f64_8x64_t8x8_r32_p2},
  {"tsm_atb_f64_16x1_t8x1_r128_p2", 0, 16, 1, 8, 1, 128, 2, 34816, launch_tsm_atb_f64_16x1_t8x1_r128_p2},
  {"tsm_atb_f64_16x2_t8x2_r128_p2", 0, 16, 2, 8, 2, 128, 2, 36864, launch_tsm_atb_f64_16x2_t8x2_r128_p2},
  {"tsm_atb_f64_16x4_t8x4_r128_p2", 0, 16, 4, 8, 4, 128, 2, 40960, launch_tsm_atb_f64_16x4_t8x4_r128_p2},
  {"tsm_atb_f64_16x8_t8x8_r128_p2", 0, 16, 8, 8, 8, 128, 2, 49152, launch_tsm_atb_f64_16x8_t8x8_r128_p2},
  {"tsm_atb_f64_16x16_t2x2_r128_p1", 0, 16, 16, 2, 2, 128, 1, 65536, launch_tsm_atb_f64_16x16_t2x2_r128_p1},
  {"tsm_atb_f64_16x16_t2x2_r128_p2", 0, 16, 16, 2, 2, 128, 2, 65536, launch_tsm_atb_f64_16x16_t2x2_r128_p2},
  {"tsm_atb_f64_16x16_t4x4_r64_p1", 0, 16, 16, 4, 4, 64, 1, 32768, launch_tsm_atb_f64_16x16_t4x4_r64_p1},
  {"tsm_atb_f64_16x16_t4x4_r64_p2", 0, 16, 16, 4, 4, 64, 2, 32768, launch_tsm_atb_f64_16x16_t4x4_r64_p2},
  {"tsm_atb_f64_16x16_t4x4_r128_p1", 0, 16, 16, 4, 4, 128, 1, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p1},
  {"tsm_atb_f64_16x16_t4x4_r128_p2", 0, 16, 16, 4, 4, 128, 2, 65536, launch_tsm_atb_f64_16x16_t4x4_r128_p2},
  {"tsm_atb_f64_16x16_t8x8_r64_p1", 0, 16, 16, 8, 8, 64, 1, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p1},
  {"tsm_atb_f64_16x16_t8x8_r64_p2", 0, 16, 16, 8, 8, 64, 2, 32768, launch_tsm_atb_f64_16x16_t8x8_r64_p2},
  {"tsm_atb_f64_16x16_t8x8_r128_p1", 0, 16, 16, 8, 8, 128, 1, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p1},
  {"tsm_atb_f64_16x16_t8x8_r128_p2", 0, 16, 16, 8, 8, 128, 2, 65536, launch_tsm_atb_f64_16x16_t8x8_r128_p2},
  {"tsm_atb_f64_16x32_t8x8_r64_p2", 0, 16, 32, 8, 8, 64, 2, 49152, launch_tsm_atb_f64_16x32_t8x8_r64_p2},
  {"tsm_atb_f64_16x64_t8x8_r32_p2", 0, 16, 64, 8, 8, 32, 2, 40960, launch_tsm_atb_f64_16x64_t8x8_r32_p2},
  {"tsm_atb_f64_32x1_t8x1_r64_p2", 0, 32, 1, 8, 1, 64, 2, 33792, launch_tsm_atb_f64_32x1_t8x1_r64_p2},
  {"tsm_atb_f64_32x2_t8x2_r64_p2", 0, 32, 2, 8, 2, 64, 2, 34816, launch_tsm_atb_f64_32x2_t8x2_r64_p2},
  {"tsm_atb_f64_32x4_t4x4_r32_p1", 0, 32, 4, 4, 4, 32, 1, 18432, launch_tsm_atb_f64_32x4_t4x4_r32_p1},
  {"tsm_atb_f64_32x4_t4x4_r32_p2", 0, 32, 4, 4, 4, 32, 2, 18432, launch_tsm_atb_f64_32x4_t4x4_r32_p2},
  {"tsm_atb_f64_32x4_t4x4_r64_p1", 0, 32, 4, 4, 4, 64, 1, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p1},
  {"tsm_atb_f64_32x4_t4x4_r64_p2", 0, 32, 4, 4, 4, 64, 2, 36864, launch_tsm_atb_f64_32x4_t4x4_r64_p2},
  {"tsm_atb_f64_32x4_t8x4_r64_p1", 0, 32, 4, 8, 4, 64, 1, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p1},
  {"tsm_atb_f64_32x4_t8x4_r64_p2", 0, 32, 4, 8, 4, 64, 2, 36864, launch_tsm_atb_f64_32x4_t8x4_r64_p2},
  {"tsm_atb_f64_32x8_t8x8_r64_p2", 0, 32, 8, 8, 8, 64, 2, 40960, launch_tsm_atb_f64_32x8_t8x8_r64_p2},
  {"tsm_atb_f64_32x16_t8x8_r64_p2", 0, 32, 16, 8, 8, 64, 2, 49152, launch_tsm_atb_f64_32x16_t8x8_r64_p2},
  {"tsm_atb_f64_32x32_t4x4_r32_p1", 0, 32, 32, 4, 4, 32, 1, 32768, launch_tsm_atb_f64_32x32_t4x4_r32_p1},
  {"tsm_atb_f64_32x32_t4x4_r32_p2", 0, 32, 32, 4, 4, 32, 2, 32768, launch_tsm_atb_f64_32x32_t4x4_r32_p2},
  {"tsm_atb_f64_32x32_t4x4_r64_p1", 0, 32, 32, 4, 4, 64, 1, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p1},
  {"tsm_atb_f64_32x32_t4x4_r64_p2", 0, 32, 32, 4, 4, 64, 2, 65536, launch_tsm_atb_f64_32x32_t4x4_r64_p2},
  {"tsm_atb_f64_32x32_t8x8_r32_p1", 0, 32, 32, 8, 8, 32, 1, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p1},
  {"tsm_atb_f64_32x32_t8x8_r32_p2", 0, 32, 32, 8, 8, 32, 2, 32768, launch_tsm_atb_f64_32x32_t8x8_r32_p2},
  {"tsm_atb_f64_32x32_t8x8_r64_p1", 0, 32, 32, 8, 8, 64, 1, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p1},
  {"tsm_atb_f64_32x32_t8x8_r64_p2", 0, 32, 32, 8, 8, 64, 2, 65536, launch_tsm_atb_f64_32x32_t8x8_r64_p2},
  {"tsm_atb_f64_32x64_t8x8_r32_p2", 0, 32, 64, 8, 8, 32, 2, 49152, launch_tsm_atb_f64_32x64_t8x8_r32_p2},
  {"tsm_atb_f64_64x1_t8x1_r32_p2", 0, 64, 1, 8, 1, 32, 2, 33280, launch_tsm_atb_f64_64x1_t8x1_r32_p2},
  {"tsm_atb_f64_64x2_t8x2_r32_p2", 0, 64, 2, 8, 2, 32, 2, 33792, launch_tsm_atb_f64_64x2_t8x2_r32_p2},
  {"tsm_atb_f64_64x4_t8x4_r32_p2", 0, 64, 4, 8, 4, 32, 2, 34816, launch_tsm_atb_f64_64x4_t8x4_r32_p2},
  {"tsm_atb_f64_64x8_t8x8_r32_p2", 0, 64, 8, 8, 8, 32, 2, 36864, launch_tsm_atb_f64_64x8_t8x8_r32_p2},
  {"tsm_atb_f64_64x16_t8x8_r32_p2", 0, 64, 16, 8, 8, 32, 2, 40960, launch_tsm_atb_f64_64x16_t8x8_r32_p2},
  {"tsm_atb_f64_64x32_t8x8_r32_p2", 0, 64, 32, 8, 8, 32, 2, 49152, launch_tsm_atb_f64_64x32_t8x8_r32_p2},
  {"tsm_atb_f64_64x64_t8x8_r16_p1", 0, 64, 64, 8, 8, 16, 1, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p1},
  {"tsm_atb_f64_64x64_t8x8_r16_p2", 0, 64, 64, 8, 8, 16, 2, 32768, launch_tsm_atb_f64_64x64_t8x8_r16_p2},
  {"tsm_atb_f64_64x64_t8x8_r32_p1", 0, 64, 64, 8, 8, 32, 1, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p1},
  {"tsm_atb_f64_64x64_t8x8_r32_p2", 0, 64, 64, 8, 8, 32, 2, 65536, launch_tsm_atb_f64_64x64_t8x8_r32_p2},
};

static const tsm::TunedEntry kTuned[] = {
  {2, 1, 1, 0, 2},  // tsm_atb_bf16_1x1_t1x1_r8192_p2
  {2, 1, 2, 1, 2},  // tsm_atb_bf16_1x2_t1x2_r4096_p2
  {2, 1, 4, 2, 2},  // tsm_atb_bf16_1x4_t1x4_r2048_p2
  {2, 1, 8, 3, 2},  // tsm_atb_bf16_1x8_t1x4_r1024_p2
  {2, 1, 16, 4, 2},  // tsm_atb_bf16_1x16_t1x4_r512_p2
  {2, 1, 32, 5, 2},  // tsm_atb_bf16_1x32_t1x4_r256_p2
  {2, 1, 64, 6, 2},  // tsm_atb_bf16_1x64_t1x4_r128_p2
  {2, 2, 1, 7, 2},  // tsm_atb_bf16_2x1_t2x1_r4096_p2
  {2, 2, 2, 8, 2},  // tsm_atb_bf16_2x2_t2x2_r4096_p2
  {2, 2, 4, 9, 2},  // tsm_atb_bf16_2x4_t2x4_r2048_p2
  {2, 2, 8, 10, 2},  // tsm_atb_bf16_2x8_t2x4_r1024_p2
  {2, 2, 16, 11, 2},  // tsm_atb_bf16_2x16_t2x4_r512_p2
  {2, 2, 32, 12, 2},  // tsm_atb_bf16_2x32_t2x4_r256_p2
  {2, 2, 64, 13, 2},  // tsm_atb_bf16_2x64_t2x4_r128_p2
  {2, 4, 1, 14, 2},  // tsm_atb_bf16_4x1_t4x1_r2048_p2
  {2, 4, 2, 15, 2},  // tsm_atb_bf16_4x2_t4x2_r2048_p2
  {2, 4, 4, 16, 2},  // tsm_atb_bf16_4x4_t4x4_r2048_p2
  {2, 4, 8, 17, 2},  // tsm_atb_bf16_4x8_t4x4_r1024_p2
  {2, 4, 16, 18, 2},  // tsm_atb_bf16_4x16_t4x4_r512_p2
  {2, 4, 32, 19, 2},  // tsm_atb_bf16_4x32_t4x4_r256_p2
  {2, 4, 64, 20, 2},  // tsm_atb_bf16_4x64_t4x4_r128_p2
  {2, 8, 1, 21, 2},  // tsm_atb_bf16_8x1_t8x1_r1024_p2
  {2, 8, 2, 22, 2},  // tsm_atb_bf16_8x2_t8x2_r1024_p2
  {2, 8, 4, 23, 2},  // tsm_atb_bf16_8x4_t8x4_r1024_p2
  {2, 8, 8, 31, 2},  // tsm_atb_bf16_8x8_t8x4_r1024_p2
  {2, 8, 16, 36, 2},  // tsm_atb_bf16_8x16_t8x4_r512_p2
  {2, 8, 32, 37, 2},  // tsm_atb_bf16_8x32_t8x4_r256_p2
  {2, 8, 64, 38, 2},  // tsm_atb_bf16_8x64_t8x4_r128_p2
  {2, 16, 1, 39, 2},  // tsm_atb_bf16_16x1_t8x1_r512_p2
  {2, 16, 2, 40, 2},  // tsm_atb_bf16_16x2_t8x2_r512_p2
  {2, 16, 4, 41, 2},  // tsm_atb_bf16_16x4_t8x4_r512_p2
  {2, 16, 8, 42, 2},  // tsm_atb_bf16_16x8_t8x4_r512_p2
  {2, 16, 16, 50, 2},  // tsm_atb_bf16_16x16_t8x4_r512_p2
  {2, 16, 32, 55, 2},  // tsm_atb_bf16_16x32_t8x4_r256_p2
  {2, 16, 64, 56, 2},  // tsm_atb_bf16_16x64_t8x4_r128_p2
  {2, 32, 1, 57, 2},  // tsm_atb_bf16_32x1_t8x1_r256_p2
  {2, 32, 2, 58, 2},  // tsm_atb_bf16_32x2_t8x2_r256_p2
  {2, 32, 4, 59, 2},  // tsm_atb_bf16_32x4_t8x4_r256_p2
  {2, 32, 8, 60, 2},  // tsm_atb_bf16_32x8_t8x4_r256_p2
  {2, 32, 16, 61, 2},  // tsm_atb_bf16_32x16_t8x4_r256_p2
  {2, 32, 32, 69, 2},  // tsm_atb_bf16_32x32_t8x4_r256_p2
  {2, 32, 64, 74, 2},  // tsm_atb_bf16_32x64_t8x4_r128_p2
  {2, 64, 1, 75, 2},  // tsm_atb_bf16_64x1_t8x1_r128_p2
  {2, 64, 2, 76, 2},  // tsm_atb_bf16_64x2_t8x2_r128_p2
  {2, 64, 4, 77, 2},  // tsm_atb_bf16_64x4_t8x4_r128_p2
  {2, 64, 8, 78, 2},  // tsm_atb_bf16_64x8_t8x4_r128_p2
  {2, 64, 16, 79, 2},  // tsm_atb_bf16_64x16_t8x4_r128_p2
  {2, 64, 32, 80, 2},  // tsm_atb_bf16_64x32_t8x4_r128_p2
  {2, 64, 64, 81, 2},  // tsm_atb_bf16_64x64_t8x8_r128_p2
  {1, 1, 1, 82, 2},  // tsm_atb_f32_1x1_t1x1_r4096_p2
  {1, 1, 2, 83, 2},  // tsm_atb_f32_1x2_t1x2_r2048_p2
  {1, 1, 4, 84, 2},  // tsm_atb_f32_1x4_t1x4_r1024_p2
  {1, 1, 8, 85, 2},  // tsm_atb_f32_1x8_t1x4_r512_p2
  {1, 1, 16, 86, 2},  // tsm_atb_f32_1x16_t1x4_r256_p2
  {1, 1, 32, 87, 2},  // tsm_atb_f32_1x32_t1x4_r128_p2
  {1, 1, 64, 88, 2},  // tsm_atb_f32_1x64_t1x4_r64_p2
  {1, 2, 1, 89, 2},  // tsm_atb_f32_2x1_t2x1_r2048_p2
  {1, 2, 2, 90, 2},  // tsm_atb_f32_2x2_t2x2_r2048_p2
  {1, 2, 4, 91, 2},  // tsm_atb_f32_2x4_t2x4_r1024_p2
  {1, 2, 8, 92, 2},  // tsm_atb_f32_2x8_t2x4_r512_p2
  {1, 2, 16, 93, 2},  // tsm_atb_f32_2x16_t2x4_r256_p2
  {1, 2, 32, 94, 2},  // tsm_atb_f32_2x32_t2x4_r128_p2
  {1, 2, 64, 95, 2},  // tsm_atb_f32_2x64_t2x4_r64_p2
  {1, 4, 1, 96, 2},  // tsm_atb_f32_4x1_t4x1_r1024_p2
  {1, 4, 2, 97, 2},  // tsm_atb_f32_4x2_t4x2_r1024_p2
  {1, 4, 4, 98, 2},  // tsm_atb_f32_4x4_t4x4_r1024_p2
  {1, 4, 8, 99, 2},  // tsm_atb_f32_4x8_t4x4_r512_p2
  {1, 4, 16, 100, 2},  // tsm_atb_f32_4x16_t4x4_r256_p2
  {1, 4, 32, 101, 2},  // tsm_atb_f32_4x32_t4x4_r128_p2
  {1, 4, 64, 102, 2},  // tsm_atb_f32_4x64_t4x4_r64_p2
  {1, 8, 1, 103, 2},  // tsm_atb_f32_8x1_t8x1_r512_p2
  {1, 8, 2, 104, 2},  // tsm_atb_f32_8x2_t8x2_r512_p2
  {1, 8, 4, 105, 2},  // tsm_atb_f32_8x4_t8x4_r512_p2
  {1, 8, 8, 113, 2},  // tsm_atb_f32_8x8_t8x4_r512_p2
  {1, 8, 16, 118, 2},  // tsm_atb_f32_8x16_t8x4_r256_p2
  {1, 8, 32, 119, 2},  // tsm_atb_f32_8x32_t8x4_r128_p2
  {1, 8, 64, 120, 2},  // tsm_atb_f32_8x64_t8x4_r64_p2
  {1, 16, 1, 121, 2},  // tsm_atb_f32_16x1_t8x1_r256_p2
  {1, 16, 2, 122, 2},  // tsm_atb_f32_16x2_t8x2_r256_p2
  {1, 16, 4, 123, 2},  // tsm_atb_f32_16x4_t8x4_r256_p2
  {1, 16, 8, 124, 2},  // tsm_atb_f32_16x8_t8x4_r256_p2
  {1, 16, 16, 132, 2},  // tsm_atb_f32_16x16_t8x4_r256_p2
  {1, 16, 32, 137, 2},  // tsm_atb_f32_16x32_t8x4_r128_p2
  {1, 16, 64, 138, 2},  // tsm_atb_f32_16x64_t8x4_r64_p2
  {1, 32, 1, 139, 2},  // tsm_atb_f32_32x1_t8x1_r128_p2
  {1, 32, 2, 140, 2},  // tsm_atb_f32_32x2_t8x2_r128_p2
  {1, 32, 4, 141, 2},  // tsm_atb_f32_32x4_t8x4_r128_p2
  {1, 32, 8, 142, 2},  // tsm_atb_f32_32x8_t8x4_r128_p2
  {1, 32, 16, 143, 2},  // tsm_atb_f32_32x16_t8x4_r128_p2
  {1, 32, 32, 151, 2},  // tsm_atb_f32_32x32_t8x4_r128_p2
  {1, 32, 64, 156, 2},  // tsm_atb_f32_32x64_t8x4_r64_p2
  {1, 64, 1, 157, 2},  // tsm_atb_f32_64x1_t8x1_r64_p2
  {1, 64, 2, 158, 2},  // tsm_atb_f32_64x2_t8x2_r64_p2
  {1, 64, 4, 159, 2},  // tsm_atb_f32_64x4_t8x4_r64_p2
  {1, 64, 8, 160, 2},  // tsm_atb_f32_64x8_t8x4_r64_p2
  {1, 64, 16, 161, 2},  // tsm_atb_f32_64x16_t8x4_r64_p2
  {1, 64, 32, 162, 2},  // tsm_atb_f32_64x32_t8x4_r64_p2
  {1, 64, 64, 163, 2},  // tsm_atb_f32_64x64_t8x8_r64_p2
  {0, 1, 1, 167, 2},  // tsm_atb_f64_1x1_t1x1_r2048_p2
  {0, 1, 2, 168, 2},  // tsm_atb_f64_1x2_t1x2_r1024_p2
  {0, 1, 4, 169, 2},  // tsm_atb_f64_1x4_t1x4_r512_p2
  {0, 1, 8, 170, 2},  // tsm_atb_f64_1x8_t1x8_r256_p2
  {0, 1, 16, 171, 2},  // tsm_atb_f64_1x16_t1x8_r128_p2
  {0, 1, 32, 172, 2},  // tsm_atb_f64_1x32_t1x8_r64_p2
  {0, 1, 64, 173, 2},  // tsm_atb_f64_1x64_t1x8_r32_p2
  {0, 2, 1, 174, 2},  // tsm_atb_f64_2x1_t2x1_r1024_p2
  {0, 2, 2, 178, 2},  // tsm_atb_f64_2x2_t2x2_r1024_p2
  {0, 2, 4, 179, 2},  // tsm_atb_f64_2x4_t2x4_r512_p2
  {0, 2, 8, 180, 2},  // tsm_atb_f64_2x8_t2x8_r256_p2
  {0, 2, 16, 181, 2},  // tsm_atb_f64_2x16_t2x8_r128_p2
  {0, 2, 32, 182, 2},  // tsm_atb_f64_2x32_t2x8_r64_p2
  {0, 2, 64, 183, 2},  // tsm_atb_f64_2x64_t2x8_r32_p2
  {0, 4, 1, 184, 2},  // tsm_atb_f64_4x1_t4x1_r512_p2
  {0, 4, 2, 185, 2},  // tsm_atb_f64_4x2_t4x2_r512_p2
  {0, 4, 4, 191, 2},  // tsm_atb_f64_4x4_t4x4_r512_p2
  {0, 4, 8, 192, 2},  // tsm_atb_f64_4x8_t4x8_r256_p2
  {0, 4, 16, 193, 2},  // tsm_atb_f64_4x16_t4x8_r128_p2
  {0, 4, 32, 199, 2},  // tsm_atb_f64_4x32_t4x8_r64_p2
  {0, 4, 64, 200, 2},  // tsm_atb_f64_4x64_t4x8_r32_p2
  {0, 8, 1, 201, 2},  // tsm_atb_f64_8x1_t8x1_r256_p2
  {0, 8, 2, 202, 2},  // tsm_atb_f64_8x2_t8x2_r256_p2
  {0, 8, 4, 203, 2},  // tsm_atb_f64_8x4_t8x4_r256_p2
  {0, 8, 8, 211, 2},  // tsm_atb_f64_8x8_t8x8_r256_p2
  {0, 8, 16, 212, 2},  // tsm_atb_f64_8x16_t8x8_r128_p2
  {0, 8, 32, 213, 2},  // tsm_atb_f64_8x32_t8x8_r64_p2
  {0, 8, 64, 214, 2},  // tsm_atb_f64_8x64_t8x8_r32_p2
  {0, 16, 1, 215, 2},  // tsm_atb_f64_16x1_t8x1_r128_p2
  {0, 16, 2, 216, 2},  // tsm_atb_f64_16x2_t8x2_r128_p2
  {0, 16, 4, 217, 2},  // tsm_atb_f64_16x4_t8x4_r128_p2
  {0, 16, 8, 218, 2},  // tsm_atb_f64_16x8_t8x8_r128_p2
  {0, 16, 16, 228, 2},  // tsm_atb_f64_16x16_t8x8_r128_p2
  {0, 16, 32, 229, 2},  // tsm_atb_f64_16x32_t8x8_r64_p2
  {0, 16, 64, 230, 2},  // tsm_atb_f64_16x64_t8x8_r32_p2
  {0, 32, 1, 231, 2},  // tsm_atb_f64_32x1_t8x1_r64_p2
  {0, 32, 2, 232, 2},  // tsm_atb_f64_32x2_t8x2_r64_p2
  {0, 32, 4, 238, 2},  // tsm_atb_f64_32x4_t8x4_r64_p2
  {0, 32, 8, 239, 2},  // tsm_atb_f64_32x8_t8x8_r64_p2
  {0, 32, 16, 240, 2},  // tsm_atb_f64_32x16_t8x8_r64_p2
  {0, 32, 32, 248, 2},  // tsm_atb_f64_32x32_t8x8_r64_p2
  {0, 32, 64, 249, 2},  // tsm_atb_f64_32x64_t8x8_r32_p2
  {0, 64, 1, 250, 2},  // tsm_atb_f64_64x1_t8x1_r32_p2
  {0, 64, 2, 251, 2},  // tsm_atb_f64_64x2_t8x2_r32_p2
  {0, 64, 4, 252, 2},  // tsm_atb_f64_64x4_t8x4_r32_p2
  {0, 64, 8, 253, 2},  // tsm_atb_f64_64x8_t8x8_r32_p2
  {0, 64, 16, 254, 2},  // tsm_atb_f64_64x16_t8x8_r32_p2
  {0, 64, 32, 255, 2},  // tsm_atb_f64_64x32_t8x8_r32_p2
  {0, 64, 64, 259, 2},  // tsm_atb_f64_64x64_t8x8_r32_p2
};

extern "C" const tsm::KernelEntry* tsm_gen_kernel_table(int* count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
extern "C" const tsm::TunedEntry* tsm_gen_tuned_table(int* count) {
  *count = (int)(sizeof(kTuned) / sizeof(kTuned[0]));
  return kTuned;
}
