// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_batched_f64: 49 width classes, LDS 81920 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_batched_f64(tsm::BatchArgs ba) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[81920];
  __shared__ int s_flag;
  __shared__ unsigned s_next;
  for (int64_t i = blockIdx.x; i < ba.n_items; i = (int64_t)gridDim.x + s_next) {
    const tsm::BatchItem it = ba.items[i];
    switch (ba.probs[it.problem].cls) {
      case 0: tsm::batched_item_call<double, 1, 1, 1, 1, 2048, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_1x1_t1x1_r2048_p2_nt
      case 1: tsm::batched_item_call<double, 1, 2, 1, 2, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_1x2_t1x2_r1024_p2_nt
      case 2: tsm::batched_item_call<double, 1, 4, 1, 4, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_1x4_t1x4_r1024_p2_nt
      case 3: tsm::batched_item_call<double, 1, 8, 1, 8, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_1x8_t1x8_r512_p2_nt
      case 4: tsm::batched_item_call<double, 1, 16, 1, 8, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_1x16_t1x8_r256_p2_nt
      case 5: tsm::batched_item_call<double, 1, 32, 1, 8, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_1x32_t1x8_r128_p2_nt
      case 6: tsm::batched_item_call<double, 1, 64, 1, 8, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_1x64_t1x8_r64_p2_nt
      case 7: tsm::batched_item_call<double, 2, 1, 2, 1, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_2x1_t2x1_r1024_p2_nt
      case 8: tsm::batched_item_call<double, 2, 2, 2, 2, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_2x2_t2x2_r1024_p2_nt
      case 9: tsm::batched_item_call<double, 2, 4, 2, 4, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_2x4_t2x4_r512_p2_nt
      case 10: tsm::batched_item_call<double, 2, 8, 2, 8, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_2x8_t2x8_r512_p2_nt
      case 11: tsm::batched_item_call<double, 2, 16, 2, 8, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_2x16_t2x8_r256_p2_nt
      case 12: tsm::batched_item_call<double, 2, 32, 2, 8, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_2x32_t2x8_r128_p2_nt
      case 13: tsm::batched_item_call<double, 2, 64, 2, 8, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_2x64_t2x8_r64_p2_nt
      case 14: tsm::batched_item_call<double, 4, 1, 4, 1, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_4x1_t4x1_r1024_p2_nt
      case 15: tsm::batched_item_call<double, 4, 2, 4, 2, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_4x2_t4x2_r512_p2_nt
      case 16: tsm::batched_item_call<double, 4, 4, 4, 4, 512, 1, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_4x4_t4x4_r512_p1_nt
      case 17: tsm::batched_item_call<double, 4, 8, 4, 8, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_4x8_t4x8_r256_p2_nt
      case 18: tsm::batched_item_call<double, 4, 16, 4, 8, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_4x16_t4x8_r256_p2_nt
      case 19: tsm::batched_item_call<double, 4, 32, 4, 8, 128, 1, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_4x32_t4x8_r128_p1_nt
      case 20: tsm::batched_item_call<double, 4, 64, 4, 8, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_4x64_t4x8_r64_p2_nt
      case 21: tsm::batched_item_call<double, 8, 1, 8, 1, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_8x1_t8x1_r512_p2_nt
      case 22: tsm::batched_item_call<double, 8, 2, 8, 2, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_8x2_t8x2_r512_p2_nt
      case 23: tsm::batched_item_call<double, 8, 4, 8, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_8x4_t8x4_r256_p2_nt
      case 24: tsm::batched_item_call<double, 8, 8, 4, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_8x8_t4x4_r256_p2_nt
      case 25: tsm::batched_item_call<double, 8, 16, 8, 8, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_8x16_t8x8_r128_p2_nt
      case 26: tsm::batched_item_call<double, 8, 32, 8, 8, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_8x32_t8x8_r128_p2_nt
      case 27: tsm::batched_item_call<double, 8, 64, 8, 8, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_8x64_t8x8_r64_p2_nt
      case 28: tsm::batched_item_call<double, 16, 1, 8, 1, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_16x1_t8x1_r256_p2_nt
      case 29: tsm::batched_item_call<double, 16, 2, 8, 2, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_16x2_t8x2_r256_p2_nt
      case 30: tsm::batched_item_call<double, 16, 4, 8, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_16x4_t8x4_r256_p2_nt
      case 31: tsm::batched_item_call<double, 16, 8, 8, 8, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_16x8_t8x8_r128_p2_nt
      case 32: tsm::batched_item_call<double, 16, 16, 0, 0, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_16x16_mfma_r128_p2_nt
      case 33: tsm::batched_item_call<double, 16, 32, 8, 8, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_16x32_t8x8_r64_p2_nt
      case 34: tsm::batched_item_call<double, 16, 64, 0, 0, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_16x64_mfma_r64_p2_nt
      case 35: tsm::batched_item_call<double, 32, 1, 8, 1, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_32x1_t8x1_r128_p2_nt
      case 36: tsm::batched_item_call<double, 32, 2, 8, 2, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_32x2_t8x2_r128_p2_nt
      case 37: tsm::batched_item_call<double, 32, 4, 4, 4, 128, 4, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_32x4_t4x4_r128_p4_nt
      case 38: tsm::batched_item_call<double, 32, 8, 8, 8, 128, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_32x8_t8x8_r128_p2_nt
      case 39: tsm::batched_item_call<double, 32, 16, 8, 8, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_32x16_t8x8_r64_p2_nt
      case 40: tsm::batched_item_call<double, 32, 32, 0, 0, 64, 1, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_32x32_mfma_r64_p1_nt
      case 41: tsm::batched_item_call<double, 32, 64, 0, 0, 32, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_32x64_mfma_r32_p2_nt
      case 42: tsm::batched_item_call<double, 64, 1, 8, 1, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_64x1_t8x1_r64_p2_nt
      case 43: tsm::batched_item_call<double, 64, 2, 8, 2, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_64x2_t8x2_r64_p2_nt
      case 44: tsm::batched_item_call<double, 64, 4, 8, 4, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_64x4_t8x4_r64_p2_nt
      case 45: tsm::batched_item_call<double, 64, 8, 8, 8, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_64x8_t8x8_r64_p2_nt
      case 46: tsm::batched_item_call<double, 64, 16, 0, 0, 64, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_64x16_mfma_r64_p2_nt
      case 47: tsm::batched_item_call<double, 64, 32, 0, 0, 32, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_64x32_mfma_r32_p2_nt
      case 48: tsm::batched_item_call<double, 64, 64, 0, 0, 32, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_f64_64x64_mfma_r32_p2_nt
      default: tsm::batched_skip(ba, &s_next); break;
    }
  }
  tsm::batched_exit(ba);
}

static const tsm::BatchClass kClasses[] = {
  {1, 1, 2048},
  {1, 2, 1024},
  {1, 4, 1024},
  {1, 8, 512},
  {1, 16, 256},
  {1, 32, 128},
  {1, 64, 64},
  {2, 1, 1024},
  {2, 2, 1024},
  {2, 4, 512},
  {2, 8, 512},
  {2, 16, 256},
  {2, 32, 128},
  {2, 64, 64},
  {4, 1, 1024},
  {4, 2, 512},
  {4, 4, 512},
  {4, 8, 256},
  {4, 16, 256},
  {4, 32, 128},
  {4, 64, 64},
  {8, 1, 512},
  {8, 2, 512},
  {8, 4, 256},
  {8, 8, 256},
  {8, 16, 128},
  {8, 32, 128},
  {8, 64, 64},
  {16, 1, 256},
  {16, 2, 256},
  {16, 4, 256},
  {16, 8, 128},
  {16, 16, 128},
  {16, 32, 64},
  {16, 64, 64},
  {32, 1, 128},
  {32, 2, 128},
  {32, 4, 128},
  {32, 8, 128},
  {32, 16, 64},
  {32, 32, 64},
  {32, 64, 32},
  {64, 1, 64},
  {64, 2, 64},
  {64, 4, 64},
  {64, 8, 64},
  {64, 16, 64},
  {64, 32, 32},
  {64, 64, 32},
};

hipError_t launch_tsm_atb_batched_f64(const tsm::BatchArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_batched_f64, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

const tsm::BatchClass* classes_tsm_atb_batched_f64(int* count) {
  *count = 49;
  return kClasses;
}

int occupancy_tsm_atb_batched_f64() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(tsm_atb_batched_f64), 256, 0) !=
          hipSuccess || nb < 1)
    nb = 1;
  return nb;
}
