// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.
// Regenerate with: python -m tsm_benchmark_amd.codegen --library

#include "../tsm_internal.hpp"

// tsm_atb_batched_bf16: 36 width classes, LDS 65536 B
extern "C" __global__ void __launch_bounds__(256) tsm_atb_batched_bf16(tsm::BatchArgs ba) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[65536];
  __shared__ int s_flag;
  __shared__ unsigned s_next;
  for (int64_t i = blockIdx.x; i < ba.n_items; i = (int64_t)gridDim.x + s_next) {
    const tsm::BatchItem it = ba.items[i];
    switch (ba.probs[it.problem].cls) {
      case 0: tsm::batched_item_call<tsm::bf16, 1, 1, 1, 1, 4096, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_1x1_t1x1_r4096_p2_nt
      case 1: tsm::batched_item_call<tsm::bf16, 1, 2, 1, 2, 4096, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_1x2_t1x2_r4096_p2_nt
      case 2: tsm::batched_item_call<tsm::bf16, 1, 4, 1, 4, 2048, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_1x4_t1x4_r2048_p2_nt
      case 3: tsm::batched_item_call<tsm::bf16, 1, 8, 1, 4, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_1x8_t1x4_r1024_p2_nt
      case 4: tsm::batched_item_call<tsm::bf16, 1, 16, 1, 4, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_1x16_t1x4_r512_p2_nt
      case 5: tsm::batched_item_call<tsm::bf16, 1, 32, 1, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_1x32_t1x4_r256_p2_nt
      case 6: tsm::batched_item_call<tsm::bf16, 2, 1, 2, 1, 4096, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_2x1_t2x1_r4096_p2_nt
      case 7: tsm::batched_item_call<tsm::bf16, 2, 2, 2, 2, 4096, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_2x2_t2x2_r4096_p2_nt
      case 8: tsm::batched_item_call<tsm::bf16, 2, 4, 2, 4, 2048, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_2x4_t2x4_r2048_p2_nt
      case 9: tsm::batched_item_call<tsm::bf16, 2, 8, 2, 4, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_2x8_t2x4_r1024_p2_nt
      case 10: tsm::batched_item_call<tsm::bf16, 2, 16, 2, 4, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_2x16_t2x4_r512_p2_nt
      case 11: tsm::batched_item_call<tsm::bf16, 2, 32, 2, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_2x32_t2x4_r256_p2_nt
      case 12: tsm::batched_item_call<tsm::bf16, 4, 1, 4, 1, 2048, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_4x1_t4x1_r2048_p2_nt
      case 13: tsm::batched_item_call<tsm::bf16, 4, 2, 4, 2, 2048, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_4x2_t4x2_r2048_p2_nt
      case 14: tsm::batched_item_call<tsm::bf16, 4, 4, 4, 4, 2048, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_4x4_t4x4_r2048_p2_nt
      case 15: tsm::batched_item_call<tsm::bf16, 4, 8, 4, 4, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_4x8_t4x4_r1024_p2_nt
      case 16: tsm::batched_item_call<tsm::bf16, 4, 16, 4, 4, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_4x16_t4x4_r512_p2_nt
      case 17: tsm::batched_item_call<tsm::bf16, 4, 32, 4, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_4x32_t4x4_r256_p2_nt
      case 18: tsm::batched_item_call<tsm::bf16, 8, 1, 8, 1, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_8x1_t8x1_r1024_p2_nt
      case 19: tsm::batched_item_call<tsm::bf16, 8, 2, 8, 2, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_8x2_t8x2_r1024_p2_nt
      case 20: tsm::batched_item_call<tsm::bf16, 8, 4, 8, 4, 1024, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_8x4_t8x4_r1024_p2_nt
      case 21: tsm::batched_item_call<tsm::bf16, 8, 8, 8, 4, 1024, 1, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_8x8_t8x4_r1024_p1_nt
      case 22: tsm::batched_item_call<tsm::bf16, 8, 16, 8, 4, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_8x16_t8x4_r512_p2_nt
      case 23: tsm::batched_item_call<tsm::bf16, 8, 32, 8, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_8x32_t8x4_r256_p2_nt
      case 24: tsm::batched_item_call<tsm::bf16, 16, 1, 8, 1, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_16x1_t8x1_r512_p2_nt
      case 25: tsm::batched_item_call<tsm::bf16, 16, 2, 8, 2, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_16x2_t8x2_r512_p2_nt
      case 26: tsm::batched_item_call<tsm::bf16, 16, 4, 8, 4, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_16x4_t8x4_r512_p2_nt
      case 27: tsm::batched_item_call<tsm::bf16, 16, 8, 8, 4, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_16x8_t8x4_r512_p2_nt
      case 28: tsm::batched_item_call<tsm::bf16, 16, 16, 0, 0, 512, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_16x16_mfma_r512_p2_nt
      case 29: tsm::batched_item_call<tsm::bf16, 16, 32, 0, 0, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_16x32_mfma_r256_p2_nt
      case 30: tsm::batched_item_call<tsm::bf16, 32, 1, 8, 1, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_32x1_t8x1_r256_p2_nt
      case 31: tsm::batched_item_call<tsm::bf16, 32, 2, 8, 2, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_32x2_t8x2_r256_p2_nt
      case 32: tsm::batched_item_call<tsm::bf16, 32, 4, 8, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_32x4_t8x4_r256_p2_nt
      case 33: tsm::batched_item_call<tsm::bf16, 32, 8, 8, 4, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_32x8_t8x4_r256_p2_nt
      case 34: tsm::batched_item_call<tsm::bf16, 32, 16, 0, 0, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_32x16_mfma_r256_p2_nt
      case 35: tsm::batched_item_call<tsm::bf16, 32, 32, 0, 0, 256, 2, 1>(ba, it, smem, &s_flag, &s_next); break;  // tsm_atb_bf16_32x32_mfma_r256_p2_nt
      default: tsm::batched_skip(ba, &s_next); break;
    }
  }
  tsm::batched_exit(ba);
}

static const tsm::BatchClass kClasses[] = {
  {1, 1, 4096},
  {1, 2, 4096},
  {1, 4, 2048},
  {1, 8, 1024},
  {1, 16, 512},
  {1, 32, 256},
  {2, 1, 4096},
  {2, 2, 4096},
  {2, 4, 2048},
  {2, 8, 1024},
  {2, 16, 512},
  {2, 32, 256},
  {4, 1, 2048},
  {4, 2, 2048},
  {4, 4, 2048},
  {4, 8, 1024},
  {4, 16, 512},
  {4, 32, 256},
  {8, 1, 1024},
  {8, 2, 1024},
  {8, 4, 1024},
  {8, 8, 1024},
  {8, 16, 512},
  {8, 32, 256},
  {16, 1, 512},
  {16, 2, 512},
  {16, 4, 512},
  {16, 8, 512},
  {16, 16, 512},
  {16, 32, 256},
  {32, 1, 256},
  {32, 2, 256},
  {32, 4, 256},
  {32, 8, 256},
  {32, 16, 256},
  {32, 32, 256},
};

hipError_t launch_tsm_atb_batched_bf16(const tsm::BatchArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(tsm_atb_batched_bf16, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

const tsm::BatchClass* classes_tsm_atb_batched_bf16(int* count) {
  *count = 36;
  return kClasses;
}

int occupancy_tsm_atb_batched_bf16() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(tsm_atb_batched_bf16), 256, 0) !=
          hipSuccess || nb < 1)
    nb = 1;
  return nb;
}
