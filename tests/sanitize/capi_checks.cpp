// The argument checks of every C-ABI entry point (include/tsm/tsm.h) in a
// host build with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/sanitize/Makefile).  Every call here is rejected, or has nothing to
// do, before the library touches the HIP runtime, so it runs without a GPU.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "tsm/tsm.h"

static int failures = 0;
#define EXPECT(call, want)                                                              \
  do {                                                                                  \
    const int got_ = (int)(call);                                                       \
    if (got_ != (int)(want)) {                                                          \
      std::fprintf(stderr, "FAIL line %d: %s returned %d, expected %d\n", __LINE__, #call, got_, \
                   (int)(want));                                                        \
      ++failures;                                                                       \
    }                                                                                   \
  } while (0)

static void* P(uintptr_t off = 0) { return reinterpret_cast<void*>(0x10000 + off); }  // never dereferenced

int main() {
  const tsm_dtype F64 = TSM_F64, F32 = TSM_F32, BF16 = TSM_BF16;
  // tsm_atb
  EXPECT(tsm_atb(F64, F64, -1, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb(F64, F64, 10, 0, 8, P(), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb(F64, F64, 10, 8, 65, P(), 8, P(), 65, P(), 65, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb(F64, F64, 10, 8, 8, P(), 7, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb(F64, F64, 10, 8, 8, P(), 8, P(), 8, P(), 7, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb(F64, F64, 10, 8, 8, nullptr, 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb(F64, F64, 10, 8, 8, P(), 8, P(), 8, nullptr, 8, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb(F64, F64, 10, 8, 8, P(4), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_MISALIGNED);
  EXPECT(tsm_atb(F32, F64, 10, 8, 8, P(), 8, P(2), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_MISALIGNED);
  EXPECT(tsm_atb((tsm_dtype)7, F64, 10, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_UNSUPPORTED);
  EXPECT(tsm_atb(F64, BF16, 10, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_UNSUPPORTED);
  EXPECT(tsm_atb(F64, F64, (int64_t)1 << 62, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr),
         TSM_ERR_INVALID_ARGUMENT);
  // tsm_atb_ex: invalid kernel requests
  tsm_kernel_config bad;
  std::memset(&bad, 0, sizeof bad);
  bad.pf = 5;
  EXPECT(tsm_atb_ex(F64, F64, 10, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr, &bad), TSM_ERR_INVALID_ARGUMENT);
  bad.pf = -1;
  EXPECT(tsm_atb_ex(F64, F64, 10, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr, &bad), TSM_ERR_INVALID_ARGUMENT);
  std::memset(&bad, 0, sizeof bad);
  bad.engine = 9;
  EXPECT(tsm_atb_ex(F64, F64, 10, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr, &bad), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb_workspace_size(F64, 10, 8, 8, &bad) == 0, 1);
  EXPECT(tsm_atb_workspace_size(F64, -1, 8, 8, nullptr) == 0, 1);
  // tsm_atb_noalloc
  EXPECT(tsm_atb_noalloc(F64, F64, -1, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb_noalloc(F64, BF16, 10, 8, 8, P(), 8, P(), 8, P(), 8, 0, nullptr), TSM_ERR_UNSUPPORTED);
  EXPECT(tsm_atb_noalloc(F64, F64, 10, 8, 8, P(8), 8, P(4), 8, P(), 8, 0, nullptr), TSM_ERR_MISALIGNED);
  // tsm_ata
  EXPECT(tsm_ata(F64, F64, 10, 0, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_ata(F64, F64, 10, 8, P(), 7, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_ata(F64, F64, 10, 8, P(4), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_MISALIGNED);
  EXPECT(tsm_ata((tsm_dtype)9, F64, 10, 8, P(), 8, P(), 8, 0, nullptr, 0, nullptr), TSM_ERR_UNSUPPORTED);
  EXPECT(tsm_ata_workspace_size(F64, 10, 65, nullptr) == 0, 1);
  // tsm_tsmm
  EXPECT(tsm_tsmm(F64, F64, 10, 4, 8, P(), 3, P(), 8, P(), 8, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_tsmm(F64, F64, 10, 4, 8, P(), 4, nullptr, 8, P(), 8, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_tsmm(F64, F64, 10, 4, 8, P(), 4, P(), 8, P(4), 8, nullptr), TSM_ERR_MISALIGNED);
  EXPECT(tsm_tsmm(F64, BF16, 10, 4, 8, P(), 4, P(), 8, P(), 8, nullptr), TSM_ERR_UNSUPPORTED);
  EXPECT(tsm_tsmm(F64, F64, 0, 4, 8, P(), 4, P(), 8, P(), 8, nullptr), TSM_SUCCESS);
  // tsm_atb_backward
  EXPECT(tsm_atb_backward(F64, F64, 10, 0, 4, P(), 4, P(), 4, P(), 4, P(), 4, P(), 4, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb_backward(F64, F64, 10, 4, 4, P(), 4, P(), 4, nullptr, 4, P(), 4, P(), 4, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb_backward(F64, F64, 10, 4, 4, P(), 4, P(), 4, P(), 4, P(), 3, P(), 4, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb_backward(F64, F64, 10, 4, 4, P(), 4, P(), 4, P(), 4, P(), 4, P(2), 4, nullptr), TSM_ERR_MISALIGNED);
  EXPECT(tsm_atb_backward(F64, BF16, 10, 4, 4, P(), 4, P(), 4, P(), 4, P(), 4, P(), 4, nullptr), TSM_ERR_UNSUPPORTED);
  EXPECT(tsm_atb_backward(F64, F64, 10, 4, 4, P(), 4, P(), 4, P(), 4, nullptr, 4, nullptr, 4, nullptr), TSM_SUCCESS);
  EXPECT(tsm_atb_backward(F64, F64, 0, 4, 4, P(), 4, P(), 4, P(), 4, P(), 4, P(), 4, nullptr), TSM_SUCCESS);
  // batched
  tsm_problem pr[2];
  std::memset(pr, 0, sizeof pr);
  pr[0] = tsm_problem{P(), P(), P(), 10, 8, 8, 8, 8, 8};
  pr[1] = tsm_problem{P(), P(), P(), 10, 8, 8, 8, 8, 0};
  EXPECT(tsm_atb_batched(F64, F64, 2, pr, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb_batched(F64, F64, -1, pr, 0, nullptr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_atb_batched_workspace_size(F64, 2, pr) == 0, 1);
  EXPECT(tsm_atb_batched(F64, F64, 0, pr, 0, nullptr, 0, nullptr), TSM_SUCCESS);
  tsm_batched_plan plan = nullptr;
  EXPECT(tsm_batched_plan_create(F64, F64, 2, pr, 0, &plan), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_batched_plan_create(F64, F64, 1, pr, 0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_batched_plan_execute(nullptr, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_batched_plan_info(nullptr, nullptr, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_batched_plan_destroy(nullptr), TSM_ERR_INVALID_ARGUMENT);
  // fill and introspection
  EXPECT(tsm_fill_random(F64, (tsm_dist)5, 1, 1, P(), 4, 4, 4, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_fill_random(F64, TSM_DIST_UNIFORM, 1, 1, P(), 4, 4, 3, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_fill_random(F32, TSM_DIST_UNIFORM, 1, 1, P(2), 4, 4, 4, nullptr), TSM_ERR_MISALIGNED);
  EXPECT(tsm_fill_random(F64, TSM_DIST_UNIFORM, 1, 1, nullptr, 0, 4, 4, nullptr), TSM_SUCCESS);
  tsm_kernel_desc d;
  EXPECT(tsm_kernel_info(-1, &d), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_kernel_info(tsm_num_kernels(), &d), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_kernel_info(0, nullptr), TSM_ERR_INVALID_ARGUMENT);
  EXPECT(tsm_kernel_info(0, &d), TSM_SUCCESS);
  EXPECT(std::strlen(d.name) > 0, 1);
  for (int s = 0; s <= 7; ++s) EXPECT(tsm_strerror((tsm_status)s) != nullptr, 1);
  EXPECT(tsm_version(), 100);
  int idx = 0;
  int64_t grid = 0;
  EXPECT(tsm_atb_plan((tsm_dtype)7, 10, 8, 8, nullptr, &idx, &grid), TSM_ERR_UNSUPPORTED);
  EXPECT(tsm_atb_plan(F64, -1, 8, 8, nullptr, &idx, &grid), TSM_ERR_INVALID_ARGUMENT);
  if (failures) {
    std::fprintf(stderr, "%d C-ABI checks failed\n", failures);
    return 1;
  }
  std::printf("C-ABI checks passed\n");
  return 0;
}
