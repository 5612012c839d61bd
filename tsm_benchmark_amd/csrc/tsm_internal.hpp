// Host-side view of the generated kernel set (shared by the generated tables
// and the C-ABI implementation).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsm/tsm.h"

#include "tsm_batched.hpp"
#include "tsm_device.hpp"

namespace tsm {

using LaunchFn = hipError_t (*)(const AtbArgs& args, unsigned grid, hipStream_t stream);

struct KernelEntry {
  const char* name;
  int dtype;  // tsm_dtype
  int kc, mc, tm, tn, rows, pf;
  int nt;  // 1: non-temporal loads
  int lds_bytes;
  LaunchFn launch;
  int engine;  // 1 VALU register tiles, 2 matrix cores through LDS, 3 direct matrix cores
};

struct TunedEntry {
  int dtype, kc, mc;
  int kernel;      // index into kernel_table()
  int wgs_per_cu;  // persistent workgroups per compute unit
  int interleave;  // 1: tiles dealt round-robin to workgroups
  // Kernel (and tile order) for inputs the default cannot read: a direct-
  // engine default needs dense rows; for other defaults this is the default.
  int fallback;
  int fallback_interleave;
};

// Width class handled by a batched kernel: the class's rows per tile fix the
// problem's tiling on the host.
struct BatchClass {
  int kc, mc, rows;
};
using BatchLaunchFn = hipError_t (*)(const BatchArgs& args, unsigned grid, hipStream_t stream);
struct BatchedKernel {
  const char* name;
  int dtype;
  const BatchClass* classes;
  int n_classes;
  int lds_bytes;
  BatchLaunchFn launch;
  int (*occupancy)();  // workgroups of the kernel resident per CU
};

}  // namespace tsm

// Records a HIP error for tsm_last_hip_error() and maps it to a status
// (TSM_SUCCESS for hipSuccess, else TSM_ERR_HIP).  Defined in tsm_capi.hip.
extern "C" int tsm_internal_hip_status(hipError_t e);
// Argument checks of tsm_atb, the device's compute-unit count and the
// library-owned per-(device, stream) workspace (all in tsm_capi.hip).
extern "C" tsm_status tsm_internal_check_atb(tsm_dtype in_type, tsm_dtype out_type, int64_t n,
                                             int k, int m, const void* A, int64_t lda,
                                             const void* B, int64_t ldb, const void* C,
                                             int64_t ldc);
extern "C" tsm_status tsm_internal_device_cus(int* cus);
extern "C" tsm_status tsm_internal_pooled_workspace(size_t bytes, void* stream, void** out);

// Defined by the generated tsm_table.hip.
extern "C" const tsm::KernelEntry* tsm_gen_kernel_table(int* count);
extern "C" const tsm::TunedEntry* tsm_gen_tuned_table(int* count);
// A^T A kernels (kc == mc) and the default per (dtype, class).
extern "C" const tsm::KernelEntry* tsm_gen_sym_table(int* count);
extern "C" const tsm::TunedEntry* tsm_gen_sym_tuned_table(int* count);
extern "C" const tsm::BatchedKernel* tsm_gen_batched_table(int* count);
