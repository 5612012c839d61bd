// Host-side view of the generated kernel set (shared by the generated tables
// and the C-ABI implementation).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
};

struct TunedEntry {
  int dtype, kc, mc;
  int kernel;      // index into kernel_table()
  int wgs_per_cu;  // persistent workgroups per compute unit
  int interleave;  // 1: tiles dealt round-robin to workgroups
};

}  // namespace tsm

// Records a HIP error for tsm_last_hip_error() and maps it to a status
// (TSM_SUCCESS for hipSuccess, else TSM_ERR_HIP).  Defined in tsm_capi.hip.
extern "C" int tsm_internal_hip_status(hipError_t e);

// Defined by the generated tsm_table.hip.
extern "C" const tsm::KernelEntry* tsm_gen_kernel_table(int* count);
extern "C" const tsm::TunedEntry* tsm_gen_tuned_table(int* count);
