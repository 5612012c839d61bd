// Shader-clock timeline of a run of back-to-back tsm_atb launches.
//
//   tsm_clock [--n rows] [--k k] [--m m] [--sets S] [--launches L] [--idle-ms I]
//             [--interval-us U]
//
// After I ms of idle, one single-wave sampler kernel on a second stream
// records every U microseconds the pair (s_memrealtime, s_memtime): the
// 100 MHz constant counter and the shader-clock cycle counter, so each
// interval gives the shader clock.  Meanwhile the first stream runs L
// back-to-back FP64 tsm_atb launches over S rotated input sets (the headline
// workload by default), with an event after every launch.  Prints one line
// per launch (start and duration in microseconds from the first launch) and
// one line per sample (time from the first sample, shader clock in GHz).
// The sampler occupies one wave of one compute unit; its loop ends after a
// fixed number of samples.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <chrono>
#include <thread>
#include <vector>

#include "tsm/tsm.h"

#define HIP_OK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      std::exit(1);                                                                 \
    }                                                                               \
  } while (0)

// count samples, interval ticks of the 100 MHz counter apart, into out[2 * i].
__global__ void __launch_bounds__(64) clock_sampler(unsigned long long* out, int count,
                                                    unsigned long long interval) {
  if (threadIdx.x != 0) return;
  unsigned long long next = __builtin_amdgcn_s_memrealtime();
  for (int i = 0; i < count; ++i) {
    unsigned long long rt;
    do {
      rt = __builtin_amdgcn_s_memrealtime();
    } while (rt < next);
    const unsigned long long ct = __builtin_amdgcn_s_memtime();
    out[2 * i] = rt;
    out[2 * i + 1] = ct;
    next = rt + interval;
  }
}

int main(int argc, char** argv) {
  int64_t n = 1 << 22;
  int k = 8, m = 8, sets = 3, launches = 400, idle_ms = 1000, interval_us = 20;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> long long {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "%s needs a value\n", a.c_str());
        std::exit(2);
      }
      return std::atoll(argv[++i]);
    };
    if (a == "--n") n = next();
    else if (a == "--k") k = (int)next();
    else if (a == "--m") m = (int)next();
    else if (a == "--sets") sets = (int)next();
    else if (a == "--launches") launches = (int)next();
    else if (a == "--idle-ms") idle_ms = (int)next();
    else if (a == "--interval-us") interval_us = (int)next();
    else {
      std::fprintf(stderr, "usage: tsm_clock [--n rows] [--k k] [--m m] [--sets S] "
                           "[--launches L] [--idle-ms I] [--interval-us U]\n");
      return 2;
    }
  }
  if (n < 1 || k < 1 || m < 1 || k > 64 || m > 64 || sets < 1 || launches < 1 ||
      interval_us < 1 || idle_ms < 0)
    return 2;
  hipStream_t s, ss;
  HIP_OK(hipStreamCreate(&s));
  HIP_OK(hipStreamCreate(&ss));
  std::vector<void*> A(sets), B(sets);
  for (int i = 0; i < sets; ++i) {
    HIP_OK(hipMalloc(&A[i], (size_t)n * k * 8));
    HIP_OK(hipMalloc(&B[i], (size_t)n * m * 8));
    if (tsm_fill_random(TSM_F64, TSM_DIST_UNIFORM, 20261203 + i, 1, A[i], n, k, k, s) ||
        tsm_fill_random(TSM_F64, TSM_DIST_UNIFORM, 20261203 + i, 2, B[i], n, m, m, s))
      return 1;
  }
  void* C = nullptr;
  HIP_OK(hipMalloc(&C, (size_t)k * m * 8));
  const size_t ws_bytes = tsm_atb_workspace_size(TSM_F64, n, k, m, nullptr);
  void* ws = nullptr;
  HIP_OK(hipMalloc(&ws, ws_bytes));
  HIP_OK(hipMemset(ws, 0, ws_bytes));
  // one launch to load the code objects, then idle
  if (tsm_atb(TSM_F64, TSM_F64, n, k, m, A[0], k, B[0], m, C, m, 0, ws, ws_bytes, s)) return 1;
  HIP_OK(hipDeviceSynchronize());
  // samples cover the launches at 200 us each (an upper bound) plus 5 ms
  const int count = (int)(((long long)launches * 200 + 5000) / interval_us);
  unsigned long long* samples = nullptr;
  HIP_OK(hipMalloc(&samples, (size_t)count * 2 * sizeof(unsigned long long)));
  std::vector<hipEvent_t> ev(launches + 1);
  for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  std::this_thread::sleep_for(std::chrono::milliseconds(idle_ms));
  hipLaunchKernelGGL(clock_sampler, dim3(1), dim3(64), 0, ss, samples, count,
                     (unsigned long long)interval_us * 100);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(ev[0], s));
  for (int i = 0; i < launches; ++i) {
    if (tsm_atb(TSM_F64, TSM_F64, n, k, m, A[i % sets], k, B[i % sets], m, C, m, 0, ws, ws_bytes,
                s))
      return 1;
    HIP_OK(hipEventRecord(ev[i + 1], s));
  }
  HIP_OK(hipDeviceSynchronize());
  const double bytes = 8.0 * n * (k + m) + 8.0 * k * m;
  double total = 0;
  for (int i = 0; i < launches; ++i) {
    float t0 = 0, d = 0;
    HIP_OK(hipEventElapsedTime(&t0, ev[0], ev[i]));
    HIP_OK(hipEventElapsedTime(&d, ev[i], ev[i + 1]));
    total += d;
    std::printf("launch %d start_us %.1f dur_us %.2f GBps %.1f\n", i, t0 * 1e3, d * 1e3,
                bytes / (d * 1e-3) / 1e9);
  }
  std::printf("launches %d mean_us %.2f\n", launches, total * 1e3 / launches);
  std::vector<unsigned long long> h((size_t)count * 2);
  HIP_OK(hipMemcpy(h.data(), samples, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int i = 1; i < count; ++i) {
    const double drt = (double)(h[2 * i] - h[2 * i - 2]);
    const double dct = (double)(h[2 * i + 1] - h[2 * i - 1]);
    std::printf("sample %d t_us %.1f clock_GHz %.3f\n", i, (h[2 * i] - h[0]) / 100.0,
                drt > 0 ? dct / drt * 0.1 : 0.0);
  }
  return 0;
}
