// Benchmark harness over the libtsm C ABI: the roles of the reference's
// main.cc (hand-written variants 4, 6, 9 over 18 (M, stride, tile) configs,
// main.cc:157-195) and main_codegen.cc + the generated run_all_benchmarks
// (sycl_codegen_new.py:417-521).
//
//   tsm_bench [--suite main|generated|atb] [--K rows] [--M m] [--dtype f64|f32|bf16]
//             [--warmup W] [--runs R] [--json]
//
// Inputs come from tsm_fill_random (counter-based, seed 20261203, uniform on
// [-1, 1)) instead of the reference's per-thread minstd_rand0 (main.cc:72-86).
// Correctness uses the reference's host loop (main.cc:14-27, here over all K
// rows) and acceptance rule (main.cc:29-41).  Timing brackets each call with
// HIP events on the launch stream, zeroing C outside the timed region
// (main.cc:118-126); the table adds GB/s and the fraction of 8 TB/s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iterator>
#include <numeric>
#include <utility>
#include <string>
#include <vector>

#include "tsm/tsmttsm_kernels.hh"

namespace {

constexpr double kHbmPeak = 8.0e12;
constexpr uint64_t kSeed = 20261203;

#define HIP_OK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      std::exit(1);                                                                 \
    }                                                                               \
  } while (0)

struct Stats {
  double avg_ns, min_ns, max_ns, stddev_ns;
};

struct Row {
  std::string name;
  bool correct;
  Stats s;
  double gflops, gbps;
};

// main.cc:14-27, generalised to M x N.
std::vector<double> compute_naive(int64_t K, int M, int N, const double* A, const double* B) {
  std::vector<double> C((size_t)M * N, 0.0);
  for (int m = 0; m < M; ++m)
    for (int n = 0; n < N; ++n) {
      double acc = 0.0;
      for (int64_t k = 0; k < K; ++k) acc += A[k * M + m] * B[k * N + n];
      C[(size_t)m * N + n] = acc;
    }
  return C;
}

// main.cc:29-41
bool is_correct(const std::vector<double>& ref, const std::vector<double>& got,
                double rel_tol = 1e-6, double abs_tol = 1e-10) {
  for (size_t i = 0; i < ref.size(); ++i) {
    const double diff = std::fabs(ref[i] - got[i]);
    const double mag = std::max(std::fabs(ref[i]), std::fabs(got[i]));
    if ((diff > abs_tol && diff > rel_tol * mag) || diff != diff) {
      std::fprintf(stderr, "Mismatch at index %zu: expected %.17g, got %.17g\n", i, ref[i], got[i]);
      return false;
    }
  }
  return true;
}

// FP32 and BF16 storage: the reference checks FP64 only, with a rule relative
// to each entry, which entries that cancel to near zero cannot meet at FP32
// accumulation.  These use the library's documented bound instead:
// |C - C_exact| <= 4 sqrt(K) 2^-23 (|A|^T |B|)_ij elementwise.
bool within_bound(const std::vector<double>& ref, const std::vector<double>& abs_ref,
                  const std::vector<double>& got, int64_t K) {
  const double c = 4.0 * std::sqrt((double)(K > 1 ? K : 1)) * std::ldexp(1.0, -23);
  for (size_t i = 0; i < ref.size(); ++i) {
    const double diff = std::fabs(ref[i] - got[i]);
    if (!(diff <= c * abs_ref[i])) {
      std::fprintf(stderr, "Mismatch at index %zu: expected %.17g, got %.17g (bound %.3g)\n", i,
                   ref[i], got[i], c * abs_ref[i]);
      return false;
    }
  }
  return true;
}

struct Problem {
  int64_t K;
  int M, N;
  tsm_dtype type;
  void* A = nullptr;
  void* B = nullptr;
  double* C = nullptr;
  std::vector<double> Ah, Bh, Cref, Cabs;  // Cabs = |A|^T |B| (FP32/BF16 storage)
  hipStream_t s = nullptr;
  int esize() const { return type == TSM_F64 ? 8 : type == TSM_F32 ? 4 : 2; }

  Problem(int64_t K_, int M_, int N_, tsm_dtype t) : K(K_), M(M_), N(N_), type(t) {
    HIP_OK(hipStreamCreate(&s));
    HIP_OK(hipMalloc(&A, (size_t)K * M * esize()));
    HIP_OK(hipMalloc(&B, (size_t)K * N * esize()));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&C), (size_t)M * N * sizeof(double)));
    if (tsm_fill_random(type, TSM_DIST_UNIFORM, kSeed, 1, A, K, M, M, s) != TSM_SUCCESS ||
        tsm_fill_random(type, TSM_DIST_UNIFORM, kSeed, 2, B, K, N, N, s) != TSM_SUCCESS) {
      std::fprintf(stderr, "tsm_fill_random failed\n");
      std::exit(1);
    }
    HIP_OK(hipStreamSynchronize(s));
    Ah = to_host(A, (size_t)K * M);
    Bh = to_host(B, (size_t)K * N);
    Cref = compute_naive(K, M, N, Ah.data(), Bh.data());
    if (type != TSM_F64) {
      std::vector<double> Aa(Ah), Ba(Bh);
      for (double& x : Aa) x = std::fabs(x);
      for (double& x : Ba) x = std::fabs(x);
      Cabs = compute_naive(K, M, N, Aa.data(), Ba.data());
    }
  }
  ~Problem() {
    (void)hipFree(A);
    (void)hipFree(B);
    (void)hipFree(C);
    (void)hipStreamDestroy(s);
  }
  std::vector<double> to_host(const void* d, size_t count) const {
    std::vector<double> out(count);
    if (type == TSM_F64) {
      HIP_OK(hipMemcpy(out.data(), d, count * 8, hipMemcpyDeviceToHost));
    } else if (type == TSM_F32) {
      std::vector<float> t(count);
      HIP_OK(hipMemcpy(t.data(), d, count * 4, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < count; ++i) out[i] = t[i];
    } else {
      std::vector<uint16_t> t(count);
      HIP_OK(hipMemcpy(t.data(), d, count * 2, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < count; ++i) {
        uint32_t u = (uint32_t)t[i] << 16;
        float f;
        std::memcpy(&f, &u, 4);
        out[i] = f;
      }
    }
    return out;
  }
  void zero_C() { HIP_OK(hipMemsetAsync(C, 0, (size_t)M * N * sizeof(double), s)); }
  double bytes() const { return (double)esize() * K * (M + N) + 8.0 * M * N; }
  double flops() const { return 2.0 * K * M * N; }
};

// run_and_profile (main.cc:109-139) with events on the launch stream.
Row run(Problem& P, const std::string& name, const std::function<tsm_status()>& kernel, int warmup,
        int runs) {
  Row r{name, false, {}, 0, 0};
  P.zero_C();
  if (kernel() != TSM_SUCCESS) {
    std::fprintf(stderr, "%s: launch failed\n", name.c_str());
    return r;
  }
  HIP_OK(hipStreamSynchronize(P.s));
  std::vector<double> got((size_t)P.M * P.N);
  HIP_OK(hipMemcpy(got.data(), P.C, got.size() * 8, hipMemcpyDeviceToHost));
  r.correct = P.type == TSM_F64 ? is_correct(P.Cref, got) : within_bound(P.Cref, P.Cabs, got, P.K);
  if (!r.correct) return r;  // skipped in timing, as main.cc:97-99
  for (int i = 0; i < warmup; ++i) {
    P.zero_C();
    kernel();
  }
  HIP_OK(hipStreamSynchronize(P.s));
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  std::vector<double> t(runs);
  for (int i = 0; i < runs; ++i) {
    P.zero_C();
    HIP_OK(hipEventRecord(e0, P.s));
    kernel();
    HIP_OK(hipEventRecord(e1, P.s));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    t[i] = ms * 1e6;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double sum = std::accumulate(t.begin(), t.end(), 0.0);
  const double avg = sum / runs;
  const double sq = std::inner_product(t.begin(), t.end(), t.begin(), 0.0);
  r.s = {avg, *std::min_element(t.begin(), t.end()), *std::max_element(t.begin(), t.end()),
         std::sqrt(std::max(0.0, sq / runs - avg * avg))};
  r.gflops = P.flops() / (avg * 1e-9) / 1e9;
  r.gbps = P.bytes() / (avg * 1e-9) / 1e9;
  return r;
}

void print(const std::vector<Row>& rows, bool json) {
  if (json) {
    std::printf("[\n");
    for (size_t i = 0; i < rows.size(); ++i) {
      const Row& r = rows[i];
      std::printf("  {\"kernel\": \"%s\", \"correct\": %s, \"avg_ms\": %.6f, \"min_ms\": %.6f, "
                  "\"stddev_ms\": %.6f, \"gflops\": %.2f, \"gbps\": %.2f, \"hbm_frac\": %.4f}%s\n",
                  r.name.c_str(), r.correct ? "true" : "false", r.s.avg_ns / 1e6, r.s.min_ns / 1e6,
                  r.s.stddev_ns / 1e6, r.gflops, r.gbps, r.gbps * 1e9 / kHbmPeak,
                  i + 1 < rows.size() ? "," : "");
    }
    std::printf("]\n");
    return;
  }
  std::printf("%-34s %8s %11s %11s %10s %11s %10s %7s\n", "Kernel", "Correct", "Avg (ms)", "Min (ms)",
              "Stddev", "GFLOP/s", "GB/s", "%HBM");
  for (const Row& r : rows) {
    std::printf("%-34s %8s", r.name.c_str(), r.correct ? "PASS" : "FAIL");
    if (r.correct)
      std::printf(" %11.4f %11.4f %10.4f %11.2f %10.1f %6.1f%%", r.s.avg_ns / 1e6, r.s.min_ns / 1e6,
                  r.s.stddev_ns / 1e6, r.gflops, r.gbps, 100.0 * r.gbps * 1e9 / kHbmPeak);
    std::printf("\n");
  }
}

// main.cc:171-192: (M, stride, local, tile).  The reference runs variants 4,
// 6 and 9 there; every hand-written name (tsmttsm.hh:15-586) runs here, with
// the library entry point on the same problem as the speed reference.
template <int M, int stride, int local, int tile>
void main_suite_one(std::vector<Row>& rows, int64_t K, int warmup, int runs) {
  Problem P(K, M, M, TSM_F64);
  const double* A = static_cast<const double*>(P.A);
  const double* B = static_cast<const double*>(P.B);
  const std::string tag = "M" + std::to_string(M) + "_s" + std::to_string(stride) + "_t" +
                          std::to_string(tile);
  using namespace tsm_compat;
  const std::pair<const char*, std::function<tsm_status()>> names[] = {
      {"Variant1_", [&] { return tsmttsm<double, M, stride>(P.s, K, A, B, P.C); }},
      {"Variant2_", [&] { return tsmttsm2<double, M, stride, tile>(P.s, K, A, B, P.C); }},
      {"Variant3_", [&] { return tsmttsm3<double, M, stride, tile>(P.s, K, A, B, P.C); }},
      {"Variant4_", [&] { return tsmttsm4<double, M, stride, tile>(P.s, K, A, B, P.C); }},
      {"Variant5_", [&] { return tsmttsm5<double, M, stride, tile>(P.s, K, A, B, P.C); }},
      {"Variant6_", [&] { return tsmttsm6<double, M, stride, local, tile>(P.s, K, A, B, P.C); }},
      {"Variant7_", [&] { return tsmttsm7<double, M, stride, tile>(P.s, K, A, B, P.C); }},
      {"Variant8_", [&] { return tsmttsm8<double, M>(P.s, K, A, B, P.C); }},
      {"Variant9_", [&] { return tsmttsm9<double, M, stride, local, tile>(P.s, K, A, B, P.C); }},
      {"tsm_atb_", [&] {
         return tsm_atb(TSM_F64, TSM_F64, K, M, M, A, M, B, M, P.C, M, 1, nullptr, 0, P.s);
       }},
  };
  for (const auto& nm : names) rows.push_back(run(P, nm.first + tag, nm.second, warmup, runs));
}

void main_suite(std::vector<Row>& rows, int64_t K, int warmup, int runs) {
  main_suite_one<4, 4096, 32, 2>(rows, K, warmup, runs);
  main_suite_one<4, 4096, 32, 4>(rows, K, warmup, runs);
  main_suite_one<4, 8192, 32, 2>(rows, K, warmup, runs);
  main_suite_one<4, 8192, 32, 4>(rows, K, warmup, runs);
  main_suite_one<4, 16384, 32, 2>(rows, K, warmup, runs);
  main_suite_one<4, 16384, 32, 4>(rows, K, warmup, runs);
  main_suite_one<8, 4096, 32, 2>(rows, K, warmup, runs);
  main_suite_one<8, 4096, 32, 4>(rows, K, warmup, runs);
  main_suite_one<8, 8192, 32, 2>(rows, K, warmup, runs);
  main_suite_one<8, 8192, 32, 4>(rows, K, warmup, runs);
  main_suite_one<8, 16384, 32, 2>(rows, K, warmup, runs);
  main_suite_one<8, 16384, 32, 4>(rows, K, warmup, runs);
  main_suite_one<16, 8192, 32, 2>(rows, K, warmup, runs);
  main_suite_one<16, 8192, 32, 4>(rows, K, warmup, runs);
  main_suite_one<16, 16384, 32, 2>(rows, K, warmup, runs);
  main_suite_one<16, 16384, 32, 4>(rows, K, warmup, runs);
  main_suite_one<16, 32768, 32, 2>(rows, K, warmup, runs);
  main_suite_one<16, 32768, 32, 4>(rows, K, warmup, runs);
}

// run_all_benchmarks<double> (sycl_codegen_new.py:417-521).
void generated_suite(std::vector<Row>& rows, int64_t K, int warmup, int runs) {
  int lastM = -1;
  Problem* P = nullptr;
  for (const auto& g : tsm_compat::kGeneratedKernels) {
    if (g.M != lastM) {
      delete P;
      P = new Problem(K, g.M, g.M, TSM_F64);
      lastM = g.M;
    }
    const double* A = static_cast<const double*>(P->A);
    const double* B = static_cast<const double*>(P->B);
    rows.push_back(run(*P, g.name, [&] { return g.fn(P->s, K, A, B, P->C); }, warmup, runs));
    const bool last_of_M = (&g + 1 == std::end(tsm_compat::kGeneratedKernels)) || (&g + 1)->M != g.M;
    if (last_of_M)  // the library entry point on the same problem, as the speed reference
      rows.push_back(run(*P, "tsm_atb_M" + std::to_string(g.M), [&] {
        return tsm_atb(TSM_F64, TSM_F64, K, g.M, g.M, A, g.M, B, g.M, P->C, g.M, 1, nullptr, 0,
                       P->s);
      }, warmup, runs));
  }
  delete P;
}

// The tuned library entry point on one shape.
void atb_suite(std::vector<Row>& rows, int64_t K, int M, int N, tsm_dtype t, int warmup, int runs) {
  Problem P(K, M, N, t);
  const std::string name = std::string("tsm_atb_") + (t == TSM_F64 ? "f64" : t == TSM_F32 ? "f32" : "bf16") +
                           "_" + std::to_string(M) + "x" + std::to_string(N);
  rows.push_back(run(P, name, [&] {
    return tsm_atb(t, TSM_F64, K, M, N, P.A, M, P.B, N, P.C, N, 1, nullptr, 0, P.s);
  }, warmup, runs));
}

}  // namespace

int main(int argc, char** argv) {
  std::string suite = "generated", dtype = "f64";
  int64_t K = 1 << 20;
  int M = 8, N = -1, warmup = 10, runs = 30;
  bool json = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", a.c_str());
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--suite") suite = next();
    else if (a == "--K") K = std::atoll(next());
    else if (a == "--M") M = std::atoi(next());
    else if (a == "--N") N = std::atoi(next());
    else if (a == "--dtype") dtype = next();
    else if (a == "--warmup") warmup = std::atoi(next());
    else if (a == "--runs") runs = std::atoi(next());
    else if (a == "--json") json = true;
    else {
      std::fprintf(stderr, "usage: tsm_bench [--suite main|generated|atb] [--K n] [--M k] [--N m] "
                           "[--dtype f64|f32|bf16] [--warmup W] [--runs R] [--json]\n");
      return 2;
    }
  }
  if (N < 0) N = M;
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, 0));
  if (!json) std::printf("Device: %s (%d CUs), K=%lld\n", prop.gcnArchName, prop.multiProcessorCount,
                         (long long)K);
  std::vector<Row> rows;
  if (suite == "main") {
    main_suite(rows, K, warmup, runs);
  } else if (suite == "generated") {
    generated_suite(rows, K, warmup, runs);
  } else if (suite == "atb") {
    const tsm_dtype t = dtype == "f32" ? TSM_F32 : dtype == "bf16" ? TSM_BF16 : TSM_F64;
    atb_suite(rows, K, M, N, t, warmup, runs);
  } else {
    std::fprintf(stderr, "unknown suite %s\n", suite.c_str());
    return 2;
  }
  print(rows, json);
  for (const Row& r : rows)
    if (!r.correct) return 1;
  return 0;
}
