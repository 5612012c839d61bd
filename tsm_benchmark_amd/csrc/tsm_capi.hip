// C ABI of libtsm (include/tsm/tsm.h): argument checks, kernel choice,
// workspace handling and launches.  Every entry point validates its
// arguments before touching the HIP runtime, so argument errors are reported
// the same way with or without a GPU, and no entry point throws.
#include "tsm/tsm.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "tsm_internal.hpp"

namespace {

thread_local int g_last_hip_error = 0;

constexpr size_t kCounterBytes = 256;  // arrival counters at the head of a workspace

tsm_status hip_fail(hipError_t e) {
  g_last_hip_error = (int)e;
  return TSM_ERR_HIP;
}

int elem_bytes(tsm_dtype t) { return t == TSM_F64 ? 8 : t == TSM_F32 ? 4 : 2; }

bool valid_in_type(int t) { return t == TSM_F64 || t == TSM_F32 || t == TSM_BF16; }
bool valid_out_type(int t) { return t == TSM_F64 || t == TSM_F32; }

int width_class(int x) {
  int c = 1;
  while (c < x) c <<= 1;
  return c;
}

// Compute units of the current device (cached per device).
tsm_status device_cus(int* cus) {
  static std::mutex mu;
  static std::map<int, int> cache;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e);
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) {
    *cus = it->second;
    return TSM_SUCCESS;
  }
  int n = 0;
  e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_fail(e);
  cache[dev] = n;
  *cus = n;
  return TSM_SUCCESS;
}

const tsm::KernelEntry* table(int* count) { return tsm_gen_kernel_table(count); }

// ---- kernel choice --------------------------------------------------------------
struct Plan {
  int kernel = -1;
  const tsm::KernelEntry* entry = nullptr;
  // Staged kernel for inputs the direct engine cannot read (== kernel otherwise).
  int fallback = -1;
  int fallback_interleave = 0;
  int64_t grid = 1;
  int64_t tiles_per_wg = 1;
  int interleave = 0;
  int kcmc = 0;
  bool finish_in_kernel = true;
  size_t workspace = 0;
};

// The kernels of one operation: A^T B (sym = false) or A^T A (sym = true).
struct Tables {
  const tsm::KernelEntry* ks;
  int nk;
  const tsm::TunedEntry* tuned;
  int ntuned;
};
Tables tables(bool sym) {
  Tables t{};
  if (sym) {
    t.ks = tsm_gen_sym_table(&t.nk);
    t.tuned = tsm_gen_sym_tuned_table(&t.ntuned);
  } else {
    t.ks = table(&t.nk);
    t.tuned = tsm_gen_tuned_table(&t.ntuned);
  }
  return t;
}

tsm_status find_kernel(const Tables& tb, tsm_dtype in_type, int kc, int mc,
                       const tsm_kernel_config* cfg, int* index, int* wgs_per_cu,
                       int* interleave, int* fallback, int* fallback_interleave) {
  const int ntuned = tb.ntuned, nk = tb.nk;
  const tsm::TunedEntry* tuned = tb.tuned;
  const tsm::KernelEntry* ks = tb.ks;
  int def = -1, def_wpc = 2, def_il = 0;
  for (int i = 0; i < ntuned; ++i)
    if (tuned[i].dtype == (int)in_type && tuned[i].kc == kc && tuned[i].mc == mc) {
      def = tuned[i].kernel;
      def_wpc = tuned[i].wgs_per_cu;
      def_il = tuned[i].interleave;
      *fallback = tuned[i].fallback;
      *fallback_interleave = tuned[i].fallback_interleave;
    }
  if (def < 0) return TSM_ERR_NO_KERNEL;
  *wgs_per_cu = (cfg && cfg->wgs_per_cu > 0) ? cfg->wgs_per_cu : def_wpc;
  *interleave = (cfg && cfg->tile_order > 0) ? (cfg->tile_order == 2 ? 1 : 0) : def_il;
  if (!cfg || (cfg->tm == 0 && cfg->tn == 0 && cfg->rows == 0 && cfg->pf == 0 &&
               cfg->load_policy == 0 && cfg->engine == 0)) {
    *index = def;
    return TSM_SUCCESS;
  }
  // Explicit request: every non-zero field filters the class's kernels; among
  // the matches the one closest to the tuned default wins (lowest index on
  // ties), so the choice is deterministic.
  const tsm::KernelEntry& d = ks[def];
  int best = -1, best_score = -1;
  for (int i = 0; i < nk; ++i) {
    const tsm::KernelEntry& e = ks[i];
    if (e.dtype != (int)in_type || e.kc != kc || e.mc != mc) continue;
    if (cfg->engine > 0 && e.engine != cfg->engine) continue;
    if (cfg->tm > 0 && e.tm != cfg->tm) continue;
    if (cfg->tn > 0 && e.tn != cfg->tn) continue;
    if (cfg->rows > 0 && e.rows != cfg->rows) continue;
    if (cfg->pf > 0 && e.pf != cfg->pf) continue;
    if (cfg->load_policy > 0 && e.nt != (cfg->load_policy == 2 ? 1 : 0)) continue;
    const int score = (e.tm == d.tm) + (e.tn == d.tn) + (e.rows == d.rows) + (e.pf == d.pf) +
                      (e.nt == d.nt);
    if (score > best_score) {
      best = i;
      best_score = score;
    }
  }
  if (best < 0) return TSM_ERR_NO_KERNEL;
  *index = best;
  return TSM_SUCCESS;
}

tsm_status run_plan(const Plan& p, tsm::LaunchFn launch, int kc, int mc, tsm_dtype out_type,
                    int64_t n, int k, int m, const void* A, int64_t lda, const void* B,
                    int64_t ldb, void* C, int64_t ldc, int accumulate, bool dense,
                    void* workspace, size_t workspace_bytes, hipStream_t s);

// Grid, tile split and workspace of a launch over n rows in tiles of T rows.
tsm_status set_grid(Plan* p, int64_t n, int64_t T, int wpc, int interleave, int64_t cfg_grid,
                    int kc, int mc);

// Kernel and launch shape of A^T B (sym = false, k x m) or A^T A (sym = true,
// m == k).
tsm_status make_plan(tsm_dtype in_type, int64_t n, int k, int m, const tsm_kernel_config* cfg,
                     Plan* p, bool sym = false) {
  const int kc = width_class(k), mc = width_class(m);
  int wpc = 2, interleave = 0;
  const Tables tb = tables(sym);
  tsm_status st = find_kernel(tb, in_type, kc, mc, cfg, &p->kernel, &wpc, &interleave,
                              &p->fallback, &p->fallback_interleave);
  if (st != TSM_SUCCESS) return st;
  p->entry = &tb.ks[p->kernel];
  if (p->entry->engine != 3) {
    p->fallback = p->kernel;
    p->fallback_interleave = interleave;
  }
  return set_grid(p, n, p->entry->rows, wpc, interleave, cfg ? cfg->grid : 0, kc, mc);
}

// A direct-engine plan for inputs that are not dense: the class's staged
// kernel on at most the same grid, so the workspace the plan asked for still
// suffices.
tsm_status use_fallback(Plan* p, int64_t n, bool sym) {
  if (p->entry->engine != 3) return TSM_SUCCESS;
  const Tables tb = tables(sym);
  if (p->fallback < 0 || p->fallback >= tb.nk || tb.ks[p->fallback].engine == 3)
    return TSM_ERR_NO_KERNEL;
  const int64_t grid = p->grid;
  p->kernel = p->fallback;
  p->entry = &tb.ks[p->kernel];
  return set_grid(p, n, p->entry->rows, 1, p->fallback_interleave, grid, p->entry->kc,
                  p->entry->mc);
}

bool valid_config(const tsm_kernel_config* c) {
  // pf: 1 to 4 tiles in flight (staged kernels), 4, 8 or 16 steps (direct engine)
  return !c || !((c->pf < 0 || (c->pf > 4 && c->pf != 8 && c->pf != 16)) || c->tm < 0 || c->tn < 0 || c->rows < 0 ||
                 c->wgs_per_cu < 0 || c->grid < 0 || c->load_policy < 0 || c->load_policy > 2 ||
                 c->tile_order < 0 || c->tile_order > 2 || c->engine < 0 || c->engine > 3);
}

// Final-sum protocol and workspace of a plan whose grid and class are set.
// The last workgroup to arrive adds the partials inside the streaming kernel
// (finish_small / finish_outputs) only for short sums: fewer than 256
// workgroups and fewer than 16384 partial words (128 KiB), at most 512
// outputs.  Otherwise a finisher kernel of ceil(KC*MC / 16) workgroups adds
// them after the streaming kernel: its workgroups then exit as soon as their
// partials are stored, with no arrival atomics (256 of them on one counter
// serialise at about 12 ns each when the workgroups arrive together) and no
// single workgroup reading every partial at the end (128 KiB into one compute
// unit takes 2.7-2.9 us).  Measured per call, finisher against in-kernel sum:
// FP64 8 x 8 at n = 2^22 79.2 -> 78.5 us, 16 x 16 at 2^21 85.6 -> 79.6 us,
// 1 x 1 at 2^20 9.0 -> 6.8 us; below 256 workgroups the second launch costs
// more than it saves (8 x 8 at n = 16384: 7.0 against 8.8 us).
constexpr int64_t kFinisherGrid = 256;
constexpr int64_t kFinisherWords = 16384;
void set_finish(Plan* p) {
  p->finish_in_kernel =
      p->kcmc <= 512 && p->grid < kFinisherGrid && p->grid * p->kcmc < kFinisherWords;
  p->workspace = kCounterBytes + (size_t)p->grid * p->kcmc * sizeof(double);
}

tsm_status set_grid(Plan* p, int64_t n, int64_t T, int wpc, int interleave, int64_t cfg_grid,
                    int kc, int mc) {
  int64_t want;
  if (cfg_grid > 0) {
    want = cfg_grid;
  } else {
    int cus = 0;
    tsm_status st = device_cus(&cus);
    if (st != TSM_SUCCESS) return st;
    want = (int64_t)wpc * cus;
  }
  p->kcmc = kc * mc;
  if (p->entry && p->entry->engine == 3) {
    // Direct engine: steps of T rows (4 for FP64/FP32, 32 for BF16); each
    // of the 4 waves of a workgroup takes q = tiles_per_wg steps (a multiple
    // of the depth D), at least D when n allows; the kernel deals the rest
    // one step at a time.
    const int64_t F = n / T, D = p->entry->pf;
    const int64_t cap = F / (4 * D);
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    p->grid = want;
    p->tiles_per_wg = F / (4 * want * D) * D;
    p->interleave = 0;
    set_finish(p);
    return TSM_SUCCESS;
  }
  const int64_t tiles = n > 0 ? (n + T - 1) / T : 1;
  if (want > tiles) want = tiles;
  if (want < 1) want = 1;
  p->interleave = interleave;
  if (interleave) {
    p->tiles_per_wg = (tiles + want - 1) / want;
    p->grid = want;
  } else {
    p->tiles_per_wg = (tiles + want - 1) / want;
    p->grid = (tiles + p->tiles_per_wg - 1) / p->tiles_per_wg;
  }
  set_finish(p);
  return TSM_SUCCESS;
}

tsm_status check_atb(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, int m,
                     const void* A, int64_t lda, const void* B, int64_t ldb, const void* C,
                     int64_t ldc) {
  if (!valid_in_type(in_type) || !valid_out_type(out_type)) return TSM_ERR_UNSUPPORTED;
  if (n < 0 || k < 1 || m < 1 || k > TSM_MAX_WIDTH || m > TSM_MAX_WIDTH)
    return TSM_ERR_INVALID_ARGUMENT;
  if (lda < k || ldb < m || ldc < m) return TSM_ERR_INVALID_ARGUMENT;
  if (!C || (n > 0 && (!A || !B))) return TSM_ERR_INVALID_ARGUMENT;
  if (n > ((int64_t)1 << 62) / (lda > ldb ? lda : ldb)) return TSM_ERR_INVALID_ARGUMENT;
  const uintptr_t es = (uintptr_t)elem_bytes(in_type);
  if (((uintptr_t)A % es) || ((uintptr_t)B % es)) return TSM_ERR_MISALIGNED;
  if ((uintptr_t)C % (uintptr_t)elem_bytes(out_type)) return TSM_ERR_MISALIGNED;
  return TSM_SUCCESS;
}

// ---- library-owned workspaces: one per (device, stream) --------------------------------
struct Pool {
  std::mutex mu;
  std::map<std::pair<int, hipStream_t>, std::pair<void*, size_t>> bufs;
};
Pool& pool() {
  static Pool* p = new Pool();  // never destroyed: the HIP runtime may be gone at exit
  return *p;
}

tsm_status pooled_workspace(size_t bytes, hipStream_t stream, void** out) {
  // A graph captured now would keep the pool's buffer, which a later, larger
  // call on this stream frees (and the first call allocates and clears it,
  // which has no place in a graph): capture needs a caller workspace.
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(stream, &cs);
  if (e != hipSuccess) return hip_fail(e);
  if (cs != hipStreamCaptureStatusNone) return TSM_ERR_WORKSPACE;
  int dev = 0;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e);
  Pool& P = pool();
  std::lock_guard<std::mutex> lock(P.mu);
  auto& slot = P.bufs[{dev, stream}];
  if (slot.second < bytes) {
    if (slot.first) {
      // Earlier work on this stream may still use the old buffer.
      e = hipStreamSynchronize(stream);
      if (e != hipSuccess) return hip_fail(e);
      e = hipFree(slot.first);
      if (e != hipSuccess) return hip_fail(e);
      slot = {nullptr, 0};
    }
    size_t want = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
    void* p = nullptr;
    e = hipMalloc(&p, want);
    if (e != hipSuccess) return hip_fail(e);
    e = hipMemsetAsync(p, 0, kCounterBytes > want ? want : kCounterBytes, stream);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return hip_fail(e);
    }
    slot = {p, want};
  }
  *out = slot.first;
  return TSM_SUCCESS;
}

// ---- the workspace of tsm_atb_noalloc: a static array per GPU --------------------
// Arrival counters (the first 256 bytes) and workgroup partials of every
// default plan fit in it.  In host code `scratch` names a host-side stand-in,
// not device memory: host code uses only the per-device address that
// hipGetSymbolAddress returns for it.
static __device__ __align__(256) unsigned char scratch[TSM_NOALLOC_SCRATCH_BYTES];

__global__ void tsm_scratch_clear(unsigned int* p, int words) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < words) p[i] = 0u;
}

// Device address of the current device's copy of `scratch`, looked up once
// per device with that device current.  The lookup also clears the arrival
// counters with a kernel on `stream` (the kernels leave them zero after every
// call), so it is refused on a stream that is being captured.
tsm_status scratch_for_device(hipStream_t stream, void** out) {
  static std::mutex mu;
  static std::map<int, void*> ptrs;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e);
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) {
    *out = it->second;
    return TSM_SUCCESS;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  e = hipStreamIsCapturing(stream, &cs);
  if (e != hipSuccess) return hip_fail(e);
  if (cs != hipStreamCaptureStatusNone) return TSM_ERR_UNSUPPORTED;
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e);
  if (!ptr || (uintptr_t)ptr % kCounterBytes) return TSM_ERR_MISALIGNED;
  const int words = (int)(kCounterBytes / sizeof(unsigned int));
  hipLaunchKernelGGL(tsm_scratch_clear, dim3(1), dim3(words), 0, stream,
                     static_cast<unsigned int*>(ptr), words);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e);
  ptrs[dev] = ptr;
  *out = ptr;
  return TSM_SUCCESS;
}

}  // namespace

// ---- finisher kernel (wide classes, large grids) ---------------------------------------
extern "C" __global__ void __launch_bounds__(256) tsm_atb_finish(tsm::AtbArgs a, int G, int KC, int MC) {
  tsm::finish_body(a, G, KC, MC);
}

// ---- synthetic inputs -------------------------------------------------------------------
namespace {
__device__ __forceinline__ uint64_t rng_word(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t z = seed + stream * 0xd1b54a32d192ed03ULL + (idx + 1) * 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double rng_value(int dist, uint64_t w) {
  if (dist == TSM_DIST_UNIFORM) return (double)(w >> 11) * 0x1p-52 - 1.0;
  const int e = (int)((w >> 11) % 41u) - 20;
  const uint64_t bits = ((uint64_t)(1023 + e) << 52) | (w & 0x8000000000000000ULL);
  return __longlong_as_double((long long)bits);
}
// Single round-to-nearest-even from binary64 to bfloat16 (8-bit significand);
// below 2^-126 to the nearest multiple of 2^-133 (the bf16 subnormals).
__device__ __forceinline__ uint16_t round_bf16(double x) {
  uint64_t b = (uint64_t)__double_as_longlong(x);
  if ((b & 0x7fffffffffffffffULL) == 0) return (uint16_t)((b >> 48) & 0x8000);
  if (__builtin_fabs(x) < 0x1p-126) {
    const double q = __builtin_rint(__builtin_fabs(x) * 0x1p133);  // exact scaling, ties to even
    return (uint16_t)(((b >> 48) & 0x8000) | (uint16_t)q);
  }
  const uint64_t lsb = (b >> 45) & 1u;
  b += 0x00000fffffffffffULL + lsb;
  b &= ~0x00001fffffffffffULL;
  const float f = (float)__longlong_as_double((long long)b);
  return (uint16_t)(__float_as_uint(f) >> 16);
}
}  // namespace

extern "C" __global__ void __launch_bounds__(256)
    tsm_fill_kernel(int type, int dist, uint64_t seed, uint64_t stream_id, void* X, int64_t rows,
                    int64_t cols, int64_t ld) {
  const int64_t total = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols, c = i - r * cols;
    const double v = rng_value(dist, rng_word(seed, stream_id, (uint64_t)i));
    const int64_t off = r * ld + c;
    if (type == TSM_F64)
      static_cast<double*>(X)[off] = v;
    else if (type == TSM_F32)
      static_cast<float*>(X)[off] = (float)v;
    else
      static_cast<uint16_t*>(X)[off] = round_bf16(v);
  }
}

// ---- entry points ------------------------------------------------------------------------
extern "C" {

size_t tsm_atb_workspace_size(tsm_dtype in_type, int64_t n, int k, int m,
                              const tsm_kernel_config* config) {
  if (!valid_in_type(in_type) || n < 0 || k < 1 || m < 1 || k > TSM_MAX_WIDTH ||
      m > TSM_MAX_WIDTH || !valid_config(config))
    return 0;
  Plan p;
  if (make_plan(in_type, n, k, m, config, &p) != TSM_SUCCESS) return 0;
  return p.workspace;
}

tsm_status tsm_atb_plan(tsm_dtype in_type, int64_t n, int k, int m,
                        const tsm_kernel_config* config, int* kernel_index, int64_t* grid) {
  if (!valid_in_type(in_type)) return TSM_ERR_UNSUPPORTED;
  if (n < 0 || k < 1 || m < 1 || k > TSM_MAX_WIDTH || m > TSM_MAX_WIDTH)
    return TSM_ERR_INVALID_ARGUMENT;
  Plan p;
  tsm_status st = make_plan(in_type, n, k, m, config, &p);
  if (st != TSM_SUCCESS) return st;
  if (kernel_index) *kernel_index = p.kernel;
  if (grid) *grid = p.grid;
  return TSM_SUCCESS;
}

tsm_status tsm_atb_ex(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, int m,
                      const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                      int64_t ldc, int accumulate, void* workspace, size_t workspace_bytes,
                      tsm_stream_t stream, const tsm_kernel_config* config) {
  tsm_status st = check_atb(in_type, out_type, n, k, m, A, lda, B, ldb, C, ldc);
  if (st != TSM_SUCCESS) return st;
  if (!valid_config(config)) return TSM_ERR_INVALID_ARGUMENT;
  Plan p;
  st = make_plan(in_type, n, k, m, config, &p);
  if (st != TSM_SUCCESS) return st;
  const bool dense = k == p.entry->kc && m == p.entry->mc && lda == k && ldb == m &&
                     ((uintptr_t)A % 16) == 0 && ((uintptr_t)B % 16) == 0;
  if (!dense) {
    st = use_fallback(&p, n, false);
    if (st != TSM_SUCCESS) return st;
  }
  const tsm::KernelEntry& ke = *p.entry;
  return run_plan(p, ke.launch, ke.kc, ke.mc, out_type, n, k, m, A, lda, B, ldb, C, ldc,
                  accumulate, dense, workspace, workspace_bytes,
                  static_cast<hipStream_t>(stream));
}

tsm_status tsm_atb_noalloc(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, int m,
                           const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                           int64_t ldc, int accumulate, tsm_stream_t stream) {
  tsm_status st = check_atb(in_type, out_type, n, k, m, A, lda, B, ldb, C, ldc);
  if (st != TSM_SUCCESS) return st;
  Plan p;
  st = make_plan(in_type, n, k, m, nullptr, &p);
  if (st != TSM_SUCCESS) return st;
  const bool dense = k == p.entry->kc && m == p.entry->mc && lda == k && ldb == m &&
                     ((uintptr_t)A % 16) == 0 && ((uintptr_t)B % 16) == 0;
  if (!dense) {
    st = use_fallback(&p, n, false);
    if (st != TSM_SUCCESS) return st;
  }
  if (p.workspace > TSM_NOALLOC_SCRATCH_BYTES) return TSM_ERR_WORKSPACE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* ws = nullptr;
  st = scratch_for_device(s, &ws);
  if (st != TSM_SUCCESS) return st;
  const tsm::KernelEntry& ke = *p.entry;
  return run_plan(p, ke.launch, ke.kc, ke.mc, out_type, n, k, m, A, lda, B, ldb, C, ldc,
                  accumulate, dense, ws, TSM_NOALLOC_SCRATCH_BYTES, s);
}

}  // extern "C"

namespace {
// Launches a planned A^T B (or A^T A) kernel and, for wide classes, the finisher.
tsm_status run_plan(const Plan& p, tsm::LaunchFn launch, int kc, int mc, tsm_dtype out_type,
                    int64_t n, int k, int m, const void* A, int64_t lda, const void* B,
                    int64_t ldb, void* C, int64_t ldc, int accumulate, bool dense,
                    void* workspace, size_t workspace_bytes, hipStream_t s) {
  if (workspace) {
    if (workspace_bytes < p.workspace) return TSM_ERR_WORKSPACE;
    if ((uintptr_t)workspace % kCounterBytes) return TSM_ERR_MISALIGNED;
  } else {
    tsm_status st = pooled_workspace(p.workspace, s, &workspace);
    if (st != TSM_SUCCESS) return st;
  }
  tsm::AtbArgs a{};
  a.A = A;
  a.B = B;
  a.C = C;
  a.counter = static_cast<unsigned int*>(workspace);
  a.partials = reinterpret_cast<double*>(static_cast<char*>(workspace) + kCounterBytes);
  a.n = n;
  a.lda = lda;
  a.ldb = ldb;
  a.ldc = ldc;
  a.tiles_per_wg = p.tiles_per_wg;
  a.interleave = p.interleave;
  a.k = k;
  a.m = m;
  a.accumulate = accumulate ? 1 : 0;
  a.c_f32 = out_type == TSM_F32 ? 1 : 0;
  a.dense = dense ? 1 : 0;
  a.finish = p.finish_in_kernel ? 1 : 0;
  hipError_t e = launch(a, (unsigned)p.grid, s);
  if (e != hipSuccess) return hip_fail(e);
  if (!p.finish_in_kernel) {
    const unsigned fg = (unsigned)((p.kcmc + tsm::kFinishOutputs - 1) / tsm::kFinishOutputs);
    hipLaunchKernelGGL(tsm_atb_finish, dim3(fg), dim3(256), 0, s, a, (int)p.grid, kc, mc);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
  }
  return TSM_SUCCESS;
}
}  // namespace

extern "C" {

size_t tsm_ata_workspace_size(tsm_dtype in_type, int64_t n, int k,
                              const tsm_kernel_config* config) {
  if (!valid_in_type(in_type) || n < 0 || k < 1 || k > TSM_MAX_WIDTH || !valid_config(config))
    return 0;
  Plan p;
  if (make_plan(in_type, n, k, k, config, &p, true) != TSM_SUCCESS) return 0;
  return p.workspace;
}

tsm_status tsm_ata_ex(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, const void* A,
                      int64_t lda, void* C, int64_t ldc, int accumulate, void* workspace,
                      size_t workspace_bytes, tsm_stream_t stream,
                      const tsm_kernel_config* config) {
  tsm_status st = check_atb(in_type, out_type, n, k, k, A, lda, A, lda, C, ldc);
  if (st != TSM_SUCCESS) return st;
  if (!valid_config(config)) return TSM_ERR_INVALID_ARGUMENT;
  Plan p;
  st = make_plan(in_type, n, k, k, config, &p, true);
  if (st != TSM_SUCCESS) return st;
  const bool dense = k == p.entry->kc && lda == k && ((uintptr_t)A % 16) == 0;
  if (!dense) {
    st = use_fallback(&p, n, true);
    if (st != TSM_SUCCESS) return st;
  }
  const tsm::KernelEntry& e = *p.entry;
  return run_plan(p, e.launch, e.kc, e.kc, out_type, n, k, k, A, lda, A, lda, C, ldc,
                  accumulate, dense, workspace, workspace_bytes,
                  static_cast<hipStream_t>(stream));
}

tsm_status tsm_ata(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, const void* A,
                   int64_t lda, void* C, int64_t ldc, int accumulate, void* workspace,
                   size_t workspace_bytes, tsm_stream_t stream) {
  return tsm_ata_ex(in_type, out_type, n, k, A, lda, C, ldc, accumulate, workspace,
                    workspace_bytes, stream, nullptr);
}

tsm_status tsm_atb(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, int m, const void* A,
                   int64_t lda, const void* B, int64_t ldb, void* C, int64_t ldc, int accumulate,
                   void* workspace, size_t workspace_bytes, tsm_stream_t stream) {
  return tsm_atb_ex(in_type, out_type, n, k, m, A, lda, B, ldb, C, ldc, accumulate, workspace,
                    workspace_bytes, stream, nullptr);
}

tsm_status tsm_fill_random(tsm_dtype type, tsm_dist dist, uint64_t seed, uint64_t stream_id,
                           void* X, int64_t rows, int64_t cols, int64_t ld, tsm_stream_t stream) {
  if (!valid_in_type(type)) return TSM_ERR_UNSUPPORTED;
  if (dist != TSM_DIST_UNIFORM && dist != TSM_DIST_POW2) return TSM_ERR_INVALID_ARGUMENT;
  if (rows < 0 || cols < 1 || ld < cols) return TSM_ERR_INVALID_ARGUMENT;
  if (rows > ((int64_t)1 << 62) / ld) return TSM_ERR_INVALID_ARGUMENT;
  if (rows > 0 && !X) return TSM_ERR_INVALID_ARGUMENT;
  if ((uintptr_t)X % (uintptr_t)elem_bytes(type)) return TSM_ERR_MISALIGNED;
  if (rows == 0) return TSM_SUCCESS;
  const int64_t total = rows * cols;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(tsm_fill_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), (int)type, (int)dist, seed, stream_id, X,
                     rows, cols, ld);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e);
  return TSM_SUCCESS;
}

const char* tsm_strerror(tsm_status s) {
  switch (s) {
    case TSM_SUCCESS: return "success";
    case TSM_ERR_INVALID_ARGUMENT: return "invalid argument";
    case TSM_ERR_UNSUPPORTED: return "unsupported data type or request";
    case TSM_ERR_MISALIGNED: return "misaligned pointer";
    case TSM_ERR_WORKSPACE:
      return "workspace too small, or none given for a call on a capturing stream";
    case TSM_ERR_HIP: return "HIP runtime error";
    case TSM_ERR_NO_KERNEL: return "no generated kernel matches the configuration";
    default: break;
  }
  return "unknown status";
}

int tsm_last_hip_error(void) { return g_last_hip_error; }

tsm_status tsm_internal_check_atb(tsm_dtype in_type, tsm_dtype out_type, int64_t n, int k, int m,
                                  const void* A, int64_t lda, const void* B, int64_t ldb,
                                  const void* C, int64_t ldc) {
  return check_atb(in_type, out_type, n, k, m, A, lda, B, ldb, C, ldc);
}

tsm_status tsm_internal_device_cus(int* cus) { return device_cus(cus); }

tsm_status tsm_internal_pooled_workspace(size_t bytes, void* stream, void** out) {
  return pooled_workspace(bytes, static_cast<hipStream_t>(stream), out);
}

int tsm_internal_hip_status(hipError_t e) {
  if (e == hipSuccess) return TSM_SUCCESS;
  g_last_hip_error = (int)e;
  return TSM_ERR_HIP;
}

int tsm_version(void) { return TSM_VERSION_MAJOR * 10000 + TSM_VERSION_MINOR * 100; }

int tsm_num_kernels(void) {
  int n = 0;
  table(&n);
  return n;
}

tsm_status tsm_kernel_info(int index, tsm_kernel_desc* out) {
  int n = 0;
  const tsm::KernelEntry* ks = table(&n);
  if (!out || index < 0 || index >= n) return TSM_ERR_INVALID_ARGUMENT;
  const tsm::KernelEntry& e = ks[index];
  out->name = e.name;
  out->in_type = e.dtype;
  out->kc = e.kc;
  out->mc = e.mc;
  out->tm = e.tm;
  out->tn = e.tn;
  out->rows = e.rows;
  out->pf = e.pf;
  out->load_policy = e.nt ? 2 : 1;
  out->lds_bytes = e.lds_bytes;
  out->engine = e.engine;
  return TSM_SUCCESS;
}

}  // extern "C"

