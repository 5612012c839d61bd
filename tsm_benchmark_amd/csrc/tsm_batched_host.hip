// Host side of the batched C_i (+)= A_i^T B_i: validation, the work plan
// (per-problem chunks of up to 4 MiB of A + B, sorted largest first), the
// workspace layout, and the launches (tsm_atb_batched and the plan API).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "tsm/tsm.h"
#include "tsm_internal.hpp"

namespace {

constexpr size_t kAlign = 256;
// Bytes of A + B per work item.  Each item costs about a microsecond to start
// and to reduce; the mixed 512-pair FP64 batch (28 GB) ran at 6.45, 6.62,
// 6.72 and 6.73 TB/s with items of 1, 2, 4 and 8 MiB (FP32: 6.24, 6.43, 6.59,
// 6.58).  Smaller batches take smaller items, down to kMinItemBytes, so that
// there are about kItemsPerWorkgroup items for each workgroup of a grid of
// kNominalGrid (MI355X's compute units) for the queue to balance.
constexpr int64_t kMaxItemBytes = 4 << 20;
constexpr int64_t kMinItemBytes = 256 << 10;
constexpr int64_t kItemsPerWorkgroup = 8;
constexpr int64_t kNominalGrid = 256;
// At most this many items per problem (larger items beyond; the persistent
// grid has 256 workgroups on MI355X): the last item of a problem adds all of
// the problem's item partials in one workgroup, so a problem cut into
// thousands of items spent hundreds of microseconds in that final sum.  A
// batch holding one 4 GiB FP64 32 x 32 problem (2048 items of 2 MiB, 16 MiB of
// partials) ran at 3.9 TB/s; with 256 items 6.2 (128 items: 4.4, the grid
// then idles half the CUs; 512: 5.6).
constexpr int64_t kMaxItemsPerProblem = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
int elem_bytes(tsm_dtype t) { return t == TSM_F64 ? 8 : t == TSM_F32 ? 4 : 2; }
int width_class(int x) {
  int c = 1;
  while (c < x) c <<= 1;
  return c;
}

const tsm::BatchedKernel* batched_kernel(tsm_dtype t) {
  int n = 0;
  const tsm::BatchedKernel* all = tsm_gen_batched_table(&n);
  for (int i = 0; i < n; ++i)
    if (all[i].dtype == (int)t) return &all[i];
  return nullptr;
}

// The plan of one batched call: problem descriptors, work items and the
// byte layout of the device buffer
//   [counters | problems | items | partials]
// whose first three parts are uploaded from `header` (counters as zeros).
struct Layout {
  std::vector<tsm::BatchProblem> probs;
  std::vector<tsm::BatchItem> items;
  size_t probs_off = 0, items_off = 0, partials_off = 0, total = 0;
  std::vector<unsigned char> header;
};

tsm_status build_layout(tsm_dtype in_type, tsm_dtype out_type, int count,
                        const tsm_problem* problems, Layout* L) {
  if (in_type != TSM_F64 && in_type != TSM_F32 && in_type != TSM_BF16) return TSM_ERR_UNSUPPORTED;
  if (out_type != TSM_F64 && out_type != TSM_F32) return TSM_ERR_UNSUPPORTED;
  if (count < 0 || (count > 0 && !problems)) return TSM_ERR_INVALID_ARGUMENT;
  const tsm::BatchedKernel* bk = batched_kernel(in_type);
  if (!bk) return TSM_ERR_NO_KERNEL;
  const int es = elem_bytes(in_type), out_es = elem_bytes(out_type);
  // Problems wider than the kernel's widest class (FP32 and BF16 stop at 32:
  // their (hi, lo) accumulators for 64 would not fit its register budget) run
  // as column blocks of at most `widest` x `widest`: block (i, j) is
  // A[:, i w ...]^T B[:, j w ...] into C's block (i, j), read through strided
  // views, so A and B are read once per block column of C and row of C.
  int widest = 0;
  for (int c = 0; c < bk->n_classes; ++c) widest = std::max(widest, bk->classes[c].kc);
  std::vector<tsm_problem> parts;
  parts.reserve((size_t)std::max(count, 0));
  for (int p = 0; p < count; ++p) {
    const tsm_problem& q = problems[p];
    tsm_status st = tsm_internal_check_atb(in_type, out_type, q.n, q.k, q.m, q.A, q.lda, q.B,
                                           q.ldb, q.C, q.ldc);
    if (st != TSM_SUCCESS) return st;
    if (q.k <= widest && q.m <= widest) {
      parts.push_back(q);
      continue;
    }
    for (int i0 = 0; i0 < q.k; i0 += widest)
      for (int j0 = 0; j0 < q.m; j0 += widest) {
        tsm_problem b = q;
        b.A = q.n > 0 ? static_cast<const unsigned char*>(q.A) + (size_t)i0 * es : q.A;
        b.B = q.n > 0 ? static_cast<const unsigned char*>(q.B) + (size_t)j0 * es : q.B;
        b.C = static_cast<unsigned char*>(q.C) + ((size_t)i0 * q.ldc + j0) * out_es;
        b.k = std::min(widest, q.k - i0);
        b.m = std::min(widest, q.m - j0);
        parts.push_back(b);
      }
  }
  if (parts.size() > (size_t)0x7fffffff) return TSM_ERR_INVALID_ARGUMENT;
  count = (int)parts.size();
  problems = parts.data();
  auto class_of = [bk](const tsm_problem& q) {
    const int kc = width_class(q.k), mc = width_class(q.m);
    int cls = -1;
    for (int c = 0; c < bk->n_classes; ++c)
      if (bk->classes[c].kc == kc && bk->classes[c].mc == mc) cls = c;
    return cls;
  };
  int64_t total_bytes = 0;  // tiles x tile bytes over the batch
  for (int p = 0; p < count; ++p) {
    const int cls = class_of(problems[p]);
    if (cls < 0) return TSM_ERR_UNSUPPORTED;
    const int64_t T = bk->classes[cls].rows;
    total_bytes += (problems[p].n + T - 1) / T * T *
                   (bk->classes[cls].kc + bk->classes[cls].mc) * es;
  }
  const int64_t item_bytes = std::min(
      kMaxItemBytes,
      std::max(kMinItemBytes, total_bytes / (kItemsPerWorkgroup * kNominalGrid)));
  int64_t partial_words = 0;
  L->probs.resize(count);
  for (int p = 0; p < count; ++p) {
    const tsm_problem& q = problems[p];
    tsm_status st = tsm_internal_check_atb(in_type, out_type, q.n, q.k, q.m, q.A, q.lda, q.B,
                                           q.ldb, q.C, q.ldc);
    if (st != TSM_SUCCESS) return st;
    const int kc = width_class(q.k), mc = width_class(q.m);
    const int cls = class_of(q);
    const int64_t T = bk->classes[cls].rows;
    const int64_t tiles = (q.n + T - 1) / T;
    const int64_t tile_bytes = T * (kc + mc) * es;
    const int64_t chunk = std::max<int64_t>(
        std::max<int64_t>(1, item_bytes / tile_bytes),
        (tiles + kMaxItemsPerProblem - 1) / kMaxItemsPerProblem);
    const int64_t nitems = std::max<int64_t>(1, (tiles + chunk - 1) / chunk);
    if (nitems > 0x7fffffff) return TSM_ERR_INVALID_ARGUMENT;
    tsm::BatchProblem& b = L->probs[p];
    b = tsm::BatchProblem{};
    b.A = q.A;
    b.B = q.B;
    b.C = q.C;
    b.n = q.n;
    b.lda = q.lda;
    b.ldb = q.ldb;
    b.ldc = q.ldc;
    b.k = q.k;
    b.m = q.m;
    b.cls = cls;
    b.dense = (q.k == kc && q.m == mc && q.lda == q.k && q.ldb == q.m &&
               ((uintptr_t)q.A % 16) == 0 && ((uintptr_t)q.B % 16) == 0)
                  ? 1
                  : 0;
    b.partial_offset = partial_words;
    b.n_items = (unsigned)nitems;
    partial_words += nitems * kc * mc;
    for (int64_t s = 0; s < nitems; ++s) {
      tsm::BatchItem it{};
      it.problem = p;
      it.slot = (int)s;
      it.tile_begin = s * chunk;
      it.tile_end = std::min(tiles, (s + 1) * chunk);
      L->items.push_back(it);
    }
  }
  if (L->items.size() > 0x7fffffffu) return TSM_ERR_INVALID_ARGUMENT;  // 32-bit queue tickets
  // Largest items first (the first grid-full by workgroup index, the rest from
  // the queue in this order; by bytes instead of tiles measured the same).
  // Among items of one size, slot-major order: the workgroups running at the
  // same time read items of different problems.  Problem-major order had all
  // workgroups streaming neighbouring 2 MiB items of one problem at once,
  // which measured 3.9 instead of 6.3 TB/s for a batch of equal FP64 32 x 32
  // problems (the summation order within a problem is by slot either way).
  std::stable_sort(L->items.begin(), L->items.end(),
                   [](const tsm::BatchItem& x, const tsm::BatchItem& y) {
                     const int64_t sx = x.tile_end - x.tile_begin, sy = y.tile_end - y.tile_begin;
                     if (sx != sy) return sx > sy;
                     if (x.slot != y.slot) return x.slot < y.slot;
                     return x.problem < y.problem;
                   });
  // one arrival counter per problem, then the item queue's two words
  const size_t counters = align_up(sizeof(unsigned) * ((size_t)count + 2));
  L->probs_off = counters;
  L->items_off = L->probs_off + align_up(sizeof(tsm::BatchProblem) * (size_t)count);
  L->partials_off = L->items_off + align_up(sizeof(tsm::BatchItem) * L->items.size());
  L->total = L->partials_off + (size_t)partial_words * sizeof(double);
  L->header.assign(L->partials_off, 0);
  if (count)
    std::memcpy(L->header.data() + L->probs_off, L->probs.data(),
                sizeof(tsm::BatchProblem) * (size_t)count);
  if (!L->items.empty())
    std::memcpy(L->header.data() + L->items_off, L->items.data(),
                sizeof(tsm::BatchItem) * L->items.size());
  return TSM_SUCCESS;
}

tsm::BatchArgs make_args(const Layout& L, void* base, int accumulate, tsm_dtype out_type) {
  unsigned char* b = static_cast<unsigned char*>(base);
  tsm::BatchArgs a{};
  a.counters = reinterpret_cast<unsigned*>(b);
  a.queue = a.counters + L.probs.size();
  a.probs = reinterpret_cast<const tsm::BatchProblem*>(b + L.probs_off);
  a.items = reinterpret_cast<const tsm::BatchItem*>(b + L.items_off);
  a.partials = reinterpret_cast<double*>(b + L.partials_off);
  a.n_items = (int64_t)L.items.size();
  a.accumulate = accumulate ? 1 : 0;
  a.c_f32 = out_type == TSM_F32 ? 1 : 0;
  return a;
}

// Persistent grid: as many workgroups as are resident at once (the batched
// kernels carry every class's registers, so usually one per CU), at most one
// per item.
tsm_status grid_for(const tsm::BatchedKernel* bk, int64_t n_items, unsigned* grid) {
  int cus = 0;
  tsm_status st = tsm_internal_device_cus(&cus);
  if (st != TSM_SUCCESS) return st;
  // per dtype, filled on first use (concurrent first calls store the same value)
  static std::atomic<int> resident[3];
  int per_cu = resident[bk->dtype].load(std::memory_order_relaxed);
  if (per_cu == 0) {
    per_cu = bk->occupancy();
    resident[bk->dtype].store(per_cu, std::memory_order_relaxed);
  }
  int64_t g = std::min<int64_t>(n_items, (int64_t)cus * per_cu);
  *grid = (unsigned)std::max<int64_t>(g, 1);
  return TSM_SUCCESS;
}

// Pinned staging for the per-call upload of tsm_atb_batched, one per stream;
// an event guards its reuse until the previous copy out of it has finished.
struct Staging {
  void* host = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;
};
std::mutex g_staging_mu;
std::map<hipStream_t, Staging>& stagings() {
  static auto* m = new std::map<hipStream_t, Staging>();  // lives to process exit
  return *m;
}

tsm_status upload(const Layout& L, void* dst, hipStream_t s) {
  // A captured copy would replay from the staging buffer, which later calls
  // overwrite: a graph takes a plan (tsm_batched_plan_execute) instead.
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(s, &cs);
  if (e != hipSuccess) return (tsm_status)tsm_internal_hip_status(e);
  if (cs != hipStreamCaptureStatusNone) return TSM_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_staging_mu);
  Staging& st = stagings()[s];
  if (st.done) {
    e = hipEventSynchronize(st.done);
    if (e != hipSuccess) return (tsm_status)tsm_internal_hip_status(e);
  } else {
    e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
    if (e != hipSuccess) return (tsm_status)tsm_internal_hip_status(e);
  }
  if (st.bytes < L.header.size()) {
    if (st.host) (void)hipHostFree(st.host);
    st.host = nullptr;
    st.bytes = 0;
    e = hipHostMalloc(&st.host, L.header.size(), hipHostMallocDefault);
    if (e != hipSuccess) return (tsm_status)tsm_internal_hip_status(e);
    st.bytes = L.header.size();
  }
  std::memcpy(st.host, L.header.data(), L.header.size());
  e = hipMemcpyAsync(dst, st.host, L.header.size(), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return (tsm_status)tsm_internal_hip_status(e);
  return (tsm_status)tsm_internal_hip_status(hipEventRecord(st.done, s));
}

}  // namespace

struct tsm_batched_plan_s {
  tsm_dtype in_type;
  void* buffer;
  tsm::BatchArgs args;
  unsigned grid;
  const tsm::BatchedKernel* kernel;
  // One event per stream the plan ran on (outside graph capture), recorded
  // after each launch: destroy waits for these launches only.
  std::mutex mu;
  std::map<hipStream_t, hipEvent_t> launched;
};

extern "C" {

size_t tsm_atb_batched_workspace_size(tsm_dtype in_type, int count, const tsm_problem* problems) {
  Layout L;
  if (build_layout(in_type, TSM_F64, count, problems, &L) != TSM_SUCCESS) return 0;
  return L.total;
}

tsm_status tsm_atb_batched(tsm_dtype in_type, tsm_dtype out_type, int count,
                           const tsm_problem* problems, int accumulate, void* workspace,
                           size_t workspace_bytes, tsm_stream_t stream) {
  Layout L;
  tsm_status st = build_layout(in_type, out_type, count, problems, &L);
  if (st != TSM_SUCCESS) return st;
  if (count == 0) return TSM_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (workspace) {
    if (workspace_bytes < L.total) return TSM_ERR_WORKSPACE;
    if ((uintptr_t)workspace % kAlign) return TSM_ERR_MISALIGNED;
  } else {
    st = tsm_internal_pooled_workspace(L.total, stream, &workspace);
    if (st != TSM_SUCCESS) return st;
  }
  const tsm::BatchedKernel* bk = batched_kernel(in_type);
  unsigned grid = 1;
  st = grid_for(bk, (int64_t)L.items.size(), &grid);
  if (st != TSM_SUCCESS) return st;
  st = upload(L, workspace, s);
  if (st != TSM_SUCCESS) return st;
  return (tsm_status)tsm_internal_hip_status(
      bk->launch(make_args(L, workspace, accumulate, out_type), grid, s));
}

tsm_status tsm_batched_plan_create(tsm_dtype in_type, tsm_dtype out_type, int count,
                                   const tsm_problem* problems, int accumulate,
                                   tsm_batched_plan* plan) {
  if (!plan) return TSM_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  Layout L;
  tsm_status st = build_layout(in_type, out_type, count, problems, &L);
  if (st != TSM_SUCCESS) return st;
  unsigned grid = 1;
  st = grid_for(batched_kernel(in_type), (int64_t)L.items.size(), &grid);
  if (st != TSM_SUCCESS) return st;
  void* buf = nullptr;
  hipError_t e = hipMalloc(&buf, std::max<size_t>(L.total, kAlign));
  if (e != hipSuccess) return (tsm_status)tsm_internal_hip_status(e);
  e = hipMemcpy(buf, L.header.data(), L.header.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(buf);
    return (tsm_status)tsm_internal_hip_status(e);
  }
  auto* p = new (std::nothrow) tsm_batched_plan_s;
  if (!p) {
    (void)hipFree(buf);
    return TSM_ERR_INVALID_ARGUMENT;
  }
  p->in_type = in_type;
  p->buffer = buf;
  p->args = make_args(L, buf, accumulate, out_type);
  p->grid = grid;
  p->kernel = batched_kernel(in_type);
  *plan = p;
  return TSM_SUCCESS;
}

tsm_status tsm_batched_plan_execute(tsm_batched_plan plan, tsm_stream_t stream) {
  if (!plan) return TSM_ERR_INVALID_ARGUMENT;
  if (plan->args.n_items == 0) return TSM_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = plan->kernel->launch(plan->args, plan->grid, s);
  if (e != hipSuccess) return (tsm_status)tsm_internal_hip_status(e);
  // A launch captured into a graph runs when the graph does: the caller keeps
  // the plan alive as long as such a graph may replay.
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  e = hipStreamIsCapturing(s, &cs);
  if (e != hipSuccess || cs != hipStreamCaptureStatusNone)
    return (tsm_status)tsm_internal_hip_status(e);
  std::lock_guard<std::mutex> lock(plan->mu);
  hipEvent_t& ev = plan->launched[s];
  if (!ev) {
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) {
      ev = nullptr;
      plan->launched.erase(s);
      return (tsm_status)tsm_internal_hip_status(e);
    }
  }
  return (tsm_status)tsm_internal_hip_status(hipEventRecord(ev, s));
}

tsm_status tsm_batched_plan_info(tsm_batched_plan plan, int64_t* items, int64_t* grid) {
  if (!plan) return TSM_ERR_INVALID_ARGUMENT;
  if (items) *items = plan->args.n_items;
  if (grid) *grid = plan->grid;
  return TSM_SUCCESS;
}

tsm_status tsm_batched_plan_destroy(tsm_batched_plan plan) {
  if (!plan) return TSM_ERR_INVALID_ARGUMENT;
  // The plan's own launches may still be queued: wait for each stream's last
  // one (not for the whole device).
  hipError_t e = hipSuccess;
  for (auto& kv : plan->launched) {
    const hipError_t w = hipEventSynchronize(kv.second);
    if (e == hipSuccess) e = w;
    (void)hipEventDestroy(kv.second);
  }
  const hipError_t f = hipFree(plan->buffer);
  delete plan;
  if (e == hipSuccess) e = f;
  return (tsm_status)tsm_internal_hip_status(e);
}

}  // extern "C"
