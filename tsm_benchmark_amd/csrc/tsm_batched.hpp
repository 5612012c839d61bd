// Batched C_i (+)= A_i^T B_i for many independent problems in one persistent
// launch.  The host cuts every problem into work items (runs of row tiles of
// up to 4 MiB of A + B each) and sorts them by size.  Workgroup g starts with
// item g; every later item is taken from a queue (an atomic ticket, drawn one
// item ahead, so its latency hides behind the item being streamed), which
// evens out compute units that stream slower than others.  A workgroup streams
// its item with the same body as the single-problem kernel (atb_partial,
// specialised for the problem's width class), stores the item's FP64 partial
// in the item's fixed slot, and arrives at the problem's counter; the last
// item of a problem to arrive adds that problem's partials in slot order and
// writes C_i.  The summation tree depends only on the plan, not on which
// workgroup took which item, so results are deterministic.
#pragma once

#include "tsm_device.hpp"

namespace tsm {

struct BatchProblem {
  const void* A;
  const void* B;
  void* C;
  int64_t n, lda, ldb, ldc;
  int k, m;
  int cls;          // width-class index of the batched kernel's class list
  int dense;        // k == KC, m == MC, lda == k, ldb == m, 16-byte aligned A and B
  int64_t partial_offset;  // first FP64 word of this problem's partial slots
  unsigned int n_items;
  unsigned int pad_;
};

struct BatchItem {
  int problem;
  int slot;             // partial slot within the problem
  int64_t tile_begin;   // tiles [tile_begin, tile_end) of the problem (the last
  int64_t tile_end;     // item may include the partial tile)
};

struct BatchArgs {
  const BatchProblem* probs;
  const BatchItem* items;
  int64_t n_items;
  double* partials;
  unsigned int* counters;  // one per problem, zero between calls
  unsigned int* queue;     // [next ticket, workgroups done], zero between calls
  int accumulate;
  int c_f32;
};

// The last item's sum of a problem's item partials for classes of R kThreads
// outputs (R = 2, 4, 8, 16): thread t takes outputs t + r kThreads and loads
// UG = 64 / R slots of all R at once, so 64 loads are in flight instead of 16
// (each output is still added in slot order: the bits of finish_outputs).  A
// separate function so that only the batched kernels carry its registers.
template <int KCMC>
__device__ __forceinline__ void finish_outputs_joint(const AtbArgs& a, int G, int MC) {
  constexpr int R = KCMC / kThreads, UG = 64 / R;
  static_assert(R * kThreads == KCMC && R >= 2 && R <= 16, "2..16 outputs per thread");
  const int tid = threadIdx.x;
  const double* P = a.partials;
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  for (int g = 0; g < G; g += UG) {
    double v[UG][R];
#pragma unroll
    for (int x = 0; x < UG; ++x)
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[x][r] = g + x < G ? P[(int64_t)(g + x) * KCMC + tid + r * kThreads] : 0.0;
#pragma unroll
    for (int x = 0; x < UG; ++x)
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] += v[x][r];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) store_output(a, MC, tid + r * kThreads, acc[r]);
}

// Ticket of the item after the current one: item gridDim.x + ticket.  Thread
// 0 draws it before the current item's loads and stores it to LDS at the end
// of the item; every thread reads it after the item's final barrier.
__device__ __forceinline__ unsigned draw_ticket(const BatchArgs& ba) {
  return __hip_atomic_fetch_add(ba.queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class S, int KC, int MC, int TM, int TN, int T, int PF, int NT>
__device__ __forceinline__ void batched_item(const BatchArgs& ba, const BatchItem& it,
                                             unsigned char* smem, int* s_flag,
                                             unsigned* s_next) {
  unsigned ticket = 0;
  if (threadIdx.x == 0) ticket = draw_ticket(ba);
  const BatchProblem& pr = ba.probs[it.problem];
  constexpr int KCMC = KC * MC;
  AtbArgs a{};
  a.A = pr.A;
  a.B = pr.B;
  a.C = pr.C;
  a.partials = ba.partials + pr.partial_offset;
  a.counter = ba.counters + it.problem;
  a.n = pr.n;
  a.lda = pr.lda;
  a.ldb = pr.ldb;
  a.ldc = pr.ldc;
  a.k = pr.k;
  a.m = pr.m;
  a.accumulate = ba.accumulate;
  a.c_f32 = ba.c_f32;
  a.dense = pr.dense;
  a.finish = 1;
  const int64_t full = pr.n / T;
  const int tail_rows = (int)(pr.n - full * T);
  const int64_t end_full = it.tile_end < full ? it.tile_end : full;
  const int64_t count = end_full > it.tile_begin ? end_full - it.tile_begin : 0;
  const bool owns_tail = tail_rows > 0 && it.tile_begin <= full && it.tile_end > full;
  atb_partial<S, KC, MC, TM, TN, T, PF, NT>(a, smem, it.tile_begin, 1, count, owns_tail, full,
                                            tail_rows, a.partials + (int64_t)it.slot * KCMC);
  if (arrive_last(a.counter, pr.n_items, s_flag)) {
    if constexpr (KCMC >= 2 * kThreads)
      finish_outputs_joint<KCMC>(a, (int)pr.n_items, MC);
    else
      finish_outputs(a, (int)pr.n_items, KCMC, MC, 0, KCMC, reinterpret_cast<double*>(smem));
    if (threadIdx.x == 0)
      __hip_atomic_store(a.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x == 0) *s_next = ticket;
  __syncthreads();  // the next item reuses the LDS this one read last; s_next is visible
}

// An item of no compiled class (the host never plans one): only the ticket.
__device__ __forceinline__ void batched_skip(const BatchArgs& ba, unsigned* s_next) {
  if (threadIdx.x == 0) *s_next = draw_ticket(ba);
  __syncthreads();
}

// After a workgroup's last item: the last workgroup to finish (every other
// one has drawn its final ticket, and its draw has returned) zeroes the queue
// for the next call.
__device__ __forceinline__ void batched_exit(const BatchArgs& ba) {
  if (threadIdx.x != 0) return;
  if (__hip_atomic_fetch_add(ba.queue + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
      gridDim.x - 1) {
    __hip_atomic_store(ba.queue, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(ba.queue + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// One class body behind a call boundary: the batched kernel switches over many
// classes, and keeping each body out of line gives it its own register frame
// (inlined together, the bodies share one allocation and spill).
template <class S, int KC, int MC, int TM, int TN, int T, int PF, int NT>
__device__ __noinline__ void batched_item_call(const BatchArgs& ba, const BatchItem& it,
                                               unsigned char* smem, int* s_flag,
                                               unsigned* s_next) {
  batched_item<S, KC, MC, TM, TN, T, PF, NT>(ba, it, smem, s_flag, s_next);
}

}  // namespace tsm
