// Persistent workgroup loops over items with a per-round compaction of the
// active ones (shared by the element and vector kernels of a solve).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace apl {

// Workgroup-uniform loop over items item_of(j), j = 0 .. n_cand-1.  When
// `filter` is set and there are at most NT candidates (the usual case: a
// persistent grid gives each workgroup a few dozen), is_active is evaluated
// for all of them in one parallel step and only the active ones are visited,
// in candidate order, so a pass over a batch whose instances have mostly
// converged pays two loads instead of a chain of loads per skipped item.
// With more candidates each one is tested as it comes.  body(item) or
// body(item, next_item) is called by every thread of the workgroup (it may
// synchronise).  list: NT ints of
// LDS (only touched when filter is set).
template <int NT, class ItemOf, class IsActive, class Body>
__device__ inline void for_active(int n_cand, bool filter, int* list, ItemOf item_of, IsActive is_active,
                                  Body body) {
  __shared__ int wave_cnt[NT / 64];
  const int tid = threadIdx.x;
  const bool listed = filter && n_cand <= NT;
  int n = n_cand;
  if (listed) {
    const int it = tid < n_cand ? item_of(tid) : 0;
    const bool on = tid < n_cand && is_active(it);
    const unsigned long long m = __ballot(on);
    const int lane = tid & 63, w = tid >> 6;
    if (lane == 0) wave_cnt[w] = __popcll(m);
    __syncthreads();
    int off = 0;
    n = 0;
#pragma unroll
    for (int ww = 0; ww < NT / 64; ++ww) {
      off += ww < w ? wave_cnt[ww] : 0;
      n += wave_cnt[ww];
    }
    if (on) list[off + __popcll(m & ((1ull << lane) - 1ull))] = it;
    __syncthreads();
  }
  for (int q = 0; q < n; ++q) {
    const int it = listed ? list[q] : item_of(q);
    if (filter && !listed && !is_active(it)) continue;
    if constexpr (std::is_invocable_v<Body, int, int>) {
      // body(item, next): next is the item visited after this one, or -1
      // when it is unknown without testing (unlisted filtered loops) or there
      // is none; lets the body prefetch for the next item
      const int next = q + 1 < n ? (listed ? list[q + 1] : (filter ? -1 : item_of(q + 1))) : -1;
      body(it, next);
    } else {
      body(it);
    }
  }
}

}  // namespace apl
