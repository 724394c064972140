// Device helper shared by the element finish pass and the solver kernels:
// the sum of one vertex's slot partials, in ascending block order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace apl {

// The partials of vertex v are the contiguous run [off[v], off[v+1]) of the
// vertex-major partial buffer (NCH values each).  The first four are fetched
// together -- that covers almost every vertex of a block-partitioned mesh --
// and the additions run in ascending order, bitwise those of a plain loop.
template <class R, int NCH>
__device__ __forceinline__ void vertex_partial_sum(const int32_t* __restrict__ off, const R* __restrict__ partial,
                                                   long long v, R acc[NCH]) {
  const int k0 = off[v], k1 = off[v + 1];
  R val[4][NCH];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch)
      val[j][ch] = k0 + j < k1 ? partial[(unsigned)(k0 + j) * NCH + ch] : R(0);
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) acc[ch] = R(0);
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (k0 + j < k1)
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) acc[ch] += val[j][ch];
  for (int k = k0 + 4; k < k1; ++k) {
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) acc[ch] += partial[(unsigned)k * NCH + ch];
  }
}

}  // namespace apl
