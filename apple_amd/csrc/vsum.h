// Device helpers shared by the element finish pass and the solver kernels:
// the sum of one vertex's slot partials, in ascending block order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace apl {

// LDS window for a workgroup's partials (bytes)
constexpr int VSUM_LDS_BYTES = 24576;

// Partial sums of the vertices [v0, v1) of one workgroup.  Their partials
// form one contiguous run of the vertex-major buffer (a vertex's partials
// are the run [off[v], off[v+1]), NCH values each, in ascending block
// order); when it fits the LDS window (cap values) the workgroup copies it
// with fully coalesced loads and the threads then add their entries from
// LDS.  Otherwise they read them from global memory.  Either way a sum runs
// from 0 over the run in ascending order, so both give the same bits.
template <class R, int NCH, int NT>
struct VertexSums {
  const int32_t* off;
  const R* partial;
  R* lds;
  int k_lo = 0;
  bool staged = false;

  __device__ void stage(int v0, int v1, int cap) {
    const int k0 = off[v0], k1 = off[v1];
    k_lo = k0;
    const int n = (k1 - k0) * NCH;
    staged = n <= cap;
    if (staged) {
      const R* src = partial + (size_t)k0 * NCH;
#pragma unroll 4
      for (int i = threadIdx.x; i < n; i += NT) lds[i] = src[i];
    }
    __syncthreads();
  }

  // all NCH channels of the vertex whose partial run is [k0, k1)
  __device__ void sum(int k0, int k1, R acc[NCH]) const {
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) acc[ch] = R(0);
    if (staged) {
      for (int k = k0 - k_lo; k < k1 - k_lo; ++k)
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) acc[ch] += lds[k * NCH + ch];
    } else {
      for (int k = k0; k < k1; ++k)
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) acc[ch] += partial[(size_t)k * NCH + ch];
    }
  }

  // component c of each 3-channel group (NCH / 3 groups) of one vertex: the
  // sums of one degree of freedom
  __device__ void sum_dof(int k0, int k1, int c, R acc[NCH / 3]) const {
#pragma unroll
    for (int g = 0; g < NCH / 3; ++g) acc[g] = R(0);
    if (staged) {
      for (int k = k0 - k_lo; k < k1 - k_lo; ++k)
#pragma unroll
        for (int g = 0; g < NCH / 3; ++g) acc[g] += lds[k * NCH + 3 * g + c];
    } else {
      for (int k = k0; k < k1; ++k)
#pragma unroll
        for (int g = 0; g < NCH / 3; ++g) acc[g] += partial[(size_t)k * NCH + 3 * g + c];
    }
  }
};

// Degrees of freedom of a vertex block, three per thread: DOF i = 3 v + c
// runs over [3 v0, 3 v1) with stride NT, so every vector access of the
// block is a fully coalesced run (vertex blocks hold NT vertices).
constexpr int DOF_PER_THREAD = 3;

}  // namespace apl
