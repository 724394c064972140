// Fixed-order workgroup sums of FP64 tables (per-block energies, per-vertex-
// block dot products), shared by the finish pass and the solver's scalar
// kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace apl {

// Column sums of a row-major table of W doubles per row, p[0 .. n) with n a
// multiple of W, by one workgroup of NT threads (W divides 64).  Thread t
// adds the entries t, t + NT, t + 2 NT, ... -- all of column t % W -- with
// eight loads in flight; lanes of equal column then combine by a fixed
// shuffle tree and the waves in order.  out[0 .. W) is valid in every
// thread on return.  The order depends only on n, so a table gives the same
// bits wherever it sits in a batch.
template <int W, int NT>
__device__ inline void flat_column_sums(const double* __restrict__ p, long long n, double* sm, double* out) {
  static_assert(64 % W == 0 && NT % 64 == 0, "column width must divide the wave");
  const int t = threadIdx.x;
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long f = t; f < n; f += 8LL * NT) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = f + (long long)u * NT < n ? p[f + (long long)u * NT] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += v[u];
  }
  double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
#pragma unroll
  for (int off = 32; off >= W; off >>= 1) s += __shfl_down(s, off, 64);
  const int lane = t & 63, w = t >> 6;
  if (lane < W) sm[w * W + lane] = s;
  __syncthreads();
  if (t < W) {
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) r += sm[i * W + t];
    sm[(NT / 64) * W + t] = r;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < W; ++k) out[k] = sm[(NT / 64) * W + k];
  __syncthreads();
}

}  // namespace apl
