// ExternalForce potential (warp/potential/_ext_force.py:17-39,63-90):
// E = -sum_k f_k . u[idx_k], grad[idx_k] -= f_k, no Hessian.
// Indices are unique (GLOBAL_POINT_ID in the reference), so the gradient
// scatter needs no atomics; the energy is a fixed-order FP64 two-level sum.
// An index outside [0, V) contributes nothing and touches no memory (the
// Python layer rejects such indices before they get here).
#include "../../include/apple_amd.h"
#include <algorithm>

#include "internal.h"

namespace apl {

constexpr int EXT_PER_BLOCK = NTH * 8;

template <class R>
__global__ __launch_bounds__(NTH) void ext_energy_partial(long long n, long long V, const int32_t* idx, const R* f,
                                                          const R* u, double* part) {
  __shared__ double red[NTH / 64];
  const long long base = (long long)blockIdx.x * EXT_PER_BLOCK;
  double acc = 0.0;
  for (int k = threadIdx.x; k < EXT_PER_BLOCK; k += NTH) {
    const long long i = base + k;
    const long long v = i < n ? idx[i] : -1;
    if (v >= 0 && v < V) {
      acc -= (double)f[3 * i] * (double)u[3 * v] + (double)f[3 * i + 1] * (double)u[3 * v + 1] +
             (double)f[3 * i + 2] * (double)u[3 * v + 2];
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < NTH / 64; ++w) s += red[w];
    part[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(NTH) void ext_energy_final(long long nparts, const double* part, double* energy,
                                                        int accumulate) {
  __shared__ double red[NTH / 64];
  double acc = 0.0;
  for (long long i = threadIdx.x; i < nparts; i += NTH) acc += part[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < NTH / 64; ++w) s += red[w];
    energy[0] = accumulate ? energy[0] + s : s;
  }
}

template <class R>
__global__ __launch_bounds__(NTH) void ext_grad(long long n, long long V, const int32_t* idx, const R* f, R* out,
                                                int accumulate) {
  const long long i = (long long)blockIdx.x * NTH + threadIdx.x;
  if (i >= n) return;
  const long long v = idx[i];
  if (v < 0 || v >= V) return;
  if (accumulate) {
    out[3 * v] -= f[3 * i];
    out[3 * v + 1] -= f[3 * i + 1];
    out[3 * v + 2] -= f[3 * i + 2];
  } else {
    out[3 * v] = -f[3 * i];
    out[3 * v + 1] = -f[3 * i + 1];
    out[3 * v + 2] = -f[3 * i + 2];
  }
}

template <class R>
static int ext_eval(apl_mesh* m, long long n, const int32_t* idx, const void* f, const void* u, double* energy,
                    void* grad, int accumulate, hipStream_t s) {
  if (n == 0) {
    if (energy != nullptr && !accumulate) APL_CHECK_HIP(hipMemsetAsync(energy, 0, sizeof(double), s));
    return 0;
  }
  if (energy != nullptr) {
    const long long nparts = (n + EXT_PER_BLOCK - 1) / EXT_PER_BLOCK;
    const long long cap = (long long)N_DOTS * (long long)std::max<size_t>(m->L.vblk_inst.size(), 1);
    if (nparts > cap) {
      set_error("too many external-force entries for this mesh");
      return -1;
    }
    hipLaunchKernelGGL((ext_energy_partial<R>), dim3((unsigned)nparts), dim3(NTH), 0, s, n, (long long)m->L.V, idx,
                       (const R*)f, (const R*)u, m->d.vblk_dots);
    hipLaunchKernelGGL(ext_energy_final, dim3(1), dim3(NTH), 0, s, nparts, m->d.vblk_dots, energy, accumulate);
    APL_CHECK_HIP(hipGetLastError());
  }
  if (grad != nullptr) {
    hipLaunchKernelGGL((ext_grad<R>), dim3((unsigned)((n + NTH - 1) / NTH)), dim3(NTH), 0, s, n, (long long)m->L.V, idx,
                       (const R*)f, (R*)grad, accumulate);
    APL_CHECK_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace apl

using namespace apl;

extern "C" int apl_ext_force_eval(apl_mesh* m, int64_t n, const int32_t* indices, const void* force, const void* u,
                                  double* energy, void* grad, int accumulate, void* stream) {
  if (m == nullptr || (n > 0 && (indices == nullptr || force == nullptr || (energy != nullptr && u == nullptr)))) {
    set_error("NULL argument");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(m->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  if (m->dtype == DT_F32)
    return ext_eval<float>(m, n, indices, force, u, energy, grad, accumulate, (hipStream_t)stream);
  return ext_eval<double>(m, n, indices, force, u, energy, grad, accumulate, (hipStream_t)stream);
}
