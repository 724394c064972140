// Setup on the device: the per-tet element data of a mesh from its rest
// positions, and its read-back.
//
// Replaces Region's shape-gradient computation (jax/fem/region/_region.py:
// 84-108: dXdr = X_cell^T dhdr, drdX = inv(dXdr), dV = det(dXdr) w with the
// one-point weight w = 1/6, dhdX = dhdr drdX) and the Fraction factor of
// warp/fem/_base.py:104, which the host otherwise forms with batched numpy
// inverses.  One thread per tet, FP64 arithmetic, results rounded once to
// the mesh precision and stored SoA (row k of the inverse at k T + t), the
// layout the element kernels read.
#include "internal.h"

#include <vector>

namespace apl {

template <class R>
__global__ __launch_bounds__(NTH) void region_kernel(const int4* __restrict__ cells, const double* __restrict__ X,
                                                     const double* __restrict__ fraction, R* __restrict__ dminv,
                                                     R* __restrict__ dV, long long T) {
  for (long long t = (long long)blockIdx.x * NTH + threadIdx.x; t < T; t += (long long)gridDim.x * NTH) {
    const int4 c = cells[t];
    const int id[4] = {c.x, c.y, c.z, c.w};
    double p[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) p[a][i] = X[3 * (long long)id[a] + i];
    // dXdr: column j = X_{j+1} - X_0 (dhdr rows 1..3 are the unit vectors)
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) M[i][j] = p[j + 1][i] - p[0][i];
    // inverse by the adjugate: inv(M)_ij = cof(M)_ji / det M
    double C[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        C[i][j] = M[i1][j1] * M[i2][j2] - M[i1][j2] * M[i2][j1];
      }
    const double det = M[0][0] * C[0][0] + M[0][1] * C[0][1] + M[0][2] * C[0][2];
    const double inv_det = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) dminv[(long long)(3 * i + j) * T + t] = (R)(C[j][i] * inv_det);
    const double f = fraction != nullptr ? fraction[t] : 1.0;
    dV[t] = (R)(det / 6.0 * f);
  }
}

template <class R>
__global__ __launch_bounds__(NTH) void region_readback_kernel(const R* __restrict__ dminv, const R* __restrict__ dV,
                                                              R* __restrict__ dminv_out, R* __restrict__ dv_out,
                                                              long long T) {
  for (long long t = (long long)blockIdx.x * NTH + threadIdx.x; t < T; t += (long long)gridDim.x * NTH) {
    if (dminv_out != nullptr)
#pragma unroll
      for (int k = 0; k < 9; ++k) dminv_out[9 * t + k] = dminv[(long long)k * T + t];
    if (dv_out != nullptr) dv_out[t] = dV[t];
  }
}

static unsigned grid_for(long long T) {
  const long long g = (T + NTH - 1) / NTH;
  return (unsigned)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

int region_from_points(apl_mesh* m, const int32_t* cells, const double* points, const double* fraction) {
  const long long T = m->L.T, V = m->L.V;
  int4* d_cells = nullptr;
  double* d_x = nullptr;
  double* d_f = nullptr;
  int rc = 0;
  auto fail = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == 0) {
      set_error(std::string(what) + ": " + hipGetErrorString(e));
      (void)hipGetLastError();
      rc = -1;
    }
  };
  fail(hipMalloc((void**)&d_cells, (size_t)T * sizeof(int4)), "hipMalloc");
  if (rc == 0) fail(hipMalloc((void**)&d_x, (size_t)std::max<long long>(V, 1) * 3 * sizeof(double)), "hipMalloc");
  if (rc == 0 && fraction != nullptr) fail(hipMalloc((void**)&d_f, (size_t)T * sizeof(double)), "hipMalloc");
  if (rc == 0) fail(hipMemcpy(d_cells, cells, (size_t)T * sizeof(int4), hipMemcpyHostToDevice), "hipMemcpy");
  if (rc == 0 && V > 0) fail(hipMemcpy(d_x, points, (size_t)V * 3 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  if (rc == 0 && d_f != nullptr)
    fail(hipMemcpy(d_f, fraction, (size_t)T * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  if (rc == 0) {
    if (m->dtype == DT_F32)
      hipLaunchKernelGGL(region_kernel<float>, dim3(grid_for(T)), dim3(NTH), 0, nullptr, d_cells, d_x, d_f,
                         (float*)m->d.dminv, (float*)m->d.dV, T);
    else
      hipLaunchKernelGGL(region_kernel<double>, dim3(grid_for(T)), dim3(NTH), 0, nullptr, d_cells, d_x, d_f,
                         (double*)m->d.dminv, (double*)m->d.dV, T);
    fail(hipGetLastError(), "region kernel launch");
    if (rc == 0) fail(hipDeviceSynchronize(), "region kernel");
  }
  for (void* p : {(void*)d_cells, (void*)d_x, (void*)d_f})
    if (p != nullptr) drop_hip_error(hipFree(p));
  return rc;
}

int region_readback(apl_mesh* m, void* dminv_out, void* dv_out, hipStream_t s) {
  const long long T = m->L.T;
  if (T == 0) return 0;
  if (m->dtype == DT_F32)
    hipLaunchKernelGGL(region_readback_kernel<float>, dim3(grid_for(T)), dim3(NTH), 0, s, (const float*)m->d.dminv,
                       (const float*)m->d.dV, (float*)dminv_out, (float*)dv_out, T);
  else
    hipLaunchKernelGGL(region_readback_kernel<double>, dim3(grid_for(T)), dim3(NTH), 0, s,
                       (const double*)m->d.dminv, (const double*)m->d.dV, (double*)dminv_out, (double*)dv_out, T);
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace apl
