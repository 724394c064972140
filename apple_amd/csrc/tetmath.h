// Per-tet elasticity algebra for the element kernels (host + device).
//
// One tet = 4 vertices, 1 quadrature point.  Inputs are the 3x3 inverse rest
// shape matrix B = inv(dX/dr) (rows 1..3 of the reference's dhdX; row 0 is
// minus their sum, jax/fem/region/_region.py:84-108), the quadrature volume
// dV and the 4 vertex displacements.  Everything is expressed through the
// displacement gradient G = F - I, which keeps FP32 accurate near the rest
// state: the energies and stresses below contain no "1 - 1" cancellation.
//
//   stable Neo-Hookean (warp/fem/_stable_neo_hookean.py:17-103)
//     Psi = mu/2 (|F|^2 - 3) - mu (J - 1) + lam/2 (J - 1)^2
//         = mu/2 |G|^2 - mu (c2(G) + det G) + lam/2 (J - 1)^2,
//     J - 1 = tr G + c2(G) + det G, c2 = sum of principal 2x2 minors,
//     cof F = I + tr(G) I - G^T + cof G,
//     P = mu (F - cof F) + lam (J-1) cof F,  F - cof F = G + G^T - tr(G) I - cof G.
//   SNH muscle (warp/fem/_stable_neo_hookean_muscle.py:18-113): the SNH law on
//     G_A = F A with A = I + S, S the symmetric matrix of the per-tet
//     activation (a0 a3 a4; a3 a1 a5; a4 a5 a2) (warp/fem/func/_misc.py:48-53).
//     With d'_a = A d_a (the rows of dhdX A), G_A - I = S + sum_a u_a d'_a^T,
//     so it is the SNH evaluation on the geometry d' plus the offset S.
//   ARAP (warp/fem/_arap.py:17-76, func/_misc.py:31-70)
//     F^T F = I + E with E = G + G^T + G^T G; E's eigenpairs give V and
//     sigma_i^2 = 1 + e_i, so sigma_i - 1 = e_i / (1 + sqrt(1 + e_i)) without
//     cancellation.  U follows by Gram-Schmidt on F V, the sign of det F
//     goes on the smallest singular value (rotation-variant SVD, U, V in
//     SO(3)), R = U V^T and F - R = U diag(sigma - 1) V^T.
//     The eigen-solver is cyclic Jacobi with exact 2x2 rotations
//     (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.5.2-8.5.3).
//   Hessian pieces follow func/_hess_diag.py, _hess_prod.py, _hess_quad.py:
//     SNH  H = lam g3 g3^T + mu I + (-mu + lam (J-1)) d2(det),
//     ARAP H = mu (I - sum_k lambda_k q_k q_k^T), lambda_k = 2/max(s_i+s_j, 2).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace apl {

enum Material : int { MAT_SNH = 0, MAT_ARAP = 1, MAT_MUSCLE = 2, N_MAT = 3 };

template <class R>
struct V3 {
  R x, y, z;
};

template <class R>
__host__ __device__ inline V3<R> v3(R x, R y, R z) {
  return V3<R>{x, y, z};
}
template <class R>
__host__ __device__ inline V3<R> operator+(V3<R> a, V3<R> b) {
  return v3(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class R>
__host__ __device__ inline V3<R> operator-(V3<R> a, V3<R> b) {
  return v3(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class R>
__host__ __device__ inline V3<R> operator*(R s, V3<R> a) {
  return v3(s * a.x, s * a.y, s * a.z);
}
template <class R>
__host__ __device__ inline R dot(V3<R> a, V3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
__host__ __device__ inline V3<R> cross(V3<R> a, V3<R> b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// 3x3 matrix, row-major m[i][j]
template <class R>
struct M3 {
  R m[3][3];
};

template <class R>
__host__ __device__ inline V3<R> col(const M3<R>& A, int j) {
  return v3(A.m[0][j], A.m[1][j], A.m[2][j]);
}
template <class R>
__host__ __device__ inline void set_col(M3<R>& A, int j, V3<R> c) {
  A.m[0][j] = c.x;
  A.m[1][j] = c.y;
  A.m[2][j] = c.z;
}
template <class R>
__host__ __device__ inline V3<R> mul(const M3<R>& A, V3<R> d) {
  return v3(A.m[0][0] * d.x + A.m[0][1] * d.y + A.m[0][2] * d.z,
            A.m[1][0] * d.x + A.m[1][1] * d.y + A.m[1][2] * d.z,
            A.m[2][0] * d.x + A.m[2][1] * d.y + A.m[2][2] * d.z);
}
template <class R>
__host__ __device__ inline R ddot(const M3<R>& A, const M3<R>& B) {
  R s = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) s += A.m[i][j] * B.m[i][j];
  return s;
}
template <class R>
__host__ __device__ inline R det3(const M3<R>& A) {
  return A.m[0][0] * (A.m[1][1] * A.m[2][2] - A.m[1][2] * A.m[2][1]) -
         A.m[0][1] * (A.m[1][0] * A.m[2][2] - A.m[1][2] * A.m[2][0]) +
         A.m[0][2] * (A.m[1][0] * A.m[2][1] - A.m[1][1] * A.m[2][0]);
}
// cofactor matrix: columns [c1 x c2, c2 x c0, c0 x c1] (func/_gradient.py:30-40)
template <class R>
__host__ __device__ inline M3<R> cof3(const M3<R>& A) {
  M3<R> C;
  V3<R> c0 = col(A, 0), c1 = col(A, 1), c2 = col(A, 2);
  set_col(C, 0, cross(c1, c2));
  set_col(C, 1, cross(c2, c0));
  set_col(C, 2, cross(c0, c1));
  return C;
}

// Per-tet geometry: rows d[0..3] of dhdX (d0 = -(d1+d2+d3)) and dV.
template <class R>
struct TetGeom {
  V3<R> d[4];
  R dV;
};

template <class R>
__host__ __device__ inline TetGeom<R> make_geom(const R B[9], R dV) {
  TetGeom<R> g;
  g.d[1] = v3(B[0], B[1], B[2]);
  g.d[2] = v3(B[3], B[4], B[5]);
  g.d[3] = v3(B[6], B[7], B[8]);
  g.d[0] = v3(-(B[0] + B[3] + B[6]), -(B[1] + B[4] + B[7]), -(B[2] + B[5] + B[8]));
  g.dV = dV;
  return g;
}

// G = sum_a u_a (x) d_a  (func/_deformation.py:9-25: u^T dhdX, also the jvp)
template <class R>
__host__ __device__ inline M3<R> outer_sum(const TetGeom<R>& g, const V3<R> u[4]) {
  M3<R> G;
  const R* uc[4] = {&u[0].x, &u[1].x, &u[2].x, &u[3].x};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R ui0 = uc[0][i], ui1 = uc[1][i], ui2 = uc[2][i], ui3 = uc[3][i];
    G.m[i][0] = ui0 * g.d[0].x + ui1 * g.d[1].x + ui2 * g.d[2].x + ui3 * g.d[3].x;
    G.m[i][1] = ui0 * g.d[0].y + ui1 * g.d[1].y + ui2 * g.d[2].y + ui3 * g.d[3].y;
    G.m[i][2] = ui0 * g.d[0].z + ui1 * g.d[1].z + ui2 * g.d[2].z + ui3 * g.d[3].z;
  }
  return G;
}

// G from edge differences u_a - u_0 (a = 1..3; the d_a sum to zero, so the
// corner-0 term drops out)
template <class R>
__host__ __device__ inline M3<R> outer_sum_edges(const TetGeom<R>& g, const V3<R> e[4]) {
  M3<R> G;
  const R* ec[4] = {&e[0].x, &e[1].x, &e[2].x, &e[3].x};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R e1 = ec[1][i], e2 = ec[2][i], e3 = ec[3][i];
    G.m[i][0] = e1 * g.d[1].x + e2 * g.d[2].x + e3 * g.d[3].x;
    G.m[i][1] = e1 * g.d[1].y + e2 * g.d[2].y + e3 * g.d[3].y;
    G.m[i][2] = e1 * g.d[1].z + e2 * g.d[2].z + e3 * g.d[3].z;
  }
  return G;
}

// M d_a for the four corners; corner 0 as minus the sum of the others (the
// d_a sum to zero), which keeps each tet's corner forces summing to zero.
template <class R>
__host__ __device__ inline void corner_products(const M3<R>& M, const TetGeom<R>& g, V3<R> out[4]) {
#pragma unroll
  for (int a = 1; a < 4; ++a) out[a] = mul(M, g.d[a]);
  out[0] = v3(-(out[1].x + out[2].x + out[3].x), -(out[1].y + out[2].y + out[3].y),
              -(out[1].z + out[2].z + out[3].z));
}

// Invariants of F = I + G without cancellation.
template <class R>
struct Inv {
  R trG, c2, detG, Jm1, G2;
};

template <class R>
__host__ __device__ inline Inv<R> invariants(const M3<R>& G) {
  Inv<R> v;
  v.trG = G.m[0][0] + G.m[1][1] + G.m[2][2];
  v.c2 = (G.m[0][0] * G.m[1][1] - G.m[0][1] * G.m[1][0]) +
         (G.m[1][1] * G.m[2][2] - G.m[1][2] * G.m[2][1]) +
         (G.m[0][0] * G.m[2][2] - G.m[0][2] * G.m[2][0]);
  v.detG = det3(G);
  v.Jm1 = v.trG + (v.c2 + v.detG);
  v.G2 = ddot(G, G);
  return v;
}

// cof F with F = I + G: I + tr(G) I - G^T + cof G
template <class R>
__host__ __device__ inline M3<R> cofF_from_G(const M3<R>& G, R trG) {
  M3<R> C = cof3(G);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C.m[i][j] -= G.m[j][i];
  C.m[0][0] += R(1) + trG;
  C.m[1][1] += R(1) + trG;
  C.m[2][2] += R(1) + trG;
  return C;
}

// F - cof F = G + G^T - tr(G) I - cof G
template <class R>
__host__ __device__ inline M3<R> F_minus_cofF(const M3<R>& G, R trG) {
  M3<R> C = cof3(G);
  M3<R> D;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) D.m[i][j] = G.m[i][j] + G.m[j][i] - C.m[i][j];
  D.m[0][0] -= trG;
  D.m[1][1] -= trG;
  D.m[2][2] -= trG;
  return D;
}

// ---------------------------------------------------------------- SVD (ARAP)

template <class R>
__host__ __device__ inline void jacobi_rotate(R (&a)[3][3], M3<R>& V, int p, int q) {
  R apq = a[p][q];
  R scale = fabs(a[p][p]) + fabs(a[q][q]);
  if (!(fabs(apq) > R(1e-30) + scale * (sizeof(R) == 4 ? R(1e-9) : R(1e-18)))) {
    a[p][q] = a[q][p] = R(0);
    return;
  }
  R tau = (a[q][q] - a[p][p]) / (R(2) * apq);
  R t = (tau >= R(0) ? R(1) : R(-1)) / (fabs(tau) + sqrt(R(1) + tau * tau));
  R c = R(1) / sqrt(R(1) + t * t);
  R s = t * c;
  a[p][p] -= t * apq;
  a[q][q] += t * apq;
  a[p][q] = a[q][p] = R(0);
  int r = 3 - p - q;
  R arp = a[r][p], arq = a[r][q];
  a[r][p] = a[p][r] = c * arp - s * arq;
  a[r][q] = a[q][r] = s * arp + c * arq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R vp = V.m[i][p], vq = V.m[i][q];
    V.m[i][p] = c * vp - s * vq;
    V.m[i][q] = s * vp + c * vq;
  }
}

// FP32 rotation with the hardware reciprocal / square-root approximations
// (about 1 ulp): used only to find a basis that FP64 sweeps then finish.
__host__ __device__ inline float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
__host__ __device__ inline float fast_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
__host__ __device__ inline float fast_rsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

// 1/x and 1/sqrt(x): the hardware approximations (about 1 ulp) in FP32,
// correctly rounded division and square root in FP64
__host__ __device__ inline float recip(float x) { return fast_rcp(x); }
__host__ __device__ inline double recip(double x) { return 1.0 / x; }
__host__ __device__ inline float rsqrt_r(float x) { return fast_rsq(x); }
__host__ __device__ inline double rsqrt_r(double x) { return 1.0 / sqrt(x); }

__host__ __device__ inline void jacobi_rotate_fast(float (&a)[3][3], M3<float>& V, int p, int q) {
  const float apq = a[p][q];
  const float scale = fabsf(a[p][p]) + fabsf(a[q][q]);
  if (!(fabsf(apq) > 1e-30f + scale * 1e-9f)) {
    a[p][q] = a[q][p] = 0.0f;
    return;
  }
  const float tau = (a[q][q] - a[p][p]) * fast_rcp(2.0f * apq);
  const float t = (tau >= 0.0f ? 1.0f : -1.0f) * fast_rcp(fabsf(tau) + fast_sqrt(1.0f + tau * tau));
  const float c = fast_rsq(1.0f + t * t);
  const float s = t * c;
  a[p][p] -= t * apq;
  a[q][q] += t * apq;
  a[p][q] = a[q][p] = 0.0f;
  const int r = 3 - p - q;
  const float arp = a[r][p], arq = a[r][q];
  a[r][p] = a[p][r] = c * arp - s * arq;
  a[r][q] = a[q][r] = s * arp + c * arq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float vp = V.m[i][p], vq = V.m[i][q];
    V.m[i][p] = c * vp - s * vq;
    V.m[i][q] = s * vp + c * vq;
  }
}

template <class R>
__host__ __device__ inline void swap_cols(M3<R>& V, R (&e)[3], int i, int j) {
  R t = e[i];
  e[i] = e[j];
  e[j] = t;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    R a = V.m[r][i];
    V.m[r][i] = V.m[r][j];
    V.m[r][j] = -a;  // keep det V = +1
  }
}

// sigma and sigma - 1 from the eigenvalue e of F^T F - I and the QR value q:
// near sigma = 1, sigma - 1 = e / (1 + sqrt(1 + e)) (no cancellation).
template <class R>
__host__ __device__ inline void sigma_from(R e, R q, R& s, R& sm1) {
  if (fabs(e) < R(0.25) && q > R(0)) {
    sm1 = e * recip(R(1) + sqrt(R(1) + e));
    s = R(1) + sm1;
  } else {
    s = q;
    sm1 = q - R(1);
  }
}

template <class R>
struct Svd {
  M3<R> U, V;
  R s[3];    // signed singular values, |s0| >= |s1| >= |s2|, sign(s2) = sign(det F)
  R sm1[3];  // s_i - 1 computed without cancellation
};

// Symmetric A = F^T F - I = G + G^T + G^T G (no cancellation near F = I).
template <class R>
__host__ __device__ inline void ftf_minus_i(const M3<R>& G, R (&a)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R gtg = G.m[0][i] * G.m[0][j] + G.m[1][i] * G.m[1][j] + G.m[2][i] * G.m[2][j];
      a[i][j] = G.m[i][j] + G.m[j][i] + gtg;
    }
}

template <class R>
__host__ __device__ inline void jacobi_sweeps(R (&a)[3][3], M3<R>& V, int sweeps) {
  for (int sw = 0; sw < sweeps; ++sw) {
    jacobi_rotate(a, V, 0, 1);
    jacobi_rotate(a, V, 0, 2);
    jacobi_rotate(a, V, 1, 2);
  }
}

// Rotation-variant SVD of F = I + G (see file header), from the symmetric
// eigen-decomposition a = V diag(e) V^T of F^T F - I that the caller has
// already brought to diagonal form.
template <class R>
__host__ __device__ inline void sort_eig(const R (&a)[3][3], M3<R>& V, R (&e)[3]) {
  e[0] = a[0][0];
  e[1] = a[1][1];
  e[2] = a[2][2];
  if (e[0] < e[1]) swap_cols(V, e, 0, 1);
  if (e[0] < e[2]) swap_cols(V, e, 0, 2);
  if (e[1] < e[2]) swap_cols(V, e, 1, 2);
}

// U and sigma from the sorted eigenpairs (e, V) of F^T F - I.
template <class R>
__host__ __device__ inline Svd<R> qr_sigma(const M3<R>& G, const M3<R>& V0, const R (&e)[3]) {
  Svd<R> r;
  M3<R>& V = r.V;
  V = V0;
  // U and the singular values by Gram-Schmidt QR of F V, F = I + G
  // (the QR step of McAdams et al., "Computing the SVD of 3x3 matrices with
  // minimal branching", 2011): sigma = diag(R) keeps absolute accuracy
  // eps*|F| even for near-singular F, unlike sqrt(1 + e_i).
  V3<R> v0 = col(V, 0), v1 = col(V, 1), v2 = col(V, 2);
  V3<R> f0 = v0 + mul(G, v0);
  V3<R> f1 = v1 + mul(G, v1);
  V3<R> f2 = v2 + mul(G, v2);
  // squared-norm floors: normal numbers in either precision
  const R tiny = sizeof(R) == 4 ? R(1e-36) : R(1e-60);
  const R d0 = dot(f0, f0);
  R n0 = R(0);
  V3<R> u0 = v0;
  if (d0 > tiny) {
    const R r0 = rsqrt_r(d0);
    n0 = d0 * r0;
    u0 = r0 * f0;
  }
  V3<R> w = f1 - dot(u0, f1) * u0;
  const R d1 = dot(w, w);
  R n1 = R(0);
  V3<R> u1;
  if (d1 > tiny * (R(1) + n0) * (R(1) + n0)) {
    const R r1 = rsqrt_r(d1);
    n1 = d1 * r1;
    u1 = r1 * w;
  } else {
    // any unit vector orthogonal to u0
    V3<R> t = fabs(u0.x) < R(0.6) ? v3(R(1), R(0), R(0)) : v3(R(0), R(1), R(0));
    w = t - dot(u0, t) * u0;
    u1 = rsqrt_r(dot(w, w)) * w;
  }
  V3<R> u2 = cross(u0, u1);
  set_col(r.U, 0, u0);
  set_col(r.U, 1, u1);
  set_col(r.U, 2, u2);
  sigma_from(e[0], n0, r.s[0], r.sm1[0]);
  sigma_from(e[1], n1, r.s[1], r.sm1[1]);
  sigma_from(e[2], dot(u2, f2), r.s[2], r.sm1[2]);
  return r;
}

template <class R>
__host__ __device__ inline Svd<R> svd_rv(const M3<R>& G) {
  R a[3][3];
  ftf_minus_i(G, a);
  M3<R> V;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) V.m[i][j] = (i == j) ? R(1) : R(0);
  jacobi_sweeps(a, V, sizeof(R) == 4 ? 4 : 6);
  R e[3];
  sort_eig(a, V, e);
  return qr_sigma(G, V, e);
}

// The same SVD to FP64 accuracy at a fraction of the FP64 work: FP32 Jacobi
// sweeps (hardware-approximate rotations) find the eigenvectors to about
// FP32 accuracy, the FP64 matrix is rotated into that basis (nearly
// diagonal), and FP64 sweeps -- whose rotations are tiny, so quadratic
// convergence finishes them -- complete it.  Three FP32 sweeps give the
// same results as four on random and nearly repeated singular values up to
// |G| ~ 10 (host harness); two do not.
#ifndef REFINE_F32_SWEEPS
#define REFINE_F32_SWEEPS 3
#endif
#ifndef REFINE_F64_SWEEPS
#define REFINE_F64_SWEEPS 1
#endif
// FP64 finish from an FP32 eigenbasis Vf of F^T F - I: re-orthonormalise
// Vf in FP64, rotate the FP64 F^T F - I into it (nearly diagonal), one FP64
// sweep, then the FP64 sort and QR; results rounded to R.
template <class R>
__host__ __device__ inline Svd<R> svd_finish_f64(const M3<R>& G, const M3<float>& Vf) {
  M3<double> Gd;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Gd.m[i][j] = (double)G.m[i][j];
  double ad[3][3], t[3][3], b[3][3];
  ftf_minus_i(Gd, ad);
  V3<double> c0 = v3((double)Vf.m[0][0], (double)Vf.m[1][0], (double)Vf.m[2][0]);
  V3<double> c1 = v3((double)Vf.m[0][1], (double)Vf.m[1][1], (double)Vf.m[2][1]);
  c0 = (1.0 / sqrt(dot(c0, c0))) * c0;
  c1 = c1 - dot(c0, c1) * c0;
  c1 = (1.0 / sqrt(dot(c1, c1))) * c1;
  const V3<double> c2 = cross(c0, c1);
  M3<double> V;
  set_col(V, 0, c0);
  set_col(V, 1, c1);
  set_col(V, 2, c2);
  // b = V^T ad V
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) t[i][j] = ad[i][0] * V.m[0][j] + ad[i][1] * V.m[1][j] + ad[i][2] * V.m[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) b[i][j] = V.m[0][i] * t[0][j] + V.m[1][i] * t[1][j] + V.m[2][i] * t[2][j];
  jacobi_sweeps(b, V, REFINE_F64_SWEEPS);
  double ed[3];
  sort_eig(b, V, ed);
  const Svd<double> Sd = qr_sigma(Gd, V, ed);
  Svd<R> S;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    S.s[i] = (R)Sd.s[i];
    S.sm1[i] = (R)Sd.sm1[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      S.U.m[i][j] = (R)Sd.U.m[i][j];
      S.V.m[i][j] = (R)Sd.V.m[i][j];
    }
  }
  return S;
}

template <class R>
__host__ __device__ inline Svd<R> svd_rv_refined(const M3<R>& G) {
  M3<float> Gf;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Gf.m[i][j] = (float)G.m[i][j];
  float af[3][3];
  ftf_minus_i(Gf, af);
  M3<float> Vf;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Vf.m[i][j] = (i == j) ? 1.0f : 0.0f;
  for (int sw = 0; sw < REFINE_F32_SWEEPS; ++sw) {
    jacobi_rotate_fast(af, Vf, 0, 1);
    jacobi_rotate_fast(af, Vf, 0, 2);
    jacobi_rotate_fast(af, Vf, 1, 2);
  }
  return svd_finish_f64(G, Vf);
}

// The same SVD with FP32 arithmetic except one step: F^T F - I is formed
// in FP64 from the FP32 G (exact products, one rounding to FP32), which
// removes the cancellation that G + G^T + G^T G suffers in FP32 under a
// large rotation (|G| ~ 1 while the strain is small).  Then FP32 Jacobi
// sweeps with hardware-approximate rotations, the FP32 QR of F V for U and
// the FP32 singular values (sigma - 1 from the eigenvalue near 1).  The
// stress mu U diag(sigma - 1) V^T and the twist-mode projection are smooth
// functions of F^T F, so the eigenvectors' FP32 error within nearly
// degenerate pairs does not reach them.
#ifndef E32_SWEEPS
#define E32_SWEEPS 4
#endif
// F^T F - I in FP64 from the FP32 G, rounded once to FP32
__host__ __device__ inline void ftf_minus_i_f64(const M3<float>& G, float (&a)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      const double gtg = (double)G.m[0][i] * G.m[0][j] + (double)G.m[1][i] * G.m[1][j] +
                         (double)G.m[2][i] * G.m[2][j];
      a[i][j] = a[j][i] = (float)((double)G.m[i][j] + (double)G.m[j][i] + gtg);
    }
}

// FP32 Jacobi sweeps on a (FP32 F^T F - I, diagonalised in place) from V = I
__host__ __device__ inline void jacobi_fast_sweeps(float (&a)[3][3], M3<float>& V) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) V.m[i][j] = (i == j) ? 1.0f : 0.0f;
  for (int sw = 0; sw < E32_SWEEPS; ++sw) {
    jacobi_rotate_fast(a, V, 0, 1);
    jacobi_rotate_fast(a, V, 0, 2);
    jacobi_rotate_fast(a, V, 1, 2);
  }
}

// The ARAP kernels' SVD.  Every element gets the FP32 basis (F^T F - I
// formed in FP64, FP32 Jacobi sweeps); elements with det F > 1/64 (large
// rotations and large strains included) finish in FP32 (sort, QR), the
// inverted and nearly flat ones in FP64 from that basis (svd_finish_f64):
// there the rotation-variant sign convention hinges on the sign of det F and
// on comparisons of nearly equal |sigma| that FP32 can get the other way.
// The threshold keeps the FP32 det3's rounding (~1e-7 |F|^3) far from
// deciding the branch; only the finish is divergent within a wave.
__host__ __device__ inline Svd<float> svd_rv_fast(const M3<float>& G) {
  float a[3][3];
  ftf_minus_i_f64(G, a);
  M3<float> V;
  jacobi_fast_sweeps(a, V);
  M3<float> F = G;
  F.m[0][0] += 1.0f;
  F.m[1][1] += 1.0f;
  F.m[2][2] += 1.0f;
  if (det3(F) > 0.015625f) {
    float e[3];
    sort_eig(a, V, e);
    return qr_sigma(G, V, e);
  }
  return svd_finish_f64(G, V);
}

template <class R>
__host__ __device__ inline Svd<R> arap_svd(const M3<R>& G) {
  if constexpr (sizeof(R) == 4)
    return svd_rv_fast(G);
  else
    return svd_rv<R>(G);
}

// U diag(d) V^T
template <class R>
__host__ __device__ inline M3<R> udv(const Svd<R>& S, const R d[3]) {
  M3<R> M;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      M.m[i][j] = S.U.m[i][0] * d[0] * S.V.m[j][0] + S.U.m[i][1] * d[1] * S.V.m[j][1] +
                  S.U.m[i][2] * d[2] * S.V.m[j][2];
  return M;
}

// lambda_k = 2 / max(s_i + s_j, 2) over pairs (0,1), (1,2), (2,0)  (func/_misc.py:31-43)
template <class R>
__host__ __device__ inline void arap_lambdas(const R s[3], R lam[3]) {
  R p0 = s[0] + s[1], p1 = s[1] + s[2], p2 = s[2] + s[0];
  lam[0] = R(2) * recip(p0 > R(2) ? p0 : R(2));
  lam[1] = R(2) * recip(p1 > R(2) ? p1 : R(2));
  lam[2] = R(2) * recip(p2 > R(2) ? p2 : R(2));
}

// Twist-mode coordinates of a vector d: (Q0 d, Q1 d, Q2 d) (func/_misc.py:56-70)
//   Q0 = (U1 V0^T - U0 V1^T)/sqrt2, Q1 = (U1 V2^T - U2 V1^T)/sqrt2, Q2 = (U0 V2^T - U2 V0^T)/sqrt2
template <class R>
__host__ __device__ inline void twist_apply(const Svd<R>& S, V3<R> d, V3<R> out[3]) {
  const R r2 = R(0.70710678118654752440);
  R w0 = S.V.m[0][0] * d.x + S.V.m[1][0] * d.y + S.V.m[2][0] * d.z;
  R w1 = S.V.m[0][1] * d.x + S.V.m[1][1] * d.y + S.V.m[2][1] * d.z;
  R w2 = S.V.m[0][2] * d.x + S.V.m[1][2] * d.y + S.V.m[2][2] * d.z;
  V3<R> U0 = col(S.U, 0), U1 = col(S.U, 1), U2 = col(S.U, 2);
  out[0] = r2 * (w0 * U1 - w1 * U0);
  out[1] = r2 * (w2 * U1 - w1 * U2);
  out[2] = r2 * (w2 * U0 - w0 * U2);
}

// Q_k : dF for k = 0,1,2 via M = U^T dF V
template <class R>
__host__ __device__ inline void twist_coords(const Svd<R>& S, const M3<R>& dF, R q[3], R diagM[3]) {
  R M[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    V3<R> Ua = col(S.U, a);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      V3<R> Vb = col(S.V, b);
      M[a][b] = dot(Ua, mul(dF, Vb));
    }
  }
  const R r2 = R(0.70710678118654752440);
  q[0] = r2 * (M[1][0] - M[0][1]);
  q[1] = r2 * (M[1][2] - M[2][1]);
  q[2] = r2 * (M[0][2] - M[2][0]);
  diagM[0] = M[0][0];
  diagM[1] = M[1][1];
  diagM[2] = M[2][2];
}

// sum_k c_k Q_k as a matrix
template <class R>
__host__ __device__ inline M3<R> twist_combine(const Svd<R>& S, const R c[3]) {
  const R r2 = R(0.70710678118654752440);
  M3<R> M;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R q0 = S.U.m[i][1] * S.V.m[j][0] - S.U.m[i][0] * S.V.m[j][1];
      R q1 = S.U.m[i][1] * S.V.m[j][2] - S.U.m[i][2] * S.V.m[j][1];
      R q2 = S.U.m[i][0] * S.V.m[j][2] - S.U.m[i][2] * S.V.m[j][0];
      M.m[i][j] = r2 * (c[0] * q0 + c[1] * q1 + c[2] * q2);
    }
  return M;
}

// ---------------------------------------------------------------- per-tet evaluation

enum : unsigned {
  W_E = 1u,      // energy  sum Psi dV
  W_G = 2u,      // gradient (4 x vec3)
  W_H = 4u,      // clamped Hessian diagonal (4 x vec3)
  W_HP = 8u,     // Hessian-vector product (4 x vec3)
  W_HQ = 16u,    // clamped quadratic form v^T H v
  W_MIX = 32u,   // mixed derivatives p^T dg/dmu, p^T dg/dlam (per tet)
  W_MIXA = 64u,  // mixed derivatives p^T dg/da_k, k = 0..5 (muscle activation, per tet)
};

template <class R>
struct TetOut {
  R e;          // Psi dV
  V3<R> g[4];   // dV P d_a
  V3<R> h[4];   // max(dV diag H, 0)
  V3<R> hp[4];  // dV H v
  R hq;         // max(dV v^T H v, 0)
  R dmu, dlam;  // dV (dP/dmu):dF(v), dV (dP/dlam):dF(v)
  R dact[6];    // dV d(P:dF(v))/da_k (muscle)
};

// Evaluate the requested quantities for one tet.
// u: edge differences (u[a] = displacement of corner a minus that of corner
// 0; u[0] is not read); v: the direction at the four corners for W_HP /
// W_HQ / W_MIX / W_MIXA; act: the six activation values (muscle only).
template <int MAT, unsigned WHAT, class R>
__host__ __device__ inline void tet_eval(const TetGeom<R>& g0, R mu, R lam, const R* act, const V3<R> u[4],
                                         const V3<R> v[4], TetOut<R>& o) {
  constexpr bool needV = (WHAT & (W_HP | W_HQ | W_MIX | W_MIXA)) != 0;
  static_assert(MAT == MAT_MUSCLE || (WHAT & W_MIXA) == 0, "activation derivatives need the muscle law");
  // muscle: the SNH law on the geometry d'_a = A d_a, G offset by S = A - I
  TetGeom<R> g = g0;
  M3<R> S;
  if constexpr (MAT == MAT_MUSCLE) {
    S.m[0][0] = act[0];
    S.m[1][1] = act[1];
    S.m[2][2] = act[2];
    S.m[0][1] = S.m[1][0] = act[3];
    S.m[0][2] = S.m[2][0] = act[4];
    S.m[1][2] = S.m[2][1] = act[5];
#pragma unroll
    for (int a = 1; a < 4; ++a) g.d[a] = g0.d[a] + mul(S, g0.d[a]);
    g.d[0] = v3(-(g.d[1].x + g.d[2].x + g.d[3].x), -(g.d[1].y + g.d[2].y + g.d[3].y),
                -(g.d[1].z + g.d[2].z + g.d[3].z));
  }
  M3<R> G = outer_sum_edges(g, u);
  if constexpr (MAT == MAT_MUSCLE) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) G.m[i][j] += S.m[i][j];
  }
  const Inv<R> I = invariants(G);
  M3<R> dF;  // dF(v) A for the muscle (the direction in G)
  if constexpr (needV) dF = outer_sum(g, v);
  // dV folded into the moduli once: every output below is per unit volume
  // times dV (warp/fem/_base.py multiplies each result by dV separately)
  const R muV = g.dV * mu;
  if constexpr (MAT != MAT_ARAP) {
    const R lamV = g.dV * lam;
    if constexpr ((WHAT & W_E) != 0)
      o.e = R(0.5) * muV * I.G2 - muV * (I.c2 + I.detG) + R(0.5) * lamV * I.Jm1 * I.Jm1;
    constexpr bool needC = (WHAT & (W_G | W_H | W_HP | W_HQ | W_MIX | W_MIXA)) != 0;
    if constexpr (needC) {
      const M3<R> C = cofF_from_G(G, I.trG);  // g3 = cof F
      M3<R> P;   // first Piola-Kirchhoff stress times dV
      M3<R> Tm;  // Hessian applied to dF, times dV
      if constexpr ((WHAT & (W_G | W_MIXA)) != 0) {
        const M3<R> D = F_minus_cofF(G, I.trG);
        const R lj = lamV * I.Jm1;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) P.m[i][j] = muV * D.m[i][j] + lj * C.m[i][j];
        // muscle: P_F = P_G A^T, and P_F d_a = P_G d'_a
        if constexpr ((WHAT & W_G) != 0) corner_products(P, g, o.g);
      }
      if constexpr ((WHAT & W_H) != 0) {
        V3<R> wc[4];
        corner_products(C, g, wc);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const V3<R> w = wc[a];
          R t = muV * dot(g.d[a], g.d[a]);
          R hx = lamV * w.x * w.x + t;
          R hy = lamV * w.y * w.y + t;
          R hz = lamV * w.z * w.z + t;
          o.h[a] = v3(hx > R(0) ? hx : R(0), hy > R(0) ? hy : R(0), hz > R(0) ? hz : R(0));
        }
      }
      if constexpr ((WHAT & (W_HP | W_HQ | W_MIX | W_MIXA)) != 0) {
        const R sV = lamV * I.Jm1 - muV;
        const R cd = ddot(C, dF);
        if constexpr ((WHAT & W_HQ) != 0) {
          // 2 F : cof(dF) = d^2/dt^2 det(F + t dF)  (h6_quad, func/_hess_quad.py:59-74)
          M3<R> cdF = cof3(dF);
          R h6 = R(2) * (cdF.m[0][0] + cdF.m[1][1] + cdF.m[2][2] + ddot(G, cdF));
          R q = lamV * cd * cd + muV * ddot(dF, dF) + sV * h6;
          o.hq = q > R(0) ? q : R(0);
        }
        if constexpr ((WHAT & (W_HP | W_MIXA)) != 0) {
          // Hp = d/dt cof(F + t dF) (h6_prod, func/_hess_prod.py:61-77)
          V3<R> f0 = v3(R(1) + G.m[0][0], G.m[1][0], G.m[2][0]);
          V3<R> f1 = v3(G.m[0][1], R(1) + G.m[1][1], G.m[2][1]);
          V3<R> f2 = v3(G.m[0][2], G.m[1][2], R(1) + G.m[2][2]);
          V3<R> p0 = col(dF, 0), p1 = col(dF, 1), p2 = col(dF, 2);
          M3<R> Hs;
          set_col(Hs, 0, cross(f1, p2) - cross(f2, p1));
          set_col(Hs, 1, cross(f2, p0) - cross(f0, p2));
          set_col(Hs, 2, cross(f0, p1) - cross(f1, p0));
          const R lc = lamV * cd;
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Tm.m[i][j] = lc * C.m[i][j] + muV * dF.m[i][j] + sV * Hs.m[i][j];
          if constexpr ((WHAT & W_HP) != 0) corner_products(Tm, g, o.hp);
        }
        if constexpr ((WHAT & W_MIX) != 0) {
          const M3<R> D = F_minus_cofF(G, I.trG);
          o.dmu = g.dV * ddot(D, dF);
          o.dlam = g.dV * I.Jm1 * cd;
        }
        if constexpr ((WHAT & W_MIXA) != 0) {
          // d/da_k [P_G(F A) : (dF0 A)] = (F^T H[dG] + dF0^T P_G) : E_k with
          // E_k = dA/da_k and F, dF0 in the rest geometry (H symmetric:
          // H[F E_k] : dG = H[dG] : F E_k)
          const M3<R> F0 = outer_sum_edges(g0, u);
          const M3<R> dF0 = outer_sum(g0, v);
          M3<R> N;
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              N.m[i][j] = Tm.m[i][j] + (F0.m[0][i] * Tm.m[0][j] + F0.m[1][i] * Tm.m[1][j] + F0.m[2][i] * Tm.m[2][j]) +
                          (dF0.m[0][i] * P.m[0][j] + dF0.m[1][i] * P.m[1][j] + dF0.m[2][i] * P.m[2][j]);
          o.dact[0] = N.m[0][0];
          o.dact[1] = N.m[1][1];
          o.dact[2] = N.m[2][2];
          o.dact[3] = N.m[0][1] + N.m[1][0];
          o.dact[4] = N.m[0][2] + N.m[2][0];
          o.dact[5] = N.m[1][2] + N.m[2][1];
        }
      }
    }
  } else {  // ARAP
    const Svd<R> S = arap_svd(G);
    if constexpr ((WHAT & W_E) != 0)
      o.e = R(0.5) * muV * (S.sm1[0] * S.sm1[0] + S.sm1[1] * S.sm1[1] + S.sm1[2] * S.sm1[2]);
    if constexpr ((WHAT & W_G) != 0) {
      R d[3] = {muV * S.sm1[0], muV * S.sm1[1], muV * S.sm1[2]};
      const M3<R> P = udv(S, d);
      corner_products(P, g, o.g);
    }
    R lk[3];
    if constexpr ((WHAT & (W_H | W_HP | W_HQ)) != 0) arap_lambdas(S.s, lk);
    if constexpr ((WHAT & W_H) != 0) {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        V3<R> qd[3];
        twist_apply(S, g.d[a], qd);
        R t = dot(g.d[a], g.d[a]);
        R hx = muV * (t - (lk[0] * qd[0].x * qd[0].x + lk[1] * qd[1].x * qd[1].x + lk[2] * qd[2].x * qd[2].x));
        R hy = muV * (t - (lk[0] * qd[0].y * qd[0].y + lk[1] * qd[1].y * qd[1].y + lk[2] * qd[2].y * qd[2].y));
        R hz = muV * (t - (lk[0] * qd[0].z * qd[0].z + lk[1] * qd[1].z * qd[1].z + lk[2] * qd[2].z * qd[2].z));
        o.h[a] = v3(hx > R(0) ? hx : R(0), hy > R(0) ? hy : R(0), hz > R(0) ? hz : R(0));
      }
    }
    if constexpr ((WHAT & (W_HP | W_HQ | W_MIX)) != 0) {
      R q[3], dM[3];
      twist_coords(S, dF, q, dM);
      if constexpr ((WHAT & W_HQ) != 0) {
        R val = muV * (ddot(dF, dF) - (lk[0] * q[0] * q[0] + lk[1] * q[1] * q[1] + lk[2] * q[2] * q[2]));
        o.hq = val > R(0) ? val : R(0);
      }
      if constexpr ((WHAT & W_HP) != 0) {
        R c[3] = {lk[0] * q[0], lk[1] * q[1], lk[2] * q[2]};
        M3<R> Qc = twist_combine(S, c);
        M3<R> T;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) T.m[i][j] = muV * (dF.m[i][j] - Qc.m[i][j]);
        corner_products(T, g, o.hp);
      }
      if constexpr ((WHAT & W_MIX) != 0) {
        o.dmu = g.dV * (S.sm1[0] * dM[0] + S.sm1[1] * dM[1] + S.sm1[2] * dM[2]);
        o.dlam = R(0);
      }
    }
  }
}

// SNH and ARAP need no activation
template <int MAT, unsigned WHAT, class R>
__host__ __device__ inline void tet_eval(const TetGeom<R>& g, R mu, R lam, const V3<R> u[4], const V3<R> v[4],
                                         TetOut<R>& o) {
  static_assert(MAT != MAT_MUSCLE, "the muscle law needs the activation");
  tet_eval<MAT, WHAT, R>(g, mu, lam, static_cast<const R*>(nullptr), u, v, o);
}

}  // namespace apl
