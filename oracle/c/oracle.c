/* C restatement of liblaf/apple's stable Neo-Hookean element passes and the
 * PNCG loop -- TEST INFRASTRUCTURE ONLY (the CPU baseline that bench.py times
 * beside the GPU path; checked against the numpy oracle in tests/).
 *
 * Structure follows the reference's kernel templates (warp/fem/_base.py:243-383):
 * separate passes for fun, grad and hess_diag, each one thread per tet that
 * re-gathers u (func/_misc.py:22-28), recomputes F = I + u^T dhdX
 * (func/_deformation.py:15-19) and scatters with atomic adds.  The SNH
 * algebra follows warp/fem/_stable_neo_hookean.py:17-103 literally
 * (I2 - 3, J - 1, g3 = cof F, h3/h5/h6 terms), in FP64.  The PNCG loop
 * restates benches/bench_pncg_branching_backends.py:254-330,393-456,606-679
 * over the free DOFs with the approximate-Wolfe acceptance described in
 * oracle/pncg.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double m[3][3];
} mat3;

static inline void gather(const double* u, const int32_t* c, double uc[4][3]) {
  for (int a = 0; a < 4; ++a)
    for (int i = 0; i < 3; ++i) uc[a][i] = u[3 * (int64_t)c[a] + i];
}

/* F = I + u^T dhdX: F_ij = delta_ij + sum_a u_a,i dhdX_a,j */
static inline mat3 defgrad(double uc[4][3], const double* dh) {
  mat3 F;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = (i == j) ? 1.0 : 0.0;
      for (int a = 0; a < 4; ++a) s += uc[a][i] * dh[3 * a + j];
      F.m[i][j] = s;
    }
  return F;
}

static inline double det3(const mat3* F) {
  const double(*m)[3] = F->m;
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

/* g3 = [f1 x f2, f2 x f0, f0 x f1] as columns (func/_gradient.py:30-40) */
static inline mat3 g3(const mat3* F) {
  mat3 C;
  for (int j = 0; j < 3; ++j) {
    int a = (j + 1) % 3, b = (j + 2) % 3;
    double fa[3] = {F->m[0][a], F->m[1][a], F->m[2][a]};
    double fb[3] = {F->m[0][b], F->m[1][b], F->m[2][b]};
    C.m[0][j] = fa[1] * fb[2] - fa[2] * fb[1];
    C.m[1][j] = fa[2] * fb[0] - fa[0] * fb[2];
    C.m[2][j] = fa[0] * fb[1] - fa[1] * fb[0];
  }
  return C;
}

static inline double snh_psi(const mat3* F, double mu, double lam) {
  double I2 = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) I2 += F->m[i][j] * F->m[i][j];
  double J = det3(F);
  return 0.5 * mu * (I2 - 3.0) - mu * (J - 1.0) + 0.5 * lam * (J - 1.0) * (J - 1.0);
}

static inline void atomic_add(double* p, double v) {
#pragma omp atomic
  *p += v;
}

/* fun_kernel: sum Psi dV (_base.py:243-263) */
double oracle_snh_fun(int64_t T, const int32_t* cells, const double* dhdX, const double* dV, const double* mu,
                      const double* lam, const double* u) {
  double e = 0.0;
#pragma omp parallel for reduction(+ : e) schedule(static)
  for (int64_t t = 0; t < T; ++t) {
    double uc[4][3];
    gather(u, cells + 4 * t, uc);
    mat3 F = defgrad(uc, dhdX + 12 * t);
    e += snh_psi(&F, mu[t], lam[t]) * dV[t];
  }
  return e;
}

/* grad_kernel: scatter dV dhdX P^T (_base.py:265-291) */
void oracle_snh_grad(int64_t T, const int32_t* cells, const double* dhdX, const double* dV, const double* mu,
                     const double* lam, const double* u, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < T; ++t) {
    double uc[4][3];
    const int32_t* c = cells + 4 * t;
    const double* dh = dhdX + 12 * t;
    gather(u, c, uc);
    mat3 F = defgrad(uc, dh);
    double J = det3(&F);
    mat3 G3 = g3(&F);
    double s = -mu[t] + lam[t] * (J - 1.0);
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i) {
        double v = 0;
        for (int j = 0; j < 3; ++j) v += dh[3 * a + j] * (mu[t] * F.m[i][j] + s * G3.m[i][j]);
        atomic_add(out + 3 * (int64_t)c[a] + i, v * dV[t]);
      }
  }
}

/* hess_diag_kernel: scatter max(dV (lam h3 + mu/2 h5), 0) (_base.py:293-324) */
void oracle_snh_hess_diag(int64_t T, const int32_t* cells, const double* dhdX, const double* dV, const double* mu,
                          const double* lam, const double* u, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < T; ++t) {
    double uc[4][3];
    const int32_t* c = cells + 4 * t;
    const double* dh = dhdX + 12 * t;
    gather(u, c, uc);
    mat3 F = defgrad(uc, dh);
    mat3 G3 = g3(&F);
    for (int a = 0; a < 4; ++a) {
      double t5 = 0;
      for (int j = 0; j < 3; ++j) t5 += dh[3 * a + j] * dh[3 * a + j];
      for (int i = 0; i < 3; ++i) {
        double w = 0;
        for (int j = 0; j < 3; ++j) w += dh[3 * a + j] * G3.m[i][j];
        double h = (lam[t] * w * w + 0.5 * mu[t] * 2.0 * t5) * dV[t];
        atomic_add(out + 3 * (int64_t)c[a] + i, h > 0 ? h : 0.0);
      }
    }
  }
}

/* hess_quad_kernel: sum max(dV (lam h3q + mu h5q/2 + s h6q), 0) (_base.py:353-383) */
double oracle_snh_hess_quad(int64_t T, const int32_t* cells, const double* dhdX, const double* dV, const double* mu,
                            const double* lam, const double* u, const double* p) {
  double q = 0.0;
#pragma omp parallel for reduction(+ : q) schedule(static)
  for (int64_t t = 0; t < T; ++t) {
    double uc[4][3], pc[4][3];
    const int32_t* c = cells + 4 * t;
    const double* dh = dhdX + 12 * t;
    gather(u, c, uc);
    gather(p, c, pc);
    mat3 F = defgrad(uc, dh);
    mat3 G3 = g3(&F);
    mat3 dF;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0;
        for (int a = 0; a < 4; ++a) s += pc[a][i] * dh[3 * a + j];
        dF.m[i][j] = s;
      }
    double J = det3(&F);
    double s = -mu[t] + lam[t] * (J - 1.0);
    double cd = 0, dd = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        cd += G3.m[i][j] * dF.m[i][j];
        dd += dF.m[i][j] * dF.m[i][j];
      }
    /* h6_quad = p0.(f1 x p2 - f2 x p1) + p1.(f2 x p0 - f0 x p2) + p2.(f0 x p1 - f1 x p0) */
    double h6 = 0;
    for (int k = 0; k < 3; ++k) {
      int a = (k + 1) % 3, b = (k + 2) % 3;
      double fa[3] = {F.m[0][a], F.m[1][a], F.m[2][a]}, fb[3] = {F.m[0][b], F.m[1][b], F.m[2][b]};
      double pa[3] = {dF.m[0][a], dF.m[1][a], dF.m[2][a]}, pb[3] = {dF.m[0][b], dF.m[1][b], dF.m[2][b]};
      double pk[3] = {dF.m[0][k], dF.m[1][k], dF.m[2][k]};
      double v1[3] = {fa[1] * pb[2] - fa[2] * pb[1], fa[2] * pb[0] - fa[0] * pb[2], fa[0] * pb[1] - fa[1] * pb[0]};
      double v2[3] = {fb[1] * pa[2] - fb[2] * pa[1], fb[2] * pa[0] - fb[0] * pa[2], fb[0] * pa[1] - fb[1] * pa[0]};
      for (int i = 0; i < 3; ++i) h6 += pk[i] * (v1[i] - v2[i]);
    }
    double v = (lam[t] * cd * cd + 0.5 * mu[t] * 2.0 * dd + s * h6) * dV[t];
    q += v > 0 ? v : 0.0;
  }
  return q;
}

/* one "pass A" as the reference runs it: three separate passes */
double oracle_snh_pass(int64_t T, int64_t V, const int32_t* cells, const double* dhdX, const double* dV,
                       const double* mu, const double* lam, const double* u, double* grad, double* hdiag) {
  memset(grad, 0, sizeof(double) * 3 * V);
  memset(hdiag, 0, sizeof(double) * 3 * V);
  double e = oracle_snh_fun(T, cells, dhdX, dV, mu, lam, u);
  oracle_snh_grad(T, cells, dhdX, dV, mu, lam, u, grad);
  oracle_snh_hess_diag(T, cells, dhdX, dV, mu, lam, u, hdiag);
  return e;
}

int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* PNCG on the full layout with a free mask; returns the number of steps,
 * writes status (1 success, 2 max steps, 3 non-finite, 4 stagnation). */
int oracle_snh_pncg(int64_t T, int64_t V, const int32_t* cells, const double* dhdX, const double* dV,
                    const double* mu, const double* lam, const double* fext, const uint8_t* free_dof, double* x,
                    double rtol, int max_steps, int max_backtracking, double wolfe_delta, int* status_out,
                    double* gnorm_out) {
  const int64_t n = 3 * V;
  double *g = calloc(n, 8), *gp = calloc(n, 8), *p = calloc(n, 8), *pp = calloc(n, 8), *hd = calloc(n, 8),
         *xt = calloc(n, 8), *gt = calloc(n, 8), *P = calloc(n, 8);
  double f, gn0, gn;
  int status = 2, steps = 0, restarts = 0;
#define EVAL(xx, gg, ff)                                                                    \
  do {                                                                                      \
    memset(gg, 0, 8 * n);                                                                   \
    ff = oracle_snh_fun(T, cells, dhdX, dV, mu, lam, xx);                                   \
    oracle_snh_grad(T, cells, dhdX, dV, mu, lam, xx, gg);                                   \
    for (int64_t k = 0; k < n; ++k) {                                                       \
      ff -= fext[k] * xx[k];                                                                \
      gg[k] = free_dof[k] ? gg[k] - fext[k] : 0.0;                                          \
    }                                                                                       \
  } while (0)
#define NORM(v) ({ double s_ = 0; for (int64_t k = 0; k < n; ++k) s_ += v[k] * v[k]; sqrt(s_); })
  EVAL(x, g, f);
  gn0 = NORM(g);
  gn = gn0;
  for (;;) {
    gn = NORM(g);
    if (!isfinite(gn) || !isfinite(f)) { status = 3; break; }
    if (gn <= rtol * gn0) { status = 1; break; }
    if (steps >= max_steps) { status = 2; break; }
    memset(hd, 0, 8 * n);
    oracle_snh_hess_diag(T, cells, dhdX, dV, mu, lam, x, hd);
    double hs = 0; int64_t hc = 0;
    for (int64_t k = 0; k < n; ++k) if (free_dof[k] && fabs(hd[k]) > 0) { hs += fabs(hd[k]); ++hc; }
    double mean = hc ? hs / hc : 1.0;
    for (int64_t k = 0; k < n; ++k) P[k] = free_dof[k] ? 1.0 / (fabs(hd[k]) > 0 ? fabs(hd[k]) : mean) : 0.0;
    double beta = 0.0;
    if (steps > 0 && restarts == 0) {
      double yp = 0, gPy = 0, yPy = 0, pg = 0;
      for (int64_t k = 0; k < n; ++k) {
        double y = g[k] - gp[k];
        yp += y * pp[k]; gPy += g[k] * P[k] * y; yPy += y * P[k] * y; pg += pp[k] * g[k];
      }
      beta = fabs(yp) > 1e-300 ? gPy / yp - (yPy / yp) * (pg / yp) : INFINITY;
      if (!isfinite(beta) || fabs(beta) > 10.0) beta = 0.0;
    }
    double gd = 0;
    for (int64_t k = 0; k < n; ++k) { p[k] = -P[k] * g[k] + beta * pp[k]; gd += g[k] * p[k]; }
    if (!(isfinite(gd) && gd < 0)) {
      gd = 0;
      for (int64_t k = 0; k < n; ++k) { p[k] = -P[k] * g[k]; gd += g[k] * p[k]; }
    }
    double pHp = oracle_snh_hess_quad(T, cells, dhdX, dV, mu, lam, x, p);
    double alpha = -gd / pHp;
    if (!(alpha > 0 && isfinite(alpha))) alpha = 1.0;
    int accepted = 0;
    double ft = 0;
    for (int bt = 0; bt <= max_backtracking; ++bt) {
      for (int64_t k = 0; k < n; ++k) xt[k] = x[k] + alpha * p[k];
      EVAL(xt, gt, ft);
      double gtp = 0;
      for (int64_t k = 0; k < n; ++k) gtp += gt[k] * p[k];
      if (isfinite(ft) && (ft <= f + 1e-4 * alpha * gd ||
                           (wolfe_delta > 0 && ft <= f + 1e-12 * fabs(f) && gtp <= (2 * wolfe_delta - 1) * gd))) {
        accepted = 1;
        break;
      }
      alpha *= 0.5;
    }
    if (accepted) {
      memcpy(x, xt, 8 * n);
      memcpy(gp, g, 8 * n);
      memcpy(pp, p, 8 * n);
      memcpy(g, gt, 8 * n);
      f = ft;
      ++steps;
      restarts = 0;
    } else {
      if (restarts >= 1) { status = 4; break; }
      ++restarts;
    }
  }
  *status_out = status;
  *gnorm_out = gn / (gn0 > 0 ? gn0 : 1.0);
  free(g); free(gp); free(p); free(pp); free(hd); free(xt); free(gt); free(P);
  return steps;
}
