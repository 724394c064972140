// Host harness for the per-tet algebra of apple_amd/csrc/tetmath.h: checks
// the rotation-variant SVDs of the FP32 ARAP kernels against the plain FP64
// Jacobi SVD, on random deformation gradients, on ones with (nearly)
// repeated singular values, and on large rotations times small strains
// (where F^T F - I formed in FP32 would cancel).  Prints one line per case
// and SVD (column 0: 0 = svd_rv_refined, 1 = svd_rv_fast, the kernels'
// choice): the case, the worst relative error of mu (F - R), of the ARAP
// energy density, the worst |sigma| error, the worst orthogonality defect of
// U and V, and the worst relative error of the projected twist-mode part
// sum_k lambda_k (Q_k : dF) Q_k of the Hessian for random dF.
#include "../../apple_amd/csrc/tetmath.h"

#include <cmath>
#include <cstdio>
#include <random>

using namespace apl;

template <class S>
static double proj_err(const Svd<double>& Sd, const S& Sf, std::mt19937& rng) {
  std::normal_distribution<double> N(0, 1);
  double lamd[3];
  float lamf[3];
  arap_lambdas(Sd.s, lamd);
  arap_lambdas(Sf.s, lamf);
  double worst = 0;
  for (int t = 0; t < 3; ++t) {
    M3<double> dFd;
    M3<float> dFf;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        dFf.m[i][j] = (float)N(rng);
        dFd.m[i][j] = dFf.m[i][j];
      }
    double qd[3], md[3], cd[3];
    float qf[3], mf[3], cf[3];
    twist_coords(Sd, dFd, qd, md);
    twist_coords(Sf, dFf, qf, mf);
    for (int k = 0; k < 3; ++k) {
      cd[k] = lamd[k] * qd[k];
      cf[k] = lamf[k] * qf[k];
    }
    const M3<double> Md = twist_combine(Sd, cd);
    const M3<float> Mf = twist_combine(Sf, cf);
    double n = 1e-300, e = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        n = std::fmax(n, std::fabs(dFd.m[i][j]));
        e = std::fmax(e, std::fabs(Md.m[i][j] - Mf.m[i][j]));
      }
    worst = std::fmax(worst, e / n);
  }
  return worst;
}

// mode 0: random G of the given scale; 1: nearly repeated singular values;
// 2: F = R0 (I + S), R0 a random rotation of up to several radians, S a
// random symmetric strain of the given scale
static M3<double> sample(std::mt19937& rng, double scale, int mode) {
  std::normal_distribution<double> N(0, 1);
  M3<double> G;
  if (mode == 2) {
    double ax[3] = {N(rng), N(rng), N(rng)};
    const double nn = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    for (double& c : ax) c /= nn;
    const double th = 3.0 * std::fabs(N(rng)), c = std::cos(th), s = std::sin(th);
    const double K[3][3] = {{0, -ax[2], ax[1]}, {ax[2], 0, -ax[0]}, {-ax[1], ax[0], 0}};
    double R0[3][3], S[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double kk = 0;
        for (int l = 0; l < 3; ++l) kk += K[i][l] * K[l][j];
        R0[i][j] = (i == j) + s * K[i][j] + (1 - c) * kk;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) S[i][j] = S[j][i] = scale * N(rng);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double f = 0;
        for (int l = 0; l < 3; ++l) f += R0[i][l] * ((l == j) + S[l][j]);
        G.m[i][j] = f - (i == j);
      }
    return G;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) G.m[i][j] = mode ? 0.0 : scale * N(rng);
  if (mode == 1) {
    const double a = scale * N(rng);
    G.m[0][0] = a;
    G.m[1][1] = a + 1e-6 * N(rng);
    G.m[2][2] = scale * N(rng);
  }
  return G;
}

static void run(double scale, int mode, int n) {
  std::mt19937 rng(7);
  double wp[2] = {0, 0}, we[2] = {0, 0}, ws[2] = {0, 0}, wo[2] = {0, 0}, wq[2] = {0, 0};
  for (int it = 0; it < n; ++it) {
    M3<double> Gd = sample(rng, scale, mode);
    M3<float> Gf;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        Gf.m[i][j] = (float)Gd.m[i][j];
        Gd.m[i][j] = (double)Gf.m[i][j];
      }
    const Svd<double> Sd = svd_rv<double>(Gd);
    const double dd[3] = {Sd.sm1[0], Sd.sm1[1], Sd.sm1[2]};
    const M3<double> Pd = udv(Sd, dd);
    const Svd<float> Sv[2] = {svd_rv_refined<float>(Gf), svd_rv_fast(Gf)};
    for (int k = 0; k < 2; ++k) {
      const Svd<float>& Sr = Sv[k];
      const float df[3] = {Sr.sm1[0], Sr.sm1[1], Sr.sm1[2]};
      const M3<float> Pf = udv(Sr, df);
      double pn = 1e-300, pe = 0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          pn = std::fmax(pn, std::fabs(Pd.m[i][j]));
          pe = std::fmax(pe, std::fabs(Pd.m[i][j] - (double)Pf.m[i][j]));
        }
      wp[k] = std::fmax(wp[k], pe / pn);
      const double ed = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
      const double ef = (double)df[0] * df[0] + (double)df[1] * df[1] + (double)df[2] * df[2];
      if (ed > 1e-30) we[k] = std::fmax(we[k], std::fabs(ed - ef) / ed);
      for (int q = 0; q < 3; ++q) ws[k] = std::fmax(ws[k], std::fabs(Sd.s[q] - Sr.s[q]));
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double u = 0, v = 0;
          for (int q = 0; q < 3; ++q) {
            u += (double)Sr.U.m[q][a] * Sr.U.m[q][b];
            v += (double)Sr.V.m[q][a] * Sr.V.m[q][b];
          }
          wo[k] = std::fmax(wo[k], std::fmax(std::fabs(u - (a == b)), std::fabs(v - (a == b))));
        }
      wq[k] = std::fmax(wq[k], proj_err(Sd, Sr, rng));
    }
  }
  for (int k = 0; k < 2; ++k)
    std::printf("%d %d %g %.3e %.3e %.3e %.3e %.3e\n", k, mode, scale, wp[k], we[k], ws[k], wo[k], wq[k]);
}

int main() {
  for (double sc : {1e-4, 1e-2, 0.3, 1.0}) {
    run(sc, 0, 20000);
    run(sc, 1, 20000);
  }
  for (double sc : {1e-5, 1e-4, 1e-2}) run(sc, 2, 20000);
  return 0;
}
