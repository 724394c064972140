// Host harness for the per-tet algebra of apple_amd/csrc/tetmath.h: checks
// the rotation-variant SVD used by the FP32 ARAP kernels (FP32 Jacobi basis
// refined by an FP64 sweep) against the plain FP64 Jacobi SVD, on random
// deformation gradients and on ones with (nearly) repeated singular values.
// Prints one line per case: the worst relative error of mu (F - R) and of
// the ARAP energy density, and the worst |sigma| error.
#include "../../apple_amd/csrc/tetmath.h"

#include <cmath>
#include <cstdio>
#include <random>

using namespace apl;

static void run(double scale, int repeated, int n) {
  std::mt19937 rng(7);
  std::normal_distribution<double> N(0, 1);
  double worst_p = 0, worst_e = 0, worst_s = 0, worst_orth = 0;
  for (int it = 0; it < n; ++it) {
    M3<double> Gd;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Gd.m[i][j] = repeated ? 0.0 : scale * N(rng);
    if (repeated) {
      const double a = scale * N(rng);
      Gd.m[0][0] = a;
      Gd.m[1][1] = a + 1e-6 * N(rng);
      Gd.m[2][2] = scale * N(rng);
    }
    M3<float> Gf;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        Gf.m[i][j] = (float)Gd.m[i][j];
        Gd.m[i][j] = (double)Gf.m[i][j];
      }
    const Svd<double> Sd = svd_rv<double>(Gd);
    const Svd<float> Sr = svd_rv_refined<float>(Gf);
    const double dd[3] = {Sd.sm1[0], Sd.sm1[1], Sd.sm1[2]};
    const float df[3] = {Sr.sm1[0], Sr.sm1[1], Sr.sm1[2]};
    const M3<double> Pd = udv(Sd, dd);
    const M3<float> Pf = udv(Sr, df);
    double pn = 1e-300, pe = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        pn = std::fmax(pn, std::fabs(Pd.m[i][j]));
        pe = std::fmax(pe, std::fabs(Pd.m[i][j] - (double)Pf.m[i][j]));
      }
    worst_p = std::fmax(worst_p, pe / pn);
    const double ed = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
    const double ef = (double)df[0] * df[0] + (double)df[1] * df[1] + (double)df[2] * df[2];
    if (ed > 1e-30) worst_e = std::fmax(worst_e, std::fabs(ed - ef) / ed);
    for (int k = 0; k < 3; ++k) worst_s = std::fmax(worst_s, std::fabs(Sd.s[k] - Sr.s[k]));
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double u = 0, v = 0;
        for (int k = 0; k < 3; ++k) {
          u += (double)Sr.U.m[k][a] * Sr.U.m[k][b];
          v += (double)Sr.V.m[k][a] * Sr.V.m[k][b];
        }
        worst_orth = std::fmax(worst_orth, std::fmax(std::fabs(u - (a == b)), std::fabs(v - (a == b))));
      }
  }
  std::printf("%g %d %.3e %.3e %.3e %.3e\n", scale, repeated, worst_p, worst_e, worst_s, worst_orth);
}

int main() {
  for (double sc : {1e-4, 1e-2, 0.3, 1.0}) {
    run(sc, 0, 20000);
    run(sc, 1, 20000);
  }
  return 0;
}
