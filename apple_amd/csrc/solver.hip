// Device-resident PNCG static solver and PCG adjoint solver.
//
// PNCG restates the loop of liblaf-peach as apple states it in
// benches/bench_pncg_branching_backends.py:254-330,393-456,606-679 (the
// external peach package itself is not vendored): Jacobi preconditioner
// P = 1/|diag H| with the mean of the positive entries for zeros, Dai-Kou
// beta reset when non-finite or |beta| > beta_max, steepest-descent fallback
// when g.p >= 0, alpha = -g.p / p^T H p (x overstep), halving backtracking,
// stop at ||g|| <= rtol ||g_0|| (or atol).  The sufficient-decrease test is
// Armijo OR the approximate Wolfe conditions of Hager & Zhang (2005): energy
// differences stop resolving the decrease long before ||g|| reaches 1e-6
// relative in FP32, the directional-derivative test does not.
//
// One iteration is six launches, all with device-side control (no host sync):
//   A  element pass at the trial point x + alpha p: energy, gradient, Hdiag
//   R  vertex pass: sums of shared-vertex partials, masks, FP64 partial dots
//   S  per-instance scalar step: energy/dot totals, accept/reject, beta, gd
//   D  vertex pass: commit x += alpha p, g_cur = g, p = -P g + beta p
//   B  element pass: p^T H p (clamped per tet)
//   S2 per-instance: alpha = -gd / pHp
// A rejected trial skips D/B/S2 work (flags) and the next A re-evaluates at
// the halved step.  K iterations are captured in one HIP graph; the host
// checks the per-instance status once per graph launch.  Every instance of a
// batch runs in the same launches; finished instances drop out via flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/apple_amd.h"
#include "internal.h"
#include "flatsum.h"
#include "persist.h"
#include "vsum.h"
#include "tetmath.h"

namespace apl {

// threads of the per-instance scalar kernels: each sums whole tables of
// per-block values, so more loads in flight shorten them
constexpr int SCAL_TH = 1024;

enum : int { D_G2 = 0, D_GPG, D_GPG0, D_GPGC, D_GPGC0, D_GCPGC, D_GCPGC0, D_GTP, D_GCP, D_HSUM, D_HCNT, D_EXT, D_PP, D_E };
static_assert(D_E < N_DOTS, "dot table too small");
static_assert(D_PP < N_DOTS, "dot slots");

struct StatePtrs {
  double *f, *gd, *alpha, *alpha_commit, *beta, *gnorm0, *gnorm, *pmean, *f_first, *pHp, *gbest, *f_mark, *f0;
  int32_t *status, *act_eval, *act_dir, *commit, *n_steps, *n_evals, *ls_step, *n_backtracks, *phase, *restarts,
      *best_step, *ls_last;
};
constexpr int N_SD = 13, N_SI = 12;

// One instance's state in registers.  The scalar kernels load it whole (all
// loads independent, issued together and overlapped with the table sums),
// decide in registers and store it back whole: the state arrays may alias as
// far as the compiler knows, so reading and writing them field by field
// chains one memory round trip per access.
struct InstState {
  double f, gd, alpha, alpha_commit, beta, gnorm0, gnorm, pmean, f_first, pHp, gbest, f_mark, f0;
  int32_t status, act_eval, act_dir, commit, n_steps, n_evals, ls_step, n_backtracks, phase, restarts, best_step,
      ls_last;
};

__device__ inline InstState load_state(const StatePtrs& st, int i) {
  InstState s;
  s.f = st.f[i], s.gd = st.gd[i], s.alpha = st.alpha[i], s.alpha_commit = st.alpha_commit[i];
  s.beta = st.beta[i], s.gnorm0 = st.gnorm0[i], s.gnorm = st.gnorm[i], s.pmean = st.pmean[i];
  s.f_first = st.f_first[i], s.pHp = st.pHp[i], s.gbest = st.gbest[i], s.f_mark = st.f_mark[i];
  s.status = st.status[i], s.act_eval = st.act_eval[i], s.act_dir = st.act_dir[i], s.commit = st.commit[i];
  s.n_steps = st.n_steps[i], s.n_evals = st.n_evals[i], s.ls_step = st.ls_step[i];
  s.n_backtracks = st.n_backtracks[i], s.phase = st.phase[i], s.restarts = st.restarts[i];
  s.best_step = st.best_step[i];
  s.f0 = st.f0[i], s.ls_last = st.ls_last[i];
  return s;
}

__device__ inline void store_state(const StatePtrs& st, int i, const InstState& s) {
  st.f[i] = s.f, st.gd[i] = s.gd, st.alpha[i] = s.alpha, st.alpha_commit[i] = s.alpha_commit;
  st.beta[i] = s.beta, st.gnorm0[i] = s.gnorm0, st.gnorm[i] = s.gnorm, st.pmean[i] = s.pmean;
  st.f_first[i] = s.f_first, st.pHp[i] = s.pHp, st.gbest[i] = s.gbest, st.f_mark[i] = s.f_mark;
  st.status[i] = s.status, st.act_eval[i] = s.act_eval, st.act_dir[i] = s.act_dir, st.commit[i] = s.commit;
  st.n_steps[i] = s.n_steps, st.n_evals[i] = s.n_evals, st.ls_step[i] = s.ls_step;
  st.n_backtracks[i] = s.n_backtracks, st.phase[i] = s.phase, st.restarts[i] = s.restarts;
  st.best_step[i] = s.best_step;
  st.f0[i] = s.f0, st.ls_last[i] = s.ls_last;
}

struct Params {
  double rtol, atol, overstep, c1, wolfe_delta, wolfe_eps, beta_max;
  int max_steps, max_backtracking, stall_steps;
};

// Workgroup sums of the 16 per-thread dots, in a fixed order, to out[0..16).
// Recursive halving across the wave (offsets 32, 16, 8, 4: each lane keeps
// half of its entries and adds its partner's copy of them) leaves dot
// lane / 4 in lanes 4j..4j+3; two more exchanges complete those sums, and
// the waves are added in order.  17 shuffles instead of 16 x 6.
template <int HALF, int OFF>
__device__ __forceinline__ void halving_step(double (&v)[N_DOTS], int lane) {
  const bool up = (lane & OFF) != 0;
#pragma unroll
  for (int k = 0; k < HALF; ++k) {
    const double send = up ? v[k] : v[k + HALF];
    const double keep = up ? v[k + HALF] : v[k];
    v[k] = keep + __shfl_xor(send, OFF, 64);
  }
}

__device__ inline void block_dots(double (&v)[N_DOTS], double* sm, double* out) {
  static_assert(N_DOTS == 16, "halving tree is written for 16 dots");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  halving_step<8, 32>(v, lane);
  halving_step<4, 16>(v, lane);
  halving_step<2, 8>(v, lane);
  halving_step<1, 4>(v, lane);
  v[0] += __shfl_xor(v[0], 2, 64);
  v[0] += __shfl_xor(v[0], 1, 64);
  if ((lane & 3) == 0) sm[w * N_DOTS + (lane >> 2)] = v[0];
  __syncthreads();
  if ((int)threadIdx.x < N_DOTS) {
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < NTH / 64; ++ww) s += sm[ww * N_DOTS + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// This vertex block's share of its instance's element-block energies: the
// instance's blocks split into equal consecutive runs, one per vertex block
// (both counted from the instance's start, so a batch sums the same way as
// the instance alone).
__device__ inline void energy_share(const double* blk_red, const int64_t* inst_blk_off,
                                    const int64_t* inst_vblk_off, int inst, int vb, double& acc) {
  const long long b0 = inst_blk_off[inst], nb = inst_blk_off[inst + 1] - b0;
  const long long vb0 = inst_vblk_off[inst], nvb = inst_vblk_off[inst + 1] - vb0;
  const long long per = (nb + nvb - 1) / nvb;
  const long long lo = b0 + (vb - vb0) * per;
  const long long hi = lo + per < b0 + nb ? lo + per : b0 + nb;
  for (long long b = lo + threadIdx.x; b < hi; b += NTH) acc += blk_red[RED_STRIDE * b];
}

// Persistent loop of a vector kernel over its vertex blocks blockIdx.x +
// j gridDim.x (compacted to the active ones in a batch).
template <class IsActive, class Body>
__device__ inline void vertex_blocks(int nvblk, bool filter, int* list, IsActive is_active, Body body) {
  const int b0 = blockIdx.x, G = gridDim.x;
  const int n_cand = b0 < nvblk ? (nvblk - b0 + G - 1) / G : 0;
  for_active<NTH>(n_cand, filter, list, [&](int j) { return b0 + j * G; }, is_active, body);
}

template <class R>
struct VecArgs {
  const int32_t* vblk_off;
  const int32_t* vblk_inst;
  const int32_t* vtx_slot_off;
  const R* partial;
  const uint8_t* mask;
  const R* fext;
  double* x;  // iterate, always FP64 (the line-search step is taken in FP64)
  R* p;
  R* gc;
  R* gE;
  R* hd;
  double* vblk_dots;
  const double* blk_red;          // per element block: energy, quadratic form
  const int64_t* inst_blk_off;
  const int64_t* inst_vblk_off;
  StatePtrs st;
  int nvblk;   // vertex blocks of the mesh
  int filter;  // visit only blocks of active instances (batches)
};

template <class R>
__global__ __launch_bounds__(NTH) void fill_kernel(R* a, long long n, R v) {
  for (long long i = (long long)blockIdx.x * NTH + threadIdx.x; i < n; i += (long long)gridDim.x * NTH) a[i] = v;
}

template <class A, class B>
__global__ __launch_bounds__(NTH) void convert_kernel(const A* src, B* dst, long long n) {
  for (long long i = (long long)blockIdx.x * NTH + threadIdx.x; i < n; i += (long long)gridDim.x * NTH)
    dst[i] = (B)src[i];
}

__global__ void pncg_init_kernel(StatePtrs st, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  st.f[i] = 0.0;
  st.gd[i] = 0.0;
  st.alpha[i] = 0.0;
  st.alpha_commit[i] = 0.0;
  st.beta[i] = 0.0;
  st.gnorm0[i] = 0.0;
  st.gnorm[i] = 0.0;
  st.pmean[i] = 1.0;
  st.f_first[i] = 0.0;
  st.pHp[i] = 0.0;
  st.gbest[i] = INFINITY;
  st.f_mark[i] = INFINITY;
  st.best_step[i] = 0;
  st.f0[i] = 0.0;
  st.ls_last[i] = 0;
  st.status[i] = APL_RUNNING;
  st.act_eval[i] = 1;
  st.act_dir[i] = 0;
  st.commit[i] = 0;
  st.n_steps[i] = 0;
  st.n_evals[i] = 0;
  st.ls_step[i] = 0;
  st.n_backtracks[i] = 0;
  st.phase[i] = 0;
  st.restarts[i] = 0;
}

// R: vertex sums of the slot partials + FP64 partial dots of one vertex
// block (one DOF per thread and stride, coalesced), plus the block's share of
// the instance's element-block energies
template <class R>
__device__ inline void pncg_reduce_block(const VecArgs<R>& A, const int vb, double* sm, R* win) {
  const int inst = A.vblk_inst[vb];
  if (A.st.act_eval[inst] == 0) return;
  const int v0 = A.vblk_off[vb], v1 = A.vblk_off[vb + 1];
  const int i0 = 3 * v0, i1 = 3 * v1;
  const double alpha = A.st.alpha[inst];
  // per-DOF inputs, loaded before the partial window so the loads overlap
  int k0[DOF_PER_THREAD], k1[DOF_PER_THREAD];
  double xv[DOF_PER_THREAD];
  R pv[DOF_PER_THREAD], gcv[DOF_PER_THREAD], fv[DOF_PER_THREAD];
  bool fixed[DOF_PER_THREAD];
#pragma unroll
  for (int j = 0; j < DOF_PER_THREAD; ++j) {
    const int i = i0 + (int)threadIdx.x + j * NTH;
    k0[j] = k1[j] = 0;
    xv[j] = 0.0;
    pv[j] = gcv[j] = fv[j] = R(0);
    fixed[j] = true;
    if (i < i1) {
      const int v = i / 3, c = i - 3 * v;
      k0[j] = A.vtx_slot_off[v];
      k1[j] = A.vtx_slot_off[v + 1];
      xv[j] = A.x[i];
      pv[j] = A.p[i];
      gcv[j] = A.gc[i];
      fv[j] = A.fext != nullptr ? A.fext[i] : R(0);
      fixed[j] = A.mask != nullptr && ((A.mask[v] >> c) & 1u) != 0;
    }
  }
  VertexSums<R, 6, NTH> vs{A.vtx_slot_off, A.partial, win};
  vs.stage(v0, v1, VSUM_LDS_BYTES / (int)sizeof(R));
  double d[N_DOTS];
#pragma unroll
  for (int k = 0; k < N_DOTS; ++k) d[k] = 0.0;
#pragma unroll
  for (int j = 0; j < DOF_PER_THREAD; ++j) {
    const int i = i0 + (int)threadIdx.x + j * NTH;
    if (i >= i1) continue;
    R gh[2];
    vs.sum_dof(k0[j], k1[j], i - 3 * (i / 3), gh);
    A.gE[i] = gh[0];
    A.hd[i] = gh[1];
    const double pc = (double)pv[j];
    d[D_EXT] -= (double)fv[j] * (xv[j] + alpha * pc);
    if (fixed[j]) continue;
    const double gt = (double)(gh[0] - fv[j]);
    const double gc = (double)gcv[j];
    const R h = gh[1];
    d[D_G2] += gt * gt;
    if (h > R(0)) {
      const double P = (double)(R(1) / h);  // the preconditioner the direction update uses
      d[D_GPG] += gt * gt * P;
      d[D_GPGC] += gt * gc * P;
      d[D_GCPGC] += gc * gc * P;
      d[D_HSUM] += (double)h;
      d[D_HCNT] += 1.0;
    } else {
      d[D_GPG0] += gt * gt;
      d[D_GPGC0] += gt * gc;
      d[D_GCPGC0] += gc * gc;
    }
    d[D_GTP] += gt * pc;
    d[D_GCP] += gc * pc;
    d[D_PP] += pc * pc;
  }
  energy_share(A.blk_red, A.inst_blk_off, A.inst_vblk_off, inst, vb, d[D_E]);
  block_dots(d, sm, A.vblk_dots + (long long)vb * N_DOTS);
}

template <class R>
__global__ __launch_bounds__(NTH) void pncg_reduce_kernel(VecArgs<R> A) {
  __shared__ double sm[(NTH / 64) * N_DOTS];
  __shared__ R win[VSUM_LDS_BYTES / sizeof(R)];
  __shared__ int list[NTH];
  vertex_blocks(A.nvblk, A.filter, list, [&](int vb) { return A.st.act_eval[A.vblk_inst[vb]] != 0; },
                [&](int vb) { pncg_reduce_block<R>(A, vb, sm, win); });
}

// S: per-instance decision from the instance's dot totals (one thread)
__device__ inline void pncg_scalar_step(InstState& s, const Params& P, const double (&tot)[N_DOTS]) {
  const double fE = tot[D_E];
  const double ft = fE + tot[D_EXT];
  const double gn = sqrt(tot[D_G2]);
  const int phase = s.phase;
  s.n_evals += 1;
  bool accept;
  if (phase != 1) {
    accept = true;
  } else {
    const double f0 = s.f, a = s.alpha, gd = s.gd;
    const bool fin = isfinite(ft);
    const bool armijo = fin && ft <= f0 + P.c1 * a * gd;
    const bool wolfe = P.wolfe_delta > 0.0 && fin && ft <= f0 + P.wolfe_eps * fabs(f0) &&
                       tot[D_GTP] <= (2.0 * P.wolfe_delta - 1.0) * gd;
    accept = armijo || wolfe;
  }
  if (accept) {
    s.alpha_commit = phase == 1 ? s.alpha : 0.0;
    s.f0 = phase == 0 ? ft : s.f;  // the line search's start and end energies
    s.ls_last = s.ls_step;
    s.f = ft;
    s.gnorm = gn;
    s.commit = 1;
    if (phase == 0) {
      s.gnorm0 = gn;
      s.f_first = ft;
    }
    if (phase == 1) {
      s.n_steps += 1;
      s.restarts = 0;
    }
    // progress marker: the last step that lowered the best gradient norm by
    // 10% or the energy, since the previous marker, by more than the energy
    // noise the approximate-Wolfe test tolerates (10 wolfe_eps relative)
    const bool g_prog = gn < 0.9 * s.gbest;
    const double f_tol = fmax(10.0 * P.wolfe_eps, 1e-12) * fabs(ft);
    const bool f_prog = !(s.f_mark - ft <= f_tol);
    if (g_prog) s.gbest = gn;
    if (g_prog || f_prog) {
      s.best_step = s.n_steps;
      s.f_mark = ft;
    }
    int status = APL_RUNNING;
    if (!isfinite(gn) || !isfinite(ft))
      status = APL_NON_FINITE;
    else if (gn <= P.rtol * s.gnorm0 || gn <= P.atol)
      status = APL_PRIMARY_SUCCESS;
    else if (s.n_steps >= P.max_steps)
      status = APL_MAX_STEPS_REACHED;
    else if (P.stall_steps > 0 && s.n_steps - s.best_step >= P.stall_steps && s.gbest <= 1e3 * P.rtol * s.gnorm0)
      status = APL_STAGNATION;  // within 3 decades of the target: the rounding floor of the gradient
    s.status = status;
    if (status == APL_RUNNING) {
      const double pm = tot[D_HCNT] > 0.0 ? tot[D_HSUM] / tot[D_HCNT] : 1.0;
      s.pmean = pm;
      const double gPg = tot[D_GPG] + tot[D_GPG0] / pm;
      double beta = 0.0;
      if (phase == 1) {
        const double gPgc = tot[D_GPGC] + tot[D_GPGC0] / pm;
        const double gcPgc = tot[D_GCPGC] + tot[D_GCPGC0] / pm;
        const double yp = tot[D_GTP] - tot[D_GCP];
        if (fabs(yp) > 1e-300) {
          const double gPy = gPg - gPgc;
          const double yPy = gPg - 2.0 * gPgc + gcPgc;
          beta = gPy / yp - (yPy / yp) * (tot[D_GTP] / yp);
        } else {
          beta = INFINITY;
        }
        if (!isfinite(beta) || fabs(beta) > P.beta_max) beta = 0.0;
      }
      double gd = -gPg + beta * tot[D_GTP];
      if (!(isfinite(gd) && gd < 0.0)) {
        beta = 0.0;
        gd = -gPg;
      }
      s.beta = beta;
      s.gd = gd;
      s.act_dir = 1;
    } else {
      s.act_dir = 0;
      s.act_eval = 0;
    }
    s.ls_step = 0;
    s.phase = 1;
  } else {
    s.commit = 0;
    s.act_dir = 0;
    s.ls_step += 1;
    s.n_backtracks += 1;
    const double a = s.alpha * 0.5;
    if (s.ls_step > P.max_backtracking || !(a > 0.0)) {
      if (s.restarts >= 1) {
        s.status = APL_STAGNATION;
        s.act_eval = 0;
      } else {
        // restart from x with steepest descent: re-evaluate at alpha = 0
        s.restarts += 1;
        s.phase = 2;
        s.alpha = 0.0;
        s.ls_step = 0;
      }
    } else {
      s.alpha = a;
    }
  }
}

// S: per-instance decision, one workgroup per instance
__global__ __launch_bounds__(SCAL_TH) void pncg_scalar_kernel(StatePtrs st, Params P, const double* vblk_dots,
                                                          const int64_t* inst_vblk_off) {
  __shared__ double sm[(SCAL_TH / 64 + 1) * N_DOTS];
  const int i = blockIdx.x;
  // the active flag, the table bounds and the state are independent: all are
  // loaded before the first wait (an inactive instance sums an empty table)
  const bool act = st.act_eval[i] != 0;
  const long long vb0 = inst_vblk_off[i], vb1 = inst_vblk_off[i + 1];
  InstState s;
  if (threadIdx.x == 0) s = load_state(st, i);
  double tot[N_DOTS];
  flat_column_sums<N_DOTS, SCAL_TH>(vblk_dots + vb0 * N_DOTS, act ? (vb1 - vb0) * N_DOTS : 0, sm, tot);
  if (threadIdx.x != 0 || !act) return;
  pncg_scalar_step(s, P, tot);
  store_state(st, i, s);
}

// D: commit and new direction
template <class R>
__device__ inline void pncg_commit_block(const VecArgs<R>& A, const int vb) {
  const int inst = A.vblk_inst[vb];
  if (A.st.commit[inst] == 0) return;
  const double a = A.st.alpha_commit[inst];
  const R beta = (R)A.st.beta[inst];
  const bool dir = A.st.act_dir[inst] != 0;
  const R pm = (R)A.st.pmean[inst];
  const int i0 = 3 * A.vblk_off[vb], i1 = 3 * A.vblk_off[vb + 1];
#pragma unroll
  for (int j = 0; j < DOF_PER_THREAD; ++j) {
    const int i = i0 + (int)threadIdx.x + j * NTH;
    if (i >= i1) continue;
    const int v = i / 3, c = i - 3 * v;
    const bool fixed = A.mask != nullptr && ((A.mask[v] >> c) & 1u) != 0;
    const R pc = A.p[i];
    A.x[i] += a * (double)pc;
    const R gt = fixed ? R(0) : A.gE[i] - (A.fext != nullptr ? A.fext[i] : R(0));
    A.gc[i] = gt;
    if (dir) {
      const R h = A.hd[i];
      const R Pc = fixed ? R(0) : (h > R(0) ? R(1) / h : R(1) / pm);
      A.p[i] = -Pc * gt + beta * pc;
    }
  }
}

template <class R>
__global__ __launch_bounds__(NTH) void pncg_commit_kernel(VecArgs<R> A) {
  __shared__ int list[NTH];
  vertex_blocks(A.nvblk, A.filter, list, [&](int vb) { return A.st.commit[A.vblk_inst[vb]] != 0; },
                [&](int vb) { pncg_commit_block<R>(A, vb); });
}

// S2: alpha from the quadratic form; clears the per-iteration flags
__global__ __launch_bounds__(SCAL_TH) void pncg_alpha_kernel(StatePtrs st, Params P, const double* blk_red,
                                                         const int64_t* inst_blk_off) {
  __shared__ double red[2 * (SCAL_TH / 64 + 1)];
  const int i = blockIdx.x;
  // all loads issued before the first wait, as in pncg_scalar_kernel
  const bool dir = st.act_dir[i] != 0;
  const long long b0 = inst_blk_off[i], b1 = inst_blk_off[i + 1];
  const double gd = threadIdx.x == 0 ? st.gd[i] : 0.0;
  double eq[2];
  flat_column_sums<2, SCAL_TH>(blk_red + RED_STRIDE * b0, dir ? RED_STRIDE * (b1 - b0) : 0, red, eq);
  if (dir) {
    if (threadIdx.x == 0) {
      const double pHp = eq[1];
      st.pHp[i] = pHp;
      double a = -gd / pHp;
      if (isnan(a)) a = 0.0;
      else if (a == -INFINITY) a = 0.0;
      else if (a == INFINITY) a = 1.0;
      if (!(a > 0.0 && isfinite(a))) a = 1.0;
      st.alpha[i] = a * P.overstep;
    }
  }
  if (threadIdx.x == 0) {
    st.commit[i] = 0;
    st.act_dir[i] = 0;
  }
}


// ------------------------------------------------------------------ PCG (adjoint)
// State reuse: gd = r.z, gnorm0 = ||b||, gnorm = ||r||, alpha = a, beta,
// pHp = d.Hd, pmean = mean positive diagonal, f = ||r|| / ||b||.

template <class R>
struct CgArgs {
  const int32_t* vblk_off;
  const int32_t* vblk_inst;
  const int32_t* vtx_slot_off;
  const R* partial;
  const uint8_t* mask;
  const R* rhs;
  R* y;
  R* r;
  R* z;
  R* d;
  R* q;
  R* hd;
  R* t;     // normal equations: H d (or the masked right-hand side while it is formed)
  R* tout;  // CG_T / CG_MASKCOPY output
  double* vblk_dots;
  StatePtrs st;
  int nvblk;   // vertex blocks of the mesh
  int filter;  // visit only blocks of active instances (batches)
  int normal;  // CG on the normal equations H^2 y = H b, Jacobi with diag(H)^2
};

// CG_T: tout = H v summed from the partials, fixed DOFs 0 (no dots);
// CG_MASKCOPY: tout = rhs with fixed DOFs 0
enum CgStage : int { CG_DIAG = 0, CG_PRECOND = 1, CG_Q = 2, CG_UPDATE = 3, CG_DIR = 4, CG_T = 5, CG_MASKCOPY = 6 };

template <class R, int STAGE>
__device__ inline void cg_vec_block(const CgArgs<R>& A, const int vb, double* sm, R* win) {
  constexpr bool SUMS = STAGE == CG_DIAG || STAGE == CG_Q || STAGE == CG_T;
  constexpr bool PRE = STAGE == CG_DIAG || STAGE == CG_PRECOND || STAGE == CG_MASKCOPY;  // before the status
  const int inst = A.vblk_inst[vb];
  if (!PRE && A.st.act_eval[inst] == 0) return;
  const int v0 = A.vblk_off[vb], v1 = A.vblk_off[vb + 1];
  const int i0 = 3 * v0, i1 = 3 * v1;
  double dd[N_DOTS];
#pragma unroll
  for (int k = 0; k < N_DOTS; ++k) dd[k] = 0.0;
  int k0[DOF_PER_THREAD], k1[DOF_PER_THREAD];
  if constexpr (SUMS) {
#pragma unroll
    for (int j = 0; j < DOF_PER_THREAD; ++j) {
      const int i = i0 + (int)threadIdx.x + j * NTH;
      k0[j] = k1[j] = 0;
      if (i < i1) {
        k0[j] = A.vtx_slot_off[i / 3];
        k1[j] = A.vtx_slot_off[i / 3 + 1];
      }
    }
  }
  VertexSums<R, 3, NTH> vs{A.vtx_slot_off, A.partial, win};
  if constexpr (SUMS) vs.stage(v0, v1, VSUM_LDS_BYTES / (int)sizeof(R));
#pragma unroll
  for (int j = 0; j < DOF_PER_THREAD; ++j) {
    const int i = i0 + (int)threadIdx.x + j * NTH;
    if (i >= i1) continue;
    const int v = i / 3, c = i - 3 * v;
    const bool fixed = A.mask != nullptr && ((A.mask[v] >> c) & 1u) != 0;
    if constexpr (SUMS) {
      R s1[1];
      vs.sum_dof(k0[j], k1[j], c, s1);
      if constexpr (STAGE == CG_DIAG) {
        A.hd[i] = s1[0];
        const R b = fixed ? R(0) : A.rhs[i];
        A.r[i] = b;
        A.y[i] = R(0);
        A.d[i] = R(0);
        A.q[i] = R(0);
        if (!fixed) {
          dd[D_G2] += (double)b * (double)b;
          if (s1[0] > R(0)) {
            dd[D_HSUM] += (double)s1[0];
            dd[D_HCNT] += 1.0;
          }
        }
      } else if constexpr (STAGE == CG_T) {
        A.tout[i] = fixed ? R(0) : s1[0];
      } else {
        const R qv = fixed ? R(0) : s1[0];
        A.q[i] = qv;
        dd[D_GTP] += (double)A.d[i] * (double)qv;
      }
    } else if constexpr (STAGE == CG_MASKCOPY) {
      A.tout[i] = fixed ? R(0) : A.rhs[i];
    } else if constexpr (STAGE == CG_PRECOND || STAGE == CG_UPDATE) {
      const R a = (R)A.st.alpha[inst];
      const R pm = (R)A.st.pmean[inst];
      R rv = A.r[i];
      if constexpr (STAGE == CG_UPDATE) {
        A.y[i] += a * A.d[i];
        rv -= a * A.q[i];
        A.r[i] = rv;
      }
      const R h = A.hd[i];
      // Jacobi: 1/diag(H), or 1/diag(H)^2 for the normal equations (diag(H^2) >= diag(H)^2)
      const R hh = h > R(0) ? h : pm;
      const R zv = fixed ? R(0) : (A.normal ? rv / (hh * hh) : rv / hh);
      A.z[i] = zv;
      if constexpr (STAGE == CG_PRECOND) A.d[i] = zv;
      dd[D_GPG] += (double)rv * (double)zv;
      dd[D_G2] += (double)rv * (double)rv;
    } else {  // CG_DIR
      A.d[i] = A.z[i] + (R)A.st.beta[inst] * A.d[i];
    }
  }
  if constexpr (STAGE != CG_DIR && STAGE != CG_T && STAGE != CG_MASKCOPY)
    block_dots(dd, sm, A.vblk_dots + (long long)vb * N_DOTS);
}

template <class R, int STAGE>
__global__ __launch_bounds__(NTH) void cg_vec_kernel(CgArgs<R> A) {
  __shared__ double sm[(NTH / 64) * N_DOTS];
  constexpr bool SUMS = STAGE == CG_DIAG || STAGE == CG_Q || STAGE == CG_T;
  __shared__ R win[SUMS ? VSUM_LDS_BYTES / sizeof(R) : 1];
  __shared__ int list[NTH];
  constexpr bool ALL = STAGE == CG_DIAG || STAGE == CG_PRECOND || STAGE == CG_MASKCOPY;  // before the status
  vertex_blocks(A.nvblk, A.filter && !ALL, list, [&](int vb) { return A.st.act_eval[A.vblk_inst[vb]] != 0; },
                [&](int vb) { cg_vec_block<R, STAGE>(A, vb, sm, win); });
}

// per-instance CG scalars (one thread); stage 0: after DIAG, 1: after
// PRECOND, 2: after Q, 3: after UPDATE
__device__ inline void cg_scalar_step(InstState& s, int stage, double rtol, int max_iters,
                                      const double (&tot)[N_DOTS]) {
  if (stage == 0) {
    s.pmean = tot[D_HCNT] > 0.0 ? tot[D_HSUM] / tot[D_HCNT] : 1.0;
    const double bn = sqrt(tot[D_G2]);
    s.gnorm0 = bn;
    s.gnorm = bn;
    s.n_steps = 0;
    s.f = bn > 0.0 ? 1.0 : 0.0;
    s.status = bn > 0.0 ? APL_RUNNING : APL_PRIMARY_SUCCESS;
    s.act_eval = bn > 0.0 ? 1 : 0;
    s.alpha = 0.0;
    s.beta = 0.0;
  } else if (stage == 1) {
    s.gd = tot[D_GPG];
  } else if (stage == 2) {
    const double dq = tot[D_GTP];
    s.pHp = dq;
    if (!(dq > 0.0) || !isfinite(dq)) {
      s.status = APL_NON_FINITE;
      s.act_eval = 0;
      s.alpha = 0.0;
    } else {
      s.alpha = s.gd / dq;
    }
  } else {
    const double rz = tot[D_GPG];
    const double rn = sqrt(tot[D_G2]);
    s.gnorm = rn;
    s.n_steps += 1;
    s.f = s.gnorm0 > 0.0 ? rn / s.gnorm0 : 0.0;
    if (rn <= rtol * s.gnorm0) {
      s.status = APL_PRIMARY_SUCCESS;
      s.act_eval = 0;
    } else if (!isfinite(rn)) {
      s.status = APL_NON_FINITE;
      s.act_eval = 0;
    } else if (s.n_steps >= max_iters) {
      s.status = APL_MAX_STEPS_REACHED;
      s.act_eval = 0;
    } else {
      s.beta = rz / s.gd;
      s.gd = rz;
    }
  }
}

// the CG's scalar stages, one workgroup per instance
__global__ __launch_bounds__(SCAL_TH) void cg_scalar_kernel(StatePtrs st, int stage, double rtol, int max_iters,
                                                        const double* vblk_dots, const int64_t* inst_vblk_off) {
  __shared__ double sm[(SCAL_TH / 64 + 1) * N_DOTS];
  const int i = blockIdx.x;
  const bool act = stage < 2 || st.act_eval[i] != 0;  // loaded with the rest, as in pncg_scalar_kernel
  const long long vb0 = inst_vblk_off[i], vb1 = inst_vblk_off[i + 1];
  InstState s;
  if (threadIdx.x == 0) s = load_state(st, i);
  double tot[N_DOTS];
  flat_column_sums<N_DOTS, SCAL_TH>(vblk_dots + vb0 * N_DOTS, act ? (vb1 - vb0) * N_DOTS : 0, sm, tot);
  if (threadIdx.x != 0 || !act) return;
  cg_scalar_step(s, stage, rtol, max_iters, tot);
  store_state(st, i, s);
}

}  // namespace apl

using namespace apl;

struct apl_solver {
  apl_mesh* mesh = nullptr;  // holds one reference of the mesh
  int device = 0;
  int material = 0;
  void* vecs = nullptr;  // p, gc, gE, hd, r, z, d, q, hd (9 x V x 3)
  void* x64 = nullptr;   // FP64 iterate of FP32 meshes (V x 3)
  double* sd = nullptr;
  int32_t* si = nullptr;
  StatePtrs st{};
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<char> gkey;
  hipGraphExec_t cg_gexec = nullptr;  // K PCG iterations
  hipGraph_t cg_graph = nullptr;
  std::vector<char> cg_gkey;
  int32_t* host_status = nullptr;
  int32_t* host_steps = nullptr;
  struct apl_solver_run* run = nullptr;  // the step-wise solve in progress
};

namespace apl {

template <class R>
static VecArgs<R> vec_args(apl_solver* s, double* x, const void* fext, const uint8_t* mask) {
  apl_mesh* m = s->mesh;
  const size_t nv = 3 * (size_t)m->L.V;
  R* base = (R*)s->vecs;
  VecArgs<R> A;
  A.vblk_off = m->d.vblk_off;
  A.vblk_inst = m->d.vblk_inst;
  A.vtx_slot_off = m->d.vtx_slot_off;
  A.partial = (const R*)m->d.partial;
  A.mask = mask;
  A.fext = (const R*)fext;
  A.x = x;
  A.p = base;
  A.gc = base + nv;
  A.gE = base + 2 * nv;
  A.hd = base + 3 * nv;
  A.vblk_dots = m->d.vblk_dots;
  A.blk_red = m->d.blk_red;
  A.inst_blk_off = m->d.inst_blk_off;
  A.inst_vblk_off = m->d.inst_vblk_off;
  A.st = s->st;
  A.nvblk = (int)m->L.vblk_inst.size();
  A.filter = m->L.n_inst > 1 ? 1 : 0;
  return A;
}

static Params to_params(const apl_pncg_params* p) {
  Params P;
  std::memset(&P, 0, sizeof P);  // the bytes (padding included) key the cached graph
  P.rtol = p->rtol;
  P.atol = p->atol;
  P.overstep = p->overstep;
  P.c1 = p->c1;
  P.wolfe_delta = p->wolfe_delta;
  P.wolfe_eps = p->wolfe_eps;
  P.beta_max = p->beta_max;
  P.max_steps = p->max_steps;
  P.max_backtracking = p->max_backtracking;
  P.stall_steps = p->stall_steps;
  return P;
}

// Persistent grid of the vector kernels: every vertex block its own
// workgroup up to sixteen per CU, beyond that a grid-stride loop.
static unsigned vec_grid(const apl_mesh* m) {
  const long long nvb = (long long)m->L.vblk_inst.size();
  const long long cap = 16LL * m->n_cu;
  return (unsigned)(nvb < cap ? nvb : cap);
}

template <class R>
static int pncg_iteration(apl_solver* s, const void* mu, const void* lam, double* x, const void* fext,
                          const uint8_t* mask, const Params& P, int grid, hipStream_t st) {
  const int x_f64 = sizeof(R) == 4 ? 1 : 0;
  apl_mesh* m = s->mesh;
  VecArgs<R> A = vec_args<R>(s, x, fext, mask);
  const unsigned nvb = (unsigned)m->L.vblk_inst.size();
  const unsigned ni = (unsigned)m->L.n_inst;
  if (launch_elem(m, s->material, W_E | W_G | W_H, mu, lam, x, A.p, s->st.alpha, s->st.act_eval, A.gE, A.hd, 0,
                  grid, st, x_f64))
    return -1;
  const unsigned vg = vec_grid(m);
  if (nvb) hipLaunchKernelGGL((pncg_reduce_kernel<R>), dim3(vg), dim3(NTH), 0, st, A);
  hipLaunchKernelGGL(pncg_scalar_kernel, dim3(ni), dim3(SCAL_TH), 0, st, s->st, P, m->d.vblk_dots, m->d.inst_vblk_off);
  if (nvb) hipLaunchKernelGGL((pncg_commit_kernel<R>), dim3(vg), dim3(NTH), 0, st, A);
  if (launch_elem(m, s->material, W_HQ, mu, lam, x, A.p, nullptr, s->st.act_dir, nullptr, nullptr, 0, grid, st,
                  x_f64))
    return -1;
  hipLaunchKernelGGL(pncg_alpha_kernel, dim3(ni), dim3(SCAL_TH), 0, st, s->st, P, m->d.blk_red, m->d.inst_blk_off);
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

template <class R>
static int fill(R* a, long long n, R v, hipStream_t st) {
  if (n <= 0) return 0;
  long long g = (n + NTH - 1) / NTH;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL((fill_kernel<R>), dim3((unsigned)g), dim3(NTH), 0, st, a, n, v);
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

// A solve in progress: what apl_pncg_begin bound, for the following
// advance / end calls.
struct PncgRun {
  bool open = false;
  const void* mu = nullptr;
  const void* lam = nullptr;
  void* x = nullptr;  // the caller's iterate (mesh dtype)
  const void* fext = nullptr;
  const uint8_t* mask = nullptr;
  double* xs = nullptr;  // the FP64 iterate (the caller's x for FP64 meshes)
  Params P{};
  int K = 16, grid = 0;
  bool use_graph = true;
  long long done = 0, guard = 0;
};

template <class R>
static int pncg_begin_t(apl_solver* s, const void* mu, const void* lam, void* x, const void* fext,
                        const uint8_t* mask, const apl_pncg_params* prm, hipStream_t st);
template <class R>
static int pncg_advance_t(apl_solver* s, long long max_iters, int steps, hipStream_t st);
template <class R>
static int pncg_end_t(apl_solver* s, apl_pncg_stats* stats, hipStream_t st);

}  // namespace apl

struct apl_solver_run {
  apl::PncgRun r;
};

namespace apl {

template <class R>
static int pncg_begin_t(apl_solver* s, const void* mu, const void* lam, void* x, const void* fext,
                        const uint8_t* mask, const apl_pncg_params* prm, hipStream_t st) {
  apl_mesh* m = s->mesh;
  PncgRun& run = s->run->r;
  run.open = false;
  const int n = (int)m->L.n_inst;
  const Params P = to_params(prm);
  const int K = prm->iters_per_launch > 0 ? prm->iters_per_launch : 16;
  // reset state and work vectors (p and g_cur must start at zero)
  hipLaunchKernelGGL(pncg_init_kernel, dim3((n + 255) / 256), dim3(256), 0, st, s->st, n);
  if (fill<R>((R*)s->vecs, 4 * 3 * (long long)m->L.V, R(0), st)) return -1;
  APL_CHECK_HIP(hipGetLastError());
  // the iterate lives in FP64; FP32 meshes keep a private FP64 copy of x
  const long long nx = 3 * (long long)m->L.V;
  const int x_f64 = sizeof(R) == 4 ? 1 : 0;
  double* xs = x_f64 ? (double*)s->x64 : (double*)x;
  if (x_f64 && nx > 0) {
    hipLaunchKernelGGL((convert_kernel<R, double>), dim3((unsigned)std::min<long long>((nx + NTH - 1) / NTH, 4096)),
                       dim3(NTH), 0, st, (const R*)x, xs, nx);
    APL_CHECK_HIP(hipGetLastError());
  }
  const int grid = prm->grid_override;
  // occupancy-derived grids are queried (and cached) outside any capture
  (void)elem_grid(m, s->material, W_E | W_G | W_H, x_f64, true);
  (void)elem_grid(m, s->material, W_HQ, x_f64, true);
  // cached graph of K iterations for these pointers and parameters
  bool use_graph = prm->use_graph != 0;
  if (use_graph) {
    std::vector<char> key(sizeof(void*) * 6 + sizeof(Params) + 3 * sizeof(int));
    char* k = key.data();
    const void* ptrs[6] = {mu, lam, x, fext, mask, m->activation};
    std::memcpy(k, ptrs, sizeof ptrs);
    k += sizeof ptrs;
    std::memcpy(k, &P, sizeof P);
    k += sizeof P;
    std::memcpy(k, &K, sizeof K);
    k += sizeof K;
    std::memcpy(k, &grid, sizeof grid);
    k += sizeof grid;
    std::memcpy(k, &m->grid_pct, sizeof m->grid_pct);
    if (s->gexec == nullptr || key != s->gkey) {
      if (s->gexec) {
        drop_hip_error(hipGraphExecDestroy(s->gexec));
        s->gexec = nullptr;
      }
      if (s->graph) {
        drop_hip_error(hipGraphDestroy(s->graph));
        s->graph = nullptr;
      }
      if (s->cap == nullptr) APL_CHECK_HIP(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
      APL_CHECK_HIP(hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal));
      int rc = 0;
      for (int it = 0; it < K && rc == 0; ++it) rc = pncg_iteration<R>(s, mu, lam, xs, fext, mask, P, grid, s->cap);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s->cap, &g);
      if (rc != 0 || e != hipSuccess) {
        if (g) drop_hip_error(hipGraphDestroy(g));
        if (rc == 0) set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
        return -1;
      }
      s->graph = g;
      APL_CHECK_HIP(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
      s->gkey = key;
    }
  }
  run.mu = mu;
  run.lam = lam;
  run.x = x;
  run.fext = fext;
  run.mask = mask;
  run.xs = xs;
  run.P = P;
  run.K = K;
  run.grid = grid;
  run.use_graph = use_graph;
  run.done = 0;
  run.guard = ((long long)prm->max_steps + 2) * ((long long)prm->max_backtracking + 2) * 2 + 8;
  run.open = true;
  return 0;
}

// Run device iterations (trial evaluations) until no instance is running,
// the guard is reached, max_iters more have run (max_iters > 0) or, when
// steps > 0, every running instance has taken that many more accepted steps
// (one trial per launch, with a status read after each).
template <class R>
static int pncg_advance_t(apl_solver* s, long long max_iters, int steps, hipStream_t st) {
  apl_mesh* m = s->mesh;
  PncgRun& run = s->run->r;
  const int n = (int)m->L.n_inst;
  std::vector<int32_t> target;
  if (steps > 0) {
    APL_CHECK_HIP(hipMemcpyAsync(s->host_steps, s->st.n_steps, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipStreamSynchronize(st));
    target.assign(s->host_steps, s->host_steps + n);
    for (int32_t& t : target) t += steps;
  }
  long long ran = 0;
  while (true) {
    // one trial per launch when counting steps; eager launches for a
    // remainder shorter than the captured graph
    const long long left = max_iters > 0 ? max_iters - ran : run.K;
    const bool graph = run.use_graph && steps <= 0 && left >= run.K;
    const int K = steps > 0 ? 1 : (graph ? run.K : (int)std::min<long long>(run.K, left));
    if (graph) {
      APL_CHECK_HIP(hipGraphLaunch(s->gexec, st));
    } else {
      for (int it = 0; it < K; ++it)
        if (pncg_iteration<R>(s, run.mu, run.lam, run.xs, run.fext, run.mask, run.P, run.grid, st)) return -1;
    }
    run.done += K;
    ran += K;
    APL_CHECK_HIP(hipMemcpyAsync(s->host_status, s->st.status, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (steps > 0)
      APL_CHECK_HIP(hipMemcpyAsync(s->host_steps, s->st.n_steps, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipStreamSynchronize(st));
    bool running = false, pending = false;
    for (int i = 0; i < n; ++i) {
      const bool r = s->host_status[i] == APL_RUNNING;
      running = running || r;
      if (steps > 0 && r && s->host_steps[i] < target[i]) pending = true;
    }
    if (!running || run.done > run.guard || (max_iters > 0 && ran >= max_iters) || (steps > 0 && !pending)) break;
  }
  return 0;
}

// Copy the FP64 iterate into the caller's x (FP32 meshes) and read the stats.
template <class R>
static int pncg_end_t(apl_solver* s, apl_pncg_stats* stats, hipStream_t st) {
  apl_mesh* m = s->mesh;
  PncgRun& run = s->run->r;
  const int n = (int)m->L.n_inst;
  const long long nx = 3 * (long long)m->L.V;
  if (sizeof(R) == 4 && nx > 0) {
    hipLaunchKernelGGL((convert_kernel<double, R>), dim3((unsigned)std::min<long long>((nx + NTH - 1) / NTH, 4096)),
                       dim3(NTH), 0, st, (const double*)run.xs, (R*)run.x, nx);
    APL_CHECK_HIP(hipGetLastError());
  }
  if (stats != nullptr) {
    std::vector<double> hd((size_t)N_SD * n);
    std::vector<int32_t> hi((size_t)N_SI * n);
    APL_CHECK_HIP(hipMemcpyAsync(hd.data(), s->sd, hd.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipMemcpyAsync(hi.data(), s->si, hi.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) {
      apl_pncg_stats& o = stats[i];
      o.f = hd[0 * n + i];
      o.alpha = hd[3 * n + i];
      o.beta = hd[4 * n + i];
      o.grad_norm_first = hd[5 * n + i];
      o.grad_norm = hd[6 * n + i];
      o.f_first = hd[8 * n + i];
      o.status = hi[0 * n + i];
      o.n_steps = hi[4 * n + i];
      o.n_evals = hi[5 * n + i];
      o.n_backtracks = hi[7 * n + i];
      if (o.status == APL_RUNNING) o.status = APL_MAX_STEPS_REACHED;
    }
  }
  return 0;
}

template <class R>
static int pncg_solve_t(apl_solver* s, const void* mu, const void* lam, void* x, const void* fext,
                        const uint8_t* mask, const apl_pncg_params* prm, apl_pncg_stats* stats, hipStream_t st) {
  if (pncg_begin_t<R>(s, mu, lam, x, fext, mask, prm, st)) return -1;
  if (pncg_advance_t<R>(s, 0, 0, st)) return -1;
  const int rc = pncg_end_t<R>(s, stats, st);
  s->run->r.open = false;
  return rc;
}

template <class R>
static CgArgs<R> cg_args(apl_solver* s, const void* rhs, void* y, const uint8_t* mask) {
  apl_mesh* m = s->mesh;
  const size_t nv = 3 * (size_t)m->L.V;
  R* base = (R*)s->vecs;
  CgArgs<R> A;
  A.vblk_off = m->d.vblk_off;
  A.vblk_inst = m->d.vblk_inst;
  A.vtx_slot_off = m->d.vtx_slot_off;
  A.partial = (const R*)m->d.partial;
  A.mask = mask;
  A.rhs = (const R*)rhs;
  A.y = (R*)y;
  A.r = base + 4 * nv;
  A.z = base + 5 * nv;
  A.d = base + 6 * nv;
  A.q = base + 7 * nv;
  A.hd = base + 8 * nv;
  A.t = base;  // the PNCG's direction buffer is free during a CG solve
  A.tout = base;
  A.vblk_dots = m->d.vblk_dots;
  A.st = s->st;
  A.nvblk = (int)s->mesh->L.vblk_inst.size();
  A.filter = s->mesh->L.n_inst > 1 ? 1 : 0;
  A.normal = 0;
  return A;
}

template <class R>
static int cg_iteration(apl_solver* s, const void* mu, const void* lam, const void* x, CgArgs<R>& A, double rtol,
                        int max_iters, hipStream_t st) {
  apl_mesh* m = s->mesh;
  const unsigned nvb = (unsigned)m->L.vblk_inst.size();
  const unsigned ni = (unsigned)m->L.n_inst;
  if (A.normal) {
    // q = H (H d): t = H d with fixed DOFs zeroed, then the pass on t
    if (launch_elem(m, s->material, W_HP, mu, lam, x, A.d, nullptr, s->st.act_eval, nullptr, nullptr, 0, 0, st))
      return -1;
    CgArgs<R> At = A;
    At.tout = A.t;
    if (nvb) hipLaunchKernelGGL((cg_vec_kernel<R, CG_T>), dim3(vec_grid(m)), dim3(NTH), 0, st, At);
    if (launch_elem(m, s->material, W_HP, mu, lam, x, A.t, nullptr, s->st.act_eval, nullptr, nullptr, 0, 0, st))
      return -1;
  } else {
    if (launch_elem(m, s->material, W_HP, mu, lam, x, A.d, nullptr, s->st.act_eval, nullptr, nullptr, 0, 0, st))
      return -1;
  }
  if (nvb) hipLaunchKernelGGL((cg_vec_kernel<R, CG_Q>), dim3(vec_grid(m)), dim3(NTH), 0, st, A);
  hipLaunchKernelGGL(cg_scalar_kernel, dim3(ni), dim3(SCAL_TH), 0, st, s->st, 2, rtol, max_iters, m->d.vblk_dots,
                     m->d.inst_vblk_off);
  if (nvb) hipLaunchKernelGGL((cg_vec_kernel<R, CG_UPDATE>), dim3(vec_grid(m)), dim3(NTH), 0, st, A);
  hipLaunchKernelGGL(cg_scalar_kernel, dim3(ni), dim3(SCAL_TH), 0, st, s->st, 3, rtol, max_iters, m->d.vblk_dots,
                     m->d.inst_vblk_off);
  if (nvb) hipLaunchKernelGGL((cg_vec_kernel<R, CG_DIR>), dim3(vec_grid(m)), dim3(NTH), 0, st, A);
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

template <class R>
static int cg_solve_t(apl_solver* s, const void* mu, const void* lam, const void* x, const void* rhs, void* y,
                      const uint8_t* mask, double rtol, int max_iters, apl_pncg_stats* stats, hipStream_t st,
                      int normal) {
  apl_mesh* m = s->mesh;
  const int n = (int)m->L.n_inst;
  const unsigned nvb = (unsigned)m->L.vblk_inst.size();
  CgArgs<R> A = cg_args<R>(s, rhs, y, mask);
  hipLaunchKernelGGL(pncg_init_kernel, dim3((n + 255) / 256), dim3(256), 0, st, s->st, n);
  if (normal) {
    // the normal equations' right-hand side H b (b with fixed DOFs zeroed),
    // kept in the PNCG's g_cur buffer; the CG then runs on it
    A.normal = 1;
    R* b2 = (R*)s->vecs + 3 * (size_t)m->L.V;
    CgArgs<R> Ab = A;
    Ab.tout = A.t;
    if (nvb) hipLaunchKernelGGL((cg_vec_kernel<R, CG_MASKCOPY>), dim3(vec_grid(m)), dim3(NTH), 0, st, Ab);
    if (launch_elem(m, s->material, W_HP, mu, lam, x, A.t, nullptr, nullptr, nullptr, nullptr, 0, 0, st)) return -1;
    Ab.tout = b2;
    if (nvb) hipLaunchKernelGGL((cg_vec_kernel<R, CG_T>), dim3(vec_grid(m)), dim3(NTH), 0, st, Ab);
    A.rhs = b2;
  }
  // zero the diagonal buffer (vertices with no tet keep 0) and run the Hdiag pass at x
  if (fill<R>(A.hd, 3 * (long long)m->L.V, R(0), st)) return -1;
  if (launch_elem(m, s->material, W_H, mu, lam, x, nullptr, nullptr, nullptr, nullptr, A.hd, 0, 0, st)) return -1;
  if (nvb) hipLaunchKernelGGL((cg_vec_kernel<R, CG_DIAG>), dim3(vec_grid(m)), dim3(NTH), 0, st, A);
  hipLaunchKernelGGL(cg_scalar_kernel, dim3(n), dim3(SCAL_TH), 0, st, s->st, 0, rtol, max_iters, m->d.vblk_dots,
                     m->d.inst_vblk_off);
  if (nvb) hipLaunchKernelGGL((cg_vec_kernel<R, CG_PRECOND>), dim3(vec_grid(m)), dim3(NTH), 0, st, A);
  hipLaunchKernelGGL(cg_scalar_kernel, dim3(n), dim3(SCAL_TH), 0, st, s->st, 1, rtol, max_iters, m->d.vblk_dots,
                     m->d.inst_vblk_off);
  APL_CHECK_HIP(hipGetLastError());
  (void)elem_grid(m, s->material, W_HP, 0, true);
  const int K = 16;
  // K iterations captured once per set of pointers and parameters
  {
    std::vector<char> key(sizeof(void*) * 7 + sizeof(double) + 3 * sizeof(int));
    char* k = key.data();
    const void* ptrs[7] = {mu, lam, x, rhs, y, mask, m->activation};
    std::memcpy(k, ptrs, sizeof ptrs);
    k += sizeof ptrs;
    std::memcpy(k, &rtol, sizeof rtol);
    k += sizeof rtol;
    std::memcpy(k, &max_iters, sizeof max_iters);
    k += sizeof max_iters;
    std::memcpy(k, &normal, sizeof normal);
    k += sizeof normal;
    std::memcpy(k, &m->grid_pct, sizeof m->grid_pct);
    if (s->cg_gexec == nullptr || key != s->cg_gkey) {
      if (s->cg_gexec) {
        drop_hip_error(hipGraphExecDestroy(s->cg_gexec));
        s->cg_gexec = nullptr;
      }
      if (s->cg_graph) {
        drop_hip_error(hipGraphDestroy(s->cg_graph));
        s->cg_graph = nullptr;
      }
      if (s->cap == nullptr) APL_CHECK_HIP(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
      APL_CHECK_HIP(hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal));
      int rc = 0;
      for (int it = 0; it < K && rc == 0; ++it) rc = cg_iteration<R>(s, mu, lam, x, A, rtol, max_iters, s->cap);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s->cap, &g);
      if (rc != 0 || e != hipSuccess) {
        if (g) drop_hip_error(hipGraphDestroy(g));
        if (rc == 0) set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
        return -1;
      }
      s->cg_graph = g;
      APL_CHECK_HIP(hipGraphInstantiate(&s->cg_gexec, g, nullptr, nullptr, 0));
      s->cg_gkey = key;
    }
  }
  long long done = 0;
  while (true) {
    APL_CHECK_HIP(hipGraphLaunch(s->cg_gexec, st));
    done += K;
    APL_CHECK_HIP(hipMemcpyAsync(s->host_status, s->st.status, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipStreamSynchronize(st));
    bool running = false;
    for (int i = 0; i < n; ++i) running = running || s->host_status[i] == APL_RUNNING;
    if (!running || done > (long long)max_iters + K) break;
  }
  if (stats != nullptr) {
    std::vector<double> hd((size_t)N_SD * n);
    std::vector<int32_t> hi((size_t)N_SI * n);
    APL_CHECK_HIP(hipMemcpyAsync(hd.data(), s->sd, hd.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipMemcpyAsync(hi.data(), s->si, hi.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) {
      apl_pncg_stats& o = stats[i];
      o.f = hd[0 * n + i];
      o.alpha = hd[2 * n + i];
      o.beta = hd[4 * n + i];
      o.grad_norm_first = hd[5 * n + i];
      o.grad_norm = hd[6 * n + i];
      o.f_first = 1.0;
      o.status = hi[0 * n + i] == APL_RUNNING ? APL_MAX_STEPS_REACHED : hi[0 * n + i];
      o.n_steps = hi[4 * n + i];
      o.n_evals = hi[4 * n + i];
      o.n_backtracks = 0;
    }
  }
  return 0;
}

}  // namespace apl

extern "C" {

int apl_solver_create(apl_mesh* m, int material, apl_solver** out) {
  if (m == nullptr || out == nullptr) {
    set_error("NULL argument");
    return -1;
  }
  if (!material_ok(m, material)) {
    set_error(m->model ? "a model mesh takes material APL_MODEL"
                       : "material must be APL_SNH, APL_ARAP or APL_SNH_MUSCLE");
    return -1;
  }
  *out = nullptr;
  APL_CHECK_HIP(hipSetDevice(m->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  apl_solver* s = new apl_solver();
  s->run = new apl_solver_run();
  s->mesh = m;
  ++m->refs;
  s->device = m->device;
  s->material = material;
  const int n = (int)m->L.n_inst;
  const size_t vb = 9 * 3 * (size_t)std::max<int64_t>(m->L.V, 1) * real_size(m);
  const size_t xb = m->dtype == DT_F32 ? 3 * (size_t)std::max<int64_t>(m->L.V, 1) * sizeof(double) : 0;
  if (hipMalloc(&s->vecs, vb) != hipSuccess || (xb && hipMalloc(&s->x64, xb) != hipSuccess) ||
      hipMalloc((void**)&s->sd, N_SD * (size_t)n * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&s->si, N_SI * (size_t)n * sizeof(int32_t)) != hipSuccess ||
      hipHostMalloc((void**)&s->host_status, (size_t)n * sizeof(int32_t), 0) != hipSuccess ||
      hipHostMalloc((void**)&s->host_steps, (size_t)n * sizeof(int32_t), 0) != hipSuccess) {
    set_error("out of device memory creating the solver");
    apl_solver_destroy(s);
    return -1;
  }
  if (hipMemset(s->vecs, 0, vb) != hipSuccess) {
    set_error("hipMemset failed creating the solver");
    apl_solver_destroy(s);
    return -1;
  }
  double* d = s->sd;
  s->st.f = d + 0 * n;
  s->st.gd = d + 1 * n;
  s->st.alpha = d + 2 * n;
  s->st.alpha_commit = d + 3 * n;
  s->st.beta = d + 4 * n;
  s->st.gnorm0 = d + 5 * n;
  s->st.gnorm = d + 6 * n;
  s->st.pmean = d + 7 * n;
  s->st.f_first = d + 8 * n;
  s->st.pHp = d + 9 * n;
  s->st.gbest = d + 10 * n;
  s->st.f_mark = d + 11 * n;
  s->st.f0 = d + 12 * n;
  int32_t* ii = s->si;
  s->st.status = ii + 0 * n;
  s->st.act_eval = ii + 1 * n;
  s->st.act_dir = ii + 2 * n;
  s->st.commit = ii + 3 * n;
  s->st.n_steps = ii + 4 * n;
  s->st.n_evals = ii + 5 * n;
  s->st.ls_step = ii + 6 * n;
  s->st.n_backtracks = ii + 7 * n;
  s->st.phase = ii + 8 * n;
  s->st.restarts = ii + 9 * n;
  s->st.best_step = ii + 10 * n;
  s->st.ls_last = ii + 11 * n;
  *out = s;
  return 0;
}

int apl_solver_destroy(apl_solver* s) {
  if (s == nullptr) return 0;
  drop_hip_error(hipSetDevice(s->device));
  if (s->gexec) drop_hip_error(hipGraphExecDestroy(s->gexec));
  if (s->graph) drop_hip_error(hipGraphDestroy(s->graph));
  if (s->cg_gexec) drop_hip_error(hipGraphExecDestroy(s->cg_gexec));
  if (s->cg_graph) drop_hip_error(hipGraphDestroy(s->cg_graph));
  if (s->cap) drop_hip_error(hipStreamDestroy(s->cap));
  if (s->vecs) drop_hip_error(hipFree(s->vecs));
  if (s->x64) drop_hip_error(hipFree(s->x64));
  if (s->sd) drop_hip_error(hipFree(s->sd));
  if (s->si) drop_hip_error(hipFree(s->si));
  if (s->host_status) drop_hip_error(hipHostFree(s->host_status));
  if (s->host_steps) drop_hip_error(hipHostFree(s->host_steps));
  delete s->run;
  if (s->mesh) mesh_release(s->mesh);
  delete s;
  return 0;
}

int apl_pncg_solve(apl_solver* s, const void* mu, const void* lam, void* x, const void* f_ext,
                   const uint8_t* fixed_mask, const apl_pncg_params* params, apl_pncg_stats* stats, void* stream) {
  if (s == nullptr || params == nullptr || (s->mesh->L.V > 0 && x == nullptr) ||
      (s->mesh->L.T > 0 && (mu == nullptr || (needs_lambda(s->mesh, s->material) && lam == nullptr)))) {
    set_error("NULL argument");
    return -1;
  }
  if (params->max_steps < 0 || params->max_backtracking < 0 || !(params->rtol >= 0) || !(params->atol >= 0)) {
    set_error("invalid PNCG parameters");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(s->mesh->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  if (s->mesh->dtype == DT_F32)
    return pncg_solve_t<float>(s, mu, lam, x, f_ext, fixed_mask, params, stats, (hipStream_t)stream);
  return pncg_solve_t<double>(s, mu, lam, x, f_ext, fixed_mask, params, stats, (hipStream_t)stream);
}

int apl_pncg_begin(apl_solver* s, const void* mu, const void* lam, void* x, const void* f_ext,
                   const uint8_t* fixed_mask, const apl_pncg_params* params, void* stream) {
  if (s == nullptr || params == nullptr || (s->mesh->L.V > 0 && x == nullptr) ||
      (s->mesh->L.T > 0 && (mu == nullptr || (needs_lambda(s->mesh, s->material) && lam == nullptr)))) {
    set_error("NULL argument");
    return -1;
  }
  if (params->max_steps < 0 || params->max_backtracking < 0 || !(params->rtol >= 0) || !(params->atol >= 0)) {
    set_error("invalid PNCG parameters");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(s->mesh->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  if (s->mesh->dtype == DT_F32) return pncg_begin_t<float>(s, mu, lam, x, f_ext, fixed_mask, params, (hipStream_t)stream);
  return pncg_begin_t<double>(s, mu, lam, x, f_ext, fixed_mask, params, (hipStream_t)stream);
}

int apl_pncg_advance(apl_solver* s, int32_t steps, int64_t max_iters, void* x_out, apl_pncg_state* state,
                     void* stream) {
  if (s == nullptr || !s->run->r.open) {
    set_error("no PNCG run in progress: call apl_pncg_begin first");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(s->mesh->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  hipStream_t st = (hipStream_t)stream;
  apl_mesh* m = s->mesh;
  const bool f32 = m->dtype == DT_F32;
  if ((f32 ? pncg_advance_t<float>(s, max_iters, steps, st) : pncg_advance_t<double>(s, max_iters, steps, st)))
    return -1;
  const long long nx = 3 * (long long)m->L.V;
  if (x_out != nullptr && nx > 0) {
    if (f32) {
      hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)std::min<long long>((nx + NTH - 1) / NTH, 4096)),
                         dim3(NTH), 0, st, (const double*)s->run->r.xs, (float*)x_out, nx);
      APL_CHECK_HIP(hipGetLastError());
    } else if (x_out != s->run->r.xs) {
      APL_CHECK_HIP(hipMemcpyAsync(x_out, s->run->r.xs, nx * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
  }
  if (state != nullptr) {
    const int n = (int)m->L.n_inst;
    std::vector<double> hd((size_t)N_SD * n);
    std::vector<int32_t> hi((size_t)N_SI * n);
    APL_CHECK_HIP(hipMemcpyAsync(hd.data(), s->sd, hd.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipMemcpyAsync(hi.data(), s->si, hi.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    APL_CHECK_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) {
      apl_pncg_state& o = state[i];
      o.status = hi[0 * n + i];
      o.n_steps = hi[4 * n + i];
      o.n_evals = hi[5 * n + i];
      o.ls_steps = hi[11 * n + i];
      o.ok = hi[9 * n + i] == 0 ? 1 : 0;
      o.alpha = hd[3 * n + i];
      o.f0 = hd[12 * n + i];
      o.f_alpha = hd[0 * n + i];
      o.grad_norm = hd[6 * n + i];
      o.grad_norm_first = hd[5 * n + i];
      o.hess_diag_mean = hd[7 * n + i];
      o.beta = hd[4 * n + i];
    }
  }
  return 0;
}

int apl_pncg_end(apl_solver* s, apl_pncg_stats* stats, void* stream) {
  if (s == nullptr || !s->run->r.open) {
    set_error("no PNCG run in progress: call apl_pncg_begin first");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(s->mesh->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  hipStream_t st = (hipStream_t)stream;
  const int rc = s->mesh->dtype == DT_F32 ? pncg_end_t<float>(s, stats, st) : pncg_end_t<double>(s, stats, st);
  s->run->r.open = false;
  return rc;
}

int apl_cg_solve(apl_solver* s, const void* mu, const void* lam, const void* x, const void* rhs, void* y,
                 const uint8_t* fixed_mask, double rtol, int max_iters, apl_pncg_stats* stats, void* stream) {
  if (s == nullptr || (s->mesh->L.V > 0 && (x == nullptr || rhs == nullptr || y == nullptr)) ||
      (s->mesh->L.T > 0 && (mu == nullptr || (needs_lambda(s->mesh, s->material) && lam == nullptr)))) {
    set_error("NULL argument");
    return -1;
  }
  if (!(rtol >= 0) || max_iters < 0) {
    set_error("invalid CG parameters");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(s->mesh->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  if (s->mesh->dtype == DT_F32)
    return cg_solve_t<float>(s, mu, lam, x, rhs, y, fixed_mask, rtol, max_iters, stats, (hipStream_t)stream, 0);
  return cg_solve_t<double>(s, mu, lam, x, rhs, y, fixed_mask, rtol, max_iters, stats, (hipStream_t)stream, 0);
}

int apl_cg_solve_normal(apl_solver* s, const void* mu, const void* lam, const void* x, const void* rhs, void* y,
                        const uint8_t* fixed_mask, double rtol, int max_iters, apl_pncg_stats* stats, void* stream) {
  if (s == nullptr || (s->mesh->L.V > 0 && (x == nullptr || rhs == nullptr || y == nullptr)) ||
      (s->mesh->L.T > 0 && (mu == nullptr || (needs_lambda(s->mesh, s->material) && lam == nullptr)))) {
    set_error("NULL argument");
    return -1;
  }
  if (!(rtol >= 0) || max_iters < 0) {
    set_error("invalid CG parameters");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(s->mesh->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  if (s->mesh->dtype == DT_F32)
    return cg_solve_t<float>(s, mu, lam, x, rhs, y, fixed_mask, rtol, max_iters, stats, (hipStream_t)stream, 1);
  return cg_solve_t<double>(s, mu, lam, x, rhs, y, fixed_mask, rtol, max_iters, stats, (hipStream_t)stream, 1);
}

}  // extern "C"
