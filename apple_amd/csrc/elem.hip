// Element kernels: one pass over the tets of every block computes any of
// energy, gradient, clamped Hessian diagonal, Hessian-vector product,
// clamped quadratic form and mixed parameter derivatives, and reduces the
// per-vertex quantities deterministically.
//
// Replaces the fun/grad/hess_diag/hess_prod/hess_quad kernel templates of
// warp/fem/_base.py:243-383 (one launch and one re-gather of u per quantity,
// float atomics for every reduction) by one fused launch per pass:
//
//  * each workgroup (256 threads) owns a run of consecutive tets (a block) and
//    stages the displacement of every vertex the block touches into LDS once
//    (slots listed in loc_gid), so the 4 corner gathers per tet read LDS;
//  * per-tet element data are SoA in HBM (9 x T inverse rest shapes, dV, mu,
//    lambda, 4 x uint16 local corner slots): every load is coalesced;
//  * per-corner contributions go to LDS in [channel][corner][thread] order
//    (bank-conflict free writes) and each local vertex sums its incidences
//    over an ascending CSR list -- a fixed order, so results are bitwise
//    reproducible; owned vertices are written once, shared vertices leave a
//    partial that the vertex pass sums in ascending block order;
//  * energies and quadratic forms are summed per thread in FP64, then by a
//    fixed wave-shuffle tree and a per-instance pass over the block partials;
//  * the grid is persistent and XCD-aware: workgroup i runs on XCD i % 8, and
//    XCD x walks the contiguous x-th eighth of the blocks, so neighbouring
//    blocks (which share vertices) hit the same L2.
#include "internal.h"
#include "tetmath.h"

namespace apl {

template <class R>
struct ElemArgs {
  const R* dminv;
  const R* dV;
  const ushort4* slot;
  const int32_t* blk_tet_off;
  const int32_t* blk_loc_off;
  const int32_t* blk_nown;
  const int32_t* blk_sh_base;
  const int32_t* blk_inst;
  const int32_t* loc_gid;
  const uint16_t* loc_inc_off;
  const uint16_t* loc_inc;
  const R* mu;
  const R* lam;
  const R* x;
  const R* p;
  const double* alpha;
  const int32_t* active;
  R* out_a;
  R* out_b;
  R* partial;
  double* blk_red;
  long long T;
  long long n_partials;
  int nblocks;
  int nloc_max;
  int accumulate;
};

__device__ inline double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NTH / 64; ++i) r += red[i];
  }
  __syncthreads();
  return r;
}

template <int MAT, unsigned WHAT, class R>
__global__ __launch_bounds__(NTH) void elem_kernel(ElemArgs<R> A) {
  constexpr bool HAS_G = (WHAT & W_G) != 0;
  constexpr bool HAS_H = (WHAT & W_H) != 0;
  constexpr bool HAS_HP = (WHAT & W_HP) != 0;
  constexpr int NCH = (HAS_G ? 3 : 0) + (HAS_H ? 3 : 0) + (HAS_HP ? 3 : 0);
  constexpr bool SCATTER = NCH > 0;
  constexpr int NCHA = NCH > 0 ? NCH : 1;
  constexpr bool NEEDV = (WHAT & (W_HP | W_HQ | W_MIX)) != 0;
  constexpr bool RED = (WHAT & (W_E | W_HQ)) != 0;
  static_assert(NCH <= MAX_CH, "too many channels");

  extern __shared__ double smem_d[];
  double* red = smem_d;
  R* u_lds = reinterpret_cast<R*>(smem_d + 8);
  R* v_lds = u_lds + 3 * A.nloc_max;
  R* contrib = v_lds + (NEEDV ? 3 * A.nloc_max : 0);

  const int tid = threadIdx.x;
  const int pid = blockIdx.x;
  const int xcd = pid & 7, k0 = pid >> 3, K = gridDim.x >> 3;
  const long long nb = A.nblocks;
  const int lo = (int)((xcd * nb) / 8), hi = (int)(((xcd + 1) * nb) / 8);

  for (int b = lo + k0; b < hi; b += K) {
    const int inst = A.blk_inst[b];
    if (A.active != nullptr && A.active[inst] == 0) continue;
    const int t0 = A.blk_tet_off[b];
    const int nt = A.blk_tet_off[b + 1] - t0;
    const int sb = A.blk_loc_off[b];
    const int nloc = A.blk_loc_off[b + 1] - sb;
    const int nown = A.blk_nown[b];
    const R alpha = A.alpha != nullptr ? (R)A.alpha[inst] : R(0);

    // stage the block's vertex window (u = x + alpha p for a trial point)
    for (int s = tid; s < nloc; s += NTH) {
      const long long gid = A.loc_gid[sb + s];
      R ux = A.x[3 * gid], uy = A.x[3 * gid + 1], uz = A.x[3 * gid + 2];
      if (A.alpha != nullptr) {
        ux += alpha * A.p[3 * gid];
        uy += alpha * A.p[3 * gid + 1];
        uz += alpha * A.p[3 * gid + 2];
      }
      u_lds[3 * s] = ux;
      u_lds[3 * s + 1] = uy;
      u_lds[3 * s + 2] = uz;
      if constexpr (NEEDV) {
        v_lds[3 * s] = A.p[3 * gid];
        v_lds[3 * s + 1] = A.p[3 * gid + 1];
        v_lds[3 * s + 2] = A.p[3 * gid + 2];
      }
    }
    __syncthreads();

    const uint16_t* inc_off = A.loc_inc_off + sb + b;
    const uint16_t* inc = A.loc_inc + 4LL * t0;
    R acc[SPT][NCHA];
    int cur[SPT], endi[SPT];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int s = tid + k * NTH;
      cur[k] = 0;
      endi[k] = 0;
      if (SCATTER && s < nloc) {
        cur[k] = inc_off[s];
        endi[k] = inc_off[s + 1];
      }
#pragma unroll
      for (int ch = 0; ch < NCHA; ++ch) acc[k][ch] = R(0);
    }
    double e_acc = 0.0, q_acc = 0.0;
    const int nchunks = (nt + NTH - 1) / NTH;
    for (int c = 0; c < nchunks; ++c) {
      const int tl = c * NTH + tid;
      if (tl < nt) {
        const long long t = t0 + tl;
        R B[9];
#pragma unroll
        for (int kk = 0; kk < 9; ++kk) B[kk] = A.dminv[kk * A.T + t];
        const TetGeom<R> g = make_geom(B, A.dV[t]);
        const ushort4 sl = A.slot[t];
        const int sidx[4] = {sl.x, sl.y, sl.z, sl.w};
        V3<R> u[4], v[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          u[a] = v3(u_lds[3 * sidx[a]], u_lds[3 * sidx[a] + 1], u_lds[3 * sidx[a] + 2]);
          if constexpr (NEEDV)
            v[a] = v3(v_lds[3 * sidx[a]], v_lds[3 * sidx[a] + 1], v_lds[3 * sidx[a] + 2]);
          else
            v[a] = v3(R(0), R(0), R(0));
        }
        const R mu = A.mu[t];
        const R lam = (MAT == MAT_SNH) ? A.lam[t] : R(0);
        TetOut<R> o;
        tet_eval<MAT, WHAT, R>(g, mu, lam, u, v, o);
        if constexpr ((WHAT & W_E) != 0) e_acc += (double)o.e;
        if constexpr ((WHAT & W_HQ) != 0) q_acc += (double)o.hq;
        if constexpr ((WHAT & W_MIX) != 0) {
          if (A.accumulate) {
            A.out_a[t] += o.dmu;
            if (MAT == MAT_SNH) A.out_b[t] += o.dlam;
          } else {
            A.out_a[t] = o.dmu;
            if (MAT == MAT_SNH) A.out_b[t] = o.dlam;
          }
        }
        if constexpr (SCATTER) {
          int ch = 0;
          if constexpr (HAS_G) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              contrib[((ch + 0) * 4 + a) * NTH + tid] = o.g[a].x;
              contrib[((ch + 1) * 4 + a) * NTH + tid] = o.g[a].y;
              contrib[((ch + 2) * 4 + a) * NTH + tid] = o.g[a].z;
            }
            ch += 3;
          }
          if constexpr (HAS_H) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              contrib[((ch + 0) * 4 + a) * NTH + tid] = o.h[a].x;
              contrib[((ch + 1) * 4 + a) * NTH + tid] = o.h[a].y;
              contrib[((ch + 2) * 4 + a) * NTH + tid] = o.h[a].z;
            }
            ch += 3;
          }
          if constexpr (HAS_HP) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              contrib[((ch + 0) * 4 + a) * NTH + tid] = o.hp[a].x;
              contrib[((ch + 1) * 4 + a) * NTH + tid] = o.hp[a].y;
              contrib[((ch + 2) * 4 + a) * NTH + tid] = o.hp[a].z;
            }
          }
        }
      }
      if constexpr (SCATTER) {
        __syncthreads();
        const int lim = (c + 1) * 4 * NTH;
        const int base = c * 4 * NTH;
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
          while (cur[k] < endi[k]) {
            const int id = inc[cur[k]];
            if (id >= lim) break;
            const int j = id - base;
            const int tlc = j >> 2, a = j & 3;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) acc[k][ch] += contrib[(ch * 4 + a) * NTH + tlc];
            ++cur[k];
          }
        }
        __syncthreads();
      }
    }

    if constexpr (SCATTER) {
      R* outs[2] = {A.out_a, A.out_b};
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        const int s = tid + k * NTH;
        if (s >= nloc) continue;
        if (s < nown) {
          const long long gid = A.loc_gid[sb + s];
#pragma unroll
          for (int grp = 0; grp < NCH / 3; ++grp) {
            // group -> output: (G,H) -> (a,b); H alone -> b; HP -> a
            R* out = (HAS_G || HAS_HP) ? outs[grp] : outs[1];
            if (A.accumulate) {
              out[3 * gid] += acc[k][3 * grp];
              out[3 * gid + 1] += acc[k][3 * grp + 1];
              out[3 * gid + 2] += acc[k][3 * grp + 2];
            } else {
              out[3 * gid] = acc[k][3 * grp];
              out[3 * gid + 1] = acc[k][3 * grp + 1];
              out[3 * gid + 2] = acc[k][3 * grp + 2];
            }
          }
        } else {
          const long long pidx = (long long)A.blk_sh_base[b] + (s - nown);
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch) A.partial[ch * A.n_partials + pidx] = acc[k][ch];
        }
      }
    }
    if constexpr (RED) {
      const double e = block_sum(e_acc, red);
      const double q = block_sum(q_acc, red);
      if (tid == 0) {
        A.blk_red[2 * b] = e;
        A.blk_red[2 * b + 1] = q;
      }
    }
    __syncthreads();
  }
}

// Sum the partials of every shared vertex (ascending block order).
template <class R, int NCH>
__global__ __launch_bounds__(NTH) void shared_reduce_kernel(
    const int32_t* sh_vtx, const int32_t* sh_off, const int32_t* sh_src, const R* partial,
    long long n_partials, long long n_sh, const int64_t* inst_vtx_off, long long n_inst,
    const int32_t* active, R* out_a, R* out_b, int h_only, int accumulate) {
  const long long i = (long long)blockIdx.x * NTH + threadIdx.x;
  if (i >= n_sh) return;
  const long long v = sh_vtx[i];
  if (active != nullptr) {
    // instance of v by binary search over the vertex offsets
    long long lo = 0, hi = n_inst;
    while (hi - lo > 1) {
      long long mid = (lo + hi) >> 1;
      if (inst_vtx_off[mid] <= v) lo = mid; else hi = mid;
    }
    if (active[lo] == 0) return;
  }
  R acc[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) acc[ch] = R(0);
  for (int k = sh_off[i]; k < sh_off[i + 1]; ++k) {
    const long long pidx = sh_src[k];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) acc[ch] += partial[ch * n_partials + pidx];
  }
#pragma unroll
  for (int grp = 0; grp < NCH / 3; ++grp) {
    R* out = h_only ? out_b : (grp == 0 ? out_a : out_b);
    if (accumulate) {
      out[3 * v] += acc[3 * grp];
      out[3 * v + 1] += acc[3 * grp + 1];
      out[3 * v + 2] += acc[3 * grp + 2];
    } else {
      out[3 * v] = acc[3 * grp];
      out[3 * v + 1] = acc[3 * grp + 1];
      out[3 * v + 2] = acc[3 * grp + 2];
    }
  }
}

// Per-instance sums of the block energies and quadratic forms, fixed order.
__global__ __launch_bounds__(NTH) void inst_reduce_kernel(const double* blk_red, const int64_t* inst_blk_off,
                                                          const int32_t* active, double* out_e,
                                                          double* out_q, int accumulate) {
  __shared__ double red[8];
  const int i = blockIdx.x;
  if (active != nullptr && active[i] == 0) return;
  const long long b0 = inst_blk_off[i], b1 = inst_blk_off[i + 1];
  double e = 0.0, q = 0.0;
  for (long long b = b0 + threadIdx.x; b < b1; b += NTH) {
    e += blk_red[2 * b];
    q += blk_red[2 * b + 1];
  }
  e = block_sum(e, red);
  q = block_sum(q, red);
  if (threadIdx.x == 0) {
    if (out_e != nullptr) out_e[i] = accumulate ? out_e[i] + e : e;
    if (out_q != nullptr) out_q[i] = accumulate ? out_q[i] + q : q;
  }
}

// ------------------------------------------------------------------ launchers

template <class R>
static ElemArgs<R> make_args(apl_mesh* m, const void* mu, const void* lam, const void* x, const void* p,
                             const double* alpha, const int32_t* active, void* out_a, void* out_b,
                             int accumulate) {
  ElemArgs<R> A;
  A.dminv = (const R*)m->d.dminv;
  A.dV = (const R*)m->d.dV;
  A.slot = (const ushort4*)m->d.slot;
  A.blk_tet_off = m->d.blk_tet_off;
  A.blk_loc_off = m->d.blk_loc_off;
  A.blk_nown = m->d.blk_nown;
  A.blk_sh_base = m->d.blk_sh_base;
  A.blk_inst = m->d.blk_inst;
  A.loc_gid = m->d.loc_gid;
  A.loc_inc_off = m->d.loc_inc_off;
  A.loc_inc = m->d.loc_inc;
  A.mu = (const R*)mu;
  A.lam = (const R*)lam;
  A.x = (const R*)x;
  A.p = (const R*)p;
  A.alpha = alpha;
  A.active = active;
  A.out_a = (R*)out_a;
  A.out_b = (R*)out_b;
  A.partial = (R*)m->d.partial;
  A.blk_red = m->d.blk_red;
  A.T = m->L.T;
  A.n_partials = m->L.n_partials;
  A.nblocks = (int)m->L.nblocks;
  A.nloc_max = m->L.nloc_max;
  A.accumulate = accumulate;
  return A;
}

template <int MAT, unsigned WHAT, class R>
static size_t smem_bytes(const apl_mesh* m) {
  constexpr bool NEEDV = (WHAT & (W_HP | W_HQ | W_MIX)) != 0;
  constexpr int NCH = ((WHAT & W_G) ? 3 : 0) + ((WHAT & W_H) ? 3 : 0) + ((WHAT & W_HP) ? 3 : 0);
  size_t n = 8 * sizeof(double) + sizeof(R) * 3 * (size_t)m->L.nloc_max * (NEEDV ? 2 : 1);
  n += sizeof(R) * (size_t)NCH * 4 * NTH;
  return n;
}

template <int MAT, unsigned WHAT, class R>
static int grid_for(apl_mesh* m) {
  static_assert(WHAT < 64, "pass variant index");
  int& cached = m->grid_cache[MAT][WHAT];
  if (cached > 0) return cached;
  const size_t smem = smem_bytes<MAT, WHAT, R>(m);
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, elem_kernel<MAT, WHAT, R>, NTH, smem) != hipSuccess ||
      occ < 1)
    occ = 1;
  long long g = (long long)occ * m->n_cu;
  long long nb8 = ((m->L.nblocks + 7) / 8) * 8;
  if (g > nb8) g = nb8;
  g = ((g + 7) / 8) * 8;
  if (g < 8) g = 8;
  cached = (int)g;
  return cached;
}

template <int MAT, unsigned WHAT, class R>
static int launch_one(apl_mesh* m, const void* mu, const void* lam, const void* x, const void* p,
                      const double* alpha, const int32_t* active, void* out_a, void* out_b,
                      int accumulate, int grid_override, hipStream_t s) {
  if (m->L.nblocks == 0) return 0;
  const size_t smem = smem_bytes<MAT, WHAT, R>(m);
  if (smem > 65536) {
    set_error("element block needs more than 64 KiB of LDS; rebuild the mesh with a smaller nloc_cap");
    return -1;
  }
  const int grid = grid_override > 0 ? ((grid_override + 7) / 8) * 8 : grid_for<MAT, WHAT, R>(m);
  ElemArgs<R> A = make_args<R>(m, mu, lam, x, p, alpha, active, out_a, out_b, accumulate);
  hipLaunchKernelGGL((elem_kernel<MAT, WHAT, R>), dim3(grid), dim3(NTH), smem, s, A);
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

// The pass variants the library instantiates.
#define APL_WHAT_LIST(X)         \
  X(W_E)                         \
  X(W_E | W_G | W_H)             \
  X(W_G)                         \
  X(W_H)                         \
  X(W_HP)                        \
  X(W_HQ)                        \
  X(W_MIX)

template <int MAT, class R>
static int dispatch_what(apl_mesh* m, unsigned what, const void* mu, const void* lam, const void* x,
                         const void* p, const double* alpha, const int32_t* active, void* out_a,
                         void* out_b, int accumulate, int grid_override, hipStream_t s) {
#define APL_CASE(W) \
  if (what == (W)) return launch_one<MAT, (W), R>(m, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s);
  APL_WHAT_LIST(APL_CASE)
#undef APL_CASE
  set_error("unsupported combination of outputs for one element pass");
  return -1;
}

int launch_elem(apl_mesh* m, int material, unsigned what, const void* mu, const void* lam, const void* x,
                const void* p, const double* alpha, const int32_t* active, void* out_a, void* out_b,
                int accumulate, int grid_override, hipStream_t s) {
  if (material == MAT_SNH) {
    if (m->dtype == DT_F32)
      return dispatch_what<MAT_SNH, float>(m, what, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s);
    return dispatch_what<MAT_SNH, double>(m, what, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s);
  }
  if (material == MAT_ARAP) {
    if (m->dtype == DT_F32)
      return dispatch_what<MAT_ARAP, float>(m, what, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s);
    return dispatch_what<MAT_ARAP, double>(m, what, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s);
  }
  set_error("unknown material");
  return -1;
}

template <int MAT>
static int grid_mat(apl_mesh* m, unsigned what) {
#define APL_CASE(W) \
  if (what == (W)) return m->dtype == DT_F32 ? grid_for<MAT, (W), float>(m) : grid_for<MAT, (W), double>(m);
  APL_WHAT_LIST(APL_CASE)
#undef APL_CASE
  return 0;
}

int elem_grid(apl_mesh* m, int material, unsigned what) {
  return material == MAT_SNH ? grid_mat<MAT_SNH>(m, what) : grid_mat<MAT_ARAP>(m, what);
}

template <class R, int NCH>
static int shared_one(apl_mesh* m, void* out_a, void* out_b, int h_only, int accumulate, const int32_t* active,
                      hipStream_t s) {
  const long long n = m->L.n_shared_vtx;
  if (n == 0) return 0;
  const int grid = (int)((n + NTH - 1) / NTH);
  hipLaunchKernelGGL((shared_reduce_kernel<R, NCH>), dim3(grid), dim3(NTH), 0, s, m->d.sh_vtx, m->d.sh_off,
                     m->d.sh_src, (const R*)m->d.partial, (long long)m->L.n_partials, n, m->d.inst_vtx_off,
                     (long long)m->L.n_inst, active, (R*)out_a, (R*)out_b, h_only, accumulate);
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_shared_reduce(apl_mesh* m, unsigned what, void* out_a, void* out_b, int accumulate,
                         const int32_t* active, hipStream_t s) {
  const bool g = what & W_G, h = what & W_H, hp = what & W_HP;
  const int nch = (g ? 3 : 0) + (h ? 3 : 0) + (hp ? 3 : 0);
  const int h_only = (h && !g && !hp) ? 1 : 0;
  if (nch == 0) return 0;
  if (m->dtype == DT_F32)
    return nch == 6 ? shared_one<float, 6>(m, out_a, out_b, h_only, accumulate, active, s)
                    : shared_one<float, 3>(m, out_a, out_b, h_only, accumulate, active, s);
  return nch == 6 ? shared_one<double, 6>(m, out_a, out_b, h_only, accumulate, active, s)
                  : shared_one<double, 3>(m, out_a, out_b, h_only, accumulate, active, s);
}

int launch_inst_reduce(apl_mesh* m, double* out_e, double* out_q, int accumulate, const int32_t* active,
                       hipStream_t s) {
  hipLaunchKernelGGL(inst_reduce_kernel, dim3((unsigned)m->L.n_inst), dim3(NTH), 0, s, m->d.blk_red,
                     m->d.inst_blk_off, active, out_e, out_q, accumulate);
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace apl
