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
//    (padded to 4 lanes, one wide LDS read per corner);
//  * per-tet element data are SoA in HBM (9 x T inverse rest shapes, dV, mu,
//    lambda, and one 16-byte record of 4 local corner slots and 4 corner
//    positions), addressed by 32-bit offsets from per-block uniform base
//    pointers: every load is coalesced;
//  * each corner's contribution goes to LDS at its position (corners ranked
//    by slot, then tet, then corner; channels packed into 16/8-byte records),
//    so the corners of a local vertex form a contiguous run; one to four lanes sum each run in a fixed order
//    (bitwise reproducible) and one partial per slot goes into a
//    vertex-major buffer (a vertex's partials contiguous, in block order);
//    the vertex pass adds each vertex's contiguous run;
//  * energies and quadratic forms are summed per thread, then in FP64 by a
//    fixed wave-shuffle tree into one partial per wave (no block barrier),
//    and a per-instance pass adds the wave partials in a fixed order;
//  * the grid is persistent and XCD-aware: workgroup i runs on XCD i % 8, and
//    XCD x walks every eighth chunk of up to 16 consecutive blocks, so
//    neighbouring blocks (which share vertices) mostly hit the same L2 while
//    any instance of more than 8 chunks spreads over all XCDs.
#include "internal.h"
#include "tetmath.h"
#include "flatsum.h"
#include "persist.h"
#include "vsum.h"

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>

namespace apl {

constexpr int XCD_CHUNK = 16;  // consecutive blocks per XCD chunk
constexpr size_t ELEM_LDS_MAX = 65536;  // LDS one element workgroup may use

template <class R, class X>
struct ElemArgs {
  const R* dminv;
  const R* dV;
  const uint4* corner;  // per tet: 4 x u16 local slots, 4 x u16 chunk positions
  const int32_t* blk_tet_off;
  const int32_t* blk_loc_off;
  const int32_t* blk_inst;
  const int32_t* loc_gid;
  const int32_t* loc_dst;
  const int32_t* blk_co_off;
  const uint16_t* chunk_off;
  const R* mu;
  const R* lam;
  const R* act;  // (T, 6) activation of the muscle law
  const X* x;  // displacements (X = R, or double when the solver keeps them in FP64)
  const R* p;
  const double* alpha;
  const int32_t* active;
  R* out_a;  // per-tet outputs of the mixed pass
  R* out_b;
  R* out_c;  // (T, 6) activation derivatives of the mixed pass
  R* partial;  // vertex-major partial sums (scatter passes)
  double* blk_red;
  const int32_t* blk_list;  // model meshes: this law's blocks (else NULL: every block)
  int T;
  int nblocks;  // blocks this launch visits (the list's length with a list)
  int nloc_max;
  int chunks;
  int accumulate;
};

// Cross-lane moves by DPP (VALU; no LDS round trip as with ds_bpermute).
// Lanes whose DPP source does not exist (row_shr at a row start, rows the
// row mask leaves out) read `old` = 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xF, false));
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <class T, int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ T dpp(T v) {
  if constexpr (sizeof(T) == 4)
    return dpp_f32<CTRL, ROW_MASK>(v);
  else
    return dpp_f64<CTRL, ROW_MASK>(v);
}
// gfx9 DPP controls
constexpr int DPP_QUAD_XOR1 = 0xB1;  // quad_perm [1,0,3,2]: the partner lane ^ 1
constexpr int DPP_QUAD_XOR2 = 0x4E;  // quad_perm [2,3,0,1]: the partner lane ^ 2
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

// Sum over the 64 lanes of a wave by a fixed DPP tree: prefix sums within
// each row of 16 lanes, then row 0's total into row 1 and row 2's into
// row 3, then lane 31's into rows 2-3.  The total is in lane 63.
__device__ inline double wave_sum(double v) {
  v += dpp_f64<DPP_ROW_SHR1>(v);
  v += dpp_f64<DPP_ROW_SHR2>(v);
  v += dpp_f64<DPP_ROW_SHR4>(v);
  v += dpp_f64<DPP_ROW_SHR8>(v);
  v += dpp_f64<DPP_ROW_BCAST15, 0xA>(v);
  v += dpp_f64<DPP_ROW_BCAST31, 0xC>(v);
  return v;
}

// Element i of a uniform base pointer through a 32-bit byte offset: maps to
// the scalar-base + vector-offset load/store form (no per-lane 64-bit
// address arithmetic).  Mesh sizes are limited so that offsets fit 32 bits.
template <class T>
__device__ __forceinline__ const T& at(const T* base, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ T& at(T* base, unsigned i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(T));
}

// Raw buffer loads: a scalar resource (base address) plus a per-lane and a
// uniform byte offset.  Used for the per-tet record so that no lane does
// 64-bit address arithmetic.  The resource's fourth word is the gfx9 raw
// buffer format; the record count is left unbounded (indices are exact).
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, -1, 0x00020000);
}
template <class R>
__device__ __forceinline__ R bload(rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float bload<float>(rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double bload<double>(rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ uint4 bload_u4(rsrc_t r, unsigned voff, unsigned soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return make_uint4(v[0], v[1], v[2], v[3]);
}

// Per-tet record, loaded into registers ahead of the block's LDS staging.
template <class R>
struct TetRec {
  R B[9];
  R dV, mu, lam;
  R act[6];  // muscle activation
  uint4 cr;  // local slots (x, y), positions (z, w): two u16 per word
};

// Resources of one block's tet data.  dminv is SoA (row kk at kk * T), so its
// base moves with the block and the row offset kk * T * sizeof(R) stays
// below 2^32 bytes (mesh creation limits T accordingly).
struct BlockRes {
  rsrc_t dm, dv, mu, lam, act, cr;
  unsigned row, t0b;  // row stride and block start, in bytes of R
};

template <int MAT, class R, class X>
__device__ inline BlockRes block_res(const ElemArgs<R, X>& A, int t0) {
  BlockRes br;
  br.dm = make_rsrc(A.dminv + t0);
  br.dv = make_rsrc(A.dV);
  br.mu = make_rsrc(A.mu);
  br.lam = make_rsrc(MAT != MAT_ARAP ? A.lam : A.mu);
  br.act = make_rsrc(MAT == MAT_MUSCLE ? A.act + 6 * (size_t)t0 : A.mu);  // block-relative (T x 24 B may pass 2^32)
  br.cr = make_rsrc(A.corner);
  br.row = (unsigned)A.T * (unsigned)sizeof(R);
  br.t0b = (unsigned)t0 * (unsigned)sizeof(R);
  return br;
}

template <int MAT, class R>
__device__ inline void load_rec(const BlockRes& br, unsigned tl, TetRec<R>& r) {
  const unsigned vo = tl * (unsigned)sizeof(R);
#pragma unroll
  for (int kk = 0; kk < 9; ++kk) r.B[kk] = bload<R>(br.dm, vo, (unsigned)kk * br.row);
  r.dV = bload<R>(br.dv, vo, br.t0b);
  r.mu = bload<R>(br.mu, vo, br.t0b);
  r.lam = (MAT != MAT_ARAP) ? bload<R>(br.lam, vo, br.t0b) : R(0);
  if constexpr (MAT == MAT_MUSCLE) {
#pragma unroll
    for (int k = 0; k < 6; ++k) r.act[k] = bload<R>(br.act, 6u * vo + (unsigned)k * (unsigned)sizeof(R), 0u);
  }
  r.cr = bload_u4(br.cr, tl * 16u, br.t0b / (unsigned)sizeof(R) * 16u);
}

// 4-lane padded vector (one 16/32-byte LDS access)
template <class T>
struct alignas(4 * sizeof(T)) Vec4 {
  T v[4];
};
template <class T>
struct alignas(2 * sizeof(T)) Vec2 {
  T v[2];
};

// Channel packing of the contribution buffer: one region per group of 4
// channels ([position][4]: one 16/32-byte LDS access per corner and per
// position), then a region for the remainder (a pair, a single channel, or
// 3 channels padded to 4).  Pass A's 6 channels take a 16-byte and an 8-byte
// access instead of six 4-byte ones.
template <int NCH>
struct Pack {
  static constexpr int G4 = NCH / 4;
  static constexpr int REM = NCH % 4;
  static constexpr int REMW = REM == 3 ? 4 : REM;  // stored width of the remainder
  static constexpr int NCHP = 4 * G4 + REMW;       // stored channels per position
};

// SPT: slot items per thread.  A block with nloc slots splits each slot's
// position range over tps = 4, 2 or 1 adjacent lanes (the most that fit
// SPT * NTH items), so all four waves share the sums; the lanes' interleaved
// partial sums are combined by a fixed xor-shuffle tree.
// LIST: the launch visits a list of blocks (a model mesh's law), else every
// block.  A separate instantiation: the list's lane-varying load in the
// block gather otherwise costs the scatter-free passes their next-block
// prefetch (the compiler's conservative memory waits, measured 17.5 ->
// 20.2 us for pass B).
template <int MAT, unsigned WHAT, class R, class X, int SPT, bool LIST>
__global__ __launch_bounds__(NTH) void elem_kernel(ElemArgs<R, X> A) {
  constexpr bool HAS_G = (WHAT & W_G) != 0;
  constexpr bool HAS_H = (WHAT & W_H) != 0;
  constexpr bool HAS_HP = (WHAT & W_HP) != 0;
  constexpr int NCH = (HAS_G ? 3 : 0) + (HAS_H ? 3 : 0) + (HAS_HP ? 3 : 0);
  constexpr bool SCATTER = NCH > 0;
  constexpr int NCHA = NCH > 0 ? NCH : 1;
  constexpr bool NEEDV = (WHAT & (W_HP | W_HQ | W_MIX | W_MIXA)) != 0;
  constexpr bool RED = (WHAT & (W_E | W_HQ)) != 0;
  constexpr int CS = 4 * NTH;  // channel stride of the contribution buffer
  static_assert(NCH <= MAX_CH, "too many channels");

  extern __shared__ double smem_d[];
  double* red = smem_d;                          // wave partials of the block sums
  X* u_lds = reinterpret_cast<X*>(smem_d + 8);  // 4 per slot (padded)
  R* v_lds = reinterpret_cast<R*>(u_lds + 4 * A.nloc_max);
  R* contrib = v_lds + (NEEDV ? 4 * A.nloc_max : 0);  // [channel group][position][channel in group]
  using PK = Pack<NCHA>;

  const int tid = threadIdx.x;
  const int pid = blockIdx.x;
  const int xcd = pid & 7, k0 = pid >> 3, K = gridDim.x >> 3;
  // XCD x owns the chunks x, x + 8, x + 16, ... of C consecutive blocks
  // (C up to XCD_CHUNK, smaller for small meshes so that every XCD gets
  // work): neighbouring blocks (which share vertices) mostly meet in one L2,
  // and every instance larger than 8 chunks spreads over all XCDs (a batch
  // whose few unconverged instances sat in one XCD's range would idle the
  // others).  The XCD's j-th block is xcd_block(j).
  const int nb = A.nblocks;
  const int C = nb / 64 < 1 ? 1 : (nb / 64 > XCD_CHUNK ? XCD_CHUNK : nb / 64);  // >= 8 chunks per XCD if possible
  const int nchunk = (nb + C - 1) / C;
  const int my_chunks = xcd < nchunk ? (nchunk - 1 - xcd) / 8 + 1 : 0;
  const bool owns_last = my_chunks > 0 && (nchunk - 1) % 8 == xcd;
  const int nb_x = my_chunks * C - (owns_last ? nchunk * C - nb : 0);
  auto xcd_block = [&](int j) {
    const int k = ((j / C) * 8 + xcd) * C + j % C;
    if constexpr (LIST)
      return A.blk_list[k];
    else
      return k;
  };

  // This workgroup's blocks are xcd_block(k0 + j K), j < n_cand; in a solve
  // only those of still-active instances.
  const int n_cand = k0 < nb_x ? (nb_x - k0 + K - 1) / K : 0;
  // the active-block list (solves only) sits after the contribution buffer
  int* act_list = reinterpret_cast<int*>(contrib + (SCATTER ? PK::NCHP * CS : 0));
  // Passes without a scatter phase have registers to spare: each thread
  // loads the next block's staging inputs (its slot tid) while the current
  // block finishes, so that block's staging waits on no global load.  Not in
  // the scatter passes nor with FP64 displacements: there the registers
  // cost a wave per SIMD and the gain disappears (measured).
  constexpr bool PREFETCH = !SCATTER && sizeof(X) == sizeof(R);
  int pf_b = -1;
  X pf_x[3] = {X(0), X(0), X(0)};
  R pf_p[3] = {R(0), R(0), R(0)};
  for_active<NTH>(
      n_cand, A.active != nullptr, act_list, [&](int j) { return xcd_block(k0 + j * K); },
      [&](int bb) { return A.active[A.blk_inst[bb]] != 0; }, [&](const int b, const int bn) {
      const int inst = A.blk_inst[b];
      const int t0 = A.blk_tet_off[b];
      const int nt = A.blk_tet_off[b + 1] - t0;
      const int sb = A.blk_loc_off[b];
      const int nloc = A.blk_loc_off[b + 1] - sb;
      const X alpha = A.alpha != nullptr ? (X)A.alpha[inst] : X(0);
      const int lg = (4 * nloc <= SPT * NTH) ? 2 : ((2 * nloc <= SPT * NTH) ? 1 : 0);
      const int tps = 1 << lg;
      // destinations of this thread's slot partials, loaded early
      int dst_pre[SPT];
      if constexpr (SCATTER) {
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
          const int item = tid + k * NTH;
          const int s = item >> lg;
          dst_pre[k] = (s < nloc && (item & (tps - 1)) == 0) ? at(A.loc_dst + sb, (unsigned)s) : 0;
        }
      }

      // this thread's first tet record, loaded before the staging barrier so the
      // two global-load latencies overlap
      const BlockRes br = block_res<MAT>(A, t0);
      TetRec<R> rec;
      if (tid < nt) load_rec<MAT>(br, (unsigned)tid, rec);

      // stage the block's vertex window (u = x + alpha p for a trial point)
      const int32_t* gids = A.loc_gid + sb;
      for (int s = tid; s < nloc; s += NTH) {
        X ux, uy, uz;
        R px = R(0), py = R(0), pz = R(0);
        if (PREFETCH && s == tid && pf_b == b) {
          ux = pf_x[0];
          uy = pf_x[1];
          uz = pf_x[2];
          px = pf_p[0];
          py = pf_p[1];
          pz = pf_p[2];
        } else {
          const unsigned g3 = 3u * (unsigned)at(gids, (unsigned)s);
          ux = at(A.x, g3);
          uy = at(A.x, g3 + 1);
          uz = at(A.x, g3 + 2);
          if (NEEDV || A.alpha != nullptr) {
            px = at(A.p, g3);
            py = at(A.p, g3 + 1);
            pz = at(A.p, g3 + 2);
          }
        }
        if (A.alpha != nullptr) {
          ux += alpha * (X)px;
          uy += alpha * (X)py;
          uz += alpha * (X)pz;
        }
        Vec4<X> u4 = {{ux, uy, uz, X(0)}};
        *reinterpret_cast<Vec4<X>*>(u_lds + 4 * s) = u4;
        if constexpr (NEEDV) {
          Vec4<R> v4 = {{px, py, pz, R(0)}};
          *reinterpret_cast<Vec4<R>*>(v_lds + 4 * s) = v4;
        }
      }
      __syncthreads();

      R acc[SPT][NCHA];
#pragma unroll
      for (int k = 0; k < SPT; ++k)
#pragma unroll
        for (int ch = 0; ch < NCHA; ++ch) acc[k][ch] = R(0);
      R e_acc = R(0), q_acc = R(0);
      const uint16_t* co = A.chunk_off + A.blk_co_off[b];
      const int nchunks = (nt + NTH - 1) / NTH;
      for (int c = 0; c < nchunks; ++c) {
        const int tl = c * NTH + tid;
        // this chunk's position ranges of the thread's slot items
        int q0[SPT], q1[SPT];
        if constexpr (SCATTER) {
          const uint16_t* cc = co + c * (nloc + 1);
#pragma unroll
          for (int k = 0; k < SPT; ++k) {
            const int item = tid + k * NTH;
            const int s = item >> lg;
            q0[k] = 0;
            q1[k] = 0;
            if (s < nloc) {
              q0[k] = at(cc, (unsigned)s) + (item & (tps - 1));
              q1[k] = at(cc, (unsigned)s + 1);
            }
          }
        }
        if (tl < nt) {
          if (c > 0) load_rec<MAT>(br, (unsigned)tl, rec);
          const TetGeom<R> g = make_geom(rec.B, rec.dV);
          const int sidx[4] = {(int)(rec.cr.x & 0xffffu), (int)(rec.cr.x >> 16), (int)(rec.cr.y & 0xffffu),
                               (int)(rec.cr.y >> 16)};
          // edge differences u_a - u_0 in the storage precision X, then R: F - I
          // depends only on them, and forming them before rounding keeps F
          // accurate relative to the strain rather than to |u|.
          Vec4<X> uc[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) uc[a] = *reinterpret_cast<const Vec4<X>*>(u_lds + 4 * sidx[a]);
          V3<R> u[4], v[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            u[a] = v3((R)(uc[a].v[0] - uc[0].v[0]), (R)(uc[a].v[1] - uc[0].v[1]), (R)(uc[a].v[2] - uc[0].v[2]));
            if constexpr (NEEDV) {
              const Vec4<R> vv = *reinterpret_cast<const Vec4<R>*>(v_lds + 4 * sidx[a]);
              v[a] = v3(vv.v[0], vv.v[1], vv.v[2]);
            } else {
              v[a] = v3(R(0), R(0), R(0));
            }
          }
          TetOut<R> o;
          tet_eval<MAT, WHAT, R>(g, rec.mu, rec.lam, rec.act, u, v, o);
          if constexpr ((WHAT & W_E) != 0) e_acc += o.e;
          if constexpr ((WHAT & W_HQ) != 0) q_acc += o.hq;
          if constexpr ((WHAT & W_MIX) != 0) {
            R* oa = A.out_a + t0;
            R* ob = A.out_b + t0;
            if (A.accumulate) {
              at(oa, (unsigned)tl) += o.dmu;
              if (MAT != MAT_ARAP) at(ob, (unsigned)tl) += o.dlam;
            } else {
              at(oa, (unsigned)tl) = o.dmu;
              if (MAT != MAT_ARAP) at(ob, (unsigned)tl) = o.dlam;
            }
          }
          if constexpr ((WHAT & W_MIXA) != 0) {
            R* oc = A.out_c + 6 * (size_t)t0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
              if (A.accumulate)
                at(oc, 6u * (unsigned)tl + (unsigned)k) += o.dact[k];
              else
                at(oc, 6u * (unsigned)tl + (unsigned)k) = o.dact[k];
            }
          }
          if constexpr (SCATTER) {
            const int pos[4] = {(int)(rec.cr.z & 0xffffu), (int)(rec.cr.z >> 16), (int)(rec.cr.w & 0xffffu),
                                (int)(rec.cr.w >> 16)};
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              R vals[NCHA];
              int ch = 0;
              if constexpr (HAS_G) {
                vals[ch + 0] = o.g[a].x;
                vals[ch + 1] = o.g[a].y;
                vals[ch + 2] = o.g[a].z;
                ch += 3;
              }
              if constexpr (HAS_H) {
                vals[ch + 0] = o.h[a].x;
                vals[ch + 1] = o.h[a].y;
                vals[ch + 2] = o.h[a].z;
                ch += 3;
              }
              if constexpr (HAS_HP) {
                vals[ch + 0] = o.hp[a].x;
                vals[ch + 1] = o.hp[a].y;
                vals[ch + 2] = o.hp[a].z;
              }
#pragma unroll
              for (int gi = 0; gi < PK::G4; ++gi) {
                const Vec4<R> t = {{vals[4 * gi], vals[4 * gi + 1], vals[4 * gi + 2], vals[4 * gi + 3]}};
                *reinterpret_cast<Vec4<R>*>(contrib + gi * 4 * CS + 4 * pos[a]) = t;
              }
              R* rem = contrib + PK::G4 * 4 * CS;
              constexpr int c0 = 4 * PK::G4;
              if constexpr (PK::REM == 3) {
                const Vec4<R> t = {{vals[c0], vals[c0 + 1], vals[c0 + 2], R(0)}};
                *reinterpret_cast<Vec4<R>*>(rem + 4 * pos[a]) = t;
              } else if constexpr (PK::REM == 2) {
                const Vec2<R> t = {{vals[c0], vals[c0 + 1]}};
                *reinterpret_cast<Vec2<R>*>(rem + 2 * pos[a]) = t;
              } else if constexpr (PK::REM == 1) {
                rem[pos[a]] = vals[c0];
              }
            }
          }
        }
        if constexpr (SCATTER) {
          __syncthreads();
#pragma unroll
          for (int k = 0; k < SPT; ++k) {
            // one position's channels: a 4-channel group per 16/32-byte load, then the remainder
            auto load_pos = [&](int q, R (&t)[PK::NCHP]) {
#pragma unroll
              for (int gi = 0; gi < PK::G4; ++gi) {
                const Vec4<R> w = *reinterpret_cast<const Vec4<R>*>(contrib + gi * 4 * CS + 4 * q);
#pragma unroll
                for (int i = 0; i < 4; ++i) t[4 * gi + i] = w.v[i];
              }
              const R* rem = contrib + PK::G4 * 4 * CS;
              constexpr int c0 = 4 * PK::G4;
              if constexpr (PK::REM == 3) {
                const Vec4<R> w = *reinterpret_cast<const Vec4<R>*>(rem + 4 * q);
#pragma unroll
                for (int i = 0; i < 4; ++i) t[c0 + i] = w.v[i];
              } else if constexpr (PK::REM == 2) {
                const Vec2<R> w = *reinterpret_cast<const Vec2<R>*>(rem + 2 * q);
                t[c0] = w.v[0];
                t[c0 + 1] = w.v[1];
              } else if constexpr (PK::REM == 1) {
                t[c0] = rem[q];
              }
            };
            int q = q0[k];
            // two positions per step, both loads first; the sums keep the sequential order
            for (; q + tps < q1[k]; q += 2 * tps) {
              R ta[PK::NCHP], tb[PK::NCHP];
              load_pos(q, ta);
              load_pos(q + tps, tb);
#pragma unroll
              for (int ch = 0; ch < NCH; ++ch) acc[k][ch] = (acc[k][ch] + ta[ch]) + tb[ch];
            }
            if (q < q1[k]) {
              R ta[PK::NCHP];
              load_pos(q, ta);
#pragma unroll
              for (int ch = 0; ch < NCH; ++ch) acc[k][ch] += ta[ch];
            }
          }
          __syncthreads();
        }
      }

      if constexpr (PREFETCH) {
        if (bn >= 0) {
          const int sb_n = A.blk_loc_off[bn];
          const int nloc_n = A.blk_loc_off[bn + 1] - sb_n;
          pf_b = bn;
          if (tid < nloc_n) {
            const unsigned g3 = 3u * (unsigned)at(A.loc_gid + sb_n, (unsigned)tid);
            pf_x[0] = at(A.x, g3);
            pf_x[1] = at(A.x, g3 + 1);
            pf_x[2] = at(A.x, g3 + 2);
            if (NEEDV || A.alpha != nullptr) {
              pf_p[0] = at(A.p, g3);
              pf_p[1] = at(A.p, g3 + 1);
              pf_p[2] = at(A.p, g3 + 2);
            }
          }
        }
      }
      if constexpr (SCATTER) {
        // combine the lanes of a slot, then one partial per slot, at the
        // slot's place in the vertex-major order
        if (lg > 0) {
#pragma unroll
          for (int k = 0; k < SPT; ++k)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) acc[k][ch] += dpp<R, DPP_QUAD_XOR1>(acc[k][ch]);
        }
        if (lg > 1) {
#pragma unroll
          for (int k = 0; k < SPT; ++k)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) acc[k][ch] += dpp<R, DPP_QUAD_XOR2>(acc[k][ch]);
        }
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
          const int item = tid + k * NTH;
          const int s = item >> lg;
          if (s < nloc && (item & (tps - 1)) == 0) {
            const unsigned j = (unsigned)dst_pre[k] * NCH;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) at(A.partial, j + ch) = acc[k][ch];
          }
        }
      }
      if constexpr (RED) {
        // one FP64 partial per wave and column into LDS; thread 0 adds them
        // in a fixed order after one barrier (the block's closing barrier
        // keeps red from being overwritten before it has read it)
        const int lane = tid & 63, w = tid >> 6;
        if constexpr ((WHAT & W_E) != 0) {
          const double e = wave_sum((double)e_acc);
          if (lane == 63) red[2 * w] = e;
        }
        if constexpr ((WHAT & W_HQ) != 0) {
          const double q = wave_sum((double)q_acc);
          if (lane == 63) red[2 * w + 1] = q;
        }
        __syncthreads();
        if (tid == 0) {
#pragma unroll
          for (int col = 0; col < 2; ++col) {
            if (col == 0 && (WHAT & W_E) == 0) continue;
            if (col == 1 && (WHAT & W_HQ) == 0) continue;
            double r = 0.0;
#pragma unroll
            for (int ww = 0; ww < NTH / 64; ++ww) r += red[2 * ww + col];
            A.blk_red[(size_t)RED_STRIDE * b + col] = r;
          }
        }
      }
      __syncthreads();
      });
}

// Finish pass of an element pass, one launch: workgroups [0, ninst) sum one
// instance's block energies / quadratic forms in a fixed order (first, so
// their longer loops start early); workgroups [ninst, ninst + nvblk) form
// the vertex sums of the slot partials over an instance-aligned vertex
// block, one degree of freedom per thread and stride (ascending block order
// per vertex).
template <class R>
struct FinishArgs {
  const int32_t* vblk_off;
  const int32_t* vblk_inst;
  const int32_t* vtx_slot_off;
  const R* partial;
  const int32_t* active;
  R* out_a;
  R* out_b;
  const double* blk_red;
  const int64_t* inst_blk_off;
  double* out_e;
  double* out_q;
  int ninst;  // instance workgroups (0: no energies / quadratic forms)
  int h_only;
  int accumulate;
};

template <class R, int NCH>
__global__ __launch_bounds__(NTH) void finish_kernel(FinishArgs<R> A) {
  __shared__ double red[2 * (NTH / 64 + 1)];
  if ((int)blockIdx.x < A.ninst) {
    const int i = (int)blockIdx.x;
    if (A.active != nullptr && A.active[i] == 0) return;
    const long long b0 = A.inst_blk_off[i];
    double eq[2];
    flat_column_sums<2, NTH>(A.blk_red + RED_STRIDE * b0, RED_STRIDE * (A.inst_blk_off[i + 1] - b0), red, eq);
    if (threadIdx.x == 0) {
      if (A.out_e != nullptr) A.out_e[i] = A.accumulate ? A.out_e[i] + eq[0] : eq[0];
      if (A.out_q != nullptr) A.out_q[i] = A.accumulate ? A.out_q[i] + eq[1] : eq[1];
    }
    return;
  }
  if constexpr (NCH > 0) {
    const int vb = (int)blockIdx.x - A.ninst;
    if (A.active != nullptr && A.active[A.vblk_inst[vb]] == 0) return;
    __shared__ R win[VSUM_LDS_BYTES / sizeof(R)];
    const int v0 = A.vblk_off[vb], v1 = A.vblk_off[vb + 1];
    const int i0 = 3 * v0, i1 = 3 * v1;
    int k0[DOF_PER_THREAD], k1[DOF_PER_THREAD];
#pragma unroll
    for (int j = 0; j < DOF_PER_THREAD; ++j) {
      const int i = i0 + (int)threadIdx.x + j * NTH;
      k0[j] = k1[j] = 0;
      if (i < i1) {
        k0[j] = at(A.vtx_slot_off, (unsigned)(i / 3));
        k1[j] = at(A.vtx_slot_off, (unsigned)(i / 3) + 1);
      }
    }
    VertexSums<R, NCH, NTH> vs{A.vtx_slot_off, A.partial, win};
    vs.stage(v0, v1, VSUM_LDS_BYTES / (int)sizeof(R));
#pragma unroll
    for (int j = 0; j < DOF_PER_THREAD; ++j) {
      const int i = i0 + (int)threadIdx.x + j * NTH;
      if (i >= i1) continue;
      R acc[NCH / 3];
      vs.sum_dof(k0[j], k1[j], i - 3 * (i / 3), acc);
#pragma unroll
      for (int grp = 0; grp < NCH / 3; ++grp) {
        R* out = A.h_only ? A.out_b : (grp == 0 ? A.out_a : A.out_b);
        if (A.accumulate)
          at(out, (unsigned)i) += acc[grp];
        else
          at(out, (unsigned)i) = acc[grp];
      }
    }
  }
}

// ------------------------------------------------------------------ launchers

template <class R, class X>
static ElemArgs<R, X> make_args(apl_mesh* m, const void* mu, const void* lam, const void* x, const void* p,
                                const double* alpha, const int32_t* active, void* out_a, void* out_b,
                                int accumulate, void* out_c, const Workspace* ws) {
  ElemArgs<R, X> A;
  A.dminv = (const R*)m->d.dminv;
  A.dV = (const R*)m->d.dV;
  A.corner = (const uint4*)m->d.corner;
  A.blk_tet_off = m->d.blk_tet_off;
  A.blk_loc_off = m->d.blk_loc_off;
  A.blk_inst = m->d.blk_inst;
  A.loc_gid = m->d.loc_gid;
  A.loc_dst = m->d.loc_dst;
  A.blk_co_off = m->d.blk_co_off;
  A.chunk_off = m->d.chunk_off;
  A.mu = (const R*)mu;
  A.lam = (const R*)lam;
  A.act = (const R*)m->activation;
  A.x = (const X*)x;
  A.p = (const R*)p;
  A.alpha = alpha;
  A.active = active;
  A.out_a = (R*)out_a;
  A.out_b = (R*)out_b;
  A.out_c = (R*)out_c;
  A.partial = (R*)(ws != nullptr ? ws->partial : m->d.partial);
  A.blk_red = ws != nullptr ? ws->blk_red : m->d.blk_red;
  A.T = (int)m->L.T;
  A.blk_list = nullptr;
  A.nblocks = (int)m->L.nblocks;
  A.nloc_max = m->L.nloc_max;
  A.chunks = m->L.chunks;
  A.accumulate = accumulate;
  return A;
}

// LDS of one element workgroup: wave partials, the vertex window (u in X,
// v in R when the pass needs it, 4 lanes per slot), the packed contribution
// buffer of the scatter passes and, in solves, the active-block list.
template <unsigned WHAT, class R, class X>
constexpr size_t smem_bytes_n(int nloc_max, bool filtered) {
  constexpr bool NEEDV = (WHAT & (W_HP | W_HQ | W_MIX | W_MIXA)) != 0;
  constexpr int NCH = ((WHAT & W_G) ? 3 : 0) + ((WHAT & W_H) ? 3 : 0) + ((WHAT & W_HP) ? 3 : 0);
  size_t n = 8 * sizeof(double) + 4 * (size_t)nloc_max * (sizeof(X) + (NEEDV ? sizeof(R) : 0));
  n += sizeof(R) * (size_t)(NCH > 0 ? Pack<NCH>::NCHP : 0) * 4 * NTH;
  if (filtered) n += sizeof(int) * NTH;  // active-block list
  return n;
}

template <int MAT, unsigned WHAT, class R, class X>
static size_t smem_bytes(const apl_mesh* m, bool filtered) {
  return smem_bytes_n<WHAT, R, X>(m->L.nloc_max, filtered);
}

// The largest block window (distinct vertices per block) every pass variant
// of a mesh of this dtype can stage within ELEM_LDS_MAX, filtered launches
// included: 1024 for FP32 meshes (also with the solver's FP64 displacements),
// 478 for FP64 meshes (pass A: 6 FP64 channels x 1024 positions + 32 B per slot).
template <class R, class X>
static bool fits(int n) {
  return smem_bytes_n<W_E | W_G | W_H, R, X>(n, true) <= ELEM_LDS_MAX &&
         smem_bytes_n<W_HP, R, R>(n, true) <= ELEM_LDS_MAX && smem_bytes_n<W_HQ, R, X>(n, true) <= ELEM_LDS_MAX &&
         smem_bytes_n<W_MIX | W_MIXA, R, R>(n, true) <= ELEM_LDS_MAX && smem_bytes_n<W_G, R, R>(n, true) <= ELEM_LDS_MAX &&
         smem_bytes_n<W_H, R, R>(n, true) <= ELEM_LDS_MAX;
}

int nloc_limit(int dtype) {
  int n = 1024;
  if (dtype == DT_F64) {
    while (n > 4 && !fits<double, double>(n)) --n;
  } else {
    while (n > 4 && !(fits<float, float>(n) && fits<float, double>(n))) --n;
  }
  return n;
}

// blocks one launch of law MAT visits (model meshes: the law's own blocks)
inline int64_t launch_blocks(const apl_mesh* m, int mat) { return m->model ? m->law_nblk[mat] : m->L.nblocks; }

// Default element grid per pass, percent of occupancy x CUs: one resident
// round of persistent workgroups, except for the VALU-heavy ARAP gradient
// passes (4 waves per SIMD) and the Hessian-vector pass, which run faster
// with a second, non-resident round that the hardware dispatches as
// workgroups of the first finish (a dynamic balance of the last blocks;
// measured on the 64^3 cube: ARAP pass A 58.6 -> 55.1 us, SNH H v 30.3 ->
// 27.4 us, while SNH pass A and the quadratic-form passes are fastest at 100).
constexpr int default_grid_pct(int mat, unsigned what) {
  return ((what & W_HP) != 0 || (mat == MAT_ARAP && (what & (W_G | W_H)) != 0)) ? 200 : 100;
}

template <int MAT, unsigned WHAT, class R, class X, int SPT>
static int grid_for_spt(apl_mesh* m, bool filtered) {
  static_assert(WHAT < 128, "pass variant index");
  int& cached = m->grid_cache[MAT][WHAT + (sizeof(X) != sizeof(R) ? 128 : 0) + (filtered ? 256 : 0)];
  if (cached > 0) return cached;
  const size_t smem = smem_bytes<MAT, WHAT, R, X>(m, filtered);
  int occ = 0;
  const hipError_t oe =
      m->model ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, elem_kernel<MAT, WHAT, R, X, SPT, true>, NTH, smem)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, elem_kernel<MAT, WHAT, R, X, SPT, false>, NTH, smem);
  if (oe != hipSuccess || occ < 1)
    occ = 1;
  const int pct = m->grid_pct > 0 ? m->grid_pct : default_grid_pct(MAT, WHAT);
  long long g = (long long)occ * m->n_cu * pct / 100;
  long long nb8 = ((launch_blocks(m, MAT) + 7) / 8) * 8;
  if (g > nb8) g = nb8;
  g = ((g + 7) / 8) * 8;
  if (g < 8) g = 8;
  cached = (int)g;
  return cached;
}

// slots per thread for this mesh's largest block window
inline int spt_for(const apl_mesh* m) { return m->L.nloc_max <= NTH ? 1 : (m->L.nloc_max <= 2 * NTH ? 2 : 4); }

template <int MAT, unsigned WHAT, class R, class X>
static int grid_for(apl_mesh* m, bool filtered) {
  constexpr bool SCATTER = (WHAT & (W_G | W_H | W_HP)) != 0;
  if (!SCATTER) return grid_for_spt<MAT, WHAT, R, X, 1>(m, filtered);
  const int spt = spt_for(m);
  if (spt == 1) return grid_for_spt<MAT, WHAT, R, X, 1>(m, filtered);
  if (spt == 2) return grid_for_spt<MAT, WHAT, R, X, 2>(m, filtered);
  return grid_for_spt<MAT, WHAT, R, X, 4>(m, filtered);
}

template <int MAT, unsigned WHAT, class R, class X>
static int launch_one(apl_mesh* m, const void* mu, const void* lam, const void* x, const void* p,
                      const double* alpha, const int32_t* active, void* out_a, void* out_b,
                      int accumulate, int grid_override, hipStream_t s, void* out_c = nullptr,
                      const Workspace* ws = nullptr) {
  if (launch_blocks(m, MAT) == 0) return 0;
  const bool filtered = active != nullptr;
  const size_t smem = smem_bytes<MAT, WHAT, R, X>(m, filtered);
  if (smem > ELEM_LDS_MAX) {
    set_error("element block needs more than 64 KiB of LDS (block window above the mesh's nloc cap)");
    return -1;
  }
  const int grid = grid_override > 0 ? ((grid_override + 7) / 8) * 8 : grid_for<MAT, WHAT, R, X>(m, filtered);
  ElemArgs<R, X> A = make_args<R, X>(m, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, out_c, ws);
  if (m->model) {
    A.blk_list = m->d.law_blk[MAT];
    A.nblocks = (int)m->law_nblk[MAT];
  }
  constexpr bool SCATTER = (WHAT & (W_G | W_H | W_HP)) != 0;
  const int spt = SCATTER ? spt_for(m) : 1;
  auto go = [&](auto list_c) {
    constexpr bool LIST = decltype(list_c)::value;
    if (spt == 1)
      hipLaunchKernelGGL((elem_kernel<MAT, WHAT, R, X, 1, LIST>), dim3(grid), dim3(NTH), smem, s, A);
    else if (spt == 2)
      hipLaunchKernelGGL((elem_kernel<MAT, WHAT, R, X, SCATTER ? 2 : 1, LIST>), dim3(grid), dim3(NTH), smem, s, A);
    else
      hipLaunchKernelGGL((elem_kernel<MAT, WHAT, R, X, SCATTER ? 4 : 1, LIST>), dim3(grid), dim3(NTH), smem, s, A);
  };
  if (m->model)
    go(std::integral_constant<bool, true>{});
  else
    go(std::integral_constant<bool, false>{});
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
                         void* out_b, int accumulate, int grid_override, int x_f64, hipStream_t s, void* out_c,
                         const Workspace* ws) {
  if constexpr (MAT == MAT_MUSCLE) {
    // activation derivatives, alone or with the mu / lambda ones
    if (what == W_MIXA)
      return launch_one<MAT, W_MIXA, R, R>(m, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s, out_c);
    if (what == (W_MIX | W_MIXA))
      return launch_one<MAT, W_MIX | W_MIXA, R, R>(m, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s, out_c);
  }
  if (x_f64 && sizeof(R) == 4) {
    // solver passes over FP64 displacements
    if (what == (W_E | W_G | W_H))
      return launch_one<MAT, W_E | W_G | W_H, R, double>(m, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s);
    if (what == W_HQ)
      return launch_one<MAT, W_HQ, R, double>(m, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s);
    set_error("FP64 displacements are supported for the solver passes only");
    return -1;
  }
#define APL_CASE(W) \
  if (what == (W)) return launch_one<MAT, (W), R, R>(m, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s, nullptr, ws);
  APL_WHAT_LIST(APL_CASE)
#undef APL_CASE
  set_error("unsupported combination of outputs for one element pass");
  return -1;
}

int launch_elem(apl_mesh* m, int material, unsigned what, const void* mu, const void* lam, const void* x,
                const void* p, const double* alpha, const int32_t* active, void* out_a, void* out_b,
                int accumulate, int grid_override, hipStream_t s, int x_f64, void* out_c, const Workspace* ws) {
  auto go = [&](auto mat_c) -> int {
    constexpr int MAT = decltype(mat_c)::value;
    if (m->dtype == DT_F32)
      return dispatch_what<MAT, float>(m, what, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, x_f64, s, out_c, ws);
    return dispatch_what<MAT, double>(m, what, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, 0, s, out_c, ws);
  };
  if (material == MAT_MODEL) {
    // a model mesh: every law present, each over its own blocks (the muscle
    // law's activation derivatives only exist for that law)
    if (!m->model) {
      set_error("APL_MODEL needs a mesh created by apl_model_create");
      return -1;
    }
    for (int law = 0; law < N_MAT; ++law) {
      if (m->law_nblk[law] == 0) continue;
      unsigned w = what;
      if (law != MAT_MUSCLE) w &= ~W_MIXA;
      if (w == 0) continue;
      if (launch_elem(m, law, w, mu, lam, x, p, alpha, active, out_a, out_b, accumulate, grid_override, s, x_f64,
                      out_c, ws))
        return -1;
    }
    return 0;
  }
  if (material == MAT_SNH) return go(std::integral_constant<int, MAT_SNH>{});
  if (material == MAT_ARAP) return go(std::integral_constant<int, MAT_ARAP>{});
  if (material == MAT_MUSCLE) {
    if (m->activation == nullptr && launch_blocks(m, MAT_MUSCLE) > 0) {
      set_error("the muscle law needs the activation: call apl_mesh_set_activation first");
      return -1;
    }
    return go(std::integral_constant<int, MAT_MUSCLE>{});
  }
  set_error("unknown material");
  return -1;
}

template <int MAT>
static int grid_mat(apl_mesh* m, unsigned what, int x_f64, bool filtered) {
  if (x_f64 && m->dtype == DT_F32) {
    if (what == (W_E | W_G | W_H)) return grid_for<MAT, W_E | W_G | W_H, float, double>(m, filtered);
    if (what == W_HQ) return grid_for<MAT, W_HQ, float, double>(m, filtered);
    return 0;
  }
#define APL_CASE(W) \
  if (what == (W)) \
    return m->dtype == DT_F32 ? grid_for<MAT, (W), float, float>(m, filtered) : grid_for<MAT, (W), double, double>(m, filtered);
  APL_WHAT_LIST(APL_CASE)
#undef APL_CASE
  return 0;
}

int elem_grid(apl_mesh* m, int material, unsigned what, int x_f64, bool filtered) {
  if (material == MAT_MODEL) {
    int g = 0;
    for (int law = 0; law < N_MAT; ++law)
      if (m->law_nblk[law] > 0) g = std::max(g, elem_grid(m, law, what, x_f64, filtered));
    return g;
  }
  if (material == MAT_MUSCLE) return grid_mat<MAT_MUSCLE>(m, what, x_f64, filtered);
  return material == MAT_SNH ? grid_mat<MAT_SNH>(m, what, x_f64, filtered)
                             : grid_mat<MAT_ARAP>(m, what, x_f64, filtered);
}

int launch_finish(apl_mesh* m, unsigned what, void* out_a, void* out_b, double* out_e, double* out_q,
                  int accumulate, const int32_t* active, hipStream_t s, const Workspace* ws) {
  const bool g = what & W_G, h = what & W_H, hp = what & W_HP;
  const int nch = (g ? 3 : 0) + (h ? 3 : 0) + (hp ? 3 : 0);
  const bool inst = out_e != nullptr || out_q != nullptr;
  const int nvblk = nch > 0 ? (int)m->L.vblk_inst.size() : 0;
  const int ninst = inst ? (int)m->L.n_inst : 0;
  const int grid = nvblk + ninst;
  if (grid == 0) return 0;
  auto go = [&](auto* tag, auto nch_c) {
    using R = std::remove_pointer_t<decltype(tag)>;
    FinishArgs<R> A;
    A.vblk_off = m->d.vblk_off;
    A.vblk_inst = m->d.vblk_inst;
    A.vtx_slot_off = m->d.vtx_slot_off;
    A.partial = (const R*)(ws != nullptr ? ws->partial : m->d.partial);
    A.active = active;
    A.out_a = (R*)out_a;
    A.out_b = (R*)out_b;
    A.blk_red = ws != nullptr ? ws->blk_red : m->d.blk_red;
    A.inst_blk_off = m->d.inst_blk_off;
    A.out_e = out_e;
    A.out_q = out_q;
    A.ninst = ninst;
    A.h_only = (h && !g && !hp) ? 1 : 0;
    A.accumulate = accumulate;
    hipLaunchKernelGGL((finish_kernel<R, decltype(nch_c)::value>), dim3((unsigned)grid), dim3(NTH), 0, s, A);
  };
  using I0 = std::integral_constant<int, 0>;
  using I3 = std::integral_constant<int, 3>;
  using I6 = std::integral_constant<int, 6>;
  if (m->dtype == DT_F32) {
    float* t = nullptr;
    if (nch == 6) go(t, I6{}); else if (nch == 3) go(t, I3{}); else go(t, I0{});
  } else {
    double* t = nullptr;
    if (nch == 6) go(t, I6{}); else if (nch == 3) go(t, I3{}); else go(t, I0{});
  }
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace apl

// ------------------------------------------------------------ allocation-free entry point
//
// apl_elem_eval_static: apl_elem_eval with the element passes' scratch (slot
// partials and block sums) in a file-scope device array instead of the
// mesh's workspace.  Each GPU the library runs on holds its own copy of the
// array until the process ends; its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and only that
// address is handed to kernels.  Every call on one GPU shares the array, so
// calls that overlap on one GPU race.  Nothing in the array needs to start
// at zero: a pass writes every partial and block sum it later reads.

namespace apl {

constexpr size_t STATIC_SCRATCH_BYTES = size_t(32) << 20;  // 32 MiB per GPU
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MAX_DEVICES] = {};

// device address of this GPU's scratch array (m->device is current)
static int static_scratch(int device, void** out) {
  if (device < 0 || device >= MAX_DEVICES) {
    set_error("device index out of range for the static scratch table");
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      (void)hipGetLastError();
      return -1;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

}  // namespace apl

extern "C" size_t apl_static_scratch_bytes(void) { return apl::STATIC_SCRATCH_BYTES; }

extern "C" int apl_elem_eval_static(apl_mesh* m, int material, const void* mu, const void* lam, const void* u,
                                    const void* p, unsigned what, double* energy, void* grad, void* hess_diag,
                                    void* hess_prod, double* hess_quad, void* dmu, void* dlam, void* stream) {
  using namespace apl;
  if (m == nullptr) {
    set_error("mesh is NULL");
    return -1;
  }
  const size_t need = eval_scratch_bytes(m);
  if (need > STATIC_SCRATCH_BYTES) {
    set_error("mesh needs " + std::to_string(need) + " bytes of element-pass scratch; apl_elem_eval_static holds " +
              std::to_string(STATIC_SCRATCH_BYTES) + " (use apl_elem_eval)");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(m->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  void* base = nullptr;
  if (static_scratch(m->device, &base)) return -1;
  const size_t part = (size_t)MAX_CH * (size_t)std::max<int64_t>(m->L.n_slots, 1) * real_size(m);
  Workspace ws;
  ws.partial = base;
  ws.blk_red = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + ((part + 255) / 256) * 256);
  return elem_eval_impl(m, material, mu, lam, u, p, what, energy, grad, hess_diag, hess_prod, hess_quad, dmu, dlam,
                        (hipStream_t)stream, &ws);
}
