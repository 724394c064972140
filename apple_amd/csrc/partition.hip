// Block partition on the device: the Layout that build_layout (layout.cpp)
// computes on the host, built by kernels from the cells.  Same arrays, bit
// for bit (tests compare every one).  layout.h describes the structures.
//
//   check_cells_kernel   every corner inside its instance's vertex range
//   greedy_blocks_kernel block ends where the distinct-vertex cap binds
//                        (one workgroup per instance, windows in sequence)
//   block_slots_kernel   first-touch slots of each block (LDS hash of the
//                        block's corners, minimal corner index per vertex)
//   chunk_rank_kernel    per chunk: slot offsets and the (slot, tet, corner)
//                        rank of every corner; the block's slot -> vertex list
//   vertex_count_kernel  slots per vertex; a scan gives the vertex CSR
//   loc_dst              stable radix sort of (vertex, slot) pairs, inverted
//
// The host keeps the bookkeeping whose size is the number of blocks,
// instances or segments (offsets, prefix sums over blocks); nothing on the
// host touches the cells, the slots or the vertices.  Also here: the Morton
// order of a mesh (apl_morton_order) and the lumped vertex volumes
// (apl_mesh_lumped_volume), the other setup steps of a mesh on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <vector>

#include "../../include/apple_amd.h"
#include "internal.h"

namespace apl {

namespace {

constexpr int PNT = 256;  // threads of the partition workgroups
constexpr int MAXK = 4;   // tets per thread of a greedy window (max_tets <= 4 PNT)

// open-addressing table of a window's / block's vertices in LDS: tv holds the
// vertex (-1 empty), ti the minimal corner index at which it occurs
__device__ inline unsigned ht_hash(int v, unsigned ts) { return ((unsigned)v * 2654435761u) % ts; }

__device__ inline void ht_insert(int* tv, int* ti, unsigned ts, int v, int i) {
  unsigned h = ht_hash(v, ts);
  for (;;) {
    const int old = atomicCAS(&tv[h], -1, v);
    if (old == -1 || old == v) {
      atomicMin(&ti[h], i);
      return;
    }
    h = h + 1 == ts ? 0u : h + 1;
  }
}

// index of vertex v (present) in the table
__device__ inline unsigned ht_find(const int* tv, unsigned ts, int v) {
  unsigned h = ht_hash(v, ts);
  while (tv[h] != v) h = h + 1 == ts ? 0u : h + 1;
  return h;
}

// table entries for n corners: at least n + 1 (probing ends), 1.5 n
__host__ __device__ inline unsigned ht_size(int n) { return (unsigned)(n + n / 2 + 1); }

// exclusive scan of one value per thread over the workgroup (ascending
// thread order); returns the thread's offset, *total the sum.  A shuffle
// scan per wave, then the wave totals through LDS (sm: one int per wave).
__device__ inline int wg_exclusive_scan(int x, int* sm, int* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) sm[w] = incl;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < PNT / 64; ++k) {
    const int v = sm[k];
    off += k < w ? v : 0;
    tot += v;
  }
  *total = tot;
  __syncthreads();  // sm is free for the next scan
  return off + incl - x;
}

// instance of tet t: the last i with inst_tet_off[i] <= t (binary search)
__device__ inline int instance_of(const int64_t* off, int n_inst, long long t) {
  int lo = 0, hi = n_inst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(PNT) void check_cells_kernel(const int32_t* __restrict__ cells, long long T,
                                                          const int64_t* __restrict__ inst_tet_off,
                                                          const int64_t* __restrict__ inst_vtx_off, int n_inst,
                                                          unsigned long long* bad) {
  for (long long t = (long long)blockIdx.x * PNT + threadIdx.x; t < T; t += (long long)gridDim.x * PNT) {
    const int i = instance_of(inst_tet_off, n_inst, t);
    const long long lo = inst_vtx_off[i], hi = inst_vtx_off[i + 1];
    for (int c = 0; c < 4; ++c) {
      const long long v = cells[4 * t + c];
      if (v < lo || v >= hi) atomicMin(bad, (unsigned long long)t);
    }
  }
}

// Greedy blocks of one instance (one workgroup each), segment by segment:
// from tet t, the window of max_tets tets (or up to the segment end) is
// hashed; D(k), the distinct vertices of tets t..t+k, is the prefix count of
// first occurrences; the block takes every k with D(k) <= cap.  This is the
// host's tet-by-tet greedy (a tet joins while the slot count stays within
// the cap), evaluated a window at a time.
__global__ __launch_bounds__(PNT) void greedy_blocks_kernel(const int32_t* __restrict__ cells,
                                                            const int64_t* __restrict__ seg_off,
                                                            const int64_t* __restrict__ inst_seg_off,
                                                            const int64_t* __restrict__ inst_tet_off, int max_tets,
                                                            int cap, int32_t* __restrict__ ends,
                                                            int64_t* __restrict__ counts) {
  extern __shared__ int sm_g[];
  __shared__ int scan_sm[PNT];
  __shared__ int bcast;
  const int inst = blockIdx.x;
  const int tid = threadIdx.x;
  const int max_n = 4 * max_tets;
  const unsigned ts_max = ht_size(max_n);
  int* tv = sm_g;
  int* ti = sm_g + ts_max;
  long long nb = 0;
  const long long out0 = inst_tet_off[inst];
  for (long long sg = inst_seg_off[inst]; sg < inst_seg_off[inst + 1]; ++sg) {
    long long t = seg_off[sg];
    const long long tend = seg_off[sg + 1];
    while (t < tend) {
      const int nw = (int)std::min<long long>(max_tets, tend - t);
      const int n = 4 * nw;
      const unsigned ts = ht_size(n);
      // tets [k0, k1) of this thread (at most MAXK), contiguous so the scan
      // runs in tet order; their corners are loaded once, into registers
      const int per = (nw + PNT - 1) / PNT;
      const int k0 = std::min(nw, tid * per), k1 = std::min(nw, k0 + per);
      int vv[4 * MAXK];
#pragma unroll
      for (int j = 0; j < 4 * MAXK; ++j) vv[j] = 4 * (k0 + j / 4) + (j & 3) < 4 * k1 ? cells[4 * (t + k0) + j] : -1;
      for (unsigned h = tid; h < ts; h += PNT) {
        tv[h] = -1;
        ti[h] = INT_MAX;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 4 * MAXK; ++j)
        if (vv[j] >= 0) ht_insert(tv, ti, ts, vv[j], 4 * k0 + j);
      __syncthreads();
      // first occurrences per tet (bit j of the thread's corners)
      unsigned first = 0;
      int mine = 0;
#pragma unroll
      for (int j = 0; j < 4 * MAXK; ++j)
        if (vv[j] >= 0 && ti[ht_find(tv, ts, vv[j])] == 4 * k0 + j) {
          first |= 1u << j;
          ++mine;
        }
      int total = 0;
      int run = wg_exclusive_scan(mine, scan_sm, &total);
      int fits = 0;
      for (int k = k0; k < k1; ++k) {
        run += __popc((first >> (4 * (k - k0))) & 15u);
        fits += run <= cap ? 1 : 0;  // D(k) is non-decreasing: the fitting k form a prefix
      }
      int len = 0;
      (void)wg_exclusive_scan(fits, scan_sm, &len);
      if (tid == 0) {
        ends[out0 + nb] = (int32_t)(t + len);
        bcast = len;
      }
      __syncthreads();
      t += bcast;
      ++nb;
      __syncthreads();
    }
  }
  if (tid == 0) counts[inst] = nb;
}

// First-touch slots of one block: corner i of the block (i = 4 (t - t0) + c)
// opens a slot when no earlier corner of the block has its vertex; slots are
// numbered in corner order.  Writes each corner's slot into the corner
// record, the block's slot count, and its vertices (slot order) into
// gid_tmp at 4 t0 (a block has at most 4 (t1 - t0) slots).
__global__ __launch_bounds__(PNT) void block_slots_kernel(const int32_t* __restrict__ cells,
                                                          const int32_t* __restrict__ blk_tet_off,
                                                          uint16_t* __restrict__ corner, int32_t* __restrict__ nloc,
                                                          int32_t* __restrict__ gid_tmp, int max_tets) {
  extern __shared__ int sm_s[];
  __shared__ int scan_sm[PNT];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const long long t0 = blk_tet_off[b], t1 = blk_tet_off[b + 1];
  const int n = (int)(4 * (t1 - t0));
  const unsigned ts_max = ht_size(4 * max_tets);
  const unsigned ts = ht_size(n);
  int* tv = sm_s;
  int* ti = sm_s + ts_max;
  for (unsigned h = tid; h < ts; h += PNT) {
    tv[h] = -1;
    ti[h] = INT_MAX;
  }
  __syncthreads();
  const int32_t* cb = cells + 4 * t0;
  for (int i = tid; i < n; i += PNT) ht_insert(tv, ti, ts, cb[i], i);
  __syncthreads();
  // first occurrences of this thread's contiguous corners (at most 16 with
  // max_tets <= 1024), counted, then numbered in corner order
  const int per = (n + PNT - 1) / PNT;
  const int i0 = std::min(n, tid * per), i1 = std::min(n, i0 + per);
  unsigned first = 0;
  int mine = 0;
  for (int i = i0; i < i1; ++i)
    if (ti[ht_find(tv, ts, cb[i])] == i) {
      first |= 1u << (i - i0);
      ++mine;
    }
  int total = 0;
  int s = wg_exclusive_scan(mine, scan_sm, &total);
  __syncthreads();
  // the table's index field now carries the slot of each vertex
  for (int i = i0; i < i1; ++i)
    if (first & (1u << (i - i0))) {
      const int v = cb[i];
      ti[ht_find(tv, ts, v)] = s;
      gid_tmp[4 * t0 + s] = v;
      ++s;
    }
  __syncthreads();
  for (int i = tid; i < n; i += PNT) {
    const long long t = t0 + i / 4;
    corner[8 * t + (i & 3)] = (uint16_t)ti[ht_find(tv, ts, cb[i])];
  }
  if (tid == 0) nloc[b] = total;
}

// Per chunk of one block: slot offsets into the corner positions (counts,
// then their exclusive scan) and the positions, corners ranked by (slot,
// tet, corner); also the block's slot -> vertex list at its final place.
__global__ __launch_bounds__(PNT) void chunk_rank_kernel(const int32_t* __restrict__ blk_tet_off,
                                                         const int32_t* __restrict__ blk_loc_off,
                                                         const int32_t* __restrict__ blk_co_off,
                                                         const int32_t* __restrict__ gid_tmp,
                                                         uint16_t* __restrict__ corner,
                                                         uint16_t* __restrict__ chunk_off,
                                                         int32_t* __restrict__ loc_gid, int nth) {
  __shared__ int cnt[1025];
  __shared__ int off[1025];
  __shared__ int bucket[1024];
  __shared__ int scan_sm[PNT];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const long long t0 = blk_tet_off[b], t1 = blk_tet_off[b + 1];
  const int nloc = blk_loc_off[b + 1] - blk_loc_off[b];
  for (int s = tid; s < nloc; s += PNT) loc_gid[blk_loc_off[b] + s] = gid_tmp[4 * t0 + s];
  uint16_t* co = chunk_off + blk_co_off[b];
  int c = 0;
  for (long long c0 = t0; c0 < t1; c0 += nth, ++c) {
    const int nc = (int)(4 * (std::min<long long>(t1, c0 + nth) - c0));
    for (int s = tid; s <= nloc; s += PNT) cnt[s] = 0;
    __syncthreads();
    for (int i = tid; i < nc; i += PNT) {
      const int sl = corner[8 * (c0 + i / 4) + (i & 3)];
      if (sl < nloc) atomicAdd(&cnt[sl], 1);
    }
    __syncthreads();
    // exclusive scan of cnt[0..nloc] (nloc + 1 <= 1025 entries, contiguous per thread)
    const int per = (nloc + 1 + PNT - 1) / PNT;
    const int s0 = std::min(nloc + 1, tid * per), s1 = std::min(nloc + 1, s0 + per);
    int mine = 0;
    for (int s = s0; s < s1; ++s) mine += cnt[s];
    int total = 0;
    int run = wg_exclusive_scan(mine, scan_sm, &total);
    for (int s = s0; s < s1; ++s) {
      off[s] = run;
      run += cnt[s];
    }
    __syncthreads();
    uint16_t* cc = co + (long long)c * (nloc + 1);
    for (int s = tid; s <= nloc; s += PNT) {
      cc[s] = (uint16_t)off[s];
      cnt[s] = off[s];  // cursor
    }
    __syncthreads();
    for (int i = tid; i < nc; i += PNT) {
      const int sl = corner[8 * (c0 + i / 4) + (i & 3)];
      if (sl < nloc) {
        const int p = atomicAdd(&cnt[sl], 1);
        if (p < nc) bucket[p] = i;
      }
    }
    __syncthreads();
    // each slot's corners in corner order (insertion sort of a short run)
    for (int s = tid; s < nloc; s += PNT) {
      const int a = off[s], e = off[s + 1];
      for (int p = a + 1; p < e; ++p) {
        const int x = bucket[p];
        int q = p - 1;
        while (q >= a && bucket[q] > x) {
          bucket[q + 1] = bucket[q];
          --q;
        }
        bucket[q + 1] = x;
      }
    }
    __syncthreads();
    for (int p = tid; p < nc; p += PNT) {
      const int i = bucket[p];
      if ((unsigned)i < (unsigned)nc) corner[8 * (c0 + i / 4) + 4 + (i & 3)] = (uint16_t)p;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(PNT) void vertex_count_kernel(const int32_t* __restrict__ loc_gid, long long n_slots,
                                                           int32_t* __restrict__ cnt) {
  for (long long k = (long long)blockIdx.x * PNT + threadIdx.x; k < n_slots; k += (long long)gridDim.x * PNT)
    atomicAdd(&cnt[loc_gid[k]], 1);
}

__global__ __launch_bounds__(PNT) void shared_count_kernel(const int32_t* __restrict__ cnt, long long V,
                                                           unsigned long long* __restrict__ n_shared) {
  __shared__ int sm[PNT];
  int mine = 0;
  for (long long v = (long long)blockIdx.x * PNT + threadIdx.x; v < V; v += (long long)gridDim.x * PNT)
    mine += cnt[v] > 1 ? 1 : 0;
  int total = 0;
  (void)wg_exclusive_scan(mine, sm, &total);
  if (threadIdx.x == 0 && total > 0) atomicAdd(n_shared, (unsigned long long)total);
}

__global__ __launch_bounds__(PNT) void iota_kernel(int32_t* __restrict__ a, long long n) {
  for (long long k = (long long)blockIdx.x * PNT + threadIdx.x; k < n; k += (long long)gridDim.x * PNT) a[k] = (int32_t)k;
}

__global__ __launch_bounds__(PNT) void invert_kernel(const int32_t* __restrict__ order, long long n,
                                                     int32_t* __restrict__ dst) {
  for (long long j = (long long)blockIdx.x * PNT + threadIdx.x; j < n; j += (long long)gridDim.x * PNT)
    dst[order[j]] = (int32_t)j;
}

unsigned grid_n(long long n) {
  const long long g = (n + PNT - 1) / PNT;
  return (unsigned)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

int bits_for(long long n) {
  int b = 1;
  while (b < 62 && (1LL << b) < n) ++b;
  return b;
}

// device allocations of one build, freed together unless kept for the mesh
struct Arena {
  std::vector<void*> tmp;
  hipError_t err = hipSuccess;
  template <class T>
  T* get(size_t n) {
    void* p = nullptr;
    if (err == hipSuccess) err = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
    if (err == hipSuccess) tmp.push_back(p);
    return (T*)p;
  }
  ~Arena() {
    for (void* p : tmp) drop_hip_error(hipFree(p));
  }
};

}  // namespace

bool device_layout_supported(int chunks, int nloc_cap) { return chunks >= 1 && chunks <= 4 && nloc_cap <= 1024; }

#define PART_CHECK(expr)                                                        \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      (void)hipGetLastError();                                                  \
      return std::string(#expr) + ": " + hipGetErrorString(e_);                 \
    }                                                                           \
  } while (0)

std::string build_layout_device(apl_mesh* m, const int32_t* cells, int64_t T, int64_t V, int64_t n_inst,
                                const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int chunks, int nloc_cap,
                                int64_t n_seg_total, const int64_t* seg_off) {
  char msg[256];
  const int nth = NTH;
  if (!device_layout_supported(chunks, nloc_cap)) return "the device partition takes 1..4 chunks";
  if (inst_tet_off[0] != 0 || inst_tet_off[n_inst] != T) return "instance tet offsets must span [0, T]";
  if (inst_vtx_off[0] != 0 || inst_vtx_off[n_inst] != V) return "instance vertex offsets must span [0, V]";
  for (int64_t i = 0; i < n_inst; ++i)
    if (inst_tet_off[i + 1] < inst_tet_off[i] || inst_vtx_off[i + 1] < inst_vtx_off[i])
      return "instance offsets must be non-decreasing";
  // segments: the model's, else one per instance; every instance boundary is one
  std::vector<int64_t> segs;
  std::vector<int64_t> inst_seg(n_inst + 1, 0);
  if (seg_off != nullptr) {
    if (n_seg_total < n_inst || seg_off[0] != 0 || seg_off[n_seg_total] != T) return "segment offsets must span [0, T]";
    for (int64_t k = 0; k < n_seg_total; ++k)
      if (seg_off[k + 1] < seg_off[k]) return "segment offsets must be non-decreasing";
    int64_t k = 0;
    for (int64_t i = 0; i <= n_inst; ++i) {
      while (k < n_seg_total && seg_off[k] < inst_tet_off[i]) ++k;
      if (seg_off[k] != inst_tet_off[i]) return "instance boundaries must be segment boundaries";
      inst_seg[i] = k;
    }
    segs.assign(seg_off, seg_off + n_seg_total + 1);
  } else {
    segs.assign(inst_tet_off, inst_tet_off + n_inst + 1);
    for (int64_t i = 0; i <= n_inst; ++i) inst_seg[i] = i;
  }
  const int64_t n_seg = (int64_t)segs.size() - 1;
  Layout& L = m->L;
  L = Layout();
  L.nth = nth;
  L.chunks = chunks;
  L.nloc_cap = nloc_cap;
  L.T = T;
  L.V = V;
  L.n_inst = n_inst;
  const int max_tets = nth * chunks;
  DevLayout d;  // assigned to the mesh on success (the arena frees everything on an error)
  Arena A;
  auto keep = [&](void* p, size_t bytes) {
    const auto it = std::find(A.tmp.begin(), A.tmp.end(), p);
    if (it != A.tmp.end()) A.tmp.erase(it);
    m->device_bytes += bytes;
  };
  int32_t* d_cells = A.get<int32_t>(4 * (size_t)T);
  int64_t* d_ito = A.get<int64_t>((size_t)n_inst + 1);
  int64_t* d_ivo = A.get<int64_t>((size_t)n_inst + 1);
  PART_CHECK(A.err);
  if (T > 0) PART_CHECK(hipMemcpy(d_cells, cells, 4 * (size_t)T * sizeof(int32_t), hipMemcpyHostToDevice));
  PART_CHECK(hipMemcpy(d_ito, inst_tet_off, ((size_t)n_inst + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  PART_CHECK(hipMemcpy(d_ivo, inst_vtx_off, ((size_t)n_inst + 1) * sizeof(int64_t), hipMemcpyHostToDevice));

  // 0. every corner inside its instance's vertex range
  if (T > 0) {
    unsigned long long* d_bad = A.get<unsigned long long>(1);
    PART_CHECK(A.err);
    const unsigned long long none = ULLONG_MAX;
    PART_CHECK(hipMemcpy(d_bad, &none, sizeof none, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(check_cells_kernel, dim3(grid_n(T)), dim3(PNT), 0, nullptr, d_cells, (long long)T, d_ito, d_ivo,
                       (int)n_inst, d_bad);
    PART_CHECK(hipGetLastError());
    unsigned long long bad = none;
    PART_CHECK(hipMemcpy(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost));
    if (bad != none) {
      int64_t v = 0;
      for (int c = 0; c < 4; ++c) {
        const int64_t w = cells[4 * (int64_t)bad + c];
        const int64_t i = std::upper_bound(inst_tet_off, inst_tet_off + n_inst + 1, (int64_t)bad) - inst_tet_off - 1;
        if (w < inst_vtx_off[i] || w >= inst_vtx_off[i + 1]) {
          v = w;
          break;
        }
      }
      if (v < 0 || v >= V) return "cell index out of range [0, n_verts)";
      snprintf(msg, sizeof msg, "tet %lld references vertex %lld outside its instance's range", (long long)bad,
               (long long)v);
      return msg;
    }
  }

  // 1. blocks.  Runs of max_tets within each segment are the greedy's blocks
  //    whenever none of them holds more than cap distinct vertices (a prefix
  //    of a run never holds more than the run): always so when 4 max_tets <=
  //    cap, else checked by computing their slots first, which a
  //    locality-ordered mesh passes.  Otherwise the greedy kernel's ends.
  auto fixed_blocks = [&]() {
    L.blk_tet_off.assign(1, 0);
    L.inst_blk_off.assign(1, 0);
    L.blk_inst.clear();
    L.seg_of_blk.clear();
    for (int64_t i = 0; i < n_inst; ++i) {
      for (int64_t k = inst_seg[i]; k < inst_seg[i + 1]; ++k)
        for (int64_t t = segs[k]; t < segs[k + 1]; t += max_tets) {
          L.blk_tet_off.push_back((int32_t)std::min<int64_t>(t + max_tets, segs[k + 1]));
          L.blk_inst.push_back((int32_t)i);
          L.seg_of_blk.push_back(seg_off != nullptr ? (int32_t)k : 0);
        }
      L.inst_blk_off.push_back((int64_t)L.blk_inst.size());
    }
  };
  // 2a. slots of the current blocks: the block tables to the device, the
  //     first-touch slots into the corner records, the slot counts back
  d.corner = A.get<uint16_t>(8 * (size_t)T);
  int32_t* d_gid_tmp = A.get<int32_t>(4 * (size_t)T);
  PART_CHECK(A.err);
  std::vector<int32_t> nloc;
  auto block_slots = [&]() -> std::string {
    const int64_t nb = (int64_t)L.blk_inst.size();
    d.blk_tet_off = A.get<int32_t>((size_t)nb + 1);
    d.blk_inst = A.get<int32_t>((size_t)nb);
    d.inst_blk_off = A.get<int64_t>((size_t)n_inst + 1);
    int32_t* d_nloc = A.get<int32_t>((size_t)nb);
    PART_CHECK(A.err);
    PART_CHECK(hipMemcpy(d.blk_tet_off, L.blk_tet_off.data(), ((size_t)nb + 1) * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    if (nb > 0)
      PART_CHECK(hipMemcpy(d.blk_inst, L.blk_inst.data(), (size_t)nb * sizeof(int32_t), hipMemcpyHostToDevice));
    PART_CHECK(hipMemcpy(d.inst_blk_off, L.inst_blk_off.data(), ((size_t)n_inst + 1) * sizeof(int64_t),
                         hipMemcpyHostToDevice));
    nloc.assign((size_t)nb, 0);
    if (nb > 0) {
      const size_t lds = 2 * (size_t)ht_size(4 * max_tets) * sizeof(int);
      hipLaunchKernelGGL(block_slots_kernel, dim3((unsigned)nb), dim3(PNT), lds, nullptr, d_cells, d.blk_tet_off,
                         d.corner, d_nloc, d_gid_tmp, max_tets);
      PART_CHECK(hipGetLastError());
      PART_CHECK(hipMemcpy(nloc.data(), d_nloc, (size_t)nb * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    return std::string();
  };
  fixed_blocks();
  std::string err = block_slots();
  if (!err.empty()) return err;
  int fixed_max = 0;
  for (const int32_t x : nloc) fixed_max = std::max(fixed_max, (int)x);
  const bool fixed_ok = fixed_max <= nloc_cap;
  if (!fixed_ok) {
    int64_t* d_segs = A.get<int64_t>(segs.size());
    int64_t* d_iseg = A.get<int64_t>(inst_seg.size());
    int32_t* d_ends = A.get<int32_t>((size_t)T);
    int64_t* d_counts = A.get<int64_t>((size_t)n_inst);
    PART_CHECK(A.err);
    PART_CHECK(hipMemcpy(d_segs, segs.data(), segs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    PART_CHECK(hipMemcpy(d_iseg, inst_seg.data(), inst_seg.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    const size_t lds = 2 * (size_t)ht_size(4 * max_tets) * sizeof(int);
    hipLaunchKernelGGL(greedy_blocks_kernel, dim3((unsigned)n_inst), dim3(PNT), lds, nullptr, d_cells, d_segs, d_iseg,
                       d_ito, max_tets, nloc_cap, d_ends, d_counts);
    PART_CHECK(hipGetLastError());
    std::vector<int64_t> counts(n_inst);
    std::vector<int32_t> ends((size_t)T);
    PART_CHECK(hipMemcpy(counts.data(), d_counts, (size_t)n_inst * sizeof(int64_t), hipMemcpyDeviceToHost));
    PART_CHECK(hipMemcpy(ends.data(), d_ends, (size_t)T * sizeof(int32_t), hipMemcpyDeviceToHost));
    L.blk_tet_off.assign(1, 0);
    L.inst_blk_off.assign(1, 0);
    L.blk_inst.clear();
    L.seg_of_blk.clear();
    for (int64_t i = 0; i < n_inst; ++i) {
      int64_t k = inst_seg[i];
      for (int64_t j = 0; j < counts[i]; ++j) {
        const int64_t t0 = L.blk_tet_off.back();
        while (k < inst_seg[i + 1] && segs[k + 1] <= t0) ++k;
        L.blk_tet_off.push_back(ends[inst_tet_off[i] + j]);
        L.blk_inst.push_back((int32_t)i);
        L.seg_of_blk.push_back(seg_off != nullptr ? (int32_t)k : 0);
      }
      L.inst_blk_off.push_back((int64_t)L.blk_inst.size());
    }
    err = block_slots();
    if (!err.empty()) return err;
  }
  L.nblocks = (int64_t)L.blk_inst.size();
  const int64_t nb = L.nblocks;

  // 2b. the offsets over blocks (host prefix sums over the block counts),
  //     then chunk offsets, positions and the slot list
  L.blk_loc_off.push_back(0);
  L.blk_co_off.push_back(0);
  int64_t n_slots = 0, n_co = 0;
  for (int64_t b = 0; b < nb; ++b) {
    n_slots += nloc[b];
    if (n_slots >= (int64_t(1) << 31)) return "too many local slots";
    L.nloc_max = std::max(L.nloc_max, (int)nloc[b]);
    L.blk_loc_off.push_back((int32_t)n_slots);
    const int64_t nt = L.blk_tet_off[b + 1] - L.blk_tet_off[b];
    n_co += ((nt + nth - 1) / nth) * ((int64_t)nloc[b] + 1);
    if (n_co >= (int64_t(1) << 31)) return "too many slot offsets";
    L.blk_co_off.push_back((int32_t)n_co);
  }
  L.n_slots = n_slots;
  d.blk_loc_off = A.get<int32_t>((size_t)nb + 1);
  d.blk_co_off = A.get<int32_t>((size_t)nb + 1);
  d.chunk_off = A.get<uint16_t>((size_t)n_co);
  d.loc_gid = A.get<int32_t>((size_t)n_slots);
  PART_CHECK(A.err);
  PART_CHECK(hipMemcpy(d.blk_loc_off, L.blk_loc_off.data(), ((size_t)nb + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  PART_CHECK(hipMemcpy(d.blk_co_off, L.blk_co_off.data(), ((size_t)nb + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  if (nb > 0) {
    hipLaunchKernelGGL(chunk_rank_kernel, dim3((unsigned)nb), dim3(PNT), 0, nullptr, d.blk_tet_off, d.blk_loc_off,
                       d.blk_co_off, d_gid_tmp, d.corner, d.chunk_off, d.loc_gid, nth);
    PART_CHECK(hipGetLastError());
  }

  // 3. vertex CSR over slots and each slot's place in it
  d.vtx_slot_off = A.get<int32_t>((size_t)V + 1);
  d.loc_dst = A.get<int32_t>((size_t)n_slots);
  int32_t* d_cnt = A.get<int32_t>((size_t)V + 1);
  unsigned long long* d_shared = A.get<unsigned long long>(1);
  PART_CHECK(A.err);
  PART_CHECK(hipMemset(d_cnt, 0, ((size_t)V + 1) * sizeof(int32_t)));
  PART_CHECK(hipMemset(d_shared, 0, sizeof(unsigned long long)));
  PART_CHECK(hipMemset(d.vtx_slot_off, 0, sizeof(int32_t)));
  if (n_slots > 0) {
    hipLaunchKernelGGL(vertex_count_kernel, dim3(grid_n(n_slots)), dim3(PNT), 0, nullptr, d.loc_gid,
                       (long long)n_slots, d_cnt);
    PART_CHECK(hipGetLastError());
  }
  if (V > 0) {
    hipLaunchKernelGGL(shared_count_kernel, dim3(grid_n(V)), dim3(PNT), 0, nullptr, d_cnt, (long long)V, d_shared);
    PART_CHECK(hipGetLastError());
    size_t tb = 0;
    PART_CHECK(rocprim::inclusive_scan(nullptr, tb, d_cnt, d.vtx_slot_off + 1, (size_t)V, rocprim::plus<int32_t>()));
    void* tmp = A.get<char>(tb);
    PART_CHECK(A.err);
    PART_CHECK(rocprim::inclusive_scan(tmp, tb, d_cnt, d.vtx_slot_off + 1, (size_t)V, rocprim::plus<int32_t>()));
  }
  if (n_slots > 0) {
    uint32_t* keys_out = A.get<uint32_t>((size_t)n_slots);
    int32_t* vals_in = A.get<int32_t>((size_t)n_slots);
    int32_t* vals_out = A.get<int32_t>((size_t)n_slots);
    PART_CHECK(A.err);
    hipLaunchKernelGGL(iota_kernel, dim3(grid_n(n_slots)), dim3(PNT), 0, nullptr, vals_in, (long long)n_slots);
    PART_CHECK(hipGetLastError());
    const uint32_t* keys_in = reinterpret_cast<const uint32_t*>(d.loc_gid);
    const int end_bit = bits_for(V);
    size_t tb = 0;
    PART_CHECK(rocprim::radix_sort_pairs(nullptr, tb, keys_in, keys_out, vals_in, vals_out, (size_t)n_slots, 0,
                                         end_bit));
    void* tmp = A.get<char>(tb);
    PART_CHECK(A.err);
    PART_CHECK(rocprim::radix_sort_pairs(tmp, tb, keys_in, keys_out, vals_in, vals_out, (size_t)n_slots, 0, end_bit));
    hipLaunchKernelGGL(invert_kernel, dim3(grid_n(n_slots)), dim3(PNT), 0, nullptr, vals_out, (long long)n_slots,
                       d.loc_dst);
    PART_CHECK(hipGetLastError());
  }
  unsigned long long n_shared = 0;
  PART_CHECK(hipMemcpy(&n_shared, d_shared, sizeof n_shared, hipMemcpyDeviceToHost));
  L.n_shared_vtx = (int64_t)n_shared;

  // 4. vertex blocks aligned to instances (bookkeeping over the instance offsets)
  L.vblk_off.push_back(0);
  L.inst_vblk_off.push_back(0);
  for (int64_t i = 0; i < n_inst; ++i) {
    for (int64_t v = inst_vtx_off[i]; v < inst_vtx_off[i + 1]; v += L.vb) {
      L.vblk_off.push_back((int32_t)std::min<int64_t>(v + L.vb, inst_vtx_off[i + 1]));
      L.vblk_inst.push_back((int32_t)i);
    }
    L.inst_vblk_off.push_back((int64_t)L.vblk_inst.size());
  }
  const size_t nvb = L.vblk_inst.size();
  d.vblk_off = A.get<int32_t>(nvb + 1);
  d.vblk_inst = A.get<int32_t>(nvb);
  d.inst_vblk_off = A.get<int64_t>((size_t)n_inst + 1);
  PART_CHECK(A.err);
  PART_CHECK(hipMemcpy(d.vblk_off, L.vblk_off.data(), (nvb + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  if (nvb > 0) PART_CHECK(hipMemcpy(d.vblk_inst, L.vblk_inst.data(), nvb * sizeof(int32_t), hipMemcpyHostToDevice));
  PART_CHECK(hipMemcpy(d.inst_vblk_off, L.inst_vblk_off.data(), ((size_t)n_inst + 1) * sizeof(int64_t),
                       hipMemcpyHostToDevice));
  PART_CHECK(hipDeviceSynchronize());

  // keep the mesh's arrays, free the temporaries
  keep(d.corner, 8 * (size_t)T * sizeof(uint16_t));
  keep(d.blk_tet_off, ((size_t)nb + 1) * sizeof(int32_t));
  keep(d.blk_inst, (size_t)nb * sizeof(int32_t));
  keep(d.inst_blk_off, ((size_t)n_inst + 1) * sizeof(int64_t));
  keep(d.blk_loc_off, ((size_t)nb + 1) * sizeof(int32_t));
  keep(d.blk_co_off, ((size_t)nb + 1) * sizeof(int32_t));
  keep(d.chunk_off, (size_t)n_co * sizeof(uint16_t));
  keep(d.loc_gid, (size_t)n_slots * sizeof(int32_t));
  keep(d.vtx_slot_off, ((size_t)V + 1) * sizeof(int32_t));
  keep(d.loc_dst, (size_t)n_slots * sizeof(int32_t));
  keep(d.vblk_off, (nvb + 1) * sizeof(int32_t));
  keep(d.vblk_inst, nvb * sizeof(int32_t));
  keep(d.inst_vblk_off, ((size_t)n_inst + 1) * sizeof(int64_t));
  DevLayout& md = m->d;
  md.corner = d.corner;
  md.blk_tet_off = d.blk_tet_off;
  md.blk_inst = d.blk_inst;
  md.inst_blk_off = d.inst_blk_off;
  md.blk_loc_off = d.blk_loc_off;
  md.blk_co_off = d.blk_co_off;
  md.chunk_off = d.chunk_off;
  md.loc_gid = d.loc_gid;
  md.vtx_slot_off = d.vtx_slot_off;
  md.loc_dst = d.loc_dst;
  md.vblk_off = d.vblk_off;
  md.vblk_inst = d.vblk_inst;
  md.inst_vblk_off = d.inst_vblk_off;
  return std::string();
}

// ------------------------------------------------------------------ Morton order

namespace {

// per-axis minimum and maximum of the points: one partial per workgroup
__global__ __launch_bounds__(PNT) void minmax_kernel(const double* __restrict__ P, long long V,
                                                     double* __restrict__ part) {
  __shared__ double sm[6][PNT];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (long long v = (long long)blockIdx.x * PNT + threadIdx.x; v < V; v += (long long)gridDim.x * PNT)
    for (int a = 0; a < 3; ++a) {
      const double x = P[3 * v + a];
      lo[a] = fmin(lo[a], x);
      hi[a] = fmax(hi[a], x);
    }
  for (int a = 0; a < 3; ++a) {
    sm[a][threadIdx.x] = lo[a];
    sm[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int d = PNT / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d)
      for (int a = 0; a < 3; ++a) {
        sm[a][threadIdx.x] = fmin(sm[a][threadIdx.x], sm[a][threadIdx.x + d]);
        sm[3 + a][threadIdx.x] = fmax(sm[3 + a][threadIdx.x], sm[3 + a][threadIdx.x + d]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = sm[k][0];
}

// the workgroup partials reduced to (lo, hi) per axis, in place at part[0..5]
__global__ __launch_bounds__(PNT) void minmax_final_kernel(double* __restrict__ part, int nparts) {
  if (threadIdx.x != 0) return;
  double r[6];
  for (int k = 0; k < 6; ++k) r[k] = part[k];
  for (int j = 1; j < nparts; ++j)
    for (int a = 0; a < 3; ++a) {
      r[a] = fmin(r[a], part[6 * j + a]);
      r[3 + a] = fmax(r[3 + a], part[6 * j + 3 + a]);
    }
  for (int k = 0; k < 6; ++k) part[k] = r[k];
}

// the 11-bit quantisation of x over [lo, lo + ext] and the 3-D interleave
// (bit 3b + a = bit b of axis a), as apple_amd.mesh._morton3
__device__ inline uint64_t morton_of(const double x[3], const double* lohi) {
  uint64_t code = 0;
  for (int a = 0; a < 3; ++a) {
    const double lo = lohi[a], hi = lohi[3 + a];
    const double ext = fmax(hi - lo, 1e-30);
    const long long q = (long long)((x[a] - lo) / ext * 2047.0);
    for (int b = 0; b < 11; ++b) code |= (uint64_t)((q >> b) & 1) << (3 * b + a);
  }
  return code;
}

__global__ __launch_bounds__(PNT) void morton_vertex_kernel(const double* __restrict__ P, long long V,
                                                            const double* __restrict__ lohi,
                                                            uint64_t* __restrict__ codes, int32_t* __restrict__ idx) {
  for (long long v = (long long)blockIdx.x * PNT + threadIdx.x; v < V; v += (long long)gridDim.x * PNT) {
    const double x[3] = {P[3 * v], P[3 * v + 1], P[3 * v + 2]};
    codes[v] = morton_of(x, lohi);
    idx[v] = (int32_t)v;
  }
}

// tet centroids ((p0 + p1) + p2) + p3, then / 4, of the original points.
// A tet with a corner outside [0, V) reads no point: it gets code 0 and
// *bad the smallest such tet (the caller turns that into an error).
__global__ __launch_bounds__(PNT) void morton_cell_kernel(const double* __restrict__ P, long long V,
                                                          const int32_t* __restrict__ cells, long long T,
                                                          const double* __restrict__ lohi,
                                                          uint64_t* __restrict__ codes, int32_t* __restrict__ idx,
                                                          unsigned long long* __restrict__ bad) {
  for (long long t = (long long)blockIdx.x * PNT + threadIdx.x; t < T; t += (long long)gridDim.x * PNT) {
    long long id[4];
    bool ok = true;
    for (int c = 0; c < 4; ++c) {
      id[c] = cells[4 * t + c];
      ok = ok && id[c] >= 0 && id[c] < V;
    }
    uint64_t code = 0;
    if (ok) {
      double x[3];
      for (int a = 0; a < 3; ++a) {
        double s = P[3 * id[0] + a];
        for (int c = 1; c < 4; ++c) s += P[3 * id[c] + a];
        x[a] = s / 4.0;
      }
      code = morton_of(x, lohi);
    } else {
      atomicMin(bad, (unsigned long long)t);
    }
    codes[t] = code;
    idx[t] = (int32_t)t;
  }
}

int sort_codes(Arena& A, uint64_t* codes, int32_t* idx, long long n, int32_t* perm_out, hipStream_t s) {
  uint64_t* keys_out = A.get<uint64_t>((size_t)n);
  if (A.err != hipSuccess) return -1;
  size_t tb = 0;
  if (rocprim::radix_sort_pairs(nullptr, tb, codes, keys_out, idx, perm_out, (size_t)n, 0, 33, s) != hipSuccess) return -1;
  void* tmp = A.get<char>(tb);
  if (A.err != hipSuccess) return -1;
  if (rocprim::radix_sort_pairs(tmp, tb, codes, keys_out, idx, perm_out, (size_t)n, 0, 33, s) != hipSuccess) return -1;
  return 0;
}

// ------------------------------------------------------------------ lumped vertex volumes

// per block: sum_t density_t dV_t / 4 over each slot's corners (FP64, in
// position order, chunk after chunk), one partial per slot at its place in
// the vertex-major order
template <class R>
__global__ __launch_bounds__(PNT) void lumped_block_kernel(const int32_t* __restrict__ blk_tet_off,
                                                           const int32_t* __restrict__ blk_loc_off,
                                                           const int32_t* __restrict__ blk_co_off,
                                                           const uint16_t* __restrict__ chunk_off,
                                                           const uint16_t* __restrict__ corner,
                                                           const int32_t* __restrict__ loc_dst,
                                                           const R* __restrict__ dV, const R* __restrict__ rho,
                                                           double* __restrict__ partial, int nth) {
  __shared__ double val[1024];
  __shared__ double acc[1024];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const long long t0 = blk_tet_off[b], t1 = blk_tet_off[b + 1];
  const int sb = blk_loc_off[b];
  const int nloc = blk_loc_off[b + 1] - sb;
  for (int s = tid; s < nloc; s += PNT) acc[s] = 0.0;
  const uint16_t* co = chunk_off + blk_co_off[b];
  int c = 0;
  for (long long c0 = t0; c0 < t1; c0 += nth, ++c) {
    const long long n_t = std::min<long long>(t1, c0 + nth) - c0;
    for (long long k = tid; k < n_t; k += PNT) {
      const long long t = c0 + k;
      const double w = (rho != nullptr ? (double)rho[t] * (double)dV[t] : (double)dV[t]) / 4.0;
      for (int q = 0; q < 4; ++q) val[corner[8 * t + 4 + q]] = w;
    }
    __syncthreads();
    const uint16_t* cc = co + (long long)c * (nloc + 1);
    for (int s = tid; s < nloc; s += PNT)
      for (int p = cc[s]; p < cc[s + 1]; ++p) acc[s] += val[p];
    __syncthreads();
  }
  for (int s = tid; s < nloc; s += PNT) partial[loc_dst[sb + s]] = acc[s];
}

__global__ __launch_bounds__(PNT) void lumped_vertex_kernel(const int32_t* __restrict__ vtx_slot_off, long long V,
                                                            const double* __restrict__ partial,
                                                            double* __restrict__ out) {
  for (long long v = (long long)blockIdx.x * PNT + threadIdx.x; v < V; v += (long long)gridDim.x * PNT) {
    double s = 0.0;
    for (int j = vtx_slot_off[v]; j < vtx_slot_off[v + 1]; ++j) s += partial[j];
    out[v] = s;
  }
}

}  // namespace

}  // namespace apl

using namespace apl;

extern "C" {

int apl_morton_order(const double* points, int64_t n_verts, const int32_t* cells, int64_t n_tets,
                     int32_t* vertex_perm, int32_t* cell_perm, int device, void* stream) {
  if (n_verts < 0 || n_tets < 0 || n_verts >= (int64_t(1) << 31) || n_tets >= (int64_t(1) << 31) ||
      (n_verts > 0 && (points == nullptr || vertex_perm == nullptr)) ||
      (n_tets > 0 && (cells == nullptr || cell_perm == nullptr))) {
    set_error("invalid Morton-order arguments");
    return -1;
  }
  if (n_verts == 0 && n_tets > 0) {
    set_error("cell index out of range [0, n_verts)");
    return -1;
  }
  if (n_verts == 0) return 0;
  APL_CHECK_HIP(hipSetDevice(device));
  (void)hipGetLastError();
  hipStream_t s = (hipStream_t)stream;
  Arena A;
  const unsigned np = std::min<unsigned>(grid_n(n_verts), 1024u);
  double* part = A.get<double>(6 * (size_t)np);
  uint64_t* codes = A.get<uint64_t>((size_t)std::max(n_verts, n_tets));
  int32_t* idx = A.get<int32_t>((size_t)std::max(n_verts, n_tets));
  unsigned long long* bad = A.get<unsigned long long>(1);
  if (A.err != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(A.err));
    return -1;
  }
  hipLaunchKernelGGL(minmax_kernel, dim3(np), dim3(PNT), 0, s, points, (long long)n_verts, part);
  hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(PNT), 0, s, part, (int)np);
  hipLaunchKernelGGL(morton_vertex_kernel, dim3(grid_n(n_verts)), dim3(PNT), 0, s, points, (long long)n_verts, part,
                     codes, idx);
  APL_CHECK_HIP(hipGetLastError());
  if (sort_codes(A, codes, idx, n_verts, vertex_perm, s)) {
    set_error("Morton sort of the vertices failed");
    return -1;
  }
  const unsigned long long none = ULLONG_MAX;
  unsigned long long bad_t = none;
  if (n_tets > 0) {
    APL_CHECK_HIP(hipMemcpyAsync(bad, &none, sizeof none, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(morton_cell_kernel, dim3(grid_n(n_tets)), dim3(PNT), 0, s, points, (long long)n_verts, cells,
                       (long long)n_tets, part, codes, idx, bad);
    APL_CHECK_HIP(hipGetLastError());
    if (sort_codes(A, codes, idx, n_tets, cell_perm, s)) {
      set_error("Morton sort of the tets failed");
      return -1;
    }
    APL_CHECK_HIP(hipMemcpyAsync(&bad_t, bad, sizeof bad_t, hipMemcpyDeviceToHost, s));
  }
  APL_CHECK_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
  if (bad_t != none) {
    char msg[128];
    snprintf(msg, sizeof msg, "cell index out of range [0, n_verts) in tet %llu", bad_t);
    set_error(msg);
    return -1;
  }
  return 0;
}

int apl_mesh_lumped_volume(apl_mesh* mesh, const void* density, double* out, void* stream) {
  if (mesh == nullptr || (mesh->L.V > 0 && out == nullptr)) {
    set_error("mesh and out must be non-NULL");
    return -1;
  }
  if (mesh->L.V == 0) return 0;
  APL_CHECK_HIP(hipSetDevice(mesh->device));
  (void)hipGetLastError();
  hipStream_t s = (hipStream_t)stream;
  const DevLayout& d = mesh->d;
  double* partial = (double*)d.partial;  // the mesh's pass scratch: MAX_CH values per slot >= one double
  if (mesh->L.nblocks > 0) {
    if (mesh->dtype == DT_F32)
      hipLaunchKernelGGL(lumped_block_kernel<float>, dim3((unsigned)mesh->L.nblocks), dim3(PNT), 0, s, d.blk_tet_off,
                         d.blk_loc_off, d.blk_co_off, d.chunk_off, d.corner, d.loc_dst, (const float*)d.dV,
                         (const float*)density, partial, mesh->L.nth);
    else
      hipLaunchKernelGGL(lumped_block_kernel<double>, dim3((unsigned)mesh->L.nblocks), dim3(PNT), 0, s,
                         d.blk_tet_off, d.blk_loc_off, d.blk_co_off, d.chunk_off, d.corner, d.loc_dst,
                         (const double*)d.dV, (const double*)density, partial, mesh->L.nth);
    APL_CHECK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(lumped_vertex_kernel, dim3(grid_n(mesh->L.V)), dim3(PNT), 0, s, d.vtx_slot_off,
                     (long long)mesh->L.V, partial, out);
  APL_CHECK_HIP(hipGetLastError());
  return 0;
}

int apl_mesh_layout(const apl_mesh* mesh, int which, void* out, int64_t n) {
  if (mesh == nullptr || out == nullptr) {
    set_error("mesh and out must be non-NULL");
    return -1;
  }
  const Layout& L = mesh->L;
  const DevLayout& d = mesh->d;
  const int64_t nb = L.nblocks, ni = L.n_inst, nvb = (int64_t)L.vblk_inst.size();
  const int64_t n_co = L.blk_co_off.empty() ? 0 : L.blk_co_off.back();
  const void* src = nullptr;
  int64_t cnt = 0;
  size_t es = 4;
  switch (which) {
    case APL_LAYOUT_CORNER: src = d.corner, cnt = 8 * L.T, es = 2; break;
    case APL_LAYOUT_BLK_TET_OFF: src = d.blk_tet_off, cnt = nb + 1; break;
    case APL_LAYOUT_BLK_LOC_OFF: src = d.blk_loc_off, cnt = nb + 1; break;
    case APL_LAYOUT_BLK_INST: src = d.blk_inst, cnt = nb; break;
    case APL_LAYOUT_INST_BLK_OFF: src = d.inst_blk_off, cnt = ni + 1, es = 8; break;
    case APL_LAYOUT_LOC_GID: src = d.loc_gid, cnt = L.n_slots; break;
    case APL_LAYOUT_BLK_CO_OFF: src = d.blk_co_off, cnt = nb + 1; break;
    case APL_LAYOUT_CHUNK_OFF: src = d.chunk_off, cnt = n_co, es = 2; break;
    case APL_LAYOUT_VTX_SLOT_OFF: src = d.vtx_slot_off, cnt = L.V + 1; break;
    case APL_LAYOUT_LOC_DST: src = d.loc_dst, cnt = L.n_slots; break;
    case APL_LAYOUT_VBLK_OFF: src = d.vblk_off, cnt = nvb + 1; break;
    case APL_LAYOUT_VBLK_INST: src = d.vblk_inst, cnt = nvb; break;
    case APL_LAYOUT_INST_VBLK_OFF: src = d.inst_vblk_off, cnt = ni + 1, es = 8; break;
    default:
      set_error("unknown layout array");
      return -1;
  }
  if (n != cnt) {
    set_error("layout array has " + std::to_string(cnt) + " entries, not " + std::to_string(n));
    return -1;
  }
  if (cnt == 0) return 0;
  if (src == nullptr) {
    set_error("layout array not on the device");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(mesh->device));
  APL_CHECK_HIP(hipMemcpy(out, src, (size_t)cnt * es, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
