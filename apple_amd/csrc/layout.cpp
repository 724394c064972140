// Host-side block partition; see layout.h.
#include "layout.h"

#include <algorithm>
#include <cstdio>

namespace apl {

std::string build_layout(const int32_t* cells, int64_t T, int64_t V, int64_t n_inst,
                         const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int nth,
                         int chunks, int nloc_cap, Layout& L, int64_t n_seg_total, const int64_t* seg_off) {
  char msg[256];
  if (T < 0 || V < 0 || n_inst < 1) return "invalid mesh sizes";
  if (nth != 256) return "threads per block must be 256";
  if (L.vb != nth) return "vertex blocks must hold one vertex per thread";
  if (chunks < 1 || (int64_t)nth * chunks > 8192) return "chunks must be in [1, 32]";
  if (nloc_cap < 4 || nloc_cap > 1024) return "nloc_cap must be in [4, 1024]";
  // 32-bit element offsets: 9 T and 3 V (8-byte values: 24 V bytes) must fit
  if (T >= (int64_t(1) << 27) || V >= (int64_t(1) << 27)) return "mesh too large (at most 2^27 tets and vertices)";
  if (inst_tet_off[0] != 0 || inst_tet_off[n_inst] != T) return "instance tet offsets must span [0, T]";
  if (inst_vtx_off[0] != 0 || inst_vtx_off[n_inst] != V) return "instance vertex offsets must span [0, V]";
  for (int64_t i = 0; i < n_inst; ++i) {
    if (inst_tet_off[i + 1] < inst_tet_off[i] || inst_vtx_off[i + 1] < inst_vtx_off[i])
      return "instance offsets must be non-decreasing";
    for (int64_t t = inst_tet_off[i]; t < inst_tet_off[i + 1]; ++t)
      for (int c = 0; c < 4; ++c) {
        int64_t v = cells[t * 4 + c];
        if (v < inst_vtx_off[i] || v >= inst_vtx_off[i + 1]) {
          snprintf(msg, sizeof msg, "tet %lld references vertex %lld outside its instance's range",
                   (long long)t, (long long)v);
          return msg;
        }
      }
  }
  if (seg_off != nullptr) {
    if (n_seg_total < n_inst || seg_off[0] != 0 || seg_off[n_seg_total] != T) return "segment offsets must span [0, T]";
    for (int64_t k = 0; k < n_seg_total; ++k)
      if (seg_off[k + 1] < seg_off[k]) return "segment offsets must be non-decreasing";
    // every instance boundary must be a segment boundary
    int64_t k = 0;
    for (int64_t i = 0; i <= n_inst; ++i) {
      while (k < n_seg_total && seg_off[k] < inst_tet_off[i]) ++k;
      if (seg_off[k] != inst_tet_off[i]) return "instance boundaries must be segment boundaries";
    }
  }
  L = Layout();
  L.nth = nth;
  L.chunks = chunks;
  L.nloc_cap = nloc_cap;
  L.T = T;
  L.V = V;
  L.n_inst = n_inst;
  const int max_tets = nth * chunks;

  // 1. greedy blocks of consecutive tets within each instance
  std::vector<int64_t> stamp(V, -1);
  L.blk_tet_off.push_back(0);
  L.inst_blk_off.push_back(0);
  int64_t seg = 0;  // segment of tet t (segments are consecutive)
  for (int64_t i = 0; i < n_inst; ++i) {
    int64_t t = inst_tet_off[i], t1 = inst_tet_off[i + 1];
    while (t < t1) {
      int64_t b = (int64_t)L.blk_inst.size();
      int nloc = 0, nt = 0;
      // this block ends at the instance's or the segment's end
      int64_t tend = t1;
      if (seg_off != nullptr) {
        while (seg_off[seg + 1] <= t) ++seg;
        tend = std::min<int64_t>(tend, seg_off[seg + 1]);
      }
      L.seg_of_blk.push_back(seg_off != nullptr ? (int32_t)seg : 0);
      while (t < tend && nt < max_tets) {
        int add = 0;
        for (int c = 0; c < 4; ++c) {
          int64_t v = cells[t * 4 + c];
          bool dup = stamp[v] == b;
          for (int c2 = 0; c2 < c; ++c2) dup = dup || cells[t * 4 + c2] == cells[t * 4 + c];
          add += dup ? 0 : 1;
        }
        if (nloc + add > nloc_cap) break;
        for (int c = 0; c < 4; ++c) stamp[cells[t * 4 + c]] = b;
        nloc += add;
        ++nt;
        ++t;
      }
      L.blk_tet_off.push_back((int32_t)t);
      L.blk_inst.push_back((int32_t)i);
    }
    L.inst_blk_off.push_back((int64_t)L.blk_inst.size());
  }
  L.nblocks = (int64_t)L.blk_inst.size();

  // 2. local slots in first-touch order; per chunk, corner positions ranked
  //    by (slot, tet, corner) and the slot offsets into them
  L.corner.resize(8 * T);
  L.blk_loc_off.push_back(0);
  L.blk_co_off.push_back(0);
  std::vector<int32_t> local_of(V, -1);
  std::fill(stamp.begin(), stamp.end(), -1);
  std::vector<int32_t> cnt;
  std::vector<int32_t> nslot_of_v(V + 1, 0);
  for (int64_t b = 0; b < L.nblocks; ++b) {
    const int64_t t0 = L.blk_tet_off[b], t1 = L.blk_tet_off[b + 1];
    const int64_t base = (int64_t)L.loc_gid.size();
    for (int64_t t = t0; t < t1; ++t)
      for (int c = 0; c < 4; ++c) {
        const int64_t v = cells[t * 4 + c];
        if (stamp[v] != b) {
          stamp[v] = b;
          local_of[v] = (int32_t)(L.loc_gid.size() - base);
          L.loc_gid.push_back((int32_t)v);
          ++nslot_of_v[v + 1];
        }
      }
    const int nloc = (int)(L.loc_gid.size() - base);
    if ((int64_t)L.loc_gid.size() >= (int64_t(1) << 31)) return "too many local slots";
    L.nloc_max = std::max(L.nloc_max, nloc);
    for (int64_t c0 = t0; c0 < t1; c0 += nth) {
      const int64_t c1 = std::min<int64_t>(t1, c0 + nth);
      cnt.assign(nloc + 1, 0);
      for (int64_t t = c0; t < c1; ++t)
        for (int c = 0; c < 4; ++c) {
          const int s = local_of[cells[t * 4 + c]];
          L.corner[8 * t + c] = (uint16_t)s;
          ++cnt[s + 1];
        }
      for (int s = 0; s < nloc; ++s) cnt[s + 1] += cnt[s];
      for (int s = 0; s <= nloc; ++s) L.chunk_off.push_back((uint16_t)cnt[s]);
      for (int64_t t = c0; t < c1; ++t)
        for (int c = 0; c < 4; ++c) L.corner[8 * t + 4 + c] = (uint16_t)cnt[L.corner[8 * t + c]]++;
    }
    if (L.chunk_off.size() >= (size_t(1) << 31)) return "too many slot offsets";
    L.blk_co_off.push_back((int32_t)L.chunk_off.size());
    L.blk_loc_off.push_back((int32_t)L.loc_gid.size());
  }
  L.n_slots = (int64_t)L.loc_gid.size();

  // 3. vertex CSR over slots: a vertex's slots in ascending slot id, i.e. in
  //    block order, each slot's partial at its place in that order; vertices
  //    no tet touches get empty ranges (their sums are 0)
  L.vtx_slot_off.assign(V + 1, 0);
  for (int64_t v = 0; v < V; ++v) L.vtx_slot_off[v + 1] = L.vtx_slot_off[v] + nslot_of_v[v + 1];
  L.loc_dst.assign(L.n_slots, 0);
  {
    std::vector<int32_t> fill(L.vtx_slot_off.begin(), L.vtx_slot_off.end() - 1);
    for (int64_t k = 0; k < L.n_slots; ++k) L.loc_dst[k] = fill[L.loc_gid[k]]++;
  }
  L.n_shared_vtx = 0;
  for (int64_t v = 0; v < V; ++v) L.n_shared_vtx += nslot_of_v[v + 1] > 1;

  // 5. vertex blocks aligned to instances
  L.vblk_off.push_back(0);
  L.inst_vblk_off.push_back(0);
  for (int64_t i = 0; i < n_inst; ++i) {
    for (int64_t v = inst_vtx_off[i]; v < inst_vtx_off[i + 1]; v += L.vb) {
      L.vblk_off.push_back((int32_t)std::min<int64_t>(v + L.vb, inst_vtx_off[i + 1]));
      L.vblk_inst.push_back((int32_t)i);
    }
    L.inst_vblk_off.push_back((int64_t)L.vblk_inst.size());
  }
  return std::string();
}

}  // namespace apl
