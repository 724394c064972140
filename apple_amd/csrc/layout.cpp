// Host-side block partition; see layout.h.
#include "layout.h"

#include <algorithm>
#include <cstdio>

namespace apl {

std::string build_layout(const int32_t* cells, int64_t T, int64_t V, int64_t n_inst,
                         const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int nth,
                         int chunks, int nloc_cap, Layout& L) {
  char msg[256];
  if (T < 0 || V < 0 || n_inst < 1) return "invalid mesh sizes";
  if (nth != 256) return "threads per block must be 256";
  if (chunks < 1 || (int64_t)nth * chunks > 8192) return "chunks must be in [1, 32]";
  if (nloc_cap < 4 || nloc_cap > 1024) return "nloc_cap must be in [4, 1024]";
  if (T >= (int64_t(1) << 31) || V >= (int64_t(1) << 31)) return "mesh too large for int32 indices";
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
  for (int64_t i = 0; i < n_inst; ++i) {
    int64_t t = inst_tet_off[i], t1 = inst_tet_off[i + 1];
    while (t < t1) {
      int64_t b = (int64_t)L.blk_inst.size();
      int nloc = 0, nt = 0;
      while (t < t1 && nt < max_tets) {
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

  // 2. number of distinct blocks touching each vertex
  std::vector<int32_t> nblk_of_v(V, 0);
  std::fill(stamp.begin(), stamp.end(), -1);
  for (int64_t b = 0; b < L.nblocks; ++b)
    for (int64_t t = L.blk_tet_off[b]; t < L.blk_tet_off[b + 1]; ++t)
      for (int c = 0; c < 4; ++c) {
        int64_t v = cells[t * 4 + c];
        if (stamp[v] != b) {
          stamp[v] = b;
          ++nblk_of_v[v];
        }
      }

  // 3. local slots (owned first, then shared, each in first-touch order),
  //    corner slots, in-block CSR and partial indices
  L.slot.resize(4 * T);
  L.loc_inc.resize(4 * T);
  L.blk_loc_off.push_back(0);
  std::vector<int32_t> local_of(V, -1);
  std::fill(stamp.begin(), stamp.end(), -1);
  std::vector<int32_t> order;
  std::vector<int32_t> cnt;
  int64_t n_partials = 0;
  std::vector<std::pair<int32_t, int32_t>> sh_pairs;  // (vertex, partial index)
  for (int64_t b = 0; b < L.nblocks; ++b) {
    int64_t t0 = L.blk_tet_off[b], t1 = L.blk_tet_off[b + 1];
    order.clear();
    for (int64_t t = t0; t < t1; ++t)
      for (int c = 0; c < 4; ++c) {
        int64_t v = cells[t * 4 + c];
        if (stamp[v] != b) {
          stamp[v] = b;
          order.push_back((int32_t)v);
        }
      }
    int nloc = (int)order.size();
    int nown = 0;
    for (int32_t v : order) nown += nblk_of_v[v] == 1;
    int io = 0, is = nown;
    int64_t base = (int64_t)L.loc_gid.size();
    L.loc_gid.resize(base + nloc);
    for (int32_t v : order) {
      int s = nblk_of_v[v] == 1 ? io++ : is++;
      local_of[v] = s;
      L.loc_gid[base + s] = v;
    }
    L.blk_nown.push_back(nown);
    L.blk_sh_base.push_back((int32_t)n_partials);
    for (int s = nown; s < nloc; ++s) sh_pairs.emplace_back(L.loc_gid[base + s], (int32_t)(n_partials + (s - nown)));
    n_partials += nloc - nown;
    if (n_partials >= (int64_t(1) << 31)) return "too many partial sums";
    L.nloc_max = std::max(L.nloc_max, nloc);
    // corner slots and CSR by slot
    cnt.assign(nloc + 1, 0);
    for (int64_t t = t0; t < t1; ++t)
      for (int c = 0; c < 4; ++c) {
        int s = local_of[cells[t * 4 + c]];
        L.slot[t * 4 + c] = (uint16_t)s;
        ++cnt[s + 1];
      }
    for (int s = 0; s < nloc; ++s) cnt[s + 1] += cnt[s];
    size_t ob = L.loc_inc_off.size();
    for (int s = 0; s <= nloc; ++s) L.loc_inc_off.push_back((uint16_t)cnt[s]);
    for (int64_t t = t0; t < t1; ++t)
      for (int c = 0; c < 4; ++c) {
        int s = local_of[cells[t * 4 + c]];
        int pos = cnt[s]++;
        L.loc_inc[4 * t0 + pos] = (uint16_t)((t - t0) * 4 + c);
      }
    (void)ob;
    L.blk_loc_off.push_back((int32_t)(base + nloc));
  }
  L.n_partials = n_partials;

  // 4. shared-vertex CSR over partials, ascending vertex, ascending partial.
  //    Vertices no tet touches are listed too, with no partials, so that the
  //    vertex pass writes their (zero) sums.
  std::vector<int32_t> npart(V + 1, 0);
  for (const auto& pr : sh_pairs) ++npart[pr.first + 1];
  std::vector<int32_t> start(V + 1, 0);
  for (int64_t v = 0; v < V; ++v) start[v + 1] = start[v] + npart[v + 1];
  L.sh_src.assign(sh_pairs.size(), 0);
  {
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (const auto& pr : sh_pairs) L.sh_src[fill[pr.first]++] = pr.second;  // block order kept
  }
  L.sh_off.push_back(0);
  for (int64_t v = 0; v < V; ++v) {
    if (npart[v + 1] > 0 || nblk_of_v[v] == 0) {
      L.sh_vtx.push_back((int32_t)v);
      L.sh_off.push_back(start[v + 1]);
    }
  }
  L.n_shared_vtx = (int64_t)L.sh_vtx.size();

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
