// Host harness for the block partition (apple_amd/csrc/layout.cpp) and the C
// oracle (oracle/c/oracle.c), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py.  It runs the
// partition over edge-case meshes -- empty, one tet, a block window at the
// slot cap, single-tet batch instances, empty segments -- and checks its
// invariants; then the oracle's element passes and PNCG over the one-tet
// and empty meshes.  Exit status 0 = every check passed (the sanitizers
// abort on the first invalid access).
#include "../../apple_amd/csrc/layout.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

extern "C" {
double oracle_snh_fun(int64_t T, const int32_t* cells, const double* dhdX, const double* dV, const double* mu,
                      const double* lam, const double* u);
void oracle_snh_grad(int64_t T, const int32_t* cells, const double* dhdX, const double* dV, const double* mu,
                     const double* lam, const double* u, double* out);
double oracle_snh_pass(int64_t T, int64_t V, const int32_t* cells, const double* dhdX, const double* dV,
                       const double* mu, const double* lam, const double* u, double* grad, double* hdiag);
int oracle_snh_pncg(int64_t T, int64_t V, const int32_t* cells, const double* dhdX, const double* dV,
                    const double* mu, const double* lam, const double* fext, const uint8_t* free_dof, double* x,
                    double rtol, int max_steps, int max_backtracking, double wolfe_delta, int* status_out,
                    double* gnorm_out);
}

static int failures = 0;
#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      std::printf("FAIL %s: ", #cond);    \
      std::printf(__VA_ARGS__);           \
      std::printf("\n");                  \
      ++failures;                         \
    }                                     \
  } while (0)

// Invariants of a built layout (see layout.h).
static void check_layout(const char* name, const std::vector<int32_t>& cells, int64_t T, int64_t V,
                         const std::vector<int64_t>& ito, const apl::Layout& L, const int64_t* seg_off,
                         int64_t n_seg) {
  CHECK(L.blk_tet_off.front() == 0 && L.blk_tet_off.back() == T, "%s: block tet offsets span", name);
  CHECK((int64_t)L.blk_inst.size() == L.nblocks && (int64_t)L.seg_of_blk.size() == L.nblocks, "%s: sizes", name);
  std::vector<int> seen_slot(L.n_slots, 0);
  for (int64_t b = 0; b < L.nblocks; ++b) {
    const int64_t t0 = L.blk_tet_off[b], t1 = L.blk_tet_off[b + 1];
    CHECK(t1 > t0 && t1 - t0 <= (int64_t)L.nth * L.chunks, "%s: block %lld size", name, (long long)b);
    const int inst = L.blk_inst[b];
    CHECK(t0 >= ito[inst] && t1 <= ito[inst + 1], "%s: block %lld straddles an instance", name, (long long)b);
    if (seg_off != nullptr) {
      const int64_t sg = L.seg_of_blk[b];
      CHECK(sg >= 0 && sg < n_seg && t0 >= seg_off[sg] && t1 <= seg_off[sg + 1], "%s: block %lld segment", name,
            (long long)b);
    }
    const int nloc = L.blk_loc_off[b + 1] - L.blk_loc_off[b];
    CHECK(nloc <= L.nloc_cap && nloc <= L.nloc_max, "%s: block %lld nloc %d", name, (long long)b, nloc);
    // every corner's slot maps back to its vertex; positions are a permutation per chunk
    for (int64_t c0 = t0; c0 < t1; c0 += L.nth) {
      const int64_t c1 = std::min<int64_t>(t1, c0 + L.nth);
      std::vector<int> pos_seen(4 * (c1 - c0), 0);
      for (int64_t t = c0; t < c1; ++t)
        for (int c = 0; c < 4; ++c) {
          const int s = L.corner[8 * t + c];
          CHECK(s < nloc, "%s: slot range", name);
          if (s < nloc) CHECK(L.loc_gid[L.blk_loc_off[b] + s] == cells[4 * t + c], "%s: slot vertex", name);
          const int p = L.corner[8 * t + 4 + c];
          CHECK(p < 4 * (c1 - c0), "%s: position range", name);
          if (p < 4 * (c1 - c0)) ++pos_seen[p];
        }
      for (int v : pos_seen) CHECK(v == 1, "%s: positions are not a permutation", name);
    }
  }
  // the vertex-major destinations are a permutation, in ascending block order per vertex
  for (int64_t k = 0; k < L.n_slots; ++k) {
    const int d = L.loc_dst[k];
    CHECK(d >= 0 && d < L.n_slots, "%s: destination range", name);
    if (d >= 0 && d < L.n_slots) ++seen_slot[d];
    const int v = L.loc_gid[k];
    CHECK(d >= L.vtx_slot_off[v] && d < L.vtx_slot_off[v + 1], "%s: destination outside its vertex's run", name);
  }
  for (int c : seen_slot) CHECK(c == 1, "%s: destinations are not a permutation", name);
  CHECK(L.vtx_slot_off.front() == 0 && L.vtx_slot_off.back() == L.n_slots && (int64_t)L.vtx_slot_off.size() == V + 1,
        "%s: vertex CSR", name);
  CHECK(L.vblk_off.back() == V, "%s: vertex blocks", name);
}

static apl::Layout build(const char* name, const std::vector<int32_t>& cells, int64_t V, std::vector<int64_t> ito,
                         std::vector<int64_t> ivo, int chunks, int cap, const std::vector<int64_t>* seg = nullptr) {
  apl::Layout L;
  const int64_t T = (int64_t)cells.size() / 4;
  const int64_t n_seg = seg ? (int64_t)seg->size() - 1 : 0;
  const std::string err = apl::build_layout(cells.data(), T, V, (int64_t)ito.size() - 1, ito.data(), ivo.data(), 256,
                                            chunks, cap, L, n_seg, seg ? seg->data() : nullptr);
  CHECK(err.empty(), "%s: %s", name, err.c_str());
  if (err.empty()) check_layout(name, cells, T, V, ito, L, seg ? seg->data() : nullptr, n_seg);
  return L;
}

// n^3 grid of cells, 5 tets each (any split: the partition does not read geometry)
static std::vector<int32_t> grid_tets(int n, int32_t base = 0) {
  auto id = [&](int x, int y, int z) { return base + (int32_t)(x + (n + 1) * (y + (n + 1) * z)); };
  std::vector<int32_t> c;
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int32_t v[8] = {id(x, y, z),         id(x + 1, y, z),         id(x, y + 1, z),
                              id(x + 1, y + 1, z), id(x, y, z + 1),         id(x + 1, y, z + 1),
                              id(x, y + 1, z + 1), id(x + 1, y + 1, z + 1)};
        const int t[5][4] = {{0, 1, 2, 4}, {3, 2, 1, 7}, {5, 4, 7, 1}, {6, 7, 4, 2}, {1, 2, 4, 7}};
        for (auto& tt : t)
          for (int k : tt) c.push_back(v[k]);
      }
  return c;
}

int main() {
  // empty mesh, vertices without tets
  build("empty", {}, 0, {0, 0}, {0, 0}, 1, 1024);
  build("no tets, 5 vertices", {}, 5, {0, 0}, {0, 5}, 1, 1024);
  // one tet
  build("one tet", {0, 1, 2, 3}, 4, {0, 1}, {0, 4}, 1, 1024);
  // a grid in order and shuffled (windows at the cap), several caps / chunks
  const int n = 12;
  std::vector<int32_t> g = grid_tets(n);
  const int64_t V = (int64_t)(n + 1) * (n + 1) * (n + 1), T = (int64_t)g.size() / 4;
  for (int chunks : {1, 2, 4})
    for (int cap : {1024, 478, 131, 4}) {
      char nm[64];
      std::snprintf(nm, sizeof nm, "grid c%d cap%d", chunks, cap);
      apl::Layout L = build(nm, g, V, {0, T}, {0, V}, chunks, cap);
      (void)L;
    }
  std::vector<int64_t> perm(T);
  for (int64_t t = 0; t < T; ++t) perm[t] = t;
  std::shuffle(perm.begin(), perm.end(), std::mt19937(3));
  std::vector<int32_t> gs(g.size());
  for (int64_t t = 0; t < T; ++t)
    for (int c = 0; c < 4; ++c) gs[4 * t + c] = g[4 * perm[t] + c];
  apl::Layout Ls = build("shuffled cap 478", gs, V, {0, T}, {0, V}, 1, 478);
  CHECK(Ls.nloc_max >= 475 && Ls.nloc_max <= 478, "shuffled: window %d not at the cap", Ls.nloc_max);
  build("shuffled cap 1024 x4", gs, V, {0, T}, {0, V}, 4, 1024);
  // a batch of one-tet instances and an empty instance
  {
    std::vector<int32_t> c;
    std::vector<int64_t> ito{0}, ivo{0};
    for (int i = 0; i < 7; ++i) {
      for (int k = 0; k < 4; ++k) c.push_back(4 * i + k);
      ito.push_back(i + 1);
      ivo.push_back(4 * (i + 1));
    }
    ito.push_back(7);  // an instance without tets or vertices
    ivo.push_back(28);
    build("one-tet instances", c, 28, ito, ivo, 1, 1024);
  }
  // segments: three potentials over one grid, one empty, across two instances
  {
    std::vector<int32_t> a = grid_tets(4), b = grid_tets(3, 125);
    std::vector<int32_t> c;
    const int64_t Ta = (int64_t)a.size() / 4, Tb = (int64_t)b.size() / 4;
    for (int s = 0; s < 3; ++s) c.insert(c.end(), a.begin(), a.end());
    for (int s = 0; s < 2; ++s) c.insert(c.end(), b.begin(), b.end());
    std::vector<int64_t> seg{0, Ta, 2 * Ta, 3 * Ta, 3 * Ta, 3 * Ta + Tb, 3 * Ta + 2 * Tb};
    std::vector<int64_t> ito{0, 3 * Ta, 3 * Ta + 2 * Tb}, ivo{0, 125, 125 + 64};
    build("segments", c, 189, ito, ivo, 2, 1024, &seg);
  }
  // rejected inputs return an error, touching nothing out of range
  {
    apl::Layout L;
    std::vector<int32_t> c{0, 1, 2, 9};
    int64_t ito[2] = {0, 1}, ivo[2] = {0, 4};
    CHECK(!apl::build_layout(c.data(), 1, 4, 1, ito, ivo, 256, 1, 1024, L).empty(), "out-of-range vertex accepted");
    int64_t ivo2[2] = {0, 3};
    CHECK(!apl::build_layout(c.data(), 0, 4, 1, ito, ivo2, 256, 1, 1024, L).empty(), "bad offsets accepted");
    std::vector<int32_t> c1{0, 1, 2, 3};
    int64_t seg_bad[3] = {0, 2, 1};
    CHECK(!apl::build_layout(c1.data(), 1, 4, 1, ito, ivo, 256, 1, 1024, L, 2, seg_bad).empty(),
          "bad segments accepted");
  }
  // the C oracle over the one-tet and empty meshes
  {
    const int32_t cells[4] = {0, 1, 2, 3};
    const double dhdX[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    const double dV[1] = {1.0 / 6}, mu[1] = {1.0}, lam[1] = {1.0};
    double u[12] = {0, 0, 0, 0.1, 0, 0, 0, 0, 0, 0, 0, 0.05}, g[12], h[12];
    const double e = oracle_snh_pass(1, 4, cells, dhdX, dV, mu, lam, u, g, h);
    CHECK(e > 0, "one-tet energy %g", e);
    CHECK(oracle_snh_pass(0, 4, cells, dhdX, dV, mu, lam, u, g, h) == 0.0, "empty energy");
    double f[12] = {0}, x[12] = {0, 0, 0, 0.1, 0, 0, 0, 0, 0, 0, 0, 0.05};
    uint8_t fr[12] = {0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};
    int status = 0;
    double gn = 0;
    oracle_snh_pncg(1, 4, cells, dhdX, dV, mu, lam, f, fr, x, 1e-10, 500, 8, 0.1, &status, &gn);
    CHECK(status == 1, "one-tet PNCG status %d", status);
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
