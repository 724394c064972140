// C ABI: handles, setup upload, element evaluation, error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/apple_amd.h"
#include "internal.h"
#include "tetmath.h"

namespace apl {

static thread_local std::string g_err;
void set_error(const std::string& s) { g_err = s; }
// 0: block partitions on the device where supported, 1: on the host
static int g_layout_mode = 0;

template <class T>
static int upload(apl_mesh* m, T** dst, const T* src, size_t n) {
  *dst = nullptr;
  if (n == 0) return 0;
  APL_CHECK_HIP(hipMalloc((void**)dst, n * sizeof(T)));
  m->device_bytes += n * sizeof(T);
  APL_CHECK_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

static int alloc_bytes(apl_mesh* m, void** dst, size_t n) {
  *dst = nullptr;
  if (n == 0) return 0;
  APL_CHECK_HIP(hipMalloc(dst, n));
  m->device_bytes += n;
  APL_CHECK_HIP(hipMemset(*dst, 0, n));
  return 0;
}

static void free_mesh(apl_mesh* m) {
  drop_hip_error(hipSetDevice(m->device));
  DevLayout& d = m->d;
  void* ptrs[] = {d.dminv,        d.dV,           d.corner,      d.blk_tet_off,  d.blk_loc_off,
                  d.blk_inst,     d.inst_blk_off, d.loc_gid,     d.blk_co_off,   d.chunk_off,
                  d.vtx_slot_off, d.loc_dst,      d.vblk_off,    d.vblk_inst,    d.inst_vblk_off,
                  d.inst_vtx_off, d.partial,      d.blk_red,     d.vblk_dots,    d.law_blk[0],
                  d.law_blk[1],   d.law_blk[2]};
  for (void* p : ptrs)
    if (p) drop_hip_error(hipFree(p));
  if (m->bench_exec) drop_hip_error(hipGraphExecDestroy(m->bench_exec));
  delete m;
}

void mesh_release(apl_mesh* m) {
  if (--m->refs == 0) free_mesh(m);
}

static unsigned to_internal(unsigned what) {
  unsigned w = 0;
  if (what & APL_ENERGY) w |= W_E;
  if (what & APL_GRAD) w |= W_G;
  if (what & APL_HESS_DIAG) w |= W_H;
  if (what & APL_HESS_PROD) w |= W_HP;
  if (what & APL_HESS_QUAD) w |= W_HQ;
  if (what & APL_MIXED) w |= W_MIX;
  return w;
}

}  // namespace apl

using namespace apl;

namespace apl {

bool material_ok(const apl_mesh* m, int material) {
  if (m->model) return material == APL_MODEL;
  return material == APL_SNH || material == APL_ARAP || material == APL_SNH_MUSCLE;
}

bool needs_lambda(const apl_mesh* m, int material) {
  if (material == APL_MODEL) return m->law_nblk[MAT_SNH] + m->law_nblk[MAT_MUSCLE] > 0;
  return material != APL_ARAP;
}

}  // namespace apl

extern "C" {

const char* apl_last_error(void) { return g_err.c_str(); }

int apl_layout_mode(int mode) {
  if (mode != 0 && mode != 1) {
    set_error("layout mode must be 0 (device) or 1 (host)");
    return -1;
  }
  const int prev = g_layout_mode;
  g_layout_mode = mode;
  return prev;
}

int apl_version(void) { return 10000; }

}  // extern "C"

namespace apl {

// Everything of a mesh except the per-tet element data (inverse rest shapes
// and volumes): validation, block partition, layout upload, workspaces.
static int create_common(int dtype, int64_t n_tets, int64_t n_verts, const int32_t* cells, int64_t n_inst,
                         const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int chunks, int device,
                         apl_mesh** out, int n_seg = 0, const int32_t* seg_law = nullptr,
                         const int64_t* seg_tet_off = nullptr) {
  if (out == nullptr) {
    set_error("out is NULL");
    return -1;
  }
  *out = nullptr;
  if (dtype != APL_F32 && dtype != APL_F64) {
    set_error("dtype must be APL_F32 or APL_F64");
    return -1;
  }
  // 32-bit byte offsets in the kernels: at most 2^27 tets and vertices
  if (n_tets < 0 || n_verts < 0 || n_tets >= (int64_t(1) << 27) || n_verts >= (int64_t(1) << 27)) {
    set_error("mesh too large or negative size (at most 2^27 tets and vertices)");
    return -1;
  }
  // the element kernels address the SoA inverse rest shapes (9 rows of T)
  // by a 32-bit byte offset from each block's start: 8 T + 8192 values of
  // the mesh dtype must stay below 2^32 bytes (2^26 - 1024 FP64 tets)
  const uint64_t rs_bytes = dtype == APL_F64 ? 8 : 4;
  if ((8 * (uint64_t)n_tets + 8192) * rs_bytes >= (uint64_t(1) << 32)) {
    set_error("mesh too large for 32-bit element offsets: at most 2^26 - 1024 FP64 tets (2^27 - 1024 FP32)");
    return -1;
  }
  if (n_tets > 0 && cells == nullptr) {
    set_error("cells must be non-NULL");
    return -1;
  }
  // LDS budget of the element workgroups: FP64 blocks hold at most 478 slots and one chunk
  const int nloc_cap = nloc_limit(dtype);
  if (chunks < 1 || dtype == APL_F64) chunks = 1;
  // the partition runs on the device unless the host pass is asked for (or
  // the blocks span more chunks than the device pass takes)
  const bool dev_layout = g_layout_mode == 0 && device_layout_supported(chunks, nloc_cap);
  if (!dev_layout)
    for (int64_t i = 0; i < 4 * n_tets; ++i)
      if (cells[i] < 0 || cells[i] >= n_verts) {
        set_error("cell index out of range [0, n_verts)");
        return -1;
      }
  int64_t one_t[2] = {0, n_tets}, one_v[2] = {0, n_verts};
  if (n_inst <= 1 || inst_tet_off == nullptr || inst_vtx_off == nullptr) {
    if (n_inst > 1) {
      set_error("instance offsets are required for n_inst > 1");
      return -1;
    }
    n_inst = 1;
    inst_tet_off = one_t;
    inst_vtx_off = one_v;
  }
  APL_CHECK_HIP(hipSetDevice(device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  apl_mesh* m = new apl_mesh();
  m->dtype = dtype;
  m->device = device;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    m->n_cu = ncu;
  const int64_t n_seg_total = seg_tet_off != nullptr ? n_inst * (int64_t)n_seg : 0;
  std::string err = dev_layout ? build_layout_device(m, cells, n_tets, n_verts, n_inst, inst_tet_off, inst_vtx_off,
                                                     chunks, nloc_cap, n_seg_total, seg_tet_off)
                               : build_layout(cells, n_tets, n_verts, n_inst, inst_tet_off, inst_vtx_off, NTH, chunks,
                                              nloc_cap, m->L, n_seg_total, seg_tet_off);
  if (!err.empty()) {
    set_error(err);
    free_mesh(m);
    return -1;
  }
  Layout& L = m->L;
  DevLayout& d = m->d;
  const size_t rs = dtype == APL_F64 ? 8 : 4;
  // the element kernels address the slot partials by 32-bit byte offsets
  if ((uint64_t)L.n_slots * MAX_CH * rs >= (uint64_t(1) << 32)) {
    set_error("too many block-local vertex slots for 32-bit partial offsets (reorder the mesh for locality)");
    free_mesh(m);
    return -1;
  }
  int rc = 0;
  rc |= alloc_bytes(m, &d.dminv, 9 * (size_t)n_tets * rs);
  rc |= alloc_bytes(m, &d.dV, (size_t)n_tets * rs);
  if (!dev_layout) {
    rc |= upload(m, &d.corner, L.corner.data(), L.corner.size());
    rc |= upload(m, &d.blk_tet_off, L.blk_tet_off.data(), L.blk_tet_off.size());
    rc |= upload(m, &d.blk_loc_off, L.blk_loc_off.data(), L.blk_loc_off.size());
    rc |= upload(m, &d.blk_inst, L.blk_inst.data(), L.blk_inst.size());
    rc |= upload(m, &d.inst_blk_off, L.inst_blk_off.data(), L.inst_blk_off.size());
    rc |= upload(m, &d.loc_gid, L.loc_gid.data(), L.loc_gid.size());
    rc |= upload(m, &d.blk_co_off, L.blk_co_off.data(), L.blk_co_off.size());
    rc |= upload(m, &d.chunk_off, L.chunk_off.data(), L.chunk_off.size());
    rc |= upload(m, &d.vtx_slot_off, L.vtx_slot_off.data(), L.vtx_slot_off.size());
    rc |= upload(m, &d.loc_dst, L.loc_dst.data(), L.loc_dst.size());
    rc |= upload(m, &d.vblk_off, L.vblk_off.data(), L.vblk_off.size());
    rc |= upload(m, &d.vblk_inst, L.vblk_inst.data(), L.vblk_inst.size());
    rc |= upload(m, &d.inst_vblk_off, L.inst_vblk_off.data(), L.inst_vblk_off.size());
  }
  rc |= upload(m, &d.inst_vtx_off, inst_vtx_off, (size_t)n_inst + 1);
  rc |= alloc_bytes(m, &d.partial, (size_t)MAX_CH * std::max<int64_t>(L.n_slots, 1) * rs);
  rc |= alloc_bytes(m, (void**)&d.blk_red, RED_STRIDE * (size_t)std::max<int64_t>(L.nblocks, 1) * sizeof(double));
  rc |= alloc_bytes(m, (void**)&d.vblk_dots,
                    (size_t)N_DOTS * std::max<size_t>(L.vblk_inst.size(), 1) * sizeof(double));
  if (seg_tet_off != nullptr) {
    // model mesh: the blocks of each law, ascending (segment s of every
    // instance evaluates seg_law[s])
    m->model = 1;
    std::vector<int32_t> lists[N_MAT];
    for (int64_t b = 0; b < L.nblocks; ++b) lists[seg_law[L.seg_of_blk[b] % n_seg]].push_back((int32_t)b);
    for (int law = 0; law < N_MAT; ++law) {
      m->law_nblk[law] = (int64_t)lists[law].size();
      rc |= upload(m, &d.law_blk[law], lists[law].data(), lists[law].size());
    }
  }
  if (rc) {
    free_mesh(m);
    return -1;
  }
  // free the host-side tables the device now holds; keep the sizes
  L.corner = std::vector<uint16_t>();
  L.chunk_off = std::vector<uint16_t>();
  L.loc_gid = std::vector<int32_t>();
  L.loc_dst = std::vector<int32_t>();
  L.vtx_slot_off = std::vector<int32_t>();
  *out = m;
  return 0;
}

}  // namespace apl

namespace apl {

static int mesh_create_impl(int dtype, int64_t n_tets, int64_t n_verts, const int32_t* cells, const double* dminv,
                            const double* dv, int64_t n_inst, const int64_t* inst_tet_off,
                            const int64_t* inst_vtx_off, int chunks, int device, apl_mesh** out, int n_seg,
                            const int32_t* seg_law, const int64_t* seg_tet_off) {
  if (n_tets > 0 && (dminv == nullptr || dv == nullptr)) {
    set_error("cells, dminv and dv must be non-NULL");
    if (out != nullptr) *out = nullptr;
    return -1;
  }
  apl_mesh* m = nullptr;
  if (create_common(dtype, n_tets, n_verts, cells, n_inst, inst_tet_off, inst_vtx_off, chunks, device, &m, n_seg,
                    seg_law, seg_tet_off))
    return -1;
  // element data, SoA, in the mesh precision
  const size_t rs = dtype == APL_F64 ? 8 : 4;
  std::vector<char> buf(9 * (size_t)n_tets * rs);
  for (int64_t t = 0; t < n_tets; ++t)
    for (int k = 0; k < 9; ++k) {
      double v = dminv[t * 9 + k];
      if (rs == 8)
        reinterpret_cast<double*>(buf.data())[k * n_tets + t] = v;
      else
        reinterpret_cast<float*>(buf.data())[k * n_tets + t] = (float)v;
    }
  std::vector<char> bv((size_t)n_tets * rs);
  for (int64_t t = 0; t < n_tets; ++t) {
    if (rs == 8)
      reinterpret_cast<double*>(bv.data())[t] = dv[t];
    else
      reinterpret_cast<float*>(bv.data())[t] = (float)dv[t];
  }
  if ((!buf.empty() && hipMemcpy(m->d.dminv, buf.data(), buf.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      (!bv.empty() && hipMemcpy(m->d.dV, bv.data(), bv.size(), hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("upload of the element data failed");
    (void)hipGetLastError();
    free_mesh(m);
    return -1;
  }
  *out = m;
  return 0;
}

}  // namespace apl

extern "C" {

int apl_mesh_create(int dtype, int64_t n_tets, int64_t n_verts, const int32_t* cells, const double* dminv,
                    const double* dv, int64_t n_inst, const int64_t* inst_tet_off, const int64_t* inst_vtx_off,
                    int chunks, int device, apl_mesh** out) {
  return mesh_create_impl(dtype, n_tets, n_verts, cells, dminv, dv, n_inst, inst_tet_off, inst_vtx_off, chunks,
                          device, out, 0, nullptr, nullptr);
}

int apl_mesh_create_from_points(int dtype, int64_t n_tets, int64_t n_verts, const int32_t* cells,
                                const double* points, const double* fraction, int64_t n_inst,
                                const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int chunks, int device,
                                apl_mesh** out) {
  if (n_tets > 0 && points == nullptr) {
    set_error("points must be non-NULL");
    if (out != nullptr) *out = nullptr;
    return -1;
  }
  apl_mesh* m = nullptr;
  if (create_common(dtype, n_tets, n_verts, cells, n_inst, inst_tet_off, inst_vtx_off, chunks, device, &m)) return -1;
  if (n_tets > 0 && region_from_points(m, cells, points, fraction) != 0) {
    free_mesh(m);
    return -1;
  }
  *out = m;
  return 0;
}

int apl_model_create(int dtype, int64_t n_tets, int64_t n_verts, const int32_t* cells, const double* dminv,
                     const double* dv, int64_t n_inst, const int64_t* inst_tet_off, const int64_t* inst_vtx_off,
                     int n_seg, const int32_t* seg_law, const int64_t* seg_tet_off, int chunks, int device,
                     apl_mesh** out) {
  if (out != nullptr) *out = nullptr;
  if (n_seg < 1 || n_seg > 1024 || seg_law == nullptr || seg_tet_off == nullptr) {
    set_error("a model needs 1..1024 segments with their laws and tet offsets");
    return -1;
  }
  for (int k = 0; k < n_seg; ++k)
    if (seg_law[k] != APL_SNH && seg_law[k] != APL_ARAP && seg_law[k] != APL_SNH_MUSCLE) {
      set_error("segment laws must be APL_SNH, APL_ARAP or APL_SNH_MUSCLE");
      return -1;
    }
  const int64_t ni = n_inst <= 1 || inst_tet_off == nullptr ? 1 : n_inst;
  int64_t one_t[2] = {0, n_tets};
  const int64_t* ito = ni == 1 ? one_t : inst_tet_off;
  for (int64_t i = 0; i < ni; ++i)
    if (seg_tet_off[i * n_seg] != ito[i]) {
      set_error("segment 0 of every instance must start at the instance's first tet");
      return -1;
    }
  if (seg_tet_off[ni * n_seg] != n_tets) {
    set_error("segment offsets must end at n_tets");
    return -1;
  }
  return mesh_create_impl(dtype, n_tets, n_verts, cells, dminv, dv, n_inst, inst_tet_off, inst_vtx_off, chunks,
                          device, out, n_seg, seg_law, seg_tet_off);
}

int apl_mesh_region(apl_mesh* mesh, void* dminv, void* dv, void* stream) {
  if (mesh == nullptr || (mesh->L.T > 0 && dminv == nullptr && dv == nullptr)) {
    set_error("mesh and at least one output must be non-NULL");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(mesh->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  return region_readback(mesh, dminv, dv, (hipStream_t)stream);
}

int apl_mesh_destroy(apl_mesh* mesh) {
  if (mesh == nullptr) return 0;
  mesh_release(mesh);
  return 0;
}

int apl_mesh_set_schedule(apl_mesh* mesh, int grid_pct) {
  if (mesh == nullptr || (grid_pct != 0 && (grid_pct < 25 || grid_pct > 800))) {
    set_error("grid_pct must be 0 (the per-pass default) or in [25, 800]");
    return -1;
  }
  mesh->grid_pct = grid_pct;
  for (auto& row : mesh->grid_cache)
    for (int& g : row) g = 0;
  return 0;
}

int apl_mesh_info(const apl_mesh* mesh, apl_mesh_info_t* info) {
  if (mesh == nullptr || info == nullptr) {
    set_error("NULL argument");
    return -1;
  }
  apl_mesh* m = const_cast<apl_mesh*>(mesh);
  drop_hip_error(hipSetDevice(m->device));
  const Layout& L = m->L;
  info->n_tets = L.T;
  info->n_verts = L.V;
  info->n_inst = L.n_inst;
  info->n_blocks = L.nblocks;
  info->n_partials = L.n_slots;
  info->n_shared_verts = L.n_shared_vtx;
  info->n_vertex_blocks = (int64_t)L.vblk_inst.size();
  info->nloc_max = L.nloc_max;
  info->chunks = L.chunks;
  info->dtype = m->dtype;
  info->elem_grid_snh = elem_grid(m, MAT_SNH, W_E | W_G | W_H);
  info->elem_grid_arap = elem_grid(m, MAT_ARAP, W_E | W_G | W_H);
  info->device_bytes = (int64_t)m->device_bytes;
  return 0;
}

}  // extern "C"

namespace apl {

size_t eval_scratch_bytes(const apl_mesh* m) {
  const size_t part = (size_t)MAX_CH * (size_t)std::max<int64_t>(m->L.n_slots, 1) * real_size(m);
  const size_t red = RED_STRIDE * (size_t)std::max<int64_t>(m->L.nblocks, 1) * sizeof(double);
  return ((part + 255) / 256) * 256 + red;
}

int elem_eval_impl(apl_mesh* m, int material, const void* mu, const void* lam, const void* u, const void* p,
                   unsigned what, double* energy, void* grad, void* hess_diag, void* hess_prod, double* hess_quad,
                   void* dmu, void* dlam, hipStream_t s, const Workspace* ws) {
  if (m == nullptr) {
    set_error("mesh is NULL");
    return -1;
  }
  if (!material_ok(m, material)) {
    set_error(m->model ? "a model mesh takes material APL_MODEL"
                       : "material must be APL_SNH, APL_ARAP or APL_SNH_MUSCLE");
    return -1;
  }
  const unsigned w = to_internal(what);
  const int acc = (what & APL_ACCUMULATE) ? 1 : 0;
  if (w == 0) return 0;
  if (m->L.T > 0 && (mu == nullptr || u == nullptr || (needs_lambda(m, material) && lam == nullptr))) {
    set_error("mu, lam (SNH, SNH_MUSCLE) and u must be non-NULL");
    return -1;
  }
  if (m->L.T > 0 && (w & (W_HP | W_HQ | W_MIX)) && p == nullptr) {
    set_error("p must be non-NULL for hess_prod, hess_quad and mixed derivatives");
    return -1;
  }
  const bool has_v = m->L.V > 0, has_t = m->L.T > 0;
  if (((w & W_E) && !energy) || ((w & W_G) && has_v && !grad) || ((w & W_H) && has_v && !hess_diag) ||
      ((w & W_HP) && has_v && !hess_prod) || ((w & W_HQ) && !hess_quad) ||
      ((w & W_MIX) && has_t && (!dmu || (needs_lambda(m, material) && !dlam)))) {
    set_error("a requested output pointer is NULL");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(m->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  const unsigned egh = w & (W_E | W_G | W_H);
  if (egh) {
    // instantiated pass variants: E, G, H, E|G|H; other subsets run as separate passes
    unsigned passes[3];
    int np = 0;
    if (egh == (W_E | W_G | W_H) || egh == W_E || egh == W_G || egh == W_H)
      passes[np++] = egh;
    else
      for (unsigned b : {W_E, W_G, W_H})
        if (egh & b) passes[np++] = b;
    for (int i = 0; i < np; ++i) {
      const unsigned pw = passes[i];
      if (launch_elem(m, material, pw, mu, lam, u, nullptr, nullptr, nullptr, grad, hess_diag, acc, 0, s, 0, nullptr,
                      ws))
        return -1;
      if (launch_finish(m, pw, grad, hess_diag, (pw & W_E) ? energy : nullptr, nullptr, acc, nullptr, s, ws))
        return -1;
    }
  }
  if (w & W_HP) {
    if (launch_elem(m, material, W_HP, mu, lam, u, p, nullptr, nullptr, hess_prod, nullptr, acc, 0, s, 0, nullptr, ws))
      return -1;
    if (launch_finish(m, W_HP, hess_prod, nullptr, nullptr, nullptr, acc, nullptr, s, ws)) return -1;
  }
  if (w & W_HQ) {
    if (launch_elem(m, material, W_HQ, mu, lam, u, p, nullptr, nullptr, nullptr, nullptr, acc, 0, s, 0, nullptr, ws))
      return -1;
    if (launch_finish(m, W_HQ, nullptr, nullptr, nullptr, hess_quad, acc, nullptr, s, ws)) return -1;
  }
  if (w & W_MIX) {
    if (launch_elem(m, material, W_MIX, mu, lam, u, p, nullptr, nullptr, dmu, dlam, acc, 0, s, 0, nullptr, ws))
      return -1;
  }
  return 0;
}

}  // namespace apl

extern "C" {

int apl_elem_eval(apl_mesh* m, int material, const void* mu, const void* lam, const void* u, const void* p,
                  unsigned what, double* energy, void* grad, void* hess_diag, void* hess_prod, double* hess_quad,
                  void* dmu, void* dlam, void* stream) {
  return elem_eval_impl(m, material, mu, lam, u, p, what, energy, grad, hess_diag, hess_prod, hess_quad, dmu, dlam,
                        (hipStream_t)stream, nullptr);
}

int apl_mesh_set_activation(apl_mesh* mesh, const void* activation) {
  if (mesh == nullptr) {
    set_error("mesh is NULL");
    return -1;
  }
  mesh->activation = activation;
  return 0;
}

int apl_elem_mixed_activation(apl_mesh* m, const void* mu, const void* lam, const void* u, const void* p,
                              void* dact, void* dmu, void* dlam, int accumulate, void* stream) {
  if (m == nullptr) {
    set_error("mesh is NULL");
    return -1;
  }
  if (m->L.T == 0) return 0;
  if (mu == nullptr || lam == nullptr || u == nullptr || p == nullptr || dact == nullptr ||
      ((dmu == nullptr) != (dlam == nullptr))) {
    set_error("mu, lam, u, p and dact must be non-NULL, and dmu / dlam both given or both NULL");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(m->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  const unsigned w = dmu != nullptr ? (W_MIX | W_MIXA) : W_MIXA;
  // a model mesh: dact of its muscle-law tets, dmu / dlam of every tet
  return launch_elem(m, m->model ? MAT_MODEL : MAT_MUSCLE, w, mu, lam, u, p, nullptr, nullptr, dmu, dlam,
                     accumulate ? 1 : 0, 0, (hipStream_t)stream, 0, dact);
}

}  // extern "C"

namespace apl {

// Timing events of the measurement entry points: created on a device's
// first call and kept for the process (concurrent calls on one device race).
static hipEvent_t* bench_events(int device) {
  static std::mutex mu;
  static hipEvent_t cache[64][4];
  if (device < 0 || device >= 64) {
    set_error("device index out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  for (int k = 0; k < 4; ++k)
    if (cache[device][k] == nullptr && hipEventCreate(&cache[device][k]) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipEventCreate failed");
      return nullptr;
    }
  return cache[device];
}

// Spin until `ev` has completed (a blocking wait adds its wake-up latency).
static int spin_wait(hipEvent_t ev) {
  hipError_t q;
  while ((q = hipEventQuery(ev)) == hipErrorNotReady) {
  }
  if (q != hipSuccess) {
    (void)hipGetLastError();
    set_error(std::string("hipEventQuery: ") + hipGetErrorString(q));
    return -1;
  }
  return 0;
}

}  // namespace apl

extern "C" {

int apl_bench_elem_pass(apl_mesh* m, int material, unsigned what, const void* mu, const void* lam, const void* u,
                        const void* p, double* energy, void* out_a, void* out_b, int iters, double* ms_kernel,
                        double* ms_pass, void* stream) {
  if (m == nullptr || iters < 1 || (ms_kernel == nullptr && ms_pass == nullptr)) {
    set_error("invalid argument");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(m->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  const unsigned w = to_internal(what);
  if (w != (W_E | W_G | W_H) && w != W_HQ && w != W_HP) {
    set_error("bench pass must be ENERGY|GRAD|HESS_DIAG, HESS_QUAD or HESS_PROD");
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  // ms_kernel (if non-NULL): iters element launches back to back between two
  // events (the kernel's duration plus the launch gap); ms_pass (if
  // non-NULL): iters full passes (element kernel + finish kernel) the same way.
  // The events are created once per device and kept, so that a timed call
  // holds the launches and little else.
  hipEvent_t* ev = bench_events(m->device);
  if (ev == nullptr) return -1;
  int rc = 0;
  // only the events a caller reads are recorded (each is a marker packet on
  // the queue); the host then spins on the last one rather than sleeping in
  // a blocking wait, so the call returns as soon as the work is done
  if (ms_kernel != nullptr) {
    APL_CHECK_HIP(hipEventRecord(ev[0], s));
    for (int it = 0; it < iters && rc == 0; ++it)
      rc = launch_elem(m, material, w, mu, lam, u, p, nullptr, nullptr, out_a, out_b, 0, 0, s);
    APL_CHECK_HIP(hipEventRecord(ev[1], s));
  }
  if (ms_pass != nullptr && rc == 0) {
    APL_CHECK_HIP(hipEventRecord(ev[2], s));
    for (int it = 0; it < iters && rc == 0; ++it) {
      rc = launch_elem(m, material, w, mu, lam, u, p, nullptr, nullptr, out_a, out_b, 0, 0, s);
      if (rc == 0)
        rc = launch_finish(m, w, out_a, out_b, (w & W_E) ? energy : nullptr, (w & W_HQ) ? energy : nullptr, 0,
                           nullptr, s);
    }
    APL_CHECK_HIP(hipEventRecord(ev[3], s));
  }
  if (rc == 0) {
    if (spin_wait(ms_pass != nullptr ? ev[3] : ev[1])) return -1;
    float a = 0, b = 0;
    if (ms_kernel != nullptr) APL_CHECK_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
    if (ms_pass != nullptr) APL_CHECK_HIP(hipEventElapsedTime(&b, ev[2], ev[3]));
    if (ms_kernel != nullptr) *ms_kernel = a / iters;
    if (ms_pass != nullptr) *ms_pass = b / iters;
  }
  return rc;
}

int apl_bench_elem_graph(apl_mesh* m, int material, unsigned what, const void* mu, const void* lam, const void* u,
                         const void* p, double* energy, void* out_a, void* out_b, int iters, double* ms_pass,
                         void* stream) {
  if (m == nullptr || iters < 1 || ms_pass == nullptr) {
    set_error("invalid argument");
    return -1;
  }
  APL_CHECK_HIP(hipSetDevice(m->device));
  (void)hipGetLastError();  // drop errors left by unrelated earlier HIP calls
  const unsigned w = to_internal(what);
  if (w != (W_E | W_G | W_H) && w != W_HQ && w != W_HP) {
    set_error("bench pass must be ENERGY|GRAD|HESS_DIAG, HESS_QUAD or HESS_PROD");
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t* ev = bench_events(m->device);
  if (ev == nullptr) return -1;
  const uintptr_t key[12] = {(uintptr_t)material, (uintptr_t)w, (uintptr_t)mu, (uintptr_t)lam, (uintptr_t)u,
                             (uintptr_t)p, (uintptr_t)energy, (uintptr_t)out_a, (uintptr_t)out_b,
                             (uintptr_t)iters, (uintptr_t)m->grid_pct, (uintptr_t)m->activation};
  if (m->bench_exec == nullptr || std::memcmp(key, m->bench_key, sizeof key) != 0) {
    // capture `iters` whole passes on a private stream (the caller's may be
    // the null stream, which cannot be captured)
    if (m->bench_exec) {
      drop_hip_error(hipGraphExecDestroy(m->bench_exec));
      m->bench_exec = nullptr;
    }
    hipStream_t cs = nullptr;
    APL_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    int rc = 0;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = -1;
    for (int it = 0; it < iters && rc == 0; ++it) {
      rc = launch_elem(m, material, w, mu, lam, u, p, nullptr, nullptr, out_a, out_b, 0, 0, cs);
      if (rc == 0)
        rc = launch_finish(m, w, out_a, out_b, (w & W_E) ? energy : nullptr, (w & W_HQ) ? energy : nullptr, 0,
                           nullptr, cs);
    }
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    if (rc == 0 && ec == hipSuccess && g != nullptr &&
        hipGraphInstantiate(&m->bench_exec, g, nullptr, nullptr, 0) != hipSuccess)
      m->bench_exec = nullptr;
    if (g) drop_hip_error(hipGraphDestroy(g));
    drop_hip_error(hipStreamDestroy(cs));
    if (rc != 0 || ec != hipSuccess || m->bench_exec == nullptr) {
      (void)hipGetLastError();
      if (rc == 0) set_error("capture of the benchmark graph failed");
      return -1;
    }
    std::memcpy(m->bench_key, key, sizeof key);
  }
  APL_CHECK_HIP(hipEventRecord(ev[2], s));
  APL_CHECK_HIP(hipGraphLaunch(m->bench_exec, s));
  APL_CHECK_HIP(hipEventRecord(ev[3], s));
  if (spin_wait(ev[3])) return -1;
  float b = 0;
  APL_CHECK_HIP(hipEventElapsedTime(&b, ev[2], ev[3]));
  *ms_pass = b / iters;
  return 0;
}

}  // extern "C"
