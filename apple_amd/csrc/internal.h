// Library-internal types shared by the element kernels, the vector kernels
// and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "layout.h"
#include "tetmath.h"

namespace apl {

constexpr int NTH = 256;       // threads per element / vector block
constexpr int MAX_CH = 6;      // scatter channels per incidence (grad + hdiag)
constexpr int N_DOTS = 16;     // per vertex-block partial dot products
constexpr int RED_STRIDE = 2;  // per element block: energy, quadratic form
// material id of a model mesh's launches: every law over its own blocks
constexpr int MAT_MODEL = 3;

enum Dtype : int { DT_F32 = 0, DT_F64 = 1 };

void set_error(const std::string& s);

#define APL_CHECK_HIP(expr)                                                      \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::apl::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
      (void)hipGetLastError(); /* reported through apl_last_error instead */   \
      return -1;                                                               \
    }                                                                          \
  } while (0)

struct DevLayout {
  void* dminv = nullptr;  // 9*T, SoA (k*T + t)
  void* dV = nullptr;     // T
  uint16_t* corner = nullptr;  // 8*T: 4 local slots, then 4 chunk positions per tet
  int32_t* blk_tet_off = nullptr;
  int32_t* blk_loc_off = nullptr;
  int32_t* blk_inst = nullptr;
  int64_t* inst_blk_off = nullptr;
  int32_t* loc_gid = nullptr;
  int32_t* blk_co_off = nullptr;   // nblocks+1 into chunk_off
  uint16_t* chunk_off = nullptr;   // per block and chunk: nloc+1 slot offsets into positions
  int32_t* vtx_slot_off = nullptr;  // V+1: slot range of every vertex
  int32_t* loc_dst = nullptr;       // n_slots: vertex-major partial index of every slot
  int32_t* vblk_off = nullptr;
  int32_t* vblk_inst = nullptr;
  int64_t* inst_vblk_off = nullptr;
  int64_t* inst_vtx_off = nullptr;
  int32_t* law_blk[N_MAT] = {};  // model meshes: the blocks of each material law, ascending
  // workspace
  void* partial = nullptr;   // n_slots x MAX_CH, vertex-major (one partial sum per local slot)
  double* blk_red = nullptr; // RED_STRIDE * nblocks
  double* vblk_dots = nullptr;  // N_DOTS * nvblk
};

}  // namespace apl

struct apl_mesh {
  int dtype = 0;
  int device = 0;
  int n_cu = 256;
  apl::Layout L;
  apl::DevLayout d;
  size_t device_bytes = 0;
  int grid_cache[3][512] = {};  // persistent grid per (material, pass variant, FP64 displacements, filtered)
  int refs = 1;                 // the caller's handle plus one per solver built on the mesh
  int grid_pct = 0;             // persistent element grid, percent of occupancy x CUs (autotuned; 0: per-pass default)
  const void* activation = nullptr;  // caller-owned (T, 6) muscle activation, read at every muscle launch
  // model meshes (apl_model_create): several potentials' tets over one
  // vertex set, in segments of one material law each
  int model = 0;
  int64_t law_nblk[3] = {0, 0, 0};  // blocks per law (MAT_SNH, MAT_ARAP, MAT_MUSCLE)
  // measurement support: the last graph of back-to-back passes that
  // apl_bench_elem_graph captured on this mesh, and the arguments it holds
  hipGraphExec_t bench_exec = nullptr;
  uintptr_t bench_key[12] = {};
};

namespace apl {

// Drop one reference; the last one frees the mesh.  A solver keeps its mesh
// alive, so handles may be destroyed in any order (garbage collectors
// finalise cycles in no particular order).
void mesh_release(apl_mesh* m);

// Clear the calling thread's HIP error state after a call whose failure is
// deliberately ignored, so that the next unrelated caller of
// hipGetLastError() (e.g. PyTorch after its own launches) does not see it.
inline void drop_hip_error(hipError_t) { (void)hipGetLastError(); }

inline size_t real_size(const apl_mesh* m) { return m->dtype == DT_F64 ? 8 : 4; }

// Scratch of the element passes: the vertex-major slot partials and the
// per-block energy / quadratic-form sums.  By default the mesh's own
// workspace (allocated at creation); the allocation-free entry point passes
// a static device array instead.
struct Workspace {
  void* partial = nullptr;
  double* blk_red = nullptr;
};
// bytes of scratch one apl_elem_eval call may use on this mesh (partials of
// up to MAX_CH channels, then the block sums at a 256-byte aligned offset)
size_t eval_scratch_bytes(const apl_mesh* m);
// apl_elem_eval against workspace ws (nullptr: the mesh's)
int elem_eval_impl(apl_mesh* m, int material, const void* mu, const void* lam, const void* u, const void* p,
                   unsigned what, double* energy, void* grad, void* hess_diag, void* hess_prod, double* hess_quad,
                   void* dmu, void* dlam, hipStream_t s, const Workspace* ws);

// element pass: per-slot partial sums into the partial workspace (scatter
// passes), per-block energy/quadratic-form sums, or per-tet outputs (mixed);
// x_f64: x holds FP64 displacements while the mesh is FP32 (solver passes)
int launch_elem(apl_mesh* m, int material, unsigned what, const void* mu, const void* lam,
                const void* x, const void* p, const double* alpha, const int32_t* active,
                void* out_a, void* out_b, int accumulate, int grid_override, hipStream_t s, int x_f64 = 0,
                void* out_c = nullptr, const Workspace* ws = nullptr);
// vertex sums of the slot partials into the (V,3) outputs
// Finish pass of an element pass in one launch: vertex sums of the slot
// partials into out_a / out_b (what: the scatter channels of the pass) and,
// when out_e / out_q are given, per-instance sums of the block energies /
// quadratic forms.
int launch_finish(apl_mesh* m, unsigned what, void* out_a, void* out_b, double* out_e, double* out_q,
                  int accumulate, const int32_t* active, hipStream_t s, const Workspace* ws = nullptr);
// material id valid for this mesh (APL_MODEL exactly for model meshes)
bool material_ok(const apl_mesh* m, int material);
// whether launches of this material read lambda
bool needs_lambda(const apl_mesh* m, int material);
// largest block window (distinct vertices per block) whose LDS every element
// pass variant of a mesh of this dtype fits
int nloc_limit(int dtype);
int elem_grid(apl_mesh* m, int material, unsigned what, int x_f64 = 0, bool filtered = false);
// device setup (setup.hip): the element data (inverse rest shapes, dV *
// fraction) of a created mesh from host cells / points / fraction, and the
// read-back of the element data in the caller's tet order (T x 9, T)
int region_from_points(apl_mesh* m, const int32_t* cells, const double* points, const double* fraction);
// device block partition (partition.hip): the Layout of build_layout from
// host cells, built by kernels; fills the mesh's layout arrays (device) and
// the sizes and block / vertex-block tables (host).  "" or an error.
bool device_layout_supported(int chunks, int nloc_cap);
std::string build_layout_device(apl_mesh* m, const int32_t* cells, int64_t T, int64_t V, int64_t n_inst,
                                const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int chunks, int nloc_cap,
                                int64_t n_seg_total, const int64_t* seg_off);
int region_readback(apl_mesh* m, void* dminv_out, void* dv_out, hipStream_t s);

}  // namespace apl
