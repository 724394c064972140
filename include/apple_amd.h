/* apple_amd -- C ABI of the MI355X (gfx950) FEM elasticity hot path.
 *
 * The drop-in boundary for liblaf/apple's element potentials and static
 * solver.  Every array argument is a caller-owned device pointer (e.g. a
 * PyTorch-ROCm tensor's data_ptr()) unless marked "host".  Real-valued arrays
 * are float32 or float64 according to the dtype the mesh was created with;
 * energies, quadratic forms and solver scalars are always float64.  Vertex
 * vectors are (V,3) row-major, per-tet arrays (T,), both in the caller's
 * order.  All work is enqueued on `stream` (a hipStream_t, NULL = default
 * stream); a mesh or solver handle must be used by one stream at a time.
 * Functions return 0 on success and -1 on error; apl_last_error() then
 * describes the error (thread-local).  No entry point allocates device memory
 * except the *_create functions.
 */
#ifndef APPLE_AMD_H
#define APPLE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct apl_mesh apl_mesh;
typedef struct apl_solver apl_solver;

enum { APL_F32 = 0, APL_F64 = 1 };
/* materials: warp/fem/_stable_neo_hookean.py, warp/fem/_arap.py,
 * warp/fem/_stable_neo_hookean_muscle.py (SNH on F A, A = I + sym(activation);
 * the activation is bound to the mesh by apl_mesh_set_activation) */
enum { APL_SNH = 0, APL_ARAP = 1, APL_SNH_MUSCLE = 2,
       /* a model mesh (apl_model_create): each segment's own law, summed */
       APL_MODEL = 3 };

/* output selection for apl_elem_eval */
enum {
  APL_ENERGY = 1,      /* WarpPotentialFem.fun        (warp/fem/_base.py:151-158, 243-263) */
  APL_GRAD = 2,        /* WarpPotentialFem.grad       (_base.py:160-167, 265-291) */
  APL_HESS_DIAG = 4,   /* WarpPotentialFem.hess_diag  (_base.py:169-176, 293-324) */
  APL_HESS_PROD = 8,   /* WarpPotentialFem.hess_prod  (_base.py:178-185, 326-351) */
  APL_HESS_QUAD = 16,  /* WarpPotentialFem.hess_quad  (_base.py:187-194, 353-383) */
  APL_MIXED = 32,      /* p^T d(grad)/d(mu), p^T d(grad)/d(lambda) per tet: the
                          mixed_derivative_prod of the adjoint
                          (exp/2025/08/27/inverse/src/30-inverse.py:96-97) */
  APL_ACCUMULATE = 1 << 16 /* add into the outputs (the reference's += semantics,
                              warp/model/_model.py:13-36) instead of overwriting */
};

/* solver result codes (liblaf-peach Result names as used by apple:
 * tests/forward/test_static_simulation.py:85) */
enum {
  APL_RUNNING = 0,
  APL_PRIMARY_SUCCESS = 1,
  APL_MAX_STEPS_REACHED = 2,
  APL_NON_FINITE = 3,
  APL_STAGNATION = 4
};

const char* apl_last_error(void);
int apl_version(void);

/* Setup: WarpPotentialFem.from_region (warp/fem/_base.py:94-111) plus the
 * block partition of the deterministic reduction.  cells (host, T x 4),
 * dminv (host, T x 3 x 3: inverse of dX/dr, i.e. rows 1..3 of the region's
 * dhdX, jax/fem/region/_region.py:84-108), dv (host, T: dV including
 * Fraction).  A batch of independent meshes is one mesh whose instances own
 * the tet ranges inst_tet_off[i]..inst_tet_off[i+1] and vertex ranges
 * inst_vtx_off[i]..; pass n_inst = 1 and NULL offsets for a single mesh.
 * chunks (>= 1) is the number of 256-tet chunks one block may span. */
int apl_mesh_create(int dtype, int64_t n_tets, int64_t n_verts, const int32_t* cells, const double* dminv,
                    const double* dv, int64_t n_inst, const int64_t* inst_tet_off, const int64_t* inst_vtx_off,
                    int chunks, int device, apl_mesh** out);
/* Setup on the device: as apl_mesh_create, but the element data are computed
 * by a kernel from the rest positions, replacing Region's shape gradients
 * (jax/fem/region/_region.py:84-108: dXdr = X_cell^T dhdr, its inverse,
 * dV = det(dXdr) / 6) and the Fraction factor (warp/fem/_base.py:104).
 * points (host, V x 3), fraction (host, T, or NULL for 1).  Tets are used as
 * given (an inverted tet gets dV < 0, as in the reference). */
int apl_mesh_create_from_points(int dtype, int64_t n_tets, int64_t n_verts, const int32_t* cells,
                                const double* points, const double* fraction, int64_t n_inst,
                                const int64_t* inst_tet_off, const int64_t* inst_vtx_off, int chunks, int device,
                                apl_mesh** out);
/* A sum of element potentials over one vertex set: WarpModel's sum of its
 * potentials (warp/model/_model.py:13-36), e.g. the three FEM energies of one
 * mesh with their own law, Fraction and materials in
 * exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:256-279.
 * As apl_mesh_create, plus n_seg segments per instance: segment s of
 * instance i owns the tets seg_tet_off[i * n_seg + s] ..
 * seg_tet_off[i * n_seg + s + 1] (host, n_inst * n_seg + 1 offsets) and
 * evaluates the law seg_law[s] (APL_SNH, APL_ARAP, APL_SNH_MUSCLE).  Tets of
 * different segments may repeat the same cell (each carries its own dV, e.g.
 * its potential's Fraction).  Element passes and solvers on a model mesh take
 * material APL_MODEL with per-tet material arrays over all its tets (lambda
 * is read for SNH and muscle tets; the activation, bound by
 * apl_mesh_set_activation, for muscle tets); each law runs its own element
 * launch over its own blocks and all of them add into one vertex reduction. */
int apl_model_create(int dtype, int64_t n_tets, int64_t n_verts, const int32_t* cells, const double* dminv,
                     const double* dv, int64_t n_inst, const int64_t* inst_tet_off, const int64_t* inst_vtx_off,
                     int n_seg, const int32_t* seg_law, const int64_t* seg_tet_off, int chunks, int device,
                     apl_mesh** out);

/* Read back the element data in the caller's tet order: dminv (T x 3 x 3,
 * the inverse of dX/dr = rows 1..3 of the region's dhdX) and dv (T, dV
 * including Fraction), device arrays in the mesh dtype; either may be NULL. */
int apl_mesh_region(apl_mesh* mesh, void* dminv, void* dv, void* stream);

/* The rest of the setup path on the device (mesh setup done by kernels:
 * CSR incidence and reordering, lumped-mass loads).
 *
 * The block partition (the vertex-major reduction layout of every element
 * pass) is built by kernels from the cells when a mesh is created with at
 * most 4 chunks (the default), else by the host pass; both give the same
 * arrays.  apl_layout_mode selects it for later creations: 0 the device
 * (default), 1 the host; returns the previous mode (no reference
 * counterpart: a test and fallback knob, process-wide). */
int apl_layout_mode(int mode);

/* Read back one layout array of a mesh (host out, n entries, which =
 * APL_LAYOUT_*): u16 corner records (8 T), i32 block tet / slot / chunk
 * offsets (n_blocks + 1), i32 block instances (n_blocks), i64 instance block
 * offsets (n_inst + 1), i32 slot vertices and vertex-major places
 * (n_partials), u16 chunk slot offsets, i32 vertex slot offsets (V + 1),
 * i32 vertex blocks and their instances, i64 instance vertex-block
 * offsets.  No reference counterpart (inspection and tests). */
enum {
  APL_LAYOUT_CORNER = 0,
  APL_LAYOUT_BLK_TET_OFF = 1,
  APL_LAYOUT_BLK_LOC_OFF = 2,
  APL_LAYOUT_BLK_INST = 3,
  APL_LAYOUT_INST_BLK_OFF = 4,
  APL_LAYOUT_LOC_GID = 5,
  APL_LAYOUT_BLK_CO_OFF = 6,
  APL_LAYOUT_CHUNK_OFF = 7,
  APL_LAYOUT_VTX_SLOT_OFF = 8,
  APL_LAYOUT_LOC_DST = 9,
  APL_LAYOUT_VBLK_OFF = 10,
  APL_LAYOUT_VBLK_INST = 11,
  APL_LAYOUT_INST_VBLK_OFF = 12
};
int apl_mesh_layout(const apl_mesh* mesh, int which, void* out /* host */, int64_t n);

/* Morton (Z-curve) order of a mesh on the device, the locality the block
 * partition relies on for arbitrary meshes: points (device, V x 3 FP64),
 * cells (device, T x 4).  Each point is quantised to 11 bits per axis over
 * the bounding box, each tet by its centroid; vertex_perm (device, V) and
 * cell_perm (device, T) are the stable sorts by code (new = old[perm]); a
 * cell index outside [0, n_verts) is an error (nothing out of range is read).
 * The reference takes meshes in the order they are given; this replaces the
 * host reorder of apple_amd.mesh. */
int apl_morton_order(const double* points, int64_t n_verts, const int32_t* cells, int64_t n_tets,
                     int32_t* vertex_perm, int32_t* cell_perm, int device, void* stream);

/* Lumped vertex volumes: out[v] (device, V, FP64) = sum over the tets t of v
 * of density[t] dV_t / 4 (density: device, T in the mesh dtype, or NULL for
 * 1; dV includes Fraction), summed in a fixed order through the mesh's
 * block partition.  times g it is the lumped-mass gravity load that the
 * experiments hand to ExternalForce (warp/potential/_ext_force.py:42-61).
 * Uses the mesh's pass scratch: stream-ordered with its other passes. */
int apl_mesh_lumped_volume(apl_mesh* mesh, const void* density, double* out, void* stream);

/* Releases the caller's handle.  Solvers built on the mesh hold their own
 * reference: the device memory is freed when the last of them is destroyed,
 * so the handles may be destroyed in any order. */
int apl_mesh_destroy(apl_mesh* mesh);

typedef struct {
  int64_t n_tets, n_verts, n_inst, n_blocks, n_partials, n_shared_verts, n_vertex_blocks;
  int32_t nloc_max, chunks, dtype, elem_grid_snh, elem_grid_arap;
  int64_t device_bytes;
} apl_mesh_info_t;
int apl_mesh_info(const apl_mesh* mesh, apl_mesh_info_t* info /* host */);

/* Schedule knob (no reference counterpart; set by the autotuner,
 * apple_amd/autotune.py): the persistent element-kernel grid as a percentage
 * (25..800) of occupancy x compute units, or 0 for a per-pass default (the
 * initial setting: 200 for the ARAP gradient passes and the Hessian-vector
 * pass, 100 otherwise).
 * Together with the block size
 * chosen at apl_mesh_create (chunks) it is the schedule the autotuner picks.
 * Takes effect for later launches; a solver's captured graph keeps its grid
 * until the solver's parameters change. */
int apl_mesh_set_schedule(apl_mesh* mesh, int grid_pct);

/* Activation of the muscle law (StableNeoHookeanMuscle.Materials.activation,
 * warp/fem/_stable_neo_hookean_muscle.py:120-133, warp/fem/utils/_material.py:15-16):
 * a device array (T, 6) in the mesh dtype, row t = (a0..a5) with
 * A_t = I + [[a0, a3, a4], [a3, a1, a5], [a4, a5, a2]] (warp/fem/func/_misc.py:48-53).
 * The mesh keeps the pointer, not a copy: every later APL_SNH_MUSCLE launch on
 * the mesh (element passes, solves) reads the array's current contents, so it
 * must stay allocated while such work is queued.  NULL unbinds it. */
int apl_mesh_set_activation(apl_mesh* mesh, const void* activation);

/* One fused element pass (replaces the five launches of warp/fem/_base.py:151-194).
 * u, p: (V,3); mu, lam: (T,) (lam ignored for ARAP); energy, hess_quad:
 * (n_inst,) float64; grad, hess_diag, hess_prod: (V,3); dmu, dlam: (T,).
 * Unused outputs may be NULL. */
int apl_elem_eval(apl_mesh* mesh, int material, const void* mu, const void* lam, const void* u, const void* p,
                  unsigned what, double* energy, void* grad, void* hess_diag, void* hess_prod, double* hess_quad,
                  void* dmu, void* dlam, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  As apl_elem_eval, but the element passes' scratch
 * (slot partials and block sums) lives in a static per-GPU device array of
 * apl_static_scratch_bytes() (32 MiB) instead of the mesh's workspace; the
 * call allocates nothing.  Meshes whose scratch exceeds the array are
 * rejected with an error (64^3 x 5 tets fits in FP32 and FP64).  Every call
 * on one GPU shares the array: calls that overlap on one GPU race. */
int apl_elem_eval_static(apl_mesh* mesh, int material, const void* mu, const void* lam, const void* u,
                         const void* p, unsigned what, double* energy, void* grad, void* hess_diag, void* hess_prod,
                         double* hess_quad, void* dmu, void* dlam, void* stream);
size_t apl_static_scratch_bytes(void);

/* Mixed derivatives of the muscle law for the activation adjoint (on a model
 * mesh: dact of its muscle-law tets, other rows untouched, and dmu / dlam of
 * every tet)
 * (mixed_derivative_prod of exp/2026/01/28/smas/src/
 * 31-inverse-activation-stable-neo-hookean.py:472-486): per tet,
 * dact[t, k] = p^T d(grad)/d(a_k) (T, 6), and optionally dmu, dlam (T,) as
 * APL_MIXED computes them.  accumulate: add into the outputs. */
int apl_elem_mixed_activation(apl_mesh* mesh, const void* mu, const void* lam, const void* u, const void* p,
                              void* dact, void* dmu, void* dlam, int accumulate, void* stream);

/* ExternalForce (warp/potential/_ext_force.py:17-39): E -= f . u[idx],
 * grad[idx] -= f for n entries with unique indices; an index outside
 * [0, n_verts) contributes nothing (no memory is touched for it).
 * energy: (1,) float64. */
int apl_ext_force_eval(apl_mesh* mesh, int64_t n, const int32_t* indices, const void* force, const void* u,
                       double* energy, void* grad, int accumulate, void* stream);

/* Measurement support (no reference counterpart).  ms_kernel (host, may be
 * NULL): mean time of `iters` element-kernel launches run back to back
 * between two HIP events on `stream`; ms_pass (host, may be NULL): mean
 * time of `iters` whole passes (element kernel + vertex and energy sums)
 * measured the same way.  what: ENERGY|GRAD|HESS_DIAG (pass A), HESS_QUAD or
 * HESS_PROD.  The timing events are created on a device's first call and kept
 * for the process: concurrent calls on one device race. */
int apl_bench_elem_pass(apl_mesh* mesh, int material, unsigned what, const void* mu, const void* lam, const void* u,
                        const void* p, double* energy, void* out_a, void* out_b, int iters, double* ms_kernel,
                        double* ms_pass, void* stream);

/* Measurement support (no reference counterpart): the same `iters` whole
 * passes as apl_bench_elem_pass, captured once into a HIP graph and replayed
 * as one launch between two events on `stream` (ms_pass: the mean pass
 * time).  The first call for a set of arguments captures (and runs) the
 * graph; later calls with the same arguments only replay it.  The mesh keeps
 * the last graph until it is destroyed. */
int apl_bench_elem_graph(apl_mesh* mesh, int material, unsigned what, const void* mu, const void* lam, const void* u,
                         const void* p, double* energy, void* out_a, void* out_b, int iters, double* ms_pass,
                         void* stream);

/* ---------------------------------------------------------------- solvers */

/* PNCG parameters.  Defaults used by the Python layer: rtol 1e-6, atol 0,
 * max_steps 1500 (forward/_forward.py:26-27), max_backtracking 8, overstep 1,
 * c1 1e-4, wolfe_delta 0.1, wolfe_eps 1e-6 (FP32) / 1e-12 (FP64), beta_max 10
 * (benches/bench_pncg_branching_backends.py:288-289), iters_per_launch 16,
 * use_graph 1, grid_override 0 (occupancy-derived persistent grid),
 * stall_steps 2000: an instance whose best relative gradient norm is within
 * three decades of rtol and that has lowered neither that norm by 10% nor
 * its energy by 10 wolfe_eps relative within that many accepted steps ends with
 * APL_STAGNATION (0 = never); this is where an FP32 solve stops when its
 * gradient reaches the rounding floor of the element forces. */
typedef struct {
  double rtol, atol;
  int32_t max_steps, max_backtracking;
  double overstep, c1, wolfe_delta, wolfe_eps, beta_max;
  int32_t iters_per_launch, use_graph, grid_override, stall_steps;
} apl_pncg_params;

/* Per-instance result: status (APL_* code), accepted steps, energy
 * evaluations, backtracks, final energy, gradient norms, last step, last beta. */
typedef struct {
  int32_t status, n_steps, n_evals, n_backtracks;
  double f, grad_norm, grad_norm_first, alpha, beta, f_first;
} apl_pncg_stats;

/* A solver owns the work vectors and per-instance state for one mesh and
 * material (the Forward/PNCG pair of forward/_forward.py:18-56). */
int apl_solver_create(apl_mesh* mesh, int material, apl_solver** out);
int apl_solver_destroy(apl_solver* solver);

/* Forward.step(): minimise E(x) = sum_t Psi_t dV_t - f_ext . x over the free
 * DOFs, independently per instance.  x (V,3) holds the initial guess with
 * the fixed DOFs at their prescribed values (DofMap.to_full,
 * forward/dof_map/_dof_map.py:39-43) and is overwritten by the solution.
 * f_ext (V,3) or NULL; fixed_mask (V,) uint8, bit c set = component c fixed
 * (DofMapBuilder, forward/dof_map/_builder.py:34-64), or NULL.  stats: host
 * array of n_inst entries or NULL.  Synchronises `stream` once per
 * iters_per_launch iterations to test for termination. */
int apl_pncg_solve(apl_solver* solver, const void* mu, const void* lam, void* x, const void* f_ext,
                   const uint8_t* fixed_mask, const apl_pncg_params* params, apl_pncg_stats* stats, void* stream);

/* Step-wise PNCG: liblaf-peach's Optimizer.init / step / terminate /
 * postprocess as apple's experiments drive them
 * (exp/2026/05/06/toy/src/21-smas-prestrain-stable-neo-hookean-muscle.py:497-564).
 * apl_pncg_begin binds the arguments of apl_pncg_solve and resets the state;
 * apl_pncg_advance runs until every running instance has taken `steps` more
 * accepted steps (steps > 0; one trial evaluation per launch, the status read
 * after each), or else until none runs or max_iters more trial evaluations
 * (max_iters > 0; graph launches of iters_per_launch) have run; x_out (V,3,
 * nullable) receives the current iterate and state (host, n_inst, nullable)
 * the per-instance optimizer state (the reference's opt_state fields);
 * apl_pncg_end writes the iterate into the x given to apl_pncg_begin and
 * fills stats.  apl_pncg_solve is begin + advance to the end + end. */
typedef struct {
  int32_t status, n_steps, n_evals;
  int32_t ls_steps; /* line_search_state.step: backtracks of the last accepted step */
  int32_t ok;       /* line_search_state.ok: 0 while the last line search failed (restart) */
  double alpha;     /* line_search_state.alpha: the last accepted step */
  double f0;        /* line_search_state.f0: energy before it */
  double f_alpha;   /* line_search_state.f_alpha: energy after it */
  double grad_norm, grad_norm_first; /* convergence_state */
  double hess_diag_mean;             /* hess_damping_state.hess_diag_mean: mean positive diag H */
  double beta;
} apl_pncg_state;
int apl_pncg_begin(apl_solver* solver, const void* mu, const void* lam, void* x, const void* f_ext,
                   const uint8_t* fixed_mask, const apl_pncg_params* params, void* stream);
int apl_pncg_advance(apl_solver* solver, int32_t steps, int64_t max_iters, void* x_out, apl_pncg_state* state,
                     void* stream);
int apl_pncg_end(apl_solver* solver, apl_pncg_stats* stats, void* stream);

/* Adjoint linear solve (exp/2025/08/27/inverse/src/30-inverse.py:85-93):
 * Jacobi-preconditioned CG on H(x) y = rhs over the free DOFs, H applied by
 * the unclamped Hessian-vector pass; y (V,3) is overwritten (zero at fixed
 * DOFs).  Stops at ||r|| <= rtol ||rhs|| per instance; stats as above with
 * f = final relative residual.  Reports APL_NON_FINITE if d^T H d <= 0. */
int apl_cg_solve(apl_solver* solver, const void* mu, const void* lam, const void* x, const void* rhs, void* y,
                 const uint8_t* fixed_mask, double rtol, int max_iters, apl_pncg_stats* stats, void* stream);

/* The adjoint solve for an indefinite Hessian: CG on the normal equations
 * H^2 y = H rhs over the free DOFs (lineax.NormalCG, the adjoint solver of
 * exp/2025/08/27/inverse/src/30-inverse.py:85-93), two Hessian-vector passes
 * per iteration, Jacobi on diag(H)^2.  Arguments as apl_cg_solve; stops at
 * ||H (rhs - H y)|| <= rtol ||H rhs|| (stats.f is that ratio). */
int apl_cg_solve_normal(apl_solver* solver, const void* mu, const void* lam, const void* x, const void* rhs, void* y,
                        const uint8_t* fixed_mask, double rtol, int max_iters, apl_pncg_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* APPLE_AMD_H */
