"""PNCG / PCG on the GPU vs the oracle (end states), through the C ABI.

* The reference's patch test (tests/forward/test_static_simulation.py:58-93)
  through ModelBuilder / Forward.step(): interior vertex at
  (0.05, -0.025, 0.0375) within 1e-8 (FP64) and PRIMARY_SUCCESS.
* Gravity-loaded cubes (SNH and ARAP, z = 0 pinned, seeded noisy start):
  equilibrium positions within 1e-5 of the bounding-box diagonal of the
  Newton-polished FP64 oracle equilibrium; the oracle gradient there is ~0.
* Graph replay == eager launches, bitwise; batched == one-by-one, bitwise.
* PCG: H y = b vs a sparse direct solve of the oracle's exact Hessian.
* Failure paths: max_steps exhaustion, non-finite materials.
"""

import numpy as np
import pytest
import torch

import oracle.fem as of
import oracle.pncg as op
from apple_amd import batch, fem, forward, potential
from apple_amd.mesh import Mesh, Region, cube
from apple_amd.solver import DeviceSolver, PncgConfig, Result, mask_bits
from helpers import patch_test

pytestmark = pytest.mark.gpu

PATCH_POINTS, PATCH_CELLS, PATCH_MASK, PATCH_VALUES, PATCH = patch_test()


@pytest.mark.parametrize("dtype,atol", [(torch.float64, 1e-8), (torch.float32, 2e-6)])
@pytest.mark.parametrize("cls", [fem.Arap, fem.StableNeoHookean])
def test_patch_test_known_answer(dtype, atol, cls):
    print(f"[solver] patch {cls.__name__} {dtype}", flush=True)
    mesh = Mesh(PATCH_POINTS.copy(), PATCH_CELLS.copy())
    n = PATCH_CELLS.shape[0]
    mesh.cell_data["mu"] = np.full(n, PATCH["mu"])
    mesh.cell_data["lambda"] = np.ones(n)
    builder = forward.ModelBuilder()
    builder.add_vertices(mesh)
    fixed_mask = PATCH_MASK.copy()
    fixed_value = PATCH_VALUES.copy()
    mesh.point_data["FixedMask"] = fixed_mask
    mesh.point_data["FixedValue"] = fixed_value
    builder.add_fixed(mesh)
    builder.add_potential(cls.from_pyvista(mesh, dtype=dtype, device="cuda"))
    model = builder.finalize()
    fwd = forward.Forward(model, forward.PNCG(rtol=1e-9 if dtype == torch.float64 else 1e-5))
    e0 = float(fwd.problem.fun(fwd.state))
    sol = fwd.step()
    e1 = float(fwd.problem.fun(fwd.state))
    exp = PATCH["expected"]
    assert sol.result.name == exp["result"]
    assert model.n_free == exp["n_free"]
    assert np.isfinite(e0) and np.isfinite(e1) and e1 < e0
    u = fwd.state.u.double().cpu().numpy()
    np.testing.assert_allclose(u[fixed_mask[:, 0]], fixed_value[fixed_mask[:, 0]], atol=1e-7)
    np.testing.assert_allclose(u[exp["interior_vertex"]], exp["interior_displacement"], atol=max(atol, exp["atol"]))


def _gravity_case(n, material, seed=0, noise_scale=0.1):
    m = cube(n)
    rng = np.random.default_rng(seed)
    mu = 1e4 * np.exp(0.5 * rng.standard_normal(m.n_cells))
    lam = 10.0 * mu
    region = Region.from_mesh(m)
    f = np.zeros((m.n_points, 3))
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8 * 1e3 * region.dV / 4.0, 4))
    fm = np.zeros((m.n_points, 3), bool)
    fm[m.points[:, 2] == 0.0] = True
    h = 1.0 / n
    x0 = noise_scale * h * np.random.default_rng(seed + 100).standard_normal((m.n_points, 3))
    x0[fm] = 0.0
    return m, region, mu, lam, f, fm, x0


def _oracle_equilibrium(region, material, mu, lam, f, fm, x_start):
    mat = of.SNH if material == "snh" else of.ARAP
    pot = of.FemPotential(mat, region.cells, region.dhdX, region.dV, mu, lam if mat == of.SNH else None)
    ext = of.ExternalForce(np.arange(f.shape[0]), f)
    dof = of.DofMap(fm, np.zeros_like(f))
    prob = op.Problem(of.Model([pot, ext]), dof)
    x = op.newton(prob, dof.to_free(x_start), tol=1e-10, max_steps=30)
    gn = np.linalg.norm(prob.grad(x))
    return dof.to_full(x), gn


@pytest.mark.parametrize("material", ["snh", "arap"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gravity_equilibrium_matches_oracle(material, dtype):
    print(f"[solver] gravity {material} {dtype}", flush=True)
    m, region, mu, lam, f, fm, x0 = _gravity_case(6, material)
    cls = fem.StableNeoHookean if material == "snh" else fem.Arap
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = lam
    pot = cls.from_region(region, dtype=dtype, device="cuda")
    solver = DeviceSolver(pot)
    x = torch.as_tensor(x0, dtype=dtype, device="cuda").contiguous()
    ft = torch.as_tensor(f, dtype=dtype, device="cuda").contiguous()
    mask = mask_bits(fm, x.device)
    st = solver.solve(x, ft, mask, PncgConfig(rtol=1e-6, max_steps=5000))[0]
    assert st.result == Result.PRIMARY_SUCCESS, st
    assert st.relative_grad_norm <= 1e-6
    xg = x.double().cpu().numpy()
    # SNH: the oracle's Newton from the same start, so a solve that lands on
    # another equilibrium would fail; ARAP's exact Hessian is indefinite far
    # from equilibrium and Newton from the noisy start does not converge, so
    # it polishes the device's answer (the batched test checks ARAP answers
    # by the oracle's gradient instead)
    x_ref, gn = _oracle_equilibrium(region, material, mu, lam, f, fm, x0 if material == "snh" else xg)
    assert gn <= 1e-6 * max(st.grad_norm_first, 1.0)
    diag = np.sqrt(3.0)
    err = np.abs(xg - x_ref).max() / diag
    print(f"  steps {st.n_steps} evals {st.n_evals} rel|g| {st.relative_grad_norm:.2e} pos err/diag {err:.2e}")
    assert err <= 1e-5


def test_graph_replay_equals_eager_launches():
    m, region, mu, lam, f, fm, x0 = _gravity_case(5, "snh", seed=2)
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = lam
    pot = fem.StableNeoHookean.from_region(region, dtype=torch.float32, device="cuda")
    solver = DeviceSolver(pot)
    ft = torch.as_tensor(f, dtype=torch.float32, device="cuda")
    mask = mask_bits(fm, "cuda")
    xs, sts = [], []
    for graph in (True, False, True):
        x = torch.as_tensor(x0, dtype=torch.float32, device="cuda").contiguous()
        sts.append(solver.solve(x, ft, mask, PncgConfig(use_graph=graph, iters_per_launch=8))[0])
        xs.append(x.cpu())
    assert torch.equal(xs[0], xs[1]) and torch.equal(xs[0], xs[2])
    assert sts[0].n_steps == sts[1].n_steps == sts[2].n_steps


def test_batched_equals_one_by_one():
    insts = []
    for k, n in enumerate((3, 5, 4)):
        m, region, mu, lam, f, fm, x0 = _gravity_case(n, "snh", seed=10 + k)
        insts.append(batch.Instance(region, mu, lam, f, fm, x0))
    b = batch.Batch(insts, 0, dtype=torch.float32, device="cuda")
    xb, stb = b.solve(config=PncgConfig(rtol=1e-6))
    parts = b.split(xb)
    for i, inst in enumerate(insts):
        single = batch.Batch([inst], 0, dtype=torch.float32, device="cuda")
        xs, sts = single.solve(config=PncgConfig(rtol=1e-6))
        assert stb[i].result == Result.PRIMARY_SUCCESS
        assert stb[i].n_steps == sts[0].n_steps
        assert torch.equal(parts[i].cpu(), xs.cpu())


@pytest.mark.parametrize("material", ["snh", "arap"])
def test_batched_instances_match_the_oracle(material):
    """Three cubes of different sizes and materials in one FP64 batch, each
    checked by the oracle on its own: SNH against the oracle's Newton
    equilibrium from the same start (not from the device's answer),
    positions within 1e-9 of the cube's diagonal; ARAP, whose exact Hessian
    is indefinite far from equilibrium (Newton from the noisy start wanders
    off), by the oracle's gradient at the device's answer, below 1e-9 of the
    oracle's gradient at the start."""
    from apple_amd import _lib

    print(f"[solver] batch vs oracle {material}", flush=True)
    insts, cases = [], []
    for k, n in enumerate((3, 5, 4)):
        m, region, mu, lam, f, fm, x0 = _gravity_case(n, material, seed=30 + k)
        insts.append(batch.Instance(region, mu, lam if material == "snh" else None, f, fm, x0))
        cases.append((region, mu, lam, f, fm, x0))
    law = _lib.APL_SNH if material == "snh" else _lib.APL_ARAP
    b = batch.Batch(insts, law, dtype=torch.float64, device="cuda")
    xb, stb = b.solve(config=PncgConfig(rtol=1e-11, max_steps=50000))
    parts = b.split(xb)
    for i, (region, mu, lam, f, fm, x0) in enumerate(cases):
        assert stb[i].result == Result.PRIMARY_SUCCESS, (i, stb[i])
        x = parts[i].cpu().numpy()
        if material == "snh":
            xo, gn = _oracle_equilibrium(region, material, mu, lam, f, fm, x0)
            err = np.abs(x - xo).max() / np.sqrt(3.0)
            print(f"  instance {i}: |x - x_newton| / diag {err:.2e} (oracle |g| {gn:.1e})")
            assert err <= 1e-9, (i, err)
        else:
            pot = of.FemPotential(of.ARAP, region.cells, region.dhdX, region.dV, mu, None)
            dof = of.DofMap(fm, np.zeros_like(f))
            prob = op.Problem(of.Model([pot, of.ExternalForce(np.arange(f.shape[0]), f)]), dof)
            g0 = np.linalg.norm(prob.grad(dof.to_free(x0)))
            g = np.linalg.norm(prob.grad(dof.to_free(x)))
            print(f"  instance {i}: oracle |g| at the device's answer {g:.2e}, at the start {g0:.2e}")
            assert g <= 1e-9 * g0, (i, g, g0)


def test_fp32_refined_batch_matches_oracle_newton():
    """FP32 element kernels with FP64-residual refinement (refine_fp64 =
    "newton") on a batch of soft and stiff SNH cubes (median mu 4e3 to 1e5):
    every instance's refined positions against the oracle's FP64 Newton
    equilibrium from the same start, to 1e-8 of the cube's diagonal; the
    plain FP32 answers, for comparison, stop at the FP32 floor."""
    print("[solver] FP32 + refinement vs oracle Newton", flush=True)
    insts, cases = [], []
    for k, (n, scale) in enumerate(((3, 0.4), (4, 1.0), (5, 10.0))):
        m, region, mu, lam, f, fm, x0 = _gravity_case(n, "snh", seed=40 + k)
        mu, lam = mu * scale, lam * scale  # median mu 4e3 (30 % strain under gravity), 1e4, 1e5
        insts.append(batch.Instance(region, mu, lam, f, fm, x0))
        cases.append((region, mu, lam, f, fm, x0))
    b = batch.Batch(insts, _lib_snh(), dtype=torch.float32, device="cuda")
    cfg = PncgConfig(rtol=1e-6, max_steps=50000)
    x32, _ = b.solve(config=cfg)
    xr, st = b.solve(config=cfg, refine_fp64="newton", refine_rtol=1e-10)
    plain, refined = b.split(x32), b.split(b.x64)
    for i, (region, mu, lam, f, fm, x0) in enumerate(cases):
        xo, gn = _oracle_equilibrium(region, "snh", mu, lam, f, fm, x0)
        e32 = np.abs(plain[i].double().cpu().numpy() - xo).max() / np.sqrt(3.0)
        er = np.abs(refined[i].cpu().numpy() - xo).max() / np.sqrt(3.0)
        print(f"  instance {i} (mu median {np.median(mu):.0e}): FP32 {e32:.2e}, refined {er:.2e} of the diagonal "
              f"(oracle |g| {gn:.1e}, {st[i].result.name})")
        assert st[i].result == Result.PRIMARY_SUCCESS, (i, st[i])
        assert er <= 1e-8, (i, er)


@pytest.mark.parametrize("material", ["snh", "arap"])
def test_cg_matches_sparse_direct_solve(material):
    m, region, mu, lam, f, fm, x0 = _gravity_case(4, material, seed=5)
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = lam
    cls = fem.StableNeoHookean if material == "snh" else fem.Arap
    pot = cls.from_region(region, dtype=torch.float64, device="cuda")
    solver = DeviceSolver(pot)
    x = torch.as_tensor(x0, dtype=torch.float64, device="cuda").contiguous()
    ft = torch.as_tensor(f, dtype=torch.float64, device="cuda")
    mask = mask_bits(fm, "cuda")
    solver.solve(x, ft, mask, PncgConfig(rtol=1e-8))
    rng = np.random.default_rng(1)
    b = rng.standard_normal((m.n_points, 3))
    bt = torch.as_tensor(b, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(bt)
    st = solver.cg(x, bt, y, mask, rtol=1e-10, max_iters=5000)[0]
    assert st.result == Result.PRIMARY_SUCCESS, st
    mat = of.SNH if material == "snh" else of.ARAP
    opot = of.FemPotential(mat, region.cells, region.dhdX, region.dV, mu, lam if mat == of.SNH else None)
    dof = of.DofMap(fm, np.zeros_like(f))
    prob = op.Problem(of.Model([opot]), dof)
    H = prob.hessian(dof.to_free(x.cpu().numpy()))
    import scipy.sparse.linalg as spla

    y_ref = dof.to_full_grad(spla.spsolve(H.tocsc(), dof.to_free(b)))
    np.testing.assert_allclose(y.cpu().numpy(), y_ref, rtol=1e-7, atol=1e-9 * np.abs(y_ref).max())


def test_failure_codes():
    m, region, mu, lam, f, fm, x0 = _gravity_case(3, "snh", seed=7)
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = lam
    pot = fem.StableNeoHookean.from_region(region, dtype=torch.float32, device="cuda")
    solver = DeviceSolver(pot)
    ft = torch.as_tensor(f, dtype=torch.float32, device="cuda")
    mask = mask_bits(fm, "cuda")
    x = torch.as_tensor(x0, dtype=torch.float32, device="cuda").contiguous()
    st = solver.solve(x, ft, mask, PncgConfig(max_steps=3, rtol=1e-12))[0]
    assert st.result == Result.MAX_STEPS_REACHED and st.n_steps == 3
    # FP32 cannot reach 1e-9: the gradient stops improving at its rounding
    # floor and the stall test ends the solve well before max_steps
    x = torch.as_tensor(x0, dtype=torch.float32, device="cuda").contiguous()
    st = solver.solve(x, ft, mask, PncgConfig(max_steps=100000, rtol=1e-9, stall_steps=200))[0]
    assert st.result == Result.STAGNATION and st.n_steps < 100000
    assert st.relative_grad_norm < 1e-4
    bad = pot.materials["mu"].clone()
    bad[5] = float("nan")
    pot.set_materials(mu=bad)
    x = torch.as_tensor(x0, dtype=torch.float32, device="cuda").contiguous()
    st = solver.solve(x, ft, mask, PncgConfig())[0]
    assert st.result == Result.NON_FINITE


class _LoadPotential:
    """A potential the fused path does not know (the reference protocol over
    an ExternalForce): forces the host-driven PNCG."""

    def __init__(self, ext):
        self.ext, self.name, self.dtype, self.device = ext, "load", ext.dtype, ext.device
        self.n_points = ext.n_points

    def fun(self, u, output):
        self.ext.fun(u, output)

    def grad(self, u, output):
        self.ext.grad(u, output)

    def hess_diag(self, u, output):
        pass

    def hess_prod(self, u, p, output):
        pass

    def hess_quad(self, u, p, output):
        pass


@pytest.mark.parametrize("fused", [True, False])
def test_two_fem_potentials_fused_and_host_driven(fused):
    """Two FEM potentials on one mesh: the fused device model (DeviceModel)
    and, with a potential the fused path does not know, the host-driven PNCG
    reach the oracle's equilibrium."""
    m, region, mu, lam, f, fm, x0 = _gravity_case(4, "snh", seed=3)
    m.cell_data["mu"] = 0.5 * mu
    m.cell_data["lambda"] = 0.5 * lam
    builder = forward.ModelBuilder()
    builder.add_vertices(m)
    m.point_data["FixedMask"] = fm
    m.point_data["FixedValue"] = np.zeros_like(f)
    builder.add_fixed(m)
    a = fem.StableNeoHookean.from_region(region, dtype=torch.float64, device="cuda")
    b2 = fem.StableNeoHookean.from_region(region, dtype=torch.float64, device="cuda")
    builder.add_potential(a)
    builder.add_potential(b2)
    ext = potential.ExternalForce(np.arange(m.n_points), f, m.n_points, dtype=torch.float64, device="cuda")
    builder.add_potential(ext if fused else _LoadPotential(ext))
    model = builder.finalize()
    assert (model.fused_parts() is None) == (not fused)
    fwd = forward.Forward(model, forward.PNCG(rtol=1e-8, max_steps=5000))
    sol = fwd.step()
    assert sol.success
    x_ref, gn = _oracle_equilibrium(region, "snh", mu, lam, f, fm, fwd.state.u.cpu().numpy())
    assert np.abs(fwd.state.u.cpu().numpy() - x_ref).max() <= 1e-7


def test_handles_destroyed_in_any_order_leave_no_hip_error():
    """A solver keeps its mesh alive; destroying the mesh handle first (as a
    garbage collector may, finalising a reference cycle) must neither free
    memory the solver still uses nor leave a HIP error for the next caller."""
    from apple_amd import _lib
    from apple_amd.solver import DeviceSolver

    m = cube(3)
    m.cell_data["mu"] = np.full(m.n_cells, 1e4)
    m.cell_data["lambda"] = np.full(m.n_cells, 1e5)
    pot = fem.StableNeoHookean.from_region(Region.from_mesh(m), dtype=torch.float64, device="cuda")
    solver = DeviceSolver(pot)
    pot.mesh.close()  # the solver still holds the mesh
    x = torch.zeros(m.n_points, 3, dtype=torch.float64, device="cuda")
    st = solver.solve(x, None, None, PncgConfig(rtol=1e-6, max_steps=5))
    assert st[0].result in (Result.PRIMARY_SUCCESS, Result.MAX_STEPS_REACHED)
    solver.close()
    # torch checks hipGetLastError after its own launches: a stale error would raise here
    y = torch.arange(10, device="cuda", dtype=torch.float64) * 2.0
    torch.cuda.synchronize()
    assert float(y.sum()) == 90.0
    assert _lib.lib() is not None


def test_batch_refine_fp64_replaces_stagnated_instances():
    """FP32 cannot reach rtol 1e-9: with a short stall window the instances end
    in STAGNATION, are solved again in FP64 from there, and then match an
    all-FP64 solve."""
    print("[solver] batch refine_fp64", flush=True)
    insts = []
    for k, n in enumerate((3, 4)):
        m, region, mu, lam, f, fm, x0 = _gravity_case(n, "snh", seed=30 + k)
        insts.append(batch.Instance(region, mu, lam, f, fm, x0))
    cfg = PncgConfig(rtol=1e-9, max_steps=20000, stall_steps=200)
    b32 = batch.Batch(insts, _lib_snh(), dtype=torch.float32, device="cuda")
    x32, st32 = b32.solve(config=cfg, refine_fp64=True)
    assert b32.refined, [s.result for s in st32]
    assert all(s.success for s in st32), st32
    b64 = batch.Batch(insts, _lib_snh(), dtype=torch.float64, device="cuda")
    x64, st64 = b64.solve(config=PncgConfig(rtol=1e-9, max_steps=20000))
    assert all(s.success for s in st64)
    err = (x32.double() - x64).abs().max().item() / np.sqrt(3.0)
    print(f"refined FP32 vs FP64 positions: {err:.2e} of the diagonal")
    assert err <= 1e-6


def _lib_snh():
    from apple_amd import _lib

    return _lib.APL_SNH


def test_batch_refine_fp64_checked_meets_the_fp64_gradient_target():
    """refine_fp64="checked": afterwards every instance's free-DOF gradient,
    evaluated in FP64, is within rtol |g_0| (the contract of the check)."""
    print("[solver] batch refine_fp64 checked", flush=True)
    insts = []
    for k, n in enumerate((3, 5, 4)):
        m, region, mu, lam, f, fm, x0 = _gravity_case(n, "snh", seed=40 + k)
        insts.append(batch.Instance(region, mu, lam, f, fm, x0))
    cfg = PncgConfig(rtol=1e-7, max_steps=20000)
    b32 = batch.Batch(insts, _lib_snh(), dtype=torch.float32, device="cuda")
    x32, st32 = b32.solve(config=cfg, refine_fp64="checked")
    b64 = b32._fp64_copy()
    g = torch.zeros_like(x32, dtype=torch.float64)
    b64.mesh.eval(b64.material, b64.mu, b64.lam, x32.double(), grad=g)
    g -= b64.f_ext
    free = torch.stack([(b64.mask >> c) & 1 == 0 for c in range(3)], dim=1)
    g = torch.where(free, g, torch.zeros_like(g))
    for i in range(b32.n_inst):
        sl = slice(int(b32.vtx_off[i]), int(b32.vtx_off[i + 1]))
        gn = float(torch.linalg.norm(g[sl]))
        # the FP32 rounding of the returned positions may add a little
        assert gn <= 2 * cfg.rtol * st32[i].grad_norm_first, (i, gn, st32[i].grad_norm_first, b32.refined)
    print(f"refined instances: {b32.refined}")


def test_batch_refine_fp64_follows_replaced_materials():
    """After the FP32 batch's materials are replaced (as adjoint.equilibrium
    and Inverse do), the FP64 check and re-solve use the new values widened,
    not the instances' original ones."""
    print("[solver] batch refine_fp64 after a material update", flush=True)
    insts = []
    for k, n in enumerate((3, 4)):
        m, region, mu, lam, f, fm, x0 = _gravity_case(n, "snh", seed=60 + k)
        insts.append(batch.Instance(region, mu, lam, f, fm, x0))
    cfg = PncgConfig(rtol=1e-7, max_steps=20000)
    b32 = batch.Batch(insts, _lib_snh(), dtype=torch.float32, device="cuda")
    b32.solve(config=cfg, refine_fp64="checked")  # builds the FP64 copy with the original materials
    mats = b32.potential.materials
    mats["mu"].mul_(1.5)
    x32, st32 = b32.solve(config=cfg, refine_fp64="checked")
    b64 = b32._fp64_copy()
    assert torch.equal(b64.potential.materials["mu"], mats["mu"].double())
    g = torch.zeros_like(x32, dtype=torch.float64)
    b64.mesh.eval(b64.material, b64.potential.materials["mu"], b64.potential.materials["lambda"], x32.double(), grad=g)
    g -= b64.f_ext
    free = torch.stack([(b64.mask >> c) & 1 == 0 for c in range(3)], dim=1)
    g = torch.where(free, g, torch.zeros_like(g))
    for i in range(b32.n_inst):
        sl = slice(int(b32.vtx_off[i]), int(b32.vtx_off[i + 1]))
        assert float(torch.linalg.norm(g[sl])) <= 2 * cfg.rtol * st32[i].grad_norm_first, i


def test_batch_stagnated_refine_uses_replaced_materials():
    """refine_fp64="stagnated" re-solves the unfinished instances in FP64 with
    the batch's current materials (here mu replaced by 1.5 mu after
    construction), not the instances' original ones: its positions equal an
    FP64 batch built from the replaced values, bit for bit."""
    import dataclasses

    print("[solver] batch stagnated refine after a material update", flush=True)
    insts = []
    for k, n in enumerate((3, 4)):
        m, region, mu, lam, f, fm, x0 = _gravity_case(n, "snh", seed=70 + k)
        insts.append(batch.Instance(region, mu, lam, f, fm, x0))
    b32 = batch.Batch(insts, _lib_snh(), dtype=torch.float32, device="cuda")
    b32.potential.materials["mu"].mul_(1.5)
    cfg = PncgConfig(rtol=1e-7, max_steps=3)  # every instance ends MAX_STEPS_REACHED
    x32, st32 = b32.solve(config=cfg)
    assert all(st.result == Result.MAX_STEPS_REACHED for st in st32)
    xr, _ = b32.solve(config=cfg, refine_fp64="stagnated")
    assert b32.refined == [0, 1]
    cur = b32.potential.materials["mu"].double().cpu().numpy()
    ref_insts = []
    for i, (inst, part) in enumerate(zip(insts, b32.split(x32))):
        mu_i = cur[int(b32.tet_off[i]):int(b32.tet_off[i + 1])]
        assert not np.array_equal(mu_i, inst.mu)
        np.testing.assert_array_equal(b32._current_instance(i).mu, mu_i)
        ref_insts.append(dataclasses.replace(inst, mu=mu_i, x0=part.double().cpu().numpy()))
    b64 = batch.Batch(ref_insts, _lib_snh(), dtype=torch.float64, device="cuda")
    atol = cfg.rtol * min(st.grad_norm_first for st in st32)
    x64, _ = b64.solve(config=dataclasses.replace(cfg, atol=atol, wolfe_eps=None))
    assert torch.equal(xr, x64.to(torch.float32))


@pytest.mark.parametrize("material", ["snh", "arap", "muscle"])
def test_trajectory_matches_oracle_pncg_step_by_step(material):
    """FP64: the device PNCG after k accepted steps equals the oracle's PNCG
    (oracle/pncg.py, the same Jacobi / Dai-Kou / closed-form alpha / Armijo or
    approximate-Wolfe rules) after k iterations from the same start, for k up
    to 60 -- the per-iteration logic, not only the end state.  The mesh has
    one extra free vertex that no tet touches: its Hessian diagonal is 0, so
    the Jacobi preconditioner's mean fallback is exercised.  The muscle law
    runs with a random activation (A = I + sym(a), |a| ~ 0.1)."""
    print(f"[solver] trajectory vs oracle {material}", flush=True)
    m, region, mu, lam, f, fm, x0 = _gravity_case(4, material, seed=5)
    V = m.n_points + 1  # the dangling vertex
    f = np.vstack([f, np.zeros((1, 3))])
    fm = np.vstack([fm, np.zeros((1, 3), bool)])
    x0 = np.vstack([x0, np.zeros((1, 3))])
    act = 0.1 * np.random.default_rng(6).standard_normal((m.n_cells, 6)) if material == "muscle" else None
    mat = {"snh": of.SNH, "arap": of.ARAP, "muscle": of.MUSCLE}[material]
    opot = of.FemPotential(mat, region.cells, region.dhdX, region.dV, mu, None if mat == of.ARAP else lam,
                           activation=act)
    dof = of.DofMap(fm, np.zeros_like(f))
    prob = op.Problem(of.Model([opot, of.ExternalForce(np.arange(V), f)]), dof)
    from apple_amd.runtime import DeviceMesh

    cls = {"snh": fem.StableNeoHookean, "arap": fem.Arap, "muscle": fem.StableNeoHookeanMuscle}[material]
    dmesh = DeviceMesh(region.cells, region.dhdX[:, 1:], region.dV, V, dtype=torch.float64, device="cuda")
    mats = {"mu": torch.as_tensor(mu, dtype=torch.float64, device="cuda")}
    if material != "arap":
        mats["lambda"] = torch.as_tensor(lam, dtype=torch.float64, device="cuda")
    if material == "muscle":
        mats["activation"] = torch.as_tensor(act, dtype=torch.float64, device="cuda").contiguous()
    pot = cls(dmesh, mats)
    solver = DeviceSolver(pot)
    ft = torch.as_tensor(f, dtype=torch.float64, device="cuda").contiguous()
    mask = mask_bits(fm, "cuda")
    for k in (1, 2, 5, 20, 60):
        x = torch.as_tensor(x0, dtype=torch.float64, device="cuda").contiguous()
        st = solver.solve(x, ft, mask, PncgConfig(rtol=1e-14, max_steps=k, use_graph=False))[0]
        sol = op.pncg(prob.fun, prob.grad, prob.hess_diag, prob.hess_quad, dof.to_free(x0), rtol=1e-14, max_steps=k,
                      wolfe_delta=0.1, wolfe_eps=1e-12)
        xo = dof.to_full(sol.x)
        err = np.abs(x.cpu().numpy() - xo).max() / max(np.abs(xo - x0).max(), 1e-300)
        print(f"  k={k}: device steps {st.n_steps} evals {st.n_evals}, oracle result {sol.result.name}, "
              f"position gap / displacement {err:.2e}")
        assert st.n_steps == k and err <= 1e-8, (k, err)
