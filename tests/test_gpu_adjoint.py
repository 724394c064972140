"""Adjoint gradient of the equilibrium w.r.t. per-tet materials, on the GPU.

Checked three ways: against the oracle's adjoint (sparse direct solve of the
exact FP64 Hessian at the same equilibrium), against FP64 central finite
differences of the loss through GPU equilibrium solves, and FP32 against FP64
(the error is stated in the assertion).  lambda tied to mu (lambda = 10 mu)
goes through autograd.
"""

import numpy as np
import pytest
import torch

import oracle.fem as of
import oracle.pncg as op
from apple_amd import adjoint, batch, fem
from apple_amd.mesh import Region, cube
from apple_amd.solver import PncgConfig, mask_bits
from helpers import rel

pytestmark = pytest.mark.gpu


def _problem(n, seed, dtype):
    m = cube(n)
    rng = np.random.default_rng(seed)
    mu = 1e4 * np.exp(0.5 * rng.standard_normal(m.n_cells))
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = 10 * mu
    region = Region.from_mesh(m)
    pot = fem.StableNeoHookean.from_region(region, dtype=dtype, device="cuda")
    f = np.zeros((m.n_points, 3))
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e3 * region.dV / 4, 4))
    fixed = np.zeros((m.n_points, 3), bool)
    fixed[m.points[:, 2] == 0] = True
    t = lambda a: torch.as_tensor(a, dtype=dtype, device="cuda").contiguous()
    return m, region, pot, mu, t(f), fixed, mask_bits(fixed, "cuda")


def _target(m, seed):
    rng = np.random.default_rng(seed)
    return 0.01 * rng.standard_normal((m.n_points, 3))


def test_adjoint_matches_oracle_adjoint():
    m, region, pot, mu, f, fixed, mask = _problem(3, 1, torch.float64)
    x_t = torch.as_tensor(_target(m, 2), dtype=torch.float64, device="cuda")
    mu_t = torch.as_tensor(mu, dtype=torch.float64, device="cuda").requires_grad_(True)
    lam_t = (10 * mu_t)
    x0 = torch.zeros(m.n_points, 3, dtype=torch.float64, device="cuda")
    x = adjoint.equilibrium(pot, mu_t, lam_t, f, mask, x0, PncgConfig(rtol=1e-11, max_steps=20000),
                            adjoint.AdjointConfig(rtol=1e-12))
    loss = ((x - x_t) ** 2).sum()
    loss.backward()
    g = mu_t.grad.cpu().numpy()
    # oracle: same equilibrium (Newton-polished), sparse adjoint
    opot = of.FemPotential(of.SNH, m.cells, region.dhdX, region.dV, mu, 10 * mu)
    ext = of.ExternalForce(np.arange(m.n_points), f.cpu().numpy())
    dof = of.DofMap(fixed, np.zeros((m.n_points, 3)))
    prob = op.Problem(of.Model([opot, ext]), dof)
    xe = op.newton(prob, dof.to_free(x.detach().cpu().numpy()), tol=1e-11)
    dLdx = dof.to_free(2 * (dof.to_full(xe) - x_t.cpu().numpy()))
    dmu, dlam = op.adjoint_grad(prob, opot, xe, dLdx)
    ref = dmu + 10 * dlam
    np.testing.assert_allclose(g, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max())


def test_adjoint_matches_central_differences():
    m, region, pot, mu, f, fixed, mask = _problem(4, 3, torch.float64)
    x_t = torch.as_tensor(_target(m, 4), dtype=torch.float64, device="cuda")
    cfg = PncgConfig(rtol=1e-12, max_steps=50000)
    x0 = torch.zeros(m.n_points, 3, dtype=torch.float64, device="cuda")
    mu_t = torch.as_tensor(mu, dtype=torch.float64, device="cuda").requires_grad_(True)
    x = adjoint.equilibrium(pot, mu_t, 10 * mu_t, f, mask, x0, cfg, adjoint.AdjointConfig(rtol=1e-12))
    ((x - x_t) ** 2).sum().backward()
    g = mu_t.grad.cpu().numpy()
    x_eq = x.detach()

    def loss_at(mu_np):
        mt = torch.as_tensor(mu_np, dtype=torch.float64, device="cuda")
        with torch.no_grad():
            xx = adjoint.equilibrium(pot, mt, 10 * mt, f, mask, x_eq, cfg)
        return float(((xx - x_t) ** 2).sum())

    rng = np.random.default_rng(5)
    for t in rng.choice(m.n_cells, 6, replace=False):
        h = 1e-4 * mu[t]
        mp = mu.copy()
        mp[t] += h
        mm = mu.copy()
        mm[t] -= h
        fd = (loss_at(mp) - loss_at(mm)) / (2 * h)
        assert g[t] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(g).max())


def test_fp32_adjoint_close_to_fp64():
    grads = {}
    for dtype in (torch.float64, torch.float32):
        m, region, pot, mu, f, fixed, mask = _problem(6, 7, dtype)
        x_t = torch.as_tensor(_target(m, 8), dtype=dtype, device="cuda")
        mu_t = torch.as_tensor(mu, dtype=dtype, device="cuda").requires_grad_(True)
        x0 = torch.zeros(m.n_points, 3, dtype=dtype, device="cuda")
        rtol = 1e-10 if dtype == torch.float64 else 1e-6
        x = adjoint.equilibrium(pot, mu_t, 10 * mu_t, f, mask, x0, PncgConfig(rtol=rtol, max_steps=20000),
                                adjoint.AdjointConfig(rtol=1e-10 if dtype == torch.float64 else 1e-6))
        ((x - x_t) ** 2).sum().backward()
        grads[dtype] = mu_t.grad.double().cpu().numpy()
    err = np.linalg.norm(grads[torch.float32] - grads[torch.float64]) / np.linalg.norm(grads[torch.float64])
    print(f"FP32 adjoint relative error vs FP64: {err:.2e}")
    assert err <= 1e-3


def test_batched_adjoint_equals_single():
    insts, mus, targets = [], [], []
    for k, n in enumerate((3, 4)):
        m = cube(n)
        rng = np.random.default_rng(20 + k)
        mu = 1e4 * np.exp(0.5 * rng.standard_normal(m.n_cells))
        region = Region.from_mesh(m)
        f = np.zeros((m.n_points, 3))
        np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e3 * region.dV / 4, 4))
        fixed = np.zeros((m.n_points, 3), bool)
        fixed[m.points[:, 2] == 0] = True
        insts.append(batch.Instance(region, mu, 10 * mu, f, fixed, np.zeros((m.n_points, 3))))
        mus.append(mu)
        targets.append(_target(m, 30 + k))
    cfg = PncgConfig(rtol=1e-10, max_steps=20000)
    acfg = adjoint.AdjointConfig(rtol=1e-10)

    def run(b, mu_np, tgt):
        mu_t = torch.as_tensor(mu_np, dtype=torch.float64, device="cuda").requires_grad_(True)
        x = adjoint.equilibrium(b.potential, mu_t, 10 * mu_t, b.f_ext, b.mask, b.x0, cfg, acfg)
        ((x - torch.as_tensor(tgt, dtype=torch.float64, device="cuda")) ** 2).sum().backward()
        return mu_t.grad.cpu().numpy()

    bb = batch.Batch(insts, 0, dtype=torch.float64, device="cuda")
    g_b = run(bb, np.concatenate(mus), np.concatenate(targets))
    off = 0
    for inst, mu, tgt in zip(insts, mus, targets):
        single = batch.Batch([inst], 0, dtype=torch.float64, device="cuda")
        g_s = run(single, mu, tgt)
        np.testing.assert_allclose(g_b[off : off + mu.size], g_s, rtol=1e-6, atol=1e-12)
        off += mu.size


def test_batched_fp32_vs_fp64_forward_and_adjoint():
    """Small version of the combined case: one batch, FP32 element kernels with
    the FP64 iterate and reductions against the same batch in FP64.  Errors
    per instance: energy <= 1e-5 relative, positions <= 1e-5 of the bounding-box
    diagonal, dL/d(mu, lambda) <= 2e-4 relative (measured 2e-5 .. 6e-5; the FP32
    adjoint is limited by the conditioning of H at the FP32 rounding floor)."""
    from apple_amd import _lib

    insts = []
    for k, n in enumerate((5, 7, 6, 8)):
        m = cube(n)
        region = Region.from_mesh(m)
        rng = np.random.default_rng(40 + k)
        mu = rng.uniform(1e4, 1e5) * np.ones(m.n_cells)
        nu = rng.uniform(0.3, 0.45)
        f = np.zeros((m.n_points, 3))
        np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e3 * region.dV / 4, 4))
        fixed = np.zeros((m.n_points, 3), bool)
        fixed[m.points[:, 2] == 0] = True
        x0 = 0.1 / n * rng.standard_normal((m.n_points, 3))
        x0[fixed] = 0
        insts.append(batch.Instance(region, mu, 2 * mu * nu / (1 - 2 * nu), f, fixed, x0))
    out = {}
    for dt, rtol in ((torch.float32, 1e-6), (torch.float64, 1e-10)):
        b = batch.Batch(insts, _lib.APL_SNH, dtype=dt, device="cuda")
        solver = b.solver()
        x = b.x0.clone()
        st = solver.solve(x, b.f_ext, b.mask, PncgConfig(rtol=rtol, max_steps=20000))
        p = torch.zeros_like(x)
        cs = solver.cg(x, (b.x0 - x).contiguous(), p, b.mask, rtol=rtol, max_iters=20000)
        assert all(s.result.name == "PRIMARY_SUCCESS" for s in st + cs), (st, cs)
        dmu = torch.empty(b.n_tets, dtype=dt, device="cuda")
        dlam = torch.empty_like(dmu)
        b.mesh.eval(_lib.APL_SNH, b.mu, b.lam, x, p, dmu=dmu, dlam=dlam)
        out[dt] = (b, x.double().cpu().numpy(), [s.f for s in st], torch.cat([dmu, dlam]).double().cpu().numpy())
    b, x32, f32, g32 = out[torch.float32]
    _, x64, f64, g64 = out[torch.float64]
    T = b.n_tets
    for i in range(len(insts)):
        sv = slice(int(b.vtx_off[i]), int(b.vtx_off[i + 1]))
        st_ = np.r_[b.tet_off[i]:b.tet_off[i + 1], T + b.tet_off[i]:T + b.tet_off[i + 1]]
        e_err = abs(f32[i] - f64[i]) / abs(f64[i])
        x_err = np.abs(x32[sv] - x64[sv]).max() / np.sqrt(3.0)
        g_err = np.linalg.norm(g32[st_] - g64[st_]) / np.linalg.norm(g64[st_])
        print(f"instance {i}: energy {e_err:.2e}, positions/diag {x_err:.2e}, dL/dq {g_err:.2e}")
        assert e_err <= 1e-5 and x_err <= 1e-5 and g_err <= 2e-4


def test_normal_cg_solves_an_indefinite_hessian():
    """Far from equilibrium the SNH Hessian is indefinite: the PCG meets
    negative curvature and reports NON_FINITE, and CG on the normal
    equations (apl_cg_solve_normal, lineax.NormalCG's method) solves H y = b
    to the oracle's sparse direct solution."""
    import scipy.sparse.linalg as spla

    from apple_amd.solver import DeviceSolver, Result, mask_bits

    print("[adjoint] normal CG on an indefinite Hessian", flush=True)
    m = cube(3)
    rng = np.random.default_rng(12)
    mu = np.full(m.n_cells, 1e4)
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = 10 * mu
    region = Region.from_mesh(m)
    fm = np.zeros((m.n_points, 3), bool)
    fm[m.points[:, 2] == 0] = True
    u = 0.25 * rng.standard_normal((m.n_points, 3))
    u[fm] = 0
    b = rng.standard_normal((m.n_points, 3))
    b[fm] = 0
    dof = of.DofMap(fm, np.zeros((m.n_points, 3)))
    prob = op.Problem(of.Model([of.FemPotential(of.SNH, region.cells, region.dhdX, region.dV, mu, 10 * mu)]), dof)
    H = prob.hessian(dof.to_free(u))
    ev = np.linalg.eigvalsh(H.toarray())
    assert ev.min() < 0 < ev.max(), "the test state must have an indefinite Hessian"
    y_ref = dof.to_full_grad(spla.spsolve(H.tocsc(), dof.to_free(b)))
    pot = fem.StableNeoHookean.from_region(region, dtype=torch.float64, device="cuda")
    solver = DeviceSolver(pot)
    ut = torch.as_tensor(u, device="cuda")
    bt = torch.as_tensor(b, device="cuda")
    mask = mask_bits(fm, "cuda")
    y = torch.zeros_like(ut)
    st = solver.cg(ut, bt, y, mask, rtol=1e-10, max_iters=20000)[0]
    assert st.result == Result.NON_FINITE
    y = torch.zeros_like(ut)
    st = solver.cg(ut, bt, y, mask, rtol=1e-12, max_iters=200000, normal=True)[0]
    assert st.result == Result.PRIMARY_SUCCESS, st
    assert rel(y.cpu().numpy(), y_ref) <= 1e-6
