"""Stable Neo-Hookean muscle law (G = F A) on the GPU.

Element passes against the FP64 oracle (same tolerances as the SNH / ARAP
parity tests: FP32 energies 1e-5 and gradients, Hessian terms and mixed
derivatives 1e-4 relative; FP64 1e-10), zero activation bitwise equal to the
SNH kernels, the PNCG equilibrium equal to SNH's on the activated rest shape
A^-1 X, and the activation adjoint against the oracle's sparse adjoint and
FP64 central differences.
"""

import numpy as np
import pytest
import torch

import oracle.fem as of
import oracle.pncg as op
from apple_amd import adjoint, fem
from apple_amd.mesh import Region, cube
from apple_amd.runtime import DeviceMesh
from apple_amd.solver import DeviceSolver, PncgConfig, Result, mask_bits
from helpers import rel, rel_max, snh_case

pytestmark = pytest.mark.gpu

TOL = {torch.float32: dict(e=1e-5, g=1e-4), torch.float64: dict(e=1e-10, g=1e-10)}


def _activation(rng, T, scale=0.15):
    return scale * rng.standard_normal((T, 6))


def _muscle_case(n, seed, noise=0.01, act_scale=0.15):
    m, region, u, v = snh_case(n, seed=seed, noise=noise)
    rng = np.random.default_rng(seed + 100)
    act = _activation(rng, m.n_cells, act_scale)
    region.cell_data["activation"] = act
    return m, region, u, v, act


def _outputs(pot, u, v):
    V = pot.n_points
    e = torch.zeros(1, dtype=torch.float64, device=pot.device)
    g = torch.zeros(V, 3, dtype=pot.dtype, device=pot.device)
    h = torch.zeros_like(g)
    pot.fused(u, energy=e, grad=g, hess_diag=h)
    hp = torch.zeros_like(g)
    pot.hess_prod(u, v, hp)
    hq = torch.zeros(1, dtype=torch.float64, device=pot.device)
    pot.hess_quad(u, v, hq)
    mix = pot.mixed_derivative_prod(u, v)
    torch.cuda.synchronize()
    out = dict(e=e.item(), g=g.cpu().numpy(), h=h.cpu().numpy(), hp=hp.cpu().numpy(), hq=hq.item())
    out.update({k: t.cpu().numpy() for k, t in mix.items()})
    return out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("n,chunks", [(4, 1), (10, 1), (12, 4)])
def test_muscle_element_pass_matches_oracle(dtype, n, chunks):
    print(f"[muscle elem] dtype={dtype} n={n} chunks={chunks}", flush=True)
    m, region, u64, v64, act = _muscle_case(n, seed=n)
    pot = fem.StableNeoHookeanMuscle.from_region(region, dtype=dtype, device="cuda", chunks=chunks)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    r = lambda a: np.asarray(a, np.float64).astype(npdt).astype(np.float64)
    u_r, v_r = r(u64), r(v64)
    ref = of.FemPotential(of.MUSCLE, region.cells, region.dhdX, region.dV, r(m.cell_data["mu"]),
                          r(m.cell_data["lambda"]), r(act))
    got = _outputs(pot, torch.as_tensor(u_r, dtype=dtype, device="cuda"), torch.as_tensor(v_r, dtype=dtype, device="cuda"))
    tol = TOL[dtype]
    e_ref = ref.fun(u_r)
    assert abs(got["e"] - e_ref) <= tol["e"] * abs(e_ref)
    assert rel(got["g"], ref.grad(u_r)) <= tol["g"]
    assert rel_max(got["g"], ref.grad(u_r)) <= tol["g"]
    assert rel(got["h"], ref.hess_diag(u_r)) <= tol["g"]
    assert rel_max(got["h"], ref.hess_diag(u_r)) <= tol["g"]
    assert rel(got["hp"], ref.hess_prod(u_r, v_r)) <= tol["g"]
    assert rel_max(got["hp"], ref.hess_prod(u_r, v_r)) <= tol["g"]
    hq_ref = ref.hess_quad(u_r, v_r)
    assert abs(got["hq"] - hq_ref) <= tol["g"] * abs(hq_ref)
    dmu_ref, dlam_ref = ref.mixed_cells(u_r, v_r)
    assert rel(got["mu"], dmu_ref) <= tol["g"]
    assert rel_max(got["mu"], dmu_ref) <= tol["g"]
    assert rel(got["lambda"], dlam_ref) <= tol["g"]
    assert rel_max(got["lambda"], dlam_ref) <= tol["g"]
    assert rel(got["activation"], ref.mixed_activation_cells(u_r, v_r)) <= tol["g"]
    assert rel_max(got["activation"], ref.mixed_activation_cells(u_r, v_r)) <= tol["g"]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_muscle_inverted_tets_match_oracle(dtype):
    """Displacement noise of half a cell on a 6^3 cube: over a fifth of the
    activated deformation gradients G = F A are inverted; every element
    output, the three mixed derivatives included, still matches the oracle."""
    m, region, u64, v64, act = _muscle_case(6, seed=11, noise=0.08)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    r = lambda a: np.asarray(a, np.float64).astype(npdt).astype(np.float64)
    u_r, v_r = r(u64), r(v64)
    ref = of.FemPotential(of.MUSCLE, region.cells, region.dhdX, region.dV, r(m.cell_data["mu"]),
                          r(m.cell_data["lambda"]), r(act))
    det = np.linalg.det(ref.F(u_r) @ of.activation_matrix(r(act)))
    print(f"[muscle elem] inverted {dtype}: {np.mean(det < 0):.1%} of det G < 0", flush=True)
    assert np.mean(det < 0) > 0.2
    pot = fem.StableNeoHookeanMuscle.from_region(region, dtype=dtype, device="cuda")
    got = _outputs(pot, torch.as_tensor(u_r, dtype=dtype, device="cuda"), torch.as_tensor(v_r, dtype=dtype, device="cuda"))
    tol = TOL[dtype]
    e_ref = ref.fun(u_r)
    assert abs(got["e"] - e_ref) <= tol["e"] * abs(e_ref)
    dmu_ref, dlam_ref = ref.mixed_cells(u_r, v_r)
    refs = dict(g=ref.grad(u_r), h=ref.hess_diag(u_r), hp=ref.hess_prod(u_r, v_r), mu=dmu_ref,
                **{"lambda": dlam_ref}, activation=ref.mixed_activation_cells(u_r, v_r))
    for k, want in refs.items():
        print(f"  {k}: rel {rel(got[k], want):.2e}, max-entry {rel_max(got[k], want):.2e}")
        assert rel(got[k], want) <= tol["g"], k
        assert rel_max(got[k], want) <= tol["g"], k
    hq_ref = ref.hess_quad(u_r, v_r)
    assert abs(got["hq"] - hq_ref) <= tol["g"] * abs(hq_ref)


def test_zero_activation_is_bitwise_snh():
    m, region, u64, v64, _ = _muscle_case(6, seed=2)
    region.cell_data["activation"] = np.zeros((m.n_cells, 6))
    mus = fem.StableNeoHookeanMuscle.from_region(region, dtype=torch.float32, device="cuda")
    snh = fem.StableNeoHookean.from_region(region, dtype=torch.float32, device="cuda")
    u = torch.as_tensor(u64, dtype=torch.float32, device="cuda")
    v = torch.as_tensor(v64, dtype=torch.float32, device="cuda")
    a, b = _outputs(mus, u, v), _outputs(snh, u, v)
    for k in b:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_missing_activation_is_an_error():
    m, region, u64, v64, act = _muscle_case(3, seed=1)
    pot = fem.StableNeoHookeanMuscle.from_region(region, dtype=torch.float64, device="cuda")
    pot.mesh.set_activation(None)
    g = torch.zeros(m.n_points, 3, dtype=torch.float64, device="cuda")
    u = torch.as_tensor(u64, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="activation"):
        pot.mesh.eval(pot.material, pot._mu(), pot._lam(), u, grad=g)
    pot.grad(u, g)  # the potential binds its activation again
    torch.cuda.synchronize()
    assert torch.isfinite(g).all()


def _loaded_cube(n, seed):
    m = cube(n)
    rng = np.random.default_rng(seed)
    mu = 1e4 * np.exp(0.3 * rng.standard_normal(m.n_cells))
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = 10 * mu
    region = Region.from_mesh(m)
    f = np.zeros((m.n_points, 3))
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-4e3 * region.dV / 4, 4))
    fixed = np.zeros((m.n_points, 3), bool)
    fixed[m.points[:, 2] == 0] = True
    return m, region, mu, f, fixed


def test_pncg_equals_snh_on_activated_rest_shape():
    """Uniform A: the muscle problem on rest X has the equilibrium positions of
    SNH on rest Y = A^-1 X with the same dV, load and fixed positions."""
    m, region, mu, f, fixed = _loaded_cube(5, 11)
    a = np.array([0.15, -0.1, 0.05, 0.04, -0.03, 0.02])
    act = np.tile(a, (m.n_cells, 1))
    A = of.activation_matrix(a[None])[0]
    X = m.points
    Y = X @ np.linalg.inv(A).T
    dev = "cuda"
    t = lambda z: torch.as_tensor(z, dtype=torch.float64, device=dev).contiguous()
    region.cell_data["activation"] = act
    mus = fem.StableNeoHookeanMuscle.from_region(region, dtype=torch.float64, device=dev)
    dhdY, _ = of.region(Y, m.cells)
    snh_mesh = DeviceMesh(m.cells, dhdY[:, 1:, :], region.dV, m.n_points, dtype=torch.float64, device=dev)
    snh = fem.StableNeoHookean(snh_mesh, {"mu": t(mu), "lambda": t(10 * mu)})
    mask = mask_bits(fixed, dev)
    cfg = PncgConfig(rtol=1e-10, max_steps=40000)
    xm = torch.zeros(m.n_points, 3, dtype=torch.float64, device=dev)
    sm = DeviceSolver(mus).solve(xm, t(f), mask, cfg)[0]
    xs = t(X - Y)  # same initial positions; the fixed ones stay at X
    ss = DeviceSolver(snh).solve(xs, t(f), mask, cfg)[0]
    assert sm.result == Result.PRIMARY_SUCCESS and ss.result == Result.PRIMARY_SUCCESS, (sm, ss)
    pos_m = X + xm.cpu().numpy()
    pos_s = Y + xs.cpu().numpy()
    diag = np.linalg.norm(X.max(0) - X.min(0))
    err = np.abs(pos_m - pos_s).max() / diag
    print(f"muscle vs SNH on A^-1 X: max position difference {err:.2e} of the diagonal")
    assert err <= 1e-7
    # and the oracle's free gradient at the GPU solution is at the solve's tolerance
    opot = of.FemPotential(of.MUSCLE, m.cells, region.dhdX, region.dV, mu, 10 * mu, act)
    g = opot.grad(xm.cpu().numpy()) - f
    g0 = opot.grad(np.zeros_like(X)) - f
    assert np.linalg.norm(g[~fixed]) <= 1e-8 * np.linalg.norm(g0[~fixed])


def _act_problem(n, seed, dtype):
    m, region, mu, f, fixed = _loaded_cube(n, seed)
    rng = np.random.default_rng(seed + 1)
    act = _activation(rng, m.n_cells, 0.1)
    region.cell_data["activation"] = act
    pot = fem.StableNeoHookeanMuscle.from_region(region, dtype=dtype, device="cuda")
    t = lambda z: torch.as_tensor(z, dtype=dtype, device="cuda").contiguous()
    return m, region, pot, mu, act, t(f), fixed, mask_bits(fixed, "cuda")


def test_activation_adjoint_matches_oracle_adjoint():
    m, region, pot, mu, act, f, fixed, mask = _act_problem(3, 21, torch.float64)
    x_t = torch.as_tensor(0.01 * np.random.default_rng(3).standard_normal((m.n_points, 3)), dtype=torch.float64,
                          device="cuda")
    mu_t = torch.as_tensor(mu, dtype=torch.float64, device="cuda")
    act_t = torch.as_tensor(act, dtype=torch.float64, device="cuda").requires_grad_(True)
    x0 = torch.zeros(m.n_points, 3, dtype=torch.float64, device="cuda")
    x = adjoint.equilibrium(pot, mu_t, 10 * mu_t, f, mask, x0, PncgConfig(rtol=1e-11, max_steps=20000),
                            adjoint.AdjointConfig(rtol=1e-12), activation=act_t)
    ((x - x_t) ** 2).sum().backward()
    g = act_t.grad.cpu().numpy()
    opot = of.FemPotential(of.MUSCLE, m.cells, region.dhdX, region.dV, mu, 10 * mu, act)
    ext = of.ExternalForce(np.arange(m.n_points), f.cpu().numpy())
    dof = of.DofMap(fixed, np.zeros((m.n_points, 3)))
    prob = op.Problem(of.Model([opot, ext]), dof)
    xe = op.newton(prob, dof.to_free(x.detach().cpu().numpy()), tol=1e-11)
    dLdx = dof.to_free(2 * (dof.to_full(xe) - x_t.cpu().numpy()))
    _, _, ref = op.adjoint_grad_activation(prob, opot, xe, dLdx)
    np.testing.assert_allclose(g, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max())


def test_activation_adjoint_matches_central_differences():
    m, region, pot, mu, act, f, fixed, mask = _act_problem(3, 31, torch.float64)
    x_t = torch.as_tensor(0.01 * np.random.default_rng(4).standard_normal((m.n_points, 3)), dtype=torch.float64,
                          device="cuda")
    cfg = PncgConfig(rtol=1e-12, max_steps=50000)
    mu_t = torch.as_tensor(mu, dtype=torch.float64, device="cuda")
    act_t = torch.as_tensor(act, dtype=torch.float64, device="cuda").requires_grad_(True)
    x0 = torch.zeros(m.n_points, 3, dtype=torch.float64, device="cuda")
    x = adjoint.equilibrium(pot, mu_t, 10 * mu_t, f, mask, x0, cfg, adjoint.AdjointConfig(rtol=1e-12),
                            activation=act_t)
    ((x - x_t) ** 2).sum().backward()
    g = act_t.grad.cpu().numpy()
    x_eq = x.detach()

    def loss_at(a_np):
        at = torch.as_tensor(a_np, dtype=torch.float64, device="cuda")
        with torch.no_grad():
            xx = adjoint.equilibrium(pot, mu_t, 10 * mu_t, f, mask, x_eq, cfg, activation=at)
        return float(((xx - x_t) ** 2).sum())

    rng = np.random.default_rng(6)
    h = 1e-5
    for t in rng.choice(m.n_cells, 3, replace=False):
        for k in (0, 3, 5):
            ap = act.copy()
            ap[t, k] += h
            am = act.copy()
            am[t, k] -= h
            fd = (loss_at(ap) - loss_at(am)) / (2 * h)
            assert g[t, k] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(g).max())


def test_batched_muscle_equals_single_solves():
    from apple_amd import _lib, batch

    insts = []
    for k, n in enumerate((3, 5)):
        m, region, mu, f, fixed = _loaded_cube(n, 50 + k)
        act = _activation(np.random.default_rng(60 + k), m.n_cells, 0.1)
        region.cell_data["activation"] = act
        insts.append(batch.Instance(region, mu, 10 * mu, f, fixed, np.zeros((m.n_points, 3)), activation=act))
    cfg = PncgConfig(rtol=1e-10, max_steps=20000)
    b = batch.Batch(insts, _lib.APL_SNH_MUSCLE, dtype=torch.float64, device="cuda")
    xb, stats = b.solve(config=cfg)
    assert all(s.success for s in stats), stats
    for inst, xi in zip(insts, b.split(xb)):
        pot = fem.StableNeoHookeanMuscle.from_region(inst.region, dtype=torch.float64, device="cuda")
        x = torch.zeros_like(xi)
        t = lambda z: torch.as_tensor(z, dtype=torch.float64, device="cuda").contiguous()
        st = DeviceSolver(pot).solve(x, t(inst.f_ext), mask_bits(inst.fixed_mask, "cuda"), cfg)[0]
        assert st.success
        diag = np.linalg.norm(np.ptp(inst.region.mesh.points, axis=0))
        assert (xi - x).abs().max().item() <= 1e-8 * diag
