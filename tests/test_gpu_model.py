"""Sums of FEM potentials on the fused device path (DeviceModel), on the GPU.

The reference's inverse target puts three FEM energies on one mesh, each over
every cell with its own Fraction and materials: fat (stable Neo-Hookean),
aponeurosis and muscle (stable Neo-Hookean muscle)
(exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:256-279).
The same construction here, through ModelBuilder, against the oracle's sum
of potentials (oracle.fem.Model):

* element passes of the model mesh (energy, gradient, Hessian diagonal,
  Hessian-vector product, quadratic form) = the oracle's sums;
* Forward.step() (the fused, graph-captured PNCG) reaches the oracle's
  Newton equilibrium started from the same x0;
* per-potential material gradients through adjoint.equilibrium = the
  oracle's sparse adjoint;
* a model over two meshes (global point ids) = the oracle on the joint mesh;
* the step-wise optimizer protocol (init / step / terminate / postprocess)
  ends where minimize() ends, with the reference's state field names.
"""

import numpy as np
import pytest
import torch

import oracle.fem as of
import oracle.pncg as op
from apple_amd import adjoint, fem, forward, potential
from apple_amd.mesh import Mesh, Region, cube
from apple_amd.model import DeviceModel
from apple_amd.solver import PncgConfig, Result
from helpers import rel, rel_max

pytestmark = pytest.mark.gpu


def _fractions(T, rng):
    muscle = rng.uniform(0, 1, T) * (rng.uniform(size=T) > 0.4)
    smas = np.clip(muscle + rng.uniform(0, 0.6, T), 0, 1)
    return np.clip(1.0 - smas, 0, None), np.clip(smas - muscle, 0, None), muscle


def _smas_builder(n, dtype, seed=0):
    """ModelBuilder with the three potentials of the reference's smas model
    on an n^3 cube, z = 0 fixed, gravity; returns (builder, mesh, oracle parts)."""
    m = cube(n)
    T = m.n_cells
    rng = np.random.default_rng(seed)
    fat, apo, mus = _fractions(T, rng)
    act = 0.15 * rng.standard_normal((T, 6))
    builder = forward.ModelBuilder()
    builder.add_vertices(m)
    fm = np.zeros((m.n_points, 3), bool)
    fm[m.points[:, 2] == 0] = True
    m.point_data["FixedMask"] = fm
    m.point_data["FixedValue"] = np.zeros((m.n_points, 3))
    builder.add_fixed(m)
    region = Region.from_mesh(m)
    parts = []
    m.cell_data["activation"] = act
    for name, cls, frac, mu, lam in (("fat", fem.StableNeoHookean, fat, 1e4, 1e5),
                                     ("aponeurosis", fem.StableNeoHookeanMuscle, apo, 3e4, 3e5),
                                     ("muscle", fem.StableNeoHookeanMuscle, mus, 3e4, 3e5)):
        m.cell_data["Fraction"] = frac
        m.cell_data["mu"] = np.full(T, mu)
        m.cell_data["lambda"] = np.full(T, lam)
        builder.add_potential(cls.from_pyvista(m, dtype=dtype, device="cuda", name=name))
        mat = of.SNH if cls is fem.StableNeoHookean else of.MUSCLE
        parts.append(of.FemPotential(mat, region.cells, region.dhdX, region.dV * frac, np.full(T, mu),
                                     np.full(T, lam), act if mat == of.MUSCLE else None))
    f = np.zeros((m.n_points, 3))
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e2 * region.dV / 4, 4))  # rho 100: a moderate sag
    builder.add_potential(potential.ExternalForce(np.arange(m.n_points), f, m.n_points, dtype=dtype, device="cuda"))
    return builder, m, parts, f, fm


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_model_element_passes_match_oracle_sum(dtype):
    print(f"[model] element passes dtype={dtype}", flush=True)
    builder, m, parts, f, fm = _smas_builder(6, dtype)
    model = builder.finalize()
    dm = model.device_model()
    assert isinstance(dm, DeviceModel) and dm.n_parts == 3
    npdt = np.float32 if dtype == torch.float32 else np.float64
    rng = np.random.default_rng(1)
    u64 = (0.01 * rng.standard_normal((m.n_points, 3))).astype(npdt).astype(np.float64)
    v64 = rng.standard_normal((m.n_points, 3)).astype(npdt).astype(np.float64)
    # the oracle sees the materials the device holds (rounded to dtype)
    ref = of.Model([of.FemPotential(p.material, p.cells, p.dhdX, p.dV, p.mu.astype(npdt), p.lam.astype(npdt),
                                    None if p.activation is None else p.activation.astype(npdt)) for p in parts])
    u = torch.as_tensor(u64, dtype=dtype, device="cuda")
    v = torch.as_tensor(v64, dtype=dtype, device="cuda")
    mats = dm.materials
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    g, h, hp = torch.zeros_like(u), torch.zeros_like(u), torch.zeros_like(u)
    hq = torch.zeros(1, dtype=torch.float64, device="cuda")
    dm.mesh.set_activation(mats["activation"])
    dm.mesh.eval(dm.material, mats["mu"], mats["lambda"], u, v, energy=e, grad=g, hess_diag=h, hess_prod=hp,
                 hess_quad=hq)
    torch.cuda.synchronize()
    tol_e, tol_g = (1e-5, 1e-4) if dtype == torch.float32 else (1e-10, 1e-10)
    assert abs(e.item() - ref.fun(u64)) <= tol_e * abs(ref.fun(u64))
    assert rel(g.cpu().numpy(), ref.grad(u64)) <= tol_g
    assert rel_max(g.cpu().numpy(), ref.grad(u64)) <= tol_g
    assert rel(h.cpu().numpy(), ref.hess_diag(u64)) <= tol_g
    assert rel_max(h.cpu().numpy(), ref.hess_diag(u64)) <= tol_g
    assert rel(hp.cpu().numpy(), ref.hess_prod(u64, v64)) <= tol_g
    assert rel_max(hp.cpu().numpy(), ref.hess_prod(u64, v64)) <= tol_g
    assert abs(hq.item() - ref.hess_quad(u64, v64)) <= tol_g * abs(ref.hess_quad(u64, v64))
    # the reference protocol (one call per potential, ExternalForce included)
    # gives the same sum
    gm = model.grad(u)
    assert rel(gm.cpu().numpy(), ref.grad(u64) - f) <= 2 * tol_g
    assert rel_max(gm.cpu().numpy(), ref.grad(u64) - f) <= 2 * tol_g


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-8), (torch.float32, 1e-5)])
def test_model_forward_step_matches_oracle_newton(dtype, tol):
    print(f"[model] forward dtype={dtype}", flush=True)
    builder, m, parts, f, fm = _smas_builder(5, dtype, seed=2)
    model = builder.finalize()
    rtol = 1e-10 if dtype == torch.float64 else 1e-6
    fwd = forward.Forward(model, forward.PNCG(rtol=rtol, max_steps=20000))
    sol = fwd.step()
    assert sol.result == Result.PRIMARY_SUCCESS, sol.stats
    x = fwd.state.u.double().cpu().numpy()
    dof = of.DofMap(fm, np.zeros_like(f))
    prob = op.Problem(of.Model(parts + [of.ExternalForce(np.arange(m.n_points), f)]), dof)
    x_ref = dof.to_full(op.newton(prob, np.zeros(dof.n_free), tol=1e-9, max_steps=40))
    err = np.abs(x - x_ref).max() / np.sqrt(3.0)
    print(f"  steps {sol.stats.n_steps}, pos err / diag {err:.2e}")
    assert err <= tol
    # reference result fields: the free gradient at the solution, relative decrease
    assert sol.state.grad.shape == (model.n_free,)
    assert sol.stats.relative_decrease <= rtol
    gn = float(torch.linalg.norm(sol.state.grad.double()))
    # FP32: the gradient of the FP32 rounding of the FP64 iterate, in FP32
    assert gn <= (10 * rtol if dtype == torch.float64 else 1e-3) * sol.stats.grad_norm_first


def test_model_adjoint_per_potential_gradients_match_oracle():
    """dL/d(mu, lambda, activation) of every potential through one model."""
    print("[model] adjoint", flush=True)
    dtype = torch.float64
    builder, m, parts, f, fm = _smas_builder(4, dtype, seed=3)
    model = builder.finalize()
    dm = model.device_model()
    dof = of.DofMap(fm, np.zeros_like(f))
    prob = op.Problem(of.Model(parts + [of.ExternalForce(np.arange(m.n_points), f)]), dof)
    x_ref = op.newton(prob, np.zeros(dof.n_free), tol=1e-11, max_steps=40)
    rng = np.random.default_rng(5)
    target = dof.to_full(x_ref) + 1e-3 * rng.standard_normal((m.n_points, 3))
    dLdx = dof.to_free(2 * (dof.to_full(x_ref) - target))
    H = prob.hessian(x_ref)
    import scipy.sparse.linalg as spla

    p = dof.to_full_grad(spla.spsolve(H.tocsc(), -dLdx))
    u = dof.to_full(x_ref)
    # device: per-part material tensors -> model arrays -> equilibrium -> loss
    leaves = [{k: v.detach().clone().requires_grad_(True) for k, v in part.materials.items()} for part in dm.parts]
    mats = dm.gather(leaves)
    ft = model.external_force()
    x = adjoint.equilibrium(dm, mats["mu"], mats["lambda"], ft, model.dof_map.mask_bits,
                            torch.zeros(m.n_points, 3, dtype=dtype, device="cuda"),
                            PncgConfig(rtol=1e-11, max_steps=20000), adjoint.AdjointConfig(rtol=1e-12),
                            activation=mats["activation"])
    tt = torch.as_tensor(target, dtype=dtype, device="cuda")
    ((x - tt) ** 2).sum().backward()
    for k, (part, ref) in enumerate(zip(dm.parts, parts)):
        dmu_ref, dlam_ref = ref.mixed_cells(u, p)
        assert rel(leaves[k]["mu"].grad.cpu().numpy(), dmu_ref) <= 1e-6, part.name
        assert rel_max(leaves[k]["mu"].grad.cpu().numpy(), dmu_ref) <= 1e-6, part.name
        assert rel(leaves[k]["lambda"].grad.cpu().numpy(), dlam_ref) <= 1e-6, part.name
        assert rel_max(leaves[k]["lambda"].grad.cpu().numpy(), dlam_ref) <= 1e-6, part.name
        if ref.material == of.MUSCLE:
            assert rel(leaves[k]["activation"].grad.cpu().numpy(), ref.mixed_activation_cells(u, p)) <= 1e-6
            assert rel_max(leaves[k]["activation"].grad.cpu().numpy(), ref.mixed_activation_cells(u, p)) <= 1e-6


def test_two_mesh_model_with_global_point_ids():
    """Two cubes added to one ModelBuilder (global ids offset the second), one
    SNH and one ARAP potential: the fused model solve equals the oracle on the
    joint vertex array."""
    print("[model] two meshes", flush=True)
    dtype = torch.float64
    a, b = cube(3), cube(4)
    b.points = b.points + np.array([2.0, 0.0, 0.0])
    builder = forward.ModelBuilder()
    oracle_parts, f_all, fm_all = [], [], []
    off = 0
    for mesh, cls, mu in ((a, fem.StableNeoHookean, 1e4), (b, fem.Arap, 2e4)):
        builder.add_vertices(mesh)
        fm = np.zeros((mesh.n_points, 3), bool)
        fm[mesh.points[:, 2] == 0] = True
        mesh.point_data["FixedMask"] = fm
        mesh.point_data["FixedValue"] = np.zeros((mesh.n_points, 3))
        builder.add_fixed(mesh)
        mesh.cell_data["mu"] = np.full(mesh.n_cells, mu)
        mesh.cell_data["lambda"] = np.full(mesh.n_cells, 10 * mu)
        builder.add_potential(cls.from_pyvista(mesh, dtype=dtype, device="cuda"))
        r = Region.from_mesh(mesh)
        mat = of.SNH if cls is fem.StableNeoHookean else of.ARAP
        oracle_parts.append(of.FemPotential(mat, r.cells + off, r.dhdX, r.dV, np.full(mesh.n_cells, mu),
                                            np.full(mesh.n_cells, 10 * mu) if mat == of.SNH else None))
        f = np.zeros((mesh.n_points, 3))
        np.add.at(f[:, 2], mesh.cells.reshape(-1), np.repeat(-9.8e3 * r.dV / 4, 4))
        f_all.append(f)
        fm_all.append(fm)
        off += mesh.n_points
    f = np.concatenate(f_all)
    fm = np.concatenate(fm_all)
    builder.add_potential(potential.ExternalForce(np.arange(off), f, off, dtype=dtype, device="cuda"))
    model = builder.finalize()
    assert model.n_points == off
    fwd = forward.Forward(model, forward.PNCG(rtol=1e-10, max_steps=20000))
    sol = fwd.step()
    assert sol.result == Result.PRIMARY_SUCCESS
    dof = of.DofMap(fm, np.zeros_like(f))
    prob = op.Problem(of.Model(oracle_parts + [of.ExternalForce(np.arange(off), f)]), dof)
    x_ref = dof.to_full(op.newton(prob, np.zeros(dof.n_free), tol=1e-10, max_steps=40))
    assert np.abs(fwd.state.u.cpu().numpy() - x_ref).max() <= 1e-8
    # the per-potential reference protocol over the global arrays agrees too
    u = fwd.state.u
    g = model.grad(u)
    g_ref = of.Model(oracle_parts + [of.ExternalForce(np.arange(off), f)]).grad(x_ref)
    assert np.abs(g.cpu().numpy() - g_ref).max() <= 1e-6 * np.abs(f).max()


@pytest.mark.parametrize("multi", [False, True])
def test_stepwise_optimizer_protocol_matches_minimize(multi):
    """init / step / terminate / postprocess (one accepted step per call,
    exp/2026/05/06/toy/src/21-smas-prestrain-stable-neo-hookean-muscle.py:497-564)
    end bitwise where minimize() ends; every step's state carries the fields
    that experiment logs."""
    print(f"[model] step-wise multi={multi}", flush=True)
    dtype = torch.float32
    if multi:
        builder, m, parts, f, fm = _smas_builder(4, dtype, seed=7)
    else:
        m = cube(4)
        m.cell_data["mu"] = np.full(m.n_cells, 1e4)
        m.cell_data["lambda"] = np.full(m.n_cells, 1e5)
        builder = forward.ModelBuilder()
        builder.add_vertices(m)
        fm = np.zeros((m.n_points, 3), bool)
        fm[m.points[:, 2] == 0] = True
        m.point_data["FixedMask"] = fm
        m.point_data["FixedValue"] = np.zeros((m.n_points, 3))
        builder.add_fixed(m)
        r = Region.from_mesh(m)
        builder.add_potential(fem.StableNeoHookean.from_region(r, dtype=dtype, device="cuda"))
        builder.add_potential(potential.gravity(r, 1e3, dtype=dtype, device="cuda"))
    model = builder.finalize()
    cfg = dict(rtol=1e-6, max_steps=5000)
    ref = forward.Forward(model, forward.PNCG(**cfg)).step()
    fwd = forward.Forward(model, forward.PNCG(**cfg))
    opt = fwd.optimizer
    problem = fwd.problem
    state = fwd.state
    opt_state = opt.init(problem, state, fwd.free)
    assert opt_state.n_steps == 0
    steps = 0
    while True:
        state, opt_state = opt.step(problem, state, opt_state)
        steps += 1
        ls, conv, damp = opt_state.line_search_state, opt_state.convergence_state, opt_state.hess_damping_state
        assert opt_state.n_steps == steps or opt_state.result != Result.RUNNING
        assert np.isfinite(ls.alpha) and ls.step >= 0 and isinstance(ls.ok, bool)
        assert ls.f_alpha <= ls.f0 + 1e-6 * abs(ls.f0)
        assert conv.grad_norm_first > 0 and damp.hess_diag_mean > 0 and damp.factor == 0.0
        assert opt_state.direction is not None and opt_state.direction.shape == (model.n_free,)
        done, result = opt.terminate(problem, state, opt_state)
        if done:
            break
    sol = opt.postprocess(problem, state, opt_state, result)
    assert sol.result == ref.result == Result.PRIMARY_SUCCESS
    assert sol.stats.n_steps == ref.stats.n_steps == steps
    assert torch.equal(sol.state.u, ref.state.u)
    assert sol.stats.relative_decrease == ref.stats.relative_decrease


def test_inverse_on_a_three_potential_model_recovers_the_activation():
    """Inverse (Adam over one log-activation parameter of the muscle
    potential, the smas experiment's parameterisation) on the fat +
    aponeurosis + muscle model: the loss falls by orders of magnitude and the
    parameter approaches the one that made the target."""
    print("[model] inverse", flush=True)
    from apple_amd import inverse

    dtype = torch.float32
    builder, m, parts, f, fm = _smas_builder(5, dtype, seed=11)
    model = builder.finalize()
    dm = model.device_model()
    muscle = dm.parts[2]
    base_act = muscle.materials["activation"].clone()
    prestrain = 0.9

    def make(p):
        act = inverse.muscle_x_activation(p.expand(muscle.n_cells), prestrain=prestrain)
        return [None, None, {"activation": act.to(dtype)}]

    ft = model.external_force()
    mask = model.dof_map.mask_bits
    p_true = torch.tensor([0.3], device="cuda")
    mats = dm.gather([dict(pt.materials, **(d or {})) for pt, d in zip(dm.parts, make(p_true))])
    x_true = adjoint.equilibrium(dm, mats["mu"], mats["lambda"], ft, mask,
                                 torch.zeros(m.n_points, 3, dtype=dtype, device="cuda"),
                                 PncgConfig(rtol=1e-6, max_steps=20000), activation=mats["activation"]).detach()
    idx = torch.arange(m.n_points, device="cuda")
    loss = inverse.PointToPointLoss(idx, x_true)
    inv = inverse.Inverse(dm, [loss], make, f_ext=ft, mask=mask, lr=0.05,
                          forward=PncgConfig(rtol=1e-5, max_steps=20000))
    p = inv.solve(torch.tensor([0.0], device="cuda"), steps=40)
    first, last = inv.history[0].loss, inv.history[-1].loss
    print(f"  loss {first:.3e} -> {last:.3e}, p = {float(p):.4f}")
    assert last <= 1e-2 * first
    assert abs(float(p) - 0.3) <= 0.03
    assert torch.equal(dm.parts[2].materials["activation"], base_act)  # the parts keep their own values


def test_batch_of_models_equals_each_model_alone():
    """Three fat + aponeurosis + muscle models of different sizes as one
    batch (Batch.from_models): every instance equals its model solved alone
    through Forward.step(), bitwise, and the batched adjoint runs."""
    print("[model] batch of models", flush=True)
    from apple_amd import _lib, batch

    dtype = torch.float32
    models = [_smas_builder(n, dtype, seed=20 + n)[0].finalize() for n in (3, 5, 4)]
    b = batch.Batch.from_models(models, dtype=dtype, device="cuda")
    assert b.material == _lib.APL_MODEL and b.n_inst == 3
    cfg = PncgConfig(rtol=1e-6, max_steps=20000)
    x, st = b.solve(config=cfg)
    parts = b.split(x)
    for k, model in enumerate(models):
        fwd = forward.Forward(model, forward.PNCG(cfg))
        sol = fwd.step()
        assert st[k].result == sol.result == Result.PRIMARY_SUCCESS
        assert st[k].n_steps == sol.stats.n_steps
        assert torch.equal(parts[k], fwd.state.u), k
    rhs = (b.x0 - x).contiguous()
    p, cs = b.adjoint_solve(x, rhs, rtol=1e-5)
    assert all(s.result == Result.PRIMARY_SUCCESS for s in cs)


def test_batch_of_models_refine_sees_replaced_materials():
    """A batch of models whose materials were replaced after construction
    (as Inverse does): the FP64 re-solve of refine_fp64="stagnated" builds its
    instances from the replaced values, part by part in the model's [part]
    tet order, and keeps the original FP64 values where the batch still
    holds their rounding."""
    print("[model] batch of models, refine with replaced materials", flush=True)
    from apple_amd import _lib, batch

    models = [_smas_builder(n, torch.float32, seed=40 + n)[0].finalize() for n in (3, 4)]
    b = batch.Batch.from_models(models, dtype=torch.float32, device="cuda")
    mats = b.potential.materials
    mats["mu"][::3].mul_(1.25)
    mats["activation"].add_(0.01)
    for i, inst in enumerate(b.instances):
        cur = b._current_instance(i)
        t0 = int(b.tet_off[i])
        off = t0
        for p0, p1 in zip(inst.parts, cur.parts):
            n = p0.cells.shape[0]
            sl = slice(off, off + n)
            mu_now = mats["mu"][sl].double().cpu().numpy()
            keep = mu_now == np.asarray(p0.mu, np.float64).astype(np.float32).astype(np.float64)
            assert not keep.all()
            np.testing.assert_array_equal(p1.mu, mu_now if not keep.all() else p0.mu)
            np.testing.assert_array_equal(p1.lam, p0.lam)  # unchanged: the original FP64 values
            if p0.material == _lib.APL_SNH_MUSCLE:
                np.testing.assert_array_equal(p1.activation, mats["activation"][sl].double().cpu().numpy())
            else:
                assert p1.activation is p0.activation
            off += n
        assert off == int(b.tet_off[i + 1])
    cfg = PncgConfig(rtol=1e-7, max_steps=3)
    x, st = b.solve(config=cfg, refine_fp64="stagnated")
    assert b.refined == [0, 1] and torch.isfinite(x).all()
