"""Inverse loop on the GPU: gradient of the whole loop (forward PNCG, adjoint
PCG, mixed pass, autograd through make_materials and the losses) against FP64
central differences of the loss, recovery of a known muscle activation by
Adam, and warm-started forward solves."""

import numpy as np
import pytest
import torch

from apple_amd import fem, inverse
from apple_amd.mesh import Region, cube
from apple_amd.solver import DeviceSolver, PncgConfig, mask_bits

pytestmark = pytest.mark.gpu

PRESTRAIN = 0.9


def _muscle_cube(n, dtype, seed=0):
    m = cube(n)
    rng = np.random.default_rng(seed)
    mu = 1e4 * np.exp(0.2 * rng.standard_normal(m.n_cells))
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = 10 * mu
    m.cell_data["activation"] = np.zeros((m.n_cells, 6))
    region = Region.from_mesh(m)
    pot = fem.StableNeoHookeanMuscle.from_region(region, dtype=dtype, device="cuda")
    fixed = np.zeros((m.n_points, 3), bool)
    fixed[m.points[:, 0] == 0] = True
    surface = np.flatnonzero((m.points == 0).any(1) | (m.points == 1).any(1))
    return m, pot, fixed, surface


def _target(pot, fixed, act, dtype):
    """Forward solution for the true activation (the synthetic observation)."""
    x = torch.zeros(pot.n_points, 3, dtype=dtype, device="cuda")
    st = DeviceSolver(pot).solve(x, None, mask_bits(fixed, "cuda"), PncgConfig(rtol=1e-10, max_steps=50000),
                                 activation=act)[0]
    assert st.success, st
    return x


def test_inverse_gradient_matches_central_differences():
    dtype = torch.float64
    m, pot, fixed, surface = _muscle_cube(3, dtype, seed=1)
    T = m.n_cells
    rng = np.random.default_rng(2)
    true = torch.as_tensor(0.2 + 0.05 * rng.standard_normal(T), dtype=dtype, device="cuda")
    mk = lambda p: {"activation": inverse.muscle_x_activation(p, prestrain=PRESTRAIN)}
    x_true = _target(pot, fixed, mk(true)["activation"].contiguous(), dtype)
    idx = torch.as_tensor(surface, device="cuda")
    losses = [inverse.PointToPointLoss(idx, x_true[idx]), inverse.UniformActivationLoss(weight=1e-3)]
    inv = inverse.Inverse(pot, losses, mk, mask=mask_bits(fixed, "cuda"),
                          forward=PncgConfig(rtol=1e-12, max_steps=50000))
    p0 = torch.as_tensor(0.05 * rng.standard_normal(T), dtype=dtype, device="cuda")
    L, g = inv.value_and_grad(p0)
    g = g.cpu().numpy()
    h = 1e-5
    for t in rng.choice(T, 5, replace=False):
        e = torch.zeros_like(p0)
        e[t] = h
        with torch.no_grad():
            fd = (float(inv.loss_at(p0 + e)) - float(inv.loss_at(p0 - e))) / (2 * h)
        assert g[t] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(g).max())


def test_inverse_recovers_uniform_activation():
    dtype = torch.float64
    m, pot, fixed, surface = _muscle_cube(4, dtype, seed=3)
    T = m.n_cells
    true = 0.25
    mk = lambda p: {"activation": inverse.muscle_x_activation(p.expand(T), prestrain=PRESTRAIN).contiguous()}
    x_true = _target(pot, fixed, mk(torch.tensor([true], dtype=dtype, device="cuda"))["activation"], dtype)
    idx = torch.as_tensor(surface, device="cuda")
    inv = inverse.Inverse(pot, [inverse.PointToPointLoss(idx, x_true[idx])], mk, mask=mask_bits(fixed, "cuda"),
                          lr=0.05, forward=PncgConfig(rtol=1e-10, max_steps=50000))
    p = inv.solve(torch.zeros(1, dtype=dtype, device="cuda"), steps=80)
    hist = inv.history
    print("inverse:", [(h.step, f"{h.loss:.3e}", h.forward_steps) for h in hist[::10]])
    # Adam at a fixed learning rate oscillates about the optimum: bound the
    # best and the final loss, and the recovered parameter
    assert min(h.loss for h in hist) <= 1e-3 * hist[0].loss
    assert hist[-1].loss <= 1e-2 * hist[0].loss
    assert abs(float(p[0]) - true) <= 0.02
    # warm starts: later forward solves start near the answer (the tolerance is
    # relative to each solve's first gradient, so the saving is partial)
    assert np.mean([h.forward_steps for h in hist[-10:]]) < 0.8 * hist[0].forward_steps
