"""The C restatement (CPU baseline) agrees with the numpy oracle."""

import numpy as np
import pytest

import oracle.fem as of
from oracle import c_oracle
from helpers import snh_case


@pytest.fixture(scope="module")
def case():
    m, region, u, v = snh_case(5, seed=4, noise=0.02)
    pot = of.FemPotential(of.SNH, m.cells, region.dhdX, region.dV, m.cell_data["mu"], m.cell_data["lambda"])
    cp = c_oracle.SnhProblem(m.cells, region.dhdX, region.dV, m.cell_data["mu"], m.cell_data["lambda"])
    return m, pot, cp, u, v


def test_c_passes_match_numpy(case):
    m, pot, cp, u, v = case
    assert cp.fun(u) == pytest.approx(pot.fun(u), rel=1e-12)
    np.testing.assert_allclose(cp.grad(u), pot.grad(u), rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(cp.hess_diag(u), pot.hess_diag(u), rtol=1e-10, atol=1e-9)
    assert cp.hess_quad(u, v) == pytest.approx(pot.hess_quad(u, v), rel=1e-10)
    g = np.zeros_like(u)
    h = np.zeros_like(u)
    assert cp.pass_a(u, g, h) == pytest.approx(pot.fun(u), rel=1e-12)
    np.testing.assert_allclose(g, pot.grad(u), rtol=1e-10, atol=1e-9)


def test_c_pncg_reaches_equilibrium(case):
    m, pot, cp, u, v = case
    f = np.zeros_like(u)
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e3 * pot.dV / 4, 4))
    fixed = np.zeros_like(u, dtype=bool)
    fixed[m.points[:, 2] == 0] = True
    x0 = np.where(fixed, 0.0, 0.01 * np.random.default_rng(0).standard_normal(u.shape))
    x, steps, status, rg = cp.pncg(x0, f, ~fixed, rtol=1e-8, max_steps=5000)
    assert status == 1 and rg <= 1e-8
    g = pot.grad(x) - f
    g[fixed] = 0
    assert np.linalg.norm(g) <= 1e-8 * np.linalg.norm((pot.grad(x0) - f)[~fixed]) * 1.01
