"""Pins the PNCG, Newton and adjoint oracles (oracle/pncg.py) by end states.

* The reference's known-answer patch test (tests/forward/test_static_simulation.py:58-93):
  a 5-vertex embedded tet, ARAP mu = 1, corners fixed, vertex 3 displaced by
  (0.2, -0.1, 0.15); the interior vertex must land on (0.05, -0.025, 0.0375)
  within 1e-8, with n_free == 3, success and a lower final energy.
* On a quadratic the minimiser equals the direct solve.
* The adjoint gradient equals FP64 central differences of the loss through
  Newton-solved equilibria.
"""

import numpy as np
import pytest

import oracle.fem as of
import oracle.pncg as op
from apple_amd.mesh import cube
from helpers import patch_test

PATCH_POINTS, PATCH_CELLS, PATCH_MASK, PATCH_VALUES, PATCH = patch_test()


def _patch_problem(material):
    dhdX, dV = of.region(PATCH_POINTS, PATCH_CELLS)
    assert np.all(dV > 0)
    n = PATCH_CELLS.shape[0]
    mu = np.full(n, PATCH["mu"])
    lam = np.ones(n) if material == of.SNH else None
    pot = of.FemPotential(material, PATCH_CELLS, dhdX, dV, mu, lam)
    dof = of.DofMap(PATCH_MASK, PATCH_VALUES)
    return op.Problem(of.Model([pot]), dof)


@pytest.mark.parametrize("material", [of.ARAP, of.SNH])
def test_patch_test_known_answer(material):
    prob = _patch_problem(material)
    exp = PATCH["expected"]
    assert prob.dof.n_free == exp["n_free"]
    x0 = np.zeros(exp["n_free"])
    f0 = prob.fun(x0)
    sol = prob.solve(x0, rtol=1e-9, max_steps=1500)
    assert sol.result.name == exp["result"]
    u = prob.dof.to_full(sol.x)
    np.testing.assert_allclose(u[PATCH_MASK[:, 0]], PATCH_VALUES[PATCH_MASK[:, 0]])
    np.testing.assert_allclose(u[exp["interior_vertex"]], exp["interior_displacement"], atol=exp["atol"])
    assert np.isfinite(f0) and sol.f < f0


def test_pncg_quadratic_equals_direct_solve():
    rng = np.random.default_rng(0)
    n = 40
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    b = rng.standard_normal(n)
    sol = op.pncg(
        lambda x: 0.5 * x @ A @ x - b @ x,
        lambda x: A @ x - b,
        lambda x: np.diag(A).copy(),
        lambda x, p: float(p @ A @ p),
        np.zeros(n),
        rtol=1e-8,
        max_steps=2000,
    )
    assert sol.success
    np.testing.assert_allclose(sol.x, np.linalg.solve(A, b), rtol=1e-6, atol=1e-9)


def test_jacobi_fallback_for_zero_diagonal():
    P = op.jacobi(np.array([2.0, 0.0, -4.0]))
    np.testing.assert_allclose(P, [0.5, 1.0 / 3.0, 0.25])


def _gravity_problem(n=2, material=of.SNH, seed=0):
    m = cube(n)
    dhdX, dV = of.region(m.points, m.cells)
    rng = np.random.default_rng(seed)
    mu = 1e4 * np.exp(0.5 * rng.standard_normal(m.n_cells))
    lam = 10.0 * mu if material == of.SNH else None
    pot = of.FemPotential(material, m.cells, dhdX, dV, mu, lam)
    lumped = of.lumped_vertex_volume(m.points, m.cells)
    f = np.zeros((m.n_points, 3))
    f[:, 2] = -9.8 * 1e3 * lumped
    ext = of.ExternalForce(np.arange(m.n_points), f)
    mask = np.zeros((m.n_points, 3), bool)
    mask[m.points[:, 2] == 0.0] = True
    dof = of.DofMap(mask, np.zeros((m.n_points, 3)))
    return m, pot, op.Problem(of.Model([pot, ext]), dof)


def test_pncg_reaches_newton_equilibrium():
    _, pot, prob = _gravity_problem(3)
    x0 = np.zeros(prob.dof.n_free)
    # Armijo alone (the benchmark's loop) reaches ~1e-7 and then stalls on f's rounding noise
    sol = prob.solve(x0, rtol=1e-6, max_steps=5000)
    assert sol.success
    xn = op.newton(prob, sol.x, tol=1e-9)
    assert np.linalg.norm(prob.grad(xn)) <= 1e-9
    np.testing.assert_allclose(sol.x, xn, atol=1e-5 * np.sqrt(3))
    stalled = prob.solve(x0, rtol=1e-11, max_steps=600)
    assert not stalled.success
    # the approximate Wolfe acceptance keeps converging below that floor
    sol2 = prob.solve(x0, rtol=1e-11, max_steps=5000, wolfe_delta=0.1)
    assert sol2.success
    np.testing.assert_allclose(sol2.x, xn, atol=1e-9 * np.sqrt(3))


def test_adjoint_matches_central_differences():
    m, pot, prob = _gravity_problem(2, seed=3)
    x_eq = op.newton(prob, np.zeros(prob.dof.n_free), tol=1e-12)
    rng = np.random.default_rng(4)
    x_t = x_eq + 1e-3 * rng.standard_normal(x_eq.shape)

    def loss(x):
        return float(np.sum((x - x_t) ** 2))

    dmu, dlam = op.adjoint_grad(prob, pot, x_eq, 2.0 * (x_eq - x_t))
    for t in (0, 11, 25, 39):
        for arr, g in ((pot.mu, dmu), (pot.lam, dlam)):
            h = 1e-4 * arr[t]
            v0 = arr[t]
            arr[t] = v0 + h
            lp = loss(op.newton(prob, x_eq, tol=1e-13))
            arr[t] = v0 - h
            lm = loss(op.newton(prob, x_eq, tol=1e-13))
            arr[t] = v0
            fd = (lp - lm) / (2 * h)
            assert g[t] == pytest.approx(fd, rel=2e-5, abs=1e-14)
