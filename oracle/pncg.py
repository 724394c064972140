"""numpy FP64 restatement of the PNCG loop, plus Newton and adjoint checks.

TEST INFRASTRUCTURE ONLY -- imported by ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py``; never by the product package.

liblaf/apple takes its PNCG from the external ``liblaf-peach`` package, which
is not vendored.  The only in-tree statement of the loop is the benchmark
``benches/bench_pncg_branching_backends.py``; :func:`pncg` restates it:

* Jacobi preconditioner P = 1/|diag H| with the mean of the positive entries
  substituted for zeros (:407-410);
* Dai-Kou beta (:663-679), reset to 0 on the first step, when non-finite or
  when |beta| > 10 (:288-289);
* steepest-descent fallback when g.p is not negative (:292-303);
* alpha = -g.p / p^T H p with nan/inf guards, times ``overstep`` (:606-611);
* Armijo backtracking, c = 1e-4, halving, at most ``max_backtracking``
  trials (:413-456);
* optional (``wolfe_delta``): also accept a trial that meets the
  approximate Wolfe conditions of Hager & Zhang (SIAM J. Optim. 16, 2005):
  f(x+ap) <= f(x) + eps|f| and g(x+ap).p <= (2 delta - 1) g.p.  Energy-only
  Armijo tests stall once the predicted decrease falls below the rounding
  noise of f (around ||g||/||g_0|| ~ 1e-7 in FP64 on a gravity-loaded cube);
  the gradient test does not;
* stop when ||g|| / ||g_0|| <= rtol (the ``relative-grad-norm`` of
  exp/2026/05/06/toy/src/21-smas-prestrain-stable-neo-hookean-muscle.py:382)
  or after ``max_steps`` (forward/_forward.py:26, default 1500).

Pinned by its end state only (trajectories of peach are not observable):
the returned point has ||g|| <= rtol ||g_0|| with a gradient that is itself
pinned by finite differences, on quadratics it equals the direct solve, and
on the reference's patch test it lands on (0.05, -0.025, 0.0375).
"""

from __future__ import annotations

import dataclasses
import enum

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class Result(enum.Enum):
    PRIMARY_SUCCESS = 0
    MAX_STEPS_REACHED = 1
    NON_FINITE = 2


@dataclasses.dataclass
class Solution:
    x: np.ndarray
    result: Result
    n_steps: int
    grad_norm: float
    grad_norm_first: float
    f: float

    @property
    def success(self) -> bool:
        return self.result == Result.PRIMARY_SUCCESS


def jacobi(hess_diag: np.ndarray) -> np.ndarray:
    """P = 1/|h| with zeros replaced by the mean of the positive entries (:407-410)."""
    h = np.abs(hess_diag)
    pos = h > 0.0
    mean = h[pos].mean() if pos.any() else 1.0
    h = np.where(pos, h, mean)
    return 1.0 / h


def dai_kou_beta(g, g_prev, p_prev, P):
    """Dai-Kou beta with the guard of :663-679 (inf when |y.p| <= 1e-12)."""
    y = g - g_prev
    yp = float(np.dot(y, p_prev))
    if abs(yp) <= 1e-12:
        return np.inf
    Py = P * y
    return float(np.dot(g, Py) / yp - (np.dot(y, Py) / yp) * (np.dot(p_prev, g) / yp))


def initial_alpha(g, p, pHp, overstep):
    """alpha = -g.p/pHp with nan->0, -inf->0, +inf->1, non-positive -> 1 (:606-611)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        a = -float(np.dot(g, p)) / pHp if pHp != 0 else (np.inf if -np.dot(g, p) > 0 else 0.0)
    if np.isnan(a):
        a = 0.0
    elif a == -np.inf:
        a = 0.0
    elif a == np.inf:
        a = 1.0
    if not (a > 0.0 and np.isfinite(a)):
        a = 1.0
    return a * overstep


def pncg(
    fun,
    grad,
    hess_diag,
    hess_quad,
    x0,
    *,
    rtol=1e-6,
    atol=0.0,
    max_steps=1500,
    overstep=1.0,
    max_backtracking=8,
    c1=1e-4,
    wolfe_delta=None,
    wolfe_eps=1e-12,
):
    """Minimise ``fun`` from ``x0`` (free-DOF vectors) -- see module docstring."""
    x = np.array(x0, np.float64)
    f = fun(x)
    g = grad(x)
    gn0 = float(np.linalg.norm(g))
    g_prev = g.copy()
    p_prev = np.zeros_like(x)
    for i in range(max_steps):
        gn = float(np.linalg.norm(g))
        if not np.isfinite(gn) or not np.isfinite(f):
            return Solution(x, Result.NON_FINITE, i, gn, gn0, f)
        if gn <= rtol * gn0 or gn <= atol:
            return Solution(x, Result.PRIMARY_SUCCESS, i, gn, gn0, f)
        P = jacobi(hess_diag(x))
        beta = 0.0 if i == 0 else dai_kou_beta(g, g_prev, p_prev, P)
        if not (np.isfinite(beta) and abs(beta) <= 10.0):
            beta = 0.0
        sd = -P * g
        p = sd + beta * p_prev
        gd = float(np.dot(g, p))
        if not (np.isfinite(gd) and gd < 0.0):
            p = sd
            gd = float(np.dot(g, p))
        pHp = hess_quad(x, p)
        alpha = initial_alpha(g, p, pHp, overstep)
        accepted = False
        gt = None
        for _ in range(max_backtracking + 1):
            xt = x + alpha * p
            ft = fun(xt)
            if np.isfinite(ft) and ft <= f + c1 * alpha * gd:
                accepted = True
                gt = None
                break
            if wolfe_delta is not None and np.isfinite(ft) and ft <= f + wolfe_eps * abs(f):
                gt = grad(xt)
                if float(np.dot(gt, p)) <= (2.0 * wolfe_delta - 1.0) * gd:
                    accepted = True
                    break
            alpha *= 0.5
        if accepted:
            x, f = xt, ft
        g_prev, p_prev = g, p
        g = gt if (accepted and gt is not None) else grad(x)
    gn = float(np.linalg.norm(g))
    res = Result.PRIMARY_SUCCESS if gn <= rtol * gn0 or gn <= atol else Result.MAX_STEPS_REACHED
    return Solution(x, res, max_steps, gn, gn0, f)


# --------------------------------------------------------------------------
# full-model helpers over a DofMap (forward/_problem.py:24-59)
# --------------------------------------------------------------------------


class Problem:
    """Free-DOF view of an oracle model -- forward/_problem.py:18-59."""

    def __init__(self, model, dof_map):
        self.model = model
        self.dof = dof_map

    def fun(self, x):
        return self.model.fun(self.dof.to_full(x))

    def grad(self, x):
        return self.dof.to_free(self.model.grad(self.dof.to_full(x)))

    def hess_diag(self, x):
        return self.dof.to_free(self.model.hess_diag(self.dof.to_full(x)))

    def hess_prod(self, x, p):
        return self.dof.to_free(self.model.hess_prod(self.dof.to_full(x), self.dof.to_full_grad(p)))

    def hess_quad(self, x, p):
        return self.model.hess_quad(self.dof.to_full(x), self.dof.to_full_grad(p))

    def hessian(self, x):
        """Sparse exact Hessian on free DOFs, assembled from 12x12 element blocks."""
        u = self.dof.to_full(x)
        n = self.dof.n_points * 3
        rows, cols, vals = [], [], []
        for pot in self.model.potentials:
            if not hasattr(pot, "element_hessians"):
                continue
            He = pot.element_hessians(u)
            dof = (pot.cells[:, :, None] * 3 + np.arange(3)).reshape(-1, 12)
            rows.append(np.repeat(dof, 12, axis=1).ravel())
            cols.append(np.tile(dof, (1, 12)).ravel())
            vals.append(He.ravel())
        H = sp.csr_matrix(
            (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)
        )
        fi = self.dof.free_indices
        return H[fi][:, fi]

    def solve(self, x0, **kw):
        return pncg(self.fun, self.grad, self.hess_diag, self.hess_quad, x0, **kw)


def newton(problem: Problem, x0, *, tol=1e-12, max_steps=50):
    """Newton with Armijo backtracking on the exact Hessian (sparse direct solve)."""
    x = np.array(x0, np.float64)
    f = problem.fun(x)
    for _ in range(max_steps):
        g = problem.grad(x)
        if np.linalg.norm(g) <= tol:
            break
        H = problem.hessian(x)
        d = -spla.spsolve(H.tocsc(), g)
        gd = float(np.dot(g, d))
        if not gd < 0:
            d = -g
            gd = float(np.dot(g, d))
        a = 1.0
        for _ in range(40):
            ft = problem.fun(x + a * d)
            if ft <= f + 1e-4 * a * gd:
                break
            a *= 0.5
        x = x + a * d
        f = problem.fun(x)
    return x


def adjoint_grad(problem: Problem, fem, x_eq, dLdx_free):
    """dL/dmu, dL/dlam (per tet) by the adjoint: H p = -dL/dx, then p^T dg/dq.

    Convention of exp/2025/08/27/inverse/src/30-inverse.py:85-97 and
    exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:472-510.
    """
    H = problem.hessian(x_eq)
    p = spla.spsolve(H.tocsc(), -np.asarray(dLdx_free))
    u = problem.dof.to_full(x_eq)
    pf = problem.dof.to_full_grad(p)
    return fem.mixed_cells(u, pf)


def adjoint_grad_activation(problem: Problem, fem, x_eq, dLdx_free):
    """dL/dmu, dL/dlam and dL/da (T, 6) of the muscle law by the same adjoint
    (exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:472-486)."""
    H = problem.hessian(x_eq)
    p = spla.spsolve(H.tocsc(), -np.asarray(dLdx_free))
    u = problem.dof.to_full(x_eq)
    pf = problem.dof.to_full_grad(p)
    dmu, dlam = fem.mixed_cells(u, pf)
    return dmu, dlam, fem.mixed_activation_cells(u, pf)
