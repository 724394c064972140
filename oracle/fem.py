"""numpy FP64 restatement of liblaf/apple's per-tet FEM algebra (oracle).

TEST INFRASTRUCTURE ONLY -- imported by ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py``; never by the product package.

All arrays are vectorised over tets: ``F`` is (T,3,3), ``dhdX`` is (T,4,3),
cell displacement blocks are (T,4,3).  Citations are into liblaf/apple
(``src/liblaf/apple/...``).

Pinning (see tests/test_oracle_fem.py): affine patch F = I + A, volume sums,
Psi(I) = 0, uniform-scale closed forms, FP64 central finite differences of the
energy (gradient) and of the gradient (Hessian-vector product, exact element
Hessian), ARAP closed forms Psi(sR) = 1.5 mu (s-1)^2, the reference's patch
test, and convention invariance of the SVD.  The only unpinned part is the
exact output of Warp's ``svd3`` (signs/ordering), which the ARAP formulas are
invariant to except through the rotation-variant sign convention restated in
:func:`svd_rv`.
"""

from __future__ import annotations

import numpy as np

# --------------------------------------------------------------------------
# setup: jax/fem/element/_tetra.py:36-45, quadrature/_tetra.py:14-15,
#        region/_region.py:84-108, warp/fem/_base.py:94-111
# --------------------------------------------------------------------------

#: reference element shape-function gradients (element/_tetra.py:36-45)
DHDR = np.array(
    [[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
)
#: one-point rule weight (quadrature/_tetra.py:14-15)
QUAD_WEIGHT = 1.0 / 6.0


def region(points: np.ndarray, cells: np.ndarray, fraction=None):
    """dhdX (T,4,3) and dV (T,) -- region/_region.py:84-108, _base.py:104.

    dXdr = X_cell^T dhdr, drdX = inv(dXdr), dV = det(dXdr) * w,
    dhdX = dhdr @ drdX; dV is multiplied by the per-cell Fraction.
    """
    X = np.asarray(points, np.float64)[np.asarray(cells)]  # (T,4,3)
    dXdr = np.einsum("tai,aj->tij", X, DHDR)
    drdX = np.linalg.inv(dXdr)
    dV = np.linalg.det(dXdr) * QUAD_WEIGHT
    dhdX = np.einsum("ai,tij->taj", DHDR, drdX)
    if fraction is not None:
        dV = dV * np.asarray(fraction, np.float64)
    return dhdX, dV


# --------------------------------------------------------------------------
# func/_deformation.py, func/_identity.py, func/_gradient.py, func/_misc.py
# --------------------------------------------------------------------------


def cell_values(u: np.ndarray, cells: np.ndarray) -> np.ndarray:
    """get_cell_displacements -- func/_misc.py:22-28 (gather 4 rows)."""
    return np.asarray(u, np.float64)[np.asarray(cells)]


def gradient(u_cell, dhdX):
    """du/dX = u_cell^T dhdX -- func/_deformation.py:9-12."""
    return np.einsum("tai,taj->tij", u_cell, dhdX)


def deformation_gradient(u_cell, dhdX):
    """F = du/dX + I -- func/_deformation.py:15-19."""
    return gradient(u_cell, dhdX) + np.eye(3)


def jvp(dhdX, p_cell):
    """dF = p^T dhdX -- func/_deformation.py:22-25."""
    return np.einsum("tai,taj->tij", p_cell, dhdX)


def vjp(dhdX, P):
    """dhdX P^T (4x3 per tet) -- func/_deformation.py:28-31."""
    return np.einsum("taj,tij->tai", dhdX, P)


def I2(F):
    """||F||_F^2 -- func/_identity.py:20-28."""
    return np.einsum("tij,tij->t", F, F)


def I3(F):
    """det F -- func/_identity.py:31-39."""
    return np.linalg.det(F)


def g3(F):
    """cof F = [f1 x f2, f2 x f0, f0 x f1] as columns -- func/_gradient.py:30-40."""
    f0, f1, f2 = F[:, :, 0], F[:, :, 1], F[:, :, 2]
    return np.stack([np.cross(f1, f2), np.cross(f2, f0), np.cross(f0, f1)], axis=2)


def ddot(A, B):
    """Frobenius product per tet (wp.ddot)."""
    return (A * B).reshape(A.shape[0], -1).sum(axis=1)


def h6_matrix(F, dF):
    """Columns [f1 x p2 - f2 x p1, f2 x p0 - f0 x p2, f0 x p1 - f1 x p0].

    The inner matrix of h6_prod -- func/_hess_prod.py:61-77.
    """
    f0, f1, f2 = F[:, :, 0], F[:, :, 1], F[:, :, 2]
    p0, p1, p2 = dF[:, :, 0], dF[:, :, 1], dF[:, :, 2]
    return np.stack(
        [
            np.cross(f1, p2) - np.cross(f2, p1),
            np.cross(f2, p0) - np.cross(f0, p2),
            np.cross(f0, p1) - np.cross(f1, p0),
        ],
        axis=2,
    )


def h6_quad(F, dF):
    """p0.(f1xp2 - f2xp1) + p1.(f2xp0 - f0xp2) + p2.(f0xp1 - f1xp0) -- func/_hess_quad.py:59-74."""
    return ddot(dF, h6_matrix(F, dF))


# --------------------------------------------------------------------------
# stable Neo-Hookean: warp/fem/_stable_neo_hookean.py:17-103
# --------------------------------------------------------------------------


def snh_energy_density(F, mu, lam):
    """Psi = mu/2 (I2 - 3) - mu (J - 1) + lam/2 (J - 1)^2 -- :17-27."""
    J = I3(F)
    return 0.5 * mu * (I2(F) - 3.0) - mu * (J - 1.0) + 0.5 * lam * (J - 1.0) ** 2


def snh_pk1(F, mu, lam):
    """P = (mu/2) g2 + (-mu + lam (J-1)) g3, g2 = 2F -- :30-39, func/_gradient.py:19-27."""
    J = I3(F)
    s = -mu + lam * (J - 1.0)
    return (0.5 * mu)[:, None, None] * 2.0 * F + s[:, None, None] * g3(F)


def snh_hess_diag(F, dhdX, mu, lam):
    """lam h3_diag + (mu/2) h5_diag + s h6_diag (h6_diag == 0) -- :42-67.

    h3_diag = (dhdX g3^T)^2 component-wise (func/_hess_diag.py:29-33);
    h5_diag = 2 |dhdX_a|^2 on every component (:53-66); h6_diag = 0 (:69-72).
    """
    W = vjp(dhdX, g3(F))
    h3 = W * W
    t = np.einsum("taj,taj->ta", dhdX, dhdX)
    h5 = 2.0 * np.repeat(t[:, :, None], 3, axis=2)
    return lam[:, None, None] * h3 + (0.5 * mu)[:, None, None] * h5


def snh_hess_prod(F, p_cell, dhdX, mu, lam):
    """lam h3_prod + (mu/2) h5_prod + s h6_prod -- :70-84, func/_hess_prod.py:29-77."""
    J = I3(F)
    s = -mu + lam * (J - 1.0)
    W = vjp(dhdX, g3(F))
    h3 = ddot(W, p_cell)[:, None, None] * W
    dF = jvp(dhdX, p_cell)
    h5 = 2.0 * vjp(dhdX, dF)
    h6 = vjp(dhdX, h6_matrix(F, dF))
    return (
        lam[:, None, None] * h3 + (0.5 * mu)[:, None, None] * h5 + s[:, None, None] * h6
    )


def snh_hess_quad(F, p_cell, dhdX, mu, lam):
    """lam h3_quad + (mu/2) h5_quad + s h6_quad -- :87-103, func/_hess_quad.py:30-74."""
    J = I3(F)
    s = -mu + lam * (J - 1.0)
    dF = jvp(dhdX, p_cell)
    h3 = ddot(dF, g3(F)) ** 2
    h5 = 2.0 * ddot(dF, dF)
    return lam * h3 + 0.5 * mu * h5 + s * h6_quad(F, dF)


def snh_dpk1_dparams(F):
    """dP/dmu = F - cof F and dP/dlam = (J - 1) cof F (derivative of :30-39)."""
    G3 = g3(F)
    J = I3(F)
    return F - G3, (J - 1.0)[:, None, None] * G3


# --------------------------------------------------------------------------
# stable Neo-Hookean muscle: warp/fem/_stable_neo_hookean_muscle.py:18-113
# --------------------------------------------------------------------------

#: index pairs of the activation vector (warp/fem/func/_misc.py:48-53)
ACTIVATION_PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def activation_matrix(a):
    """A = I + [[a0, a3, a4], [a3, a1, a5], [a4, a5, a2]] -- func/_misc.py:48-53."""
    a = np.asarray(a, np.float64).reshape(-1, 6)
    A = np.repeat(np.eye(3)[None], a.shape[0], axis=0)
    for k, (i, j) in enumerate(ACTIVATION_PAIRS):
        A[:, i, j] += a[:, k]
        if i != j:
            A[:, j, i] += a[:, k]
    return A


def muscle_energy_density(F, mu, lam, a):
    """SNH energy of G = F A -- :18-30."""
    return snh_energy_density(F @ activation_matrix(a), mu, lam)


def muscle_pk1(F, mu, lam, a):
    """(dPsi/dI2 g2(G) + dPsi/dI3 g3(G)) A^T -- :33-43."""
    A = activation_matrix(a)
    return snh_pk1(F @ A, mu, lam) @ np.swapaxes(A, 1, 2)


def muscle_hess_diag(F, dhdX, mu, lam, a):
    """The SNH diagonal terms with G = F A and dhdX A -- :46-75."""
    A = activation_matrix(a)
    return snh_hess_diag(F @ A, dhdX @ A, mu, lam)


def muscle_hess_prod(F, p_cell, dhdX, mu, lam, a):
    """lam h3_prod + (mu/2) h5_prod + s h6_prod on (G, dhdX A) -- :78-92."""
    A = activation_matrix(a)
    return snh_hess_prod(F @ A, p_cell, dhdX @ A, mu, lam)


def muscle_hess_quad(F, p_cell, dhdX, mu, lam, a):
    """lam h3_quad + (mu/2) h5_quad + s h6_quad on (G, dhdX A) -- :95-113."""
    A = activation_matrix(a)
    return snh_hess_quad(F @ A, p_cell, dhdX @ A, mu, lam)


def snh_hess_bilinear(F, X, Y, mu, lam):
    """X : d^2 Psi/dF^2 : Y of the SNH law (per tet), from its three terms.

    lam (g3 : X)(g3 : Y) + mu X : Y + s Y : h6(F, X), the bilinear form whose
    diagonal h6_quad and whose matrix form h6_prod are (func/_hess_prod.py,
    _hess_quad.py).
    """
    J = I3(F)
    s = -mu + lam * (J - 1.0)
    G3 = g3(F)
    return lam * ddot(G3, X) * ddot(G3, Y) + mu * ddot(X, Y) + s * ddot(Y, h6_matrix(F, X))


def muscle_mixed(F, dF, mu, lam, a):
    """p^T d(grad)/dq per tet (without dV) for q = mu, lambda, a_0..a_5.

    grad . p = P_G(F A) : (dF A); its derivatives are
    d/dmu = (dP_G/dmu) : dF A, d/dlam likewise, and for a_k with
    E_k = dA/da_k:  H_G[F E_k] : (dF A) + P_G : (dF E_k).
    """
    A = activation_matrix(a)
    G = F @ A
    dG = dF @ A
    dmu, dlam = snh_dpk1_dparams(G)
    P = snh_pk1(G, mu, lam)
    dact = np.zeros((F.shape[0], 6))
    for k, (i, j) in enumerate(ACTIVATION_PAIRS):
        E = np.zeros((3, 3))
        E[i, j] = E[j, i] = 1.0
        dact[:, k] = snh_hess_bilinear(G, F @ E, dG, mu, lam) + ddot(P, dF @ E)
    return ddot(dmu, dG), ddot(dlam, dG), dact


# --------------------------------------------------------------------------
# ARAP: warp/fem/_arap.py:17-76, warp/math/_rotation.py, func/_misc.py:31-70
# --------------------------------------------------------------------------


def svd_rv(F):
    """Rotation-variant SVD: F = U diag(sigma) V^T with U, V in SO(3).

    Restates the convention of ``wp.svd3`` as used by ``svd_rv``
    (math/_rotation.py:9-13): singular values sorted by decreasing magnitude,
    the sign of det F carried by the last one.  Built from numpy's SVD with
    the last columns flipped where det U or det V is negative.
    """
    U, s, Vt = np.linalg.svd(F)
    V = np.swapaxes(Vt, 1, 2).copy()
    s = s.copy()
    U = U.copy()
    dU = np.linalg.det(U) < 0
    U[dU, :, 2] *= -1.0
    s[dU, 2] *= -1.0
    dVn = np.linalg.det(V) < 0
    V[dVn, :, 2] *= -1.0
    s[dVn, 2] *= -1.0
    return U, s, V


def polar_rv(F):
    """R = U V^T, S = V diag(sigma) V^T -- math/_rotation.py:16-22."""
    U, s, V = svd_rv(F)
    R = U @ np.swapaxes(V, 1, 2)
    S = V @ (s[:, :, None] * np.swapaxes(V, 1, 2))
    return R, S


def arap_lambdas(sigma, clamp=True):
    """2 / max(s_i + s_j, 2) over pairs (0,1), (1,2), (2,0) -- func/_misc.py:31-43."""
    pairs = np.stack(
        [sigma[:, 0] + sigma[:, 1], sigma[:, 1] + sigma[:, 2], sigma[:, 2] + sigma[:, 0]],
        axis=1,
    )
    if clamp:
        return 2.0 / np.maximum(pairs, 2.0)
    return 2.0 / pairs


def arap_Qs(U, V):
    """Twist modes Q0, Q1, Q2 -- func/_misc.py:56-70."""
    r2 = np.sqrt(2.0)
    U0, U1, U2 = U[:, :, 0], U[:, :, 1], U[:, :, 2]
    V0, V1, V2 = V[:, :, 0], V[:, :, 1], V[:, :, 2]

    def outer(a, b):
        return np.einsum("ti,tj->tij", a, b)

    Q0 = (outer(U1, V0) - outer(U0, V1)) / r2
    Q1 = (outer(U1, V2) - outer(U2, V1)) / r2
    Q2 = (outer(U0, V2) - outer(U2, V0)) / r2
    return Q0, Q1, Q2


def arap_energy_density(F, mu):
    """Psi = mu/2 ||F - R||^2 -- _arap.py:17-21."""
    R, _ = polar_rv(F)
    D = F - R
    return 0.5 * mu * ddot(D, D)


def arap_pk1(F, mu):
    """P = mu (F - R) -- _arap.py:24-28."""
    R, _ = polar_rv(F)
    return mu[:, None, None] * (F - R)


def arap_hess_diag(F, dhdX, mu, clamp=True):
    """(mu/2)(-2 h4_diag + h5_diag) -- _arap.py:31-40, func/_hess_diag.py:36-66."""
    U, s, V = svd_rv(F)
    lam = arap_lambdas(s, clamp)
    h4 = np.zeros(dhdX.shape)
    for k, Q in enumerate(arap_Qs(U, V)):
        W = vjp(dhdX, Q)
        h4 += lam[:, k, None, None] * W * W
    t = np.einsum("taj,taj->ta", dhdX, dhdX)
    h5 = 2.0 * np.repeat(t[:, :, None], 3, axis=2)
    return (0.5 * mu)[:, None, None] * (-2.0 * h4 + h5)


def arap_hess_prod(F, p_cell, dhdX, mu, clamp=True):
    """(mu/2)(-2 h4_prod(p, dhdX) + h5_prod(p, dhdX)): the H v intended by _arap.py:43-58.

    Uses the argument order of func/_hess_prod.py:38-39,55 (p first).  The
    reference body passes (dhdX, p); that variant is
    :func:`arap_hess_prod_as_written`.
    """
    U, s, V = svd_rv(F)
    lam = arap_lambdas(s, clamp)
    h4 = np.zeros(dhdX.shape)
    for k, Q in enumerate(arap_Qs(U, V)):
        W = vjp(dhdX, Q)
        h4 += (lam[:, k] * ddot(W, p_cell))[:, None, None] * W
    dF = jvp(dhdX, p_cell)
    h5 = 2.0 * vjp(dhdX, dF)
    return (0.5 * mu)[:, None, None] * (-2.0 * h4 + h5)


def arap_hess_prod_as_written(F, p_cell, dhdX, mu, clamp=True):
    """Literal restatement of _arap.py:55-56: h4_prod(dhdX, p, ...), h5_prod(dhdX, p).

    With the callee signatures (p, dhdX, ...) this evaluates the formulas with
    the two 4x3 operands exchanged.  Kept only to document that the
    reference's Arap.hess_prod is not H v; the product implements
    :func:`arap_hess_prod`.
    """
    return arap_hess_prod(F, dhdX, p_cell, mu, clamp)


def arap_hess_quad(F, p_cell, dhdX, mu, clamp=True):
    """(mu/2)(-2 h4_quad + h5_quad) -- _arap.py:61-76, func/_hess_quad.py:37-56."""
    U, s, V = svd_rv(F)
    lam = arap_lambdas(s, clamp)
    dF = jvp(dhdX, p_cell)
    h4 = np.zeros(F.shape[0])
    for k, Q in enumerate(arap_Qs(U, V)):
        h4 += lam[:, k] * ddot(Q, dF) ** 2
    h5 = 2.0 * ddot(dF, dF)
    return 0.5 * mu * (-2.0 * h4 + h5)


def arap_dpk1_dmu(F):
    """dP/dmu = F - R (derivative of _arap.py:24-28)."""
    R, _ = polar_rv(F)
    return F - R


# --------------------------------------------------------------------------
# kernel templates (warp/fem/_base.py:243-383) and the model sum
# --------------------------------------------------------------------------

SNH = "snh"
ARAP = "arap"
MUSCLE = "snh_muscle"


class FemPotential:
    """One FEM energy over a tet mesh with the reference's five entry points.

    ``fun``/``grad``/``hess_diag``/``hess_prod``/``hess_quad`` follow the
    kernel templates of warp/fem/_base.py:243-383: evaluate per (tet, q=0),
    multiply by dV, clamp the Hessian diagonal component-wise (:314-320) and
    the quadratic form per tet (:379-380), scatter-add to vertices.
    """

    def __init__(self, material, cells, dhdX, dV, mu, lam=None, activation=None):
        self.material = material
        self.cells = np.asarray(cells)
        self.dhdX = np.asarray(dhdX, np.float64)
        self.dV = np.asarray(dV, np.float64)
        self.mu = np.asarray(mu, np.float64)
        self.lam = None if lam is None else np.asarray(lam, np.float64)
        self.activation = None if activation is None else np.asarray(activation, np.float64).reshape(-1, 6)
        if material == MUSCLE and (self.lam is None or self.activation is None):
            raise ValueError("the muscle law needs lam and activation")

    @property
    def n_cells(self):
        return self.cells.shape[0]

    def F(self, u):
        return deformation_gradient(cell_values(u, self.cells), self.dhdX)

    def energy_density(self, u):
        F = self.F(u)
        if self.material == SNH:
            return snh_energy_density(F, self.mu, self.lam)
        if self.material == MUSCLE:
            return muscle_energy_density(F, self.mu, self.lam, self.activation)
        return arap_energy_density(F, self.mu)

    def pk1(self, u):
        F = self.F(u)
        if self.material == SNH:
            return snh_pk1(F, self.mu, self.lam)
        if self.material == MUSCLE:
            return muscle_pk1(F, self.mu, self.lam, self.activation)
        return arap_pk1(F, self.mu)

    def _scatter(self, n_points, cell_vals):
        out = np.zeros((n_points, 3))
        np.add.at(out, self.cells.reshape(-1), cell_vals.reshape(-1, 3))
        return out

    def fun(self, u):
        """Sum Psi dV -- _base.py:243-263."""
        return float(np.sum(self.energy_density(u) * self.dV))

    def grad_cells(self, u):
        return vjp(self.dhdX, self.pk1(u)) * self.dV[:, None, None]

    def grad(self, u):
        """Scatter dV dhdX P^T -- _base.py:265-291."""
        return self._scatter(u.shape[0], self.grad_cells(u))

    def hess_diag_cells(self, u):
        F = self.F(u)
        if self.material == SNH:
            h = snh_hess_diag(F, self.dhdX, self.mu, self.lam)
        elif self.material == MUSCLE:
            h = muscle_hess_diag(F, self.dhdX, self.mu, self.lam, self.activation)
        else:
            h = arap_hess_diag(F, self.dhdX, self.mu)
        return np.maximum(h * self.dV[:, None, None], 0.0)

    def hess_diag(self, u):
        """Scatter max(dV diag H_e, 0) -- _base.py:293-324."""
        return self._scatter(u.shape[0], self.hess_diag_cells(u))

    def hess_prod_cells(self, u, p):
        F = self.F(u)
        pc = cell_values(p, self.cells)
        if self.material == SNH:
            h = snh_hess_prod(F, pc, self.dhdX, self.mu, self.lam)
        elif self.material == MUSCLE:
            h = muscle_hess_prod(F, pc, self.dhdX, self.mu, self.lam, self.activation)
        else:
            h = arap_hess_prod(F, pc, self.dhdX, self.mu)
        return h * self.dV[:, None, None]

    def hess_prod(self, u, p):
        """Scatter dV H_e p_e, no clamp -- _base.py:326-351."""
        return self._scatter(u.shape[0], self.hess_prod_cells(u, p))

    def hess_quad_cells(self, u, p):
        F = self.F(u)
        pc = cell_values(p, self.cells)
        if self.material == SNH:
            q = snh_hess_quad(F, pc, self.dhdX, self.mu, self.lam)
        elif self.material == MUSCLE:
            q = muscle_hess_quad(F, pc, self.dhdX, self.mu, self.lam, self.activation)
        else:
            q = arap_hess_quad(F, pc, self.dhdX, self.mu)
        return np.maximum(q * self.dV, 0.0)

    def hess_quad(self, u, p):
        """Sum max(dV p_e^T H_e p_e, 0) -- _base.py:353-383."""
        return float(np.sum(self.hess_quad_cells(u, p)))

    def mixed_cells(self, u, p):
        """Per-tet p^T d(grad)/d(mu) and p^T d(grad)/d(lambda).

        The mixed derivative the adjoint needs (exp/2025/08/27/inverse/src/
        30-inverse.py:96-97, ``mixed_derivative_prod``): dV (dP/dq) : dF(p).
        """
        F = self.F(u)
        dF = jvp(self.dhdX, cell_values(p, self.cells))
        if self.material == SNH:
            dmu, dlam = snh_dpk1_dparams(F)
            return self.dV * ddot(dmu, dF), self.dV * ddot(dlam, dF)
        if self.material == MUSCLE:
            dmu, dlam, _ = muscle_mixed(F, dF, self.mu, self.lam, self.activation)
            return self.dV * dmu, self.dV * dlam
        return self.dV * ddot(arap_dpk1_dmu(F), dF), np.zeros(self.n_cells)

    def mixed_activation_cells(self, u, p):
        """Per-tet p^T d(grad)/d(a_k), (T, 6): the muscle activation adjoint's
        mixed term (exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-
        hookean.py:472-486)."""
        F = self.F(u)
        dF = jvp(self.dhdX, cell_values(p, self.cells))
        _, _, dact = muscle_mixed(F, dF, self.mu, self.lam, self.activation)
        return self.dV[:, None] * dact

    def element_hessians(self, u):
        """Exact 12x12 element Hessians (T,12,12) from hess_prod on unit vectors."""
        F = self.F(u)
        T = self.n_cells
        H = np.zeros((T, 12, 12))
        for col in range(12):
            e = np.zeros((T, 4, 3))
            e[:, col // 3, col % 3] = 1.0
            if self.material == SNH:
                h = snh_hess_prod(F, e, self.dhdX, self.mu, self.lam)
            elif self.material == MUSCLE:
                h = muscle_hess_prod(F, e, self.dhdX, self.mu, self.lam, self.activation)
            else:
                h = arap_hess_prod(F, e, self.dhdX, self.mu)
            H[:, :, col] = (h * self.dV[:, None, None]).reshape(T, 12)
        return H


class ExternalForce:
    """E = -sum f . u[idx], g[idx] -= f, no Hessian -- warp/potential/_ext_force.py:17-39,63-90."""

    def __init__(self, indices, force):
        self.indices = np.asarray(indices)
        self.force = np.asarray(force, np.float64)

    def fun(self, u):
        return float(-np.sum(self.force * np.asarray(u)[self.indices]))

    def grad(self, u):
        out = np.zeros((u.shape[0], 3))
        np.add.at(out, self.indices, -self.force)
        return out


class Model:
    """Sum over potentials after zeroing -- warp/model/_model.py:13-36."""

    def __init__(self, potentials):
        self.potentials = list(potentials)

    def fun(self, u):
        return float(sum(p.fun(u) for p in self.potentials))

    def grad(self, u):
        return sum(p.grad(u) for p in self.potentials)

    def hess_diag(self, u):
        out = np.zeros((u.shape[0], 3))
        for p in self.potentials:
            if isinstance(p, FemPotential):
                out += p.hess_diag(u)
        return out

    def hess_prod(self, u, v):
        out = np.zeros((u.shape[0], 3))
        for p in self.potentials:
            if isinstance(p, FemPotential):
                out += p.hess_prod(u, v)
        return out

    def hess_quad(self, u, v):
        return float(
            sum(p.hess_quad(u, v) for p in self.potentials if isinstance(p, FemPotential))
        )


class DofMap:
    """free/fixed split of the flattened (V,3) array -- forward/dof_map/_builder.py:52-64, _dof_map.py:30-49."""

    def __init__(self, fixed_mask, fixed_values):
        fixed_mask = np.asarray(fixed_mask, bool)
        self.n_points = fixed_mask.shape[0]
        self.fixed_indices = np.flatnonzero(fixed_mask)
        self.fixed_values = np.asarray(fixed_values, np.float64).reshape(-1)[
            self.fixed_indices
        ]
        self.free_indices = np.flatnonzero(~fixed_mask)

    @property
    def n_free(self):
        return self.free_indices.size

    def to_free(self, full):
        return np.asarray(full).reshape(-1)[self.free_indices]

    def to_full(self, free):
        out = np.empty(self.n_points * 3)
        out[self.fixed_indices] = self.fixed_values
        out[self.free_indices] = free
        return out.reshape(self.n_points, 3)

    def to_full_grad(self, free):
        out = np.empty(self.n_points * 3)
        out[self.fixed_indices] = 0.0
        out[self.free_indices] = free
        return out.reshape(self.n_points, 3)


def lame_converter(E, nu):
    """(lambda, mu) from Young's modulus and Poisson ratio -- common/_moduli.py:5-10."""
    E = np.asarray(E, np.float64)
    nu = np.asarray(nu, np.float64)
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


def lumped_vertex_volume(points, cells):
    """Quarter-volume lumping per vertex -- benches/test_aggregation.py:30-82,89-129."""
    X = np.asarray(points, np.float64)[np.asarray(cells)]
    vol = np.abs(np.linalg.det(np.stack([X[:, 0] - X[:, 3], X[:, 1] - X[:, 3], X[:, 2] - X[:, 3]], axis=2))) / 6.0
    out = np.zeros(len(points))
    np.add.at(out, np.asarray(cells).reshape(-1), np.repeat(vol / 4.0, 4))
    return out
