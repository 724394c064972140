"""Pins the numpy oracle (oracle/fem.py) without the reference.

Every check is chosen so that a plausible slip (a dropped term, a wrong sign
or index, a transposed operand) fails: closed forms, FP64 central finite
differences of energies and gradients, the reference's volume invariant
(benches/test_aggregation.py:82,129) and SVD-convention invariance.
"""

import numpy as np
import pytest

import oracle.fem as of
from apple_amd.mesh import cube

RNG = np.random.default_rng(1234)


def _random_F(T, scale=0.3, rng=RNG):
    return np.eye(3) + scale * rng.standard_normal((T, 3, 3))


def _rot(rng, T):
    q = rng.standard_normal((T, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack(
        [
            np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
            np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
            np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1),
        ],
        1,
    )


def _activation(rng, T, scale=0.15):
    return scale * rng.standard_normal((T, 6))


def _fem(material, mesh, rng=RNG):
    dhdX, dV = of.region(mesh.points, mesh.cells)
    T = mesh.n_cells
    mu = 1e4 * np.exp(0.3 * rng.standard_normal(T))
    lam = 10.0 * mu if material != of.ARAP else None
    act = _activation(rng, T) if material == of.MUSCLE else None
    return of.FemPotential(material, mesh.cells, dhdX, dV, mu, lam, act)


# ----------------------------------------------------------------- setup


@pytest.mark.parametrize("order", ["morton", "lex"])
def test_affine_displacement_gives_constant_F(order):
    m = cube(3, order=order)
    A = RNG.standard_normal((3, 3)) * 0.2
    u = m.points @ A.T
    dhdX, _ = of.region(m.points, m.cells)
    F = of.deformation_gradient(of.cell_values(u, m.cells), dhdX)
    np.testing.assert_allclose(F, np.broadcast_to(np.eye(3) + A, F.shape), atol=1e-12)


def test_volume_sums_match_mesh_volume():
    m = cube(4, lo=-0.5, hi=1.5)
    _, dV = of.region(m.points, m.cells)
    assert dV.min() > 0
    np.testing.assert_allclose(dV.sum(), 8.0, rtol=1e-13)
    # the reference's aggregation invariant: sum of lumped quarter volumes = volume
    np.testing.assert_allclose(of.lumped_vertex_volume(m.points, m.cells).sum(), 8.0, rtol=1e-13)


def test_lame_converter_closed_form():
    lam, mu = of.lame_converter(1.0, 0.25)
    assert lam == pytest.approx(0.4) and mu == pytest.approx(0.4)
    lam, mu = of.lame_converter(3.0, 0.0)
    assert lam == 0.0 and mu == pytest.approx(1.5)


# ----------------------------------------------------------------- SNH


def test_snh_rest_state_is_stress_free():
    F = np.broadcast_to(np.eye(3), (4, 3, 3)).copy()
    mu = np.full(4, 3.0)
    lam = np.full(4, 7.0)
    np.testing.assert_allclose(of.snh_energy_density(F, mu, lam), 0.0, atol=1e-14)
    np.testing.assert_allclose(of.snh_pk1(F, mu, lam), 0.0, atol=1e-14)


def test_snh_uniform_scale_closed_form():
    s = np.array([0.7, 1.0, 1.3, 2.0])
    F = s[:, None, None] * np.eye(3)
    mu, lam = np.full(4, 2.0), np.full(4, 5.0)
    want = 1.5 * mu * (s**2 - 1) - mu * (s**3 - 1) + 0.5 * lam * (s**3 - 1) ** 2
    np.testing.assert_allclose(of.snh_energy_density(F, mu, lam), want, rtol=1e-13)


def _fd_pk1(energy, F, h=1e-6):
    P = np.zeros_like(F)
    for i in range(3):
        for j in range(3):
            E = np.zeros_like(F)
            E[:, i, j] = h
            P[:, i, j] = (energy(F + E) - energy(F - E)) / (2 * h)
    return P


def test_snh_pk1_is_energy_derivative():
    F = _random_F(32)
    mu = np.full(32, 1.3)
    lam = np.full(32, 4.1)
    P = of.snh_pk1(F, mu, lam)
    Pfd = _fd_pk1(lambda G: of.snh_energy_density(G, mu, lam), F)
    np.testing.assert_allclose(P, Pfd, rtol=1e-7, atol=1e-7)


# ----------------------------------------------------------------- ARAP


def test_svd_rv_convention():
    F = _random_F(64, 0.8)
    F[:8] *= -1.0  # inverted
    U, s, V = of.svd_rv(F)
    np.testing.assert_allclose(np.linalg.det(U), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(V), 1.0, atol=1e-12)
    np.testing.assert_allclose(U @ (s[:, :, None] * np.swapaxes(V, 1, 2)), F, atol=1e-12)
    assert np.all(np.abs(s[:, 0]) >= np.abs(s[:, 1])) and np.all(np.abs(s[:, 1]) >= np.abs(s[:, 2]))
    assert np.all(s[:, :2] >= 0)
    np.testing.assert_array_equal(np.sign(s[:, 2]), np.sign(np.linalg.det(F)))


def test_arap_closed_forms():
    R = _rot(RNG, 16)
    mu = np.full(16, 2.5)
    np.testing.assert_allclose(of.arap_energy_density(R, mu), 0.0, atol=1e-13)
    s = np.linspace(0.5, 2.0, 16)
    np.testing.assert_allclose(
        of.arap_energy_density(s[:, None, None] * R, mu), 1.5 * mu * (s - 1) ** 2, rtol=1e-11, atol=1e-14
    )
    Rp, _ = of.polar_rv(s[:, None, None] * R)
    np.testing.assert_allclose(Rp, R, atol=1e-12)


def test_arap_pk1_is_energy_derivative():
    F = _random_F(32, 0.4)
    mu = np.full(32, 1.7)
    P = of.arap_pk1(F, mu)
    Pfd = _fd_pk1(lambda G: of.arap_energy_density(G, mu), F)
    np.testing.assert_allclose(P, Pfd, rtol=1e-6, atol=1e-7)


def _fd_element_hessian(pot, u, t, h=1e-5):
    """Exact 12x12 Hessian of tet t by central differences of its gradient."""
    cells = pot.cells[t]
    H = np.zeros((12, 12))
    for col in range(12):
        a, i = divmod(col, 3)
        up, um = u.copy(), u.copy()
        up[cells[a], i] += h
        um[cells[a], i] -= h
        gp = pot.grad_cells(up)[t].reshape(12)
        gm = pot.grad_cells(um)[t].reshape(12)
        H[:, col] = (gp - gm) / (2 * h)
    return H


def _stretch_mesh_state(rng, m, A_diag=(1.3, 1.2, 1.25)):
    """Stretched + rotated affine state, small noise: all sigma_i + sigma_j > 2."""
    R = _rot(rng, 1)[0]
    A = R @ np.diag(A_diag)
    return m.points @ (A - np.eye(3)).T + 0.003 * rng.standard_normal(m.points.shape)


@pytest.mark.parametrize("material", [of.SNH, of.ARAP, of.MUSCLE])
def test_element_hessian_matches_fd(material):
    rng = np.random.default_rng(7)
    m = cube(2)
    pot = _fem(material, m, rng)
    pot.mu[:] = 1.0
    if pot.lam is not None:
        pot.lam[:] = 3.0
    if material == of.ARAP:
        u = _stretch_mesh_state(rng, m)  # no lambda clamp fires: H is the exact Hessian
    else:
        u = 0.15 * rng.standard_normal(m.points.shape)
    He = pot.element_hessians(u)
    for t in range(0, pot.n_cells, 5):
        np.testing.assert_allclose(He[t], _fd_element_hessian(pot, u, t), rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("material", [of.SNH, of.ARAP, of.MUSCLE])
def test_global_grad_and_hess_prod_fd(material):
    rng = np.random.default_rng(11)
    m = cube(2)
    pot = _fem(material, m, rng)
    u = _stretch_mesh_state(rng, m) if material == of.ARAP else 0.1 * rng.standard_normal(m.points.shape)
    v = rng.standard_normal(u.shape)
    h = 1e-6
    # directional derivative of the energy
    d_fd = (pot.fun(u + h * v) - pot.fun(u - h * v)) / (2 * h)
    assert np.dot(pot.grad(u).ravel(), v.ravel()) == pytest.approx(d_fd, rel=1e-7)
    # hess_prod = derivative of grad along v
    hv_fd = (pot.grad(u + h * v) - pot.grad(u - h * v)) / (2 * h)
    np.testing.assert_allclose(pot.hess_prod(u, v), hv_fd, rtol=1e-5, atol=1e-6 * np.abs(hv_fd).max())


@pytest.mark.parametrize("material", [of.SNH, of.MUSCLE])
def test_snh_hess_diag_is_exact_diagonal(material):
    """SNH: every diag term is >= 0 and h6_diag == 0 is exact, so the clamped
    reference diagonal equals the diagonal of the exact assembled Hessian."""
    rng = np.random.default_rng(5)
    m = cube(2)
    pot = _fem(material, m, rng)
    u = 0.1 * rng.standard_normal(m.points.shape)
    hd = pot.hess_diag(u)
    He = pot.element_hessians(u)
    diag = np.zeros_like(hd)
    np.add.at(diag, pot.cells.ravel(), np.einsum("tii->ti", He).reshape(-1, 3))
    np.testing.assert_allclose(hd, diag, rtol=1e-12)
    # and the element Hessian diagonal itself is the FD diagonal
    np.testing.assert_allclose(
        np.diag(He[3]), np.diag(_fd_element_hessian(pot, u, 3)), rtol=1e-6
    )


def test_arap_hess_diag_matches_fd_when_unclamped():
    rng = np.random.default_rng(9)
    m = cube(2)
    pot = _fem(of.ARAP, m, rng)
    u = _stretch_mesh_state(rng, m)
    hd_cells = pot.hess_diag_cells(u)
    for t in range(0, pot.n_cells, 7):
        fd = np.diag(_fd_element_hessian(pot, u, t)).reshape(4, 3)
        np.testing.assert_allclose(hd_cells[t], np.maximum(fd, 0.0), rtol=1e-5, atol=1e-8)


@pytest.mark.parametrize("material", [of.SNH, of.ARAP, of.MUSCLE])
def test_hess_quad_clamp_and_consistency(material):
    rng = np.random.default_rng(3)
    m = cube(3)
    pot = _fem(material, m, rng)
    u = 0.08 * rng.standard_normal(m.points.shape)
    v = rng.standard_normal(u.shape)
    q_cells = pot.hess_quad_cells(u, v)
    hp_cells = pot.hess_prod_cells(u, v)
    vc = of.cell_values(v, pot.cells)
    raw = np.einsum("tai,tai->t", hp_cells, vc)
    # per-tet clamp max(0, .) (_base.py:379-380); hess_prod is unclamped
    np.testing.assert_allclose(q_cells, np.maximum(raw, 0.0), rtol=1e-10, atol=1e-12 * np.abs(raw).max())
    if np.all(raw >= 0):
        assert pot.hess_quad(u, v) == pytest.approx(np.dot(pot.hess_prod(u, v).ravel(), v.ravel()), rel=1e-10)


def test_arap_hessian_invariant_to_svd_convention():
    rng = np.random.default_rng(21)
    F = _random_F(16, 0.5, rng)
    dhdX = rng.standard_normal((16, 4, 3))
    p = rng.standard_normal((16, 4, 3))
    U, s, V = of.svd_rv(F)
    # flip the sign of one (U_i, V_i) pair: still a rotation-variant SVD of F
    U2, V2 = U.copy(), V.copy()
    U2[:, :, 0] *= -1
    U2[:, :, 1] *= -1
    V2[:, :, 0] *= -1
    V2[:, :, 1] *= -1
    np.testing.assert_allclose(U2 @ (s[:, :, None] * np.swapaxes(V2, 1, 2)), F, atol=1e-12)
    lam = of.arap_lambdas(s)
    dF = of.jvp(dhdX, p)
    q1 = sum(lam[:, k] * of.ddot(Q, dF) ** 2 for k, Q in enumerate(of.arap_Qs(U, V)))
    q2 = sum(lam[:, k] * of.ddot(Q, dF) ** 2 for k, Q in enumerate(of.arap_Qs(U2, V2)))
    np.testing.assert_allclose(q1, q2, rtol=1e-12)


def test_reference_arap_hess_prod_as_written_is_not_hv():
    """Documents the argument swap of _arap.py:55-56 (the product implements H v)."""
    rng = np.random.default_rng(2)
    m = cube(2)
    pot = _fem(of.ARAP, m, rng)
    u = _stretch_mesh_state(rng, m)
    v = rng.standard_normal(u.shape)
    F = pot.F(u)
    vc = of.cell_values(v, pot.cells)
    good = of.arap_hess_prod(F, vc, pot.dhdX, pot.mu)
    bad = of.arap_hess_prod_as_written(F, vc, pot.dhdX, pot.mu)
    assert np.abs(good - bad).max() > 1e-3 * np.abs(good).max()


@pytest.mark.parametrize("material", [of.SNH, of.ARAP, of.MUSCLE])
def test_mixed_derivative_fd(material):
    """p^T dg/dmu (and dg/dlam) against FD of the gradient in the parameter."""
    rng = np.random.default_rng(17)
    m = cube(2)
    pot = _fem(material, m, rng)
    u = 0.05 * rng.standard_normal(m.points.shape)
    p = rng.standard_normal(u.shape)
    dmu, dlam = pot.mixed_cells(u, p)
    h = 1.0  # the gradient is linear in mu and lambda: central differences are exact up to rounding
    for t in (0, 7, 19):
        mu0 = pot.mu[t]
        pot.mu[t] = mu0 + h
        gp = pot.grad(u)
        pot.mu[t] = mu0 - h
        gm = pot.grad(u)
        pot.mu[t] = mu0
        assert dmu[t] == pytest.approx(np.sum((gp - gm) / (2 * h) * p), rel=1e-6)
        if material != of.ARAP:
            l0 = pot.lam[t]
            pot.lam[t] = l0 + h
            gp = pot.grad(u)
            pot.lam[t] = l0 - h
            gm = pot.grad(u)
            pot.lam[t] = l0
            assert dlam[t] == pytest.approx(np.sum((gp - gm) / (2 * h) * p), rel=1e-6, abs=1e-12)


# ----------------------------------------------------------------- SNH muscle


def test_muscle_zero_activation_is_snh():
    rng = np.random.default_rng(31)
    m = cube(2)
    snh = _fem(of.SNH, m, rng)
    mus = of.FemPotential(of.MUSCLE, m.cells, snh.dhdX, snh.dV, snh.mu, snh.lam, np.zeros((m.n_cells, 6)))
    u = 0.1 * rng.standard_normal(m.points.shape)
    v = rng.standard_normal(u.shape)
    assert mus.fun(u) == pytest.approx(snh.fun(u), rel=1e-14)
    for name in ("grad", "hess_diag"):
        np.testing.assert_allclose(getattr(mus, name)(u), getattr(snh, name)(u), rtol=1e-13, atol=1e-9)
    np.testing.assert_allclose(mus.hess_prod(u, v), snh.hess_prod(u, v), rtol=1e-13, atol=1e-9)
    assert mus.hess_quad(u, v) == pytest.approx(snh.hess_quad(u, v), rel=1e-13)


def test_muscle_uniform_scale_closed_form():
    """u = 0, A = s I: G = s I, Psi = mu/2 (3 s^2 - 3) - mu (s^3 - 1) + lam/2 (s^3 - 1)^2,
    P_F = (mu s + (lam (s^3 - 1) - mu) s^2) s I."""
    T = 3
    s = np.array([0.8, 1.0, 1.3])
    mu = np.array([2.0, 3.0, 5.0])
    lam = np.array([7.0, 11.0, 13.0])
    a = np.zeros((T, 6))
    a[:, :3] = (s - 1.0)[:, None]
    F = np.broadcast_to(np.eye(3), (T, 3, 3)).copy()
    J = s**3
    psi = 0.5 * mu * (3 * s**2 - 3) - mu * (J - 1) + 0.5 * lam * (J - 1) ** 2
    np.testing.assert_allclose(of.muscle_energy_density(F, mu, lam, a), psi, rtol=1e-14)
    P = of.muscle_pk1(F, mu, lam, a)
    c = (mu * s + (lam * (J - 1) - mu) * s**2) * s
    np.testing.assert_allclose(P, c[:, None, None] * np.eye(3), rtol=1e-14, atol=1e-14)


def test_muscle_equals_snh_on_activated_rest_shape():
    """G = F A = dx/dY for the rest shape Y = A^-1 X: tet by tet, the muscle law
    at displacement u is SNH on rest corners A^-1 X at displacement X + u - A^-1 X
    (same dV).  Pins where A enters (F A, dhdX A, the A^T of the stress)."""
    rng = np.random.default_rng(41)
    m = cube(2)
    T = m.n_cells
    # one tet per element: independent corners, so each tet may have its own A
    X = m.points[m.cells].reshape(-1, 3)
    cells = np.arange(4 * T).reshape(T, 4)
    dhdX, dV = of.region(X, cells)
    act = _activation(rng, T, 0.2)
    A = of.activation_matrix(act)
    Y = np.einsum("tij,taj->tai", np.linalg.inv(A), X.reshape(T, 4, 3)).reshape(-1, 3)
    dhdY, _ = of.region(Y, cells)
    mu = 1e4 * np.exp(0.3 * rng.standard_normal(T))
    lam = 10.0 * mu
    mus = of.FemPotential(of.MUSCLE, cells, dhdX, dV, mu, lam, act)
    snh = of.FemPotential(of.SNH, cells, dhdY, dV, mu, lam)
    u = 0.1 * rng.standard_normal(X.shape)
    uY = X + u - Y
    v = rng.standard_normal(X.shape)
    np.testing.assert_allclose(mus.energy_density(u), snh.energy_density(uY), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(mus.grad(u), snh.grad(uY), rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(mus.hess_diag(u), snh.hess_diag(uY), rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(mus.hess_prod(u, v), snh.hess_prod(uY, v), rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(mus.hess_quad_cells(u, v), snh.hess_quad_cells(uY, v), rtol=1e-10, atol=1e-8)


def test_muscle_mixed_activation_fd():
    """p^T dg/da_k against central differences of the gradient in a_k."""
    rng = np.random.default_rng(43)
    m = cube(2)
    pot = _fem(of.MUSCLE, m, rng)
    u = 0.05 * rng.standard_normal(m.points.shape)
    p = rng.standard_normal(u.shape)
    dact = pot.mixed_activation_cells(u, p)
    h = 1e-6
    for t in (0, 9, 23):
        for k in range(6):
            a0 = pot.activation[t, k]
            pot.activation[t, k] = a0 + h
            gp = pot.grad(u)
            pot.activation[t, k] = a0 - h
            gm = pot.grad(u)
            pot.activation[t, k] = a0
            fd = np.sum((gp - gm) / (2 * h) * p)
            assert dact[t, k] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(dact).max())


def test_external_force_sign_convention():
    u = RNG.standard_normal((5, 3))
    f = RNG.standard_normal((2, 3))
    ext = of.ExternalForce(np.array([1, 3]), f)
    assert ext.fun(u) == pytest.approx(-(f * u[[1, 3]]).sum())
    g = ext.grad(u)
    np.testing.assert_allclose(g[[1, 3]], -f)
    np.testing.assert_allclose(g[[0, 2, 4]], 0.0)


@pytest.mark.parametrize("sigma", [(1.5, 0.9, 0.8), (0.9, 0.8, 0.7), (1.3, 1.2, 1.1)])
def test_arap_lambda_clamp_closed_form(sigma):
    """The clamp branch of lambda_k = 2 / max(s_i + s_j, 2) (func/_misc.py:31-43)
    pinned by the mathematics: for F = U diag(s) V^T and the unit twist
    W = U K_ab V^T (K_ab = (e_a e_b^T - e_b e_a^T)/sqrt2), the polar factor
    moves as dR[W] = 2/(s_a + s_b) W, so the exact curvature of
    Psi = mu/2 |F - R|^2 along W is mu (1 - 2/(s_a + s_b)) (negative under
    compression, s_a + s_b < 2) and the projected one mu (1 - min(1, 2/(s_a + s_b)));
    along U E_aa V^T (a stretch) both are mu.  The exact one is also the
    central difference of the stress (arap_pk1) along W."""
    rng = np.random.default_rng(11)
    U, V = _rot(rng, 2)  # two rotations
    s = np.asarray(sigma, np.float64)
    F = (U * s[None, :]) @ V.T
    mu = np.array([2.0])
    dhdX = np.array([[[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]])

    def quad(W, clamp):
        p = np.zeros((1, 4, 3))
        p[0, 1:] = W.T  # corner j + 1 moves by column j of W: dF = W
        assert np.allclose(of.jvp(dhdX, p)[0], W)
        return float((p * of.arap_hess_prod(F[None], p, dhdX, mu, clamp=clamp)).sum())

    for a, b in ((0, 1), (1, 2), (2, 0)):
        K = np.zeros((3, 3))
        K[a, b], K[b, a] = 1 / np.sqrt(2), -1 / np.sqrt(2)
        W = U @ K @ V.T
        lam = 2.0 / (s[a] + s[b])
        assert quad(W, True) == pytest.approx(mu[0] * (1 - min(1.0, lam)), abs=1e-12)
        assert quad(W, False) == pytest.approx(mu[0] * (1 - lam), rel=1e-10)
        h = 1e-6
        fd = (of.arap_pk1((F + h * W)[None], mu)[0] - of.arap_pk1((F - h * W)[None], mu)[0]) / (2 * h)
        assert float((fd * W).sum()) == pytest.approx(mu[0] * (1 - lam), rel=1e-6)
        if s[a] + s[b] < 2:
            assert quad(W, True) == pytest.approx(0.0, abs=1e-12) and quad(W, False) < 0
    for a in range(3):
        E = np.zeros((3, 3))
        E[a, a] = 1.0
        S = U @ E @ V.T
        assert quad(S, True) == pytest.approx(mu[0], rel=1e-12)
        assert quad(S, False) == pytest.approx(mu[0], rel=1e-12)
