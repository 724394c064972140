"""Element pass parity: HIP kernels (through the C ABI) vs the FP64 oracle.

Tolerances (the targets stated for this path): energies 1e-5 relative,
gradients 1e-4 relative (norm-wise); the same 1e-4 for the Hessian diagonal,
Hessian-vector product, quadratic form and mixed derivatives.  FP64 mode is
held to 1e-10.  Inputs are identical: the FP32 run and the oracle see the same
float32-rounded u, p, mu and lambda; only the region data differ by their
FP32 rounding, which is part of the FP32 path's error.
"""

import numpy as np
import pytest
import torch

from apple_amd import fem
from helpers import oracle_pot, rel, rel_max, snh_case

pytestmark = pytest.mark.gpu

TOL = {torch.float32: dict(e=1e-5, g=1e-4), torch.float64: dict(e=1e-10, g=1e-10)}


def _run(pot, u, v):
    dev = pot.device
    V = pot.n_points
    out = {}
    e = torch.zeros(1, dtype=torch.float64, device=dev)
    g = torch.zeros(V, 3, dtype=pot.dtype, device=dev)
    h = torch.zeros_like(g)
    pot.fused(u, energy=e, grad=g, hess_diag=h)
    hp = torch.zeros_like(g)
    pot.hess_prod(u, v, hp)
    hq = torch.zeros(1, dtype=torch.float64, device=dev)
    pot.hess_quad(u, v, hq)
    mix = pot.mixed_derivative_prod(u, v)
    torch.cuda.synchronize()
    out.update(e=e.item(), g=g.cpu().numpy(), h=h.cpu().numpy(), hp=hp.cpu().numpy(), hq=hq.item())
    out.update({k: t.cpu().numpy() for k, t in mix.items()})
    return out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_arap_compressed_state_takes_the_clamp_branch(dtype):
    """ARAP on a cube compressed to 0.85 of its size plus noise: nearly every
    twist pair has s_i + s_j < 2, so the lambda clamp decides Hdiag, H v and
    v^T H v; the kernels match the oracle there (whose clamp branch is pinned
    by the closed form of test_oracle_fem.test_arap_lambda_clamp_closed_form)."""
    import oracle.fem as of

    m, region, _, v64 = snh_case(8, seed=3)
    u64 = -0.15 * m.points + 0.005 * np.random.default_rng(4).standard_normal(m.points.shape)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    u_r = u64.astype(npdt).astype(np.float64)
    v_r = v64.astype(npdt).astype(np.float64)
    ref = oracle_pot(region, "arap", m.cell_data["mu"], None, dtype_round=npdt)
    _, s, _ = of.svd_rv(ref.F(u_r))
    pairs = np.stack([s[:, 0] + s[:, 1], s[:, 1] + s[:, 2], s[:, 2] + s[:, 0]], 1)
    print(f"[elem] ARAP compressed {dtype}: clamped pairs {np.mean(pairs < 2):.1%}", flush=True)
    assert np.mean(pairs < 2) > 0.9
    pot = fem.Arap.from_region(region, dtype=dtype, device="cuda")
    got = _run(pot, torch.as_tensor(u_r, dtype=dtype, device="cuda"), torch.as_tensor(v_r, dtype=dtype, device="cuda"))
    tol = TOL[dtype]
    for k, r in (("g", ref.grad(u_r)), ("h", ref.hess_diag(u_r)), ("hp", ref.hess_prod(u_r, v_r))):
        assert rel(got[k], r) <= tol["g"], k
        assert rel_max(got[k], r) <= tol["g"], k
    hq_ref = ref.hess_quad(u_r, v_r)
    assert abs(got["hq"] - hq_ref) <= tol["g"] * abs(hq_ref)
    # the projected quadratic form differs from the exact one here
    assert hq_ref > 1.01 * of.arap_hess_quad(ref.F(u_r), of.cell_values(v_r, region.cells), region.dhdX,
                                             ref.mu, clamp=False).clip(min=0).dot(region.dV)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("material", ["snh", "arap"])
@pytest.mark.parametrize("n,noise,chunks", [(4, 0.01, 1), (10, 0.01, 1), (12, 0.003, 4)])
def test_element_pass_matches_oracle(dtype, material, n, noise, chunks):
    print(f"[elem] material={material} dtype={dtype} n={n} chunks={chunks}", flush=True)
    m, region, u64, v64 = snh_case(n, seed=n, noise=noise)
    cls = fem.StableNeoHookean if material == "snh" else fem.Arap
    pot = cls.from_region(region, dtype=dtype, device="cuda", chunks=chunks)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    u_r = u64.astype(npdt).astype(np.float64)
    v_r = v64.astype(npdt).astype(np.float64)
    ref = oracle_pot(region, material, m.cell_data["mu"], m.cell_data["lambda"], dtype_round=npdt)
    u = torch.as_tensor(u_r, dtype=dtype, device="cuda")
    v = torch.as_tensor(v_r, dtype=dtype, device="cuda")
    got = _run(pot, u, v)
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
    if material == "snh":
        assert rel(got["lambda"], dlam_ref) <= tol["g"]
        assert rel_max(got["lambda"], dlam_ref) <= tol["g"]


@pytest.mark.parametrize("material", ["snh", "arap"])
def test_bitwise_repeatable(material):
    m, region, u64, v64 = snh_case(8, seed=3)
    cls = fem.StableNeoHookean if material == "snh" else fem.Arap
    pot = cls.from_region(region, dtype=torch.float32, device="cuda")
    u = torch.as_tensor(u64, dtype=torch.float32, device="cuda")
    v = torch.as_tensor(v64, dtype=torch.float32, device="cuda")
    a = _run(pot, u, v)
    b = _run(pot, u, v)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_accumulate_semantics_and_zero_state():
    m, region, u64, v64 = snh_case(3, seed=1)
    pot = fem.StableNeoHookean.from_region(region, dtype=torch.float64, device="cuda")
    u = torch.zeros(m.n_points, 3, dtype=torch.float64, device="cuda")
    g = torch.ones(m.n_points, 3, dtype=torch.float64, device="cuda")
    pot.grad(u, g)  # rest state is stress free: += 0
    assert torch.equal(g, torch.ones_like(g))
    e = torch.full((1,), 2.5, dtype=torch.float64, device="cuda")
    pot.fun(u, e)
    assert e.item() == pytest.approx(2.5, abs=1e-12)
    e32 = torch.zeros(1, dtype=torch.float32, device="cuda")
    uu = torch.as_tensor(u64, dtype=torch.float64, device="cuda")
    pot.fun(uu, e32)
    ref = oracle_pot(region, "snh", m.cell_data["mu"], m.cell_data["lambda"])
    assert e32.item() == pytest.approx(ref.fun(u64), rel=1e-6)


def test_empty_mesh():
    from apple_amd.runtime import DeviceMesh

    dm = DeviceMesh(np.zeros((0, 4), np.int32), np.zeros((0, 3, 3)), np.zeros(0), 4, dtype=torch.float32, device="cuda")
    e = torch.full((1,), 7.0, dtype=torch.float64, device="cuda")
    mu = torch.zeros(0, dtype=torch.float32, device="cuda")
    u = torch.zeros(4, 3, dtype=torch.float32, device="cuda")
    g = torch.zeros(4, 3, dtype=torch.float32, device="cuda")
    dm.eval(0, mu, mu, u, energy=e, grad=g)
    torch.cuda.synchronize()
    assert e.item() == 0.0 and torch.count_nonzero(g).item() == 0


def test_graph_replayed_passes_equal_eager_passes():
    """The measurement entry point that replays K whole passes from a HIP
    graph (apl_bench_elem_graph) writes what K eager passes write, bit for
    bit, on a replay after its capture too."""
    import ctypes

    from apple_amd import _lib
    from apple_amd.runtime import current_stream, ptr

    m, region, u64, _ = snh_case(8, seed=3)
    pot = fem.StableNeoHookean.from_region(region, dtype=torch.float32, device="cuda")
    u = torch.as_tensor(u64, dtype=torch.float32, device="cuda")
    dm = pot.mesh
    what = _lib.APL_ENERGY | _lib.APL_GRAD | _lib.APL_HESS_DIAG
    outs = []
    for fn in ("apl_bench_elem_pass", "apl_bench_elem_graph", "apl_bench_elem_graph"):
        a = torch.full((dm.n_verts, 3), 7.0, dtype=torch.float32, device="cuda")
        b = torch.full_like(a, 7.0)
        e = torch.full((1,), 7.0, dtype=torch.float64, device="cuda")
        ms = ctypes.c_double()
        args = (dm._handle, int(pot.material), what, ptr(pot.materials["mu"]), ptr(pot.materials["lambda"]), ptr(u),
                None, ptr(e), ptr(a), ptr(b), 3)
        if fn == "apl_bench_elem_pass":
            _lib.check(getattr(_lib.lib(), fn)(*args, None, ctypes.byref(ms), current_stream(dm.device)))
        else:
            _lib.check(getattr(_lib.lib(), fn)(*args, ctypes.byref(ms), current_stream(dm.device)))
        torch.cuda.synchronize()
        assert ms.value > 0.0
        outs.append((e.cpu().numpy(), a.cpu().numpy(), b.cpu().numpy()))
    for got in outs[1:]:
        for x, y in zip(outs[0], got):
            assert np.array_equal(x, y)
    ref = oracle_pot(region, "snh", m.cell_data["mu"], m.cell_data["lambda"], dtype_round=np.float32)
    u_r = u64.astype(np.float32).astype(np.float64)
    assert abs(outs[0][0][0] - ref.fun(u_r)) <= 1e-5 * abs(ref.fun(u_r))
    assert rel(outs[0][1], ref.grad(u_r)) <= 1e-4


@pytest.mark.parametrize("material", ["snh", "arap"])
def test_graph_replay_recaptures_when_the_pass_changes(material):
    """One mesh, the graph entry point called for pass A, the quadratic form,
    the Hessian-vector product and pass A again: each call captures the new
    pass (its key changed) and writes what the eager passes write, bit for
    bit, and what the oracle gives."""
    import ctypes

    from apple_amd import _lib
    from apple_amd.runtime import current_stream, ptr

    m, region, u64, v64 = snh_case(7, seed=5)
    cls = fem.StableNeoHookean if material == "snh" else fem.Arap
    pot = cls.from_region(region, dtype=torch.float32, device="cuda")
    u = torch.as_tensor(u64, dtype=torch.float32, device="cuda")
    v = torch.as_tensor(v64, dtype=torch.float32, device="cuda")
    dm = pot.mesh
    lam = ptr(pot.materials["lambda"]) if material == "snh" else None
    A = _lib.APL_ENERGY | _lib.APL_GRAD | _lib.APL_HESS_DIAG
    ref = oracle_pot(region, material, m.cell_data["mu"], m.cell_data["lambda"], dtype_round=np.float32)
    u_r = u64.astype(np.float32).astype(np.float64)
    v_r = v64.astype(np.float32).astype(np.float64)

    def run(fn, what):
        a = torch.full((dm.n_verts, 3), 7.0, dtype=torch.float32, device="cuda")
        b = torch.full_like(a, 7.0)
        e = torch.full((1,), 7.0, dtype=torch.float64, device="cuda")
        ms = ctypes.c_double()
        p = ptr(v) if what != A else None
        args = (dm._handle, int(pot.material), what, ptr(pot.materials["mu"]), lam, ptr(u), p, ptr(e), ptr(a),
                ptr(b), 2)
        if fn == "apl_bench_elem_pass":
            _lib.check(getattr(_lib.lib(), fn)(*args, None, ctypes.byref(ms), current_stream(dm.device)))
        else:
            _lib.check(getattr(_lib.lib(), fn)(*args, ctypes.byref(ms), current_stream(dm.device)))
        torch.cuda.synchronize()
        assert ms.value > 0.0
        return e.cpu().numpy(), a.cpu().numpy(), b.cpu().numpy()

    for what in (A, _lib.APL_HESS_QUAD, _lib.APL_HESS_PROD, A):
        print(f"[elem] graph replay {material} what={what}", flush=True)
        eager = run("apl_bench_elem_pass", what)
        graph = run("apl_bench_elem_graph", what)
        e, a, b = graph
        if what == A:
            for x, y in zip(eager, graph):
                assert np.array_equal(x, y)
            assert abs(e[0] - ref.fun(u_r)) <= 1e-5 * abs(ref.fun(u_r))
            assert rel(a, ref.grad(u_r)) <= 1e-4 and rel(b, ref.hess_diag(u_r)) <= 1e-4
        elif what == _lib.APL_HESS_QUAD:
            assert np.array_equal(eager[0], e)
            hq = ref.hess_quad(u_r, v_r)
            assert abs(e[0] - hq) <= 1e-4 * abs(hq)
        else:
            assert np.array_equal(eager[1], a)
            assert rel(a, ref.hess_prod(u_r, v_r)) <= 1e-4


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("material", ["snh", "arap"])
def test_inverted_and_flat_tets_match_oracle(dtype, material):
    """A 6^3 cube displaced by noise of half a cell: 22 % of the tets are
    inverted (det F down to -4.5) and 0.7 % nearly flat (|det F| < 1/64, the
    ARAP kernels' FP64 finish); every output of the element passes (the
    mixed material derivatives included) still matches the oracle, the ARAP ones through the rotation-variant SVD
    convention (the sign goes to the smallest singular value)."""
    m, region, u64, v64 = snh_case(6, seed=11, noise=0.08)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    u_r = u64.astype(npdt).astype(np.float64)
    v_r = v64.astype(npdt).astype(np.float64)
    ref = oracle_pot(region, material, m.cell_data["mu"], m.cell_data["lambda"], dtype_round=npdt)
    det = np.linalg.det(ref.F(u_r))
    print(f"[elem] inverted {material} {dtype}: {np.mean(det < 0):.1%} inverted, "
          f"{np.mean(np.abs(det) < 1 / 64):.1%} flat", flush=True)
    assert np.mean(det < 0) > 0.2
    cls = fem.StableNeoHookean if material == "snh" else fem.Arap
    pot = cls.from_region(region, dtype=dtype, device="cuda")
    got = _run(pot, torch.as_tensor(u_r, dtype=dtype, device="cuda"), torch.as_tensor(v_r, dtype=dtype, device="cuda"))
    tol = TOL[dtype]
    e_ref = ref.fun(u_r)
    assert abs(got["e"] - e_ref) <= tol["e"] * abs(e_ref)
    for k, r in (("g", ref.grad(u_r)), ("h", ref.hess_diag(u_r)), ("hp", ref.hess_prod(u_r, v_r))):
        print(f"  {k}: rel {rel(got[k], r):.2e}, max-entry {rel_max(got[k], r):.2e}")
        assert rel(got[k], r) <= tol["g"], k
        assert rel_max(got[k], r) <= tol["g"], k
    hq_ref = ref.hess_quad(u_r, v_r)
    assert abs(got["hq"] - hq_ref) <= tol["g"] * abs(hq_ref)
    dmu_ref, dlam_ref = ref.mixed_cells(u_r, v_r)
    assert rel(got["mu"], dmu_ref) <= tol["g"] and rel_max(got["mu"], dmu_ref) <= tol["g"]
    if material == "snh":
        assert rel(got["lambda"], dlam_ref) <= tol["g"] and rel_max(got["lambda"], dlam_ref) <= tol["g"]
