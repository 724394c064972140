"""Meshes at the layout's limits, on the GPU.

* Tets in random order (no locality): every block's vertex window fills the
  dtype's LDS cap (478 slots for FP64 meshes, 1024 for FP32).  Element
  passes, PNCG and PCG must run on such meshes and agree with the same mesh
  in locality order and with the oracle.
"""

import numpy as np
import pytest
import torch

from apple_amd import fem
from apple_amd.mesh import Region, cube
from apple_amd.solver import DeviceSolver, PncgConfig, Result, mask_bits
from helpers import oracle_pot, rel, rel_max

pytestmark = pytest.mark.gpu


def _shuffled_case(n, seed):
    m = cube(n)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(m.n_cells)
    cells = m.cells[perm]
    mu = 1e4 * np.exp(0.5 * rng.standard_normal(m.n_cells))
    return m, cells, perm, mu


@pytest.mark.parametrize("dtype,cap", [(torch.float64, 478), (torch.float32, 1024)])
def test_unordered_tets_fill_the_lds_cap_and_still_agree(dtype, cap):
    print(f"[limits] shuffled 16^3 dtype={dtype}", flush=True)
    m, cells, perm, mu = _shuffled_case(16, seed=3)
    ms = type(m)(m.points.copy(), cells)
    ms.cell_data["mu"] = mu[perm]
    ms.cell_data["lambda"] = 10 * mu[perm]
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = 10 * mu
    rs, ro = Region.from_mesh(ms), Region.from_mesh(m)
    chunks = 4 if dtype == torch.float32 else 1  # FP32: blocks of up to 1024 tets reach the cap
    pot_s = fem.StableNeoHookean.from_region(rs, dtype=dtype, device="cuda", chunks=chunks)
    pot_o = fem.StableNeoHookean.from_region(ro, dtype=dtype, device="cuda")
    info = pot_s.mesh.info()
    # a block closes when its next tet would pass the cap (up to 4 new vertices)
    assert cap - 3 <= info["nloc_max"] <= cap, info
    npdt = np.float32 if dtype == torch.float32 else np.float64
    rng = np.random.default_rng(4)
    u64 = (0.01 * rng.standard_normal((m.n_points, 3))).astype(npdt).astype(np.float64)
    v64 = rng.standard_normal((m.n_points, 3)).astype(npdt).astype(np.float64)
    u = torch.as_tensor(u64, dtype=dtype, device="cuda")
    v = torch.as_tensor(v64, dtype=dtype, device="cuda")
    ref = oracle_pot(rs, "snh", ms.cell_data["mu"], ms.cell_data["lambda"], dtype_round=npdt)
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = torch.zeros_like(u)
    h = torch.zeros_like(u)
    hp = torch.zeros_like(u)
    pot_s.fused(u, energy=e, grad=g, hess_diag=h)
    pot_s.hess_prod(u, v, hp)
    torch.cuda.synchronize()
    tol_e, tol_g = (1e-5, 1e-4) if dtype == torch.float32 else (1e-11, 1e-11)
    assert abs(e.item() - ref.fun(u64)) <= tol_e * abs(ref.fun(u64))
    assert rel(g.cpu().numpy(), ref.grad(u64)) <= tol_g
    assert rel_max(g.cpu().numpy(), ref.grad(u64)) <= tol_g
    assert rel(h.cpu().numpy(), ref.hess_diag(u64)) <= tol_g
    assert rel_max(h.cpu().numpy(), ref.hess_diag(u64)) <= tol_g
    assert rel(hp.cpu().numpy(), ref.hess_prod(u64, v64)) <= tol_g
    assert rel_max(hp.cpu().numpy(), ref.hess_prod(u64, v64)) <= tol_g

    # PNCG and PCG on the shuffled mesh == the ordered mesh (same problem)
    f = np.zeros((m.n_points, 3))
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e3 * ro.dV / 4, 4))
    fm = np.zeros((m.n_points, 3), bool)
    fm[m.points[:, 2] == 0] = True
    ft = torch.as_tensor(f, dtype=dtype, device="cuda")
    mask = mask_bits(fm, "cuda")
    rtol = 1e-9 if dtype == torch.float64 else 1e-5  # FP32: above the rounding floor of the shuffled sums
    xs = []
    for pot in (pot_s, pot_o):
        solver = DeviceSolver(pot)
        x = torch.zeros(m.n_points, 3, dtype=dtype, device="cuda")
        st = solver.solve(x, ft, mask, PncgConfig(rtol=rtol, max_steps=20000))[0]
        assert st.result == Result.PRIMARY_SUCCESS, st
        y = torch.zeros_like(x)
        cs = solver.cg(x, ft, y, mask, rtol=1e-8 if dtype == torch.float64 else 1e-5, max_iters=20000)[0]
        assert cs.result == Result.PRIMARY_SUCCESS, cs
        xs.append((x.double().cpu().numpy(), y.double().cpu().numpy()))
    diag = np.sqrt(3.0)
    tol_x = 1e-7 if dtype == torch.float64 else 1e-5
    assert np.abs(xs[0][0] - xs[1][0]).max() / diag <= tol_x
    assert rel(xs[0][1], xs[1][1]) <= (1e-6 if dtype == torch.float64 else 1e-3)
    assert rel_max(xs[0][1], xs[1][1]) <= (1e-6 if dtype == torch.float64 else 1e-3)


def test_external_force_skips_indices_outside_the_mesh():
    """At the C ABI an ExternalForce index outside [0, V) contributes nothing
    and touches no memory; the valid entries give E = -f . u[idx] and
    grad[idx] -= f exactly."""
    from apple_amd.runtime import DeviceMesh

    print("[limits] external force with out-of-range indices", flush=True)
    m = cube(3)
    V = m.n_points
    dm = DeviceMesh(m.cells, np.tile(np.eye(3), (m.n_cells, 1, 1)), np.full(m.n_cells, 1.0 / 6.0), V,
                    dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(11)
    idx = np.array([0, V, 5, -1, V - 1, 1 << 30, 7], np.int32)
    f = rng.standard_normal((idx.size, 3))
    u = rng.standard_normal((V, 3))
    ok = (idx >= 0) & (idx < V)
    e_ref = -float(np.sum(f[ok] * u[idx[ok]]))
    g_ref = np.full((V, 3), 2.0)
    g_ref[idx[ok]] -= f[ok]
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = torch.full((V, 3), 2.0, dtype=torch.float64, device="cuda")
    dm.ext_force(torch.as_tensor(idx, device="cuda"), torch.as_tensor(f, device="cuda"),
                 torch.as_tensor(u, device="cuda"), energy=e, grad=g, accumulate=True)
    torch.cuda.synchronize()
    assert e.item() == pytest.approx(e_ref, rel=1e-14)
    np.testing.assert_array_equal(g.cpu().numpy(), g_ref)
