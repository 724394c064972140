"""Setup on the device (apl_mesh_create_from_points) against the oracle's
region (jax/fem/region/_region.py:84-108 restated in oracle/fem.py: dhdX rows
1..3 = inv(dXdr), dV = det(dXdr)/6, times Fraction), and potentials built by
device setup against those built from the host region."""

import numpy as np
import pytest
import torch

import oracle.fem as of
from apple_amd import fem, potential
from apple_amd.mesh import Mesh, Region, cube
from apple_amd.runtime import DeviceMesh

pytestmark = pytest.mark.gpu


def _irregular(n, seed):
    m = cube(n)
    rng = np.random.default_rng(seed)
    pts = m.points + 0.15 / n * rng.standard_normal(m.points.shape)
    cells = m.cells.copy()
    cells[0, [1, 2]] = cells[0, [2, 1]]  # one inverted tet: dV < 0, as the reference keeps it
    mesh = Mesh(pts, cells)
    mu = 1e4 * np.exp(0.3 * rng.standard_normal(mesh.n_cells))
    mesh.cell_data["mu"] = mu
    mesh.cell_data["lambda"] = 10 * mu
    mesh.cell_data["Fraction"] = rng.uniform(0.5, 1.0, mesh.n_cells)
    return mesh


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_device_region_matches_oracle(dtype):
    mesh = _irregular(6, 1)
    dm = DeviceMesh(mesh.cells, None, None, mesh.n_points, dtype=dtype, device="cuda", points=mesh.points,
                    fraction=mesh.cell_data["Fraction"])
    dminv, dV = dm.region()
    dhdX, dV_ref = of.region(mesh.points, mesh.cells, mesh.cell_data["Fraction"])
    assert dV_ref[0] < 0
    dminv, dV = dminv.double().cpu().numpy(), dV.double().cpu().numpy()
    tol = 1e-12 if dtype == torch.float64 else 2e-7
    np.testing.assert_allclose(dV, dV_ref, rtol=tol)
    scale = np.abs(dhdX[:, 1:]).max(axis=(1, 2), keepdims=True)
    assert np.abs(dminv - dhdX[:, 1:]).max() <= tol * scale.max()
    assert np.all(np.abs(dminv - dhdX[:, 1:]) <= tol * scale)


@pytest.mark.parametrize("cls", [fem.StableNeoHookean, fem.Arap, fem.StableNeoHookeanMuscle])
def test_device_setup_potential_equals_host_setup(cls):
    mesh = _irregular(5, 2)
    mesh.cell_data["activation"] = 0.1 * np.random.default_rng(3).standard_normal((mesh.n_cells, 6))
    a = cls.from_pyvista(mesh, setup="device", dtype=torch.float64, device="cuda")
    b = cls.from_pyvista(mesh, setup="host", dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(4)
    u = torch.as_tensor(0.01 * rng.standard_normal((mesh.n_points, 3)), dtype=torch.float64, device="cuda")
    out = []
    for pot in (a, b):
        e = torch.zeros(1, dtype=torch.float64, device="cuda")
        g = torch.zeros_like(u)
        h = torch.zeros_like(u)
        pot.fused(u, energy=e, grad=g, hess_diag=h)
        out.append((e.item(), g.cpu().numpy(), h.cpu().numpy()))
    assert out[0][0] == pytest.approx(out[1][0], rel=1e-12)
    for k in (1, 2):
        np.testing.assert_allclose(out[0][k], out[1][k], rtol=1e-10, atol=1e-12 * np.abs(out[1][k]).max())


def test_gravity_from_device_setup_equals_host():
    mesh = _irregular(4, 5)
    del mesh.cell_data["Fraction"]
    a = fem.StableNeoHookean.from_pyvista(mesh, setup="device", dtype=torch.float64, device="cuda")
    ga = potential.gravity_from_potential(a, 1e3)
    gb = potential.gravity(Region.from_mesh(mesh), 1e3, dtype=torch.float64, device="cuda")
    np.testing.assert_allclose(ga.dense().cpu().numpy(), gb.dense().cpu().numpy(), rtol=1e-12, atol=1e-12)
