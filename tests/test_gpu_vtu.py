"""A forward solve read from and written to .vtu files, as the reference's
experiments run it (exp/2026/03/11/test-inverse/src/10-gen.py:14-17 writes the
input mesh with its Dirichlet attributes, 20-forward.py:21-60 loads it, adds
materials, solves and saves the solution as point data): the reference's
patch test (tests/forward/test_static_simulation.py:58-93) goes through a
compressed binary .vtu on the way in and an appended one on the way out, and
the known answer must come back."""

import numpy as np
import pytest
import torch

from apple_amd import fem, forward, vtu
from apple_amd.common import ACTIVATION, DISPLACEMENT, FIXED_MASK, FIXED_VALUE, LAMBDA, MU
from apple_amd.mesh import Mesh

from helpers import patch_test

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cls", [fem.Arap, fem.StableNeoHookean, fem.StableNeoHookeanMuscle])
def test_patch_test_through_vtu_files(tmp_path, cls):
    # the muscle law with zero activation is SNH (bit for bit), so the same
    # known answer holds; it also reads its activation from the file
    print(f"[vtu] patch {cls.__name__} through .vtu files", flush=True)
    points, cells, mask, values, ref = patch_test()
    src = Mesh(points, cells)
    src.point_data[FIXED_MASK.vtk] = mask
    src.point_data[FIXED_VALUE.vtk] = values
    src.cell_data[MU.vtk] = np.full(cells.shape[0], ref["mu"])
    src.cell_data[LAMBDA.vtk] = np.ones(cells.shape[0])
    if cls is fem.StableNeoHookeanMuscle:
        src.cell_data[ACTIVATION.value] = np.zeros((cells.shape[0], 6))  # utils/_material.py:16
    vtu.save_vtu(tmp_path / "input.vtu", src)  # binary, zlib

    mesh = vtu.load_vtu(tmp_path / "input.vtu")
    builder = forward.ModelBuilder()
    builder.add_vertices(mesh)
    builder.add_fixed(mesh)
    builder.add_potential(cls.from_pyvista(mesh, dtype=torch.float64, device="cuda"))
    model = builder.finalize()
    fwd = forward.Forward(model, forward.PNCG(rtol=1e-9))
    sol = fwd.step()
    assert sol.result.name == ref["expected"]["result"]
    assert model.n_free == ref["expected"]["n_free"]

    vtu.save_vtu(tmp_path / "solution.vtu", vtu.with_solution(mesh, fwd.state.u), encoding="appended")
    out = vtu.load_vtu(tmp_path / "solution.vtu")
    u = out.point_data[DISPLACEMENT.vtk]
    np.testing.assert_array_equal(u, fwd.state.u.double().cpu().numpy())
    exp = ref["expected"]
    np.testing.assert_allclose(u[exp["interior_vertex"]], exp["interior_displacement"], atol=max(1e-8, exp["atol"]))
    np.testing.assert_allclose(u[mask[:, 0]], values[mask[:, 0]], atol=1e-12)
