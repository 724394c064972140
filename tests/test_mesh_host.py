"""Host-side mesh helpers: the region's dV <= 0 warning (jax/fem/region/
_region.py:98-99 warns the same way) and the numpy Morton order."""

import logging

import numpy as np

from apple_amd.mesh import Mesh, Region, cube, reorder_host


def test_inverted_tet_warns_and_keeps_negative_volume(caplog):
    m = cube(2)
    cells = m.cells.copy()
    cells[3, [1, 2]] = cells[3, [2, 1]]  # one inverted tet
    with caplog.at_level(logging.WARNING, logger="apple_amd.mesh"):
        r = Region.from_mesh(Mesh(m.points, cells))
    assert any("dV <= 0" in rec.getMessage() for rec in caplog.records)
    assert r.dV[3] < 0 and np.all(np.delete(r.dV, 3) > 0)
    caplog.clear()
    with caplog.at_level(logging.WARNING, logger="apple_amd.mesh"):
        Region.from_mesh(m)
    assert not caplog.records


def test_host_morton_order_is_a_permutation_restoring_locality():
    m = cube(6)
    rng = np.random.default_rng(0)
    vp = rng.permutation(m.n_points)
    rank = np.empty_like(vp)
    rank[vp] = np.arange(vp.size)
    shuffled = Mesh(m.points[vp], rank[m.cells][rng.permutation(m.n_cells)].astype(np.int32))
    new, vperm, cperm = reorder_host(shuffled)
    assert np.array_equal(np.sort(vperm), np.arange(m.n_points))
    assert np.array_equal(np.sort(cperm), np.arange(m.n_cells))
    # same geometry: every reordered tet is an original tet with relabelled vertices
    assert np.allclose(new.points[new.cells], shuffled.points[shuffled.cells[cperm]])
    # locality: consecutive tets share far more vertices after the reorder
    def shared(cells):
        return np.mean([len(set(cells[i]) & set(cells[i + 1])) for i in range(len(cells) - 1)])
    assert shared(new.cells) > 2 * shared(shuffled.cells)
