"""Tet meshes and per-tet setup (host side).

``Mesh`` is a small stand-in for the pyvista ``UnstructuredGrid`` that
liblaf/apple reads (points, tetra cells, ``point_data`` and ``cell_data``
dictionaries keyed by the attribute names of ``apple_amd.common``).
``Region`` restates the setup of ``jax/fem/region/_region.py:84-108``: the
shape-function gradients ``dhdX`` and quadrature volumes ``dV`` of every tet,
computed once per mesh in FP64 on the host.

``cube`` generates the benchmark meshes: n^3 cells, each split into 5 tets
with alternating parity so that neighbouring cells share face diagonals.
Cells and vertices are emitted in Morton (Z-curve) order, so that any run of
consecutive tets covers a compact patch of the mesh -- the element kernels
stage runs of consecutive tets as one block, and that locality is what keeps
their vertex windows small.
"""

from __future__ import annotations

import dataclasses
import logging
from typing import Any

import numpy as np

logger = logging.getLogger(__name__)

# reference-element gradients and one-point rule (jax/fem/element/_tetra.py:36-45,
# jax/fem/quadrature/_tetra.py:14-15)
_DHDR = np.array(
    [[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
)
_QUAD_W = 1.0 / 6.0

# 5-tet split of a unit cell; corner b = x + 2y + 4z.  Even cells use the
# centre tet on the odd-popcount corners {1,2,4,7}, odd cells the mirror.
_TETS_EVEN = np.array([[0, 1, 2, 4], [3, 2, 1, 7], [5, 4, 7, 1], [6, 7, 4, 2], [1, 2, 4, 7]])
_TETS_ODD = np.array([[1, 0, 5, 3], [2, 3, 6, 0], [4, 5, 0, 6], [7, 6, 3, 5], [0, 3, 5, 6]])
_CORNER = np.array([[b & 1, (b >> 1) & 1, (b >> 2) & 1] for b in range(8)])


@dataclasses.dataclass
class Mesh:
    """Tet mesh with point/cell attributes (pyvista-like)."""

    points: np.ndarray  # (V,3) float64
    cells: np.ndarray  # (T,4) int32
    point_data: dict[str, Any] = dataclasses.field(default_factory=dict)
    cell_data: dict[str, Any] = dataclasses.field(default_factory=dict)

    @property
    def n_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def n_cells(self) -> int:
        return int(self.cells.shape[0])

    @property
    def volume(self) -> float:
        X = self.points[self.cells]
        d = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2))
        return float(np.abs(d).sum() / 6.0)

    def copy(self) -> "Mesh":
        return Mesh(
            self.points.copy(),
            self.cells.copy(),
            {k: np.array(v, copy=True) for k, v in self.point_data.items()},
            {k: np.array(v, copy=True) for k, v in self.cell_data.items()},
        )

    @classmethod
    def from_pyvista(cls, obj: Any) -> "Mesh":
        """Accept a pyvista UnstructuredGrid of tetra cells (duck-typed)."""
        if isinstance(obj, Mesh):
            return obj
        cells = np.asarray(obj.cells_dict[10], np.int32)  # VTK_TETRA == 10
        return cls(
            np.asarray(obj.points, np.float64),
            cells,
            {k: np.asarray(obj.point_data[k]) for k in obj.point_data.keys()},
            {k: np.asarray(obj.cell_data[k]) for k in obj.cell_data.keys()},
        )


def _morton3(i: np.ndarray, j: np.ndarray, k: np.ndarray) -> np.ndarray:
    code = np.zeros(i.shape, np.int64)
    for b in range(11):
        code |= ((i >> b) & 1).astype(np.int64) << (3 * b)
        code |= ((j >> b) & 1).astype(np.int64) << (3 * b + 1)
        code |= ((k >> b) & 1).astype(np.int64) << (3 * b + 2)
    return code


def cube(n: int, lo: float = 0.0, hi: float = 1.0, *, order: str = "morton") -> Mesh:
    """Cube [lo,hi]^3 with n^3 cells x 5 tets, positively oriented.

    ``order="morton"`` emits cells and vertices along a Z-curve;
    ``order="lex"`` emits them x-fastest.
    """
    if n < 1 or n > 2000:
        raise ValueError(f"cube resolution must be in [1, 2000], got {n}")
    m = n + 1
    vi, vj, vk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    vi, vj, vk = vi.ravel(), vj.ravel(), vk.ravel()  # lexicographic id = (i*m + j)*m + k
    if order == "morton":
        vperm = np.argsort(_morton3(vi, vj, vk), kind="stable")
    elif order == "lex":
        vperm = np.lexsort((vi, vj, vk))  # x fastest
    else:
        raise ValueError(order)
    vrank = np.empty_like(vperm)
    vrank[vperm] = np.arange(vperm.size)
    h = (hi - lo) / n
    points = np.stack([vi[vperm], vj[vperm], vk[vperm]], axis=1).astype(np.float64) * h + lo

    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    if order == "morton":
        cperm = np.argsort(_morton3(ci, cj, ck), kind="stable")
    else:
        cperm = np.lexsort((ci, cj, ck))
    ci, cj, ck = ci[cperm], cj[cperm], ck[cperm]
    parity = (ci + cj + ck) & 1
    local = np.where(parity[:, None, None] == 0, _TETS_EVEN[None], _TETS_ODD[None])  # (C,5,4)
    off = _CORNER[local]  # (C,5,4,3)
    gi = ci[:, None, None] + off[..., 0]
    gj = cj[:, None, None] + off[..., 1]
    gk = ck[:, None, None] + off[..., 2]
    lex = (gi * m + gj) * m + gk
    cells = vrank[lex].reshape(-1, 4).astype(np.int32)
    cells = orient(points, cells)
    return Mesh(points, cells)


def orient(points: np.ndarray, cells: np.ndarray) -> np.ndarray:
    """Swap two corners of every negatively oriented tet."""
    X = points[cells]
    d = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2))
    cells = cells.copy()
    neg = d < 0
    cells[neg, 1], cells[neg, 2] = cells[neg, 2].copy(), cells[neg, 1].copy()
    return cells


@dataclasses.dataclass
class Region:
    """Per-tet shape gradients and volumes -- jax/fem/region/_region.py:84-108.

    ``dhdX[t]`` (4,3) is dhdr @ inv(dXdr) and ``dV[t]`` = det(dXdr) / 6
    (one-point rule).  ``dminv[t]`` (3,3) = inv(dXdr) = rows 1..3 of dhdX,
    which is what the device stores (row 0 is minus their sum).
    """

    mesh: Mesh
    dhdX: np.ndarray  # (T,4,3) float64
    dV: np.ndarray  # (T,) float64
    dminv: np.ndarray  # (T,3,3) float64

    @classmethod
    def from_mesh(cls, mesh: Mesh | Any) -> "Region":
        mesh = Mesh.from_pyvista(mesh)
        X = mesh.points[mesh.cells]
        dXdr = np.einsum("tai,aj->tij", X, _DHDR)
        drdX = np.linalg.inv(dXdr)
        dV = np.linalg.det(dXdr) * _QUAD_W
        if np.any(dV <= 0):
            logger.warning("dV <= 0")  # same warning as _region.py:98-99
        dhdX = np.einsum("ai,tij->taj", _DHDR, drdX)
        return cls(mesh, dhdX, dV, drdX)

    from_pyvista = from_mesh

    @property
    def cells(self) -> np.ndarray:
        return self.mesh.cells

    @property
    def cell_data(self) -> dict:
        return self.mesh.cell_data

    @property
    def n_cells(self) -> int:
        return self.mesh.n_cells


def reorder(mesh: Mesh, device=None) -> tuple[Mesh, np.ndarray, np.ndarray]:
    """Morton-sort the vertices and tets of an arbitrary mesh on the device.

    Returns (new_mesh, vertex_perm, cell_perm) with new = old[perm].  The
    element kernels are correct for any order; this only restores the
    locality their block staging relies on.  The codes and the stable sorts
    run on the GPU (``apl_morton_order``); the host applies the permutations
    to the mesh's arrays.  ``reorder_host`` is the same order computed with
    numpy, for preprocessing without a GPU.
    """
    import torch

    from .runtime import morton_order, normalize_device

    dev = normalize_device(device)
    pts = torch.as_tensor(np.ascontiguousarray(mesh.points, np.float64), device=dev)
    cells = torch.as_tensor(np.ascontiguousarray(mesh.cells, np.int32), device=dev)
    vp, cp = morton_order(pts, cells)
    vperm = vp.cpu().numpy().astype(np.int64)
    cperm = cp.cpu().numpy().astype(np.int64)
    return _permuted(mesh, vperm, cperm), vperm, cperm


def _permuted(mesh: Mesh, vperm: np.ndarray, cperm: np.ndarray) -> Mesh:
    vrank = np.empty_like(vperm)
    vrank[vperm] = np.arange(vperm.size)
    cells = vrank[mesh.cells]
    return Mesh(
        mesh.points[vperm].copy(),
        cells[cperm].astype(np.int32),
        {k: np.asarray(v)[vperm] for k, v in mesh.point_data.items()},
        {k: np.asarray(v)[cperm] for k, v in mesh.cell_data.items()},
    )


def reorder_host(mesh: Mesh) -> tuple[Mesh, np.ndarray, np.ndarray]:
    """``reorder`` computed on the host with numpy (same permutations)."""
    P = mesh.points
    lo, hi = P.min(0), P.max(0)
    q = ((P - lo) / np.maximum(hi - lo, 1e-30) * 2047).astype(np.int64)
    vperm = np.argsort(_morton3(q[:, 0], q[:, 1], q[:, 2]), kind="stable")
    C = P[mesh.cells].mean(1)
    qc = ((C - lo) / np.maximum(hi - lo, 1e-30) * 2047).astype(np.int64)
    cperm = np.argsort(_morton3(qc[:, 0], qc[:, 1], qc[:, 2]), kind="stable")
    return _permuted(mesh, vperm, cperm), vperm, cperm
