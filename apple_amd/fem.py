"""FEM energy potentials with liblaf/apple's interface, on the HIP kernels.

``StableNeoHookean``, ``Arap`` and ``StableNeoHookeanMuscle`` keep the methods of
``WarpPotentialFem`` (warp/fem/_base.py:39-194) and the protocol of
``WarpPotential`` (warp/model/_potential.py:9-32):

    fun(u, output)            output (1,)  += sum_t Psi_t dV_t
    grad(u, output)           output (V,3) += dE/du
    hess_diag(u, output)      output (V,3) += max(dV diag H_e, 0) scattered
    hess_prod(u, p, output)   output (V,3) += H p            (unclamped)
    hess_quad(u, p, output)   output (1,)  += sum_t max(dV p_e^T H_e p_e, 0)

with the reference's accumulate (``+=``) semantics: the caller zeroes the
output first (``Model`` does).  ``u``/``p`` are (V,3) device tensors in the
potential's dtype; scalar outputs may be float64 or the potential dtype.
Additionally ``fused`` evaluates energy, gradient and Hessian diagonal in one
pass, and ``mixed_derivative_prod`` returns p^T d(grad)/d(material) per tet
(the adjoint's mixed term).

Differences from the reference, by design:

* ``Arap.hess_prod`` returns H p.  The reference passes (dhdX, p) to
  functions taking (p, dhdX) (warp/fem/_arap.py:55-56 vs
  func/_hess_prod.py:38-39,55), so its result is not H p.
* Reductions are deterministic (fixed order) instead of float atomics.
"""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any

import numpy as np
import torch

from . import _lib
from .common import ACTIVATION, FRACTION, GLOBAL_POINT_ID, LAMBDA, MU, default_name
from .mesh import Mesh, Region
from .runtime import DeviceMesh, as_dtype


def _global_cells(mesh: Mesh) -> tuple[np.ndarray, int]:
    """Cells in the model's vertex numbering: through point_data['GlobalPointId']
    when a ModelBuilder assigned one (the reference indexes
    cells_global = GlobalPointId[cells], jax/fem/geometry/_geometry.py:40-41),
    with the vertex count up to the highest global id; else as given."""
    gid = mesh.point_data.get(GLOBAL_POINT_ID.vtk)
    if gid is None:
        return mesh.cells, mesh.n_points
    gid = np.asarray(gid, np.int64).reshape(-1)
    if gid.shape[0] != mesh.n_points:
        raise ValueError("point_data['GlobalPointId'] must have one id per point")
    if np.array_equal(gid, np.arange(mesh.n_points)):
        return mesh.cells, mesh.n_points
    return gid[mesh.cells].astype(np.int32), int(gid.max()) + 1


def _fraction(region: Region) -> np.ndarray:
    f = region.cell_data.get(FRACTION.vtk)
    if f is None:
        return np.ones(region.n_cells)
    return np.asarray(f, np.float64)


class FemPotential:
    """Base class: one material law over one tet region (a WarpPotentialFem)."""

    material: int = -1
    material_names: tuple[str, ...] = ()

    def __init__(
        self,
        mesh: DeviceMesh,
        materials: dict[str, torch.Tensor],
        *,
        name: str | None = None,
        requires_grad: Sequence[str] = (),
        region: Region | None = None,
    ):
        self.mesh = mesh
        self.materials = materials
        self.name = name if name is not None else default_name(self)
        self.requires_grad = tuple(requires_grad)
        self.region = region
        self.cells_host = region.cells if region is not None else None
        self._energy = torch.zeros(1, dtype=torch.float64, device=mesh.device)

    # ------------------------------------------------------------ construction

    @classmethod
    def from_region(
        cls,
        region: Region,
        *,
        requires_grad: Sequence[str] = (),
        name: str | None = None,
        dtype=torch.float32,
        device=None,
        chunks: int = 1,
    ):
        """Upload the region (dhdX -> inverse rest shapes, dV * Fraction) -- _base.py:94-111."""
        dtype = as_dtype(dtype)
        dV = _fraction(region) * region.dV
        cells, n_points = _global_cells(region.mesh)
        mesh = DeviceMesh(cells, region.dminv, dV, n_points, dtype=dtype, device=device, chunks=chunks)
        materials = cls.materials_from_region(region, mesh)
        pot = cls(mesh, materials, name=name, requires_grad=requires_grad, region=region)
        pot.cells_host = cells
        # the element data as uploaded (the caller may change the mesh's
        # cell data, e.g. its Fraction, before building the next potential)
        pot._host_data = (np.asarray(cells, np.int32), region.dminv, dV)
        return pot

    @classmethod
    def from_pyvista(cls, obj: Any, *, setup: str = "host", **kwargs):
        """setup="host": the region (dhdX, dV) in numpy, as the reference's
        Region; setup="device": the element data computed by a kernel from the
        rest positions (``apl_mesh_create_from_points``); no host region."""
        mesh = Mesh.from_pyvista(obj)
        if setup == "host":
            return cls.from_region(Region.from_mesh(mesh), **kwargs)
        if setup != "device":
            raise ValueError("setup must be 'host' or 'device'")
        return cls.from_points(mesh, **kwargs)

    from_mesh = from_pyvista

    @classmethod
    def from_points(
        cls,
        mesh: Mesh,
        *,
        requires_grad: Sequence[str] = (),
        name: str | None = None,
        dtype=torch.float32,
        device=None,
        chunks: int = 1,
    ):
        """Device setup: the mesh's rest positions and cells go to
        ``apl_mesh_create_from_points``; materials come from its cell data."""
        dtype = as_dtype(dtype)
        f = mesh.cell_data.get(FRACTION.vtk)
        cells, n_points = _global_cells(mesh)
        points = mesh.points
        if n_points != mesh.n_points or cells is not mesh.cells:
            # rest positions in the global numbering (rows no tet touches stay 0)
            points = np.zeros((n_points, 3))
            points[np.asarray(mesh.point_data[GLOBAL_POINT_ID.vtk], np.int64).reshape(-1)] = mesh.points
        dmesh = DeviceMesh(
            cells, None, None, n_points, dtype=dtype, device=device, chunks=chunks, points=points,
            fraction=None if f is None else np.asarray(f, np.float64),
        )
        materials = cls.materials_from_region(mesh, dmesh)
        pot = cls(dmesh, materials, name=name, requires_grad=requires_grad)
        pot.cells_host = np.asarray(cells)
        # what the element data were computed from (host element data are
        # derived from it on demand)
        pot._source = (mesh.points.copy(), np.asarray(mesh.cells).copy(),
                       None if f is None else np.asarray(f, np.float64).copy())
        return pot

    def host_element_data(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(cells (T,4) in the vertex numbering this potential evaluates in,
        dminv (T,3,3), dV (T,) including Fraction) as FP64 host arrays, as they
        were at construction (a DeviceModel and the FP64 shadow are built
        from them)."""
        hd = getattr(self, "_host_data", None)
        if hd is None:
            src = getattr(self, "_source", None)
            if src is None:
                raise ValueError("no host geometry for this potential")
            points, cells, frac = src
            r = Region.from_mesh(Mesh(points, cells))
            dV = r.dV if frac is None else frac * r.dV
            hd = (np.asarray(self.cells_host, np.int32), r.dminv, dV)
            self._host_data = hd
        return hd

    def fp64_shadow(self) -> "FemPotential":
        """The same tets with FP64 element data (from the host data, not from
        this potential's rounded device data), for the FP64 residual passes of
        mixed-precision refinement; built once.  Its materials are this
        potential's at build time, widened: callers pass current materials
        explicitly."""
        if self.dtype == torch.float64:
            return self
        sh = getattr(self, "_shadow", None)
        if sh is None:
            cells, dminv, dV = self.host_element_data()
            mesh = DeviceMesh(cells, dminv, dV, self.n_points, dtype=torch.float64, device=self.device)
            mats = {k: v.double().contiguous() for k, v in self.materials.items()}
            sh = type(self)(mesh, mats, name=f"{self.name}_fp64")
            sh._host_data = self._host_data
            self._shadow = sh
        return sh

    @classmethod
    def materials_from_region(cls, region: Region, mesh: DeviceMesh) -> dict[str, torch.Tensor]:
        raise NotImplementedError

    @staticmethod
    def _cell_tensor(region: Region, key: str, mesh: DeviceMesh) -> torch.Tensor:
        if key not in region.cell_data:
            raise KeyError(f"cell_data[{key!r}] is required")
        arr = np.asarray(region.cell_data[key], np.float64).reshape(-1)
        if arr.shape[0] != region.n_cells:
            raise ValueError(f"cell_data[{key!r}] must have one value per cell")
        return torch.as_tensor(arr, dtype=mesh.dtype, device=mesh.device).contiguous()

    def set_materials(self, **values: torch.Tensor) -> None:
        for k, v in values.items():
            if k not in self.material_names:
                raise KeyError(f"{type(self).__name__} has no material {k!r}")
            v = torch.as_tensor(v, dtype=self.mesh.dtype, device=self.mesh.device).reshape(-1).contiguous()
            if v.shape[0] != self.mesh.n_tets:
                raise ValueError(f"{k} must have one value per tet")
            self.materials[k] = v

    # ------------------------------------------------------------ properties

    @property
    def n_points(self) -> int:
        return self.mesh.n_verts

    @property
    def n_cells(self) -> int:
        return self.mesh.n_tets

    @property
    def dtype(self) -> torch.dtype:
        return self.mesh.dtype

    @property
    def device(self) -> torch.device:
        return self.mesh.device

    @property
    def launch_dim(self) -> tuple[int, int]:
        return (self.mesh.n_tets, 1)

    def _mu(self):
        return self.materials["mu"]

    def _lam(self):
        return self.materials.get("lambda")

    def _prepare(self) -> None:
        """Hook run before every element pass (binds per-mesh material arrays)."""

    def _eval(self, u, p=None, **outputs) -> None:
        self._prepare()
        self.mesh.eval(self.material, self._mu(), self._lam(), u, p, **outputs)

    # ------------------------------------------------------------ reference protocol

    def _add_scalar(self, output: torch.Tensor, which: str, u, p=None) -> None:
        if output.dtype == torch.float64 and output.is_contiguous() and output.numel() == 1:
            kw = {which: output}
            self._eval(u, p, accumulate=True, **kw)
            return
        self._energy.zero_()
        kw = {which: self._energy}
        self._eval(u, p, accumulate=True, **kw)
        output += self._energy.to(output.dtype)

    def fun(self, u: torch.Tensor, output: torch.Tensor) -> None:
        self._add_scalar(output, "energy", u)

    def grad(self, u: torch.Tensor, output: torch.Tensor) -> None:
        self._eval(u, grad=output, accumulate=True)

    def hess_diag(self, u: torch.Tensor, output: torch.Tensor) -> None:
        self._eval(u, hess_diag=output, accumulate=True)

    def hess_prod(self, u: torch.Tensor, p: torch.Tensor, output: torch.Tensor) -> None:
        self._eval(u, p, hess_prod=output, accumulate=True)

    def hess_quad(self, u: torch.Tensor, p: torch.Tensor, output: torch.Tensor) -> None:
        self._add_scalar(output, "hess_quad", u, p)

    # ------------------------------------------------------------ fused / adjoint

    def fused(self, u, energy=None, grad=None, hess_diag=None, *, accumulate=True) -> None:
        """Energy (float64 (1,)), gradient and Hessian diagonal in one element pass."""
        self._eval(
            u, energy=energy, grad=grad, hess_diag=hess_diag, accumulate=accumulate,
        )

    def mixed_derivative_prod(self, u: torch.Tensor, p: torch.Tensor) -> dict[str, torch.Tensor]:
        """Per-tet p^T d(grad)/d(mu) (and d/d(lambda) for SNH)."""
        dmu = torch.empty(self.n_cells, dtype=self.dtype, device=self.device)
        dlam = torch.empty_like(dmu) if self.material != _lib.APL_ARAP else None
        self._eval(u, p, dmu=dmu, dlam=dlam)
        out = {"mu": dmu}
        if dlam is not None:
            out["lambda"] = dlam
        return out


WarpPotentialFem = FemPotential


class StableNeoHookean(FemPotential):
    """Stable Neo-Hookean (warp/fem/_stable_neo_hookean.py:17-154):
    Psi = mu/2 (|F|^2 - 3) - mu (J - 1) + lam/2 (J - 1)^2."""

    material = _lib.APL_SNH
    material_names = ("mu", "lambda")

    @classmethod
    def materials_from_region(cls, region, mesh):
        return {
            "lambda": cls._cell_tensor(region, LAMBDA.vtk, mesh),
            "mu": cls._cell_tensor(region, MU.vtk, mesh),
        }


class Arap(FemPotential):
    """As-rigid-as-possible (warp/fem/_arap.py:17-125): Psi = mu/2 |F - R|^2."""

    material = _lib.APL_ARAP
    material_names = ("mu",)

    @classmethod
    def materials_from_region(cls, region, mesh):
        return {"mu": cls._cell_tensor(region, MU.vtk, mesh)}


class StableNeoHookeanMuscle(StableNeoHookean):
    """Stable Neo-Hookean on G = F A (warp/fem/_stable_neo_hookean_muscle.py:18-171),
    A = I + [[a0, a3, a4], [a3, a1, a5], [a4, a5, a2]] from the per-tet
    activation (warp/fem/func/_misc.py:48-53); materials activation (T, 6),
    lambda, mu (Materials, :120-133)."""

    material = _lib.APL_SNH_MUSCLE
    material_names = ("activation", "lambda", "mu")

    @classmethod
    def materials_from_region(cls, region, mesh):
        # utils/_material.py:15-16 reads cell_data[ACTIVATION] (the enum value)
        key = ACTIVATION.value
        if key not in region.cell_data:
            raise KeyError(f"cell_data[{key!r}] is required")
        act = np.asarray(region.cell_data[key], np.float64).reshape(region.n_cells, 6)
        out = super().materials_from_region(region, mesh)
        out["activation"] = torch.as_tensor(act, dtype=mesh.dtype, device=mesh.device).contiguous()
        return out

    def set_materials(self, **values: torch.Tensor) -> None:
        act = values.pop("activation", None)
        if act is not None:
            act = torch.as_tensor(act, dtype=self.mesh.dtype, device=self.mesh.device).reshape(-1, 6).contiguous()
            if act.shape[0] != self.mesh.n_tets:
                raise ValueError("activation must have one row of 6 per tet")
            self.materials["activation"] = act
        super().set_materials(**values)

    def _prepare(self) -> None:
        self.mesh.set_activation(self.materials["activation"])

    def mixed_derivative_prod(self, u: torch.Tensor, p: torch.Tensor) -> dict[str, torch.Tensor]:
        """Per-tet p^T d(grad)/d(mu), d/d(lambda) and d/d(activation) (T, 6):
        the mixed term of the activation inverse
        (exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:472-486)."""
        self._prepare()
        dmu = torch.empty(self.n_cells, dtype=self.dtype, device=self.device)
        dlam = torch.empty_like(dmu)
        dact = torch.empty(self.n_cells, 6, dtype=self.dtype, device=self.device)
        self.mesh.mixed_activation(self._mu(), self._lam(), u, p, dact, dmu=dmu, dlam=dlam)
        return {"mu": dmu, "lambda": dlam, "activation": dact}
