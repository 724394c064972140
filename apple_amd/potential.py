"""External loads: ``ExternalForce`` (warp/potential/_ext_force.py:42-90).

E = -sum_k f_k . u[idx_k], grad[idx_k] -= f_k, no Hessian contribution.
Gravity enters the elasticity problem this way; ``gravity`` builds the
lumped nodal load rho g sum_t dV_t / 4 for a region.
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from .common import FORCE, GLOBAL_POINT_ID, default_name
from .mesh import Mesh, Region
from .runtime import DeviceMesh, as_dtype, normalize_device


class ExternalForce:
    def __init__(self, indices, force, n_points: int, *, name: str | None = None, dtype=torch.float32, device=None):
        self.name = name if name is not None else default_name(self)
        self.requires_grad = ()
        self.dtype = as_dtype(dtype)
        self.device = normalize_device(device)
        self.n_points = int(n_points)
        # host arrays or device tensors (the checks run where the data are)
        idx = torch.as_tensor(np.asarray(indices) if not isinstance(indices, torch.Tensor) else indices)
        idx = idx.to(device=self.device, dtype=torch.int64).reshape(-1)
        if not isinstance(force, torch.Tensor):
            force = torch.as_tensor(np.asarray(force, np.float64))
        force = force.to(device=self.device).reshape(-1, 3)
        if idx.shape[0] != force.shape[0]:
            raise ValueError("one force per index")
        if idx.numel() and (int(idx.min()) < 0 or int(idx.max()) >= n_points):
            raise ValueError("force index out of range")
        if torch.unique(idx).numel() != idx.numel():
            raise ValueError("ExternalForce indices must be unique")
        self.indices = idx.to(torch.int32).contiguous()
        self.force = force.to(self.dtype).contiguous()
        self._mesh = DeviceMesh(np.zeros((0, 4), np.int32), np.zeros((0, 3, 3)), np.zeros(0), n_points,
                                dtype=self.dtype, device=self.device)
        self._energy = torch.zeros(1, dtype=torch.float64, device=self.device)

    @property
    def materials(self) -> dict[str, torch.Tensor]:
        return {"force": self.force}

    @classmethod
    def from_pyvista(cls, obj: Any, n_points: int | None = None, **kwargs) -> "ExternalForce":
        """point_data['Force'] at point_data['GlobalPointId'] (_ext_force.py:55-61)."""
        mesh = Mesh.from_pyvista(obj)
        force = np.asarray(mesh.point_data[FORCE.vtk])
        idx = np.asarray(mesh.point_data[GLOBAL_POINT_ID.vtk])
        return cls(idx, force, n_points if n_points is not None else int(idx.max()) + 1, **kwargs)

    def dense(self) -> torch.Tensor:
        """The load as a dense (V,3) tensor."""
        out = torch.zeros(self.n_points, 3, dtype=self.dtype, device=self.device)
        out.index_copy_(0, self.indices.long(), self.force)
        return out

    # reference protocol (accumulate semantics)
    def fun(self, u: torch.Tensor, output: torch.Tensor) -> None:
        if output.dtype == torch.float64 and output.numel() == 1 and output.is_contiguous():
            self._mesh.ext_force(self.indices, self.force, u, energy=output, accumulate=True)
            return
        self._energy.zero_()
        self._mesh.ext_force(self.indices, self.force, u, energy=self._energy, accumulate=True)
        output += self._energy.to(output.dtype)

    def grad(self, u: torch.Tensor, output: torch.Tensor) -> None:
        self._mesh.ext_force(self.indices, self.force, None, grad=output, accumulate=True)

    def hess_diag(self, u, output) -> None:
        pass

    def hess_prod(self, u, p, output) -> None:
        pass

    def hess_quad(self, u, p, output) -> None:
        pass


def gravity(region: Region, density, g=(0.0, 0.0, -9.8), **kwargs) -> ExternalForce:
    """Lumped gravity load f_v = rho g sum_{t ni v} dV_t / 4 on every vertex."""
    rho = np.broadcast_to(np.asarray(density, np.float64), (region.n_cells,))
    mass = np.zeros(region.mesh.n_points)
    np.add.at(mass, region.cells.reshape(-1), np.repeat(rho * region.dV / 4.0, 4))
    force = mass[:, None] * np.asarray(g, np.float64)[None, :]
    return ExternalForce(np.arange(region.mesh.n_points), force, region.mesh.n_points, **kwargs)


def gravity_from_potential(potential, density, g=(0.0, 0.0, -9.8), **kwargs) -> ExternalForce:
    """Lumped gravity load of a potential's device mesh: f_v = g sum_{t ni v}
    rho_t dV_t / 4 with the device's dV (which includes the tets' Fraction),
    summed on the device in a fixed order over the mesh's block partition
    (``apl_mesh_lumped_volume``); density is a scalar or one value per tet."""
    mesh = potential.mesh
    rho = None
    scale = 1.0
    if np.ndim(density) == 0:
        scale = float(density)
    else:
        rho = torch.as_tensor(np.asarray(density, np.float64).reshape(-1), dtype=mesh.dtype, device=mesh.device)
        if rho.shape[0] != mesh.n_tets:
            raise ValueError("density must be a scalar or one value per tet")
        rho = rho.contiguous()
    vol = mesh.lumped_volume(rho)
    if scale != 1.0:
        vol = vol * scale
    force = vol[:, None] * torch.as_tensor(np.asarray(g, np.float64), device=mesh.device)[None, :]
    idx = torch.arange(mesh.n_verts, dtype=torch.int32, device=mesh.device)
    return ExternalForce(idx, force, mesh.n_verts, dtype=mesh.dtype, device=mesh.device, **kwargs)
