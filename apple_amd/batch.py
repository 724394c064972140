"""Batched solves: many independent meshes of uneven sizes in one launch.

The instances are concatenated into one device mesh whose element blocks
never straddle an instance, so every launch of the PNCG (element passes,
vector passes, per-instance scalar steps) covers the whole batch; an
instance that has converged drops out of the following launches through its
status flag.  Each instance keeps its own materials, loads, boundary
conditions and per-instance FP64 solver scalars.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import _lib
from .mesh import Region
from .runtime import DeviceMesh, as_dtype, normalize_device
from .solver import DeviceSolver, PncgConfig, SolveStats, mask_bits


@dataclasses.dataclass
class Instance:
    region: Region
    mu: np.ndarray
    lam: np.ndarray | None = None
    f_ext: np.ndarray | None = None  # (V,3)
    fixed_mask: np.ndarray | None = None  # (V,3) bool
    x0: np.ndarray | None = None  # (V,3) initial guess incl. fixed values
    activation: np.ndarray | None = None  # (T, 6), muscle law


class _BatchPotential:
    """Minimal potential view (materials + mesh) for the device solver."""

    def __init__(self, mesh, material, mu, lam, activation=None):
        self.mesh = mesh
        self.material = material
        self.materials = {"mu": mu}
        if lam is not None:
            self.materials["lambda"] = lam
        if activation is not None:
            self.materials["activation"] = activation


class Batch:
    def __init__(self, instances: list[Instance], material: int, *, dtype=torch.float32, device=None, chunks=1,
                 schedule=None):
        """schedule: None (chunks as given, default grid), "auto" (apple_amd.autotune
        picks chunks and the element grid for this batch, cached per signature)
        or an autotune.Schedule."""
        self.instances = instances
        self.material = int(material)
        self.dtype = as_dtype(dtype)
        self.device = normalize_device(device)
        nv = [inst.region.mesh.n_points for inst in instances]
        nt = [inst.region.n_cells for inst in instances]
        self.vtx_off = np.concatenate([[0], np.cumsum(nv)]).astype(np.int64)
        self.tet_off = np.concatenate([[0], np.cumsum(nt)]).astype(np.int64)
        cells = np.concatenate([inst.region.cells + self.vtx_off[i] for i, inst in enumerate(instances)])
        dminv = np.concatenate([inst.region.dminv for inst in instances])
        dV = np.concatenate([inst.region.dV for inst in instances])
        t = lambda a: torch.as_tensor(np.concatenate(a), dtype=self.dtype, device=self.device).contiguous()
        self.mu = t([np.asarray(i.mu, np.float64) for i in instances])
        self.lam = (
            t([np.asarray(i.lam, np.float64) for i in instances]) if self.material != _lib.APL_ARAP else None
        )
        self.activation = (
            t([np.asarray(i.activation, np.float64).reshape(-1, 6) for i in instances])
            if self.material == _lib.APL_SNH_MUSCLE else None
        )
        if isinstance(schedule, str) and schedule == "auto":
            from . import autotune

            schedule = autotune.tune(cells, dminv, dV, int(self.vtx_off[-1]), material=self.material, mu=self.mu,
                                     lam=self.lam, activation=self.activation, dtype=self.dtype, device=self.device,
                                     inst_tet_off=self.tet_off, inst_vtx_off=self.vtx_off)
        self.schedule = schedule
        if schedule is not None:
            chunks = schedule.chunks
        self.mesh = DeviceMesh(cells, dminv, dV, int(self.vtx_off[-1]), dtype=self.dtype, device=self.device,
                               inst_tet_off=self.tet_off, inst_vtx_off=self.vtx_off, chunks=chunks)
        if schedule is not None:
            self.mesh.set_schedule(schedule.grid_pct)
        V = int(self.vtx_off[-1])
        f = np.zeros((V, 3))
        fm = np.zeros((V, 3), bool)
        x0 = np.zeros((V, 3))
        for i, inst in enumerate(instances):
            sl = slice(self.vtx_off[i], self.vtx_off[i + 1])
            if inst.f_ext is not None:
                f[sl] = inst.f_ext
            if inst.fixed_mask is not None:
                fm[sl] = inst.fixed_mask
            if inst.x0 is not None:
                x0[sl] = inst.x0
        self.f_ext = torch.as_tensor(f, dtype=self.dtype, device=self.device).contiguous()
        self.mask = mask_bits(fm, self.device)
        self.x0 = torch.as_tensor(x0, dtype=self.dtype, device=self.device).contiguous()
        self.potential = _BatchPotential(self.mesh, self.material, self.mu, self.lam, self.activation)
        self._solver = None

    @property
    def n_inst(self) -> int:
        return len(self.instances)

    @property
    def n_tets(self) -> int:
        return int(self.tet_off[-1])

    def solver(self) -> DeviceSolver:
        if self._solver is None:
            self._solver = DeviceSolver(self.potential)
        return self._solver

    def solve(self, x: torch.Tensor | None = None, config: PncgConfig | None = None, *,
              refine_fp64: bool = False) -> tuple[torch.Tensor, list[SolveStats]]:
        """PNCG over every instance.  refine_fp64 (FP32 batches): instances
        that end in STAGNATION or MAX_STEPS_REACHED -- the FP32 rounding floor
        of soft, heavily loaded instances -- are solved again in FP64 from
        their FP32 result, to the FP32 run's absolute target rtol * |g_0|
        (atol), and their positions and stats replaced.  Instances that met
        rtol on their FP32 gradient are kept as they are."""
        config = config or PncgConfig()
        x = self.x0.clone() if x is None else x
        stats = self.solver().solve(x, self.f_ext, self.mask, config)
        if refine_fp64 and self.dtype == torch.float32:
            from .solver import Result

            redo = [i for i, st in enumerate(stats) if st.result in (Result.STAGNATION, Result.MAX_STEPS_REACHED)]
            if redo:
                parts = self.split(x)
                sub = [dataclasses.replace(self.instances[i], x0=parts[i].double().cpu().numpy()) for i in redo]
                b64 = Batch(sub, self.material, dtype=torch.float64, device=self.device)
                atol = config.rtol * min(stats[i].grad_norm_first for i in redo)
                cfg64 = dataclasses.replace(config, atol=max(config.atol, atol), wolfe_eps=None)
                x64, st64 = b64.solve(config=cfg64)
                self.refine_stats = st64  # the FP64 re-solves as run (their own first gradient)
                for k, i in enumerate(redo):
                    parts[i].copy_(b64.split(x64)[k].to(self.dtype))
                    stats[i] = dataclasses.replace(st64[k], grad_norm_first=stats[i].grad_norm_first)
                self.refined = redo
        return x, stats

    def split(self, x: torch.Tensor) -> list[torch.Tensor]:
        return [x[self.vtx_off[i] : self.vtx_off[i + 1]] for i in range(self.n_inst)]
