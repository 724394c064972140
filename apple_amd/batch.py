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
              refine_fp64: bool | str = False) -> tuple[torch.Tensor, list[SolveStats]]:
        """PNCG over every instance.  For FP32 batches, refine_fp64 re-solves in
        FP64, from their FP32 result and to the FP32 run's absolute target
        rtol * |g_0| (atol), the instances whose FP32 answer is not trusted:

        * True / "stagnated": those ending in STAGNATION or MAX_STEPS_REACHED
          (the FP32 rounding floor of soft, heavily loaded instances);
        * "checked": in addition every instance whose free-DOF gradient,
          evaluated once in FP64 at its FP32 solution, exceeds that target
          (an FP32 gradient can meet rtol while the true one does not).

        "stagnated" re-solves a batch of just those instances; "checked"
        builds an FP64 copy of the whole batch on first use (kept for later
        calls) and re-solves on it with the other instances held fixed.
        Positions and stats of re-solved instances are replaced (stats keep
        the FP32 run's first gradient norm); ``self.refined`` lists them."""
        from .solver import Result

        config = config or PncgConfig()
        x = self.x0.clone() if x is None else x
        stats = self.solver().solve(x, self.f_ext, self.mask, config)
        self.refined = []
        if not refine_fp64 or self.dtype != torch.float32:
            return x, stats
        if refine_fp64 not in (True, "stagnated", "checked"):
            raise ValueError("refine_fp64 must be False, True / 'stagnated' or 'checked'")
        redo = {i for i, st in enumerate(stats) if st.result in (Result.STAGNATION, Result.MAX_STEPS_REACHED)}
        if refine_fp64 == "checked":
            b64 = self._fp64_copy()
            x64 = x.double()
            g = torch.zeros_like(x64)
            if b64.activation is not None:
                b64.mesh.set_activation(b64.activation)
            b64.mesh.eval(b64.material, b64.mu, b64.lam, x64, grad=g)
            g -= b64.f_ext
            free = torch.stack([(b64.mask >> c) & 1 == 0 for c in range(3)], dim=1)
            g = torch.where(free, g, torch.zeros_like(g))
            for i in range(self.n_inst):
                sl = slice(int(self.vtx_off[i]), int(self.vtx_off[i + 1]))
                gn = float(torch.linalg.norm(g[sl]))
                if gn > config.rtol * stats[i].grad_norm_first + config.atol:
                    redo.add(i)
        redo = sorted(redo)
        if redo and refine_fp64 != "checked":
            # a batch of just these instances, from their FP32 positions
            parts = self.split(x)
            sub = [dataclasses.replace(self.instances[i], x0=parts[i].double().cpu().numpy()) for i in redo]
            bs = Batch(sub, self.material, dtype=torch.float64, device=self.device)
            atol = config.rtol * min(stats[i].grad_norm_first for i in redo)
            cfg64 = dataclasses.replace(config, atol=max(config.atol, atol), wolfe_eps=None)
            xs, sts = bs.solve(config=cfg64)
            for k, i in enumerate(redo):
                parts[i].copy_(bs.split(xs)[k].to(self.dtype))
                stats[i] = dataclasses.replace(sts[k], grad_norm_first=stats[i].grad_norm_first)
            self.refine_stats = sts
        elif redo:
            # on the kept FP64 copy; the others are held fixed (all DOFs) and converge at once
            keep = torch.ones(self.n_verts, dtype=torch.bool, device=self.device)
            for i in redo:
                keep[int(self.vtx_off[i]):int(self.vtx_off[i + 1])] = False
            mask = torch.where(keep, torch.full_like(b64.mask, 7), b64.mask)
            atol = config.rtol * min(stats[i].grad_norm_first for i in redo)
            cfg64 = dataclasses.replace(config, atol=max(config.atol, atol), wolfe_eps=None)
            st64 = b64.solver().solve(x64, b64.f_ext, mask, cfg64)
            parts, parts64 = self.split(x), b64.split(x64)
            for i in redo:
                parts[i].copy_(parts64[i].to(self.dtype))
                stats[i] = dataclasses.replace(st64[i], grad_norm_first=stats[i].grad_norm_first)
            self.refine_stats = [st64[i] for i in redo]
        self.refined = redo
        return x, stats

    def _fp64_copy(self) -> "Batch":
        if getattr(self, "_b64", None) is None:
            self._b64 = Batch(self.instances, self.material, dtype=torch.float64, device=self.device)
        return self._b64

    @property
    def n_verts(self) -> int:
        return int(self.vtx_off[-1])

    def split(self, x: torch.Tensor) -> list[torch.Tensor]:
        return [x[self.vtx_off[i] : self.vtx_off[i + 1]] for i in range(self.n_inst)]
