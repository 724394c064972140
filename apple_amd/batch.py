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
from .solver import DeviceSolver, PncgConfig, Result, SolveStats, mask_bits


@dataclasses.dataclass
class Part:
    """One FEM potential of a model instance: its law, element data (cells in
    the instance's vertex numbering, dminv (T,3,3), dV (T,) with Fraction)
    and per-tet materials."""

    material: int
    cells: np.ndarray
    dminv: np.ndarray
    dV: np.ndarray
    mu: np.ndarray
    lam: np.ndarray | None = None
    activation: np.ndarray | None = None


@dataclasses.dataclass
class Instance:
    region: Region | None
    mu: np.ndarray | None = None
    lam: np.ndarray | None = None
    f_ext: np.ndarray | None = None  # (V,3)
    fixed_mask: np.ndarray | None = None  # (V,3) bool
    x0: np.ndarray | None = None  # (V,3) initial guess incl. fixed values
    activation: np.ndarray | None = None  # (T, 6), muscle law
    # a model instance (material APL_MODEL): several potentials over n_points vertices
    parts: list[Part] | None = None
    n_points: int | None = None

    @property
    def n_verts(self) -> int:
        return int(self.n_points) if self.parts is not None else self.region.mesh.n_points

    @property
    def n_tets(self) -> int:
        return sum(p.cells.shape[0] for p in self.parts) if self.parts is not None else self.region.n_cells


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
        nv = [inst.n_verts for inst in instances]
        nt = [inst.n_tets for inst in instances]
        self.vtx_off = np.concatenate([[0], np.cumsum(nv)]).astype(np.int64)
        self.tet_off = np.concatenate([[0], np.cumsum(nt)]).astype(np.int64)
        t = lambda a: torch.as_tensor(np.concatenate(a), dtype=self.dtype, device=self.device).contiguous()
        segments = None
        if self.material == _lib.APL_MODEL:
            cells, dminv, dV, mats, segments = self._model_arrays(instances)
            self.mu = t(mats["mu"])
            self.lam = t(mats["lambda"]) if "lambda" in mats else None
            self.activation = t(mats["activation"]) if "activation" in mats else None
            if isinstance(schedule, str):
                schedule = None  # the autotuner times single-law passes
        else:
            cells = np.concatenate([inst.region.cells + self.vtx_off[i] for i, inst in enumerate(instances)])
            dminv = np.concatenate([inst.region.dminv for inst in instances])
            dV = np.concatenate([inst.region.dV for inst in instances])
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
                               inst_tet_off=self.tet_off, inst_vtx_off=self.vtx_off, chunks=chunks,
                               segments=segments)
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

    def _model_arrays(self, instances):
        """Element data and materials of model instances in [instance][part]
        tet order, and the segments (one law per part, the same sequence in
        every instance)."""
        laws = [p.material for p in instances[0].parts]
        for inst in instances:
            if inst.parts is None or [p.material for p in inst.parts] != laws:
                raise ValueError("every model instance needs parts with the same laws in the same order")
        cells, dminv, dV = [], [], []
        mats = {"mu": []}
        if {_lib.APL_SNH, _lib.APL_SNH_MUSCLE} & set(laws):
            mats["lambda"] = []
        if _lib.APL_SNH_MUSCLE in laws:
            mats["activation"] = []
        seg_off = [0]
        for i, inst in enumerate(instances):
            for p in inst.parts:
                T = p.cells.shape[0]
                cells.append(np.asarray(p.cells, np.int64) + self.vtx_off[i])
                dminv.append(np.asarray(p.dminv, np.float64).reshape(T, 3, 3))
                dV.append(np.asarray(p.dV, np.float64).reshape(T))
                mats["mu"].append(np.asarray(p.mu, np.float64).reshape(T))
                if "lambda" in mats:
                    mats["lambda"].append(np.zeros(T) if p.lam is None or p.material == _lib.APL_ARAP
                                          else np.asarray(p.lam, np.float64).reshape(T))
                if "activation" in mats:
                    mats["activation"].append(np.zeros((T, 6)) if p.activation is None
                                              else np.asarray(p.activation, np.float64).reshape(T, 6))
                seg_off.append(seg_off[-1] + T)
        return (np.concatenate(cells).astype(np.int32), np.concatenate(dminv), np.concatenate(dV), mats,
                (np.array(laws, np.int32), np.array(seg_off, np.int64)))

    @classmethod
    def from_models(cls, models, *, dtype=torch.float32, device=None, chunks=1) -> "Batch":
        """A batch of forward.Model instances (FEM potentials + external forces,
        each model with the same sequence of material laws), solved in one
        set of launches: each model's potentials become its parts, its
        ExternalForce sum the load, its DofMap the fixed mask and start."""
        from .fem import FemPotential

        insts = []
        for model in models:
            parts = []
            for pot in model.potentials.values():
                if isinstance(pot, FemPotential):
                    c, dm, dv = pot.host_element_data()
                    mt = {k: v.double().cpu().numpy() for k, v in pot.materials.items()}
                    parts.append(Part(int(pot.material), c, dm, dv, mt["mu"], mt.get("lambda"), mt.get("activation")))
            f = model.external_force()
            x0 = model.init().u.double().cpu().numpy()
            insts.append(Instance(None, f_ext=None if f is None else f.double().cpu().numpy(),
                                  fixed_mask=model.dof_map.fixed_mask, x0=x0, parts=parts, n_points=model.n_points))
        return cls(insts, _lib.APL_MODEL, dtype=dtype, device=device, chunks=chunks)

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
              refine_fp64: bool | str = False, refine_rtol: float | None = None,
              refine_cg_rtol: float = 1e-3, refine_max_steps: int = 8) -> tuple[torch.Tensor, list[SolveStats]]:
        """PNCG over every instance.  For FP32 batches, refine_fp64 improves the
        FP32 answer with FP64 arithmetic where the FP32 rounding of the element
        forces limits it:

        * "newton": mixed-precision iterative refinement (apple_amd.refine):
          the free-DOF residual f - grad E(x) of every instance is evaluated
          in FP64 on an FP64 copy of the batch, and Newton corrections
          H(x) d = r are solved by the FP32 PCG, until the FP64 residual is
          <= refine_rtol * |g_0| (default: config.rtol).  FP32 kernels do all
          iterations; the FP64 passes run once per refinement step.
          Instances whose correction meets negative curvature (not at a
          stable equilibrium) or that miss the target after
          refine_max_steps are re-solved by the FP64 PNCG on the copy.
          ``self.x64`` holds the FP64 positions, ``self.refine_report`` the
          per-instance residuals.

        The legacy modes re-solve whole instances with the FP64 PNCG, from
        their FP32 result and to the FP32 run's absolute target
        rtol * |g_0| (atol), when the FP32 answer is not trusted:

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
        self.x64 = None
        if not refine_fp64 or self.dtype != torch.float32:
            return x, stats
        if refine_fp64 not in (True, "stagnated", "checked", "newton"):
            raise ValueError("refine_fp64 must be False, 'newton', True / 'stagnated' or 'checked'")
        if refine_fp64 == "newton":
            return self._refine_newton(x, stats, config, refine_rtol if refine_rtol is not None else config.rtol,
                                       refine_cg_rtol, refine_max_steps)
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
            sub = [dataclasses.replace(self._current_instance(i), x0=parts[i].double().cpu().numpy()) for i in redo]
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

    def _current_instance(self, i: int) -> Instance:
        """Instance i with the batch's current materials in FP64: its own
        values where the batch still holds their rounding, else the current
        values widened (adjoint.equilibrium and Inverse replace them)."""
        inst = self.instances[i]
        t0, t1 = int(self.tet_off[i]), int(self.tet_off[i + 1])
        np_dt = np.float32 if self.dtype == torch.float32 else np.float64

        def current(key, orig):
            c64 = self.potential.materials[key][t0:t1].double().cpu().numpy().reshape(orig.shape)
            return orig if np.array_equal(c64, orig.astype(np_dt).astype(np.float64)) else c64

        if inst.parts is None:
            upd = {"mu": current("mu", np.asarray(inst.mu, np.float64).reshape(-1))}
            if "lambda" in self.potential.materials and inst.lam is not None:
                upd["lam"] = current("lambda", np.asarray(inst.lam, np.float64).reshape(-1))
            if "activation" in self.potential.materials and inst.activation is not None:
                upd["activation"] = current("activation", np.asarray(inst.activation, np.float64).reshape(-1, 6))
            return dataclasses.replace(inst, **upd)
        # a model instance: the materials of its parts, in [part] tet order
        sizes = [p.cells.shape[0] for p in inst.parts]
        cuts = np.cumsum(sizes)[:-1]
        mats = self.potential.materials
        mu = current("mu", np.concatenate([np.asarray(p.mu, np.float64).reshape(-1) for p in inst.parts]))
        lam = act = None
        if "lambda" in mats:
            lam = current("lambda", np.concatenate([
                np.zeros(n) if p.lam is None or p.material == _lib.APL_ARAP else np.asarray(p.lam, np.float64).reshape(n)
                for p, n in zip(inst.parts, sizes)]))
        if "activation" in mats:
            act = current("activation", np.concatenate([
                np.zeros((n, 6)) if p.activation is None else np.asarray(p.activation, np.float64).reshape(n, 6)
                for p, n in zip(inst.parts, sizes)]))
        parts = []
        for k, (p, m_k) in enumerate(zip(inst.parts, np.split(mu, cuts))):
            upd = {"mu": m_k}
            if lam is not None and p.material != _lib.APL_ARAP:
                upd["lam"] = np.split(lam, cuts)[k]
            if act is not None and p.material == _lib.APL_SNH_MUSCLE:
                upd["activation"] = np.split(act, cuts)[k]
            parts.append(dataclasses.replace(p, **upd))
        return dataclasses.replace(inst, parts=parts)

    def _fp64_copy(self) -> "Batch":
        """FP64 copy of the batch (FP64 element data from the instances' regions),
        built on first use and kept.  Its materials follow the FP32 batch's
        current ones: the instances' FP64 values while the FP32 materials are
        their rounding, else the current FP32 values widened (they may have
        been replaced, e.g. by adjoint.equilibrium or Inverse)."""
        if getattr(self, "_b64", None) is None:
            self._b64 = Batch(self.instances, self.material, dtype=torch.float64, device=self.device)
            self._b64_src = {k: v.clone() for k, v in self._b64.potential.materials.items()}
        b64 = self._b64
        for k, cur in self.potential.materials.items():
            src = self._b64_src[k]
            val = src if torch.equal(cur, src.to(cur.dtype)) else cur.double()
            b64.potential.materials[k].copy_(val)
        return b64

    def _refine_newton(self, x, stats, config, rtol, cg_rtol, max_steps):
        from . import refine
        from .solver import Result

        b64 = self._fp64_copy()
        norms = refine.InstanceNorms(self.vtx_off, self.device)
        g0 = np.array([s.grad_norm_first for s in stats])
        target = rtol * g0 + config.atol
        mats = b64.potential.materials
        x64, rep = refine.refine_equilibrium(self.solver(), b64.potential, x, b64.f_ext, self.mask, target, norms,
                                             max_steps=max_steps, cg_rtol=cg_rtol, shadow_materials=mats)
        redo = sorted(set(rep.failed) | {i for i in range(self.n_inst) if not rep.converged[i]})
        self.refine_report = rep
        self.refine_resolved = redo
        if redo:
            # FP64 PNCG on the copy from the refined positions; the others are held fixed
            keep = torch.ones(self.n_verts, dtype=torch.bool, device=self.device)
            for i in redo:
                keep[int(self.vtx_off[i]):int(self.vtx_off[i + 1])] = False
            mask = torch.where(keep, torch.full_like(b64.mask, 7), b64.mask)
            atol = float(min(target[i] for i in redo))
            cfg64 = dataclasses.replace(config, rtol=0.0, atol=atol, wolfe_eps=None)
            st64 = b64.solver().solve(x64, b64.f_ext, mask, cfg64)
            x.copy_(x64.to(x.dtype))
            for i in redo:
                stats[i] = dataclasses.replace(st64[i], grad_norm_first=stats[i].grad_norm_first)
        # final FP64 residuals and energies of every instance
        g = torch.zeros_like(x64)
        e = torch.zeros(self.n_inst, dtype=torch.float64, device=self.device)
        b64.mesh.eval(b64.material, mats["mu"], mats.get("lambda"), x64, energy=e, grad=g)
        r = torch.where(refine._free(self.mask, g), b64.f_ext - g, torch.zeros_like(g))
        rn = norms(r).cpu().numpy()
        e = e.cpu().numpy()
        ext = (b64.f_ext * x64).sum(1)
        fx = torch.zeros(self.n_inst, dtype=torch.float64, device=self.device).index_add_(
            0, norms.inst_of_vertex, ext).cpu().numpy()
        for i in range(self.n_inst):
            if rn[i] <= target[i]:
                res = Result.PRIMARY_SUCCESS
            elif stats[i].result == Result.PRIMARY_SUCCESS:
                res = Result.STAGNATION  # the FP64 residual misses the target the FP32 one met
            else:
                res = stats[i].result
            # energy E(x) - f.x and the residual norm, both in FP64
            stats[i] = dataclasses.replace(stats[i], result=res, f=float(e[i] - fx[i]), grad_norm=float(rn[i]))
        self.refined = redo
        self.x64 = x64
        return x, stats

    def adjoint_solve(self, x: torch.Tensor, rhs: torch.Tensor, *, rtol: float = 1e-6, max_iters: int = 20000,
                      refine_fp64: bool = False, cg_rtol: float | None = None, refine_cg_rtol: float = 1e-3,
                      refine_max_steps: int = 8, x64: torch.Tensor | None = None):
        """Adjoint solve H(x) p = rhs of every instance (free DOFs), by the
        device PCG on the Hessian-vector pass, to ||rhs - H p|| <= rtol ||rhs||.
        refine_fp64 (FP32 batches): the PCG runs to cg_rtol (default
        max(rtol, 1e-6), above the FP32 floor), then iterative refinement with
        the FP64 Hessian-vector residual of the FP64 copy at x64 (default:
        self.x64 from the last refined solve, else x widened) brings the FP64
        residual to rtol.  rhs may be FP64 (the refinement then uses it
        unrounded).  Returns (p, stats); ``self.adjoint_report`` holds the
        refinement's per-instance residuals."""
        p = torch.zeros_like(x)
        first = rtol if not (refine_fp64 and self.dtype == torch.float32) else (
            cg_rtol if cg_rtol is not None else max(rtol, 1e-6))
        rhs_r = rhs.to(x.dtype).contiguous()
        st = self.solver().cg(x, rhs_r, p, self.mask, rtol=first, max_iters=max_iters)
        indefinite = [i for i, s in enumerate(st) if s.result == Result.NON_FINITE]
        if indefinite:
            # negative curvature: CG on the normal equations for those instances
            from .adjoint import AdjointConfig, _normal_fallback

            st = _normal_fallback(self.solver(), x, rhs_r, p, self.mask, AdjointConfig(rtol=first, max_iters=max_iters),
                                  st, indefinite)
        if not refine_fp64 or self.dtype != torch.float32:
            return p, st
        from . import refine

        b64 = self._fp64_copy()
        x64 = x64 if x64 is not None else (self.x64 if getattr(self, "x64", None) is not None else x.double())
        norms = refine.InstanceNorms(self.vtx_off, self.device)
        _, rep = refine.refine_linear(self.solver(), b64.potential, x64, rhs, p, self.mask, rtol, norms,
                                      max_steps=refine_max_steps, cg_rtol=refine_cg_rtol, x32=x,
                                      shadow_materials=b64.potential.materials)
        for i in range(self.n_inst):
            if rep.converged[i]:
                st[i] = dataclasses.replace(st[i], result=Result.PRIMARY_SUCCESS, f=float(rep.residual[i] /
                                            max(rep.target[i] / rtol, 1e-300)))
        self.adjoint_report = rep
        return p, st

    @property
    def n_verts(self) -> int:
        return int(self.vtx_off[-1])

    def split(self, x: torch.Tensor) -> list[torch.Tensor]:
        return [x[self.vtx_off[i] : self.vtx_off[i + 1]] for i in range(self.n_inst)]
