"""Differentiable static equilibrium: x*(mu, lambda[, activation]) with an adjoint backward.

``equilibrium(potential, mu, lam, f_ext, mask, x0)`` returns the PNCG
equilibrium x* as a tensor that autograd can differentiate with respect to
the per-tet materials (for the muscle law also the (T, 6) activation).  The backward pass follows the adjoint convention of
liblaf/apple's inverse experiments (exp/2025/08/27/inverse/src/30-inverse.py:85-97,
exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:472-510):

    solve   H(x*) p = -dL/dx        (free DOFs; PCG on the Hessian-vector pass)
    dL/dq = p^T d grad / d q         (per tet, the mixed-derivative pass)

All three stages run on the GPU: the forward PNCG, the PCG adjoint solve
(apl_cg_solve) and the mixed derivatives (the APL_MIXED element pass, or
apl_elem_mixed_activation for the muscle law).  The
same code serves a batch: with a batched potential every instance's forward
and adjoint solve share the launches.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import _lib
from .solver import DeviceSolver, PncgConfig, Result


@dataclasses.dataclass
class AdjointConfig:
    """rtol / max_iters of the adjoint PCG; normal_fallback: re-solve an
    instance whose Hessian shows negative curvature by CG on the normal
    equations (apl_cg_solve_normal) instead of failing."""

    rtol: float = 1e-8
    max_iters: int = 20000
    normal_fallback: bool = True


@dataclasses.dataclass
class RefineConfig:
    """Mixed-precision refinement of an FP32 potential's equilibrium and
    adjoint (apple_amd.refine): after the FP32 PNCG (to config.rtol) the
    forward residual is driven to rtol * |g_0| with FP64 residual passes on
    the potential's FP64 shadow and FP32 PCG corrections, and the output is
    the FP64 position; the backward PCG runs to the AdjointConfig's rtol in
    FP32, then to adjoint_rtol with FP64 Hessian-vector residuals.
    f_ext64: the load in FP64 (default: f_ext widened)."""

    rtol: float = 1e-10
    adjoint_rtol: float = 1e-10
    cg_rtol: float = 1e-3
    max_steps: int = 8
    f_ext64: torch.Tensor | None = None


class _Equilibrium(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mu, lam, activation, solver, f_ext, mask, x0, config, adjoint, refine_cfg):
        pot = solver.potential
        pot.materials["mu"] = mu.detach().contiguous()
        if lam is not None:
            pot.materials["lambda"] = lam.detach().contiguous()
        if activation is not None:
            pot.materials["activation"] = activation.detach().contiguous()
        x = x0.detach().to(pot.mesh.dtype).clone().contiguous()
        stats = solver.solve(x, f_ext, mask, config)
        x64 = None
        refine_report = None
        if refine_cfg is not None and pot.mesh.dtype == torch.float32:
            from . import refine

            shadow = pot.fp64_shadow()
            norms = refine.InstanceNorms(pot.mesh.inst_vtx_off, x.device)
            f64 = refine_cfg.f_ext64 if refine_cfg.f_ext64 is not None else (
                f_ext.double() if f_ext is not None else None)
            target = np.array([refine_cfg.rtol * s.grad_norm_first for s in stats]) + config.atol
            mats64 = _widened(pot)
            x64, refine_report = refine.refine_equilibrium(
                solver, shadow, x, f64, mask, target, norms, max_steps=refine_cfg.max_steps,
                cg_rtol=refine_cfg.cg_rtol, shadow_materials=mats64)
            redo = [i for i in range(len(stats)) if not refine_report.converged[i]]
            if redo:
                # corrections that do not converge: FP64 PNCG from the refined point
                shadow_mats = shadow.materials
                shadow.materials = mats64
                try:
                    st64, _ = refine.finish_fp64(shadow, x64, f64, mask, target, redo, norms, config)
                finally:
                    shadow.materials = shadow_mats
                x.copy_(x64.to(x.dtype))
                for i in redo:
                    refine_report.residual[i] = st64[i].grad_norm
                    refine_report.converged[i] = st64[i].success
            stats = [dataclasses.replace(s, grad_norm=float(refine_report.residual[i]),
                                         result=Result.PRIMARY_SUCCESS if refine_report.converged[i] else s.result)
                     for i, s in enumerate(stats)]
        ctx.solver = solver
        ctx.mask = mask
        ctx.adjoint = adjoint
        ctx.refine = refine_cfg if x64 is not None else None
        ctx.stats = stats
        ctx.has_lam = lam is not None
        ctx.has_act = activation is not None
        ctx.save_for_backward(
            x, x64 if x64 is not None else x, mu.detach(), lam.detach() if lam is not None else mu.detach(),
            activation.detach() if activation is not None else mu.detach(),
        )
        _Equilibrium.last_forward_stats = stats
        _Equilibrium.last_refine_report = refine_report
        return x64 if x64 is not None else x

    @staticmethod
    def backward(ctx, grad_x):
        x, x64, mu, lam, act = ctx.saved_tensors
        solver = ctx.solver
        pot = solver.potential
        mu = mu.contiguous()
        lam = lam.contiguous() if ctx.has_lam else None
        act = act.contiguous() if ctx.has_act else None
        pot.materials["mu"] = mu
        if lam is not None:
            pot.materials["lambda"] = lam
        if act is not None:
            pot.materials["activation"] = act
        rhs = (-grad_x).to(x.dtype).contiguous()
        p = torch.zeros_like(x)
        stats = solver.cg(x, rhs, p, ctx.mask, rtol=ctx.adjoint.rtol, max_iters=ctx.adjoint.max_iters)
        indefinite = [i for i, s in enumerate(stats) if s.result == Result.NON_FINITE]
        if indefinite and ctx.adjoint.normal_fallback:
            # negative curvature met: CG on the normal equations H^2 p = H rhs
            # for those instances (lineax.NormalCG, the reference's adjoint
            # solver), the others' right-hand sides zeroed
            stats = _normal_fallback(solver, x, rhs, p, ctx.mask, ctx.adjoint, stats, indefinite)
        _Equilibrium.last_adjoint_stats = stats
        bad = [s for s in stats if s.result not in (Result.PRIMARY_SUCCESS,)]
        if bad:
            raise RuntimeError(f"adjoint solve failed: {bad[0]}")
        if ctx.refine is not None:
            from . import refine

            norms = refine.InstanceNorms(pot.mesh.inst_vtx_off, x.device)
            _, rep = refine.refine_linear(solver, pot.fp64_shadow(), x64, (-grad_x).double(), p, ctx.mask,
                                          ctx.refine.adjoint_rtol, norms, max_steps=ctx.refine.max_steps,
                                          cg_rtol=ctx.refine.cg_rtol, x32=x, shadow_materials=_widened(pot))
            _Equilibrium.last_adjoint_refine_report = rep
        mesh = pot.mesh
        dmu = torch.empty(mesh.n_tets, dtype=mesh.dtype, device=mesh.device)
        dlam = torch.empty_like(dmu) if lam is not None else None
        dact = None
        if act is not None:
            # muscle law: activation, mu and lambda derivatives in one pass
            dact = torch.zeros(mesh.n_tets, 6, dtype=mesh.dtype, device=mesh.device)  # rows of other laws stay 0
            mesh.set_activation(act)
            mesh.mixed_activation(mu, lam, x, p, dact, dmu=dmu, dlam=dlam)
        else:
            mesh.eval(pot.material, mu, lam, x, p, dmu=dmu, dlam=dlam)
        return dmu, dlam, dact, None, None, None, None, None, None, None


def _normal_fallback(solver, x, rhs, p, mask, cfg, stats, redo):
    """Re-solve the instances in `redo` by CG on the normal equations and
    splice their solutions into p; returns the merged stats."""
    mesh = solver.potential.mesh
    keep = torch.ones(mesh.n_verts, dtype=torch.bool, device=x.device)
    off = mesh.inst_vtx_off
    for i in redo:
        keep[int(off[i]):int(off[i + 1])] = False
    rhs_n = torch.where(keep[:, None], torch.zeros_like(rhs), rhs).contiguous()
    pn = torch.zeros_like(p)
    st_n = solver.cg(x, rhs_n, pn, mask, rtol=cfg.rtol, max_iters=cfg.max_iters, normal=True)
    p.copy_(torch.where(keep[:, None], p, pn))
    out = list(stats)
    for i in redo:
        out[i] = dataclasses.replace(st_n[i])
    _Equilibrium.last_adjoint_normal = sorted(redo)
    return out


def _widened(pot) -> dict:
    """The potential's current materials in FP64 (for its FP64 shadow)."""
    return {k: v.double().contiguous() for k, v in pot.materials.items()}


def equilibrium(
    potential,
    mu: torch.Tensor,
    lam: torch.Tensor | None,
    f_ext: torch.Tensor | None,
    mask: torch.Tensor | None,
    x0: torch.Tensor,
    config: PncgConfig | None = None,
    adjoint: AdjointConfig | None = None,
    solver: DeviceSolver | None = None,
    activation: torch.Tensor | None = None,
    refine: RefineConfig | None = None,
) -> torch.Tensor:
    """Differentiable equilibrium of ``potential`` (one FEM potential or a
    DeviceModel over several) with materials (mu, lam) and, for the muscle
    law, the (T, 6) activation (for a model: per-tet arrays over all its
    tets, e.g. DeviceModel.gather of per-potential tensors).  With ``refine`` (FP32
    potentials) the result is the refined FP64 equilibrium and the adjoint is
    refined the same way (RefineConfig)."""
    if solver is None:
        solver = getattr(potential, "_adjoint_solver", None)
        if solver is None:
            solver = DeviceSolver(potential)
            potential._adjoint_solver = solver
    if potential.material == _lib.APL_ARAP:
        lam = None
    has_act = potential.material == _lib.APL_SNH_MUSCLE or (
        potential.material == _lib.APL_MODEL and "activation" in potential.materials)
    if has_act:
        if activation is None:
            activation = potential.materials["activation"]
    elif activation is not None:
        raise ValueError("activation applies to the muscle law only")
    return _Equilibrium.apply(
        mu, lam, activation, solver, f_ext, mask, x0, config or PncgConfig(), adjoint or AdjointConfig(), refine
    )


def last_stats():
    """(forward stats, adjoint stats) of the most recent forward / backward
    (process-wide: the last call of any thread)."""
    return getattr(_Equilibrium, "last_forward_stats", None), getattr(_Equilibrium, "last_adjoint_stats", None)


def last_refine_reports():
    """(forward, adjoint) refinement reports of the most recent refined
    forward / backward, or None."""
    return (getattr(_Equilibrium, "last_refine_report", None),
            getattr(_Equilibrium, "last_adjoint_refine_report", None))
