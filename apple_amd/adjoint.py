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

import torch

from . import _lib
from .solver import DeviceSolver, PncgConfig, Result


@dataclasses.dataclass
class AdjointConfig:
    rtol: float = 1e-8
    max_iters: int = 20000


class _Equilibrium(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mu, lam, activation, solver, f_ext, mask, x0, config, adjoint):
        pot = solver.potential
        pot.materials["mu"] = mu.detach().contiguous()
        if lam is not None:
            pot.materials["lambda"] = lam.detach().contiguous()
        if activation is not None:
            pot.materials["activation"] = activation.detach().contiguous()
        x = x0.detach().clone().contiguous()
        stats = solver.solve(x, f_ext, mask, config)
        ctx.solver = solver
        ctx.mask = mask
        ctx.adjoint = adjoint
        ctx.stats = stats
        ctx.has_lam = lam is not None
        ctx.has_act = activation is not None
        ctx.save_for_backward(
            x, mu.detach(), lam.detach() if lam is not None else mu.detach(),
            activation.detach() if activation is not None else mu.detach(),
        )
        _Equilibrium.last_forward_stats = stats
        return x

    @staticmethod
    def backward(ctx, grad_x):
        x, mu, lam, act = ctx.saved_tensors
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
        _Equilibrium.last_adjoint_stats = stats
        bad = [s for s in stats if s.result not in (Result.PRIMARY_SUCCESS,)]
        if bad:
            raise RuntimeError(f"adjoint solve failed: {bad[0]}")
        mesh = pot.mesh
        dmu = torch.empty(mesh.n_tets, dtype=mesh.dtype, device=mesh.device)
        dlam = torch.empty_like(dmu) if lam is not None else None
        dact = None
        if act is not None:
            # muscle law: activation, mu and lambda derivatives in one pass
            dact = torch.empty(mesh.n_tets, 6, dtype=mesh.dtype, device=mesh.device)
            mesh.set_activation(act)
            mesh.mixed_activation(mu, lam, x, p, dact, dmu=dmu, dlam=dlam)
        else:
            mesh.eval(pot.material, mu, lam, x, p, dmu=dmu, dlam=dlam)
        return dmu, dlam, dact, None, None, None, None, None, None


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
) -> torch.Tensor:
    """Differentiable equilibrium of ``potential`` with materials (mu, lam) and,
    for the muscle law, the (T, 6) activation."""
    if solver is None:
        solver = getattr(potential, "_adjoint_solver", None)
        if solver is None:
            solver = DeviceSolver(potential)
            potential._adjoint_solver = solver
    if potential.material == _lib.APL_ARAP:
        lam = None
    if potential.material == _lib.APL_SNH_MUSCLE:
        if activation is None:
            activation = potential.materials["activation"]
    elif activation is not None:
        raise ValueError("activation applies to the muscle law only")
    return _Equilibrium.apply(
        mu, lam, activation, solver, f_ext, mask, x0, config or PncgConfig(), adjoint or AdjointConfig()
    )


def last_stats():
    """(forward stats, adjoint stats) of the most recent forward / backward
    (process-wide: the last call of any thread)."""
    return getattr(_Equilibrium, "last_forward_stats", None), getattr(_Equilibrium, "last_adjoint_stats", None)
