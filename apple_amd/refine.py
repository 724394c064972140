"""Mixed-precision refinement: FP64 residuals, FP32 corrections.

The FP32 element kernels round every tet's force to ~6e-8 of its size.  At
an equilibrium the forces of neighbouring tets cancel to the (much smaller)
net load, so the FP32 gradient carries an absolute error of ~6e-8 times the
internal forces.  On soft, heavily loaded instances that floor lies above
rtol ||g_0||, and along soft modes (small Hessian eigenvalues) it becomes
position errors well above 1e-5 of the bounding-box diagonal.  Solving
longer in FP32 cannot remove it: the gradient the solver sees is wrong at
that level.

Classical iterative refinement removes it without FP64 element kernels in
the loop: only the *residual* is evaluated in FP64, once per refinement step,

    r   = f_ext - grad E(x)           (one FP64 gradient pass, free DOFs)
    H(x) d = r                        (Jacobi-PCG on the FP32 Hessian-vector pass)
    x  += d                           (in FP64)

and likewise for the adjoint, r = b - H(x) y with an FP64 Hessian-vector
pass and the FP32 PCG for the correction.  Each step shrinks the residual
by about max(cg_rtol, cond(H) * 1e-7), so a few steps reach the FP64
solution as long as cond(H) * 1e-7 << 1; the FP32 kernels do all the
iterations.  The FP64 passes run on an FP64 copy of the element data
(``shadow``), so they see the FP64 geometry, not its FP32 rounding.

An instance whose correction solve meets a direction of negative curvature
(it sits away from a stable equilibrium), or whose residual a correction
does not halve (the Hessian-vector pass is a projected Hessian: ARAP's
clamped twist modes on compressed elements make these corrections converge
slowly), is reported in ``failed``; ``finish_fp64`` continues such
instances with the FP64 PNCG on the shadow.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from .solver import Result


@dataclasses.dataclass
class RefineReport:
    steps: int  # refinement steps taken (FP64 residual evaluations - 1)
    residual: np.ndarray  # final FP64 residual norm per instance
    target: np.ndarray  # per-instance target (atol)
    converged: np.ndarray  # bool per instance
    failed: list[int]  # instances whose correction solve did not succeed
    cg_iterations: int  # FP32 PCG iterations over all steps (max over instances per step, summed)


def _free(mask_bits: torch.Tensor | None, like: torch.Tensor) -> torch.Tensor:
    if mask_bits is None:
        return torch.ones_like(like, dtype=torch.bool)
    return torch.stack([((mask_bits >> c) & 1) == 0 for c in range(3)], dim=1)


class InstanceNorms:
    """Per-instance 2-norms of (V,3) vectors of a (possibly batched) mesh."""

    def __init__(self, vtx_off, device):
        vtx_off = np.asarray(vtx_off, np.int64)
        self.n = len(vtx_off) - 1
        counts = np.diff(vtx_off)
        self.inst_of_vertex = torch.as_tensor(np.repeat(np.arange(self.n), counts), device=device)
        self.vtx_off = vtx_off

    def __call__(self, r: torch.Tensor) -> torch.Tensor:
        s = torch.zeros(self.n, dtype=torch.float64, device=r.device)
        s.index_add_(0, self.inst_of_vertex, (r.double() ** 2).sum(1))
        return s.sqrt()

    def zero_done(self, r: torch.Tensor, done: torch.Tensor) -> torch.Tensor:
        return torch.where(done[self.inst_of_vertex][:, None], torch.zeros_like(r), r)


def refine_equilibrium(solver, shadow, x: torch.Tensor, f64: torch.Tensor | None, mask: torch.Tensor | None,
                       target: np.ndarray, norms: InstanceNorms, *, max_steps: int = 8, cg_rtol: float = 1e-3,
                       cg_max_iters: int = 20000, mu=None, lam=None, activation=None,
                       shadow_materials: dict | None = None) -> tuple[torch.Tensor, RefineReport]:
    """Refine the equilibrium x (the FP32 solver's mesh dtype) until the FP64
    free-DOF residual of every instance is <= target[i].  ``shadow`` is the
    FP64 potential (same tets, FP64 element data); ``shadow_materials`` its
    FP64 materials (default: the FP32 ones cast).  Returns the FP64 iterate
    (x itself is updated to its rounding) and a report."""
    m64 = shadow.mesh
    dev = x.device
    free = _free(mask, x)
    x64 = x.double()
    mats = shadow_materials
    if mats is None:
        src = solver.potential.materials
        mats = {k: v.double().contiguous() for k, v in src.items()}
        if mu is not None:
            mats["mu"] = mu.double().contiguous()
        if lam is not None:
            mats["lambda"] = lam.double().contiguous()
        if activation is not None:
            mats["activation"] = activation.double().contiguous()
    if "activation" in mats:
        m64.set_activation(mats["activation"])
    target_t = torch.as_tensor(target, dtype=torch.float64, device=dev)
    failed: set[int] = set()
    cg_its = 0
    steps = 0
    prev = None
    while True:
        g = torch.zeros_like(x64)
        m64.eval(shadow.material, mats["mu"], mats.get("lambda"), x64, grad=g)
        r = -g if f64 is None else f64 - g
        r = torch.where(free, r, torch.zeros_like(r))
        rn = norms(r)
        done = rn <= target_t
        if prev is not None:
            # a correction that does not halve the residual: the Hessian the
            # CG uses is too far from the true one here (ARAP's projected
            # Hessian on compressed elements); stop refining this instance
            stalled = (rn > 0.5 * prev) & ~done
            for i in torch.nonzero(stalled).flatten().tolist():
                failed.add(i)
        prev = rn
        if failed:
            done[list(failed)] = True
        if bool(done.all()) or steps >= max_steps:
            break
        r = norms.zero_done(r, done)
        d = torch.zeros_like(x)
        st = solver.cg(x, r.to(x.dtype).contiguous(), d, mask, rtol=cg_rtol, max_iters=cg_max_iters, mu=mu, lam=lam,
                       activation=activation)
        cg_its += max(s.n_steps for s in st)
        for i, s in enumerate(st):
            if not bool(done[i]) and s.result not in (Result.PRIMARY_SUCCESS, Result.MAX_STEPS_REACHED):
                failed.add(i)
        bad = torch.zeros(norms.n, dtype=torch.bool, device=dev)
        if failed:
            bad[list(failed)] = True
        d = norms.zero_done(d, bad)
        x64 += d.double()
        x.copy_(x64.to(x.dtype))
        steps += 1
    rn_np = rn.cpu().numpy()
    rep = RefineReport(steps, rn_np, np.asarray(target, np.float64), rn_np <= np.asarray(target), sorted(failed),
                       cg_its)
    return x64, rep


def refine_linear(solver, shadow, x64: torch.Tensor, rhs: torch.Tensor, y: torch.Tensor, mask: torch.Tensor | None,
                  rtol: float, norms: InstanceNorms, *, max_steps: int = 8, cg_rtol: float = 1e-3,
                  cg_max_iters: int = 20000, x32: torch.Tensor | None = None, mu=None, lam=None, activation=None,
                  shadow_materials: dict | None = None) -> tuple[torch.Tensor, RefineReport]:
    """Refine the solution y of H(x) y = rhs on the free DOFs (the adjoint
    solve) until ||rhs - H64(x64) y|| <= rtol ||rhs|| per instance; H64 is the
    FP64 Hessian-vector pass of ``shadow`` at the FP64 point x64, the
    corrections come from the FP32 PCG at x32 (default: x64 rounded).
    Returns y in FP64 (``y`` is updated to its rounding) and a report."""
    m64 = shadow.mesh
    dev = y.device
    free = _free(mask, y)
    x32 = x64.to(y.dtype).contiguous() if x32 is None else x32
    mats = shadow_materials
    if mats is None:
        src = solver.potential.materials
        mats = {k: v.double().contiguous() for k, v in src.items()}
        if mu is not None:
            mats["mu"] = mu.double().contiguous()
        if lam is not None:
            mats["lambda"] = lam.double().contiguous()
        if activation is not None:
            mats["activation"] = activation.double().contiguous()
    if "activation" in mats:
        m64.set_activation(mats["activation"])
    b = torch.where(free, rhs.double(), torch.zeros_like(x64))
    target = (rtol * norms(b)).cpu().numpy()
    target_t = torch.as_tensor(target, dtype=torch.float64, device=dev)
    y64 = y.double()
    failed: set[int] = set()
    cg_its = 0
    steps = 0
    while True:
        hy = torch.zeros_like(x64)
        m64.eval(shadow.material, mats["mu"], mats.get("lambda"), x64, y64.contiguous(), hess_prod=hy)
        r = torch.where(free, b - hy, torch.zeros_like(b))
        rn = norms(r)
        done = rn <= target_t
        if failed:
            done[list(failed)] = True
        if bool(done.all()) or steps >= max_steps:
            break
        r = norms.zero_done(r, done)
        d = torch.zeros_like(y)
        st = solver.cg(x32, r.to(y.dtype).contiguous(), d, mask, rtol=cg_rtol, max_iters=cg_max_iters, mu=mu,
                       lam=lam, activation=activation)
        cg_its += max(s.n_steps for s in st)
        for i, s in enumerate(st):
            if not bool(done[i]) and s.result not in (Result.PRIMARY_SUCCESS, Result.MAX_STEPS_REACHED):
                failed.add(i)
        bad = torch.zeros(norms.n, dtype=torch.bool, device=dev)
        if failed:
            bad[list(failed)] = True
        d = norms.zero_done(d, bad)
        y64 += d.double()
        steps += 1
    y.copy_(y64.to(y.dtype))
    rn_np = rn.cpu().numpy()
    rep = RefineReport(steps, rn_np, target, rn_np <= target, sorted(failed), cg_its)
    return y64, rep


def finish_fp64(shadow, x64: torch.Tensor, f64: torch.Tensor | None, mask: torch.Tensor | None, target: np.ndarray,
                redo: list[int], norms: InstanceNorms, config, solver64=None):
    """FP64 PNCG on the shadow from x64 (updated in place) for the instances
    in ``redo`` to their absolute targets; the other instances are held fixed.
    Returns (per-instance stats of that solve, the solver used)."""
    import dataclasses

    from .solver import DeviceSolver

    solver64 = solver64 if solver64 is not None else DeviceSolver(shadow)
    dev = x64.device
    keep = torch.ones(x64.shape[0], dtype=torch.bool, device=dev)
    for i in redo:
        keep[int(norms.vtx_off[i]):int(norms.vtx_off[i + 1])] = False
    base = mask if mask is not None else torch.zeros(x64.shape[0], dtype=torch.uint8, device=dev)
    mask64 = torch.where(keep, torch.full_like(base, 7), base)
    atol = float(min(target[i] for i in redo))
    cfg = dataclasses.replace(config, rtol=0.0, atol=atol, wolfe_eps=None)
    st = solver64.solve(x64, f64, mask64, cfg)
    return st, solver64
