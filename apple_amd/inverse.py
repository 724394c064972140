"""Inverse loop: fit per-tet materials (e.g. muscle activations) to target positions.

Restates the inverse experiments of liblaf/apple on this package's device
path.  The explicit form is exp/2025/08/27/inverse/src/30-inverse.py:71-110:

    fun(params):  set the materials, solve the forward problem warm-started
                  from the previous solution (:71-74), evaluate the loss;
    jac(params):  dL/du by differentiating the loss, the adjoint solve
                  H p = -dL/du (:85-93), then dL/dq += p^T d grad/d q
                  (mixed_derivative_prod, :96-97);
    loss:         1/2 sum over the selected points of |u - target|^2 plus a
                  regulariser sum (a - mean a)^2 over the activations (:100-110);

and the later experiments name the two terms ``PointToPointLoss`` and
``UniformActivationLoss`` (exp/2026/03/11/test-inverse/src/30-inverse.py:68-88)
and drive them with Adam (optax.adam) over a parameter vector mapped to the
materials by ``make_materials`` (:27-34; the log-parameterised muscle
activation of exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-
hookean.py:134-147, 413-431), warm-starting every forward solve
(:432-449).

Here the forward solve is the device PNCG, the adjoint is the device PCG and
the mixed term is one element pass, all behind
``apple_amd.adjoint.equilibrium`` (a ``torch.autograd.Function``);
``make_materials`` and the losses are ordinary torch functions on device
tensors, so autograd chains dL/d(materials) to the parameters, and
``torch.optim.Adam`` takes the step on the device.
"""

from __future__ import annotations

import dataclasses
import time
from collections.abc import Callable, Sequence

import torch

from . import adjoint as _adjoint
from .solver import PncgConfig, Result


class Loss:
    """A term of the objective: ``value(u, materials)`` -> scalar tensor."""

    name: str = "loss"

    def value(self, u: torch.Tensor, materials: dict[str, torch.Tensor]) -> torch.Tensor:
        raise NotImplementedError

    def value_and_grad(self, u: torch.Tensor, materials: dict[str, torch.Tensor]):
        """(L, dL/du, {name: dL/dq}) by autograd, for callers that want the pieces."""
        u = u.detach().requires_grad_(True)
        mats = {k: v.detach().requires_grad_(True) for k, v in materials.items()}
        L = self.value(u, mats)
        names = list(mats)
        grads = torch.autograd.grad(L, [u, *mats.values()], allow_unused=True)
        dq = {k: (g if g is not None else torch.zeros_like(mats[k])) for k, g in zip(names, grads[1:])}
        return L.detach(), grads[0] if grads[0] is not None else torch.zeros_like(u), dq


@dataclasses.dataclass
class PointToPointLoss(Loss):
    """weight/2 * sum_i |u[indices[i]] - target[i]|^2 (30-inverse.py:101-103)."""

    indices: torch.Tensor
    target: torch.Tensor
    weight: float = 1.0
    name: str = "point_to_point"

    def value(self, u, materials):
        d = u[self.indices] - self.target
        return 0.5 * self.weight * (d * d).sum()


@dataclasses.dataclass
class UniformActivationLoss(Loss):
    """weight * sum over the selected tets of |a_t - mean(a)|^2 (30-inverse.py:104-108)."""

    muscle_indices: torch.Tensor | None = None
    weight: float = 1.0
    key: str = "activation"
    name: str = "uniform_activation"

    def value(self, u, materials):
        a = materials[self.key]
        if self.muscle_indices is not None:
            a = a[self.muscle_indices]
        d = a - a.mean(dim=0, keepdim=True)
        return self.weight * (d * d).sum()


def directional_activation(x, y, z) -> torch.Tensor:
    """(x - 1, y - 1, z - 1, 0, 0, 0) per tet: A = diag(x, y, z)
    (exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:106-122)."""
    zeros = torch.zeros_like(x)
    return torch.stack([x - 1.0, y - 1.0, z - 1.0, zeros, zeros, zeros], dim=-1)


def muscle_x_activation(log_relative, *, prestrain: float, log_min: float = -2.0, log_max: float = 2.0):
    """Contraction along x scaled by exp(clip(log_relative)), y = 1/prestrain^2,
    z = prestrain (the same file, :125-147)."""
    rel = torch.exp(torch.clamp(log_relative, log_min, log_max))
    x = rel * prestrain
    return directional_activation(x, torch.ones_like(x) / prestrain**2, torch.ones_like(x) * prestrain)


@dataclasses.dataclass
class InverseStep:
    step: int
    loss: float
    grad_norm: float
    forward_steps: int
    adjoint_iters: int
    seconds: float


class Inverse:
    """Adam over ``params`` of L(x*(make_materials(params)), materials).

    potential: a FEM potential (``fem.StableNeoHookeanMuscle`` for activations,
    or SNH / ARAP for mu, lambda), a ``batch.Batch``'s potential, or a
    ``model.DeviceModel`` (several potentials, e.g. fat + aponeurosis +
    muscle of exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:256-279);
    make_materials(params) -> dict with any of "mu", "lambda", "activation"
    (missing ones keep the potential's values), or for a DeviceModel a list
    with one such dict (or None) per potential.  Every forward solve starts
    from the previous solution (warm start); the first from ``x0``.  The
    solver's rtol is relative to each solve's first gradient, which is small
    after a warm start; so the first (cold) solve fixes the absolute target:
    later solves also stop at ||g|| <= rtol ||g_0 of the cold solve|| (atol).
    """

    def __init__(
        self,
        potential,
        losses: Sequence[Loss],
        make_materials: Callable[[torch.Tensor], dict[str, torch.Tensor]],
        *,
        f_ext: torch.Tensor | None = None,
        mask: torch.Tensor | None = None,
        x0: torch.Tensor | None = None,
        lr: float = 0.03,
        forward: PncgConfig | None = None,
        adjoint: _adjoint.AdjointConfig | None = None,
    ):
        self.potential = potential
        self.losses = list(losses)
        self.make_materials = make_materials
        self.f_ext = f_ext
        self.mask = mask
        mesh = potential.mesh
        x0 = x0 if x0 is not None else torch.zeros(mesh.n_verts, 3, dtype=mesh.dtype, device=mesh.device)
        self.x_init = x0.detach().clone().contiguous()
        self.x = self.x_init.clone()
        self.lr = lr
        self.forward = forward or PncgConfig()
        # the adjoint CG tolerance reachable in the mesh precision
        self.adjoint = adjoint or _adjoint.AdjointConfig(rtol=1e-6 if mesh.dtype == torch.float32 else 1e-10)
        self.history: list[InverseStep] = []
        self._warm_cfg: PncgConfig | None = None

    def _materials(self, params):
        made = self.make_materials(params)
        pot = self.potential
        if isinstance(made, (list, tuple)):
            # a DeviceModel: one dict per potential (None keeps it), gathered
            # into the model's per-tet arrays (autograd reaches each part)
            per = [dict(part.materials, **(d or {})) for part, d in zip(pot.parts, made)]
            return pot.gather(per)
        base = dict(pot.materials)
        base.update(made)
        return base

    def loss_at(self, params: torch.Tensor, *, warm: bool = True) -> torch.Tensor:
        """Loss as a differentiable function of params (one forward solve, from
        the last solution when warm, else from x0)."""
        mats = self._materials(params)
        cfg = self._warm_cfg if self._warm_cfg is not None else self.forward
        x = _adjoint.equilibrium(
            self.potential, mats["mu"], mats.get("lambda"), self.f_ext, self.mask, self.x if warm else self.x_init,
            cfg, self.adjoint, activation=mats.get("activation"),
        )
        fwd, _ = _adjoint.last_stats()
        bad = [s for s in fwd if s.result not in (Result.PRIMARY_SUCCESS, Result.STAGNATION)]
        if bad:
            raise RuntimeError(f"forward solve failed during the inverse: {bad[0]}")
        if self._warm_cfg is None:
            g0 = min(s.grad_norm_first for s in fwd)
            self._warm_cfg = dataclasses.replace(self.forward, atol=max(self.forward.atol, self.forward.rtol * g0))
        self._last_x = x
        return sum(term.value(x, mats) for term in self.losses)

    def value_and_grad(self, params: torch.Tensor) -> tuple[float, torch.Tensor]:
        p = params.detach().clone().requires_grad_(True)
        L = self.loss_at(p)
        L.backward()
        self.x = self._last_x.detach().clone().contiguous()  # warm start of the next solve
        return float(L.detach()), p.grad.detach()

    def solve(self, params: torch.Tensor, steps: int, callback: Callable[[InverseStep, torch.Tensor], None] | None = None,
              gtol: float = 0.0) -> torch.Tensor:
        """``steps`` Adam steps from ``params``; returns the final parameters."""
        p = params.detach().clone().requires_grad_(True)
        opt = torch.optim.Adam([p], lr=self.lr)
        for k in range(steps):
            t0 = time.perf_counter()
            opt.zero_grad(set_to_none=True)
            L = self.loss_at(p)
            L.backward()
            self.x = self._last_x.detach().clone().contiguous()
            fwd, adj = _adjoint.last_stats()
            gnorm = float(p.grad.norm())
            opt.step()
            rec = InverseStep(k, float(L.detach()), gnorm, sum(s.n_steps for s in fwd),
                              sum(s.n_steps for s in adj), time.perf_counter() - t0)
            self.history.append(rec)
            if callback is not None:
                callback(rec, p.detach())
            if gnorm <= gtol:
                break
        return p.detach()
