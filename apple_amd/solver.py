"""Device PNCG and PCG solvers over one element potential (C ABI wrapper).

``DeviceSolver`` runs the whole PNCG loop on the GPU: element passes, vector
kernels and per-instance scalar decisions are HIP kernels, K iterations are
replayed from one HIP graph, and the host only reads the per-instance status
once per graph launch.  It solves every instance of a batched mesh in the
same launches.
"""

from __future__ import annotations

import ctypes
import dataclasses
import enum

import torch

from . import _lib
from .runtime import current_stream, ptr


class Result(enum.Enum):
    """Solver outcome; names follow liblaf-peach's Result as used by apple
    (tests/forward/test_static_simulation.py:85)."""

    RUNNING = _lib.APL_RUNNING
    PRIMARY_SUCCESS = _lib.APL_PRIMARY_SUCCESS
    MAX_STEPS_REACHED = _lib.APL_MAX_STEPS_REACHED
    NON_FINITE = _lib.APL_NON_FINITE
    STAGNATION = _lib.APL_STAGNATION


@dataclasses.dataclass
class SolveStats:
    result: Result
    n_steps: int
    n_evals: int
    n_backtracks: int
    f: float
    grad_norm: float
    grad_norm_first: float
    alpha: float
    beta: float
    f_first: float

    @property
    def success(self) -> bool:
        return self.result == Result.PRIMARY_SUCCESS

    @property
    def relative_grad_norm(self) -> float:
        return self.grad_norm / self.grad_norm_first if self.grad_norm_first > 0 else 0.0


@dataclasses.dataclass
class PncgConfig:
    """PNCG settings (defaults: forward/_forward.py:22-30 max_steps 1500;
    benches/bench_pncg_branching_backends.py c1 1e-4, 8 backtracks, |beta|<=10)."""

    rtol: float = 1e-6
    atol: float = 0.0
    max_steps: int = 1500
    max_backtracking: int = 8
    overstep: float = 1.0
    c1: float = 1e-4
    wolfe_delta: float = 0.1
    wolfe_eps: float | None = None  # default: 1e-6 in FP32, 1e-12 in FP64
    beta_max: float = 10.0
    iters_per_launch: int = 16
    use_graph: bool = True
    grid: int = 0
    stall_steps: int = 2000  # STAGNATION: this many steps without gradient or energy progress near rtol

    def to_c(self, dtype: torch.dtype) -> _lib.PncgParams:
        eps = self.wolfe_eps
        if eps is None:
            eps = 1e-6 if dtype == torch.float32 else 1e-12
        return _lib.PncgParams(
            self.rtol, self.atol, int(self.max_steps), int(self.max_backtracking), self.overstep, self.c1,
            self.wolfe_delta, eps, self.beta_max, int(self.iters_per_launch), 1 if self.use_graph else 0,
            int(self.grid), int(self.stall_steps),
        )


def _stats(arr, n) -> list[SolveStats]:
    return [
        SolveStats(
            Result(arr[i].status), arr[i].n_steps, arr[i].n_evals, arr[i].n_backtracks, arr[i].f, arr[i].grad_norm,
            arr[i].grad_norm_first, arr[i].alpha, arr[i].beta, arr[i].f_first,
        )
        for i in range(n)
    ]


class DeviceSolver:
    """PNCG (static equilibrium) and PCG (adjoint) for one FEM potential."""

    def __init__(self, potential):
        self.potential = potential
        self.mesh = potential.mesh
        self._handle = ctypes.c_void_p()
        _lib.check(_lib.lib().apl_solver_create(self.mesh._handle, int(potential.material), ctypes.byref(self._handle)))

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            _lib.lib().apl_solver_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _mat_ptrs(self, mu=None, lam=None, activation=None):
        """Per-tet materials: the given tensors, else the potential's (the
        muscle activation is bound to the mesh)."""
        m = self.mesh
        tet = dict(dtype=m.dtype, device=m.device, shape=(m.n_tets,))
        mu = self.potential.materials["mu"] if mu is None else mu
        lam = self.potential.materials.get("lambda") if lam is None else lam
        if self.potential.material == _lib.APL_ARAP:
            lam = None
        if self.potential.material == _lib.APL_SNH_MUSCLE:
            m.set_activation(self.potential.materials["activation"] if activation is None else activation)
        return ptr(mu, name="mu", **tet), (ptr(lam, name="lambda", **tet) if lam is not None else None)

    def solve(self, x: torch.Tensor, f_ext: torch.Tensor | None = None, mask: torch.Tensor | None = None,
              config: PncgConfig | None = None, *, mu=None, lam=None, activation=None) -> list[SolveStats]:
        """Minimise in place from ``x`` (fixed DOFs already at their values);
        mu / lam / activation override the potential's materials for this call."""
        config = config or PncgConfig()
        m = self.mesh
        mu, lam = self._mat_ptrs(mu, lam, activation)
        params = config.to_c(m.dtype)
        stats = (_lib.PncgStats * m.n_inst)()
        _lib.check(
            _lib.lib().apl_pncg_solve(
                self._handle, mu, lam, ptr(x, **m.vec("x")), ptr(f_ext, **m.vec("f_ext")),
                ptr(mask, dtype=torch.uint8, device=m.device, shape=(m.n_verts,), name="mask"),
                ctypes.byref(params), stats, current_stream(m.device),
            )
        )
        return _stats(stats, m.n_inst)

    def cg(self, x: torch.Tensor, rhs: torch.Tensor, y: torch.Tensor, mask: torch.Tensor | None = None,
           rtol: float = 1e-6, max_iters: int = 10000, *, mu=None, lam=None, activation=None) -> list[SolveStats]:
        """Solve H(x) y = rhs on the free DOFs (y overwritten); materials as in solve()."""
        m = self.mesh
        mu, lam = self._mat_ptrs(mu, lam, activation)
        stats = (_lib.PncgStats * m.n_inst)()
        _lib.check(
            _lib.lib().apl_cg_solve(
                self._handle, mu, lam, ptr(x, **m.vec("x")), ptr(rhs, **m.vec("rhs")), ptr(y, **m.vec("y")),
                ptr(mask, dtype=torch.uint8, device=m.device, shape=(m.n_verts,), name="mask"),
                float(rtol), int(max_iters), stats, current_stream(m.device),
            )
        )
        return _stats(stats, m.n_inst)


def mask_bits(fixed_mask, device) -> torch.Tensor:
    """(V,3) bool -> (V,) uint8 with bit c set when component c is fixed."""
    fm = torch.as_tensor(fixed_mask, dtype=torch.bool).reshape(-1, 3)
    bits = fm[:, 0].to(torch.uint8) | (fm[:, 1].to(torch.uint8) << 1) | (fm[:, 2].to(torch.uint8) << 2)
    return bits.to(device).contiguous()
