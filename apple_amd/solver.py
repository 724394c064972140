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

    @property
    def relative_decrease(self) -> float:
        """||g|| / ||g_0||, the quantity the stopping rule compares with rtol
        (the reference's experiments log it as solution.stats.relative_decrease,
        exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:395-406;
        peach's own definition is not in the reference tree)."""
        return self.relative_grad_norm


@dataclasses.dataclass
class LineSearchState:
    """The last accepted step's line search (opt_state.line_search_state)."""

    alpha: float
    f0: float
    f_alpha: float
    step: int
    ok: bool


@dataclasses.dataclass
class ConvergenceState:
    grad_norm: float
    grad_norm_first: float


@dataclasses.dataclass
class HessDampingState:
    """factor: the Hessian damping the quadratic-form step uses (0: none);
    hess_diag_mean: the mean positive Hessian diagonal (the Jacobi fallback)."""

    factor: float
    hess_diag_mean: float


@dataclasses.dataclass
class OptState:
    """Per-instance PNCG state after a step, with the field names apple's
    experiments read (exp/2026/05/06/toy/src/21-smas-prestrain-stable-neo-hookean-muscle.py:369-391).
    direction: the last step divided by its alpha (None before the first step)."""

    n_steps: int
    n_evals: int
    result: Result
    line_search_state: LineSearchState
    convergence_state: ConvergenceState
    hess_damping_state: HessDampingState
    beta: float
    direction: torch.Tensor | None = None


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
        pot = self.potential
        mu = pot.materials["mu"] if mu is None else mu
        lam = pot.materials.get("lambda") if lam is None else lam
        if pot.material == _lib.APL_ARAP:
            lam = None
        if pot.material == _lib.APL_SNH_MUSCLE or (pot.material == _lib.APL_MODEL and
                                                   "activation" in pot.materials):
            m.set_activation(pot.materials["activation"] if activation is None else activation)
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

    # ------------------------------------------------------------ step-wise solve

    def begin(self, x: torch.Tensor, f_ext: torch.Tensor | None = None, mask: torch.Tensor | None = None,
              config: PncgConfig | None = None, *, mu=None, lam=None, activation=None) -> None:
        """Start a step-wise solve from x (Optimizer.init); x receives the
        iterate at end()."""
        config = config or PncgConfig()
        m = self.mesh
        mu, lam = self._mat_ptrs(mu, lam, activation)
        self._params = config.to_c(m.dtype)
        self._run_x = x
        self._run_keep = (mu, lam, f_ext, mask)
        _lib.check(
            _lib.lib().apl_pncg_begin(
                self._handle, mu, lam, ptr(x, **m.vec("x")), ptr(f_ext, **m.vec("f_ext")),
                ptr(mask, dtype=torch.uint8, device=m.device, shape=(m.n_verts,), name="mask"),
                ctypes.byref(self._params), current_stream(m.device),
            )
        )

    def advance(self, steps: int = 1, max_iters: int = 0, x_out: torch.Tensor | None = None,
                hess_damping: float = 0.0) -> list[OptState]:
        """Run until every running instance has taken `steps` more accepted
        steps (or, with steps = 0, until none runs or max_iters trials ran);
        x_out receives the current iterate.  Returns the per-instance state."""
        m = self.mesh
        st = (_lib.PncgState * m.n_inst)()
        _lib.check(
            _lib.lib().apl_pncg_advance(
                self._handle, int(steps), int(max_iters), ptr(x_out, **m.vec("x_out")), st,
                current_stream(m.device),
            )
        )
        return [
            OptState(
                st[i].n_steps, st[i].n_evals, Result(st[i].status),
                LineSearchState(st[i].alpha, st[i].f0, st[i].f_alpha, st[i].ls_steps, bool(st[i].ok)),
                ConvergenceState(st[i].grad_norm, st[i].grad_norm_first),
                HessDampingState(float(hess_damping), st[i].hess_diag_mean), st[i].beta,
            )
            for i in range(m.n_inst)
        ]

    def end(self) -> list[SolveStats]:
        """Write the iterate into the x given to begin() and return the stats."""
        m = self.mesh
        stats = (_lib.PncgStats * m.n_inst)()
        _lib.check(_lib.lib().apl_pncg_end(self._handle, stats, current_stream(m.device)))
        self._run_keep = None
        return _stats(stats, m.n_inst)

    def cg(self, x: torch.Tensor, rhs: torch.Tensor, y: torch.Tensor, mask: torch.Tensor | None = None,
           rtol: float = 1e-6, max_iters: int = 10000, *, mu=None, lam=None, activation=None,
           normal: bool = False) -> list[SolveStats]:
        """Solve H(x) y = rhs on the free DOFs (y overwritten); materials as in
        solve().  normal: CG on H^2 y = H rhs (for an indefinite H; stats.f is
        then the relative residual of the normal equations)."""
        m = self.mesh
        mu, lam = self._mat_ptrs(mu, lam, activation)
        stats = (_lib.PncgStats * m.n_inst)()
        entry = _lib.lib().apl_cg_solve_normal if normal else _lib.lib().apl_cg_solve
        _lib.check(
            entry(
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
