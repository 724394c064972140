"""Static forward problem with liblaf/apple's interface.

Mirrors ``forward/`` of the reference:

* ``DofMapBuilder`` / ``DofMap`` -- forward/dof_map/_builder.py:13-64,
  _dof_map.py:10-49 (free DOFs = flatnonzero(~mask) of the (V,3) array);
* ``ModelBuilder`` / ``Model`` -- forward/_builder.py:12-40, _model.py:14-74
  (a named set of potentials summed after zeroing, warp/model/_model.py:13-36);
* ``ForwardProblem`` -- forward/_problem.py:18-59 (free-vector view);
* ``Forward`` -- forward/_forward.py:18-56, ``step()`` runs ``PNCG``.

``PNCG.minimize`` runs the fused device solver (apple_amd.solver) when the
model's potentials are FEM energies plus external forces: one FEM potential
directly, several (e.g. the fat / aponeurosis / muscle energies of one mesh,
exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:256-279,
or the energies of several meshes) as one ``DeviceModel``.  Any other
potential makes it fall back to a PNCG driven from the host over the
model's potentials (same algorithm, one call per potential).
"""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np
import torch

from .common import FIXED_MASK, FIXED_VALUE, GLOBAL_POINT_ID
from .fem import FemPotential
from .mesh import Mesh
from .model import DeviceModel
from .potential import ExternalForce
from .runtime import normalize_device
from .solver import DeviceSolver, OptState, PncgConfig, Result, SolveStats, mask_bits


class DofMapBuilder:
    def __init__(self, dim: int = 3):
        self.dim = dim
        self.fixed_mask = np.zeros((0, dim), bool)
        self.full_values = np.zeros((0, dim))

    @property
    def n_points(self) -> int:
        return self.fixed_mask.shape[0]

    def add_vertices(self, obj: Any) -> None:
        mesh = Mesh.from_pyvista(obj)
        mesh.point_data[GLOBAL_POINT_ID.vtk] = np.arange(self.n_points, self.n_points + mesh.n_points)
        if obj is not mesh and hasattr(obj, "point_data"):
            obj.point_data[GLOBAL_POINT_ID.vtk] = mesh.point_data[GLOBAL_POINT_ID.vtk]
        self.fixed_mask = np.pad(self.fixed_mask, ((0, mesh.n_points), (0, 0)), constant_values=False)
        self.full_values = np.pad(self.full_values, ((0, mesh.n_points), (0, 0)), constant_values=0.0)

    def add_fixed(self, obj: Any) -> None:
        mesh = Mesh.from_pyvista(obj)
        gid = np.asarray(mesh.point_data[GLOBAL_POINT_ID.vtk])
        self.fixed_mask[gid] = np.asarray(mesh.point_data[FIXED_MASK.vtk], bool).reshape(-1, self.dim)
        self.full_values[gid] = np.asarray(mesh.point_data[FIXED_VALUE.vtk], np.float64).reshape(-1, self.dim)

    def finalize(self, device=None, dtype=torch.float32) -> "DofMap":
        return DofMap(self.fixed_mask.copy(), self.full_values.copy(), device=device, dtype=dtype)


class DofMap:
    def __init__(self, fixed_mask, full_values, *, device=None, dtype=torch.float32):
        self.dim = 3
        self.device = normalize_device(device)
        self.dtype = dtype
        self.fixed_mask = np.asarray(fixed_mask, bool).reshape(-1, 3)
        self.n_points = self.fixed_mask.shape[0]
        flat = self.fixed_mask.reshape(-1)
        fi = np.flatnonzero(flat)
        self.fixed_indices = torch.as_tensor(fi, device=self.device)
        self.fixed_values = torch.as_tensor(np.asarray(full_values, np.float64).reshape(-1)[fi], dtype=dtype,
                                            device=self.device)
        self.free_indices = torch.as_tensor(np.flatnonzero(~flat), device=self.device)
        self.mask_bits = mask_bits(self.fixed_mask, self.device)

    @property
    def n_fixed(self) -> int:
        return int(self.fixed_indices.numel())

    @property
    def n_free(self) -> int:
        return int(self.free_indices.numel())

    @property
    def n_full(self) -> int:
        return self.n_points * self.dim

    def to_free(self, full: torch.Tensor) -> torch.Tensor:
        return full.reshape(-1)[self.free_indices]

    to_free_grad = to_free
    to_free_hess_diag = to_free

    def to_full(self, free: torch.Tensor) -> torch.Tensor:
        out = torch.empty(self.n_full, dtype=free.dtype, device=free.device)
        out[self.fixed_indices] = self.fixed_values.to(free.dtype)
        out[self.free_indices] = free
        return out.reshape(self.n_points, self.dim)

    def to_full_grad(self, free: torch.Tensor) -> torch.Tensor:
        out = torch.zeros(self.n_full, dtype=free.dtype, device=free.device)
        out[self.free_indices] = free
        return out.reshape(self.n_points, self.dim)


@dataclasses.dataclass
class ModelState:
    u: torch.Tensor


class Model:
    """Named potentials summed over a full (V,3) layout (WarpModel + adapter)."""

    State = ModelState

    def __init__(self, dof_map: DofMap, potentials: dict[str, Any]):
        self.dof_map = dof_map
        self.potentials = dict(potentials)
        pots = list(self.potentials.values())
        self.dtype = pots[0].dtype if pots else dof_map.dtype
        self.device = dof_map.device

    @property
    def dim(self) -> int:
        return 3

    @property
    def n_fixed(self) -> int:
        return self.dof_map.n_fixed

    @property
    def n_free(self) -> int:
        return self.dof_map.n_free

    @property
    def n_full(self) -> int:
        return self.dof_map.n_full

    @property
    def n_points(self) -> int:
        return self.dof_map.n_points

    def init(self) -> ModelState:
        return ModelState(self.dof_map.to_full(torch.zeros(self.n_free, dtype=self.dtype, device=self.device)))

    def _zeros(self):
        return torch.zeros(self.n_points, 3, dtype=self.dtype, device=self.device)

    @staticmethod
    def _view(t: torch.Tensor, pot) -> torch.Tensor:
        """The rows a potential evaluates in: the vertex arrays are global, a
        potential of a mesh added later covers a prefix up to its highest
        global id (a contiguous view; += outputs leave the rest untouched)."""
        n = pot.n_points
        return t if t.shape[0] == n else t[:n]

    def fun(self, u: torch.Tensor) -> torch.Tensor:
        out = torch.zeros(1, dtype=torch.float64, device=self.device)
        for pot in self.potentials.values():
            pot.fun(self._view(u, pot), out)
        return out[0]

    def grad(self, u: torch.Tensor) -> torch.Tensor:
        out = self._zeros()
        for pot in self.potentials.values():
            pot.grad(self._view(u, pot), self._view(out, pot))
        return out

    def hess_diag(self, u: torch.Tensor) -> torch.Tensor:
        out = self._zeros()
        for pot in self.potentials.values():
            pot.hess_diag(self._view(u, pot), self._view(out, pot))
        return out

    def hess_prod(self, u: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        out = self._zeros()
        for pot in self.potentials.values():
            pot.hess_prod(self._view(u, pot), self._view(p, pot), self._view(out, pot))
        return out

    def hess_quad(self, u: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        out = torch.zeros(1, dtype=torch.float64, device=self.device)
        for pot in self.potentials.values():
            pot.hess_quad(self._view(u, pot), self._view(p, pot), out)
        return out[0]

    def external_force(self) -> torch.Tensor | None:
        """Sum of the ExternalForce potentials as one dense (V,3) load."""
        f = None
        for e in self.potentials.values():
            if isinstance(e, ExternalForce):
                d = e.dense()
                if d.shape[0] < self.n_points:
                    d = torch.cat([d, d.new_zeros(self.n_points - d.shape[0], 3)])
                f = d if f is None else f + d
        return f

    def device_model(self) -> DeviceModel:
        """All FEM potentials as one device model over the model's vertices
        (built once; its materials are refreshed from the potentials)."""
        fem = [p for p in self.potentials.values() if isinstance(p, FemPotential)]
        dm = getattr(self, "_device_model", None)
        if dm is None or [id(p) for p in dm.parts] != [id(p) for p in fem]:
            dm = DeviceModel(fem, self.n_points, name="model")
            self._device_model = dm
        dm.sync_materials()
        return dm

    # fused-solver view: FEM potentials plus external forces
    def fused_parts(self):
        fem = [p for p in self.potentials.values() if isinstance(p, FemPotential)]
        ext = [p for p in self.potentials.values() if isinstance(p, ExternalForce)]
        if not fem or len(fem) + len(ext) != len(self.potentials):
            return None
        f = self.external_force()
        if len(fem) == 1 and fem[0].n_points == self.n_points:
            return fem[0], f
        return self.device_model(), f


class ModelBuilder:
    def __init__(self):
        self.dof = DofMapBuilder()
        self.potentials: list[Any] = []

    def add_vertices(self, obj: Any) -> None:
        self.dof.add_vertices(obj)

    def add_fixed(self, obj: Any) -> None:
        self.dof.add_fixed(obj)

    def add_potential(self, potential: Any) -> None:
        self.potentials.append(potential)

    def finalize(self) -> Model:
        pots = self.potentials
        device = pots[0].device if pots else None
        dtype = pots[0].dtype if pots else torch.float32
        dof_map = self.dof.finalize(device=device, dtype=dtype)
        # duplicate names overwrite, as in the reference (forward/_builder.py:32-34)
        return Model(dof_map, {p.name: p for p in pots})


class ForwardProblem:
    """Free-DOF view of a model -- forward/_problem.py:18-59."""

    State = ModelState

    def __init__(self, model: Model):
        self.model = model

    def before_trial(self, state: ModelState, u_free: torch.Tensor) -> ModelState:
        return ModelState(self.model.dof_map.to_full(u_free))

    def max_step_size(self, state: ModelState, p: torch.Tensor) -> torch.Tensor:
        return torch.ones((), dtype=torch.float64, device=self.model.device)

    def fun(self, state: ModelState) -> torch.Tensor:
        return self.model.fun(state.u)

    def grad(self, state: ModelState) -> torch.Tensor:
        return self.model.dof_map.to_free_grad(self.model.grad(state.u))

    def hess_diag(self, state: ModelState) -> torch.Tensor:
        return self.model.dof_map.to_free_hess_diag(self.model.hess_diag(state.u))

    def hess_prod(self, state: ModelState, p: torch.Tensor) -> torch.Tensor:
        return self.model.dof_map.to_free_grad(self.model.hess_prod(state.u, self.model.dof_map.to_full_grad(p)))

    def hess_quad(self, state: ModelState, p: torch.Tensor) -> torch.Tensor:
        return self.model.hess_quad(state.u, self.model.dof_map.to_full_grad(p))


@dataclasses.dataclass
class SolutionState:
    """The solver's end state: ``u`` (the full displacement, as ModelState)
    and ``grad``, the free-DOF gradient there (solution.state.grad,
    exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:395-406)."""

    u: torch.Tensor
    grad: torch.Tensor
    opt: OptState | None = None


@dataclasses.dataclass
class Solution:
    result: Result
    state: SolutionState
    stats: SolveStats

    @property
    def success(self) -> bool:
        return self.result == Result.PRIMARY_SUCCESS

    @property
    def n_steps(self) -> int:
        return self.stats.n_steps


class PNCG:
    """Preconditioned nonlinear CG (see apple_amd/csrc/solver.hip).

    ``minimize`` solves to the end; ``init`` / ``step`` / ``terminate`` /
    ``postprocess`` drive the same device solve one accepted step at a time,
    as liblaf-peach's Optimizer does in apple's experiments
    (exp/2026/05/06/toy/src/21-smas-prestrain-stable-neo-hookean-muscle.py:497-564);
    the state they return carries the reference's field names (OptState)."""

    def __init__(self, config: PncgConfig | None = None, **kwargs):
        self.config = dataclasses.replace(config or PncgConfig(), **kwargs)
        self._solvers: dict[int, DeviceSolver] = {}

    def _solver_for(self, fem) -> DeviceSolver:
        key = id(fem)
        if key not in self._solvers:
            self._solvers[key] = DeviceSolver(fem)
        return self._solvers[key]

    def minimize(self, problem: ForwardProblem, state: ModelState, u_free: torch.Tensor):
        model = problem.model
        parts = model.fused_parts()
        if parts is not None:
            return self._fused(problem, parts, u_free)
        return self._host(problem, u_free)

    def _solution(self, problem: ForwardProblem, x: torch.Tensor, stats: SolveStats, opt=None) -> Solution:
        grad = problem.grad(ModelState(x))
        return Solution(stats.result, SolutionState(x, grad, opt), stats)

    def _fused(self, problem: ForwardProblem, parts, u_free):
        model = problem.model
        fem, f = parts
        x = model.dof_map.to_full(u_free.to(fem.dtype)).contiguous()
        stats = self._solver_for(fem).solve(x, f, model.dof_map.mask_bits, self.config)[0]
        return self._solution(problem, x, stats), ModelState(x)

    # ------------------------------------------------------------ step-wise (Optimizer protocol)

    def init(self, problem: ForwardProblem, state: ModelState, u_free: torch.Tensor) -> OptState:
        parts = problem.model.fused_parts()
        if parts is None:
            raise NotImplementedError("step-wise PNCG runs on the fused device path (FEM + external forces)")
        model = problem.model
        fem, f = parts
        self._run = dict(solver=self._solver_for(fem), x=model.dof_map.to_full(u_free.to(fem.dtype)).contiguous(),
                         f=f)
        self._run["solver"].begin(self._run["x"], f, model.dof_map.mask_bits, self.config)
        # evaluate the start point (the solver's first trial), no step yet
        opt = self._run["solver"].advance(steps=0, max_iters=1)[0]
        self._run["prev"] = self._run["x"].clone()
        return opt

    def step(self, problem: ForwardProblem, state: ModelState, opt_state: OptState):
        run = self._run
        x = torch.empty_like(run["x"])
        opt = run["solver"].advance(steps=1, x_out=x)[0]
        if opt.line_search_state.alpha > 0:
            opt.direction = problem.model.dof_map.to_free(x - run["prev"]) / opt.line_search_state.alpha
        run["prev"] = x
        return ModelState(x), opt

    def terminate(self, problem: ForwardProblem, state: ModelState, opt_state: OptState):
        done = opt_state.result != Result.RUNNING
        return done, opt_state.result

    def postprocess(self, problem: ForwardProblem, state: ModelState, opt_state: OptState, result) -> Solution:
        run = self._run
        stats = run["solver"].end()[0]
        self._run = None
        return self._solution(problem, run["x"], stats, opt_state)

    def _host(self, problem: ForwardProblem, u_free: torch.Tensor):
        """PNCG over the ForwardProblem protocol (any sum of potentials)."""
        cfg = self.config
        x = u_free.clone()
        st = problem.before_trial(None, x)
        f = float(problem.fun(st))
        g = problem.grad(st).to(torch.float64)
        gn0 = float(torch.linalg.norm(g))
        f_first = f
        g_prev = g.clone()
        p_prev = torch.zeros_like(g)
        n_steps = n_evals = n_bt = 0
        result = Result.MAX_STEPS_REACHED
        alpha = beta = 0.0
        restarts = 0
        while True:
            gn = float(torch.linalg.norm(g))
            if not np.isfinite(gn) or not np.isfinite(f):
                result = Result.NON_FINITE
                break
            if gn <= cfg.rtol * gn0 or gn <= cfg.atol:
                result = Result.PRIMARY_SUCCESS
                break
            if n_steps >= cfg.max_steps:
                break
            hd = problem.hess_diag(st).to(torch.float64).abs()
            pos = hd > 0
            mean = hd[pos].mean() if bool(pos.any()) else torch.tensor(1.0, dtype=torch.float64, device=hd.device)
            P = 1.0 / torch.where(pos, hd, mean)
            beta = 0.0
            if n_steps > 0 and restarts == 0:
                y = g - g_prev
                yp = float(torch.dot(y, p_prev))
                if abs(yp) > 1e-300:
                    Py = P * y
                    beta = float(torch.dot(g, Py)) / yp - (float(torch.dot(y, Py)) / yp) * (float(torch.dot(p_prev, g)) / yp)
                else:
                    beta = float("inf")
                if not (np.isfinite(beta) and abs(beta) <= cfg.beta_max):
                    beta = 0.0
            p = -P * g + beta * p_prev
            gd = float(torch.dot(g, p))
            if not (np.isfinite(gd) and gd < 0):
                p = -P * g
                gd = float(torch.dot(g, p))
                beta = 0.0
            pHp = float(problem.hess_quad(st, p.to(x.dtype)))
            alpha = -gd / pHp if pHp != 0 else (1.0 if -gd > 0 else 0.0)
            if not (alpha > 0 and np.isfinite(alpha)):
                alpha = 1.0
            alpha *= cfg.overstep
            eps = cfg.wolfe_eps if cfg.wolfe_eps is not None else (1e-6 if x.dtype == torch.float32 else 1e-12)
            accepted = False
            for _ in range(cfg.max_backtracking + 1):
                xt = x + (alpha * p).to(x.dtype)
                stt = problem.before_trial(st, xt)
                ft = float(problem.fun(stt))
                n_evals += 1
                gt = problem.grad(stt).to(torch.float64)
                if np.isfinite(ft) and (
                    ft <= f + cfg.c1 * alpha * gd
                    or (cfg.wolfe_delta > 0 and ft <= f + eps * abs(f)
                        and float(torch.dot(gt, p)) <= (2 * cfg.wolfe_delta - 1) * gd)
                ):
                    accepted = True
                    break
                alpha *= 0.5
                n_bt += 1
            if accepted:
                x, st, f = xt, stt, ft
                g_prev, p_prev, g = g, p, gt
                n_steps += 1
                restarts = 0
            else:
                if restarts >= 1:
                    result = Result.STAGNATION
                    break
                restarts += 1
        stats = SolveStats(result, n_steps, n_evals, n_bt, f, float(torch.linalg.norm(g)), gn0, alpha, beta, f_first)
        return Solution(result, SolutionState(st.u, g.to(st.u.dtype)), stats), st


class Forward:
    """forward/_forward.py:18-56: ``step()`` minimises from the current state."""

    def __init__(self, model: Model, optimizer: PNCG | None = None, state: ModelState | None = None):
        self.model = model
        self.optimizer = optimizer if optimizer is not None else PNCG(max_steps=1500)
        self.state = state if state is not None else model.init()

    @property
    def free(self) -> torch.Tensor:
        return self.model.dof_map.to_free(self.state.u)

    @property
    def problem(self) -> ForwardProblem:
        if not hasattr(self, "_problem"):
            self._problem = ForwardProblem(self.model)
        return self._problem

    def step(self) -> Solution:
        solution, self.state = self.optimizer.minimize(self.problem, self.state, self.free)
        return solution
