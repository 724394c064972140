"""A sum of FEM potentials over one vertex set, evaluated as one device model.

liblaf/apple's ``WarpModel`` sums its potentials (warp/model/_model.py:13-36),
and its inverse experiments put several FEM energies on one mesh: fat
(stable Neo-Hookean), aponeurosis and muscle (stable Neo-Hookean muscle), each
over every cell with its own Fraction (so its own dV) and materials
(exp/2026/01/28/smas/src/31-inverse-activation-stable-neo-hookean.py:256-279).

``DeviceModel`` concatenates the potentials' tets into one model mesh
(``apl_model_create``): one segment per potential, blocks never straddle a
segment, each material law runs one element launch over its own blocks, and
all of them add into one deterministic vertex reduction.  It has the
interface of a single potential (``mesh``, ``material`` = APL_MODEL,
``materials`` over all tets), so the fused device PNCG, the PCG adjoint,
``adjoint.equilibrium`` and the mixed-precision refinement take it as they
take one potential.  Per-part materials live in the parts; the model's
per-tet buffers are refreshed from them by ``sync_materials`` (in place, so
captured solver graphs stay valid), and ``split`` maps per-tet model arrays
back to the parts.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import torch

from . import _lib
from .runtime import DeviceMesh, as_dtype, normalize_device


class DeviceModel:
    material = _lib.APL_MODEL

    def __init__(self, parts: Sequence, n_points: int, *, dtype=None, device=None, chunks: int = 1,
                 name: str = "model"):
        parts = list(parts)
        if not parts:
            raise ValueError("a device model needs at least one FEM potential")
        self.parts = parts
        self.name = name
        self.dtype = as_dtype(dtype if dtype is not None else parts[0].dtype)
        self.device = normalize_device(device if device is not None else parts[0].device)
        self.n_points = int(n_points)
        self._chunks = int(chunks)
        data = [p.host_element_data() for p in parts]
        for p, (cells, _, _) in zip(parts, data):
            if cells.size and int(cells.max()) >= self.n_points:
                raise ValueError(f"potential {p.name!r} references vertices beyond the model's {n_points}")
        self._host = data
        counts = [d[0].shape[0] for d in data]
        self.tet_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.seg_law = np.array([int(p.material) for p in parts], np.int32)
        self.mesh = self._build(self.dtype)
        T = self.mesh.n_tets
        laws = set(self.seg_law.tolist())
        # the model's own per-tet buffers; ``materials`` points at them until a
        # caller (e.g. adjoint.equilibrium) binds other tensors
        self._buffers: dict[str, torch.Tensor] = {
            "mu": torch.zeros(T, dtype=self.dtype, device=self.device),
        }
        if laws & {_lib.APL_SNH, _lib.APL_SNH_MUSCLE}:
            self._buffers["lambda"] = torch.zeros(T, dtype=self.dtype, device=self.device)
        if _lib.APL_SNH_MUSCLE in laws:
            self._buffers["activation"] = torch.zeros(T, 6, dtype=self.dtype, device=self.device)
        self.materials: dict[str, torch.Tensor] = dict(self._buffers)
        self.requires_grad = tuple(sorted({k for p in parts for k in getattr(p, "requires_grad", ())}))
        self.sync_materials()

    def _build(self, dtype) -> DeviceMesh:
        cells = np.concatenate([d[0] for d in self._host]).astype(np.int32)
        dminv = np.concatenate([d[1] for d in self._host])
        dV = np.concatenate([d[2] for d in self._host])
        return DeviceMesh(cells, dminv, dV, self.n_points, dtype=dtype, device=self.device, chunks=self._chunks,
                          segments=(self.seg_law, self.tet_off))

    # ------------------------------------------------------------ potential-like view

    @property
    def n_cells(self) -> int:
        return self.mesh.n_tets

    @property
    def n_parts(self) -> int:
        return len(self.parts)

    def part_slice(self, k: int) -> slice:
        return slice(int(self.tet_off[k]), int(self.tet_off[k + 1]))

    def sync_materials(self) -> None:
        """Copy every part's current materials into the model's per-tet buffers
        (in place) and bind them as the model's materials."""
        buf = self._buffers
        for k, p in enumerate(self.parts):
            sl = self.part_slice(k)
            buf["mu"][sl].copy_(p.materials["mu"])
            if "lambda" in buf:
                lam = p.materials.get("lambda")
                if lam is not None and p.material != _lib.APL_ARAP:
                    buf["lambda"][sl].copy_(lam)
            if "activation" in buf and "activation" in p.materials:
                buf["activation"][sl].copy_(p.materials["activation"])
        self.materials.update(buf)

    def gather(self, per_part: Sequence[dict]) -> dict[str, torch.Tensor]:
        """Model-wide per-tet material tensors from per-part dicts, by torch.cat
        (autograd flows back to the parts' tensors).  Entries a part lacks
        (lambda of ARAP, the activation of non-muscle laws) are zeros."""
        out = {}
        for key in self._buffers:
            chunks = []
            for k, (p, d) in enumerate(zip(self.parts, per_part)):
                n = self.part_slice(k).stop - self.part_slice(k).start
                v = d.get(key)
                if v is None or (key == "lambda" and p.material == _lib.APL_ARAP):
                    shape = (n, 6) if key == "activation" else (n,)
                    v = torch.zeros(shape, dtype=self.dtype, device=self.device)
                chunks.append(v.to(self.dtype))
            out[key] = torch.cat(chunks).contiguous()
        return out

    def split(self, t: torch.Tensor) -> list[torch.Tensor]:
        """Per-part views of a per-tet model array."""
        return [t[self.part_slice(k)] for k in range(self.n_parts)]

    def fp64_shadow(self) -> "DeviceModel":
        """The model with FP64 element data from the parts' host data (for
        FP64 residual passes); built once."""
        if self.dtype == torch.float64:
            return self
        sh = getattr(self, "_shadow", None)
        if sh is None:
            sh = DeviceModel.__new__(DeviceModel)
            sh.parts, sh.name, sh.dtype, sh.device = self.parts, f"{self.name}_fp64", torch.float64, self.device
            sh.n_points, sh._chunks, sh._host = self.n_points, 1, self._host
            sh.tet_off, sh.seg_law = self.tet_off, self.seg_law
            sh.mesh = sh._build(torch.float64)
            sh._buffers = {k: v.double() for k, v in self._buffers.items()}
            sh.materials = dict(sh._buffers)
            sh.requires_grad = self.requires_grad
            self._shadow = sh
        return sh

    def mixed_derivative_prod(self, u: torch.Tensor, p: torch.Tensor) -> list[dict[str, torch.Tensor]]:
        """Per part, per-tet p^T d(grad)/d(material) (mu, lambda, activation
        as the part's law has them), from one mixed pass per law."""
        mats = self.materials
        dmu = torch.zeros(self.n_cells, dtype=self.dtype, device=self.device)
        dlam = torch.zeros_like(dmu)
        lam = mats.get("lambda")
        if "activation" in mats:
            dact = torch.zeros(self.n_cells, 6, dtype=self.dtype, device=self.device)
            self.mesh.set_activation(mats["activation"])
            self.mesh.mixed_activation(mats["mu"], lam, u, p, dact, dmu=dmu, dlam=dlam)
        else:
            dact = None
            self.mesh.eval(self.material, mats["mu"], lam, u, p, dmu=dmu, dlam=dlam)
        out = []
        for k, part in enumerate(self.parts):
            sl = self.part_slice(k)
            d = {"mu": dmu[sl]}
            if part.material != _lib.APL_ARAP:
                d["lambda"] = dlam[sl]
            if part.material == _lib.APL_SNH_MUSCLE:
                d["activation"] = dact[sl]
            out.append(d)
        return out
