"""Autotuner of the element-pass schedule of a mesh or batch.

The schedule has two knobs: the block size (``chunks`` of 256 tets per
element block, fixed when the device mesh is built) and the persistent
element grid (a percentage of occupancy x compute units,
``apl_mesh_set_schedule``).  ``tune`` builds the mesh for each block size,
times the two element passes of a PNCG iteration (energy + gradient +
Hessian diagonal, and the quadratic form) under each grid, and keeps the
fastest.  Results are cached on disk per mesh signature and device, so a
batch of the same shapes is tuned once.

There is no counterpart in liblaf/apple (its Warp kernels launch one thread
per tet); this replaces that fixed schedule.
"""

from __future__ import annotations

import ctypes
import dataclasses
import hashlib
import json
import os
import time

import numpy as np
import torch

from . import _lib
from .runtime import DeviceMesh, current_stream, ptr

CHUNKS = (1, 2, 4)
GRID_PCT = (0, 100, 150, 200, 300)  # 0: the library's per-pass default grids
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "apple_amd", "autotune.json")


@dataclasses.dataclass
class Schedule:
    chunks: int = 1
    grid_pct: int = 0  # 0: the per-pass default grids
    us_per_iteration: float | None = None  # pass A + pass B element kernels at the pick
    table: dict | None = None  # "chunks/grid_pct" -> microseconds


def signature(cells, n_verts, inst_tet_off, dtype, material, device) -> str:
    """Mesh shape, batch layout, dtype, material and device model."""
    cells = np.ascontiguousarray(cells, np.int32)
    h = hashlib.sha1()
    h.update(np.int64([cells.shape[0], n_verts]).tobytes())
    step = max(1, cells.shape[0] // 4096)
    h.update(cells[::step].tobytes())
    if inst_tet_off is not None:
        h.update(np.ascontiguousarray(inst_tet_off, np.int64).tobytes())
    name = torch.cuda.get_device_name(device) if torch.cuda.is_available() else "cpu"
    return f"{name}|{str(dtype)}|{int(material)}|{h.hexdigest()}"


def _time_passes(mesh: DeviceMesh, material: int, mu, lam, u, iters: int) -> float:
    """Mean microseconds of one pass-A plus one pass-B element launch."""
    V = mesh.n_verts
    a = torch.empty(V, 3, dtype=mesh.dtype, device=mesh.device)
    b = torch.empty_like(a)
    e = torch.zeros(max(mesh.n_inst, 1), dtype=torch.float64, device=mesh.device)
    total = 0.0
    for what, p in ((_lib.APL_ENERGY | _lib.APL_GRAD | _lib.APL_HESS_DIAG, None), (_lib.APL_HESS_QUAD, u)):
        k = ctypes.c_double()
        _lib.check(
            _lib.lib().apl_bench_elem_pass(
                mesh._handle, int(material), what, ptr(mu), ptr(lam) if lam is not None else None, ptr(u), ptr(p),
                ptr(e), ptr(a), ptr(b), int(iters), ctypes.byref(k), None, current_stream(mesh.device),
            )
        )
        total += k.value * 1e3
    return total


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(cells, dminv, dV, n_verts, *, material, mu, lam=None, activation=None, dtype=torch.float32, device=None,
         inst_tet_off=None, inst_vtx_off=None, iters=10, cache: str | None = DEFAULT_CACHE,
         chunks=CHUNKS, grid_pct=GRID_PCT) -> Schedule:
    """Pick (chunks, grid_pct) for this mesh; mu / lam / activation are per-tet device tensors."""
    sig = signature(cells, n_verts, inst_tet_off, dtype, material, device)
    if cache:
        hit = _load(cache).get(sig)
        if hit:
            return Schedule(**hit)
    rng = np.random.default_rng(0)
    table = {}
    best = None
    for c in chunks:
        mesh = DeviceMesh(cells, dminv, dV, n_verts, dtype=dtype, device=device, inst_tet_off=inst_tet_off,
                          inst_vtx_off=inst_vtx_off, chunks=c)
        if activation is not None:
            mesh.set_activation(activation)
        u = torch.as_tensor(1e-3 * rng.standard_normal((n_verts, 3)), dtype=mesh.dtype, device=mesh.device)
        for g in grid_pct:
            mesh.set_schedule(g)
            _time_passes(mesh, material, mu, lam, u, 2)  # warm-up (first launch, grid cache)
            t = _time_passes(mesh, material, mu, lam, u, iters)
            table[f"{c}/{g}"] = round(t, 2)
            if best is None or t < best[2]:
                best = (c, g, t)
        mesh.close()
    sched = Schedule(chunks=best[0], grid_pct=best[1], us_per_iteration=round(best[2], 2), table=table)
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        data = _load(cache)
        data[sig] = dataclasses.asdict(sched)
        tmp = f"{cache}.{os.getpid()}.{time.time_ns()}"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1)
        os.replace(tmp, cache)
    return sched
