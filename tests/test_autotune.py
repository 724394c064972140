"""Autotuner: signature and cache handling on the CPU; tuning and the tuned
batch on the GPU (a tuned schedule changes only the speed: the element grid
leaves every result bitwise unchanged, the block size within rounding)."""

import dataclasses
import json

import numpy as np
import pytest
import torch

from apple_amd import autotune
from apple_amd.mesh import cube


def test_signature_separates_meshes_dtypes_and_materials():
    a, b = cube(3), cube(4)
    s = lambda m, dt=torch.float32, mat=0, off=None: autotune.signature(m.cells, m.n_points, off, dt, mat, None)
    assert s(a) == s(a)
    assert len({s(a), s(b), s(a, torch.float64), s(a, mat=1), s(a, off=np.array([0, a.n_cells]))}) == 5


def test_schedule_round_trips_through_the_cache_file(tmp_path):
    sched = autotune.Schedule(chunks=2, grid_pct=150, us_per_iteration=12.5, table={"2/150": 12.5})
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"sig": dataclasses.asdict(sched)}))
    assert autotune.Schedule(**autotune._load(str(path))["sig"]) == sched
    assert autotune._load(str(tmp_path / "missing.json")) == {}


@pytest.mark.gpu
def test_tuned_batch_matches_default_schedule(tmp_path):
    from apple_amd import _lib, batch
    from apple_amd.mesh import Region
    from apple_amd.solver import PncgConfig

    insts = []
    for k, n in enumerate((4, 6, 5)):
        m = cube(n)
        region = Region.from_mesh(m)
        mu = 1e4 * np.exp(0.5 * np.random.default_rng(k).standard_normal(m.n_cells))
        f = np.zeros((m.n_points, 3))
        np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e3 * region.dV / 4, 4))
        fixed = np.zeros((m.n_points, 3), bool)
        fixed[m.points[:, 2] == 0] = True
        insts.append(batch.Instance(region, mu, 10 * mu, f, fixed, np.zeros((m.n_points, 3))))
    cache = str(tmp_path / "tune.json")
    b0 = batch.Batch(insts, _lib.APL_SNH, dtype=torch.float32, device="cuda")
    sched = autotune.tune(
        np.concatenate([i.region.cells + b0.vtx_off[k] for k, i in enumerate(insts)]),
        np.concatenate([i.region.dminv for i in insts]), np.concatenate([i.region.dV for i in insts]),
        int(b0.vtx_off[-1]), material=_lib.APL_SNH, mu=b0.mu, lam=b0.lam, dtype=torch.float32, device="cuda",
        inst_tet_off=b0.tet_off, inst_vtx_off=b0.vtx_off, iters=3, cache=cache)
    assert sched.chunks in autotune.CHUNKS and sched.grid_pct in autotune.GRID_PCT
    assert len(sched.table) == len(autotune.CHUNKS) * len(autotune.GRID_PCT)
    again = autotune.tune(
        np.concatenate([i.region.cells + b0.vtx_off[k] for k, i in enumerate(insts)]),
        np.concatenate([i.region.dminv for i in insts]), np.concatenate([i.region.dV for i in insts]),
        int(b0.vtx_off[-1]), material=_lib.APL_SNH, mu=b0.mu, lam=b0.lam, dtype=torch.float32, device="cuda",
        inst_tet_off=b0.tet_off, inst_vtx_off=b0.vtx_off, iters=3, cache=cache)
    assert again == sched  # served from the cache
    # the element grid never changes results: same chunks, other grid -> same bits
    cfg = PncgConfig(rtol=1e-6)
    bg = batch.Batch(insts, _lib.APL_SNH, dtype=torch.float32, device="cuda",
                     schedule=autotune.Schedule(chunks=1, grid_pct=200))
    x0, st0 = b0.solve(config=cfg)
    xg, stg = bg.solve(config=cfg)
    assert torch.equal(x0, xg) and [s.n_steps for s in st0] == [s.n_steps for s in stg]
    bt = batch.Batch(insts, _lib.APL_SNH, dtype=torch.float32, device="cuda", schedule=sched)
    xt, stt = bt.solve(config=cfg)
    assert all(s.success for s in stt)
    assert float((xt - x0).abs().max()) <= 1e-5 * np.sqrt(3.0)
