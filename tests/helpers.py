"""Shared test inputs: seeded cube meshes, materials, displacement fields."""

import json
import os

import numpy as np

import oracle.fem as of
from apple_amd.mesh import Region, cube


def snh_case(n, seed=0, noise=0.01, mu_sigma=0.5, material="snh"):
    """Cube n^3 x 5 tets, lognormal mu (median 1e4), lambda = 10 mu, noisy u, random v."""
    m = cube(n)
    rng = np.random.default_rng(seed)
    mu = 1e4 * np.exp(mu_sigma * rng.standard_normal(m.n_cells))
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = 10.0 * mu
    region = Region.from_mesh(m)
    u = noise * rng.standard_normal((m.n_points, 3))
    v = rng.standard_normal((m.n_points, 3))
    return m, region, u, v


def oracle_pot(region, material, mu, lam, dtype_round=None):
    """Oracle potential on FP64 region data with (optionally rounded) materials."""
    mu = np.asarray(mu, np.float64)
    lam = None if lam is None else np.asarray(lam, np.float64)
    if dtype_round is not None:
        mu = mu.astype(dtype_round).astype(np.float64)
        lam = None if lam is None else lam.astype(dtype_round).astype(np.float64)
    mat = of.SNH if material == "snh" else of.ARAP
    return of.FemPotential(mat, region.cells, region.dhdX, region.dV, mu, lam if mat == of.SNH else None)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def rel_max(a, b):
    """Maximum over entries: max |a - b| / max |b|."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    mb = np.abs(b).max() if b.size else 0.0
    return float(np.abs(a - b).max() / (mb if mb > 0 else 1.0)) if b.size else 0.0


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def patch_test():
    """The reference's known-answer patch test as data (golden/patch_test.json):
    points (5,3), cells (4,4), fixed mask / values (5,3) and the expected answer."""
    with open(os.path.join(GOLDEN, "patch_test.json")) as f:
        d = json.load(f)
    points = np.asarray(d["points"], np.float64)
    cells = np.asarray(d["cells"], np.int32)
    mask = np.zeros(points.shape, bool)
    values = np.zeros(points.shape)
    for v, val in zip(d["fixed_vertices"], d["fixed_values"]):
        mask[v] = True
        values[v] = val
    return points, cells, mask, values, d
