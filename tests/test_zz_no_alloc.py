"""apl_elem_eval_static (the entry point that allocates nothing) against
apl_elem_eval: same kernels, scratch in a static per-GPU array instead of the
mesh's workspace, so every output must be bitwise equal.  Two calls in a
row, the second on a smaller mesh (the array keeps the first call's
contents); a mesh whose scratch exceeds the array is rejected."""

import ctypes

import numpy as np
import pytest
import torch

from apple_amd import _lib, fem
from apple_amd.mesh import cube
from helpers import snh_case

pytestmark = pytest.mark.gpu

E, G, H, HP, HQ, MIX = (_lib.APL_ENERGY, _lib.APL_GRAD, _lib.APL_HESS_DIAG, _lib.APL_HESS_PROD,
                        _lib.APL_HESS_QUAD, _lib.APL_MIXED)


def _call(entry, pot, u, v, what):
    """One call through the C ABI; returns every output as a host array."""
    m = pot.mesh
    dev, dt, V, T = m.device, m.dtype, m.n_verts, m.n_tets
    e = torch.full((1,), 7.0, dtype=torch.float64, device=dev)
    q = torch.full((1,), 7.0, dtype=torch.float64, device=dev)
    g = torch.full((V, 3), 7.0, dtype=dt, device=dev)
    h = torch.full_like(g, 7.0)
    hp = torch.full_like(g, 7.0)
    dmu = torch.full((T,), 7.0, dtype=dt, device=dev)
    dlam = torch.full_like(dmu, 7.0)
    lam = pot.materials.get("lambda")
    p = lambda t: t.data_ptr() if t is not None else None
    rc = entry(m.handle, int(pot.material), p(pot.materials["mu"]), p(lam) if pot.material != _lib.APL_ARAP else None,
               p(u), p(v), what, p(e), p(g), p(h), p(hp), p(q), p(dmu), p(dlam),
               torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    return {k: t.cpu().numpy() for k, t in dict(e=e, q=q, g=g, h=h, hp=hp, dmu=dmu, dlam=dlam).items()}


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("material", ["snh", "arap"])
def test_static_scratch_entry_matches_elem_eval(dtype, material):
    L = _lib.lib()
    cls = fem.StableNeoHookean if material == "snh" else fem.Arap
    for n in (14, 5):  # the second call on a smaller mesh
        print(f"[no-alloc] material={material} dtype={dtype} n={n}", flush=True)
        m, region, u64, v64 = snh_case(n, seed=30 + n, noise=0.01)
        pot = cls.from_region(region, dtype=dtype, device="cuda")
        u = torch.as_tensor(u64, dtype=dtype, device="cuda")
        v = torch.as_tensor(v64, dtype=dtype, device="cuda")
        for what in (E | G | H, G, H, E | H, HP, HQ, MIX, E | G | H | HP | HQ | MIX):
            ref = _call(L.apl_elem_eval, pot, u, v, what)
            got = _call(L.apl_elem_eval_static, pot, u, v, what)
            for k in ref:
                assert np.array_equal(got[k], ref[k]), (what, k)
        # accumulate semantics are the same too
        what = E | G | H | _lib.APL_ACCUMULATE
        assert all(np.array_equal(a, b) for a, b in zip(_call(L.apl_elem_eval, pot, u, v, what).values(),
                                                        _call(L.apl_elem_eval_static, pot, u, v, what).values()))


def test_static_scratch_rejects_meshes_beyond_its_size():
    L = _lib.lib()
    cap = int(L.apl_static_scratch_bytes())
    assert cap == 32 << 20
    m = cube(80)  # 2.56M tets: about 1.2M slot partials x 6 channels x 8 B > 32 MiB in FP64
    m.cell_data["mu"] = np.full(m.n_cells, 1e4)
    m.cell_data["lambda"] = np.full(m.n_cells, 1e5)
    pot = fem.StableNeoHookean.from_points(m, dtype=torch.float64, device="cuda")
    u = torch.zeros(m.n_points, 3, dtype=torch.float64, device="cuda")
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = torch.zeros_like(u)
    rc = L.apl_elem_eval_static(pot.mesh.handle, _lib.APL_SNH, pot.materials["mu"].data_ptr(),
                                pot.materials["lambda"].data_ptr(), u.data_ptr(), None, E | G, e.data_ptr(),
                                g.data_ptr(), None, None, None, None, None, 0)
    assert rc == -1
    assert b"apl_elem_eval_static holds" in L.apl_last_error()
    # the regular entry point still evaluates it
    pot.fused(u, energy=e, grad=g, accumulate=False)
    torch.cuda.synchronize()
    assert e.item() == 0.0
