"""The setup path on the device: the block partition built by kernels
(partition.hip) against the host pass (layout.cpp), every array bit for bit;
the Morton order against its numpy restatement; the lumped vertex volumes
against numpy's np.add.at over the same dV."""

import numpy as np
import pytest
import torch

from apple_amd import _lib, fem, potential
from apple_amd.mesh import Mesh, cube, reorder, reorder_host
from apple_amd.runtime import DeviceMesh, layout_mode

pytestmark = pytest.mark.gpu


def _build(cells, V, *, dtype=torch.float32, chunks=1, inst=None, segments=None):
    T = cells.shape[0]
    dminv = np.tile(np.eye(3), (T, 1, 1))
    dV = np.full(T, 1.0 / 6.0)
    kw = dict(inst_tet_off=inst[0], inst_vtx_off=inst[1]) if inst is not None else {}
    return DeviceMesh(cells, dminv, dV, V, dtype=dtype, device="cuda", chunks=chunks, segments=segments, **kw)


def _host_and_device(*args, **kw):
    prev = layout_mode(_lib.LAYOUT_HOST)
    try:
        host = _build(*args, **kw)
    finally:
        layout_mode(prev)
    return _build(*args, **kw), host


def _assert_same(dev, host):
    a, b = dev.layout(), host.layout()
    assert a.keys() == b.keys()
    for k in b:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    ia, ib = dev.info(), host.info()
    for k in ("n_blocks", "n_partials", "n_shared_verts", "n_vertex_blocks", "nloc_max", "chunks"):
        assert ia[k] == ib[k], k
    return ia


def _batch(ns):
    cells, toff, voff = [], [0], [0]
    for n in ns:
        m = cube(n)
        cells.append(m.cells + voff[-1])
        toff.append(toff[-1] + m.n_cells)
        voff.append(voff[-1] + m.n_points)
    return np.concatenate(cells).astype(np.int32), voff[-1], (np.array(toff), np.array(voff))


@pytest.mark.parametrize("dtype,chunks", [(torch.float32, 1), (torch.float32, 2), (torch.float32, 4),
                                          (torch.float64, 1)])
def test_cube_partition_equals_host(dtype, chunks):
    m = cube(12)
    print(f"[partition] cube 12 {dtype} chunks {chunks}", flush=True)
    info = _assert_same(*_host_and_device(m.cells, m.n_points, dtype=dtype, chunks=chunks))
    assert info["n_blocks"] > 1


@pytest.mark.parametrize("dtype,chunks", [(torch.float32, 1), (torch.float32, 2), (torch.float64, 1)])
def test_shuffled_and_lex_meshes_equal_host(dtype, chunks):
    # tets in random order: windows far above the cap, the greedy block ends bind
    m = cube(14)
    cells = m.cells[np.random.default_rng(0).permutation(m.n_cells)]
    print(f"[partition] shuffled cube 14 {dtype} chunks {chunks}", flush=True)
    info = _assert_same(*_host_and_device(cells, m.n_points, dtype=dtype, chunks=chunks))
    if dtype == torch.float64:
        assert info["nloc_max"] > 256  # the FP64 slot cap decides the block ends
    lex = cube(9, order="lex")
    _assert_same(*_host_and_device(lex.cells, lex.n_points, dtype=dtype, chunks=chunks))


@pytest.mark.parametrize("dtype,chunks", [(torch.float32, 1), (torch.float32, 2), (torch.float64, 1)])
def test_batch_and_model_segments_equal_host(dtype, chunks):
    cells, V, inst = _batch([2, 5, 1, 7, 3])
    print(f"[partition] batch of 5 cubes {dtype} chunks {chunks}", flush=True)
    _assert_same(*_host_and_device(cells, V, dtype=dtype, chunks=chunks, inst=inst))
    # a model: every instance's tets twice (two potentials), segments of unequal laws
    toff, voff = inst
    mc, seg = [], [0]
    for i in range(len(toff) - 1):
        c = cells[toff[i]:toff[i + 1]]
        mc += [c, c[: max(1, c.shape[0] // 3)]]
        seg += [seg[-1] + c.shape[0], seg[-1] + c.shape[0] + max(1, c.shape[0] // 3)]
    mcells = np.concatenate(mc).astype(np.int32)
    mtoff = np.array(seg[::2])
    segments = (np.array([_lib.APL_SNH, _lib.APL_SNH_MUSCLE], np.int32), np.array(seg, np.int64))
    _assert_same(*_host_and_device(mcells, V, dtype=dtype, chunks=chunks, inst=(mtoff, voff), segments=segments))


def test_small_and_degenerate_meshes_equal_host():
    print("[partition] small meshes", flush=True)
    one = np.array([[0, 1, 2, 3]], np.int32)
    _assert_same(*_host_and_device(one, 4))
    _assert_same(*_host_and_device(one, 9))  # vertices no tet touches
    dup = np.array([[0, 0, 1, 2], [2, 3, 3, 3], [1, 2, 3, 4]], np.int32)  # repeated corners
    _assert_same(*_host_and_device(dup, 5))
    _assert_same(*_host_and_device(dup, 5, dtype=torch.float64))
    empty = np.zeros((0, 4), np.int32)
    _assert_same(*_host_and_device(empty, 3))


def test_out_of_range_cells_rejected_like_host():
    cells, V, inst = _batch([2, 2])
    bad = cells.copy()
    bad[3, 1] = V  # outside [0, V)
    msgs = []
    for mode in (_lib.LAYOUT_HOST, _lib.LAYOUT_DEVICE):
        prev = layout_mode(mode)
        try:
            with pytest.raises(RuntimeError) as e:
                _build(bad, V)
            msgs.append(str(e.value))
        finally:
            layout_mode(prev)
    assert "out of range" in msgs[1]
    bad = cells.copy()
    bad[45, 2] = 1  # instance 1's tet on instance 0's vertex
    msgs = []
    for mode in (_lib.LAYOUT_HOST, _lib.LAYOUT_DEVICE):
        prev = layout_mode(mode)
        try:
            with pytest.raises(RuntimeError) as e:
                _build(bad, V, inst=inst)
            msgs.append(str(e.value))
        finally:
            layout_mode(prev)
    assert msgs[0] == msgs[1] and "outside its instance" in msgs[1]


def _irregular(n, seed):
    m = cube(n)
    rng = np.random.default_rng(seed)
    vp = rng.permutation(m.n_points)
    rank = np.empty_like(vp)
    rank[vp] = np.arange(vp.size)
    pts = (m.points + 0.2 / n * rng.standard_normal(m.points.shape))[vp]
    cells = rank[m.cells][rng.permutation(m.n_cells)].astype(np.int32)
    mesh = Mesh(pts, cells)
    mesh.cell_data["mu"] = np.arange(mesh.n_cells, dtype=np.float64)
    return mesh


def test_morton_order_equals_numpy():
    print("[partition] Morton order", flush=True)
    for n, seed in ((7, 1), (16, 2)):
        mesh = _irregular(n, seed)
        new_d, vp_d, cp_d = reorder(mesh)
        new_h, vp_h, cp_h = reorder_host(mesh)
        np.testing.assert_array_equal(vp_d, vp_h)
        np.testing.assert_array_equal(cp_d, cp_h)
        np.testing.assert_array_equal(new_d.cells, new_h.cells)
        np.testing.assert_array_equal(new_d.cell_data["mu"], new_h.cell_data["mu"])
    # the reordered mesh blocks with far fewer slots than the shuffled one
    mesh = _irregular(16, 3)
    shuffled = _build(mesh.cells, mesh.n_points).info()["n_partials"]
    ordered = _build(reorder(mesh)[0].cells, mesh.n_points).info()["n_partials"]
    print(f"  slots: shuffled {shuffled}, Morton {ordered}")
    assert ordered < 0.6 * shuffled


def test_morton_order_rejects_out_of_range_cells():
    """A cell index outside [0, V) is an error that names the first bad tet;
    the kernel reads no point for it."""
    from apple_amd.runtime import morton_order

    print("[partition] Morton order, out-of-range cells", flush=True)
    mesh = _irregular(5, 7)
    pts = torch.as_tensor(mesh.points, device="cuda")
    for t, c, v in ((17, 2, mesh.n_points), (40, 0, -3), (3, 3, 1 << 30)):
        cells = mesh.cells.copy()
        cells[t, c] = v
        with pytest.raises(RuntimeError, match=f"out of range.*tet {t}"):
            morton_order(pts, torch.as_tensor(cells, device="cuda"))
    with pytest.raises(RuntimeError, match="out of range"):
        morton_order(pts[:0], torch.as_tensor(mesh.cells[:2], device="cuda"))
    # the same call with valid cells still works afterwards
    vp, cp = morton_order(pts, torch.as_tensor(mesh.cells, device="cuda"))
    np.testing.assert_array_equal(vp.cpu().numpy(), reorder_host(mesh)[1])
    np.testing.assert_array_equal(cp.cpu().numpy(), reorder_host(mesh)[2])


@pytest.mark.parametrize("dtype,chunks", [(torch.float64, 1), (torch.float32, 2)])
def test_lumped_volume_equals_add_at(dtype, chunks):
    print(f"[partition] lumped volume {dtype} chunks {chunks}", flush=True)
    mesh = _irregular(9, 4)
    mesh.cell_data["Fraction"] = np.random.default_rng(5).uniform(0.5, 1.0, mesh.n_cells)
    dm = DeviceMesh(mesh.cells, None, None, mesh.n_points, dtype=dtype, device="cuda", points=mesh.points,
                    fraction=mesh.cell_data["Fraction"], chunks=chunks)
    rho = torch.as_tensor(np.random.default_rng(6).uniform(500, 2000, mesh.n_cells), dtype=dtype, device="cuda")
    vol = dm.lumped_volume(rho)
    vol2 = dm.lumped_volume(rho)
    assert torch.equal(vol, vol2)  # a fixed summation order
    _, dV = dm.region()
    ref = np.zeros(mesh.n_points)
    np.add.at(ref, mesh.cells.reshape(-1), np.repeat(rho.double().cpu().numpy() * dV.double().cpu().numpy() / 4.0, 4))
    np.testing.assert_allclose(vol.cpu().numpy(), ref, rtol=1e-13, atol=0)
    # the total is the mass of the mesh
    assert float(vol.sum()) == pytest.approx(float((rho.double() * dV.double()).sum()), rel=1e-12)
    # gravity of a device-set-up potential: the same sums times g
    mesh.cell_data["mu"] = np.ones(mesh.n_cells)
    mesh.cell_data["lambda"] = 10.0 * np.ones(mesh.n_cells)
    pot = fem.StableNeoHookean.from_pyvista(mesh, setup="device", dtype=dtype, device="cuda", chunks=chunks)
    grav = potential.gravity_from_potential(pot, rho.double().cpu().numpy(), g=(0.0, 0.0, -9.8))
    f = grav.dense().double().cpu().numpy()
    np.testing.assert_allclose(f[:, 2], -9.8 * ref, rtol=1e-6 if dtype == torch.float32 else 1e-13)
    assert not f[:, :2].any()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_instance_without_tets_equals_host_and_passes_run(dtype):
    """A batch whose middle instance has vertices but no tets: both
    partitions agree, that instance owns vertex blocks but no element block,
    and an element pass gives it zero energy and a gradient of zero."""
    print(f"[partition] instance without tets {dtype}", flush=True)
    a, b = cube(2), cube(3)
    n_empty = 5
    cells = np.concatenate([a.cells, b.cells + a.n_points + n_empty]).astype(np.int32)
    V = a.n_points + n_empty + b.n_points
    toff = np.array([0, a.n_cells, a.n_cells, a.n_cells + b.n_cells])
    voff = np.array([0, a.n_points, a.n_points + n_empty, V])
    dev, host = _host_and_device(cells, V, dtype=dtype, inst=(toff, voff))
    _assert_same(dev, host)
    lay = dev.layout()
    assert 1 not in lay["blk_inst"]
    assert lay["inst_blk_off"][1] == lay["inst_blk_off"][2]
    assert lay["inst_vblk_off"][2] > lay["inst_vblk_off"][1]
    u = torch.as_tensor(0.01 * np.random.default_rng(1).standard_normal((V, 3)), dtype=dtype, device="cuda")
    mu = torch.full((cells.shape[0],), 1e3, dtype=dtype, device="cuda")
    lam = 10 * mu
    e = torch.zeros(3, dtype=torch.float64, device="cuda")
    g = torch.full((V, 3), 5.0, dtype=dtype, device="cuda")
    dev.eval(_lib.APL_SNH, mu, lam, u, energy=e, grad=g)
    torch.cuda.synchronize()
    assert e[1].item() == 0.0 and e[0].item() > 0.0 and e[2].item() > 0.0
    assert torch.count_nonzero(g[a.n_points:a.n_points + n_empty]).item() == 0
