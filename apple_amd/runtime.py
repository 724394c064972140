"""Device-side mesh handle and tensor plumbing around the C ABI.

PyTorch-ROCm provides device memory and streams; every computation on the
hot path runs in the HIP kernels of ``libapple_amd.so``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

_DT = {torch.float32: _lib.APL_F32, torch.float64: _lib.APL_F64}


def as_dtype(dtype) -> torch.dtype:
    if dtype in (torch.float32, "float32", np.float32, "f32"):
        return torch.float32
    if dtype in (torch.float64, "float64", np.float64, "f64"):
        return torch.float64
    raise ValueError(f"unsupported dtype {dtype!r}: use float32 or float64")


def ptr(t: torch.Tensor | None, *, dtype=None, device=None, shape=None, name="tensor") -> int | None:
    """Raw device pointer of a contiguous tensor, after checking it."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device (cuda) tensor")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} has dtype {t.dtype}, expected {dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t.data_ptr()


def normalize_device(device) -> torch.device:
    """torch.device with an explicit index (``cuda`` -> ``cuda:<current>``)."""
    dev = torch.device(device) if device is not None else torch.device("cuda")
    if dev.type != "cuda":
        raise ValueError("apple_amd runs on a ROCm device; got " + str(dev))
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def current_stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class DeviceMesh:
    """Owner of an ``apl_mesh`` handle: element data and the block partition.

    ``cells`` (T,4) int, ``dminv`` (T,3,3) = inverse of dX/dr, ``dV`` (T,),
    all host arrays; ``inst_tet_off`` / ``inst_vtx_off`` describe a batch of
    independent meshes concatenated into one (tet and vertex ranges).  With
    ``points`` (V,3) (and optionally ``fraction`` (T,)) instead of dminv / dV,
    the element data are computed on the device (``apl_mesh_create_from_points``).
    """

    def __init__(
        self,
        cells,
        dminv,
        dV,
        n_verts: int,
        *,
        dtype=torch.float32,
        device=None,
        inst_tet_off=None,
        inst_vtx_off=None,
        chunks: int = 1,
        points=None,
        fraction=None,
        segments=None,
    ):
        self.dtype = as_dtype(dtype)
        self.device = normalize_device(device)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        T = cells.shape[0]
        if points is not None:
            points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            if points.shape[0] != int(n_verts):
                raise ValueError("points must be (n_verts, 3)")
            if fraction is not None:
                fraction = np.ascontiguousarray(fraction, dtype=np.float64).reshape(-1)
                if fraction.shape[0] != T:
                    raise ValueError("fraction must have one value per tet")
            if cells.shape != (T, 4):
                raise ValueError("cells must be (T,4)")
        else:
            dminv = np.ascontiguousarray(dminv, dtype=np.float64).reshape(-1, 9)
            dV = np.ascontiguousarray(dV, dtype=np.float64).reshape(-1)
            if cells.shape != (T, 4) or dminv.shape[0] != T or dV.shape[0] != T:
                raise ValueError("cells (T,4), dminv (T,3,3) and dV (T,) must agree on T")
        self.n_tets = T
        self.n_verts = int(n_verts)
        if inst_tet_off is None:
            self.n_inst = 1
            self.inst_tet_off = np.array([0, T], np.int64)
            self.inst_vtx_off = np.array([0, self.n_verts], np.int64)
        else:
            self.inst_tet_off = np.ascontiguousarray(inst_tet_off, dtype=np.int64)
            self.inst_vtx_off = np.ascontiguousarray(inst_vtx_off, dtype=np.int64)
            self.n_inst = self.inst_tet_off.size - 1
        self._handle = ctypes.c_void_p()
        L = _lib.lib()
        common = (self.n_inst, self.inst_tet_off.ctypes.data, self.inst_vtx_off.ctypes.data, int(chunks),
                  self.device.index, ctypes.byref(self._handle))
        self.segments = None
        if segments is not None:
            # a model: (seg_law (n_seg,), seg_tet_off (n_inst * n_seg + 1,))
            seg_law = np.ascontiguousarray(segments[0], dtype=np.int32)
            seg_off = np.ascontiguousarray(segments[1], dtype=np.int64)
            self.segments = (seg_law, seg_off)
            _lib.check(
                L.apl_model_create(
                    _DT[self.dtype], T, self.n_verts, cells.ctypes.data, dminv.ctypes.data, dV.ctypes.data,
                    self.n_inst, self.inst_tet_off.ctypes.data, self.inst_vtx_off.ctypes.data, int(seg_law.size),
                    seg_law.ctypes.data, seg_off.ctypes.data, int(chunks), self.device.index,
                    ctypes.byref(self._handle),
                )
            )
        elif points is not None:
            _lib.check(
                L.apl_mesh_create_from_points(
                    _DT[self.dtype], T, self.n_verts, cells.ctypes.data, points.ctypes.data,
                    fraction.ctypes.data if fraction is not None else None, *common,
                )
            )
        else:
            _lib.check(
                L.apl_mesh_create(
                    _DT[self.dtype], T, self.n_verts, cells.ctypes.data, dminv.ctypes.data, dV.ctypes.data, *common,
                )
            )
        self._scratch = torch.zeros(max(self.n_inst, 1) * 2, dtype=torch.float64, device=self.device)

    @property
    def handle(self) -> int:
        return self._handle.value

    def set_schedule(self, grid_pct: int) -> None:
        """Persistent element grid as a percentage of occupancy x CUs (autotuner knob)."""
        _lib.check(_lib.lib().apl_mesh_set_schedule(self._handle, int(grid_pct)))
        self.grid_pct = int(grid_pct)

    def set_activation(self, activation: torch.Tensor | None) -> None:
        """Bind the (T, 6) activation that muscle-law launches read (the mesh
        keeps the pointer; this object keeps the tensor alive)."""
        if activation is not None:
            ptr(activation, dtype=self.dtype, device=self.device, shape=(self.n_tets, 6), name="activation")
        _lib.check(_lib.lib().apl_mesh_set_activation(self._handle, activation.data_ptr() if activation is not None else None))
        self._activation = activation

    def region(self) -> tuple[torch.Tensor, torch.Tensor]:
        """The element data in tet order: (dminv (T,3,3), dV (T,)) device tensors."""
        dminv = torch.empty(self.n_tets, 3, 3, dtype=self.dtype, device=self.device)
        dV = torch.empty(self.n_tets, dtype=self.dtype, device=self.device)
        _lib.check(_lib.lib().apl_mesh_region(self._handle, ptr(dminv), ptr(dV), current_stream(self.device)))
        return dminv, dV

    def info(self) -> dict:
        inf = _lib.MeshInfo()
        _lib.check(_lib.lib().apl_mesh_info(self._handle, ctypes.byref(inf)))
        return {k: getattr(inf, k) for k, _ in _lib.MeshInfo._fields_}

    def layout(self) -> dict[str, np.ndarray]:
        """The block partition's arrays (``apl_mesh_layout``), read back to the host."""
        inf = self.info()
        T, V, nb, ni, nvb = inf["n_tets"], inf["n_verts"], inf["n_blocks"], inf["n_inst"], inf["n_vertex_blocks"]
        L = _lib.lib()

        def get(name, n):
            which, dt = _lib.LAYOUT_ARRAYS[name]
            a = np.empty(n, dtype=dt)
            _lib.check(L.apl_mesh_layout(self._handle, which, a.ctypes.data, int(n)))
            return a

        out = {"blk_co_off": get("blk_co_off", nb + 1)}
        sizes = {"corner": 8 * T, "blk_tet_off": nb + 1, "blk_loc_off": nb + 1, "blk_inst": nb,
                 "inst_blk_off": ni + 1, "loc_gid": inf["n_partials"], "chunk_off": int(out["blk_co_off"][-1]),
                 "vtx_slot_off": V + 1, "loc_dst": inf["n_partials"], "vblk_off": nvb + 1, "vblk_inst": nvb,
                 "inst_vblk_off": ni + 1}
        for name, n in sizes.items():
            out[name] = get(name, n)
        return out

    def lumped_volume(self, density: torch.Tensor | None = None) -> torch.Tensor:
        """Per-vertex sum of density_t dV_t / 4 over the vertex's tets (FP64
        device tensor (V,)), summed in a fixed order on the device."""
        out = torch.empty(self.n_verts, dtype=torch.float64, device=self.device)
        dens = None
        if density is not None:
            dens = ptr(density, dtype=self.dtype, device=self.device, shape=(self.n_tets,), name="density")
        _lib.check(_lib.lib().apl_mesh_lumped_volume(self._handle, dens, ptr(out), current_stream(self.device)))
        return out


    def close(self) -> None:
        if getattr(self, "_handle", None) is not None and self._handle.value:
            _lib.lib().apl_mesh_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ evaluation

    def vec(self, name="u"):
        return dict(dtype=self.dtype, device=self.device, shape=(self.n_verts, 3), name=name)

    def eval(
        self,
        material: int,
        mu,
        lam,
        u,
        p=None,
        *,
        energy=None,
        grad=None,
        hess_diag=None,
        hess_prod=None,
        hess_quad=None,
        dmu=None,
        dlam=None,
        accumulate=False,
    ) -> None:
        """One fused element pass writing (or adding into) the given outputs."""
        what = 0
        if energy is not None:
            what |= _lib.APL_ENERGY
        if grad is not None:
            what |= _lib.APL_GRAD
        if hess_diag is not None:
            what |= _lib.APL_HESS_DIAG
        if hess_prod is not None:
            what |= _lib.APL_HESS_PROD
        if hess_quad is not None:
            what |= _lib.APL_HESS_QUAD
        if dmu is not None:
            what |= _lib.APL_MIXED
        if accumulate:
            what |= _lib.APL_ACCUMULATE
        tet = dict(dtype=self.dtype, device=self.device, shape=(self.n_tets,))
        sc = dict(dtype=torch.float64, device=self.device, shape=(self.n_inst,))
        _lib.check(
            _lib.lib().apl_elem_eval(
                self._handle,
                int(material),
                ptr(mu, name="mu", **tet),
                ptr(lam, name="lambda", **tet) if (material != _lib.APL_ARAP and lam is not None) else None,
                ptr(u, **self.vec("u")),
                ptr(p, **self.vec("p")),
                what,
                ptr(energy, name="energy", **sc),
                ptr(grad, **self.vec("grad")),
                ptr(hess_diag, **self.vec("hess_diag")),
                ptr(hess_prod, **self.vec("hess_prod")),
                ptr(hess_quad, name="hess_quad", **sc),
                ptr(dmu, name="dmu", **tet),
                ptr(dlam, name="dlam", **tet),
                current_stream(self.device),
            )
        )

    def mixed_activation(self, mu, lam, u, p, dact, *, dmu=None, dlam=None, accumulate=False) -> None:
        """Muscle law: per-tet p^T d(grad)/d(activation) into dact (T, 6), and
        optionally p^T d(grad)/d(mu), d/d(lambda) in the same pass."""
        tet = dict(dtype=self.dtype, device=self.device, shape=(self.n_tets,))
        _lib.check(
            _lib.lib().apl_elem_mixed_activation(
                self._handle,
                ptr(mu, name="mu", **tet),
                ptr(lam, name="lambda", **tet),
                ptr(u, **self.vec("u")),
                ptr(p, **self.vec("p")),
                ptr(dact, dtype=self.dtype, device=self.device, shape=(self.n_tets, 6), name="dact"),
                ptr(dmu, name="dmu", **tet),
                ptr(dlam, name="dlam", **tet),
                1 if accumulate else 0,
                current_stream(self.device),
            )
        )

    def ext_force(self, indices, force, u, *, energy=None, grad=None, accumulate=True) -> None:
        n = int(indices.shape[0])
        _lib.check(
            _lib.lib().apl_ext_force_eval(
                self._handle,
                n,
                ptr(indices, dtype=torch.int32, device=self.device, name="indices"),
                ptr(force, dtype=self.dtype, device=self.device, shape=(n, 3), name="force"),
                ptr(u, **self.vec("u")) if u is not None else None,
                ptr(energy, dtype=torch.float64, device=self.device, shape=(1,), name="energy"),
                ptr(grad, **self.vec("grad")),
                1 if accumulate else 0,
                current_stream(self.device),
            )
        )


def layout_mode(mode: int) -> int:
    """Where later meshes build their block partition: ``_lib.LAYOUT_DEVICE``
    (default) or ``_lib.LAYOUT_HOST``; returns the previous mode."""
    prev = _lib.lib().apl_layout_mode(int(mode))
    if prev < 0:
        _lib.check(prev)
    return prev


def morton_order(points: torch.Tensor, cells: torch.Tensor | None = None) -> tuple[torch.Tensor, torch.Tensor | None]:
    """(vertex_perm, cell_perm) of the Morton order on the device
    (``apl_morton_order``): points (V,3) FP64, cells (T,4) int32, both on
    one GPU; new = old[perm]."""
    dev = points.device
    points = points.to(torch.float64).contiguous()
    V = points.shape[0]
    vperm = torch.empty(V, dtype=torch.int32, device=dev)
    cperm = None
    T = 0
    cptr = None
    if cells is not None:
        cells = cells.to(device=dev, dtype=torch.int32).contiguous()
        T = cells.shape[0]
        cperm = torch.empty(T, dtype=torch.int32, device=dev)
        cptr = cells.data_ptr() if T else None
    _lib.check(_lib.lib().apl_morton_order(points.data_ptr() if V else None, V, cptr, T,
                                           vperm.data_ptr() if V else None,
                                           cperm.data_ptr() if cperm is not None and T else None, dev.index,
                                           current_stream(dev)))
    return vperm, cperm
