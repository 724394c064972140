"""On-disk meshes and results: VTK XML UnstructuredGrid (``.vtu``) and
ParaView file series (``.vtu.series``).

liblaf/apple reads its inputs and writes its results as ``.vtu`` files
through pyvista (``melon.load_unstructured_grid`` / ``melon.save`` in
exp/2026/03/11/test-inverse/src/10-gen.py:14-17 and 20-forward.py:21-60),
keyed by the attribute names of ``common/attr_name.py:4-44`` (``mu``,
``lambda``, ``FixedMask``, ``FixedValue``, ``GlobalPointId``; the muscle
activation under the enum value ``activation``, warp/fem/utils/_material.py:16).  This
module reads and writes the same files without VTK: ``load_vtu`` returns an
``apple_amd.mesh.Mesh`` (points, tetra cells, point and cell data), which
``from_pyvista`` / ``from_mesh`` of every potential accept, and ``save_vtu``
writes one back, e.g. with the solved displacement added as point data.

Format, as the VTK file-format documentation fixes it:

* ``<VTKFile type="UnstructuredGrid" byte_order="LittleEndian"
  header_type="UInt32|UInt64" [compressor="vtkZLibDataCompressor"]>`` with
  one ``<Piece NumberOfPoints NumberOfCells>`` holding ``PointData``,
  ``CellData``, ``Points`` and ``Cells`` (``connectivity``, ``offsets`` = end
  of each cell's run, ``types``; VTK_TETRA = 10);
* a ``DataArray`` is ``format="ascii"`` (whitespace-separated values),
  ``"binary"`` (inline base64) or ``"appended"`` (``offset`` into the
  ``<AppendedData encoding="raw|base64">`` section after its ``_`` marker);
* binary payloads: uncompressed, one header integer (the byte count) then the
  bytes, base64-encoded together; zlib-compressed, the header
  ``[n_blocks, block_size, last_block_size (0 = full), compressed sizes...]``
  then the compressed blocks, base64-encoded separately (header, then data).

Bool arrays are stored as UInt8 (VTK has no bool type).  Only tetra meshes
are read: other cell types raise ``ValueError``.
"""

from __future__ import annotations

import base64
import json
import os
import re
import xml.etree.ElementTree as ET
import zlib
from collections.abc import Iterator
from xml.sax.saxutils import quoteattr
from typing import Any

import numpy as np

from .common import DISPLACEMENT, GLOBAL_POINT_ID
from .mesh import Mesh

VTK_TETRA = 10

_VTK_TYPES = {
    "Int8": np.int8, "UInt8": np.uint8, "Int16": np.int16, "UInt16": np.uint16,
    "Int32": np.int32, "UInt32": np.uint32, "Int64": np.int64, "UInt64": np.uint64,
    "Float32": np.float32, "Float64": np.float64,
}
_NP_TO_VTK = {np.dtype(v).newbyteorder("<"): k for k, v in _VTK_TYPES.items()}

# ----------------------------------------------------------------- encoding


def _vtk_type(a: np.ndarray) -> tuple[str, np.ndarray]:
    if a.dtype == np.bool_:
        a = a.astype(np.uint8)
    a = np.ascontiguousarray(a, dtype=a.dtype.newbyteorder("<"))
    try:
        return _NP_TO_VTK[a.dtype], a
    except KeyError:
        raise TypeError(f"no VTK type for dtype {a.dtype}") from None


def _header_bytes(values, header_type: str) -> bytes:
    return np.asarray(values, dtype=_VTK_TYPES[header_type]).astype(np.dtype(_VTK_TYPES[header_type]).newbyteorder("<")).tobytes()


def _pack(raw: bytes, header_type: str, compress: bool, block_size: int, level: int) -> tuple[bytes, bytes]:
    """(header, payload) bytes of one array: the byte count then the bytes, or
    the zlib block table then the compressed blocks."""
    if not compress:
        return _header_bytes([len(raw)], header_type), raw
    blocks = [raw[i:i + block_size] for i in range(0, len(raw), block_size)] or [b""]
    comp = [zlib.compress(b, level) for b in blocks]
    last = len(blocks[-1]) if len(blocks[-1]) != block_size else 0
    head = [len(blocks), block_size, last, *[len(c) for c in comp]]
    return _header_bytes(head, header_type), b"".join(comp)


def _unpack(buf: bytes, pos: int, header_type: str, compressed: bool, b64: bool) -> bytes:
    """Decode one binary array starting at ``pos`` of ``buf`` (base64 text
    when ``b64``, else raw bytes)."""
    isz = np.dtype(_VTK_TYPES[header_type]).itemsize

    def take_b64(p: int, nbytes: int) -> tuple[bytes, int]:
        nchar = 4 * ((nbytes + 2) // 3)
        return base64.b64decode(buf[p:p + nchar]), p + nchar

    def ints(b: bytes, n: int) -> list[int]:
        return [int(x) for x in np.frombuffer(b[:n * isz], dtype=np.dtype(_VTK_TYPES[header_type]).newbyteorder("<"))]

    if not compressed:
        if b64:
            # a header encoded on its own ends in the padding of its size (two
            # '=' for 4 bytes, one for 8); otherwise header and data form one
            # base64 stream
            first, hp = take_b64(pos, isz)
            n = ints(first, 1)[0]
            pad = (3 - isz % 3) % 3
            if buf[hp - pad:hp] == b"=" * pad:
                data, _ = take_b64(hp, n)
                return data[:n]
            allb, _ = take_b64(pos, isz + n)
            return allb[isz:isz + n]
        n = ints(buf[pos:pos + isz], 1)[0]
        return bytes(buf[pos + isz:pos + isz + n])
    if b64:
        h3, _ = take_b64(pos, 3 * isz)
        nb = ints(h3, 1)[0]
        head, p = take_b64(pos, (3 + nb) * isz)
        vals = ints(head, 3 + nb)
        csz = vals[3:]
        data, _ = take_b64(p, sum(csz))
    else:
        nb = ints(buf[pos:pos + isz], 1)[0]
        vals = ints(buf[pos:pos + (3 + nb) * isz], 3 + nb)
        csz = vals[3:]
        p = pos + (3 + nb) * isz
        data = bytes(buf[p:p + sum(csz)])
    out, q = [], 0
    for c in csz:
        out.append(zlib.decompress(data[q:q + c]))
        q += c
    bsz, last = vals[1], vals[2]
    for k, b in enumerate(out):
        want = last if (k == len(out) - 1 and last) else bsz
        if len(b) != want and not (len(out) == 1 and len(b) == 0):
            raise ValueError(f"zlib block {k} holds {len(b)} bytes, header says {want}")
    return b"".join(out)


# ----------------------------------------------------------------- writing


def save_vtu(path: str | os.PathLike, mesh: Mesh, *, encoding: str = "binary", compress: bool = True,
             header_type: str = "UInt64", block_size: int = 1 << 15, level: int = 6) -> None:
    """Write ``mesh`` (points, tetra cells, point and cell data) as ``.vtu``.

    encoding: "binary" (inline base64, what pyvista writes), "appended" (one
    raw section at the end: fastest for large meshes) or "ascii".  compress:
    zlib blocks of ``block_size`` bytes (binary and appended only).
    """
    if encoding not in ("binary", "appended", "ascii"):
        raise ValueError("encoding must be 'binary', 'appended' or 'ascii'")
    if header_type not in ("UInt32", "UInt64"):
        raise ValueError("header_type must be 'UInt32' or 'UInt64'")
    if block_size < 1:
        raise ValueError("block_size must be positive")
    cells = np.asarray(mesh.cells, np.int64)
    if cells.ndim != 2 or cells.shape[1] != 4:
        raise ValueError("only tetra meshes (T, 4) are written")
    compress = compress and encoding != "ascii"
    appended: list[bytes] = []
    app_off = 0

    def data_array(a, name: str | None) -> str:
        nonlocal app_off
        a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)  # torch tensors on any device
        n = a.shape[0] if a.ndim else 1
        ncomp = int(np.prod(a.shape[1:])) if a.ndim > 1 else 1
        vtype, a = _vtk_type(a.reshape(n, ncomp) if a.ndim > 1 else a.reshape(n))
        attrs = f'type="{vtype}"' + (f" Name={quoteattr(name)}" if name is not None else "")
        attrs += f' NumberOfComponents="{ncomp}"'
        if encoding == "ascii":
            if a.dtype.kind == "f":
                body = " ".join(repr(float(x)) if a.dtype == np.float64 else repr(float(np.float32(x)))
                                for x in a.ravel())
            else:
                body = " ".join(str(int(x)) for x in a.ravel())
            return f'<DataArray {attrs} format="ascii">{body}</DataArray>'
        head, payload = _pack(a.tobytes(), header_type, compress, block_size, level)
        if encoding == "binary":
            text = (base64.b64encode(head) + base64.b64encode(payload) if compress
                    else base64.b64encode(head + payload)).decode("ascii")
            return f'<DataArray {attrs} format="binary">{text}</DataArray>'
        appended.append(head + payload)
        off = app_off
        app_off += len(head) + len(payload)
        return f'<DataArray {attrs} format="appended" offset="{off}"/>'

    def section(tag: str, data: dict[str, Any]) -> str:
        inner = "".join(data_array(v, str(k)) for k, v in data.items())
        return f"<{tag}>{inner}</{tag}>"

    T = cells.shape[0]
    parts = [
        '<?xml version="1.0"?>\n',
        f'<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="{header_type}"'
        + (' compressor="vtkZLibDataCompressor"' if compress else "") + ">\n",
        "<UnstructuredGrid>\n",
        f'<Piece NumberOfPoints="{mesh.n_points}" NumberOfCells="{T}">\n',
        section("PointData", mesh.point_data) + "\n",
        section("CellData", mesh.cell_data) + "\n",
        "<Points>" + data_array(np.asarray(mesh.points, np.float64), None) + "</Points>\n",
        "<Cells>"
        + data_array(cells.reshape(-1), "connectivity")
        + data_array(4 * np.arange(1, T + 1, dtype=np.int64), "offsets")
        + data_array(np.full(T, VTK_TETRA, np.uint8), "types")
        + "</Cells>\n",
        "</Piece>\n</UnstructuredGrid>\n",
    ]
    with open(path, "wb") as f:
        f.write("".join(parts).encode("utf-8"))
        if encoding == "appended":
            f.write(b'<AppendedData encoding="raw">\n_')
            for b in appended:
                f.write(b)
            f.write(b"\n</AppendedData>\n")
        f.write(b"</VTKFile>\n")


# ----------------------------------------------------------------- reading


def _split_appended(raw: bytes) -> tuple[bytes, bytes | None, str | None]:
    """XML text with the appended payload cut out, the payload (after the
    ``_`` marker) and its encoding."""
    i = raw.find(b"<AppendedData")
    if i < 0:
        return raw, None, None
    j = raw.find(b">", i)
    tag = raw[i:j + 1].decode("ascii", "replace")
    m = re.search(r'encoding="(\w+)"', tag)
    enc = m.group(1) if m else "raw"
    u = raw.find(b"_", j)
    k = raw.rfind(b"</AppendedData>")
    if u < 0 or k < 0:
        raise ValueError("malformed AppendedData section")
    payload = raw[u + 1:k]
    if enc == "base64":
        payload = payload.rstrip()
    xml = raw[:i] + b"<AppendedData/>" + raw[k + len(b"</AppendedData>"):]
    return xml, payload, enc


def _read_array(el: ET.Element, header_type: str, compressed: bool, app: bytes | None, app_enc: str | None) -> np.ndarray:
    vtype = el.get("type")
    if vtype not in _VTK_TYPES:
        raise ValueError(f"unsupported DataArray type {vtype!r}")
    dt = np.dtype(_VTK_TYPES[vtype]).newbyteorder("<")
    ncomp = int(el.get("NumberOfComponents", "1"))
    fmt = el.get("format", "ascii")
    if fmt == "ascii":
        text = (el.text or "").split()
        a = np.array(text, dtype=np.float64 if dt.kind == "f" else np.int64).astype(dt.newbyteorder("="))
    elif fmt == "binary":
        buf = re.sub(rb"\s+", b"", (el.text or "").encode("ascii"))
        a = np.frombuffer(_unpack(buf, 0, header_type, compressed, True), dtype=dt)
    elif fmt == "appended":
        if app is None:
            raise ValueError("appended DataArray without an AppendedData section")
        off = int(el.get("offset", "0"))
        if app_enc == "base64":
            a = np.frombuffer(_unpack(app, off, header_type, compressed, True), dtype=dt)
        else:
            a = np.frombuffer(_unpack(app, off, header_type, compressed, False), dtype=dt)
    else:
        raise ValueError(f"unsupported DataArray format {fmt!r}")
    a = a.astype(dt.newbyteorder("="))
    return a.reshape(-1, ncomp) if ncomp > 1 else a


def load_vtu(path: str | os.PathLike) -> Mesh:
    """Read a ``.vtu`` file of tetra cells into a ``Mesh`` (the arrays keep
    their stored types; bool fields come back as uint8)."""
    with open(path, "rb") as f:
        raw = f.read()
    xml, app, app_enc = _split_appended(raw)
    root = ET.fromstring(xml)
    if root.tag != "VTKFile" or root.get("type") != "UnstructuredGrid":
        raise ValueError("not a VTK XML UnstructuredGrid file")
    if root.get("byte_order", "LittleEndian") != "LittleEndian":
        raise ValueError("only little-endian files are read")
    header_type = root.get("header_type", "UInt32")
    comp = root.get("compressor")
    if comp not in (None, "", "vtkZLibDataCompressor"):
        raise ValueError(f"unsupported compressor {comp!r}")
    compressed = comp == "vtkZLibDataCompressor"
    pieces = root.findall("./UnstructuredGrid/Piece")
    if len(pieces) != 1:
        raise ValueError(f"expected one Piece, found {len(pieces)}")
    piece = pieces[0]
    nP, nC = int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))
    rd = lambda el: _read_array(el, header_type, compressed, app, app_enc)  # noqa: E731
    points = rd(piece.find("./Points/DataArray")).astype(np.float64).reshape(nP, 3)
    carr = {el.get("Name"): el for el in piece.findall("./Cells/DataArray")}
    for k in ("connectivity", "offsets", "types"):
        if k not in carr:
            raise ValueError(f"Cells has no {k} array")
    conn, offs, types = rd(carr["connectivity"]), rd(carr["offsets"]), rd(carr["types"])
    if types.shape[0] != nC or offs.shape[0] != nC:
        raise ValueError("Cells arrays disagree with NumberOfCells")
    if np.any(types != VTK_TETRA):
        bad = sorted(set(int(t) for t in np.unique(types)) - {VTK_TETRA})
        raise ValueError(f"only tetra cells (VTK type 10) are read; found types {bad}")
    if not np.array_equal(offs, 4 * np.arange(1, nC + 1)):
        raise ValueError("tetra offsets must be 4, 8, 12, ...")
    cells = conn.astype(np.int64).reshape(nC, 4)
    if cells.size and (cells.min() < 0 or cells.max() >= nP):
        raise ValueError("cell connectivity out of range")

    def data(tag: str, n: int) -> dict[str, np.ndarray]:
        out = {}
        for el in piece.findall(f"./{tag}/DataArray"):
            a = rd(el)
            if a.shape[0] != n:
                raise ValueError(f"{tag} array {el.get('Name')!r} has {a.shape[0]} tuples, expected {n}")
            out[el.get("Name")] = a
        return out

    return Mesh(points, cells.astype(np.int32), data("PointData", nP), data("CellData", nC))


# ----------------------------------------------------------------- results and series


def with_solution(mesh: Mesh, u_full, *, name: str = DISPLACEMENT.vtk) -> Mesh:
    """Copy of ``mesh`` with the solved displacement as point data.

    ``u_full`` is the (V_model, 3) full displacement of a model; a mesh with a
    ``GlobalPointId`` array (set when it was added to a model builder) picks
    its rows by it, as exp/2026/03/11/test-inverse/src/20-forward.py:57-59
    does; otherwise ``u_full`` must have one row per mesh point.
    """
    u = u_full.detach().cpu().numpy() if hasattr(u_full, "detach") else np.asarray(u_full)
    u = np.asarray(u, np.float64).reshape(-1, 3)
    gid = mesh.point_data.get(GLOBAL_POINT_ID.vtk)
    if gid is not None:
        u = u[np.asarray(gid, np.int64)]
    elif u.shape[0] != mesh.n_points:
        raise ValueError(f"{u.shape[0]} displacement rows for {mesh.n_points} points and no GlobalPointId")
    out = mesh.copy()
    out.point_data[name] = u
    return out


class SeriesWriter:
    """``<name>.vtu.series`` (ParaView's JSON file series) plus one ``.vtu``
    per frame beside it, named ``<name>_000000.vtu``, ``<name>_000001.vtu``,
    ...  The index is rewritten after every frame, so a series cut short is
    still readable."""

    def __init__(self, path: str | os.PathLike, **save_kwargs):
        path = os.fspath(path)
        if not path.endswith(".vtu.series"):
            raise ValueError("series path must end with .vtu.series")
        self.path = path
        self.stem = os.path.basename(path)[: -len(".vtu.series")]
        self.dir = os.path.dirname(path) or "."
        self.save_kwargs = save_kwargs
        self.files: list[dict[str, Any]] = []

    def append(self, mesh: Mesh, time: float | None = None) -> str:
        k = len(self.files)
        name = f"{self.stem}_{k:06d}.vtu"
        save_vtu(os.path.join(self.dir, name), mesh, **self.save_kwargs)
        self.files.append({"name": name, "time": float(k if time is None else time)})
        with open(self.path, "w") as f:
            json.dump({"file-series-version": "1.0", "files": self.files}, f, indent=1)
        return os.path.join(self.dir, name)

    def __enter__(self) -> "SeriesWriter":
        return self

    def __exit__(self, *exc) -> None:
        return None


def load_series(path: str | os.PathLike) -> list[tuple[float, str]]:
    """(time, file path) of every frame of a ``.vtu.series`` index."""
    path = os.fspath(path)
    with open(path) as f:
        idx = json.load(f)
    if "files" not in idx:
        raise ValueError("not a file-series index (no 'files')")
    base = os.path.dirname(path) or "."
    return [(float(e["time"]), os.path.join(base, e["name"])) for e in idx["files"]]


def iter_series(path: str | os.PathLike) -> Iterator[tuple[float, Mesh]]:
    for t, p in load_series(path):
        yield t, load_vtu(p)
