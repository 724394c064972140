"""The .vtu / .vtu.series format (apple_amd.vtu), host only.

Pinned by files assembled by hand from the VTK XML file-format documentation
(an ascii one-tet file; base64 payloads with the header encoded with the data
and on its own; a zlib payload whose stored-block stream and block table are
written out byte by byte; a raw appended section), not by the writer.
Round trips then cover every encoding x compression x header type, and the
reference's attribute names carry over (common/attr_name.py:4-44).
"""

import json
import os

import numpy as np
import pytest

from apple_amd import vtu
from apple_amd.common import ACTIVATION, DISPLACEMENT, FIXED_MASK, FIXED_VALUE, GLOBAL_POINT_ID, LAMBDA, MU
from apple_amd.mesh import Mesh, cube

ONE_TET_ASCII = """<?xml version="1.0"?>
<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">
  <UnstructuredGrid>
    <Piece NumberOfPoints="4" NumberOfCells="1">
      <PointData>
        <DataArray type="Float32" Name="Displacement" NumberOfComponents="3" format="ascii">
          0 0 0  0.5 0 0
          0 0.25 0  0 0 -1
        </DataArray>
        <DataArray type="UInt8" Name="FixedMask" format="ascii">1 0 0 1</DataArray>
      </PointData>
      <CellData>
        <DataArray type="Float64" Name="mu" format="ascii">1000</DataArray>
      </CellData>
      <Points>
        <DataArray type="Float32" NumberOfComponents="3" format="ascii">0 0 0 1 0 0 0 1 0 0 0 1</DataArray>
      </Points>
      <Cells>
        <DataArray type="Int32" Name="connectivity" format="ascii">0 1 2 3</DataArray>
        <DataArray type="Int32" Name="offsets" format="ascii">4</DataArray>
        <DataArray type="UInt8" Name="types" format="ascii">10</DataArray>
      </Cells>
    </Piece>
  </UnstructuredGrid>
</VTKFile>
"""


def _one_tet(mu_array: str, header_type: str | None = None, compressor: str | None = None, tail: bytes = b"") -> bytes:
    """One-tet file (ascii points and cells) whose cell array ``mu`` is the given DataArray element."""
    attrs = ""
    if header_type:
        attrs += f' header_type="{header_type}"'
    if compressor:
        attrs += f' compressor="{compressor}"'
    return (
        f'<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian"{attrs}>'
        '<UnstructuredGrid><Piece NumberOfPoints="4" NumberOfCells="1"><PointData/>'
        f"<CellData>{mu_array}</CellData>"
        '<Points><DataArray type="Float64" NumberOfComponents="3" format="ascii">0 0 0 1 0 0 0 1 0 0 0 1</DataArray></Points>'
        '<Cells><DataArray type="Int64" Name="connectivity" format="ascii">0 1 2 3</DataArray>'
        '<DataArray type="Int64" Name="offsets" format="ascii">4</DataArray>'
        '<DataArray type="UInt8" Name="types" format="ascii">10</DataArray></Cells>'
        "</Piece></UnstructuredGrid>"
    ).encode() + tail + b"</VTKFile>\n"


def _write(tmp_path, name, data: bytes | str):
    p = tmp_path / name
    p.write_bytes(data.encode() if isinstance(data, str) else data)
    return p


def test_ascii_known_file(tmp_path):
    m = vtu.load_vtu(_write(tmp_path, "t.vtu", ONE_TET_ASCII))
    np.testing.assert_array_equal(m.points, np.eye(4, 3, -1))
    np.testing.assert_array_equal(m.cells, [[0, 1, 2, 3]])
    d = m.point_data["Displacement"]
    assert d.dtype == np.float32 and d.shape == (4, 3)
    np.testing.assert_array_equal(d, [[0, 0, 0], [0.5, 0, 0], [0, 0.25, 0], [0, 0, -1]])
    np.testing.assert_array_equal(m.point_data["FixedMask"], [1, 0, 0, 1])
    assert m.cell_data["mu"].dtype == np.float64 and m.cell_data["mu"][0] == 1000.0


# mu = [1.0] as Float64: 8 data bytes 00 00 00 00 00 00 F0 3F
BIN_UINT32_JOINT = "CAAAAAAAAAAAAPA/"  # base64(08 00 00 00 | data)
BIN_UINT64_SPLIT = "CAAAAAAAAAA=AAAAAAAA8D8="  # base64(08 00 .. 00) + base64(data)


@pytest.mark.parametrize("header_type,text", [("UInt32", BIN_UINT32_JOINT), ("UInt64", BIN_UINT64_SPLIT)])
def test_binary_known_payloads(tmp_path, header_type, text):
    arr = f'<DataArray type="Float64" Name="mu" format="binary">\n  {text}\n</DataArray>'
    m = vtu.load_vtu(_write(tmp_path, "b.vtu", _one_tet(arr, header_type)))
    assert m.cell_data["mu"].tolist() == [1.0]


def test_binary_header_type_defaults_to_uint32(tmp_path):
    arr = f'<DataArray type="Float64" Name="mu" format="binary">{BIN_UINT32_JOINT}</DataArray>'
    assert vtu.load_vtu(_write(tmp_path, "b.vtu", _one_tet(arr))).cell_data["mu"].tolist() == [1.0]


def test_zlib_known_payload(tmp_path):
    # mu = [7] as Int32.  zlib stream of the 4 bytes with a stored block:
    # 78 01 | 01 (final, stored) | 04 00 FB FF (LEN, ~LEN) | 07 00 00 00 |
    # adler32 00 20 00 08 (a = 1 + 7 = 8, b = 8 + 8 + 8 + 8 = 32) -- 15 bytes.
    # Block table (UInt32): 1 block, block size 32768, last block 4 bytes, 15
    # compressed bytes; table and stream are base64-encoded one after the other.
    text = "AQAAAACAAAAEAAAADwAAAA==" + "eAEBBAD7/wcAAAAAIAAI"
    arr = f'<DataArray type="Int32" Name="mu" format="binary">{text}</DataArray>'
    m = vtu.load_vtu(_write(tmp_path, "z.vtu", _one_tet(arr, "UInt32", "vtkZLibDataCompressor")))
    assert m.cell_data["mu"].dtype == np.int32 and m.cell_data["mu"].tolist() == [7]


def test_appended_raw_known_layout(tmp_path):
    # two arrays in one raw section: offsets count from the byte after '_'
    a = b"\x08\x00\x00\x00" + b"\x00\x00\x00\x00\x00\x00\xf0\x3f"  # UInt32 count, Float64 1.0
    b = b"\x04\x00\x00\x00" + b"\xfe\xff\xff\xff"  # UInt32 count, Int32 -2
    arrs = ('<DataArray type="Float64" Name="mu" format="appended" offset="0"/>'
            f'<DataArray type="Int32" Name="tag" format="appended" offset="{len(a)}"/>')
    tail = b'<AppendedData encoding="raw">\n   _' + a + b + b"\n</AppendedData>"
    m = vtu.load_vtu(_write(tmp_path, "r.vtu", _one_tet(arrs, "UInt32", tail=tail)))
    assert m.cell_data["mu"].tolist() == [1.0] and m.cell_data["tag"].tolist() == [-2]


def test_appended_base64_known_layout(tmp_path):
    # base64 appended data: offsets count base64 characters
    arrs = ('<DataArray type="Float64" Name="mu" format="appended" offset="0"/>'
            f'<DataArray type="Float64" Name="nu" format="appended" offset="{len(BIN_UINT32_JOINT)}"/>')
    tail = b'<AppendedData encoding="base64">_' + BIN_UINT32_JOINT.encode() * 2 + b"\n</AppendedData>"
    m = vtu.load_vtu(_write(tmp_path, "a.vtu", _one_tet(arrs, "UInt32", tail=tail)))
    assert m.cell_data["mu"].tolist() == [1.0] and m.cell_data["nu"].tolist() == [1.0]


def test_appended_base64_compressed_known_layout(tmp_path):
    # the zlib payload of test_zlib_known_payload (table and stream, each
    # base64-encoded) as the first array of a base64 appended section, then
    # the same again: the second array's offset is the first one's length
    z = "AQAAAACAAAAEAAAADwAAAA==" + "eAEBBAD7/wcAAAAAIAAI"
    arrs = ('<DataArray type="Int32" Name="mu" format="appended" offset="0"/>'
            f'<DataArray type="Int32" Name="tag" format="appended" offset="{len(z)}"/>')
    tail = b'<AppendedData encoding="base64">_' + (z + z).encode() + b"\n</AppendedData>"
    m = vtu.load_vtu(_write(tmp_path, "ac.vtu", _one_tet(arrs, "UInt32", "vtkZLibDataCompressor", tail=tail)))
    assert m.cell_data["mu"].tolist() == [7] and m.cell_data["tag"].tolist() == [7]


def _attributed_cube(n=3, seed=0):
    m = cube(n)
    rng = np.random.default_rng(seed)
    m.point_data[FIXED_MASK.vtk] = np.repeat((m.points[:, 0] < 1e-9)[:, None], 3, axis=1)
    m.point_data[FIXED_VALUE.vtk] = rng.standard_normal((m.n_points, 3))
    m.point_data[GLOBAL_POINT_ID.vtk] = np.arange(m.n_points, dtype=np.int64)
    m.cell_data[MU.vtk] = rng.lognormal(np.log(1e4), 0.5, m.n_cells)
    m.cell_data[LAMBDA.vtk] = (10.0 * m.cell_data[MU.vtk]).astype(np.float32)
    m.cell_data[ACTIVATION.value] = rng.standard_normal((m.n_cells, 6))  # keyed by the enum value (utils/_material.py:16)
    m.cell_data["tag"] = rng.integers(-5, 5, m.n_cells, dtype=np.int16)
    return m


@pytest.mark.parametrize("encoding", ["ascii", "binary", "appended"])
@pytest.mark.parametrize("compress", [False, True])
@pytest.mark.parametrize("header_type", ["UInt32", "UInt64"])
def test_round_trip_bit_exact(tmp_path, encoding, compress, header_type):
    m = _attributed_cube()
    p = tmp_path / "m.vtu"
    # a small block size makes the compressed arrays span several blocks
    vtu.save_vtu(p, m, encoding=encoding, compress=compress, header_type=header_type, block_size=256)
    r = vtu.load_vtu(p)
    np.testing.assert_array_equal(r.points, m.points)
    np.testing.assert_array_equal(r.cells, m.cells)
    assert set(r.point_data) == set(m.point_data) and set(r.cell_data) == set(m.cell_data)
    for src, dst in ((m.point_data, r.point_data), (m.cell_data, r.cell_data)):
        for k, v in src.items():
            v = np.asarray(v)
            expect = v.astype(np.uint8) if v.dtype == np.bool_ else v
            assert dst[k].dtype == expect.dtype and dst[k].shape == expect.shape, k
            np.testing.assert_array_equal(dst[k], expect)


def test_zlib_block_table(tmp_path):
    # 40 cells x 8 B of mu with 100-byte blocks: 4 blocks, the last one 20 bytes
    m = cube(2)
    m.cell_data["mu"] = np.arange(m.n_cells, dtype=np.float64)
    p = tmp_path / "z.vtu"
    vtu.save_vtu(p, Mesh(m.points, m.cells, {}, {"mu": m.cell_data["mu"]}), compress=True, header_type="UInt32",
                 block_size=100)
    text = p.read_text()
    payload = text.split('Name="mu" NumberOfComponents="1" format="binary">')[1].split("<")[0]
    head = np.frombuffer(__import__("base64").b64decode(payload[:4 * ((7 * 4 + 2) // 3)])[:28], "<u4")
    assert head[:3].tolist() == [4, 100, 20]
    np.testing.assert_array_equal(vtu.load_vtu(p).cell_data["mu"], m.cell_data["mu"])


def test_rejects_other_cells_and_bad_files(tmp_path):
    tri = ONE_TET_ASCII.replace(">10<", ">5<")
    with pytest.raises(ValueError, match="tetra"):
        vtu.load_vtu(_write(tmp_path, "tri.vtu", tri))
    with pytest.raises(ValueError, match="offsets"):
        vtu.load_vtu(_write(tmp_path, "off.vtu", ONE_TET_ASCII.replace('format="ascii">4<', 'format="ascii">3<')))
    with pytest.raises(ValueError, match="range"):
        vtu.load_vtu(_write(tmp_path, "rng.vtu", ONE_TET_ASCII.replace(">0 1 2 3<", ">0 1 2 4<")))
    with pytest.raises(ValueError, match="little-endian"):
        vtu.load_vtu(_write(tmp_path, "be.vtu", ONE_TET_ASCII.replace("LittleEndian", "BigEndian")))
    with pytest.raises(ValueError, match="UnstructuredGrid"):
        vtu.load_vtu(_write(tmp_path, "pd.vtu", ONE_TET_ASCII.replace('type="UnstructuredGrid"', 'type="PolyData"')))
    with pytest.raises(ValueError, match="tuples"):
        vtu.load_vtu(_write(tmp_path, "n.vtu", ONE_TET_ASCII.replace(">1 0 0 1<", ">1 0 0<")))
    with pytest.raises(ValueError, match="encoding"):
        vtu.save_vtu(tmp_path / "x.vtu", cube(1), encoding="hdf5")
    with pytest.raises(ValueError, match="tetra"):
        vtu.save_vtu(tmp_path / "x.vtu", Mesh(np.zeros((3, 3)), np.zeros((1, 3), np.int32)))


def test_reference_attribute_names_survive(tmp_path):
    # the names a liblaf/apple input file carries: .vtk names (common/attr_name.py:39-44)
    # for mu, lambda and the point fields (utils/_material.py:27-31,
    # forward/dof_map/_builder.py:36-39), the enum value for the activation
    # (utils/_material.py:16)
    assert [a.vtk for a in (MU, LAMBDA, FIXED_MASK, FIXED_VALUE, GLOBAL_POINT_ID, DISPLACEMENT)] == [
        "mu", "lambda", "FixedMask", "FixedValue", "GlobalPointId", "Displacement"]
    assert ACTIVATION.value == "activation"
    m = _attributed_cube()
    vtu.save_vtu(tmp_path / "m.vtu", m)
    r = vtu.load_vtu(tmp_path / "m.vtu")
    assert {"mu", "lambda", "activation", "tag"} <= set(r.cell_data)
    assert {"FixedMask", "FixedValue", "GlobalPointId"} <= set(r.point_data)


def test_torch_fields_and_awkward_names(tmp_path):
    import torch

    m = cube(1)
    m.point_data['a "quoted" <name> & more'] = torch.arange(3 * m.n_points, dtype=torch.float32).reshape(-1, 3)
    vtu.save_vtu(tmp_path / "t.vtu", m)
    r = vtu.load_vtu(tmp_path / "t.vtu")
    np.testing.assert_array_equal(r.point_data['a "quoted" <name> & more'],
                                  np.arange(3 * m.n_points, dtype=np.float32).reshape(-1, 3))


def test_with_solution_picks_rows_by_global_point_id():
    m = cube(1)
    m.point_data[GLOBAL_POINT_ID.vtk] = np.arange(m.n_points)[::-1] + 3  # mesh added after 3 other points
    u_full = np.arange(3 * (m.n_points + 3), dtype=np.float64).reshape(-1, 3)
    out = vtu.with_solution(m, u_full)
    np.testing.assert_array_equal(out.point_data[DISPLACEMENT.vtk], u_full[m.point_data[GLOBAL_POINT_ID.vtk]])
    assert DISPLACEMENT.vtk not in m.point_data  # the input is left alone
    plain = cube(1)
    with pytest.raises(ValueError, match="rows"):
        vtu.with_solution(plain, u_full)


def test_series_index_and_frames(tmp_path):
    path = tmp_path / "run.vtu.series"
    m = cube(1)
    with vtu.SeriesWriter(path, encoding="appended") as w:
        for k in range(3):
            f = m.copy()
            f.point_data["Displacement"] = np.full((m.n_points, 3), float(k))
            w.append(f, time=0.5 * k)
    idx = json.loads(path.read_text())
    assert idx == {"file-series-version": "1.0",
                   "files": [{"name": f"run_{k:06d}.vtu", "time": 0.5 * k} for k in range(3)]}
    frames = list(vtu.iter_series(path))
    assert [t for t, _ in frames] == [0.0, 0.5, 1.0]
    for k, (_, fm) in enumerate(frames):
        assert np.all(fm.point_data["Displacement"] == k)
    assert os.path.dirname(vtu.load_series(path)[0][1]) == str(tmp_path)
    with pytest.raises(ValueError, match="series"):
        vtu.SeriesWriter(tmp_path / "run.json")
