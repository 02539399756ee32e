"""VTK ImageData output (glide/io.py:23-389): the byte layout of the
reference's writer, pinned by a file assembled here by hand, round trips of 2-D,
vector and 3-D fields, and the writer's static/dynamic/PVD bookkeeping (CPU)."""
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest
import torch

from glide_amd.io import VTIWriter, read_vti, write_vti


def test_bytes_of_a_small_file(tmp_path):
    """A 2 x 3 field with a time value, byte for byte: declaration, ImageData
    header, one appended Float32 block with its uint32 length, rows flipped."""
    a = np.arange(6, dtype=np.float64).reshape(2, 3)
    p = write_vti(tmp_path / "a.vti", {"H": a}, 10.0, origin=(1.0, 2.0), time_value=5.0)
    body = np.flip(a, axis=0).astype("<f4").tobytes()
    expected = (
        "<?xml version='1.0' encoding='utf-8'?>\n"
        '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian">'
        '<ImageData WholeExtent="0 2 0 1 0 0" Origin="1.0 2.0 0.0" Spacing="10.0 10.0 1.0">'
        '<Piece Extent="0 2 0 1 0 0">'
        '<FieldData><DataArray type="Float32" Name="TimeValue" NumberOfComponents="1" format="ascii">\n5.0\n'
        '</DataArray></FieldData>'
        '<PointData Scalars="H"><DataArray type="Float32" Name="H" NumberOfComponents="1" format="appended" '
        'offset="0" /></PointData></Piece></ImageData>'
    ).encode() + b'  <AppendedData encoding="raw">\n   _' + struct.pack("<I", 24) + body + \
        b"\n  </AppendedData>\n</VTKFile>\n"
    assert p.read_bytes() == expected


def test_round_trip_2d_scalar_and_vector(tmp_path):
    rng = np.random.default_rng(0)
    H = rng.standard_normal((5, 7))
    u = torch.as_tensor(rng.standard_normal((5, 7)))
    v = rng.standard_normal((5, 7))
    write_vti(tmp_path / "b.vti", {"H": H, "U": [u, v]}, 100.0, dy=50.0)
    r = read_vti(tmp_path / "b.vti")
    assert r["extent"] == [0, 6, 0, 4, 0, 0]
    assert r["spacing"] == [100.0, 50.0, 1.0]
    assert r["time"] is None
    assert np.array_equal(r["fields"]["H"], np.flip(H, 0).astype(np.float32))
    U = r["fields"]["U"]
    assert U.shape == (5, 7, 2)
    assert np.array_equal(U[..., 0], np.flip(u.numpy(), 0).astype(np.float32))
    assert np.array_equal(U[..., 1], np.flip(v, 0).astype(np.float32))
    # the second block starts after the first block and its 4-byte length
    root = ET.fromstring(open(tmp_path / "b.vti", "rb").read().split(b"  <AppendedData")[0] + b"</VTKFile>")
    offs = [int(d.get("offset")) for d in root.iter("DataArray")]
    assert offs == [0, 5 * 7 * 4 + 4]


def test_round_trip_3d_is_x_fastest(tmp_path):
    E = np.random.default_rng(1).standard_normal((4, 6, 3))
    write_vti(tmp_path / "c.vti", {"E": E}, 1000.0, dz=0.5, flip_y=False, origin=(0.0, 0.0, -1.0))
    r = read_vti(tmp_path / "c.vti")
    assert r["extent"] == [0, 5, 0, 3, 0, 2]
    assert r["origin"] == [0.0, 0.0, -1.0]
    assert np.array_equal(r["fields"]["E"], E.transpose(2, 0, 1).astype(np.float32))


def test_shape_errors(tmp_path):
    with pytest.raises(ValueError):
        write_vti(tmp_path / "d.vti", {"E": np.zeros((2, 2, 2))}, 1.0)
    with pytest.raises(ValueError):
        write_vti(tmp_path / "d.vti", {"E": np.zeros((2, 2, 2)), "H": np.zeros((2, 2))}, 1.0, dz=1.0)
    with pytest.raises(NotImplementedError):
        write_vti(tmp_path / "d.vti", {"E": np.zeros((2, 2, 2)), "U": [np.zeros((2, 2, 2))]}, 1.0, dz=1.0)


def test_writer_static_dynamic_and_manifest(tmp_path):
    from glide_amd.grid import Grid
    g = Grid(4, 6, 500.0, device="cpu")
    g.state.H.set(np.full((4, 6), 100.0))
    g.state.u.set(np.tile(np.arange(7.0), (4, 1)))
    g.state.v.set(np.zeros((5, 6)))
    calls = []
    w = VTIWriter(tmp_path / "out", static_fields={"bed": np.zeros((4, 6))},
                  dynamic_fields={"H": g.state.H, "U": [g.state.u, g.state.v],
                                  "twice_H": lambda: calls.append(1) or 2 * g.state.H.data},
                  base="run", dx=500.0)
    w.initialize(g, write_dynamic_initial=True, time=0.0)
    g.state.H.data.add_(1.0)
    w.append(g, time=10.0)
    pvd = w.write_pvd()
    assert sorted(p.name for p in (tmp_path / "out").iterdir()) == ["run.pvd", "run_0000.vti", "run_0001.vti",
                                                                    "run_static.vti"]
    assert len(calls) == 2
    r = read_vti(tmp_path / "out" / "run_0001.vti")
    assert r["time"] == 10.0
    assert np.all(r["fields"]["H"] == 101.0)
    assert np.all(r["fields"]["twice_H"] == 202.0)
    # facet velocity collocated to cell centres: u = j + 0.5
    assert np.array_equal(r["fields"]["U"][..., 0], np.tile(np.arange(6) + 0.5, (4, 1)).astype(np.float32))
    coll = ET.parse(pvd).getroot().find("Collection")
    assert [(d.get("timestep"), d.get("file")) for d in coll] == [("0.0", "run_0000.vti"), ("10.0", "run_0001.vti")]
    w.initialize(g, overwrite=True)
    assert sorted(p.name for p in (tmp_path / "out").iterdir()) == ["run_static.vti"]
