"""Result files (femgpu_write_vtu, host only) against the reference's own writer.

The golden file is write_mesh_vtk(filename, m) of LMT++ (mesh/write_mesh_vtk.h:434-656,827-831),
built from source, on a make_rect Hexa mesh of the elasticity_dep_static problem; given the same
fields (the node fields dep, f_nodal, is_on_skin, f_surf, the element fields epsilon, sigma,
tr_epsilon, the global fields elastic_modulus, density, poisson_ratio, f_vol), the library must
write the same bytes.  A second check decodes a file with Python's own XML and base64 readers:
2-component point fields padded to 3, cell fields in name order, offsets and length prefixes.
No GPU compute is called here.
"""
import base64
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from conftest import GOLDEN


def test_bytes_equal_the_reference_writer(tmp_path):
    g = np.load(os.path.join(GOLDEN, "ref_write_vtu_hexa_3_4_2.npz"))
    c, e, u = g["coords"], g["conn"], g["u"]
    assert np.array_equal(e, L.make_rect(L.HEXA, (3, 4, 2))[1])
    nn, ne = len(c), len(e)
    z3, z1 = np.zeros((nn, 3)), np.zeros((nn, 1))
    out = tmp_path / "r.vtu"
    A.write_vtu(out, L.HEXA, c, e,
                point=[("dep", u), ("f_nodal", z3), ("is_on_skin", z1), ("f_surf", z3)],
                cell=[("tr_epsilon", g["tr"]), ("sigma", g["sigma"]), ("epsilon", g["eps"])],
                glob=[("elastic_modulus", [float(g["E"])]), ("density", [7800.0]), ("poisson_ratio", [float(g["nu"])]),
                      ("f_vol", [0.0, 0.0, 0.0])])
    assert out.read_bytes() == bytes(g["vtu"])


def _decode(block, offset, dtype):
    """One base64 segment of the appended block: a 4-byte length, then that many bytes."""
    n = struct.unpack("<I", base64.b64decode(block[offset:offset + 8])[:4])[0]
    seg = 4 * ((4 + n + 2) // 3)
    raw = base64.b64decode(block[offset:offset + seg])
    assert len(raw) == 4 + n
    return np.frombuffer(raw[4:], dtype=dtype)


def test_decoded_fields(tmp_path):
    c, e = L.make_rect(L.TETRA10, (3, 2, 2))
    nn, ne = len(c), len(e)
    rng = np.random.default_rng(3)
    p2, p3 = rng.standard_normal((nn, 2)), rng.standard_normal((nn, 3))
    c2, c1 = rng.standard_normal((ne, 2)), rng.standard_normal((ne, 1))
    grp = np.arange(ne, dtype=np.uint32) % 3
    out = tmp_path / "t.vtu"
    A.write_vtu(out, L.TETRA10, c, e, point=[("b2", p2), ("a3", p3)], cell=[("zz", c1), ("aa", c2)], group=grp,
                glob=[("g2", [1.5, -2.0])])
    text = out.read_text()
    root = ET.fromstring(text)
    block = root.find("AppendedData").text.strip()[1:]       # after the leading '_'
    piece = root.find("UnstructuredGrid/Piece")
    assert (piece.get("NumberOfPoints"), piece.get("NumberOfCells")) == (str(nn), str(ne))
    pd = piece.findall("PointData/DataArray")
    assert [(d.get("Name"), d.get("NumberOfComponents")) for d in pd] == [("b2", "3"), ("a3", "3")]
    v = _decode(block, int(pd[0].get("offset")), np.float64).reshape(nn, 3)
    assert np.array_equal(v[:, :2], p2) and not v[:, 2].any()
    assert np.array_equal(_decode(block, int(pd[1].get("offset")), np.float64).reshape(nn, 3), p3)
    cd = piece.findall("CellData/DataArray")
    assert [d.get("Name") for d in cd] == ["aa", "zz", "group", "g2"]
    assert cd[0].get("NumberOfComponents") == "2"
    assert np.array_equal(_decode(block, int(cd[0].get("offset")), np.float64).reshape(ne, 2), c2)
    assert np.array_equal(_decode(block, int(cd[1].get("offset")), np.float64).reshape(ne, 1), c1)
    assert np.array_equal(np.array(cd[2].text.split(), dtype=np.uint32), grp)
    gv = _decode(block, int(cd[3].get("offset")), np.float64).reshape(ne, 3)
    assert np.array_equal(gv, np.tile([1.5, -2.0, 0.0], (ne, 1)))
    pts = piece.find("Points/DataArray")
    assert np.array_equal(_decode(block, int(pts.get("offset")), np.float64).reshape(nn, 3), c)
    cells = {d.get("Name"): np.array(d.text.split(), dtype=np.int64) for d in piece.findall("Cells/DataArray")}
    assert np.array_equal(cells["connectivity"], e.ravel())
    assert np.array_equal(cells["offsets"], 10 * np.arange(1, ne + 1))
    assert (cells["types"] == 24).all()


def test_bad_input(tmp_path):
    c, e = L.make_rect(L.HEXA, (2, 2, 2))
    bad = e.copy()
    bad[0, 0] = len(c)
    with pytest.raises(A.FemGpuError):
        A.write_vtu(tmp_path / "x.vtu", L.HEXA, c, bad)
    with pytest.raises(A.FemGpuError):
        A.write_vtu(tmp_path / "no_such_dir" / "x.vtu", L.HEXA, c, e)
