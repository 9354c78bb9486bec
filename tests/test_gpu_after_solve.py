"""GPU after-solve recovery (k_after_solve, femgpu_after_solve) against the oracle and the
reference's generated Hexa after_solve.  Tolerance: 1e-13 of each field's largest entry (the
reference's generated expression order is not reproduced bit for bit)."""
import os
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from conftest import GOLDEN
from problems import build, jitter_interior, lognormal_E, tetra10_jittered

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def close(a, b, rel=1e-13):
    return np.abs(a - b).max() <= rel * np.abs(b).max()


def test_after_solve_matches_reference_generated_hexa():
    say("after_solve vs reference fixture")
    g = np.load(os.path.join(GOLDEN, "ref_after_solve_hexa_4_3_5.npz"))
    f = L.Formulation(L.HEXA, g["coords"], g["conn"])
    f.set_material(float(g["E"]), float(g["nu"]))
    f.set_result(g["u"])
    eps, sig, tr = f.call_after_solve()
    assert close(eps, g["eps"]) and close(sig, g["sigma"]) and close(tr, g["tr"])


@pytest.mark.parametrize("kind", [L.TETRA, L.TETRA10, L.HEXA])
def test_after_solve_of_a_solution_matches_oracle(oracle, kind):
    say("after_solve", kind)
    if kind == L.TETRA10:
        c, e = tetra10_jittered(3, seed=1)
    else:
        c, e = L.make_rect(kind, (6, 5, 5))
        c = jitter_interior(c, 0.2, 0.15, 2)
    E = lognormal_E(len(e), 0.5, 9)
    f, *_ = build(kind, c, e, E, nu=0.28)
    assert f.solve(rel_tol=1e-10)
    u = f.get_result()
    eps, sig, tr = f.call_after_solve()
    oe, os_, ot = oracle.after_solve(kind, e, c, E, 0.28, u)
    assert close(eps, oe) and close(sig, os_) and close(tr, ot)


def test_batch_after_solve_matches_single():
    say("batch after_solve")
    from problems import clamp_and_traction
    items = []
    for i, n in enumerate((3, 5)):
        c, e = L.make_rect(L.HEXA, (n + 1,) * 3)
        items.append((c, e, lognormal_E(len(e), 0.5, 20 + i)))
    b = L.Batch(L.HEXA, [(c, e) for c, e, _ in items])
    for i, (c, e, E) in enumerate(items):
        faces, tr, dofs = clamp_and_traction(L.HEXA, c, e)
        b.set_material(i, E, 0.3)
        b.set_traction(i, faces, tr)
        b.add_constraints(i, dofs, None, 1e5)
    b.assemble()
    ok, _ = b.solve(rel_tol=1e-11, precision=L.FP64)
    assert ok
    for i, (c, e, E) in enumerate(items):
        f = L.Formulation(L.HEXA, c, e)
        f.set_material(E, 0.3)
        f.set_result(b.get_result(i))
        single = f.call_after_solve()
        for x, y in zip(b.call_after_solve(i), single):
            assert np.array_equal(x, y)


def test_result_file_holds_result_and_recovery(tmp_path):
    """femgpu_write_result_vtu: dep, E and the after-solve fields of the problem, decoded back."""
    import base64
    import struct
    import xml.etree.ElementTree as ET
    c, e = L.make_rect(L.TETRA10, (4, 3, 3))
    E = lognormal_E(len(e), 0.3, 9)
    f, *_ = build(L.TETRA10, c, e, E)
    assert f.solve(rel_tol=1e-12)
    path = tmp_path / "res.vtu"
    f.write_vtu(path)
    root = ET.fromstring(path.read_text())
    block = root.find("AppendedData").text.strip()[1:]

    def arr(d, rows):
        off = int(d.get("offset"))
        n = struct.unpack("<I", base64.b64decode(block[off:off + 8])[:4])[0]
        raw = base64.b64decode(block[off:off + 4 * ((4 + n + 2) // 3)])
        return np.frombuffer(raw[4:], dtype=np.float64).reshape(rows, -1)

    piece = root.find("UnstructuredGrid/Piece")
    pd = {d.get("Name"): d for d in piece.findall("PointData/DataArray")}
    cd = {d.get("Name"): d for d in piece.findall("CellData/DataArray")}
    assert np.array_equal(arr(pd["dep"], len(c)).ravel(), f.get_result())
    eps, sig, tr = f.call_after_solve()
    assert np.array_equal(arr(cd["epsilon"], len(e)), eps)
    assert np.array_equal(arr(cd["sigma"], len(e)), sig)
    assert np.array_equal(arr(cd["tr_epsilon"], len(e)).ravel(), tr)
    assert np.array_equal(arr(cd["elastic_modulus"], len(e)).ravel(), E)
    assert (arr(cd["poisson_ratio"], len(e)) == 0.3).all()
    assert np.array_equal(arr(piece.find("Points/DataArray"), len(c)), c)
    f.close()


def test_result_avs_holds_result_and_recovery(tmp_path):
    """femgpu_write_result_avs: the problem's mesh, dep, pos, E and the after-solve fields in the
    reference writer's layout, byte for byte what femgpu_write_avs writes from the same values
    read back through the API (its byte equality with the reference's writer is a CPU test)."""
    from lmtpp_amd import _capi as A
    say("write_result_avs")
    c, e = L.make_rect(L.HEXA, (5, 4, 4))
    c = jitter_interior(c, 0.2, 0.1, 4)
    E = lognormal_E(len(e), 0.3, 4)
    f, *_ = build(L.HEXA, c, e, E, bf=(0.0, 0.1, -0.2))
    assert f.solve(rel_tol=1e-11)
    eps, sig, tr = f.call_after_solve()
    u = f.get_result()
    for display in (None, ["sigma", "dep", "poisson_ratio"]):
        got, ref = tmp_path / "res.inp", tmp_path / "ref.inp"
        f.write_avs(got, display)
        A.write_avs(ref, L.HEXA, c, e, point=[("dep", u), ("pos", c)],
                    cell=[("elastic_modulus", E, 0), ("epsilon", eps, 2), ("sigma", sig, 2), ("tr_epsilon", tr)],
                    glob=[("poisson_ratio", [0.3], 0), ("f_vol", [0.0, 0.1, -0.2])],
                    display=display or ["all"])
        assert got.read_bytes() == ref.read_bytes()
    f.close()
