"""Oracle solvers against the reference's known answers and its own compiled IC(0)-PCG."""
import json
import os

import numpy as np

from conftest import GOLDEN


def _kat():
    with open(os.path.join(GOLDEN, "ref_cg_kat.json")) as f:
        return json.load(f)


def _csr(dense):
    rp, col, val = [0], [], []
    for row in dense:
        for j, v in enumerate(row):
            if v != 0:
                col.append(j)
                val.append(v)
        rp.append(len(col))
    return np.array(rp), np.array(col), np.array(val, float)


def test_cg_known_answers(oracle):
    k = _kat()
    low = k["sym3"]["lower"]
    A = np.zeros((3, 3))
    for i, row in enumerate(low):
        for j, v in enumerate(row):
            A[i, j] = A[j, i] = v
    x, it, rc = oracle.cg_csr(*_csr(A), np.array(k["sym3"]["b"], float), rel_tol=1e-15)
    assert rc == 0 and np.allclose(x, k["sym3"]["x"], rtol=1e-12, atol=0)
    A = np.array(k["dense10"]["rows"], float)
    x, it, rc = oracle.cg_csr(*_csr(A), np.array(k["dense10"]["b"], float), rel_tol=1e-15, max_iter=10000)
    xr = np.array(k["dense10"]["x"])
    err = np.abs(x - xr) / (1e-20 + np.abs(x) + np.abs(xr))  # ERROR_UNIT_TEST of the reference test
    assert rc == 0 and err.max() < 1e-10


def test_ic0_factor_and_solution_match_reference(oracle):
    g = np.load(os.path.join(GOLDEN, "ref_tinyblock_pcg_tetra11.npz"))
    rp, col, vals, F = g["row_ptr"], g["col"], g["vals"], g["F"]
    L, rc = oracle.block_ic0(rp, col, vals)
    assert rc == 0
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    low = col <= rows
    assert np.array_equal(g["present"], low)  # no block of this K is entirely zero
    assert np.abs(L[low] - g["factor"][low]).max() <= 1e-13 * np.abs(g["factor"][low]).max()
    x, it, rc = oracle.pcg_ic0(rp, col, vals, L, F, crit=float(g["crit"]))
    assert rc == 0
    assert it == int(g["iters"])
    assert np.linalg.norm(x - g["x"]) <= 1e-12 * np.linalg.norm(g["x"])


def test_oracle_config0_system_matches_golden_inputs(oracle):
    g = np.load(os.path.join(GOLDEN, "ref_tinyblock_pcg_tetra11.npz"))
    c, e = oracle.make_rect(oracle.TETRA, (11, 11, 11))
    faces = oracle.skin_faces(oracle.TETRA, len(c), e)
    fx = faces[np.all(c[faces][:, :, 0] > 1 - 1e-6, axis=1)]
    rp, col, vals, F, rc = oracle.assemble(oracle.TETRA, c, e, 210000.0, 0.3, faces=fx, face_trac=(0, 0, -1))
    clamp = np.where(c[:, 0] < 1e-6)[0]
    oracle.penalty(rp, col, vals, F, (3 * clamp[:, None] + np.arange(3)).ravel(), None, 1e5)
    assert np.array_equal(vals, g["vals"]) and np.array_equal(F, g["F"])
    assert np.allclose(F.reshape(-1, 3).sum(0), [0, 0, -1], atol=1e-14)


def test_jacobi_pcg_matches_dense_solve(oracle):
    g = np.load(os.path.join(GOLDEN, "ref_tinyblock_pcg_tetra11.npz"))
    rp, col, vals, F = g["row_ptr"], g["col"], g["vals"], g["F"]
    n = 3 * (len(rp) - 1)
    A = np.zeros((n, n))
    for r in range(len(rp) - 1):
        for p in range(rp[r], rp[r + 1]):
            A[3 * r:3 * r + 3, 3 * col[p]:3 * col[p] + 3] = vals[p]
    assert np.array_equal(A, A.T)
    xs = np.linalg.solve(A, F)
    for block in (False, True):
        x, it, rr, rc = oracle.pcg_jacobi(rp, col, vals, F, rel_tol=1e-13, block=block)
        assert rc == 0 and rr <= 1e-13
        assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)
    # the reference's IC(0)-PCG solution agrees with the dense solve to its stop test
    assert np.linalg.norm(g["x"] - xs) <= 1e-9 * np.linalg.norm(xs)
