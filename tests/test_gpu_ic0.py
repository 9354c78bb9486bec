"""GPU block IC(0) preconditioner (ic0.hip) through the C ABI.

Natural order: the factor is the reference's own chol_incomp factor (fixture written by the
reference's MatWithTinyBlocks compiled from source), and the PCG with the reference's stop test
max|r_i| <= crit reproduces the reference's iteration count and solution.  Coloured order: the
factor equals the oracle's block IC(0) of the system renumbered by the same elimination order.
Tolerances: factors to 1e-13 of their largest entry (summation order differs from the
reference's SIMD pairs); solutions as in test_gpu_parity (1e-10 FP64, 1e-8 MIXED, relative).
"""
import os
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from conftest import GOLDEN
from problems import build, lognormal_E, permute_bsr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _fixture():
    return np.load(os.path.join(GOLDEN, "ref_tinyblock_pcg_tetra11.npz"))


def _to_natural(order, rp, col, Lt):
    """Blocks of a factor in elimination numbering, keyed by natural (row, col) node pairs."""
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    return {(int(order[t]), int(order[c])): Lt[i] for i, (t, c) in enumerate(zip(rows, col)) if c <= t}


def test_ic0_factor_is_the_reference_factor():
    say("ic0 factor vs reference chol_incomp")
    g = _fixture()
    s = L.BlockSystem(g["row_ptr"], g["col"], g["vals"], g["F"])
    order, rp, col, Lt, nl = s.get_ic0_factor(L.BLOCK_IC0)
    assert nl > 1
    blocks = _to_natural(order, rp, col, Lt)
    n = len(g["row_ptr"]) - 1
    rows = np.repeat(np.arange(n), np.diff(g["row_ptr"]))
    low = g["col"] <= rows
    assert len(blocks) == int(low.sum())
    mine = np.array([blocks[(int(r), int(c))] for r, c in zip(rows[low], g["col"][low])])
    ref = g["factor"][low]
    assert np.abs(mine - ref).max() <= 1e-13 * np.abs(ref).max()


def test_ic0_pcg_reproduces_reference_solve():
    say("ic0 pcg vs reference solve")
    g = _fixture()
    s = L.BlockSystem(g["row_ptr"], g["col"], g["vals"], g["F"])
    s.precond = L.BLOCK_IC0
    s.precision = L.FP64
    assert s.solve_system(crit=float(g["crit"])), s.last_stats.status
    x = s.get_result()
    assert abs(s.last_stats.iters - int(g["iters"])) <= 1, s.last_stats.iters
    assert s.last_stats.max_abs_res <= float(g["crit"])
    assert np.linalg.norm(x - g["x"]) <= 1e-10 * np.linalg.norm(g["x"])


def test_ic0_coloured_factor_matches_oracle(oracle):
    say("ic0 coloured factor vs oracle")
    g = _fixture()
    s = L.BlockSystem(g["row_ptr"], g["col"], g["vals"], g["F"])
    order, rp, col, Lt, nl = s.get_ic0_factor(L.BLOCK_IC0_COLOURED)
    assert nl <= 16
    prp, pcol, pvals = permute_bsr(g["row_ptr"], g["col"], g["vals"], order)
    assert np.array_equal(prp, rp) and np.array_equal(pcol, col)
    Lo, rc = oracle.block_ic0(prp, pcol, pvals)
    assert rc == 0
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    low = col <= rows
    assert np.abs(Lt[low] - Lo[low]).max() <= 1e-13 * np.abs(Lo[low]).max()
    assert not Lt[~low].any()


@pytest.mark.parametrize("precond", [L.BLOCK_IC0, L.BLOCK_IC0_COLOURED])
@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_ic0_pcg_on_mesh_matches_oracle(oracle, precond, precision):
    say("ic0 mesh solve", precond, precision)
    c, e = L.make_rect(L.HEXA, (7, 6, 6))
    E = lognormal_E(len(e), 0.5, 3)
    f, faces, tr, dofs = build(L.HEXA, c, e, E)
    f.assemble()
    rp, col, vals = f.get_mat()
    F = f.get_sollicitation()
    x_ref, it_j, rr, rc = oracle.pcg_jacobi(rp, col, vals, F, rel_tol=1e-13)
    assert rc == 0
    f.precond, f.precision = L.JACOBI, L.FP64
    f.set_result(np.zeros_like(F))
    assert f.solve_system(rel_tol=1e-11)
    it_jacobi = f.last_stats.iters
    f.precond, f.precision = precond, precision
    f.set_result(np.zeros_like(F))
    assert f.solve_system(rel_tol=1e-11), f.last_stats.status
    u = f.get_result()
    assert f.last_stats.rel_res <= 1e-11
    tol = 1e-10 if precision == L.FP64 else 1e-8
    assert np.linalg.norm(u - x_ref) <= tol * np.linalg.norm(x_ref)
    assert f.last_stats.iters < it_jacobi, (f.last_stats.iters, it_jacobi)


def test_ic0_batch_matches_individual_solves():
    """The coloured schedule of a batch restricted to one sub-problem is that sub-problem's own
    schedule (first-fit colours and levels only see its nodes), so results are bitwise equal."""
    say("ic0 batch")
    from problems import clamp_and_traction
    rng = np.random.default_rng(7)
    items = []
    for i in range(3):
        n = int(rng.integers(3, 6))
        c, e = L.make_rect(L.HEXA, (n + 1,) * 3)
        items.append((c, e, lognormal_E(len(e), 0.5, 10 + i)))
    b = L.Batch(L.HEXA, [(c, e) for c, e, _ in items])
    ref = []
    for i, (c, e, E) in enumerate(items):
        faces, tr, dofs = clamp_and_traction(L.HEXA, c, e)
        b.set_material(i, E, 0.3)
        b.set_traction(i, faces, tr)
        b.add_constraints(i, dofs, None, 1e5)
        f, *_ = build(L.HEXA, c, e, E)
        f.precond = L.BLOCK_IC0_COLOURED
        assert f.solve(rel_tol=1e-12)
        ref.append((f.get_result(), f.last_stats.iters))
    b.assemble()
    ok, ms = b.solve(rel_tol=1e-12, precision=L.FP64, precond=L.BLOCK_IC0_COLOURED)
    assert ok
    for i, (u, it) in enumerate(ref):
        assert b.last_stats[i].iters == it
        assert np.array_equal(b.get_result(i), u)


def test_ic0_reports_non_positive_pivot():
    say("ic0 not spd")
    rp = np.array([0, 2, 4], np.uint32)
    col = np.array([0, 1, 0, 1], np.uint32)
    I3 = np.eye(3)
    vals = np.stack([I3, 2 * I3, 2 * I3, I3])  # [[I, 2I], [2I, I]]: indefinite
    s = L.BlockSystem(rp, col, vals, np.ones(6))
    with pytest.raises(L.FemGpuError) as ei:
        s.get_ic0_factor(L.BLOCK_IC0)
    assert ei.value.status == L._capi.NOT_SPD
    s.precond = L.BLOCK_IC0
    assert not s.solve_system(rel_tol=1e-10)
    assert s.last_status == L._capi.NOT_SPD
