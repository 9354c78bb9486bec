"""Host-side block IC(0) schedule (femgpu_ic0_schedule) and the invariance it relies on.

The GPU factorises and sweeps one level at a time, with rows renumbered by (level, elimination
position).  These CPU tests check that every level is a set of rows sharing no block, that a
row's earlier-eliminated neighbours sit in earlier levels, that the natural order keeps the
reference's relative order of every pair of coupled nodes (so its factor is the reference's),
and -- with the oracle's block IC(0) -- that the level renumbering leaves the factor unchanged.
No GPU compute is called here.
"""
import os

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from conftest import GOLDEN
from problems import permute_bsr


def _fixture():
    g = np.load(os.path.join(GOLDEN, "ref_tinyblock_pcg_tetra11.npz"))
    return g["row_ptr"], g["col"], g["vals"]


def _check_levels(rp, col, order, lv):
    n = len(rp) - 1
    assert np.array_equal(np.sort(order), np.arange(n))
    tof = np.empty(n, np.int64)
    tof[order] = np.arange(n)
    level_of_t = np.repeat(np.arange(len(lv) - 1), np.diff(lv))
    assert len(level_of_t) == n
    rows = np.repeat(np.arange(n), np.diff(rp))
    ta, tb = tof[rows], tof[col]
    off = ta != tb
    la, lb = level_of_t[ta[off]], level_of_t[tb[off]]
    assert (la != lb).all()                       # no coupling inside a level
    assert ((ta[off] < tb[off]) == (la < lb)).all()  # earlier-eliminated neighbours in earlier levels
    return tof


@pytest.mark.parametrize("precond", [A.BLOCK_IC0, A.BLOCK_IC0_COLOURED])
def test_levels_are_independent_sets(precond):
    rp, col, _ = _fixture()
    order, lv = A.ic0_schedule(rp, col, precond)
    tof = _check_levels(rp, col, order, lv)
    if precond == A.BLOCK_IC0:
        # natural order: every coupled pair (a, b), a < b, is eliminated a first
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        m = rows < col
        assert (tof[rows[m]] < tof[col[m]]).all()
    else:
        assert len(lv) - 1 <= 16


@pytest.mark.parametrize("n", [4, 6])
def test_hexa_lattice_level_counts(oracle, n):
    c, e = L.make_rect(L.HEXA, (n, n, n))
    rp, col = oracle.pattern(L.HEXA, len(c), e)
    order, lv = A.ic0_schedule(rp, col, A.BLOCK_IC0)
    _check_levels(rp, col, order, lv)
    # natural order on the 27-point lattice: level(i, j, k) = i + 2j + 4k
    assert len(lv) - 1 == 7 * (n - 1) + 1
    order, lv = A.ic0_schedule(rp, col, A.BLOCK_IC0_COLOURED)
    _check_levels(rp, col, order, lv)
    assert len(lv) - 1 == 8


def test_level_renumbering_keeps_the_factor(oracle):
    """The oracle's block IC(0) of the reference system in natural order equals, block for block,
    its IC(0) of the same system renumbered by the natural-order level schedule."""
    rp, col, vals = _fixture()
    Lnat, rc = oracle.block_ic0(rp, col, vals)
    assert rc == 0
    order, lv = A.ic0_schedule(rp, col, A.BLOCK_IC0)
    assert not np.array_equal(order, np.arange(len(rp) - 1))  # the renumbering is not trivial
    prp, pcol, pvals = permute_bsr(rp, col, vals, order)
    Lp, rc = oracle.block_ic0(prp, pcol, pvals)
    assert rc == 0
    n = len(rp) - 1
    tof = np.empty(n, np.int64)
    tof[order] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(rp))
    low = col <= rows
    # position of block (tof[r], tof[c]) in the permuted CSR
    prow = np.repeat(np.arange(n), np.diff(prp))
    key = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(prow, pcol))}
    idx = np.array([key[(int(tof[r]), int(tof[c]))] for r, c in zip(rows[low], col[low])])
    scale = np.abs(Lnat[low]).max()
    assert np.abs(Lp[idx] - Lnat[low]).max() <= 1e-13 * scale


def test_schedule_rejects_bad_patterns():
    rp = np.array([0, 1, 2], np.uint32)
    with pytest.raises(L.FemGpuError):
        A.ic0_schedule(rp, np.array([1, 0], np.uint32))  # no diagonal
    with pytest.raises(L.FemGpuError):
        A.ic0_schedule(rp, np.array([0, 1], np.uint32), precond=A.JACOBI)
