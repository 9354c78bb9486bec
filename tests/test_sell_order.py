"""Storage order of the block rows in the SELL-64 slices (femgpu_sell_order, host only).

A slice of 64 storage rows is as wide as its longest row.  Sorting the rows of every run of 1024
block rows by decreasing length groups rows of similar length into one slice.  These CPU tests
check the order the library picks -- a permutation that stays inside its 1024-row windows, rows
by decreasing length with ties in row order -- and the slice sizes, against row lengths fixed in
closed form on the Hexa lattice (27 for interior nodes, a product of 2s and 3s on the boundary)
and against the oracle's pattern on Tetra_10.  No GPU compute is called here.
"""
import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A

WINDOW = 1024


def _slices(lens):
    n = len(lens)
    pad = np.zeros(((n + 63) // 64) * 64, np.int64)
    pad[:n] = lens
    return int(pad.reshape(-1, 64).max(1).sum()) * 64


def _hexa_lengths(m):
    """Block-row lengths of the Hexa lattice with m points per side: a node couples to the
    3 x 3 x 3 box around it, cut at the boundary (2 per boundary direction)."""
    a = np.full(m, 3, np.int64)
    a[0] = a[-1] = 2
    k, j, i = np.meshgrid(a, a, a, indexing="ij")   # node index i + m j + m^2 k
    return (i * j * k).ravel()


def _check_order(lens, perm):
    n = len(lens)
    perm = perm.astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(n))
    t = np.arange(n)
    assert (perm // WINDOW == t // WINDOW).all()          # rows stay in their window
    for w0 in range(0, n, WINDOW):
        p = perm[w0:w0 + WINDOW]
        lw = lens[p]
        assert (np.diff(lw) <= 0).all()                    # decreasing length
        same = np.diff(lw) == 0
        assert (np.diff(p)[same] > 0).all()                # ties in row order


def _row_ptr(lens):
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)


def test_hexa_lattice_lengths_and_slices():
    m = 101
    lens = _hexa_lengths(m)
    assert lens.min() == 8 and lens.max() == 27 and lens.sum() == (3 * m - 2) ** 3
    perm, nat, npos = A.sell_order(_row_ptr(lens), A.ROW_ORDER_NATURAL)
    assert np.array_equal(perm, np.arange(len(lens)))
    assert nat == npos == _slices(lens) == 27508224      # the 1.009x padding of the 100^3 lattice
    perm, nat, npos = A.sell_order(_row_ptr(lens), A.ROW_ORDER_AUTO)
    assert np.array_equal(perm, np.arange(len(lens)))    # gain under 3%: natural order kept
    perm, nat, npos = A.sell_order(_row_ptr(lens), A.ROW_ORDER_BY_LENGTH)
    _check_order(lens, perm)
    assert npos == _slices(lens[perm]) <= nat


@pytest.mark.parametrize("cells", [3, 6])
def test_tetra10_order(oracle, cells):
    c, e = L.make_rect(L.TETRA10, (cells + 1,) * 3)
    rp, col = oracle.pattern(L.TETRA10, len(c), e)
    lens = np.diff(rp).astype(np.int64)
    perm, nat, npos = A.sell_order(rp, A.ROW_ORDER_BY_LENGTH)
    _check_order(lens, perm)
    assert nat == _slices(lens)
    assert npos == _slices(lens[perm])
    assert npos < 0.8 * nat                               # vertex and edge rows no longer share slices
    perm_auto, _, npos_auto = A.sell_order(rp, A.ROW_ORDER_AUTO)
    assert np.array_equal(perm_auto, perm) and npos_auto == npos


def test_small_and_uniform_profiles():
    # one row; uniform rows (nothing to gain); a profile whose sort only ties: natural order
    for lens in (np.array([5]), np.full(300, 7), np.arange(1, 2000) % 1 + 4):
        perm, nat, npos = A.sell_order(_row_ptr(lens), A.ROW_ORDER_BY_LENGTH)
        assert np.array_equal(perm, np.arange(len(lens))) and nat == npos
    # two windows with opposite trends: each sorted on its own
    lens = np.concatenate([np.arange(1024) % 50 + 1, 50 - np.arange(500) % 50])
    perm, nat, npos = A.sell_order(_row_ptr(lens), A.ROW_ORDER_BY_LENGTH)
    _check_order(lens, perm)
    assert npos < nat


def test_bad_input():
    with pytest.raises(A.FemGpuError):
        A.sell_order(np.array([0, 3, 2], np.uint32))
    with pytest.raises(A.FemGpuError):
        A.sell_order(np.array([0, 3], np.uint32), 7)
