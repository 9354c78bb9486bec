"""Node orders (femgpu_node_order, femgpu_node_order_amd, host only) against the reference's
symrcm and symamd.

The golden orders are the reference's symrcm(m) (util/symrcm.cpp:38-101 with
util/GetLinkForOrdering.h), compiled from source, on make_rect Tetra and Hexa meshes; the library
must give the same visiting order.  Further checks: a permutation of the nodes, a breadth-first
visit (every node after the first is adjacent to an earlier one within its component), a start at
the lowest degree, unreached components restarted, and a profile (sum of squared row spans, the
reference's symrcm_cost) much smaller than the one of a shuffled numbering.

The golden AMD orders are the reference's symamd(m) (util/symamd.h:27-36) and symamd(ind)
(util/symamd.cpp:16-36) with the AMD 1.1 package the reference vendors (amd/, built with -DDINT
as its SConscript does), compiled from source, on make_rect Tetra, Tetra_10 and Hexa meshes and
on random symmetric graphs with rows denser than the package's threshold and rows with no
neighbour; the library must give the same positions.  Further checks from the mathematics: a
tree is ordered with no fill at all (minimum degree eliminates leaves first), and on a mesh the
fill of the Cholesky factor (counted by the elimination game) falls far below the natural
order's.  No GPU compute.
"""
import os

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from conftest import GOLDEN


def _graph(oracle, kind, nb, shuffle_seed=None):
    c, e = L.make_rect(kind, nb)
    if shuffle_seed is not None:
        perm = np.random.default_rng(shuffle_seed).permutation(len(c))
        new_of = np.empty(len(c), np.int64)
        new_of[perm] = np.arange(len(c))
        e = new_of[e].astype(np.uint32)
    return oracle.pattern(kind, len(c), e)


def _profile(rp, col, order):
    pos = np.empty(len(order), np.int64)
    pos[order] = np.arange(len(order))
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    lo = np.full(len(order), np.iinfo(np.int64).max)
    np.minimum.at(lo, pos[rows], pos[col])
    return float(((np.arange(len(order)) - lo + 1.0) ** 2).sum())


def test_matches_reference_symrcm(oracle):
    g = np.load(os.path.join(GOLDEN, "ref_symrcm.npz"))
    for key, kind in (("tetra_7_6_5", L.TETRA), ("hexa_7_6_5", L.HEXA)):
        rp, col = _graph(oracle, kind, (7, 6, 5))
        assert np.array_equal(A.node_order(rp, col), g[key]), key


def test_breadth_first_properties(oracle):
    rp, col = _graph(oracle, L.TETRA10, (6, 5, 5), shuffle_seed=5)
    n = len(rp) - 1
    o = A.node_order(rp, col).astype(np.int64)
    assert np.array_equal(np.sort(o), np.arange(n))
    deg = np.diff(rp) - 1
    assert deg[o[0]] == deg.min() and o[0] == np.argmin(deg)
    pos = np.empty(n, np.int64)
    pos[o] = np.arange(n)
    for k in range(1, n):                       # one component: each node joins from an earlier one
        v = o[k]
        assert pos[col[rp[v]:rp[v + 1]]].min() < k
    natural = _profile(rp, col, np.arange(n))
    assert _profile(rp, col, o) < 0.1 * natural    # 0.06 on this mesh


def test_components_are_restarted():
    # two paths 0-2-4 and 1-3 (rows hold themselves): every node numbered, second component restarted
    adj = {0: [0, 2], 1: [1, 3], 2: [0, 2, 4], 3: [1, 3], 4: [2, 4]}
    rp = np.cumsum([0] + [len(adj[i]) for i in range(5)]).astype(np.uint32)
    col = np.concatenate([adj[i] for i in range(5)]).astype(np.uint32)
    assert A.node_order(rp, col).tolist() == [0, 2, 4, 1, 3]


def _positions(order):
    pos = np.empty(len(order), np.int64)
    pos[order] = np.arange(len(order))
    return pos


def _fill(rp, col, order):
    """Entries the elimination game adds to the pattern when nodes are eliminated in order."""
    adj = [set(col[rp[i]:rp[i + 1]].tolist()) - {i} for i in range(len(rp) - 1)]
    added = 0
    for v in order:
        nb = list(adj[v])
        for a in nb:
            adj[a].discard(v)
        for x in range(len(nb)):
            for y in range(x + 1, len(nb)):
                if nb[y] not in adj[nb[x]]:
                    adj[nb[x]].add(nb[y])
                    adj[nb[y]].add(nb[x])
                    added += 1
    return added


def test_amd_matches_reference_symamd(oracle):
    g = np.load(os.path.join(GOLDEN, "ref_symamd.npz"))
    for key, kind, nb in (("tetra_7_6_5", L.TETRA, (7, 6, 5)), ("hexa_7_6_5", L.HEXA, (7, 6, 5)),
                          ("tetra10_4_3_3", L.TETRA10, (4, 3, 3)), ("hexa_13_12_11", L.HEXA, (13, 12, 11))):
        rp, col = _graph(oracle, kind, nb)
        assert np.array_equal(_positions(A.node_order_amd(rp, col)), g[key]), key
    for k in range(4):
        rp, col = g["graph%d_row_ptr" % k], g["graph%d_col" % k]
        assert np.array_equal(_positions(A.node_order_amd(rp, col)), g["graph%d_pos" % k]), k


def test_amd_fill_properties(oracle):
    # a random tree: no fill in the order
    rng = np.random.default_rng(4)
    n = 200
    par = [int(rng.integers(0, i)) for i in range(1, n)]
    adj = [{i} for i in range(n)]
    for i, p in enumerate(par, start=1):
        adj[i].add(p)
        adj[p].add(i)
    rp = np.cumsum([0] + [len(a) for a in adj]).astype(np.uint32)
    col = np.concatenate([sorted(a) for a in adj]).astype(np.uint32)
    o = A.node_order_amd(rp, col)
    assert np.array_equal(np.sort(o), np.arange(n))
    assert _fill(rp, col, o) == 0
    # a mesh: far less fill than the natural order
    rp, col = _graph(oracle, L.HEXA, (8, 8, 6))
    o = A.node_order_amd(rp, col)
    assert _fill(rp, col, o) < 0.6 * _fill(rp, col, np.arange(len(rp) - 1))


def test_amd_bad_input():
    rp = np.array([0, 2, 3], np.uint32)
    with pytest.raises(A.FemGpuError):      # 0 -> 1 without 1 -> 0
        A.node_order_amd(rp, np.array([0, 1, 1], np.uint32))
    with pytest.raises(A.FemGpuError):      # unsorted row
        A.node_order_amd(np.array([0, 2, 4], np.uint32), np.array([1, 0, 0, 1], np.uint32))
