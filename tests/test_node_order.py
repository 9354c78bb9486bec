"""Cuthill-McKee node order (femgpu_node_order, host only) against the reference's symrcm.

The golden orders are the reference's symrcm(m) (util/symrcm.cpp:38-101 with
util/GetLinkForOrdering.h), compiled from source, on make_rect Tetra and Hexa meshes; the library
must give the same visiting order.  Further checks: a permutation of the nodes, a breadth-first
visit (every node after the first is adjacent to an earlier one within its component), a start at
the lowest degree, unreached components restarted, and a profile (sum of squared row spans, the
reference's symrcm_cost) much smaller than the one of a shuffled numbering.  No GPU compute.
"""
import os

import numpy as np

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
