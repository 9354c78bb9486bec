"""Oracle mesh / pattern checks against the reference's own make_rect output (golden fixtures)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

CASES = [("tetra", 0, (4, 4, 4)), ("tetra", 0, (4, 3, 5)), ("hexa", 2, (3, 4, 5)), ("tetra10", 1, (3, 3, 3)),
         ("tetra10", 1, (4, 3, 2))]


@pytest.mark.parametrize("name,kind,nb", CASES)
def test_make_rect_matches_reference(oracle, name, kind, nb):
    ref = np.load(os.path.join(GOLDEN, "ref_mesh_%s_%d_%d_%d.npz" % ((name,) + nb)))
    c, e = oracle.make_rect(kind, nb)
    # bitwise: accumulated coordinates (util/rectilinear_iterator.h:40-51) and first-appearance order
    assert np.array_equal(c, ref["coords"])
    assert np.array_equal(e, ref["conn"])


def test_tetra10_hash_equals_linear_search(oracle):
    for nb in [(2, 2, 2), (3, 2, 4), (4, 4, 3)]:
        c1, e1 = oracle.make_rect(oracle.TETRA10, nb)
        c2, e2 = oracle.make_rect_tetra10_linear_search(nb)
        assert np.array_equal(c1, c2) and np.array_equal(e1, e2)


def _brute_pattern(n_nodes, conn):
    rows = [set([i]) for i in range(n_nodes)]
    for el in conn:
        for a in el:
            rows[a].update(int(b) for b in el)
    return rows


@pytest.mark.parametrize("kind,nb", [(0, (4, 3, 5)), (1, (3, 3, 2)), (2, (4, 4, 4))])
def test_pattern_is_node_adjacency(oracle, kind, nb):
    c, e = oracle.make_rect(kind, nb)
    rp, col = oracle.pattern(kind, len(c), e)
    rows = _brute_pattern(len(c), e)
    for r in range(len(c)):
        assert list(col[rp[r]:rp[r + 1]]) == sorted(rows[r])


def test_pattern_counts_match_published_counts(oracle):
    # config 0: 17,191 full / 9,261 lower blocks; Hexa 21^3 (the published mesh): 118,121 lower
    # blocks = 1,035,306 lower scalars (include/mesh/exemple_probleme_elastique.txt:218-219)
    c, e = oracle.make_rect(oracle.TETRA, (11, 11, 11))
    rp, col = oracle.pattern(oracle.TETRA, len(c), e)
    rows = np.repeat(np.arange(len(c)), np.diff(rp))
    assert len(col) == 17191 and (col <= rows).sum() == 9261
    c, e = oracle.make_rect(oracle.HEXA, (21, 21, 21))
    rp, col = oracle.pattern(oracle.HEXA, len(c), e)
    rows = np.repeat(np.arange(len(c)), np.diff(rp))
    lower_blocks = int((col < rows).sum())
    assert lower_blocks + len(c) == 118121
    assert 9 * lower_blocks + 6 * len(c) == 1035306


def test_counts_of_configs(oracle):
    import ctypes as C
    for kind, nb, nn, ne in [(0, (11, 11, 11), 1331, 6000), (2, (101, 101, 101), 1030301, 1000000),
                             (1, (41, 41, 41), 531441, 384000), (2, (65, 65, 65), 274625, 262144)]:
        a, b = C.c_uint32(), C.c_uint32()
        oracle.lib().or_make_rect_counts(kind, (C.c_uint * 3)(*nb), C.byref(a), C.byref(b))
        assert (a.value, b.value) == (nn, ne)


def test_skin_faces_counts(oracle):
    c, e = oracle.make_rect(oracle.TETRA, (11, 11, 11))
    f = oracle.skin_faces(oracle.TETRA, len(c), e)
    assert len(f) == 6 * 200
    fx = f[np.all(c[f][:, :, 0] > 1 - 1e-6, axis=1)]
    assert len(fx) == 200
    c, e = oracle.make_rect(oracle.HEXA, (5, 4, 3))
    f = oracle.skin_faces(oracle.HEXA, len(c), e)
    assert len(f) == 2 * (4 * 3 + 4 * 2 + 3 * 2)
