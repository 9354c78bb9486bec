"""Block rows stored by length (SELL-C-sigma) against the natural order, through the C ABI.

The storage order is internal: the CSR view, K, F, the solution and the IC(0) factor must not
depend on it.  K and F agree to 1e-15 of their scale between orders and element-loop variants
(each entry is summed in the same colour order); solutions agree with the oracle to 1e-10 (FP64)
and 1e-8 (MIXED) relative; a batch sorted per sub-problem gives bitwise the same solutions as the
same meshes solved alone.
"""
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from problems import build, clamp_and_traction, jitter_interior, lognormal_E, tetra10_jittered

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _case(kind, seed):
    if kind == L.TETRA10:
        c, e = tetra10_jittered(4, seed=seed)
    else:
        c, e = L.make_rect(kind, (9, 7, 6))
        c = jitter_interior(c, 0.1, 0.15, seed)
    return c, e, lognormal_E(len(e), 0.5, seed)


def _built(kind, c, e, E, order):
    with A.row_order(order):
        f, *_ = build(kind, c, e, E, bf=(0.1, -0.2, 0.3))
    return f


@pytest.mark.parametrize("kind", [L.TETRA, L.TETRA10, L.HEXA])
def test_orders_give_the_same_system(oracle, kind):
    say("row order system", kind)
    c, e, E = _case(kind, 50 + kind)
    nat = _built(kind, c, e, E, A.ROW_ORDER_NATURAL)
    srt = _built(kind, c, e, E, A.ROW_ORDER_BY_LENGTH)
    o_nat, n_nat = nat.row_order()
    o_srt, n_srt = srt.row_order()
    assert o_nat == A.ROW_ORDER_NATURAL and o_srt == A.ROW_ORDER_BY_LENGTH and n_srt <= n_nat
    if kind == L.TETRA10:
        assert n_srt < 0.85 * n_nat
    ref = None
    for f in (nat, srt):
        _, rmax = f.assembly_rows()
        for R in (0, rmax, 8, A.ASSEMBLY_GATHER):
            f.set_assembly_rows(R)
            f.assemble()
            rp, col, vals = f.get_mat()
            F = f.get_sollicitation()
            if ref is None:
                ref = (rp, col, vals, F)
                continue
            assert np.array_equal(rp, ref[0]) and np.array_equal(col, ref[1])
            assert np.abs(vals - ref[2]).max() <= 1e-15 * np.abs(ref[2]).max(), R
            assert np.abs(F - ref[3]).max() <= 1e-15 * np.abs(ref[3]).max(), R
    x_ref, _, _, rc = oracle.pcg_jacobi(ref[0], ref[1], ref[2], ref[3], rel_tol=1e-13)
    assert rc == 0
    for f in (nat, srt):
        for fmt, ld in ((A.SPMV_FULL, A.LOADS_PLAIN), (A.SPMV_FULL, A.LOADS_NT), (A.SPMV_FULL, A.LOADS_NT_COL16),
                        (A.SPMV_SYM, A.LOADS_PLAIN)):
            f.set_spmv_config(fp64=(fmt, 2, ld), fp32=(fmt, 3, ld))
            for prec, tol in ((L.FP64, 1e-10), (L.MIXED, 1e-8)):
                f.precision, f.precond = prec, L.BLOCK_JACOBI
                f.set_result(np.zeros_like(ref[3]))
                assert f.solve_system(rel_tol=1e-11), (fmt, ld, prec, f.last_stats.status)
                assert np.linalg.norm(f.get_result() - x_ref) <= tol * np.linalg.norm(x_ref)
    nat.close()
    srt.close()


def test_ic0_factor_does_not_depend_on_the_order():
    say("row order ic0")
    c, e, E = _case(L.TETRA10, 61)
    out = []
    for order in (A.ROW_ORDER_NATURAL, A.ROW_ORDER_BY_LENGTH):
        f = _built(L.TETRA10, c, e, E, order)
        f.precond = L.BLOCK_IC0_COLOURED
        assert f.solve(rel_tol=1e-10)
        out.append((f.get_ic0_factor(L.BLOCK_IC0_COLOURED), f.get_result(), f.last_stats.iters))
        f.close()
    (fa, ua, ia), (fb, ub, ib) = out
    for a, b in zip(fa, fb):
        assert np.array_equal(a, b)
    assert abs(ia - ib) <= 2
    assert np.linalg.norm(ua - ub) <= 1e-9 * np.linalg.norm(ua)


def test_batch_sorted_per_sub_problem_matches_single_solves():
    say("row order batch")
    rng = np.random.default_rng(5)
    items = []
    for i in range(5):
        nb = tuple(int(v) for v in rng.integers(3, 9, size=3))
        c, e = L.make_rect(L.HEXA, nb)
        items.append((c, e, lognormal_E(len(e), 0.4, 70 + i), 0.25 + 0.02 * i))
    with A.row_order(A.ROW_ORDER_BY_LENGTH):
        b = L.Batch(L.HEXA, [(c, e) for c, e, *_ in items])
        singles = [L.Formulation(L.HEXA, c, e) for c, e, *_ in items]
    for i, (c, e, E, nu) in enumerate(items):
        faces, trs, dofs = clamp_and_traction(L.HEXA, c, e, (0.0, 0.5, -1.0))
        b.set_material(i, E, nu)
        b.set_traction(i, faces, trs)
        b.add_constraints(i, dofs, None, 1e5)
        f = singles[i]
        assert f.row_order()[0] == A.ROW_ORDER_BY_LENGTH
        f.set_material(E, nu)
        f.set_traction(faces, trs)
        f.add_constraints(dofs, None, 1e5)
        f.precision = L.MIXED
        assert f.solve(rel_tol=1e-11)
    b.assemble()
    ok, ms = b.solve(rel_tol=1e-11, precision=L.MIXED)
    assert ok
    for i, f in enumerate(singles):
        assert b.last_stats[i].iters == f.last_stats.iters
        assert np.array_equal(b.get_result(i), f.get_result())
        f.close()


def test_auto_order_sorts_tetra10_keeps_hexa():
    say("row order auto")
    c, e = tetra10_jittered(6, seed=1)
    f = L.Formulation(L.TETRA10, c, e)
    assert f.row_order()[0] == A.ROW_ORDER_BY_LENGTH
    f.close()
    c, e = L.make_rect(L.HEXA, (41, 41, 41))
    f = L.Formulation(L.HEXA, c, e)
    assert f.row_order() == (A.ROW_ORDER_NATURAL, f.row_order()[1])
    f.close()
