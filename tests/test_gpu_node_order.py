"""Nodes stored in Cuthill-McKee order (femgpu_set_node_order RCM, the reference's want_rcm) or in
approximate minimum degree order (AMD, the reference's want_amd) against the caller's order, through the C ABI, on meshes whose node numbering is shuffled.

The node order is internal: K (as the CSR view in the caller's numbering) and F must match the
oracle's assembly of the shuffled mesh (pattern exact, values to 1e-13 of the scale), solutions
the oracle's solve to 1e-10 (FP64) / 1e-8 (MIXED), and the reference's block IC(0) factor, the
after-solve fields, the adjoint gradient and the result file must not depend on the order (bitwise
for the factor, the after-solve fields and the file; 1e-9 for the gradient, which goes through a
solve).
"""
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from problems import clamp_and_traction, jitter_interior, lognormal_E, tetra10_jittered

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _shuffled(kind, seed):
    if kind == L.TETRA10:
        c, e = tetra10_jittered(3, seed=seed)
    else:
        c, e = L.make_rect(kind, (6, 5, 5))
        c = jitter_interior(c, 0.2, 0.15, seed)
    perm = np.random.default_rng(seed).permutation(len(c))
    new_of = np.empty(len(c), np.int64)
    new_of[perm] = np.arange(len(c))
    return c[perm].copy(), new_of[e].astype(np.uint32), lognormal_E(len(e), 0.4, seed)


def _make(kind, c, e, E, order, bf=(0.1, -0.2, 0.3)):
    faces, tr, dofs = clamp_and_traction(kind, c, e, (0.0, 0.4, -1.0))
    with A.node_order_ctx(order):
        f = L.Formulation(kind, c, e)
    f.set_material(E, 0.3, bf)
    f.set_traction(faces, tr)
    f.add_constraints(dofs, None, 1e5)
    return f, faces, tr, dofs


@pytest.mark.parametrize("kind,node_order", [(L.TETRA, A.NODE_ORDER_RCM), (L.TETRA10, A.NODE_ORDER_RCM),
                                             (L.HEXA, A.NODE_ORDER_RCM), (L.TETRA10, A.NODE_ORDER_AMD),
                                             (L.HEXA, A.NODE_ORDER_AMD)])
def test_reordered_system_and_solution_match_the_oracle(oracle, kind, node_order):
    say("node order system", kind, node_order)
    c, e, E = _shuffled(kind, 80 + kind)
    f, faces, tr, dofs = _make(kind, c, e, E, node_order)
    order, perm = f.node_order()
    assert order == node_order and np.array_equal(np.sort(perm), np.arange(len(c)))
    if node_order == A.NODE_ORDER_AMD:     # the internal order is the library's symamd order
        assert np.array_equal(perm, A.node_order_amd(*oracle.pattern(kind, len(c), e)))
    f.assemble()
    rp, col, vals = f.get_mat()
    F = f.get_sollicitation()
    orp, ocol, ovals, oF, rc = oracle.assemble(kind, c, e, E, 0.3, bf=(0.1, -0.2, 0.3), faces=faces, face_trac=tr)
    assert rc == 0
    oracle.penalty(orp, ocol, ovals, oF, dofs, None, 1e5)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.abs(vals - ovals).max() <= 1e-13 * np.abs(ovals).max()
    assert np.abs(F - oF).max() <= 1e-13 * np.abs(oF).max()
    x, _, _, rc = oracle.pcg_jacobi(orp, ocol, ovals, oF, rel_tol=1e-14, block=True)
    assert rc == 0
    for prec, tol, pc in ((L.FP64, 1e-10, L.BLOCK_JACOBI), (L.MIXED, 1e-8, L.BLOCK_JACOBI), (L.FP64, 1e-10, L.BLOCK_IC0)):
        f.precision, f.precond = prec, pc
        f.set_result(np.zeros_like(F))
        assert f.solve_system(rel_tol=1e-12), (prec, pc)
        u = f.get_result()
        assert np.linalg.norm(u - x) <= tol * np.linalg.norm(x), (prec, pc)
    # set_result / get_result round trip in the caller's order
    f.set_result(x)
    assert np.array_equal(f.get_result(), x)
    f.close()


def test_order_does_not_change_factor_recovery_gradient_or_file(tmp_path):
    say("node order invariants")
    kind = L.HEXA
    c, e, E = _shuffled(kind, 91)
    out = {}
    for order in (A.NODE_ORDER_NATURAL, A.NODE_ORDER_RCM, A.NODE_ORDER_AMD):
        f, *_ = _make(kind, c, e, E, order)
        f.precond = L.BLOCK_IC0
        assert f.solve(rel_tol=1e-12)
        u = f.get_result()
        fac = f.get_ic0_factor(L.BLOCK_IC0)
        if order == A.NODE_ORDER_NATURAL:
            u_common = u
        f.set_result(u_common)
        eps, sig, tr = f.call_after_solve()
        path = tmp_path / ("r%d.vtu" % order)
        f.write_vtu(path)
        rng = np.random.default_rng(3)
        g = f.adjoint(rng.standard_normal(u.size), rel_tol=1e-13)
        out[order] = (u, fac, eps, sig, tr, path.read_bytes(), g)
        f.close()
    a = out[A.NODE_ORDER_NATURAL]
    for b in (out[A.NODE_ORDER_RCM], out[A.NODE_ORDER_AMD]):
        assert np.linalg.norm(a[0] - b[0]) <= 1e-10 * np.linalg.norm(a[0])
        for x, y in zip(a[1], b[1]):
            assert np.array_equal(x, y)                    # the reference's factor in every order
        for x, y in zip(a[2:5], b[2:5]):
            assert np.array_equal(x, y)
        assert a[5] == b[5]
        assert np.linalg.norm(a[6] - b[6]) <= 1e-9 * np.linalg.norm(a[6])


def test_batch_members_renumbered_like_single_problems():
    say("node order batch")
    items = []
    for i in range(3):
        c, e, E = _shuffled(L.HEXA, 100 + i)
        items.append((c, e, E))
    with A.node_order_ctx(A.NODE_ORDER_RCM):
        b = L.Batch(L.HEXA, [(c, e) for c, e, _ in items])
    singles = []
    for i, (c, e, E) in enumerate(items):
        faces, tr, dofs = clamp_and_traction(L.HEXA, c, e, (0.2, 0.0, -1.0))
        b.set_material(i, E, 0.3)
        b.set_traction(i, faces, tr)
        b.add_constraints(i, dofs, None, 1e5)
        f, *_ = _make(L.HEXA, c, e, E, A.NODE_ORDER_RCM, bf=(0.0, 0.0, 0.0))
        f.set_traction(faces, tr)
        f.clear_constraints()
        f.add_constraints(dofs, None, 1e5)
        f.precision = L.MIXED
        assert f.solve(rel_tol=1e-11)
        singles.append(f)
    b.assemble()
    ok, _ = b.solve(rel_tol=1e-11, precision=L.MIXED)
    assert ok
    for i, f in enumerate(singles):
        assert np.array_equal(b.get_result(i), f.get_result())
        f.close()
