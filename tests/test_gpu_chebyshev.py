"""Chebyshev block-Jacobi preconditioner (FEMGPU_CHEBYSHEV_BLOCK_JACOBI) on the GPU, through the
C ABI.  It has no counterpart in the reference, so its bar is the system's solution: FP64
solutions agree with the oracle's PCG solution of the same assembled system to 1e-10 relative
(MIXED: 1e-8) on stored and matrix-free products, single problems and batches.  The
preconditioner itself is pinned by the iteration count of the numpy restatement
(tests/chebyshev_ref.py: the same start vector, bound and polynomial) -- within 10% plus 2 --
and by needing at most 80% of block Jacobi's iterations.
"""
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from chebyshev_ref import block_jacobi, coefficients, csr, pcg, power_bound
from lmtpp_amd import _capi as A
from problems import clamp_and_traction, jitter_interior, lognormal_E, tetra10_jittered

pytestmark = pytest.mark.gpu

MF = (A.SPMV_MATRIX_FREE, 2, A.LOADS_NT)
FULL = (A.SPMV_FULL, 2, A.LOADS_NT)
CHEB = L.CHEBYSHEV_BLOCK_JACOBI


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _mesh(kind, geometry, seed):
    if kind == L.TETRA10:
        c, e = tetra10_jittered(3, seed=seed)
    else:
        c, e = L.make_rect(kind, (9, 7, 7))
        if geometry == "jittered":
            c = jitter_interior(c, 1.0 / 8, 0.2, seed)
    return c, e, lognormal_E(len(e), 0.5, seed)


def _system(oracle, kind, c, e, E, trac=(0.0, 0.4, -1.0)):
    faces, tr, dofs = clamp_and_traction(kind, c, e, trac)
    with A.node_order_ctx(A.NODE_ORDER_NATURAL):
        f = L.Formulation(kind, c, e)
    f.set_material(E, 0.3, (0.1, -0.2, 0.3))
    f.set_traction(faces, tr)
    f.add_constraints(dofs, None, 1e5)
    f.assemble()
    rp, col, vals, F, rc = oracle.assemble(kind, c, e, E, 0.3, bf=(0.1, -0.2, 0.3), faces=faces, face_trac=tr)
    assert rc == 0
    oracle.penalty(rp, col, vals, F, dofs, None, 1e5)
    return f, (rp, col, vals, F)


@pytest.mark.parametrize("kind,geometry,products", [(L.TETRA, "jittered", (FULL,)),
                                                    (L.TETRA10, "jittered", (FULL, MF)),
                                                    (L.HEXA, "jittered", (FULL, MF)),
                                                    (L.HEXA, "lattice", (FULL, MF))])
def test_solutions_match_oracle_and_restatement(oracle, kind, geometry, products):
    say("chebyshev solve", kind, geometry)
    c, e, E = _mesh(kind, geometry, 5)
    f, (rp, col, vals, F) = _system(oracle, kind, c, e, E)
    x, _, _, rc = oracle.pcg_jacobi(rp, col, vals, F, rel_tol=1e-14, block=True)
    assert rc == 0
    Ac = csr(rp, col, vals)
    Dinv = block_jacobi(rp, col, vals)
    _, it_ref = pcg(Ac, Dinv, F, *coefficients(power_bound(Ac, Dinv)), rel_tol=1e-12)
    for cfg in products:
        f.set_spmv_config(fp64=cfg, fp32=cfg)
        iters = {}
        for pc in (L.BLOCK_JACOBI, CHEB):
            f.precond = pc
            for prec, tol in ((L.FP64, 1e-10), (L.MIXED, 1e-8)):
                say("  product", cfg[0], "precond", pc, "precision", prec)
                f.precision = prec
                f.set_result(np.zeros_like(F))
                assert f.solve_system(rel_tol=1e-12), (cfg, pc, prec, f.last_stats.status)
                u = f.get_result()
                assert np.linalg.norm(u - x) <= tol * np.linalg.norm(x), (cfg, pc, prec)
                iters[(pc, prec)] = f.last_stats.iters
        say("  iterations", iters, "restatement", it_ref)
        it = iters[(CHEB, L.FP64)]
        assert abs(it - it_ref) <= 0.1 * it_ref + 2, (it, it_ref)
        assert it <= 0.8 * iters[(L.BLOCK_JACOBI, L.FP64)], iters
        assert iters[(CHEB, L.MIXED)] <= 0.8 * iters[(L.BLOCK_JACOBI, L.MIXED)] + 5, iters
    f.close()


def test_deterministic_and_setup_reused(oracle):
    """Two solves give bitwise the same result; the second reuses the set-up (setup time 0); a
    new assembly builds it again."""
    say("chebyshev determinism")
    c, e, E = _mesh(L.HEXA, "jittered", 7)
    f, (rp, col, vals, F) = _system(oracle, L.HEXA, c, e, E)
    f.set_spmv_config(fp64=MF, fp32=MF)
    f.precond = CHEB
    f.precision = L.MIXED
    res = []
    for k in range(3):
        if k == 2:
            f.assemble()
        f.set_result(np.zeros_like(F))
        assert f.solve_system(rel_tol=1e-11)
        res.append((f.get_result(), f.last_stats.iters, f.precond_setup_ms()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][0], res[2][0])
    assert res[0][1] == res[1][1] == res[2][1]
    assert res[0][2] > 0 and res[1][2] == 0 and res[2][2] > 0
    f.close()


@pytest.mark.parametrize("precision,product", [(L.FP64, FULL), (L.MIXED, MF)])
def test_batch_matches_single_problems(oracle, precision, product):
    """Each sub-problem of a batch runs its own power iteration (start vector indexed within the
    sub-problem) and polynomial: bitwise the result and iteration count of the same problem
    solved alone, and the oracle's solution."""
    say("chebyshev batch", precision, product[0])
    rng = np.random.default_rng(23)
    items = []
    for i in range(4):
        div = tuple(int(v) for v in rng.integers(4, 8, size=3))
        c, e = L.make_rect(L.HEXA, div)
        if i % 2:
            c = jitter_interior(c, 1.0 / max(div), 0.2, 40 + i)
        items.append((c, e, lognormal_E(len(e), 0.5, 40 + i), float(rng.uniform(0.2, 0.4)), rng.standard_normal(3)))
    b = L.Batch(L.HEXA, [(c, e) for c, e, *_ in items])
    singles = []
    for i, (c, e, E, nu, tr) in enumerate(items):
        faces, trs, dofs = clamp_and_traction(L.HEXA, c, e, tr)
        b.set_material(i, E, nu)
        b.set_traction(i, faces, trs)
        b.add_constraints(i, dofs, None, 1e5)
        f = L.Formulation(L.HEXA, c, e)
        f.set_material(E, nu)
        f.set_traction(faces, trs)
        f.add_constraints(dofs, None, 1e5)
        f.assemble()
        f.set_spmv_config(fp64=product, fp32=product)
        f.precision, f.precond = precision, CHEB
        f.set_result(np.zeros(3 * len(c)))
        assert f.solve_system(rel_tol=1e-11)
        singles.append((f.get_result(), f.last_stats.iters))
        f.close()
        rp, col, vals, F, rc = oracle.assemble(L.HEXA, c, e, E, nu, faces=faces, face_trac=trs)
        oracle.penalty(rp, col, vals, F, dofs, None, 1e5)
        x, _, _, rc = oracle.pcg_jacobi(rp, col, vals, F, rel_tol=1e-14, block=True)
        assert rc == 0
        tol = 1e-9 if precision == L.FP64 else 1e-8
        assert np.linalg.norm(singles[-1][0] - x) <= tol * np.linalg.norm(x), i
    b.assemble()
    b.set_spmv_config(fp64=product, fp32=product)
    ok, ms = b.solve(rel_tol=1e-11, precision=precision, precond=CHEB)
    assert ok and ms > 0
    for i in range(len(items)):
        assert b.last_stats[i].iters == singles[i][1], i
        assert np.array_equal(b.get_result(i), singles[i][0]), i
    b.close()
