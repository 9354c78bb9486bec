"""The benchmarked workloads through the exact paths the benchmark times, against the oracle.

* Config 1 (Hexa8 lattice, lognormal E with sigma 0.5, clamp x=0, traction on x=1) at 40^3 cells,
  the largest size the oracle's FP64 block-Jacobi PCG solves in seconds on the host cores: the
  box-only matrix-free product in both precisions, block-Jacobi PCG, MIXED and FP64 at the
  benchmark's relative-residual tolerance 1e-8 -> relative L2 error <= 1e-8 against the oracle's
  solution; FP64 at 1e-11 -> <= 1e-10.
* Config 2 (jittered Tetra_10, body force + traction) at 15^3 cells: the matrix-free Tetra_10
  product, Jacobi PCG, the same bounds.
* Config 3 (Hexa8 64^3, smoothed modulus field, 5% of the top face observed): the adjoint
  gradient against central finite differences (step 1e-4 E_e) of the objective on 32 seeded
  elements, relative error <= 1e-5 per element where |g_e| >= 1e-3 max|g|; and the adjoint against
  the oracle's FP64 adjoint (its own solves and element matrices) at 16^3.  The gradient is taken
  at a modulus field near the initial 210000 (x exp(0.05 N(0,1)), seeded): at E = 210000
  everywhere every interior node of the lattice holds the largest diagonal entry, so raising any
  element's modulus raises the penalty scale max|diag K| (formulation.h:1029-1030) while lowering
  it does not -- J has a kink there, which the adjoint (penalty scale held fixed) does not see.
  Elements whose perturbation moves max|diag K| are left out for the same reason.

The oracle's solutions are checked for their own true residual before use.  Finite differences
use J(E +- h) from the perturbed solutions u+- = u0 + d+-, where d+- solves K(E +- h) d = -+h
K_e(1) u0 with the oracle's unit-modulus element matrix as the load (problems.fd_gradient): the
changes are then solved to a tight relative residual of their own, without the cancellation
of K u against F that limits a warm-started solve of the whole system.
"""
import os
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from problems import E0, NU0, build, config1, config2, config3, fd_gradient

pytestmark = pytest.mark.gpu

MF = (A.SPMV_MATRIX_FREE, 2, A.LOADS_NT)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def O(oracle):
    oracle.set_threads(min(16, os.cpu_count() or 1))
    yield oracle
    oracle.set_threads(1)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_solution(O, kind, c, e, E, bf, faces, tr, dofs, block, tol=1e-12):
    rp, col, vals, F, rc = O.assemble(kind, c, e, E, NU0, bf=bf, faces=faces, face_trac=tr)
    assert rc == 0
    O.penalty(rp, col, vals, F, dofs, None, 1e5)
    x, it, rr, rc = O.pcg_jacobi(rp, col, vals, F, rel_tol=tol, block=block)
    assert rc == 0
    true_res = np.linalg.norm(F - O.spmv(rp, col, vals, x)) / np.linalg.norm(F)
    say("oracle: %d iterations, true rel. residual %.2e" % (it, true_res))
    assert true_res <= 1e-10
    return x


def gpu_solutions(f, ndof):
    out = {}
    for name, prec, tol in (("mixed_1e-8", L.MIXED, 1e-8), ("fp64_1e-8", L.FP64, 1e-8), ("fp64_1e-11", L.FP64, 1e-11)):
        f.precision = prec
        f.set_result(np.zeros(ndof))
        assert f.solve_system(rel_tol=tol), (name, f.last_stats.status)
        assert f.last_stats.rel_res <= tol, (name, f.last_stats.rel_res)
        out[name] = (f.get_result(), f.last_stats)
    return out


def check_against(x, sols):
    for name, (u, st) in sols.items():
        err = rel(u, x)
        say("%s: %d iterations, %d true-residual checks, rel. res %.2e, rel. err vs oracle %.2e" % (
            name, st.iters, st.refreshes, st.rel_res, err))
        assert err <= (1e-10 if name == "fp64_1e-11" else 1e-8), (name, err)


def test_config1_headline_path_matches_oracle(O):
    """Box-only matrix-free product, block-Jacobi PCG, Hexa8 40^3 of the config-1 workload."""
    n = 40
    kind, c, e, E = config1(n=n)
    say("config-1 workload, Hexa8 %d^3" % n)
    f, faces, tr, dofs = build(kind, c, e, E)
    f.assemble()
    f.set_spmv_config(fp64=MF, fp32=MF)
    info = f.matrix_free_info()
    assert info["box_elements"] == len(e)          # every element a box: the box-only kernel serves
    f.precond = L.BLOCK_JACOBI
    sols = gpu_solutions(f, 3 * len(c))
    x = oracle_solution(O, kind, c, e, E, (0.0, 0.0, 0.0), faces, tr, dofs, block=True)
    check_against(x, sols)
    f.close()


def test_config2_workload_matches_oracle(O):
    """Matrix-free Tetra_10 product, Jacobi PCG, jittered Tetra_10 15^3 with body force."""
    kind, c, e, E, bf = config2(cells=15)
    say("config-2 workload, Tetra_10 15^3, %d DOF" % (3 * len(c)))
    f, faces, tr, dofs = build(kind, c, e, E, bf=bf)
    f.assemble()
    f.set_spmv_config(fp64=MF, fp32=MF)
    f.precond = L.JACOBI
    sols = gpu_solutions(f, 3 * len(c))
    x = oracle_solution(O, kind, c, e, E, bf, faces, tr, dofs, block=False)
    check_against(x, sols)
    f.close()


def _identification(n):
    """Base state of the config-3 problem at the initial guess E = 210000: the formulation, the
    solution u0, the residual at the observations and dJ/du."""
    c, e, E_true, odofs, _ = config3(n=n)
    f, faces, tr, dofs = build(L.HEXA, c, e, E_true)
    f.precond = L.BLOCK_JACOBI
    f.set_result(np.zeros(3 * len(c)))
    f.assemble()
    assert f.solve_system(rel_tol=1e-10)
    u_meas = f.get_result()[odofs]
    E = E0 * np.exp(0.05 * np.random.default_rng(11).standard_normal(len(e)))
    f.set_material(E, NU0)
    f.assemble()
    f.set_result(np.zeros(3 * len(c)))
    assert f.solve_system(rel_tol=3e-11)
    u0 = f.get_result()
    res = u0[odofs] - u_meas
    dJdu = np.zeros_like(u0)
    dJdu[odofs] = res
    return c, e, E, odofs, f, faces, tr, dofs, u0, res, dJdu


def test_config3_adjoint_matches_central_differences(O):
    n = 64
    say("config-3 workload, Hexa8 %d^3 adjoint vs finite differences" % n)
    c, e, E, odofs, f, faces, tr, dofs, u0, res, dJdu = _identification(n)
    g = f.adjoint(dJdu, rel_tol=1e-10)
    gmax = np.abs(g).max()
    rng = np.random.default_rng(7)
    big = np.where(np.abs(g) >= 1e-3 * gmax)[0]
    els = np.concatenate([rng.choice(len(e), size=16, replace=False), rng.choice(big, size=16, replace=False)])
    checked = skipped = 0
    for el in els:
        fd = fd_gradient(f, O, E, u0, res, odofs, el)
        if fd is None:
            skipped += 1
            continue
        if abs(g[el]) >= 1e-3 * gmax:
            err = abs(fd - g[el]) / abs(g[el])
            assert err <= 1e-5, (int(el), fd, g[el], err)
            checked += 1
        else:
            err = abs(fd - g[el]) / gmax
            assert err <= 1e-8, (int(el), fd, g[el], err)
    say("%d elements at |g| >= 1e-3 max|g| within 1e-5 relative, %d small ones within 1e-8 max|g|, %d next to "
        "the largest diagonal left out" % (checked, len(els) - checked - skipped, skipped))
    assert checked >= 14
    f.close()


def test_config3_adjoint_matches_oracle_adjoint(O):
    """The FP64 adjoint at 16^3 against the oracle's: its K (penalty included), its block-Jacobi
    solves of K u = F and K lambda = dJ/du, and its element gradient -lambda_e^T K_e(1) u_e."""
    n = 16
    say("config-3 workload, Hexa8 %d^3 adjoint vs the oracle's" % n)
    c, e, E, odofs, f, faces, tr, dofs, u0, res, dJdu = _identification(n)
    g = f.adjoint(dJdu, rel_tol=1e-10)
    rp, col, vals, F, rc = O.assemble(L.HEXA, c, e, E, NU0, faces=faces, face_trac=tr)
    O.penalty(rp, col, vals, F, dofs, None, 1e5)
    u, _, _, rc1 = O.pcg_jacobi(rp, col, vals, F, rel_tol=1e-12, block=True)
    ro = np.zeros_like(u)
    ro[odofs] = u[odofs] - u0[odofs] + res      # the oracle's own residual at the observations
    lam, _, _, rc2 = O.pcg_jacobi(rp, col, vals, ro, rel_tol=1e-12, block=True)
    assert rc1 == 0 and rc2 == 0
    go = O.elem_grad_E(L.HEXA, e, c, E, NU0, u, lam)
    big = np.abs(go) >= 1e-3 * np.abs(go).max()
    err = np.abs(g - go)[big] / np.abs(go)[big]
    say("adjoint vs oracle: max rel. err %.2e over %d elements, max abs %.2e of max|g|" % (
        err.max(), big.sum(), np.abs(g - go).max() / np.abs(go).max()))
    assert err.max() <= 1e-6
    assert np.abs(g - go).max() <= 1e-8 * np.abs(go).max()
    f.close()
