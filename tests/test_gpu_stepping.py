"""Time steps, the non-linear loop and preconditioner reuse (Formulation::solve, shift,
get_precond / solve_system_using_precond, ProblemAncestor::iteration) through the C ABI.

The reference semantics pinned here:
* time bookkeeping in closed form -- set_initial_time_step(ts) puts time at -2 ts and each shift
  adds ts (formulation.h:424-433, 2624); iteration() updates only the formulations whose time
  equals the problem's time and then takes the smallest time (problem_ancestor.h:61-64, 124-127);
* with assume_constant_matrix, K and its preconditioner are built once and later steps
  reassemble only F: a load ramp solved after the modulus changed still matches the oracle's
  solve with the original K (problem_ancestor.h:46-53, 70-71, 83-87);
* the non-linear loop of a linear problem stops after the second solve, whose iterate no longer
  moves, and fails after max_non_linear_iteration + 1 solves (formulation.h:1526-1563).
Solutions agree with the oracle's to 1e-10 relative (FP64, relative residual 1e-12).
"""
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from problems import clamp_and_traction, jitter_interior, lognormal_E

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _mesh(seed):
    c, e = L.make_rect(L.HEXA, (6, 5, 4))
    c = jitter_interior(c, 0.2, 0.15, seed)
    return c, e, lognormal_E(len(e), 0.4, seed)


def _formulation(c, e, E, trac_scale=1.0, precond=L.BLOCK_IC0):
    faces, tr, dofs = clamp_and_traction(L.HEXA, c, e, (0.0, 0.3, -1.0))
    f = L.Formulation(L.HEXA, c, e)
    f.set_material(E, 0.3)
    f.set_traction(faces, trac_scale * tr)
    f.add_constraints(dofs, None, 1e5)
    f.precision, f.precond = L.FP64, precond
    return f, faces, tr, dofs


def _oracle_solution(O, c, e, E, faces, tr, dofs):
    rp, col, vals, F, rc = O.assemble(L.HEXA, c, e, E, 0.3, faces=faces, face_trac=tr)
    assert rc == 0
    O.penalty(rp, col, vals, F, dofs, None, 1e5)
    x, _, _, rc = O.pcg_jacobi(rp, col, vals, F, rel_tol=1e-14, block=True)
    assert rc == 0
    return x


def _close(u, x, tol=1e-10):
    return np.linalg.norm(u - x) <= tol * np.linalg.norm(x)


def test_time_bookkeeping_and_history():
    say("stepping time")
    c, e, E = _mesh(1)
    f, *_ = _formulation(c, e, E)
    assert f.get_time() == 0.0 and f.get_next_time_step() == 1e40
    f.set_initial_time_step(0.25)
    assert f.get_time() == -0.5 and f.get_next_time_step() == 0.25
    f.shift()
    f.shift(2)
    assert f.get_time() == 0.25
    f.set_time(3.0)
    f.set_next_time_step(0.5)
    assert (f.get_time(), f.get_next_time_step()) == (3.0, 0.5)
    assert f.solve(rel_tol=1e-12)           # solve() shifts first
    assert f.get_time() == 3.5
    u1 = f.get_result()
    assert np.array_equal(f.get_previous_result(), np.zeros_like(u1))
    assert f.solve(rel_tol=1e-12)
    assert np.array_equal(f.get_previous_result(), u1)
    assert f.last_stats.iters == 0          # warm start from vectors[0]: already converged
    f.close()


def test_constant_matrix_load_ramp_reuses_k_and_preconditioner(oracle):
    say("stepping constant matrix")
    c, e, E = _mesh(2)
    f, faces, tr, dofs = _formulation(c, e, E)
    f.assume_constant_matrix = True
    f.set_initial_time_step(1.0)
    P = L.Problem([f], rel_tol=1e-12)
    x1 = _oracle_solution(oracle, c, e, E, faces, tr, dofs)
    iters = []
    for k in (1, 2, 3):
        f.set_traction(faces, k * tr)
        if k == 3:
            f.set_material(2.0 * E, 0.3)     # ignored: K is assumed constant
        assert P.iteration() == [True]
        assert P.time == f.get_time() == k - 2.0
        assert _close(f.get_result(), k * x1), k
        iters.append(f.last_stats.iters)
        if k > 1:
            assert f.precond_setup_ms() == 0.0   # get_precond's factor, reused
    assert iters[1] <= iters[0] and iters[2] <= iters[0]
    # solve_system (not the kept preconditioner) uses the current K: reassembled with 2 E
    f.assemble()
    f.set_result(np.zeros(3 * f.n_nodes))
    assert f.solve_system(rel_tol=1e-12)
    x2 = _oracle_solution(oracle, c, e, 2.0 * E, faces, 3 * tr, dofs)
    assert _close(f.get_result(), x2)
    P.close()
    f.close()


def test_changing_matrix_is_reassembled(oracle):
    say("stepping matrix")
    c, e, E = _mesh(3)
    f, faces, tr, dofs = _formulation(c, e, E, precond=L.BLOCK_JACOBI)
    P = L.Problem([f], rel_tol=1e-12)
    assert P.iteration() == [True]
    assert _close(f.get_result(), _oracle_solution(oracle, c, e, E, faces, tr, dofs))
    f.set_material(3.0 * E, 0.3)
    assert P.iteration() == [True]
    assert _close(f.get_result(), _oracle_solution(oracle, c, e, 3.0 * E, faces, tr, dofs))
    assert f.nonlinear_solves() == 1
    P.close()
    f.close()


@pytest.mark.parametrize("through", ["solve", "iteration"])
def test_nonlinear_loop(oracle, through):
    say("stepping non-linear", through)
    c, e, E = _mesh(4)
    f, faces, tr, dofs = _formulation(c, e, E)
    x = _oracle_solution(oracle, c, e, E, faces, tr, dofs)
    f.non_linear_iterative_criterium = 1e-9
    P = L.Problem([f], rel_tol=1e-12)
    if through == "solve":
        assert f.solve(rel_tol=1e-12)
    else:
        P.iteration()
    assert f.nonlinear_solves() == 2          # the second solve no longer moves the iterate
    assert _close(f.get_result(), x)
    f.max_non_linear_iteration = 0
    f.set_result(np.zeros_like(x))
    with pytest.raises(L.FemGpuError):
        if through == "solve":
            f.solve(rel_tol=1e-12)
        else:
            P.iteration()
    assert f.nonlinear_solves() == 1
    P.close()
    f.close()


def test_problem_updates_formulations_due_at_its_time():
    say("stepping two formulations")
    c, e, E = _mesh(5)
    f1, *_ = _formulation(c, e, E)
    f2, *_ = _formulation(c, e, 1.5 * E)
    f1.set_initial_time_step(1.0)
    f2.set_initial_time_step(0.5)
    P = L.Problem([f1, f2], rel_tol=1e-12)
    assert P.time == 0.0
    # (updated flags, f1 time, f2 time, problem time) after each iteration
    expect = [([True, True], -1.0, -0.5, -1.0),
              ([True, False], 0.0, -0.5, -0.5),
              ([False, True], 0.0, 0.0, 0.0),
              ([True, True], 1.0, 0.5, 0.5)]
    for upd, t1, t2, t in expect:
        assert P.iteration() == upd
        assert (f1.get_time(), f2.get_time(), P.time) == (t1, t2, t)
    P.close()
    f1.close()
    f2.close()


def test_precond_reuse_errors():
    say("stepping errors")
    c, e, E = _mesh(6)
    f, *_ = _formulation(c, e, E)
    f.assemble()
    with pytest.raises(L.FemGpuError):
        f.solve_system_using_precond(rel_tol=1e-10)      # no get_precond yet
    f.get_precond(rel_tol=1e-10)
    f.precond = L.JACOBI
    with pytest.raises(L.FemGpuError):
        f.solve_system_using_precond(rel_tol=1e-10)      # another preconditioner kind
    f.precond = L.BLOCK_IC0
    f.set_result(np.zeros(3 * f.n_nodes))
    assert f.solve_system_using_precond(rel_tol=1e-10)
    f.close()


def test_rhs_only_assembly_equals_full_assembly():
    say("stepping rhs-only assembly")
    c, e, E = _mesh(7)
    f, faces, tr, dofs = _formulation(c, e, E)
    for bf in ((0.2, -0.1, 0.3), (0.0, 0.0, 0.0)):
        f.set_material(E, 0.3, bf)
        f.assemble()
        F_full = f.get_sollicitation()
        f.set_traction(faces, 2.0 * tr)
        f.assemble(False, True)
        f.set_traction(faces, tr)
        f.assemble(False, True)
        assert np.array_equal(f.get_sollicitation(), F_full), bf
    f.close()


@pytest.mark.parametrize("precond", [L.BLOCK_JACOBI, L.BLOCK_IC0, L.CHEBYSHEV_BLOCK_JACOBI])
def test_rhs_only_assembly_after_get_precond_keeps_penalty_scale(oracle, precond):
    """assemble(); get_precond(); assemble(False, True); solve_system_using_precond with non-zero
    prescribed displacements: the F-only assembly scales the penalty right-hand side with the
    max|diag K| of the last matrix assembly (formulation.h:1029-1030,1059-1062), so the solution is
    the oracle's and the constrained DOFs hold their values to O(1/P)."""
    say("stepping rhs-only after get_precond", precond)
    c, e, E = _mesh(8)
    f, faces, tr, dofs = _formulation(c, e, E, precond=precond)
    vals = 1e-4 * np.cos(np.arange(len(dofs)))
    f.clear_constraints()
    f.add_constraints(dofs, vals, 1e5)
    f.assemble()
    f.get_precond(rel_tol=1e-12)
    f.assemble(False, True)
    rp, col, ovals, F, rc = oracle.assemble(L.HEXA, c, e, E, 0.3, faces=faces, face_trac=tr)
    oracle.penalty(rp, col, ovals, F, dofs, vals, 1e5)
    # the penalty right-hand side M P value at the constrained DOFs (zero if M were lost)
    assert np.abs(f.get_sollicitation() - F).max() <= 1e-14 * np.abs(F).max()
    f.set_result(np.zeros(3 * f.n_nodes))
    assert f.solve_system_using_precond(rel_tol=1e-12)
    u = f.get_result()
    # ||F|| is dominated by the penalty loads M P value, so a relative residual of 1e-12 leaves
    # the free DOFs about 1e-8 accurate: compare with the dense solve at that level
    x = np.linalg.solve(_dense(rp, col, ovals), F)
    assert _close(u, x, 1e-6)
    assert np.abs(u[dofs] - vals).max() <= 1e-3 * np.abs(vals).max()
    f.close()


def _dense(rp, col, vals):
    n = 3 * (len(rp) - 1)
    A_ = np.zeros((n, n))
    for r in range(len(rp) - 1):
        for p in range(rp[r], rp[r + 1]):
            A_[3 * r:3 * r + 3, 3 * col[p]:3 * col[p] + 3] = vals[p]
    return A_
