"""Members of the reference's FormulationAncestor boundary beyond assemble/solve, through the C ABI:

* add_constraint(txt, penalty) (formulation_ancestor.h:197, constraint.h:46-93): the penalty terms
  of assemble_constraints (formulation.h:1058-1070) for c u + b = 0 -- K += M P c^2, F += M P c (-b)
  -- against the oracle's penalty with value -b/c and penalty P c^2; bitwise the numeric form for
  a unit coefficient; time-dependent expressions re-evaluated at each assembly; couplings of
  several unknowns and non-linear expressions refused;
* shift / unshift and get_result(num) over the nb_vectors = 20 history (formulation.h:424-445,
  2648), time_steps bookkeeping in closed form;
* get_indice_noda (formulation_ancestor.h:188) = 3 n in the API's numbering, with either node order;
* get_nodal_forces (formulation.h:447-458) = the element and skin loads without constraint terms
  (the oracle's F before or_penalty), the assembled sollicitation untouched;
* set_sollicitation: the right-hand side of the next solve.
"""
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from problems import NU0, clamp_and_traction, jitter_interior, lognormal_E

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _problem(seed=3):
    c, e = L.make_rect(L.HEXA, (5, 4, 4))
    c = jitter_interior(c, 0.25, 0.15, seed)
    E = lognormal_E(len(e), 0.4, seed)
    faces, tr, dofs = clamp_and_traction(L.HEXA, c, e, (0.1, 0.0, -1.0))
    f = L.Formulation(L.HEXA, c, e)
    f.set_material(E, NU0, (0.0, 0.2, -0.4))
    f.set_traction(faces, tr)
    return f, c, e, E, faces, tr, dofs


def test_text_constraints_match_penalty_terms(oracle):
    say("text constraints")
    f, c, e, E, faces, tr, dofs = _problem()
    # unit coefficients: bitwise the numeric form
    g, *_ = _problem()
    for d in dofs:
        f.add_constraint("node[%d].dep[%d]" % (d // 3, d % 3), 1e5)
    g.add_constraints(dofs, None, 1e5)
    f.assemble()
    g.assemble()
    assert np.array_equal(f.get_mat()[2], g.get_mat()[2])
    assert np.array_equal(f.get_sollicitation(), g.get_sollicitation())
    g.close()
    # a coefficient and a constant: value -b/c with penalty P c^2 in the oracle's penalty
    num = f.add_constraint("2*node[7].dep[1] - 0.004", 1e5)
    assert num == len(dofs)
    f.assemble()
    rp, col, ov, oF, _ = oracle.assemble(L.HEXA, c, e, E, NU0, bf=(0.0, 0.2, -0.4), faces=faces, face_trac=tr)
    M = oracle.penalty(rp, col, ov, oF, dofs, None, 1e5)
    # u[22] = 0.002 with penalty 4e5, scaled by the same M (max|diag K| before any constraint)
    p77 = rp[7] + np.searchsorted(col[rp[7]:rp[8]], 7)
    ov[p77][1, 1] += M * 4e5
    oF[22] += M * 4e5 * 0.002
    _, _, vals = f.get_mat()
    assert np.abs(vals - ov).max() <= 1e-13 * np.abs(ov).max()
    assert np.abs(f.get_sollicitation() - oF).max() <= 1e-13 * np.abs(oF).max()
    assert abs(f.max_diag() - M) <= 1e-13 * M
    assert f.solve_system(rel_tol=1e-11)
    assert abs(f.get_result()[22] - 0.002) <= 1e-6 * 0.002
    # time: re-evaluated at each assembly at the formulation's time
    f.clear_constraints()
    f.add_constraints(dofs, None, 1e5)
    f.add_constraint("node[7].dep[1] - 0.001*time", 1e5)
    F = {}
    for t in (1.0, 3.0):
        f.set_time(t)
        f.assemble(False, True)
        F[t] = f.get_sollicitation()
    Mf = f.max_diag()
    assert F[3.0][22] - F[1.0][22] == pytest.approx(Mf * 1e5 * 0.002, rel=1e-12)
    assert np.array_equal(np.delete(F[3.0], 22), np.delete(F[1.0], 22))
    for bad in ("node[1].dep[0] + node[2].dep[1]", "node[1].dep[0]*node[1].dep[1]", "node[999].dep[0]", "time"):
        with pytest.raises(L.FemGpuError):
            f.add_constraint(bad, 1e5)
    with pytest.raises(L.FemGpuError):
        f.add_constraint("node[1].dep[0]", 0.0)      # Lagrange form: K indefinite
    f.close()


def test_history_shift_unshift_and_time_steps():
    say("history")
    f, c, *_ = _problem(4)
    n = 3 * len(c)
    f.set_initial_time_step(0.5)
    v = [np.full(n, float(k + 1)) for k in range(4)]
    for k in range(4):           # vectors[0] = v_k, then shift: vectors[1] = v_k
        f.set_result(v[k])
        f.shift()
    assert f.get_time() == pytest.approx(-1.0 + 4 * 0.5)
    for k in range(4):
        assert np.array_equal(f.get_result_num(k + 1), v[3 - k])
    assert np.array_equal(f.get_result_num(0), v[3])        # vectors[0] stays
    assert np.array_equal(f.get_result_num(5), np.zeros(n))  # never filled
    assert np.array_equal(f.get_previous_result(), v[3])
    f.set_next_time_step(0.25)
    f.shift()                    # time += 0.25; time_steps[1] = 0.25
    assert f.get_time() == pytest.approx(1.25)
    f.unshift()                  # time -= time_steps[1]
    assert f.get_time() == pytest.approx(1.0)
    assert np.array_equal(f.get_result_num(0), v[3]) and np.array_equal(f.get_result_num(1), v[3])
    f.unshift(2)                 # vectors[0] takes vectors[1] twice
    assert np.array_equal(f.get_result_num(0), v[2])
    assert np.array_equal(f.get_result_num(1), v[1])
    assert f.get_time() == pytest.approx(0.0)
    with pytest.raises(L.FemGpuError):
        f.get_result_num(20)
    f.close()


@pytest.mark.parametrize("order", [A.NODE_ORDER_NATURAL, A.NODE_ORDER_RCM])
def test_indice_noda_and_nodal_forces(oracle, order):
    say("indice_noda / nodal forces", order)
    lib = A.load()
    lib.femgpu_set_node_order(order)
    try:
        f, c, e, E, faces, tr, dofs = _problem(5)
    finally:
        lib.femgpu_set_node_order(A.NODE_ORDER_NATURAL)
    assert [f.get_indice_noda(i) for i in (0, 7, len(c) - 1)] == [0, 21, 3 * (len(c) - 1)]
    f.add_constraints(dofs, 1e-3 * np.ones(len(dofs)), 1e5)
    f.assemble()
    F_asm = f.get_sollicitation()
    Fn = f.get_nodal_forces()
    _, _, _, oF, _ = oracle.assemble(L.HEXA, c, e, E, NU0, bf=(0.0, 0.2, -0.4), faces=faces, face_trac=tr)
    assert np.abs(Fn - oF).max() <= 1e-14 * np.abs(oF).max()
    assert np.array_equal(f.get_sollicitation(), F_asm)                  # untouched
    assert np.abs(F_asm[dofs] - Fn[dofs]).min() > 1.0                    # the penalty terms M P value
    # set_sollicitation: the right-hand side of the next solve
    f.set_sollicitation(2.0 * F_asm)
    f.set_result(np.zeros(3 * len(c)))
    assert f.solve_system(rel_tol=1e-11)
    u2 = f.get_result()
    f.set_sollicitation(F_asm)
    f.set_result(np.zeros(3 * len(c)))
    assert f.solve_system(rel_tol=1e-11)
    assert np.abs(u2 - 2.0 * f.get_result()).max() <= 1e-8 * np.abs(u2).max()
    f.close()
