"""Pins of three oracle functions by what the mathematics fixes, independently of their formulas.

* or_penalty (formulation.h:1028-1070): against a dense, element-by-element K built from
  or_elem_matrix, the penalty adds exactly P * max|diag K| (max taken on that unconstrained K) to
  each constrained diagonal entry and P * max|diag K| * value to F, nothing else; the solution
  then holds the prescribed displacements to O(1/P) -- the defect shrinks 100x when P grows 100x.
* or_skin_faces (element_{Tetra,Tetra_10,Hexa}.py children, element_Triangle_6.py): on the unit
  cube every boundary face appears once, oriented outward by its first three nodes, and for
  Tetra_10 the face's nodes 3, 4, 5 sit at the midpoints of its edges (0,1), (1,2), (2,0); a
  uniform traction t on the x = 1 face gives exactly zero load at its vertex nodes and t A/3 per
  adjacent face at its mid-edge nodes (A = h^2/2).
* or_elem_grad_E: g_e = -lambda_e^T K_e(E_e) u_e / E_e with or_elem_matrix (K_e linear in E), and
  central differences (step 1e-4 E_e) of J = 1/2 ||u[obs] - u_meas||^2 with dense LAPACK solves,
  the penalty scale held fixed as the adjoint holds it.
"""
import numpy as np
import pytest

from problems import lognormal_E


def dense_from_elements(O, kind, c, e, E, nu):
    n = 3 * len(c)
    K = np.zeros((n, n))
    for k, nodes in enumerate(e):
        Ke, _, rc = O.elem_matrix(kind, c[nodes], E[k], nu)
        assert rc == 0
        d = (3 * nodes.astype(np.int64)[:, None] + np.arange(3)).ravel()
        K[np.ix_(d, d)] += Ke
    return K


def dense_from_bsr(rp, col, vals):
    n = 3 * (len(rp) - 1)
    A = np.zeros((n, n))
    for r in range(len(rp) - 1):
        for p in range(rp[r], rp[r + 1]):
            A[3 * r:3 * r + 3, 3 * col[p]:3 * col[p] + 3] = vals[p]
    return A


def clamp_dofs(c):
    nodes = np.where(np.abs(c[:, 0]) < 1e-9)[0]
    return (3 * nodes[:, None] + np.arange(3)).ravel()


@pytest.mark.parametrize("kind", [0, 2])
def test_penalty_against_dense_brute_force(oracle, kind):
    O = oracle
    c, e = O.make_rect(kind, (4, 3, 3))
    E = lognormal_E(len(e), 0.4, 3)
    nu = 0.3
    K0 = dense_from_elements(O, kind, c, e, E, nu)
    M = np.abs(np.diag(K0)).max()
    dofs = clamp_dofs(c)
    vals = 1e-3 * np.sin(1.0 + np.arange(len(dofs)))
    rp, col, bv, F, rc = O.assemble(kind, c, e, E, nu)
    assert rc == 0 and np.abs(F).max() == 0.0
    K_before = dense_from_bsr(rp, col, bv)
    assert np.abs(K_before - K0).max() <= 1e-12 * np.abs(K0).max()
    for P in (1e5, 1e7):
        bv2, F2 = bv.copy(), F.copy()
        Mo = O.penalty(rp, col, bv2, F2, dofs, vals, P)
        assert Mo == np.abs(np.diag(K_before)).max() and abs(Mo - M) <= 1e-14 * M
        M = Mo
        dK = dense_from_bsr(rp, col, bv2) - K_before
        expect = np.zeros_like(dK)
        expect[dofs, dofs] = P * M
        assert np.abs(dK - expect).max() <= 1e-15 * P * M
        fexp = np.zeros_like(F2)
        fexp[dofs] = P * M * vals
        assert np.abs(F2 - fexp).max() <= 1e-15 * P * M * np.abs(vals).max()
        u = np.linalg.solve(dense_from_bsr(rp, col, bv2), F2)
        defect = np.abs(u[dofs] - vals).max()
        if P == 1e5:
            d5 = defect
            assert defect <= 1e-3 * np.abs(vals).max()
        else:
            assert defect <= 0.012 * d5            # O(1/P): 100x larger P, 100x smaller defect
    # F-only: with no constraint values F stays zero, K only gets the diagonal terms
    bv3, F3 = bv.copy(), F.copy()
    O.penalty(rp, col, bv3, F3, dofs, None, 1e5)
    assert np.abs(F3).max() == 0.0


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_skin_faces_orientation_and_midpoints(oracle, kind):
    O = oracle
    c, e = O.make_rect(kind, (4, 3, 5))
    faces = O.skin_faces(kind, len(c), e)
    npf = faces.shape[1]
    x = c[faces]
    nrm = np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0])
    cen = x[:, :3].mean(axis=1)
    # outward: the normal points along the face's position relative to the cube centre, on the
    # axis where the face lies on the boundary
    on = np.isclose(cen, 0.0, atol=1e-9) | np.isclose(cen, 1.0, atol=1e-9)
    assert np.all(on.sum(axis=1) == 1)
    ax = np.argmax(on, axis=1)
    side = np.sign(cen[np.arange(len(cen)), ax] - 0.5)
    assert np.all(np.sign(nrm[np.arange(len(nrm)), ax]) == side)
    assert np.all(np.abs(nrm[np.arange(len(nrm)), (ax + 1) % 3]) < 1e-12)
    # each boundary face once: total area 6
    area = 0.5 * np.linalg.norm(nrm, axis=1) * (2.0 if npf == 4 else 1.0)
    assert np.isclose(area.sum(), 6.0, rtol=1e-12)
    if kind == 1:
        for m, (i, j) in zip((3, 4, 5), ((0, 1), (1, 2), (2, 0))):
            assert np.abs(x[:, m] - 0.5 * (x[:, i] + x[:, j])).max() < 1e-14


def test_tetra10_face_loads_on_vertex_and_mid_edge_nodes(oracle):
    O = oracle
    n = 4
    c, e = O.make_rect(1, (n + 1,) * 3)
    faces = O.skin_faces(1, len(c), e)
    x1 = faces[np.all(np.isclose(c[faces][:, :, 0], 1.0), axis=1)]
    t = np.array([0.3, -0.7, 1.1])
    _, _, _, F, rc = O.assemble(1, c, e, 210000.0, 0.3, faces=x1, face_trac=np.tile(t, (len(x1), 1)))
    assert rc == 0
    F = F.reshape(-1, 3)
    on = np.where(np.isclose(c[:, 0], 1.0))[0]
    # vertex nodes sit on the coarse lattice (multiples of 1/n), mid-edge nodes do not
    coarse = np.all(np.isclose(np.round(c[on] * n), c[on] * n, atol=1e-9), axis=1)
    assert np.abs(F[on[coarse]]).max() == 0.0
    A = 0.5 / n ** 2
    mid = on[~coarse]
    yz = c[mid][:, 1:]
    border = np.any(np.isclose(yz, 0.0, atol=1e-9) | np.isclose(yz, 1.0, atol=1e-9), axis=1)
    expect = np.where(border, 1.0, 2.0)[:, None] * (A / 3) * t
    assert np.abs(F[mid] - expect).max() <= 1e-15
    off = np.setdiff1d(np.arange(len(c)), on)
    assert np.abs(F[off]).max() == 0.0


def test_elem_grad_against_element_matrices_and_central_differences(oracle):
    O = oracle
    kind, nu, P = 2, 0.3, 1e5
    c, e = O.make_rect(kind, (4, 4, 4))
    E = lognormal_E(len(e), 0.3, 2)
    dofs = clamp_dofs(c)
    faces = O.skin_faces(kind, len(c), e)
    fx = faces[np.all(np.isclose(c[faces][:, :, 0], 1.0), axis=1)]
    tr = np.tile([0.0, 0.2, -1.0], (len(fx), 1))
    rp, col, bv, F, _ = O.assemble(kind, c, e, E, nu, faces=fx, face_trac=tr)
    K0 = dense_from_bsr(rp, col, bv)
    M = O.penalty(rp, col, bv, F, dofs, None, P)
    pen = np.zeros(len(F))
    pen[dofs] = P * M

    def K_of(Ev):   # unconstrained K of Ev plus the base penalty (scale held fixed)
        return dense_from_elements(O, kind, c, e, Ev, nu) + np.diag(pen)

    K = dense_from_bsr(rp, col, bv)
    assert np.abs(K - K_of(E)).max() <= 1e-12 * np.abs(K).max()
    u = np.linalg.solve(K, F)
    top = np.where(np.isclose(c[:, 2], 1.0))[0][::3]
    od = (3 * top[:, None] + np.arange(3)).ravel()
    u_meas = u[od] * (1.0 + 0.1 * np.cos(np.arange(len(od))))
    res = u[od] - u_meas
    dJdu = np.zeros_like(u)
    dJdu[od] = res
    lam = np.linalg.solve(K, dJdu)
    g = O.elem_grad_E(kind, e, c, E, nu, u, lam)
    # linearity in E: dK_e/dE_e = K_e(E_e) / E_e
    for k in range(len(e)):
        Ke, _, _ = O.elem_matrix(kind, c[e[k]], E[k], nu)
        d = (3 * e[k].astype(np.int64)[:, None] + np.arange(3)).ravel()
        assert abs(g[k] + lam[d] @ Ke @ u[d] / E[k]) <= 1e-12 * np.abs(g).max()
    # central differences: d+- = K(E +- h)^-1 (-+h K_e(1) u) exactly (dense), J+- = 1/2|res + d[od]|^2
    gmax = np.abs(g).max()
    assert np.abs(K0).max() > 0
    for k in np.random.default_rng(4).choice(len(e), size=10, replace=False):
        h = 1e-4 * E[k]
        Ke1, _, _ = O.elem_matrix(kind, c[e[k]], 1.0, nu)
        d = (3 * e[k].astype(np.int64)[:, None] + np.arange(3)).ravel()
        Js = []
        for s in (1.0, -1.0):
            E2 = E.copy()
            E2[k] += s * h
            rhs = np.zeros_like(u)
            rhs[d] = -s * h * (Ke1 @ u[d])
            dd = np.linalg.solve(K_of(E2), rhs)[od]
            Js.append(res @ dd + 0.5 * dd @ dd)
        fd = (Js[0] - Js[1]) / (2 * h)
        if abs(g[k]) >= 1e-3 * gmax:
            assert abs(fd - g[k]) <= 1e-6 * abs(g[k]), (k, fd, g[k])
        else:
            assert abs(fd - g[k]) <= 1e-9 * gmax, (k, fd, g[k])
