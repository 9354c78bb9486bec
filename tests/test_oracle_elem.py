"""Pins the oracle's element kernels by closed forms, textbook restatements and invariants.

The reference's generated element kernels cannot be run here (they need its Python-2
generator), so each check is chosen so that a plausible mistake -- a dropped term, a wrong sign
or index, a transposed operand, the wrong Gauss table -- fails it.
"""
import numpy as np
import pytest


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def voigt_D(E, nu):
    lam, mu = lame(E, nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    return D


def voigt_B(grads):
    """Textbook engineering-strain B (6 x 3n), Voigt [xx, yy, zz, yz, xz, xy]."""
    n = len(grads)
    B = np.zeros((6, 3 * n))
    for a, (bx, by, bz) in enumerate(grads):
        B[:, 3 * a:3 * a + 3] = [[bx, 0, 0], [0, by, 0], [0, 0, bz], [0, bz, by], [bz, 0, bx], [by, bx, 0]]
    return B


def rigid_modes(x):
    n = len(x)
    R = np.zeros((3 * n, 6))
    for a in range(n):
        R[3 * a:3 * a + 3, :3] = np.eye(3)
        X, Y, Z = x[a]
        R[3 * a:3 * a + 3, 3] = [-Y, X, 0]
        R[3 * a:3 * a + 3, 4] = [Z, 0, -X]
        R[3 * a:3 * a + 3, 5] = [0, -Z, Y]
    return R


rng = np.random.default_rng(7)
TET = np.array([[0.1, 0.0, 0.05], [1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [0.0, 0.2, 1.1]])


def test_tetra4_equals_volume_BtDB(oracle):
    E, nu = 210000.0, 0.3
    K, _, rc = oracle.elem_matrix(oracle.TETRA, TET.ravel(), E, nu)
    assert rc == 0
    A = np.hstack([np.ones((4, 1)), TET])
    coef = np.linalg.inv(A)  # shape-function coefficients, column a = node a
    grads = coef[1:, :].T
    V = np.linalg.det(A) / 6.0
    Kt = V * voigt_B(grads).T @ voigt_D(E, nu) @ voigt_B(grads)
    assert np.abs(K - Kt).max() <= 1e-12 * np.abs(Kt).max()


def _hexa_textbook(x, E, nu, pts):
    """Independent trilinear hexa: B-matrix quadrature over the given 1-D points (weights 1/2)."""
    K = np.zeros((24, 24))
    sgn = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    for xi in pts:
        for eta in pts:
            for ze in pts:
                p = np.array([xi, eta, ze])
                dN = np.zeros((8, 3))
                for a in range(8):
                    f = np.where(sgn[a] == 1, p, 1 - p)
                    df = np.where(sgn[a] == 1, 1.0, -1.0)
                    dN[a] = [df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]]
                J = dN.T @ x  # J[j, i] = dx_i / dxi_j
                g = np.linalg.solve(J, dN.T).T
                B = voigt_B(g)
                K += 0.125 * np.linalg.det(J) * B.T @ voigt_D(E, nu) @ B
    return K


def test_hexa_uses_twelve_digit_gauss_table(oracle):
    # distorted hexa: the 12-digit abscissae (include/mesh/element_Hexa.h:12-19) and the exact
    # Gauss points give matrices that differ by ~1e-12 relative
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    x = base * [0.7, 1.1, 0.9] + rng.uniform(-0.12, 0.12, size=(8, 3))
    K, _, rc = oracle.elem_matrix(oracle.HEXA, x.ravel(), 1.0, 0.3)
    assert rc == 0
    k12 = _hexa_textbook(x, 1.0, 0.3, [0.211324865405, 0.788675134595])
    g = 1.0 / np.sqrt(3.0)
    kex = _hexa_textbook(x, 1.0, 0.3, [(1 - g) / 2, (1 + g) / 2])
    scale = np.abs(k12).max()
    d12 = np.abs(K - k12).max() / scale
    dex = np.abs(K - kex).max() / scale
    assert d12 < 1e-14
    assert dex > 10 * d12 and dex > 1e-14


def _tet10_grads(x, p):
    L = np.array([1 - p.sum(), p[0], p[1], p[2]])
    dL = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    dN = np.zeros((10, 3))
    for a in range(4):
        dN[a] = (4 * L[a] - 1) * dL[a]
    for k, (i, j) in enumerate(edges):
        dN[4 + k] = 4 * (dL[i] * L[j] + L[i] * dL[j])
    J = dN.T @ x
    return np.linalg.solve(J, dN.T).T, np.linalg.det(J)


def test_tetra10_matches_high_order_quadrature(oracle):
    # straight-sided Tetra_10: B^T D B is quadratic, the 4-point rule is exact, so a 64-point
    # collapsed Gauss product rule must agree to round-off
    x4 = TET
    mids = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    x = np.vstack([x4, [(x4[i] + x4[j]) / 2 for i, j in mids]])
    K, _, rc = oracle.elem_matrix(oracle.TETRA10, x.ravel(), 1.0, 0.25)
    assert rc == 0
    gp, gw = np.polynomial.legendre.leggauss(4)
    gp, gw = (gp + 1) / 2, gw / 2
    Kq = np.zeros((30, 30))
    D = voigt_D(1.0, 0.25)
    for i, u in enumerate(gp):
        for j, v in enumerate(gp):
            for k, w in enumerate(gp):
                # Duffy map of the unit cube onto the reference tetrahedron
                p = np.array([u, v * (1 - u), w * (1 - u) * (1 - v)])
                jac = (1 - u) ** 2 * (1 - v)
                g, detJ = _tet10_grads(x, p)
                B = voigt_B(g)
                Kq += gw[i] * gw[j] * gw[k] * jac * detJ * B.T @ D @ B
    assert np.abs(K - Kq).max() <= 1e-12 * np.abs(Kq).max()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_invariants(oracle, kind):
    if kind == 0:
        x = TET
    elif kind == 1:
        mids = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
        x = np.vstack([TET, [(TET[i] + TET[j]) / 2 for i, j in mids]]) + rng.uniform(-0.02, 0.02, (10, 3))
    else:
        base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
        x = base + rng.uniform(-0.1, 0.1, (8, 3))
    K, F, rc = oracle.elem_matrix(kind, x.ravel(), 3.0, 0.3, bf=(0.5, -1.0, 2.0))
    assert rc == 0
    n = 3 * len(x)
    assert np.array_equal(K, K.T)
    R = rigid_modes(x)
    assert np.abs(K @ R).max() <= 1e-12 * np.abs(K).max() * np.abs(R).max()
    ev = np.linalg.eigvalsh(K)
    assert ev.min() > -1e-10 * ev.max()
    assert np.sum(ev > 1e-10 * ev.max()) == n - 6
    K2, _, _ = oracle.elem_matrix(kind, x.ravel(), 6.0, 0.3)
    assert np.abs(K2 - 2 * K).max() <= 1e-13 * np.abs(K).max()
    # total body load = f * volume (volume from the textbook quadrature of det J)
    _, F1, _ = oracle.elem_matrix(kind, x.ravel(), 1.0, 0.3, bf=(1.0, 0.0, 0.0))
    vol = F1[0::3].sum()
    assert np.allclose(F.reshape(-1, 3).sum(0), np.array([0.5, -1.0, 2.0]) * vol, rtol=1e-13)


def test_tetra10_body_load_closed_form(oracle):
    mids = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    x = np.vstack([TET, [(TET[i] + TET[j]) / 2 for i, j in mids]])
    _, F, _ = oracle.elem_matrix(oracle.TETRA10, x.ravel(), 1.0, 0.3, bf=(0.0, 0.0, 1.0))
    V = abs(np.linalg.det(np.hstack([np.ones((4, 1)), TET]))) / 6
    assert np.allclose(F[2::3][:4], -V / 20, rtol=1e-13) and np.allclose(F[2::3][4:], V / 5, rtol=1e-13)


def test_face_loads_closed_forms(oracle):
    t = np.array([0.3, -1.0, 2.0])
    tri = np.array([[0, 0, 0], [2, 0, 0.5], [0.3, 1.5, 0]])
    A = 0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    F = oracle.face_load(oracle.TRI3, tri.ravel(), t).reshape(-1, 3)
    assert np.allclose(F, np.tile(t * A / 3, (3, 1)), rtol=1e-14)
    tri6 = np.vstack([tri, (tri[0] + tri[1]) / 2, (tri[1] + tri[2]) / 2, (tri[2] + tri[0]) / 2])
    F6 = oracle.face_load(oracle.TRI6, tri6.ravel(), t).reshape(-1, 3)
    assert np.abs(F6[:3]).max() < 1e-15 and np.allclose(F6[3:], np.tile(t * A / 3, (3, 1)), rtol=1e-13)
    quad = np.array([[0, 0, 0], [1.5, 0, 0], [1.5 + 0.2, 0.8, 0.1], [0.2, 0.8, 0.1]])  # parallelogram
    Aq = np.linalg.norm(np.cross(quad[1] - quad[0], quad[3] - quad[0]))
    Fq = oracle.face_load(oracle.QUAD4, quad.ravel(), t).reshape(-1, 3)
    assert np.allclose(Fq, np.tile(t * Aq / 4, (4, 1)), rtol=1e-14)


def test_patch_test_on_assembled_mesh(oracle):
    """Affine displacement fields produce zero nodal forces at interior nodes."""
    for kind, nb in [(0, (4, 4, 4)), (2, (4, 4, 4)), (1, (3, 3, 3))]:
        c, e = oracle.make_rect(kind, nb)
        interior = np.all((c > 1e-9) & (c < 1 - 1e-9), axis=1)
        c = c.copy()
        if kind != 1:
            c[interior] += rng.uniform(-0.05, 0.05, (interior.sum(), 3))
        rp, col, vals, F, rc = oracle.assemble(kind, c, e, 7.0, 0.27)
        assert rc == 0
        A = rng.normal(size=(3, 3))
        u = (c @ A.T + [0.1, -0.2, 0.3]).ravel()
        f = oracle.spmv(rp, col, vals, u).reshape(-1, 3)
        scale = np.abs(vals).max() * np.abs(u).max()
        assert np.abs(f[interior]).max() <= 1e-12 * scale
