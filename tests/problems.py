"""Synthetic inputs of the benchmark configurations, shared by tests and bench.py.

Inputs only: meshes come from the library's make_rect, random fields from numpy seeds.
"""
import numpy as np

import lmtpp_amd as L

E0, NU0, P0 = 210000.0, 0.3, 1e5


def jitter_interior(coords, h, amp, seed):
    """Move nodes strictly inside the unit cube by U(-amp*h, amp*h) per coordinate."""
    rng = np.random.default_rng(seed)
    c = coords.copy()
    inside = np.all((c > 1e-9) & (c < 1 - 1e-9), axis=1)
    c[inside] += rng.uniform(-amp * h, amp * h, size=(int(inside.sum()), 3))
    return c


def tetra10_jittered(cells, seed=0, amp=0.2):
    """Tetra_10 mesh on cells^3 cubes; interior corner nodes jittered by +-amp*h (h = half cell),
    mid-edge nodes re-placed at the exact midpoints of their (jittered) end nodes; boundary nodes
    are not moved."""
    c, e = L.make_rect(L.TETRA10, (cells + 1,) * 3)
    h = 0.5 / cells
    rng = np.random.default_rng(seed)
    corner = np.zeros(len(c), bool)
    corner[np.unique(e[:, :4])] = True
    inside = np.all((c > 1e-9) & (c < 1 - 1e-9), axis=1) & corner
    c = c.copy()
    c[inside] += rng.uniform(-amp * h, amp * h, size=(int(inside.sum()), 3))
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    for k, (i, j) in enumerate(edges):
        c[e[:, 4 + k]] = 0.5 * (c[e[:, i]] + c[e[:, j]])
    return c, e


def lognormal_E(n_elem, sigma, seed, E=E0):
    rng = np.random.default_rng(seed)
    return E * np.exp(sigma * rng.standard_normal(n_elem))


def clamp_and_traction(kind, coords, conn, trac=(0.0, 0.0, -1.0)):
    faces = L.skin_faces(kind, len(coords), conn)
    fx = L.faces_on_plane(coords, faces, 0, 1.0)
    dofs = L.clamp_dofs(L.nodes_on_plane(coords, 0, 0.0))
    return fx, np.broadcast_to(np.asarray(trac, float), (len(fx), 3)).copy(), dofs


def build(kind, coords, conn, E, nu=NU0, bf=(0.0, 0.0, 0.0), trac=(0.0, 0.0, -1.0), penalty=P0):
    """A Formulation with the clamp on x=0 and a uniform traction on x=1."""
    f = L.Formulation(kind, coords, conn)
    f.set_material(E, nu, bf)
    faces, tr, dofs = clamp_and_traction(kind, coords, conn, trac)
    f.set_traction(faces, tr)
    f.add_constraints(dofs, None, penalty)
    return f, faces, tr, dofs


def config0():
    c, e = L.make_rect(L.TETRA, (11, 11, 11))
    return L.TETRA, c, e, E0


def config1(n=100, seed=0):
    c, e = L.make_rect(L.HEXA, (n + 1,) * 3)
    return L.HEXA, c, e, lognormal_E(len(e), 0.5, seed)


def config2(cells=40, seed=0):
    """Tetra_10 jittered mesh with body force (0, 0, -9.81 * 7.8e-9) beside the traction."""
    c, e = tetra10_jittered(cells, seed=seed)
    return L.TETRA10, c, e, E0, (0.0, 0.0, -9.81 * 7.8e-9)


def box_filter27(z, n, passes=3):
    """3 passes of the 27-point box filter on an n^3 cell field, zero-flux (replicated) edges."""
    a = z.reshape(n, n, n)
    for _ in range(passes):
        p = np.pad(a, 1, mode="edge")
        acc = np.zeros_like(a)
        for dz in range(3):
            for dy in range(3):
                for dx in range(3):
                    acc += p[dz:dz + n, dy:dy + n, dx:dx + n]
        a = acc / 27.0
    return a.reshape(-1)


def config3(n=64, seed=3):
    """Identification problem: Hexa8 n^3 lattice, E_true = 210000 exp(0.3 z~) with z~ the box-filtered
    N(0,1) field rescaled to unit std, and 5% of the top-face nodes observed (all 3 components).
    Returns (c, e, E_true, odofs, rng) -- rng continues the seeded stream."""
    c, e = L.make_rect(L.HEXA, (n + 1,) * 3)
    rng = np.random.default_rng(seed)
    zt = box_filter27(rng.standard_normal(len(e)), n)
    zt = zt / zt.std()
    E_true = E0 * np.exp(0.3 * zt)
    top = np.where(np.abs(c[:, 2] - 1.0) < 1e-9)[0]
    obs = rng.choice(top, size=int(0.05 * len(top)), replace=False)
    odofs = (3 * obs[:, None] + np.arange(3)).ravel()
    return c, e, E_true, odofs, rng


def fd_gradient(f, O, E, u, res, odofs, el, nu=NU0, rel_tol=1e-11):
    """Central difference (step h = 1e-4 E_el) of J = 1/2 ||u[odofs] - u_meas||^2 at the solution u
    of K(E) u = F, res = u[odofs] - u_meas.  Each side solves K(E +- h e_el) d = -+h K_el(1) u_el for
    the change d of the solution (K_el(1): the oracle's element matrix at unit modulus), so the
    differences are computed without the cancellation of K u against F; J+- = 1/2 ||res + d[odofs]||^2.
    None when the step moves max|diag K| (the penalty scale, which the adjoint holds fixed)."""
    F0 = f.get_sollicitation()
    M = f.max_diag()
    h = 1e-4 * E[el]
    nodes = f.conn[el]
    dofs = (3 * nodes.astype(np.int64)[:, None] + np.arange(3)).ravel()
    Ke1, _, _ = O.elem_matrix(f.kind, f.coords[nodes], 1.0, nu)
    load = Ke1 @ u[dofs]
    d = []
    for s in (+1.0, -1.0):
        E2 = E.copy()
        E2[el] += s * h
        f.set_material(E2, nu)
        f.assemble(True, False)
        if f.max_diag() != M:
            d = None
            break
        rhs = np.zeros(3 * f.n_nodes)
        rhs[dofs] = -s * h * load
        f.set_sollicitation(rhs)
        f.set_result(np.zeros(3 * f.n_nodes))
        assert f.solve_system(rel_tol=rel_tol), f.last_stats.status
        d.append(f.get_result()[odofs])
    f.set_material(E, nu)
    f.assemble()
    f.set_sollicitation(F0)
    f.set_result(u)
    if d is None:
        return None
    return (res @ (d[0] - d[1]) + 0.5 * (d[0] @ d[0] - d[1] @ d[1])) / (2 * h)


def permute_bsr(row_ptr, col, vals, order):
    """Full 3x3-block CSR renumbered by an elimination order (order[t] = node): row t holds the
    blocks of row order[t], columns renumbered and sorted."""
    n = len(row_ptr) - 1
    order = np.asarray(order, dtype=np.int64)
    tof = np.empty(n, dtype=np.int64)
    tof[order] = np.arange(n)
    rp = np.zeros(n + 1, dtype=np.uint32)
    cols, vs = [], []
    for t in range(n):
        r = order[t]
        sl = slice(int(row_ptr[r]), int(row_ptr[r + 1]))
        c = tof[col[sl]]
        k = np.argsort(c, kind="stable")
        cols.append(c[k])
        vs.append(vals[sl][k])
        rp[t + 1] = rp[t] + len(k)
    return rp, np.concatenate(cols).astype(np.uint32), np.ascontiguousarray(np.concatenate(vs))
