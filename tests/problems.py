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
