"""ctypes binding of liboracle.so, the plain-C restatement of the LMT++ path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the lmtpp_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

TETRA, TETRA10, HEXA = 0, 1, 2
TRI3, TRI6, QUAD4 = 0, 1, 2
NPE = {TETRA: 4, TETRA10: 10, HEXA: 8}
NPF = {TRI3: 3, TRI6: 6, QUAD4: 4}
FACE_OF = {TETRA: TRI3, TETRA10: TRI6, HEXA: QUAD4}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    """Host threads of assemble, spmv and pcg_jacobi (n <= 0: all); 1 is the serial restatement."""
    return int(lib().or_set_threads(int(n)))


def make_rect(kind, nb_points, X0=(0.0, 0.0, 0.0), X1=(1.0, 1.0, 1.0)):
    nb = (C.c_uint * 3)(*nb_points)
    nn, ne = C.c_uint32(), C.c_uint32()
    if lib().or_make_rect_counts(kind, nb, C.byref(nn), C.byref(ne)):
        raise ValueError("bad nb_points")
    coords = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, NPE[kind]), dtype=np.uint32)
    x0 = (C.c_double * 3)(*X0)
    x1 = (C.c_double * 3)(*X1)
    rc = lib().or_make_rect(kind, x0, x1, nb, _p(coords), _p(conn))
    if rc:
        raise RuntimeError("or_make_rect failed %d" % rc)
    return coords, conn


def make_rect_tetra10_linear_search(nb_points, X0=(0.0, 0.0, 0.0), X1=(1.0, 1.0, 1.0)):
    nb = (C.c_uint * 3)(*nb_points)
    nn, ne = C.c_uint32(), C.c_uint32()
    lib().or_make_rect_counts(TETRA10, nb, C.byref(nn), C.byref(ne))
    coords = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, 10), dtype=np.uint32)
    got = C.c_uint32()
    lib().or_make_rect_tetra10_linear_search((C.c_double * 3)(*X0), (C.c_double * 3)(*X1), nb, _p(coords),
                                             _p(conn), C.byref(got))
    return coords[: got.value], conn


def skin_faces(kind, n_nodes, conn):
    conn = _u32(conn)
    n = C.c_uint32()
    lib().or_skin_faces(kind, n_nodes, conn.shape[0], _p(conn), None, 0, C.byref(n))
    faces = np.zeros((n.value, NPF[FACE_OF[kind]]), dtype=np.uint32)
    lib().or_skin_faces(kind, n_nodes, conn.shape[0], _p(conn), _p(faces), n.value, C.byref(n))
    return faces


def elem_matrix(kind, xe, E, nu, bf=(0.0, 0.0, 0.0)):
    xe = _f64(xe)
    nd = 3 * NPE[kind]
    Ke = np.zeros((nd, nd))
    Fe = np.zeros(nd)
    dmin = C.c_double()
    rc = lib().or_elem_matrix(kind, _p(xe), C.c_double(E), C.c_double(nu), (C.c_double * 3)(*bf), _p(Ke), _p(Fe),
                              C.byref(dmin))
    return Ke, Fe, rc


def face_load(fkind, xf, t):
    xf = _f64(xf)
    Ff = np.zeros(3 * NPF[fkind])
    lib().or_face_load(fkind, _p(xf), (C.c_double * 3)(*t), _p(Ff))
    return Ff


def pattern(kind, n_nodes, conn):
    conn = _u32(conn)
    nnz = C.c_size_t()
    lib().or_pattern(kind, n_nodes, conn.shape[0], _p(conn), None, None, C.byref(nnz))
    row_ptr = np.zeros(n_nodes + 1, dtype=np.uint32)
    col = np.zeros(nnz.value, dtype=np.uint32)
    lib().or_pattern(kind, n_nodes, conn.shape[0], _p(conn), _p(row_ptr), _p(col), C.byref(nnz))
    return row_ptr, col


def assemble(kind, coords, conn, E_e, nu, bf=(0.0, 0.0, 0.0), faces=None, face_trac=None, pat=None):
    coords, conn = _f64(coords), _u32(conn)
    n_nodes = coords.shape[0]
    E_e = _f64(np.broadcast_to(np.asarray(E_e, dtype=np.float64), (conn.shape[0],)))
    row_ptr, col = pat if pat is not None else pattern(kind, n_nodes, conn)
    vals = np.zeros((len(col), 3, 3))
    F = np.zeros(3 * n_nodes)
    nf = 0
    fp = ft = None
    if faces is not None and len(faces):
        faces = _u32(faces)
        face_trac = _f64(np.broadcast_to(np.asarray(face_trac, dtype=np.float64), (faces.shape[0], 3)))
        nf, fp, ft = faces.shape[0], _p(faces), _p(face_trac)
    rc = lib().or_assemble(kind, n_nodes, _p(coords), conn.shape[0], _p(conn), _p(E_e), C.c_double(nu),
                           (C.c_double * 3)(*bf), nf, fp, ft, _p(row_ptr), _p(col), _p(vals), _p(F))
    return row_ptr, col, vals, F, rc


def penalty(row_ptr, col, vals, F, cons_dof, cons_val=None, P=1e5):
    cons_dof = _u32(cons_dof)
    cv = None if cons_val is None else _f64(cons_val)
    M = C.c_double()
    lib().or_penalty(len(row_ptr) - 1, _p(row_ptr), _p(col), _p(vals), _p(F), len(cons_dof), _p(cons_dof),
                     None if cv is None else _p(cv), C.c_double(P), C.byref(M))
    return M.value


def spmv(row_ptr, col, vals, x):
    x = _f64(x)
    y = np.zeros_like(x)
    lib().or_spmv(len(row_ptr) - 1, _p(row_ptr), _p(col), _p(vals), _p(x), _p(y))
    return y


def pcg_jacobi(row_ptr, col, vals, b, x0=None, rel_tol=1e-10, max_iter=100000, block=False):
    b = _f64(b)
    x = np.zeros_like(b) if x0 is None else _f64(x0).copy()
    it, rr = C.c_int(), C.c_double()
    rc = lib().or_pcg_jacobi(len(row_ptr) - 1, _p(row_ptr), _p(col), _p(vals), _p(b), _p(x), C.c_double(rel_tol),
                             max_iter, int(block), C.byref(it), C.byref(rr))
    return x, it.value, rr.value, rc


def pcg_jacobi_hist(row_ptr, col, vals, b, x0=None, rel_tol=1e-10, max_iter=100000, block=False):
    """pcg_jacobi, also returning per iteration (relative recurrence residual, seconds since the
    call began): an array of shape (iters, 2)."""
    b = _f64(b)
    x = np.zeros_like(b) if x0 is None else _f64(x0).copy()
    it, rr = C.c_int(), C.c_double()
    hist = np.zeros((max(1, min(max_iter, 1 << 20)), 2))
    rc = lib().or_pcg_jacobi_hist(len(row_ptr) - 1, _p(row_ptr), _p(col), _p(vals), _p(b), _p(x), C.c_double(rel_tol),
                                  max_iter, int(block), C.byref(it), C.byref(rr), _p(hist), hist.shape[0])
    return x, it.value, rr.value, rc, hist[:min(it.value, hist.shape[0])].copy()


def cg_csr(row_ptr, col, vals, b, rel_tol=1e-14, max_iter=1000):
    row_ptr, col, vals, b = _u32(row_ptr), _u32(col), _f64(vals), _f64(b)
    x = np.zeros_like(b)
    it = C.c_int()
    rc = lib().or_cg_csr(len(b), _p(row_ptr), _p(col), _p(vals), _p(b), _p(x), C.c_double(rel_tol), max_iter,
                         C.byref(it))
    return x, it.value, rc


def block_ic0(row_ptr, col, vals):
    L = np.zeros_like(vals)
    rc = lib().or_block_ic0(len(row_ptr) - 1, _p(row_ptr), _p(col), _p(vals), _p(L))
    return L, rc


def pcg_ic0(row_ptr, col, vals, L, b, x0=None, crit=1e-10, max_iter=100000):
    b = _f64(b)
    x = np.zeros_like(b) if x0 is None else _f64(x0).copy()
    it = C.c_int()
    rc = lib().or_pcg_ic0(len(row_ptr) - 1, _p(row_ptr), _p(col), _p(vals), _p(L), _p(b), _p(x), C.c_double(crit),
                          max_iter, C.byref(it))
    return x, it.value, rc


def elem_grad_E(kind, conn, coords, E_e, nu, u, lam):
    conn, coords = _u32(conn), _f64(coords)
    E_e = _f64(np.broadcast_to(np.asarray(E_e, dtype=np.float64), (conn.shape[0],)))
    g = np.zeros(conn.shape[0])
    lib().or_elem_grad_E(kind, conn.shape[0], _p(conn), _p(coords), _p(E_e), C.c_double(nu), _p(_f64(u)),
                         _p(_f64(lam)), _p(g))
    return g


def after_solve(kind, conn, coords, E_e, nu, u):
    """(eps[n,6], sigma[n,6], tr[n]) at xi = (1/2,1/2,1/2) per element."""
    conn, coords = _u32(conn), _f64(coords)
    n = conn.shape[0]
    E_e = _f64(np.broadcast_to(np.asarray(E_e, dtype=np.float64), (n,)))
    eps, sig, tr = np.zeros((n, 6)), np.zeros((n, 6)), np.zeros(n)
    rc = lib().or_after_solve(kind, n, _p(conn), _p(coords), _p(E_e), C.c_double(nu), _p(_f64(u)), _p(eps), _p(sig),
                              _p(tr))
    assert rc == 0, rc
    return eps, sig, tr


# --- reference builds (oracle/_ref), available only where the LMT++ sources are -----------------

def ref_available():
    return all(os.path.exists(os.path.join(REF_DIR, b))
               for b in ("ref_mesh_dump", "ref_tinyblock_pcg", "ref_after_solve", "ref_write_vtu", "ref_symrcm"))


def ref_mesh(kind_name, nb_points, tmpdir):
    out = os.path.join(tmpdir, "mesh_%s_%d_%d_%d.bin" % ((kind_name,) + tuple(nb_points)))
    subprocess.check_call([os.path.join(REF_DIR, "ref_mesh_dump"), kind_name] + [str(v) for v in nb_points] + [out],
                          stdout=subprocess.DEVNULL)
    with open(out, "rb") as f:
        nn, ne, npe = np.frombuffer(f.read(12), dtype=np.uint32)
        coords = np.frombuffer(f.read(8 * 3 * int(nn)), dtype=np.float64).reshape(-1, 3).copy()
        conn = np.frombuffer(f.read(4 * int(ne) * int(npe)), dtype=np.uint32).reshape(-1, int(npe)).copy()
    return coords, conn


def ref_after_solve_hexa(nb_points, coords, u, E, nu, tmpdir):
    """The reference's generated Hexa after_solve on make_rect(Hexa, nb_points) with the node
    positions replaced by coords: (eps[n,6], sigma[n,6], tr[n]) in element order."""
    inp = os.path.join(tmpdir, "as_in.bin")
    out = os.path.join(tmpdir, "as_out.bin")
    coords, u = _f64(coords), _f64(u)
    with open(inp, "wb") as f:
        f.write(np.array([coords.shape[0]], dtype=np.uint32).tobytes())
        f.write(coords.tobytes())
        f.write(u.tobytes())
        f.write(np.array([E, nu], dtype=np.float64).tobytes())
    subprocess.check_call([os.path.join(REF_DIR, "ref_after_solve")] + [str(v) for v in nb_points] + [inp, out],
                          stdout=subprocess.DEVNULL)
    with open(out, "rb") as f:
        ne = int(np.frombuffer(f.read(4), dtype=np.uint32)[0])
        d = np.frombuffer(f.read(8 * 13 * ne), dtype=np.float64).reshape(ne, 13)
    return d[:, :6].copy(), d[:, 6:12].copy(), d[:, 12].copy()


def ref_write_vtu_hexa(nb_points, coords, u, E, nu, tmpdir):
    """The reference's write_mesh_vtk (binary, appended) for make_rect(Hexa, nb_points) of the
    elasticity_dep_static mesh with node positions coords, displacements u, E, nu and the strain
    and stress of its generated after_solve: (file bytes, eps[n,6], sigma[n,6], tr[n])."""
    inp = os.path.join(tmpdir, "vtu_in.bin")
    out = os.path.join(tmpdir, "vtu_out.vtu")
    flds = os.path.join(tmpdir, "vtu_fields.bin")
    coords, u = _f64(coords), _f64(u)
    with open(inp, "wb") as f:
        f.write(np.array([coords.shape[0]], dtype=np.uint32).tobytes())
        f.write(coords.tobytes())
        f.write(u.tobytes())
        f.write(np.array([E, nu], dtype=np.float64).tobytes())
    subprocess.check_call([os.path.join(REF_DIR, "ref_write_vtu")] + [str(v) for v in nb_points] + [inp, out, flds],
                          stdout=subprocess.DEVNULL)
    with open(out, "rb") as f:
        data = f.read()
    with open(flds, "rb") as f:
        ne = int(np.frombuffer(f.read(4), dtype=np.uint32)[0])
        d = np.frombuffer(f.read(8 * 13 * ne), dtype=np.float64).reshape(ne, 13)
    return data, d[:, :6].copy(), d[:, 6:12].copy(), d[:, 12].copy()


def ref_write_avs(kind_name, nb_points, coords, u, E, nu, display, tmpdir):
    """The reference's ASCII write_avs for make_rect(kind, nb_points) with node positions coords:
    the elasticity_dep_static Hexa mesh with displacements u, E, nu and the strain and stress of
    its generated after_solve, or (tetra, tetra10) the field-less standard mesh.  Returns (file
    bytes, eps[n,6], sigma[n,6], tr[n]); the last three are None for the field-less meshes."""
    inp = os.path.join(tmpdir, "avs_in.bin")
    out = os.path.join(tmpdir, "avs_out.inp")
    flds = os.path.join(tmpdir, "avs_fields.bin")
    if os.path.exists(flds):
        os.remove(flds)
    coords, u = _f64(coords), _f64(u)
    with open(inp, "wb") as f:
        f.write(np.array([coords.shape[0]], dtype=np.uint32).tobytes())
        f.write(coords.tobytes())
        f.write(u.tobytes())
        f.write(np.array([E, nu], dtype=np.float64).tobytes())
    subprocess.check_call([os.path.join(REF_DIR, "ref_write_avs"), kind_name] + [str(v) for v in nb_points] +
                          [inp, out, flds] + list(display), stdout=subprocess.DEVNULL)
    with open(out, "rb") as f:
        data = f.read()
    if not os.path.exists(flds):
        return data, None, None, None
    with open(flds, "rb") as f:
        ne = int(np.frombuffer(f.read(4), dtype=np.uint32)[0])
        d = np.frombuffer(f.read(8 * 13 * ne), dtype=np.float64).reshape(ne, 13)
    return data, d[:, :6].copy(), d[:, 6:12].copy(), d[:, 12].copy()


def ref_symrcm(kind_name, nb_points, tmpdir):
    """The reference's symrcm(m) order on make_rect(kind, nb_points)."""
    out = os.path.join(tmpdir, "symrcm.bin")
    subprocess.check_call([os.path.join(REF_DIR, "ref_symrcm"), kind_name] + [str(v) for v in nb_points] + [out],
                          stdout=subprocess.DEVNULL)
    d = np.fromfile(out, dtype=np.uint32)
    return d[2:2 + d[1]].copy()


def ref_symamd(kind_name, nb_points, tmpdir):
    """The reference's symamd(m) on make_rect(kind, nb_points): the position of each node."""
    out = os.path.join(tmpdir, "symamd.bin")
    subprocess.check_call([os.path.join(REF_DIR, "ref_symamd"), kind_name] + [str(v) for v in nb_points] + [out],
                          stdout=subprocess.DEVNULL)
    d = np.fromfile(out, dtype=np.uint32)
    return d[2:2 + d[1]].copy()


def ref_symamd_graph(row_ptr, col, tmpdir):
    """The reference's symamd(ind) on a node graph (rows sorted, each holding its node): the
    position of each node."""
    inp = os.path.join(tmpdir, "symamd_graph.bin")
    out = os.path.join(tmpdir, "symamd.bin")
    with open(inp, "wb") as f:
        f.write(np.array([len(row_ptr) - 1], dtype=np.uint32).tobytes())
        f.write(_u32(row_ptr).tobytes())
        f.write(_u32(col).tobytes())
    subprocess.check_call([os.path.join(REF_DIR, "ref_symamd"), "graph", inp, out], stdout=subprocess.DEVNULL)
    d = np.fromfile(out, dtype=np.uint32)
    return d[2:2 + d[1]].copy()


def ref_tinyblock_pcg(row_ptr, col, vals, b, x0, crit, tmpdir):
    inp = os.path.join(tmpdir, "pcg_in.bin")
    out = os.path.join(tmpdir, "pcg_out.bin")
    n_nodes = len(row_ptr) - 1
    with open(inp, "wb") as f:
        f.write(np.array([n_nodes, len(col)], dtype=np.uint32).tobytes())
        f.write(_u32(row_ptr).tobytes())
        f.write(_u32(col).tobytes())
        f.write(_f64(vals).tobytes())
        f.write(_f64(b).tobytes())
        f.write(_f64(x0).tobytes())
        f.write(np.array([crit], dtype=np.float64).tobytes())
    subprocess.check_call([os.path.join(REF_DIR, "ref_tinyblock_pcg"), inp, out], stdout=subprocess.DEVNULL)
    with open(out, "rb") as f:
        iters = int(np.frombuffer(f.read(4), dtype=np.uint32)[0])
        x = np.frombuffer(f.read(8 * 3 * n_nodes), dtype=np.float64).copy()
        raw = f.read()
    rec = np.frombuffer(raw, dtype=np.dtype([("c", "<u4"), ("v", "<f8", (3, 3))]))
    # factor blocks arrive per row as [strictly-lower blocks..., diagonal]; place them at
    # full-BSR slots.  Blocks absent from the reference's pattern (all-zero blocks are
    # deleted by its sparse rows) are reported through the returned mask.
    L = np.zeros((len(col), 3, 3))
    present = np.zeros(len(col), dtype=bool)
    k = 0
    for r in range(n_nodes):
        cols = col[row_ptr[r]:row_ptr[r + 1]]
        while True:
            c = int(rec["c"][k])
            p = row_ptr[r] + int(np.searchsorted(cols, c))
            L[p] = rec["v"][k]
            present[p] = True
            k += 1
            if c == r:
                break
    return x, iters, L, present
