"""Mesh utilities of the C ABI: make_rect lattices and skin faces (host code in libfemgpu.so)."""
import ctypes as C

import numpy as np

from . import _capi as A


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def make_rect(kind, nb_points, X0=(0.0, 0.0, 0.0), X1=(1.0, 1.0, 1.0)):
    """make_rect(m, Elem(), X0, X1, nb_points) (mesh/make_rect.h:389-430, 443-544, 653-679)."""
    lib = A.load()
    nb = np.ascontiguousarray(nb_points, dtype=np.uint32)
    x0 = np.ascontiguousarray(X0, dtype=np.float64)
    x1 = np.ascontiguousarray(X1, dtype=np.float64)
    nn, ne = C.c_uint32(), C.c_uint32()
    A.check(lib.femgpu_make_rect(kind, _ptr(x0), _ptr(x1), _ptr(nb), None, None, C.byref(nn), C.byref(ne)), "make_rect")
    coords = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, A.NPE[kind]), dtype=np.uint32)
    A.check(lib.femgpu_make_rect(kind, _ptr(x0), _ptr(x1), _ptr(nb), _ptr(coords), _ptr(conn), C.byref(nn),
                                 C.byref(ne)), "make_rect")
    return coords, conn


def skin_faces(kind, n_nodes, conn):
    """Faces with exactly one parent element (mesh.h:151-188), reference child node order."""
    lib = A.load()
    conn = np.ascontiguousarray(conn, dtype=np.uint32)
    n = C.c_uint32()
    A.check(lib.femgpu_skin_faces(kind, n_nodes, conn.shape[0], _ptr(conn), None, C.byref(n)), "skin_faces")
    faces = np.zeros((n.value, A.NPF[A.FACE_OF[kind]]), dtype=np.uint32)
    A.check(lib.femgpu_skin_faces(kind, n_nodes, conn.shape[0], _ptr(conn), _ptr(faces), C.byref(n)), "skin_faces")
    return faces


def faces_on_plane(coords, faces, axis, value, tol=1e-6):
    """Skin faces whose nodes all lie on coordinate plane x_axis = value (within tol)."""
    on = np.abs(coords[faces][:, :, axis] - value) < tol
    return faces[np.all(on, axis=1)]


def nodes_on_plane(coords, axis, value, tol=1e-6):
    return np.where(np.abs(coords[:, axis] - value) < tol)[0]


def clamp_dofs(nodes):
    nodes = np.asarray(nodes, dtype=np.int64)
    return (3 * nodes[:, None] + np.arange(3)[None, :]).reshape(-1).astype(np.uint32)
