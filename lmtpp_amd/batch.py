"""Batched uneven problems: many independent elasticity systems assembled and solved by one call
(femgpu_batch_*).  Each sub-problem keeps its own material, tractions, constraints, CG scalars
and stop test; chunks of iterations are replayed from a captured hipGraph."""
import ctypes as C

import numpy as np

from . import _capi as A


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Batch:
    def __init__(self, kind, meshes):
        """meshes: list of (coords[n,3], conn[m,npe]) in the reference layout."""
        lib = A.load()
        self._lib = lib
        self.kind = int(kind)
        self._keep = []
        arr = (A.Mesh * len(meshes))()
        self.sizes = []
        for i, (c, e) in enumerate(meshes):
            c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1, 3)
            e = np.ascontiguousarray(e, dtype=np.uint32).reshape(-1, A.NPE[self.kind])
            self._keep += [c, e]
            arr[i] = A.Mesh(self.kind, c.shape[0], _ptr(c), e.shape[0], _ptr(e))
            self.sizes.append((c.shape[0], e.shape[0]))
        h = C.c_void_p()
        A.check(lib.femgpu_batch_create(len(meshes), arr, C.byref(h)), "batch_create")
        self._h = h
        self.n = len(meshes)
        self.last_stats = None

    def close(self):
        if getattr(self, "_h", None):
            self._lib.femgpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_material(self, i, E, nu, body_force=(0.0, 0.0, 0.0)):
        E_arr = np.asarray(E, dtype=np.float64)
        keep = None
        if E_arr.ndim == 0:
            mat = A.Material(float(nu), float(E_arr), None, (C.c_double * 3)(*body_force))
        else:
            keep = np.ascontiguousarray(E_arr.reshape(-1))
            mat = A.Material(float(nu), 0.0, _ptr(keep), (C.c_double * 3)(*body_force))
        A.check(self._lib.femgpu_batch_set_material(self._h, i, C.byref(mat)), "batch_set_material")
        del keep

    def set_traction(self, i, faces, traction):
        faces = np.ascontiguousarray(faces, dtype=np.uint32)
        tr = np.ascontiguousarray(np.broadcast_to(np.asarray(traction, dtype=np.float64), (faces.shape[0], 3)))
        A.check(self._lib.femgpu_batch_set_traction(self._h, i, faces.shape[0], _ptr(faces), _ptr(tr)), "batch_traction")

    def add_constraints(self, i, dofs, values=None, penalty=1e5):
        dofs = np.ascontiguousarray(dofs, dtype=np.uint32)
        vals = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        A.check(self._lib.femgpu_batch_add_constraints(self._h, i, dofs.shape[0], _ptr(dofs),
                                                       None if vals is None else _ptr(vals), float(penalty)),
                "batch_add_constraints")

    def assemble(self):
        A.check(self._lib.femgpu_batch_assemble(self._h), "batch_assemble")

    def assemble_ms(self):
        v = C.c_double()
        A.check(self._lib.femgpu_batch_get_assemble_ms(self._h, C.byref(v)), "batch_assemble_ms")
        return v.value

    def _opts(self, rel_tol, precision, precond, max_iter, chunk):
        o = A.SolveOpts()
        self._lib.femgpu_default_solve_opts(C.byref(o))
        o.precision, o.precond, o.crit_kind, o.tol, o.max_iter, o.chunk = (
            int(precision), int(precond), A.REL_L2, float(rel_tol), int(max_iter), int(chunk))
        return o

    def solve(self, rel_tol=1e-10, precision=A.MIXED, precond=A.JACOBI, max_iter=100000, chunk=32):
        o = self._opts(rel_tol, precision, precond, max_iter, chunk)
        st = (A.SolveStats * self.n)()
        ms = C.c_double()
        rc = self._lib.femgpu_batch_solve(self._h, C.byref(o), st, C.byref(ms))
        self.last_stats = list(st)
        if rc in (A.BAD_INPUT, A.HIP_ERROR, A.NOT_READY):
            A.check(rc, "batch_solve")
        return rc == A.OK, ms.value

    def set_spmv_config(self, fp64=(A.SPMV_FULL, 2, 1), fp32=(A.SPMV_FULL, 2, 1)):
        """Block-SpMV variant per precision: (format, unroll, loads), as Formulation."""
        f = (C.c_int * 2)(int(fp64[0]), int(fp32[0]))
        u = (C.c_int * 2)(int(fp64[1]), int(fp32[1]))
        nt = (C.c_int * 2)(int(fp64[2]), int(fp32[2]))
        A.check(self._lib.femgpu_batch_set_spmv_config(self._h, f, u, nt), "batch_set_spmv_config")

    def autotune(self, rel_tol=1e-10, precision=A.MIXED, precond=A.JACOBI, max_iter=100000, chunk=32):
        """Product variant and element loop chosen by timing whole batched solves with these
        options (femgpu_batch_autotune); returns the chosen product (A.BATCH_*).  The batch's
        results are left as they were."""
        o = self._opts(rel_tol, precision, precond, max_iter, chunk)
        prod = C.c_int()
        A.check(self._lib.femgpu_batch_autotune(self._h, C.byref(o), C.byref(prod)), "batch_autotune")
        return prod.value

    def chunk_stats(self):
        """Iteration chunks of the last solve: (chunks, device ms inside them, gaps between
        chunks with no refresh, device ms waited in those gaps)."""
        n, g = C.c_int(), C.c_int()
        cm, gm = C.c_double(), C.c_double()
        A.check(self._lib.femgpu_batch_get_chunk_stats(self._h, C.byref(n), C.byref(cm), C.byref(g), C.byref(gm)),
                "batch_get_chunk_stats")
        return n.value, cm.value, g.value, gm.value

    def set_assembly_rows(self, rows):
        """Element-loop variant, as Formulation.set_assembly_rows."""
        A.check(self._lib.femgpu_batch_set_assembly_rows(self._h, int(rows)), "batch_set_assembly_rows")

    def assembly_rows(self):
        """(current element-loop variant, largest admissible R), as Formulation.assembly_rows."""
        r, m = C.c_uint32(), C.c_uint32()
        A.check(self._lib.femgpu_batch_get_assembly_rows(self._h, C.byref(r), C.byref(m)), "batch_get_assembly_rows")
        return r.value, m.value

    def clear_result(self):
        """Zero initial guess for every problem (solves otherwise start from the previous result)."""
        A.check(self._lib.femgpu_batch_clear_result(self._h), "batch_clear_result")

    def call_after_solve(self, i):
        """Per-element (epsilon[n,6], sigma[n,6], tr_epsilon[n]) of problem i (Formulation.call_after_solve)."""
        ne = self.sizes[i][1]
        eps, sig, tr = np.zeros((ne, 6)), np.zeros((ne, 6)), np.zeros(ne)
        A.check(self._lib.femgpu_batch_after_solve(self._h, i, _ptr(eps), _ptr(sig), _ptr(tr)), "batch_after_solve")
        return eps, sig, tr

    def write_vtu(self, i, path):
        """Problem i's mesh and result as a VTK unstructured grid (Formulation.write_vtu)."""
        A.check(self._lib.femgpu_batch_write_result_vtu(self._h, int(i), str(path).encode()), "batch_write_result_vtu")

    def write_avs(self, i, path, display=None):
        """Formulation.write_avs for problem i of the batch."""
        nd, dp = A.display_names(display)
        A.check(self._lib.femgpu_batch_write_result_avs(self._h, int(i), str(path).encode(), nd, dp),
                "batch_write_result_avs")

    def get_result(self, i):
        u = np.zeros(3 * self.sizes[i][0])
        A.check(self._lib.femgpu_batch_get_result(self._h, i, _ptr(u), u.size), "batch_get_result")
        return u
