"""ctypes binding of libfemgpu.so, the C ABI declared in include/femgpu.h.

The library is the only compute path: importing this module on a machine with a GPU and
without the built library raises instead of falling back to anything else.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfemgpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "femgpu.h")

OK, BAD_INPUT, NEG_JACOBIAN, NOT_SPD, MAX_ITER, HIP_ERROR, NAN, NOT_READY, STAGNATED = range(9)
TETRA, TETRA10, HEXA = 0, 1, 2
TRI3, TRI6, QUAD4 = 0, 1, 2
FP64, MIXED = 0, 1
JACOBI, BLOCK_JACOBI, BLOCK_IC0, BLOCK_IC0_COLOURED, CHEBYSHEV_BLOCK_JACOBI = 0, 1, 2, 3, 4
REL_L2, ABS_MAXNORM = 0, 1
SPMV_FULL, SPMV_SYM, SPMV_MATRIX_FREE = 0, 1, 2
LOADS_PLAIN, LOADS_NT, LOADS_NT_COL16 = 0, 1, 2
ROW_ORDER_AUTO, ROW_ORDER_NATURAL, ROW_ORDER_BY_LENGTH = 0, 1, 2
NODE_ORDER_NATURAL, NODE_ORDER_RCM, NODE_ORDER_AMD = 0, 1, 2
BATCH_STORED_PLAIN, BATCH_STORED_NT, BATCH_STORED_NT_COL16, BATCH_MATRIX_FREE = 0, 1, 2, -1
ASSEMBLY_STAGED = 0xFFFFFFFF
ASSEMBLY_GATHER = 0xFFFFFFFE
ASSEMBLY_GATHER_BOX = 0xFFFFFFFD
NPE = {TETRA: 4, TETRA10: 10, HEXA: 8}
FACE_OF = {TETRA: TRI3, TETRA10: TRI6, HEXA: QUAD4}
NPF = {TRI3: 3, TRI6: 6, QUAD4: 4}


class Mesh(C.Structure):
    _fields_ = [("elem_kind", C.c_int), ("n_nodes", C.c_uint32), ("coords", C.c_void_p),
                ("n_elem", C.c_uint32), ("conn", C.c_void_p)]


class Material(C.Structure):
    _fields_ = [("nu", C.c_double), ("E", C.c_double), ("E_elem", C.c_void_p), ("body_force", C.c_double * 3)]


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("precond", C.c_int), ("crit_kind", C.c_int), ("tol", C.c_double),
                ("max_iter", C.c_int), ("reliable_delta", C.c_double), ("chunk", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("refreshes", C.c_int), ("rel_res", C.c_double),
                ("max_abs_res", C.c_double), ("solve_ms", C.c_double)]


class VtuField(C.Structure):
    _fields_ = [("name", C.c_char_p), ("n_comp", C.c_uint32), ("data", C.c_void_p)]


class AvsField(C.Structure):
    _fields_ = [("name", C.c_char_p), ("n_comp", C.c_uint32), ("nest", C.c_uint32), ("data", C.c_void_p)]


class FemGpuError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (status %d)" % (msg, status))
        self.status = status


_lib = None

_P = C.c_void_p
_SIGS = {
    "femgpu_create": (C.c_int, [C.POINTER(Mesh), C.POINTER(_P)]),
    "femgpu_destroy": (C.c_int, [_P]),
    "femgpu_set_material": (C.c_int, [_P, C.POINTER(Material)]),
    "femgpu_set_elastic_modulus_device": (C.c_int, [_P, _P]),
    "femgpu_set_traction": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "femgpu_add_constraint": (C.c_int, [_P, C.c_uint32, C.c_double, C.c_double]),
    "femgpu_add_constraints": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_double]),
    "femgpu_clear_constraints": (C.c_int, [_P]),
    "femgpu_assemble": (C.c_int, [_P, C.c_int, C.c_int]),
    "femgpu_solve_system": (C.c_int, [_P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "femgpu_solve": (C.c_int, [_P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "femgpu_default_solve_opts": (None, [C.POINTER(SolveOpts)]),
    "femgpu_get_result": (C.c_int, [_P, _P, C.c_size_t]),
    "femgpu_get_result_device": (C.c_int, [_P, _P]),
    "femgpu_set_result": (C.c_int, [_P, _P, C.c_size_t]),
    "femgpu_get_sollicitation": (C.c_int, [_P, _P, C.c_size_t]),
    "femgpu_set_sollicitation": (C.c_int, [_P, _P, C.c_size_t]),
    "femgpu_apply_noalloc": (C.c_int, [_P, C.c_int, _P, _P]),
    "femgpu_add_constraint_text": (C.c_int, [_P, C.c_char_p, C.c_double, C.POINTER(C.c_uint32)]),
    "femgpu_unshift": (C.c_int, [_P, C.c_int]),
    "femgpu_constraint_parse": (C.c_int, [C.c_char_p, C.c_double, C.c_uint32, _P, _P, C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_double)]),
    "femgpu_get_result_history": (C.c_int, [_P, C.c_uint32, _P, C.c_size_t]),
    "femgpu_get_indice_noda": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "femgpu_get_nodal_forces": (C.c_int, [_P, _P, C.c_size_t]),
    "femgpu_get_matrix": (C.c_int, [_P, _P, _P, _P]),
    "femgpu_get_sizes": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint32)]),
    "femgpu_get_max_diag": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "femgpu_adjoint": (C.c_int, [_P, _P, _P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "femgpu_adjoint_device": (C.c_int, [_P, _P, _P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "femgpu_elem_grad": (C.c_int, [_P, _P, _P, _P]),
    "femgpu_create_from_bsr": (C.c_int, [C.c_uint32, _P, _P, _P, _P, C.POINTER(_P)]),
    "femgpu_make_rect": (C.c_int, [C.c_int, _P, _P, _P, _P, _P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "femgpu_skin_faces": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, _P, _P, C.POINTER(C.c_uint32)]),
    "femgpu_batch_create": (C.c_int, [C.c_uint32, C.POINTER(Mesh), C.POINTER(_P)]),
    "femgpu_batch_set_material": (C.c_int, [_P, C.c_uint32, C.POINTER(Material)]),
    "femgpu_batch_set_traction": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    "femgpu_batch_add_constraints": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, C.c_double]),
    "femgpu_batch_assemble": (C.c_int, [_P]),
    "femgpu_batch_solve": (C.c_int, [_P, C.POINTER(SolveOpts), _P, C.POINTER(C.c_double)]),
    "femgpu_batch_get_result": (C.c_int, [_P, C.c_uint32, _P, C.c_size_t]),
    "femgpu_batch_destroy": (C.c_int, [_P]),
    "femgpu_status_string": (C.c_char_p, [C.c_int]),
    "femgpu_last_error": (C.c_char_p, []),
    "femgpu_get_assemble_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "femgpu_time_spmv": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "femgpu_apply": (C.c_int, [_P, C.c_int, _P, _P, C.c_size_t]),
    "femgpu_matrix_free_info": (C.c_int, [_P, _P]),
    "femgpu_time_matrix_free": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "femgpu_time_iteration": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "femgpu_autotune": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "femgpu_stream_copy": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "femgpu_stream_read": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "femgpu_set_graph_replay": (C.c_int, [_P, C.c_int]),
    "femgpu_batch_get_assemble_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "femgpu_batch_clear_result": (C.c_int, [_P]),
    "femgpu_after_solve": (C.c_int, [_P, _P, _P, _P]),
    "femgpu_after_solve_device": (C.c_int, [_P, _P, _P, _P]),
    "femgpu_get_after_solve_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "femgpu_batch_after_solve": (C.c_int, [_P, C.c_uint32, _P, _P, _P]),
    "femgpu_get_spmv_config": (C.c_int, [_P, _P, _P, _P]),
    "femgpu_set_assembly_rows": (C.c_int, [_P, C.c_uint32]),
    "femgpu_batch_set_assembly_rows": (C.c_int, [_P, C.c_uint32]),
    "femgpu_get_assembly_rows": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "femgpu_batch_get_assembly_rows": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "femgpu_set_spmv_config": (C.c_int, [_P, _P, _P, _P]),
    "femgpu_batch_autotune": (C.c_int, [_P, _P, C.POINTER(C.c_int)]),
    "femgpu_get_chunk_stats": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                         C.POINTER(C.c_double)]),
    "femgpu_batch_get_chunk_stats": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                               C.POINTER(C.c_double)]),
    "femgpu_batch_set_spmv_config": (C.c_int, [_P, _P, _P, _P]),
    "femgpu_get_precond_setup_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "femgpu_get_ic0_factor": (C.c_int, [_P, C.c_int, _P, _P, _P, _P, C.POINTER(C.c_uint32)]),
    "femgpu_ic0_schedule": (C.c_int, [C.c_uint32, _P, _P, C.c_int, _P, _P, C.POINTER(C.c_uint32)]),
    "femgpu_set_row_order": (C.c_int, [C.c_int]),
    "femgpu_node_order": (C.c_int, [C.c_uint32, _P, _P, _P]),
    "femgpu_node_order_amd": (C.c_int, [C.c_uint32, _P, _P, _P]),
    "femgpu_set_node_order": (C.c_int, [C.c_int]),
    "femgpu_get_node_order": (C.c_int, [_P, C.POINTER(C.c_int), _P]),
    "femgpu_write_vtu": (C.c_int, [C.c_char_p, C.c_int, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _P,
                                   _P, C.c_uint32, _P]),
    "femgpu_write_result_vtu": (C.c_int, [_P, C.c_char_p]),
    "femgpu_batch_write_result_vtu": (C.c_int, [_P, C.c_uint32, C.c_char_p]),
    "femgpu_write_avs": (C.c_int, [C.c_char_p, C.c_int, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _P,
                                   C.c_uint32, _P, C.c_uint32, _P]),
    "femgpu_write_result_avs": (C.c_int, [_P, C.c_char_p, C.c_uint32, _P]),
    "femgpu_batch_write_result_avs": (C.c_int, [_P, C.c_uint32, C.c_char_p, C.c_uint32, _P]),
    "femgpu_shift": (C.c_int, [_P, C.c_int]),
    "femgpu_set_initial_time_step": (C.c_int, [_P, C.c_double]),
    "femgpu_set_next_time_step": (C.c_int, [_P, C.c_double]),
    "femgpu_get_next_time_step": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "femgpu_get_time": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "femgpu_set_time": (C.c_int, [_P, C.c_double]),
    "femgpu_get_previous_result": (C.c_int, [_P, _P, C.c_size_t]),
    "femgpu_set_assume_constant_matrix": (C.c_int, [_P, C.c_int]),
    "femgpu_set_nonlinear": (C.c_int, [_P, C.c_double, C.c_int]),
    "femgpu_get_nonlinear_solves": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "femgpu_set_default_solve_opts": (C.c_int, [_P, C.POINTER(SolveOpts)]),
    "femgpu_get_precond": (C.c_int, [_P, C.POINTER(SolveOpts)]),
    "femgpu_solve_system_using_precond": (C.c_int, [_P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "femgpu_stepper_create": (C.c_int, [C.c_uint32, _P, C.POINTER(_P)]),
    "femgpu_stepper_destroy": (C.c_int, [_P]),
    "femgpu_stepper_iteration": (C.c_int, [_P, _P, _P]),
    "femgpu_stepper_get_time": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "femgpu_get_row_order": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "femgpu_sell_order": (C.c_int, [C.c_uint32, _P, C.c_int, _P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
}


def load():
    """Load libfemgpu.so (raises if it is missing: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("lmtpp_amd: %s is not built (run __graft_entry__.build())" % LIB_PATH)
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(rc, what=""):
    if rc != OK:
        lib = load()
        raise FemGpuError(rc, "%s: %s" % (what, lib.femgpu_last_error().decode()))
    return rc


def stream_copy_gbs(n_doubles=1 << 27, reps=20):
    """Measured copy bandwidth (GB/s, read + write bytes) of the library's plain copy kernel."""
    lib = load()
    ms = C.c_double()
    check(lib.femgpu_stream_copy(n_doubles, reps, C.byref(ms)), "stream_copy")
    return 16.0 * n_doubles / (ms.value * 1e-3) / 1e9, ms.value


def constraint_parse(txt, time=0.0):
    """The affine form of a text constraint at time (host only): ({dof: coefficient}, constant)."""
    import numpy as np
    lib = load()
    n, b = C.c_uint32(), C.c_double()
    check(lib.femgpu_constraint_parse(txt.encode(), float(time), 0, None, None, C.byref(n), C.byref(b)), "constraint_parse")
    dof = np.zeros(max(1, n.value), np.uint64)
    co = np.zeros(max(1, n.value))
    check(lib.femgpu_constraint_parse(txt.encode(), float(time), n.value, dof.ctypes.data, co.ctypes.data, C.byref(n),
                                      C.byref(b)), "constraint_parse")
    return {int(d): float(v) for d, v in zip(dof[:n.value], co[:n.value])}, b.value


def ic0_schedule(row_ptr, col, precond=BLOCK_IC0_COLOURED):
    """Elimination order and level sets of the block IC(0) for a full 3x3-block CSR pattern
    (host only).  Returns (order, level_ptr)."""
    import numpy as np
    lib = load()
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    cl = np.ascontiguousarray(col, dtype=np.uint32)
    n = len(rp) - 1
    order = np.zeros(n, np.uint32)
    nl = C.c_uint32(0)
    check(lib.femgpu_ic0_schedule(n, rp.ctypes.data, cl.ctypes.data, precond, order.ctypes.data, None, C.byref(nl)),
          "ic0_schedule")
    lv = np.zeros(nl.value + 1, np.uint32)
    check(lib.femgpu_ic0_schedule(n, rp.ctypes.data, cl.ctypes.data, precond, order.ctypes.data, lv.ctypes.data,
                                  C.byref(nl)), "ic0_schedule")
    return order, lv


def sell_order(row_ptr, order=ROW_ORDER_BY_LENGTH):
    """Storage order the library picks for a block-row length profile (host only).
    Returns (perm, n_natural, n_positions): perm[t] = block row stored at t."""
    import numpy as np
    lib = load()
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    n = len(rp) - 1
    perm = np.zeros(n, np.uint32)
    nat, npos = C.c_uint64(), C.c_uint64()
    check(lib.femgpu_sell_order(n, rp.ctypes.data, order, perm.ctypes.data, C.byref(nat), C.byref(npos)), "sell_order")
    return perm, nat.value, npos.value


def write_vtu(path, kind, coords, conn, point=(), cell=(), group=None, glob=()):
    """Result file in the reference writer's layout (femgpu_write_vtu, host only).  point / cell /
    glob: sequences of (name, array) with one row of components per node / element, or one row
    for a global field."""
    import numpy as np
    lib = load()
    keep = []

    def fields(items, rows):
        arr = (VtuField * max(len(items), 1))()
        for i, (name, a) in enumerate(items):
            a = np.ascontiguousarray(a, dtype=np.float64)
            a = a.reshape(rows, -1) if rows else a.reshape(1, -1)
            keep.append(a)
            arr[i] = VtuField(name.encode(), a.shape[1], a.ctypes.data)
        return arr

    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(-1, NPE[int(kind)])
    pf, cf, gf = fields(point, len(coords)), fields(cell, len(conn)), fields(glob, 0)
    g = None
    if group is not None:
        g = np.ascontiguousarray(group, dtype=np.uint32)
        keep.append(g)
    check(lib.femgpu_write_vtu(str(path).encode(), int(kind), len(coords), coords.ctypes.data, len(conn),
                               conn.ctypes.data, len(point), pf, len(cell), cf, None if g is None else g.ctypes.data,
                               len(glob), gf), "write_vtu")


def display_names(display):
    """A display list for the AVS writers: (count, char** array or None); None means "all"."""
    if display is None:
        return 0, None
    if isinstance(display, str):
        display = [display]
    arr = (C.c_char_p * max(len(display), 1))(*[d.encode() for d in display])
    return len(display), arr


def write_avs(path, kind, coords, conn, point=(), cell=(), glob=(), display=("all",)):
    """AVS UCD ASCII file in the reference writer's layout and bytes (femgpu_write_avs, host only).
    point / cell / glob: sequences of (name, array) or (name, array, nest) with one row of
    components per node / element, or one row for a global field; nest 0 is a scalar, 1 (the
    default) a Vec, 2 a Vec of one Vec (epsilon, sigma).  display: the field names to write, or
    ("all",)."""
    import numpy as np
    lib = load()
    keep = []

    def fields(items, rows):
        arr = (AvsField * max(len(items), 1))()
        for i, it in enumerate(items):
            name, a = it[0], it[1]
            nest = it[2] if len(it) > 2 else 1
            a = np.ascontiguousarray(a, dtype=np.float64)
            a = a.reshape(rows, -1) if rows else a.reshape(1, -1)
            keep.append(a)
            arr[i] = AvsField(name.encode(), a.shape[1], nest, a.ctypes.data)
        return arr

    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(-1, NPE[int(kind)])
    pf, cf, gf = fields(point, len(coords)), fields(cell, len(conn)), fields(glob, 0)
    nd, dp = display_names(display)
    check(lib.femgpu_write_avs(str(path).encode(), int(kind), len(coords), coords.ctypes.data, len(conn),
                               conn.ctypes.data, len(point), pf, len(cell), cf, len(glob), gf, nd, dp), "write_avs")


def node_order(row_ptr, col):
    """Cuthill-McKee visiting order of a node graph (femgpu_node_order, host only)."""
    import numpy as np
    lib = load()
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    cl = np.ascontiguousarray(col, dtype=np.uint32)
    o = np.zeros(len(rp) - 1, np.uint32)
    check(lib.femgpu_node_order(len(rp) - 1, rp.ctypes.data, cl.ctypes.data, o.ctypes.data), "node_order")
    return o


def node_order_amd(row_ptr, col):
    """Approximate minimum degree order of a symmetric node graph, the node at each position
    (femgpu_node_order_amd, host only; symamd's result is its inverse)."""
    import numpy as np
    lib = load()
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    cl = np.ascontiguousarray(col, dtype=np.uint32)
    o = np.zeros(len(rp) - 1, np.uint32)
    check(lib.femgpu_node_order_amd(len(rp) - 1, rp.ctypes.data, cl.ctypes.data, o.ctypes.data), "node_order_amd")
    return o


class row_order:
    """Context manager: problems created inside use the given block-row storage order."""

    def __init__(self, order):
        self.order = order

    def __enter__(self):
        check(load().femgpu_set_row_order(self.order), "set_row_order")
        return self

    def __exit__(self, *exc):
        check(load().femgpu_set_row_order(ROW_ORDER_AUTO), "set_row_order")
        return False


class node_order_ctx:
    """Context manager: problems created inside store their nodes in the given order."""

    def __init__(self, order):
        self.order = order

    def __enter__(self):
        check(load().femgpu_set_node_order(self.order), "set_node_order")
        return self

    def __exit__(self, *exc):
        check(load().femgpu_set_node_order(NODE_ORDER_NATURAL), "set_node_order")
        return False


def stream_read_gbs(n_doubles=1 << 27, reps=20):
    """Measured streaming-read bandwidth (GB/s) of the library's read kernel."""
    lib = load()
    ms = C.c_double()
    check(lib.femgpu_stream_read(n_doubles, reps, C.byref(ms)), "stream_read")
    return 8.0 * n_doubles / (ms.value * 1e-3) / 1e9, ms.value


def declared_symbols():
    """Names of every function declared in include/femgpu.h."""
    import re
    with open(HEADER) as f:
        txt = f.read()
    return sorted(set(re.findall(r"FEMGPU_API\s+[\w\s\*]+?\b(femgpu_\w+)\s*\(", txt)))
