"""Host-side mirror of LMT++'s FormulationAncestor<double> for the elasticity_iso formulation.

Same call surface as the reference (include/formulation/formulation_ancestor.h:49-64,179-237):
``allocate_matrices()``, ``assemble(assemble_mat, assemble_vec)``, ``solve_system(crit)``,
``solve(crit)``, ``get_result()``, ``get_sollicitation()``, ``get_mat()``, ``add_constraint``,
``shift()``, ``update_variables()``.  Every computation runs in libfemgpu.so on the GPU.
"""
import ctypes as C

import numpy as np

from . import _capi as A


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Formulation:
    """One elasticity_iso formulation on one mesh (formulation.h:102-2787).

    ``solve_system(crit)`` keeps the reference meaning of ``crit``: with ``crit > 0`` the solve
    stops when max_i |r_i| <= crit (containers/MatWithTinyBlocks.h:1115,1136); ``rel_tol``
    selects the relative L2 stop test instead.  Like the reference, the methods return ``True``
    on success and ``False`` when the system could not be solved (formulation.h:1394,1530-1531);
    the reason is kept in ``last_stats`` / ``last_status``.
    """

    def __init__(self, kind, coords, conn):
        lib = A.load()
        self.kind = int(kind)
        self.coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
        self.conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(-1, A.NPE[self.kind])
        m = A.Mesh(self.kind, self.coords.shape[0], _ptr(self.coords), self.conn.shape[0], _ptr(self.conn))
        h = C.c_void_p()
        A.check(lib.femgpu_create(C.byref(m), C.byref(h)), "femgpu_create")
        self._h = h
        self._lib = lib
        self.n_nodes = self.coords.shape[0]
        self.n_elem = self.conn.shape[0]
        # public knobs of FormulationAncestor (formulation_ancestor.h:244-257)
        self.default_iterative_criterium = 1e-10
        self.precision = A.FP64
        self.precond = A.JACOBI
        self.max_iter = 100000
        self.last_stats = None
        self.last_status = A.OK
        self.result = np.zeros(3 * self.n_nodes)
        self._const, self._nl_crit, self._max_nl = False, 0.0, 50   # the library's defaults

    # -- lifecycle -----------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.femgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- inputs --------------------------------------------------------------------------------
    def set_material(self, E=210000.0, nu=0.3, body_force=(0.0, 0.0, 0.0)):
        """elastic_modulus (global float or per-element array), poisson_ratio, density*f_vol."""
        E_arr = np.asarray(E, dtype=np.float64)
        keep = None
        if E_arr.ndim == 0:
            mat = A.Material(float(nu), float(E_arr), None, (C.c_double * 3)(*body_force))
        else:
            keep = np.ascontiguousarray(E_arr.reshape(-1))
            if keep.shape[0] != self.n_elem:
                raise ValueError("E must have one value per element")
            mat = A.Material(float(nu), 0.0, _ptr(keep), (C.c_double * 3)(*body_force))
        A.check(self._lib.femgpu_set_material(self._h, C.byref(mat)), "set_material")
        del keep

    def set_traction(self, faces, traction):
        faces = np.ascontiguousarray(faces, dtype=np.uint32)
        tr = np.ascontiguousarray(np.broadcast_to(np.asarray(traction, dtype=np.float64), (faces.shape[0], 3)))
        A.check(self._lib.femgpu_set_traction(self._h, faces.shape[0], _ptr(faces), _ptr(tr)), "set_traction")

    def add_constraint(self, dof, value=0.0, penalty=None):
        """Penalty constraint u[dof] = value (formulation.h:1028-1070), penalty 1e5 by default.
        With a string first, add_constraint(txt, penalty_value) as the reference's
        (formulation_ancestor.h:197): txt is an expression in node[i].dep[d] and time whose zero
        is the constraint, e.g. "node[3].dep[2] + 0.01*time"; returns the constraint's number.
        penalty_value must be > 0 (the reference's default 0, the Lagrange form, is refused)."""
        if isinstance(dof, str):
            pen = float(value) if penalty is None else float(penalty)
            num = C.c_uint32()
            A.check(self._lib.femgpu_add_constraint_text(self._h, dof.encode(), pen, C.byref(num)), "add_constraint")
            return num.value
        A.check(self._lib.femgpu_add_constraint(self._h, int(dof), float(value), 1e5 if penalty is None else float(penalty)),
                "add_constraint")

    def add_constraints(self, dofs, values=None, penalty=1e5):
        dofs = np.ascontiguousarray(dofs, dtype=np.uint32)
        vals = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        A.check(self._lib.femgpu_add_constraints(self._h, dofs.shape[0], _ptr(dofs),
                                                 None if vals is None else _ptr(vals), float(penalty)),
                "add_constraints")

    def clear_constraints(self):
        A.check(self._lib.femgpu_clear_constraints(self._h), "clear_constraints")

    # -- hot path ------------------------------------------------------------------------------
    def allocate_matrices(self, *_):
        """Pattern and DOF numbering are built once at construction (formulation.h:338-412)."""
        return None

    def assemble(self, assemble_mat=True, assemble_vec=True):
        A.check(self._lib.femgpu_assemble(self._h, int(bool(assemble_mat)), int(bool(assemble_vec))), "assemble")

    def _opts(self, crit, rel_tol):
        o = A.SolveOpts()
        self._lib.femgpu_default_solve_opts(C.byref(o))
        o.precision = self.precision
        o.precond = self.precond
        o.max_iter = self.max_iter
        if rel_tol is not None:
            o.crit_kind, o.tol = A.REL_L2, float(rel_tol)
        else:
            c = crit if crit else self.default_iterative_criterium
            o.crit_kind, o.tol = A.ABS_MAXNORM, float(c)
        return o

    def solve_system(self, crit=0.0, rel_tol=None):
        o = self._opts(crit, rel_tol)
        st = A.SolveStats()
        rc = self._lib.femgpu_solve_system(self._h, C.byref(o), C.byref(st))
        self.last_stats, self.last_status = st, rc
        if rc in (A.BAD_INPUT, A.HIP_ERROR, A.NOT_READY):
            A.check(rc, "solve_system")
        return rc == A.OK

    def chunk_stats(self):
        """Iteration chunks of the last solve: (chunks, device ms inside them, gaps between
        chunks with no refresh, device ms waited in those gaps) (femgpu_get_chunk_stats)."""
        n, g = C.c_int(), C.c_int()
        cm, gm = C.c_double(), C.c_double()
        A.check(self._lib.femgpu_get_chunk_stats(self._h, C.byref(n), C.byref(cm), C.byref(g), C.byref(gm)),
                "get_chunk_stats")
        return n.value, cm.value, g.value, gm.value

    def solve(self, crit=0.0, rel_tol=None):
        """solve(crit) (formulation.h:1517-1568): shift, then assemble + solve_system, repeated
        while non_linear_iterative_criterium is set and the iterate still moves; True on success.
        A loop that reaches max_non_linear_iteration raises (the reference throws)."""
        o = self._opts(crit, rel_tol)
        st = A.SolveStats()
        rc = self._lib.femgpu_solve(self._h, C.byref(o), C.byref(st))
        self.last_stats, self.last_status = st, rc
        if rc in (A.BAD_INPUT, A.HIP_ERROR, A.NOT_READY, A.NEG_JACOBIAN):
            A.check(rc, "solve")
        if rc == A.MAX_ITER and st.status == A.OK:   # the last solve converged: the non-linear loop ran out
            A.check(rc, "solve")
        if rc == A.OK:
            self.update_variables()
        return rc == A.OK

    # -- time steps, non-linear loop, preconditioner reuse --------------------------------------
    @property
    def assume_constant_matrix(self):
        """formulation_ancestor.h:244: Problem.iteration assembles K and its preconditioner once."""
        return self._const

    @assume_constant_matrix.setter
    def assume_constant_matrix(self, v):
        A.check(self._lib.femgpu_set_assume_constant_matrix(self._h, int(bool(v))), "assume_constant_matrix")
        self._const = bool(v)

    @property
    def non_linear_iterative_criterium(self):
        return self._nl_crit

    @non_linear_iterative_criterium.setter
    def non_linear_iterative_criterium(self, v):
        A.check(self._lib.femgpu_set_nonlinear(self._h, float(v), int(self._max_nl)), "set_nonlinear")
        self._nl_crit = float(v)

    @property
    def max_non_linear_iteration(self):
        return self._max_nl

    @max_non_linear_iteration.setter
    def max_non_linear_iteration(self, v):
        A.check(self._lib.femgpu_set_nonlinear(self._h, float(self._nl_crit), int(v)), "set_nonlinear")
        self._max_nl = int(v)

    def nonlinear_solves(self):
        """solve_system calls made by the last solve() or Problem.iteration()."""
        n = C.c_int()
        A.check(self._lib.femgpu_get_nonlinear_solves(self._h, C.byref(n)), "get_nonlinear_solves")
        return n.value

    def shift(self, nb=1):
        """shift(nb) (formulation.h:424-433): time += time step; the previous solution moves to
        vectors[1]; vectors[0] stays the next initial guess."""
        A.check(self._lib.femgpu_shift(self._h, int(nb)), "shift")

    def set_initial_time_step(self, ts):
        A.check(self._lib.femgpu_set_initial_time_step(self._h, float(ts)), "set_initial_time_step")

    def set_next_time_step(self, ts):
        A.check(self._lib.femgpu_set_next_time_step(self._h, float(ts)), "set_next_time_step")

    def get_next_time_step(self):
        v = C.c_double()
        A.check(self._lib.femgpu_get_next_time_step(self._h, C.byref(v)), "get_next_time_step")
        return v.value

    def get_time(self):
        v = C.c_double()
        A.check(self._lib.femgpu_get_time(self._h, C.byref(v)), "get_time")
        return v.value

    def set_time(self, t):
        A.check(self._lib.femgpu_set_time(self._h, float(t)), "set_time")

    def unshift(self, nb=1):
        """unshift(nb) (formulation.h:435-445): time -= time_steps[1]; the history moves down."""
        A.check(self._lib.femgpu_unshift(self._h, int(nb)), "unshift")

    def get_result_num(self, num=0):
        """get_result(num) (formulation_ancestor.h:237): vectors[num], 0 <= num < 20."""
        u = np.zeros(3 * self.n_nodes)
        A.check(self._lib.femgpu_get_result_history(self._h, int(num), _ptr(u), u.size), "get_result")
        return u

    def get_indice_noda(self, node):
        """get_indice_noda(num_node) (formulation_ancestor.h:188): index of the node's first unknown."""
        v = C.c_uint32()
        A.check(self._lib.femgpu_get_indice_noda(self._h, int(node), C.byref(v)), "get_indice_noda")
        return v.value

    def get_nodal_forces(self):
        """get_nodal_forces() (formulation.h:447-458): body and surface loads without constraint terms."""
        F = np.zeros(3 * self.n_nodes)
        A.check(self._lib.femgpu_get_nodal_forces(self._h, _ptr(F), F.size), "get_nodal_forces")
        return F

    def get_previous_result(self):
        """vectors[1]: the solution before the last shift."""
        u = np.zeros(3 * self.n_nodes)
        A.check(self._lib.femgpu_get_previous_result(self._h, _ptr(u), u.size), "get_previous_result")
        return u

    def get_precond(self, crit=0.0, rel_tol=None):
        """get_precond() (formulation.h:1655): build self.precond for the current K and keep it for
        solve_system_using_precond, through later assemblies."""
        o = self._opts(crit, rel_tol)
        A.check(self._lib.femgpu_get_precond(self._h, C.byref(o)), "get_precond")

    def solve_system_using_precond(self, crit=0.0, rel_tol=None):
        """solve_system_using_precond(crit) (formulation.h:1656): PCG with get_precond's
        preconditioner."""
        o = self._opts(crit, rel_tol)
        st = A.SolveStats()
        rc = self._lib.femgpu_solve_system_using_precond(self._h, C.byref(o), C.byref(st))
        self.last_stats, self.last_status = st, rc
        if rc in (A.BAD_INPUT, A.HIP_ERROR, A.NOT_READY):
            A.check(rc, "solve_system_using_precond")
        return rc == A.OK

    def _push_defaults(self, rel_tol=None):
        """Hand default_iterative_criterium, precision, precond and max_iter to the library for
        Problem.iteration."""
        o = self._opts(0.0, rel_tol)
        A.check(self._lib.femgpu_set_default_solve_opts(self._h, C.byref(o)), "set_default_solve_opts")

    # -- results -------------------------------------------------------------------------------
    def get_result(self):
        u = np.zeros(3 * self.n_nodes)
        A.check(self._lib.femgpu_get_result(self._h, _ptr(u), u.size), "get_result")
        return u

    def set_result(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        A.check(self._lib.femgpu_set_result(self._h, _ptr(u), u.size), "set_result")

    def update_variables(self):
        """node.dep[d] = vectors[0][3 i + d] (formulation.h:2571-2576)."""
        self.result = self.get_result()
        return self.result.reshape(-1, 3)

    def get_sollicitation(self):
        F = np.zeros(3 * self.n_nodes)
        A.check(self._lib.femgpu_get_sollicitation(self._h, _ptr(F), F.size), "get_sollicitation")
        return F

    def set_sollicitation(self, F):
        """The right-hand side of the next solve_system (the reference's get_sollicitation()
        returns a mutable Vec&, formulation.h:2647); a later vector assembly overwrites it."""
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
        A.check(self._lib.femgpu_set_sollicitation(self._h, _ptr(F), F.size), "set_sollicitation")

    def sizes(self):
        nn, ne, nz, nc = C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_uint32()
        A.check(self._lib.femgpu_get_sizes(self._h, C.byref(nn), C.byref(ne), C.byref(nz), C.byref(nc)), "sizes")
        return dict(n_nodes=nn.value, n_elem=ne.value, nnzb=nz.value, n_colors=nc.value)

    def get_mat(self):
        """Assembled K as full 3x3-block CSR: (row_ptr, col, vals[nnzb,3,3])."""
        nnzb = self.sizes()["nnzb"]
        rp = np.zeros(self.n_nodes + 1, dtype=np.uint32)
        col = np.zeros(nnzb, dtype=np.uint32)
        vals = np.zeros((nnzb, 3, 3))
        A.check(self._lib.femgpu_get_matrix(self._h, _ptr(rp), _ptr(col), _ptr(vals)), "get_mat")
        return rp, col, vals

    def get_ic0_factor(self, precond=A.BLOCK_IC0):
        """Block IC(0) factor of the assembled K (MatWithTinyBlocks::chol_incomp,
        containers/MatWithTinyBlocks.h:426-523), factorised on the GPU, in elimination numbering:
        (order, row_ptr, col, L[nnzb,3,3], n_levels); L holds L_tc for c < t, the lower-triangular
        diagonal factor for c = t and zeros above."""
        nnzb = self.sizes()["nnzb"]
        order = np.zeros(self.n_nodes, dtype=np.uint32)
        rp = np.zeros(self.n_nodes + 1, dtype=np.uint32)
        col = np.zeros(nnzb, dtype=np.uint32)
        L = np.zeros((nnzb, 3, 3))
        nl = C.c_uint32()
        A.check(self._lib.femgpu_get_ic0_factor(self._h, int(precond), _ptr(order), _ptr(rp), _ptr(col), _ptr(L),
                                                C.byref(nl)), "get_ic0_factor")
        return order, rp, col, L, nl.value

    def max_diag(self):
        v = C.c_double()
        A.check(self._lib.femgpu_get_max_diag(self._h, C.byref(v)), "max_diag")
        return v.value

    def assemble_ms(self):
        v = C.c_double()
        A.check(self._lib.femgpu_get_assemble_ms(self._h, C.byref(v)), "assemble_ms")
        return v.value

    def precond_setup_ms(self):
        """Device time of the last solve's preconditioner set-up (0 when reused)."""
        v = C.c_double()
        A.check(self._lib.femgpu_get_precond_setup_ms(self._h, C.byref(v)), "precond_setup_ms")
        return v.value

    def time_spmv(self, precision=A.FP64, reps=20):
        v = C.c_double()
        A.check(self._lib.femgpu_time_spmv(self._h, int(precision), int(reps), C.byref(v)), "time_spmv")
        return v.value

    def apply(self, x, precision=A.FP64):
        """K x (caller's node order) with the product variant the solver uses for precision
        (FP64, or MIXED for the FP32 one: x rounded to float, the result widened)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.empty_like(x)
        A.check(self._lib.femgpu_apply(self._h, int(precision), x.ctypes.data, y.ctypes.data, x.size), "apply")
        return y

    def matrix_free_info(self):
        """Layout of the matrix-free product (Hexa and Tetra_10): element blocks, sub-colours,
        block-local node entries, shared nodes, their partial slots, nodes of no element, elements
        that are axis-aligned boxes."""
        info = (C.c_uint64 * 7)()
        A.check(self._lib.femgpu_matrix_free_info(self._h, info), "matrix_free_info")
        return dict(zip(("blocks", "colors", "local_entries", "shared_nodes", "partial_slots", "isolated_nodes",
                         "box_elements"), (int(v) for v in info)))

    def time_matrix_free(self, precision=A.FP64, reps=20):
        """(element-block kernel ms, whole product ms), averaged over reps launches."""
        b, p = C.c_double(), C.c_double()
        A.check(self._lib.femgpu_time_matrix_free(self._h, int(precision), int(reps), C.byref(b), C.byref(p)),
                "time_matrix_free")
        return b.value, p.value

    def time_iteration(self, precision=A.MIXED, iters=50):
        """(element-block kernel ms, vector update ms) of the two-kernel CG iteration (block
        Jacobi, matrix-free product), averaged over iters iterations (femgpu_time_iteration)."""
        b, u = C.c_double(), C.c_double()
        A.check(self._lib.femgpu_time_iteration(self._h, int(precision), int(iters), C.byref(b), C.byref(u)),
                "time_iteration")
        return b.value, u.value

    def autotune(self):
        """Pick the fastest block-SpMV variant (storage format, SELL slots per loop trip) for this
        matrix in FP64 and FP32; returns spmv_config()."""
        A.check(self._lib.femgpu_autotune(self._h, None), "autotune")
        return self.spmv_config()

    def spmv_config(self):
        """{"fp64": (format, unroll, loads), "fp32": (...)}: format SPMV_FULL, SPMV_SYM or
        SPMV_MATRIX_FREE (unroll and loads then those of the residual products), loads
        LOADS_PLAIN, LOADS_NT or LOADS_NT_COL16."""
        f = (C.c_int * 2)()
        u = (C.c_int * 2)()
        nt = (C.c_int * 2)()
        A.check(self._lib.femgpu_get_spmv_config(self._h, f, u, nt), "get_spmv_config")
        return {"fp64": (f[0], u[0], nt[0]), "fp32": (f[1], u[1], nt[1])}

    def set_assembly_rows(self, rows):
        """Element-loop variant: 0 = colour passes, ASSEMBLY_STAGED = colour passes through a
        block-contiguous copy of K, ASSEMBLY_GATHER = element blocks then per-position gathers,
        R = row groups of R rows (see femgpu.h)."""
        A.check(self._lib.femgpu_set_assembly_rows(self._h, int(rows)), "set_assembly_rows")

    def assembly_rows(self):
        """(current variant, largest admissible R)."""
        r, m = C.c_uint32(), C.c_uint32()
        A.check(self._lib.femgpu_get_assembly_rows(self._h, C.byref(r), C.byref(m)), "get_assembly_rows")
        return r.value, m.value

    def node_order(self):
        """(NODE_ORDER_NATURAL, _RCM or _AMD, perm): perm[i] = the caller's node stored at
        internal position i."""
        o = C.c_int()
        perm = np.zeros(self.n_nodes, dtype=np.uint32)
        A.check(self._lib.femgpu_get_node_order(self._h, C.byref(o), _ptr(perm)), "get_node_order")
        return o.value, perm

    def row_order(self):
        """(storage order of the block rows: ROW_ORDER_NATURAL or ROW_ORDER_BY_LENGTH, SELL
        positions including padding)."""
        o, n = C.c_int(), C.c_uint64()
        A.check(self._lib.femgpu_get_row_order(self._h, C.byref(o), C.byref(n)), "get_row_order")
        return o.value, n.value

    def set_spmv_config(self, fp64=(A.SPMV_FULL, 2, 0), fp32=(A.SPMV_FULL, 2, 0)):
        """Force the SpMV variant per precision: (format, unroll, loads)."""
        fp64, fp32 = tuple(fp64) + (0,) * (3 - len(fp64)), tuple(fp32) + (0,) * (3 - len(fp32))
        f = (C.c_int * 2)(int(fp64[0]), int(fp32[0]))
        u = (C.c_int * 2)(int(fp64[1]), int(fp32[1]))
        nt = (C.c_int * 2)(int(fp64[2]), int(fp32[2]))
        A.check(self._lib.femgpu_set_spmv_config(self._h, f, u, nt), "set_spmv_config")

    def set_graph_replay(self, enable=True):
        A.check(self._lib.femgpu_set_graph_replay(self._h, int(bool(enable))), "set_graph_replay")

    def call_after_solve(self):
        """call_after_solve() (formulation.h:2026-2050): per-element strain, stress and trace at
        xi = (1/2, 1/2, 1/2) from the current result, as the generated after_solve of
        formulations/formulation_elasticity_iso.py:64-85.  Returns (epsilon[n_elem, 6],
        sigma[n_elem, 6], tr_epsilon[n_elem]), Voigt xx, yy, zz, xy, xz, yz with tensor shear."""
        eps = np.zeros((self.n_elem, 6))
        sig = np.zeros((self.n_elem, 6))
        tr = np.zeros(self.n_elem)
        A.check(self._lib.femgpu_after_solve(self._h, _ptr(eps), _ptr(sig), _ptr(tr)), "after_solve")
        return eps, sig, tr

    def write_vtu(self, path):
        """Mesh and current result as a VTK unstructured grid in the reference writer's layout
        (write_mesh_vtk, mesh/write_mesh_vtk.h:827-831): dep; elastic_modulus, epsilon, sigma,
        tr_epsilon; poisson_ratio."""
        A.check(self._lib.femgpu_write_result_vtu(self._h, str(path).encode()), "write_result_vtu")

    def write_avs(self, path, display=None):
        """Mesh and current result as an AVS UCD ASCII file in the reference writer's layout
        (write_avs(mesh, name, display_fields, Ascii()), mesh/write_avs.h:389-530): node fields
        dep, pos; cell fields elastic_modulus, epsilon, sigma, tr_epsilon; global fields
        poisson_ratio, f_vol; display: the names to write (None: every field)."""
        nd, dp = A.display_names(display)
        A.check(self._lib.femgpu_write_result_avs(self._h, str(path).encode(), nd, dp), "write_result_avs")

    def after_solve_ms(self):
        """Device time of the last after-solve kernel."""
        v = C.c_double()
        A.check(self._lib.femgpu_get_after_solve_ms(self._h, C.byref(v)), "after_solve_ms")
        return v.value

    # -- differentiation -----------------------------------------------------------------------
    def adjoint(self, dJdu, rel_tol=1e-12):
        dJdu = np.ascontiguousarray(dJdu, dtype=np.float64)
        g = np.zeros(self.n_elem)
        o = self._opts(0.0, rel_tol)
        st = A.SolveStats()
        A.check(self._lib.femgpu_adjoint(self._h, _ptr(dJdu), _ptr(g), C.byref(o), C.byref(st)), "adjoint")
        self.last_stats = st
        return g

    def elem_grad(self, u, lam):
        u = np.ascontiguousarray(u, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        g = np.zeros(self.n_elem)
        A.check(self._lib.femgpu_elem_grad(self._h, _ptr(u), _ptr(lam), _ptr(g)), "elem_grad")
        return g


class BlockSystem(Formulation):
    """MatWithTinyBlocks<double,Sym<3>> built from a caller's full 3x3-block CSR and solved by the
    same PCG (containers/MatWithTinyBlocks.h:224-269,1101-1151)."""

    def __init__(self, row_ptr, col, vals, rhs):  # noqa: D401  (no mesh)
        lib = A.load()
        rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        cl = np.ascontiguousarray(col, dtype=np.uint32)
        vv = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1, 9)
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        h = C.c_void_p()
        A.check(lib.femgpu_create_from_bsr(rp.shape[0] - 1, _ptr(rp), _ptr(cl), _ptr(vv), _ptr(b), C.byref(h)),
                "create_from_bsr")
        self._h, self._lib = h, lib
        self.kind = None
        self.n_nodes = rp.shape[0] - 1
        self.n_elem = 0
        self.default_iterative_criterium = 1e-10
        self.precision, self.precond, self.max_iter = A.FP64, A.JACOBI, 100000
        self.last_stats, self.last_status = None, A.OK
        self.result = np.zeros(3 * self.n_nodes)

    def assemble(self, assemble_mat=True, assemble_vec=True):
        return None
