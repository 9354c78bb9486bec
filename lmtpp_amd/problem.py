"""Time stepping over several formulations: the mirror of ProblemAncestor
(include/formulation/problem_ancestor.h:21-140) on the native stepper of libfemgpu.so
(femgpu_stepper_*).

``iteration()`` follows problem_ancestor.h:41-129.  The first call shifts and assembles every
formulation and builds the preconditioner of those with ``assume_constant_matrix``.  Later calls
update only the formulations whose time equals the problem's time: they are shifted and their
right-hand side (and K, unless constant) reassembled.  Constant-matrix formulations then solve
with the kept preconditioner, warm-started from their previous solution.  The others run the
non-linear loop: 2-norm of the change, reassembly between solves, and an error after
``max_non_linear_iteration`` + 1 solves, where the reference throws.  The problem's time becomes
the smallest formulation time.
"""
import ctypes as C

from . import _capi as A


class Problem:
    def __init__(self, formulations, rel_tol=None):
        """formulations: Formulation objects (kept alive and owned by the caller).  Each solve uses
        that formulation's default_iterative_criterium (the reference's max-norm test), or the
        relative L2 test rel_tol when given, with its precision, precond and max_iter."""
        self._lib = A.load()
        self.formulations = list(formulations)
        self.rel_tol = rel_tol
        n = len(self.formulations)
        arr = (C.c_void_p * max(n, 1))(*[f.handle for f in self.formulations])
        h = C.c_void_p()
        A.check(self._lib.femgpu_stepper_create(n, arr, C.byref(h)), "stepper_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.femgpu_stepper_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def time(self):
        v = C.c_double()
        A.check(self._lib.femgpu_stepper_get_time(self._h, C.byref(v)), "stepper_get_time")
        return v.value

    def iteration(self):
        """One step of every formulation due at the problem's time; returns the list of updated
        flags.  Raises FemGpuError when a solve fails or the non-linear loop does not settle."""
        n = len(self.formulations)
        for f in self.formulations:
            f._push_defaults(self.rel_tol)
        stats = (A.SolveStats * max(n, 1))()
        upd = (C.c_int * max(n, 1))()
        rc = self._lib.femgpu_stepper_iteration(self._h, stats, upd)
        for i, f in enumerate(self.formulations):
            if upd[i]:
                f.last_stats, f.last_status = stats[i], stats[i].status
        A.check(rc, "iteration")
        for i, f in enumerate(self.formulations):
            if upd[i]:
                f.update_variables()
        return [bool(upd[i]) for i in range(n)]

    def call_after_solve(self):
        """call_after_solve() of every formulation (problem_ancestor.h:131-135)."""
        return [f.call_after_solve() for f in self.formulations]
