"""PyTorch binding of the adjoint path: u = ElasticSolve(E) with dJ/dE from the adjoint kernels.

Forward: the per-element modulus (CUDA float64 tensor) is copied into the problem, K and F are
assembled and the system is solved on the GPU.  Backward: one adjoint solve K lambda = dJ/du and
the element kernel dJ/dE_e = -lambda_e^T (dK_e/dE_e) u_e (femgpu_adjoint_device), device pointers
only.  The reference identifies parameters with forward sensitivities, one solve per parameter
(FormulationFit.h:19-85); the adjoint needs one solve for all of them.  The penalty scale
max|diag K| is held fixed in the adjoint.

Import order: torch (this module imports it) must be loaded before libfemgpu.so, so that the
library binds to torch's HIP runtime; a process holds one HIP runtime.
"""
import ctypes as C

import torch

from . import _capi as A


def _opts(lib, rel_tol, precision=A.FP64):
    o = A.SolveOpts()
    lib.femgpu_default_solve_opts(C.byref(o))
    o.crit_kind, o.tol, o.precision = A.REL_L2, float(rel_tol), int(precision)
    return o


class ElasticSolve(torch.autograd.Function):
    """u = K(E)^-1 F for a Formulation whose material, tractions and constraints are set."""

    @staticmethod
    def forward(ctx, E, problem, rel_tol=1e-12, adj_rel_tol=1e-12, precision=A.FP64):
        if not E.is_cuda or E.dtype != torch.float64 or E.numel() != problem.n_elem:
            raise ValueError("E must be a CUDA float64 tensor with one value per element")
        lib = problem._lib
        E = E.contiguous()
        torch.cuda.current_stream(E.device).synchronize()
        A.check(lib.femgpu_set_elastic_modulus_device(problem.handle, C.c_void_p(E.data_ptr())), "set_E")
        problem.assemble(True, True)
        o = _opts(lib, rel_tol, precision)
        st = A.SolveStats()
        rc = lib.femgpu_solve_system(problem.handle, C.byref(o), C.byref(st))
        A.check(rc, "solve_system")
        u = torch.empty(3 * problem.n_nodes, dtype=torch.float64, device=E.device)
        A.check(lib.femgpu_get_result_device(problem.handle, C.c_void_p(u.data_ptr())), "get_result")
        problem.forward_stats = st
        problem._adjoint_token = getattr(problem, "_adjoint_token", 0) + 1
        ctx.problem, ctx.token, ctx.rel_tol, ctx.adj_rel_tol = problem, problem._adjoint_token, rel_tol, adj_rel_tol
        ctx.save_for_backward(E)
        return u

    @staticmethod
    def backward(ctx, grad_u):
        problem = ctx.problem
        (E,) = ctx.saved_tensors
        lib = problem._lib
        g = grad_u.contiguous().to(torch.float64)
        torch.cuda.current_stream(g.device).synchronize()
        if problem._adjoint_token != ctx.token:
            # another forward ran on this problem since: restore K(E) and u before the adjoint
            A.check(lib.femgpu_set_elastic_modulus_device(problem.handle, C.c_void_p(E.data_ptr())), "set_E")
            problem.assemble(True, True)
            o = _opts(lib, ctx.rel_tol)
            A.check(lib.femgpu_solve_system(problem.handle, C.byref(o), C.byref(A.SolveStats())), "solve_system")
            problem._adjoint_token = ctx.token
        out = torch.empty(problem.n_elem, dtype=torch.float64, device=g.device)
        o = _opts(lib, ctx.adj_rel_tol)
        st = A.SolveStats()
        A.check(lib.femgpu_adjoint_device(problem.handle, C.c_void_p(g.data_ptr()), C.c_void_p(out.data_ptr()),
                                          C.byref(o), C.byref(st)), "adjoint")
        problem.adjoint_stats = st
        return out, None, None, None, None


def elastic_solve(E, problem, rel_tol=1e-12, adj_rel_tol=1e-12, precision=A.FP64):
    return ElasticSolve.apply(E, problem, rel_tol, adj_rel_tol, precision)


def identify_modulus(problem, E_init, obs_dofs, u_meas, steps=50, lr=0.05, rel_tol=1e-10, callback=None):
    """Recalage-style identification of the elementary modulus field: Adam on log E for
    J = 1/2 ||u[obs] - u_meas||^2 (the objective of the identification configuration).
    Returns (E, history of J)."""
    dev = torch.device("cuda")
    logE = torch.log(torch.as_tensor(E_init, dtype=torch.float64, device=dev)).clone().requires_grad_(True)
    obs = torch.as_tensor(obs_dofs, dtype=torch.long, device=dev)
    meas = torch.as_tensor(u_meas, dtype=torch.float64, device=dev)
    scale = 1.0 / float(torch.linalg.vector_norm(meas)) ** 2
    opt = torch.optim.Adam([logE], lr=lr)
    hist = []
    for it in range(steps):
        opt.zero_grad()
        u = elastic_solve(torch.exp(logE), problem, rel_tol, rel_tol)
        J = 0.5 * torch.sum((u[obs] - meas) ** 2)
        (J * scale).backward()
        opt.step()
        hist.append(float(J))
        if callback:
            callback(it, float(J))
    return torch.exp(logE).detach(), hist
