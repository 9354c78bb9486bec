// Internal structures shared by the C ABI (capi.hip) and the kernels (assemble.hip, solve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "host_mesh.hpp"

namespace femgpu {

// Device-resident scalars of one CG solve.  Written by the last workgroup of each reducing
// kernel, read by the next kernel and, between chunks of iterations, by the host.
struct Scalars {
    double rho, rho_prev, pq, alpha, beta, rr, rmax;
    double bb, stop2, crit;       // stop test: rr <= stop2 (REL_L2) or rmax <= crit (ABS_MAXNORM)
    double rr_ref, delta2;        // mixed: refresh when rr <= delta2 * rr_ref
    double max_diag;              // assembly: max |diag K|
    int iter, max_iter, done, pause, status, crit_kind, mixed, pad;
    unsigned counter[8];          // last-workgroup counters, one per reducing kernel
};

enum { CTR_SPMV = 0, CTR_UPD = 1, CTR_RES = 2, CTR_REFRESH = 3, CTR_MAXDIAG = 4, CTR_INIT = 5 };

constexpr int WG = 256;   // threads per workgroup for all vector / SpMV kernels

struct SpmvConfig {
    int group = 32;       // lanes per block row
    int nwg = 2048;       // workgroups (multiple of 8: XCD-contiguous row slabs)
};

struct Problem {
    int kind = 0, npe = 0, device = 0;
    uint32_t n_nodes = 0, n_elem = 0;
    uint64_t nnzb = 0, ps = 0;  // stored blocks, plane stride (elements)
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};

    Pattern pat;
    Coloring ecol, fcol;
    std::vector<double> coords_h;
    std::vector<uint32_t> conn_h;

    // device: mesh, pattern, scatter map
    double *d_coords = nullptr, *d_E = nullptr;
    uint32_t *d_conn = nullptr, *d_rp = nullptr, *d_col = nullptr, *d_diag = nullptr, *d_slot = nullptr, *d_order = nullptr;
    // device: system
    double *d_K = nullptr, *d_F = nullptr, *d_u = nullptr;
    float *d_K32 = nullptr;
    bool k32_valid = false, assembled_mat = false, assembled_vec = false;
    // skin loads
    uint32_t n_faces = 0;
    uint32_t *d_faces = nullptr, *d_forder = nullptr;
    double *d_trac = nullptr;
    // constraints, grouped by dof in insertion order
    struct Cons { uint32_t dof; double val, pen; };
    std::vector<Cons> cons;
    bool cons_dirty = true;
    uint32_t n_cdof = 0;
    uint32_t *d_cdof = nullptr, *d_cptr = nullptr;
    double *d_cval = nullptr, *d_cpen = nullptr;
    // material
    double nu = 0.3, bf[3] = {0, 0, 0};
    bool material_set = false;
    // solver work vectors
    double *d_r = nullptr, *d_p = nullptr, *d_q = nullptr, *d_Minv = nullptr, *d_b = nullptr, *d_x = nullptr;
    double *d_aux = nullptr;  // adjoint state / scratch vector
    float *d_r32 = nullptr, *d_p32 = nullptr, *d_q32 = nullptr, *d_x32 = nullptr, *d_Minv32 = nullptr;
    double *d_partials = nullptr;
    Scalars *d_sc = nullptr, *h_sc = nullptr;
    int minv_kind = -1;  // preconditioner currently stored in d_Minv
    SpmvConfig spmv;
    uint32_t *d_wgrows = nullptr;
    int wgrows_nwg = 0;
    double last_asm_ms = 0;
    double max_diag = 0;  // max|diag K| of the last matrix assembly, before constraints
    bool from_bsr = false;
};

// ---- launchers (assemble.hip) ----
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
hipError_t launch_assemble(Problem &P, int assemble_mat, int assemble_vec, int *d_flags);
hipError_t launch_constraints(Problem &P, int assemble_mat, int assemble_vec);
hipError_t launch_to_fp32(Problem &P);
hipError_t launch_elem_grad(Problem &P, const double *d_u, const double *d_lam, double *d_g);

// ---- solver (solve.hip) ----
struct SolveOpts { int precision, precond, crit_kind; double tol; int max_iter; double delta; int chunk; };
struct SolveOut { int status, iters, refreshes; double rel_res, max_abs_res, ms; };
int solve_pcg(Problem &P, const double *d_rhs, double *d_x, const SolveOpts &o, SolveOut &out);
hipError_t ensure_wgrows(Problem &P);
double time_spmv(Problem &P, int precision, int reps);
hipError_t spmv_apply(Problem &P, const double *d_x, double *d_y);

std::string &last_error();

}  // namespace femgpu
