// The C ABI's opaque handles (include/femgpu.h): a problem (one Problem and the host state of the
// reference's Formulation members around it), the time-stepping driver, a batch.  Shared by the
// entry-point files (capi.hip, noalloc.hip).
#pragma once
#include <string>
#include <vector>

#include "../../include/femgpu.h"
#include "femgpu_internal.hpp"

namespace femgpu {
std::string &last_error();
}

struct femgpu_problem {
    femgpu::Problem P;
    std::vector<std::vector<uint32_t>> sub_faces;  // per sub-problem, local node indices
    std::vector<std::vector<double>> sub_trac;
    bool faces_dirty = false;
    // time stepping (formulation.h:124-130, 424-433, 2624-2627)
    double time = 0.0, time_step = 1e40;
    double *d_prev = nullptr;     // vectors[1]: the solution before the last shift
    bool has_prev = false;
    // FormulationAncestor knobs (formulation_ancestor.h:31-35, 244)
    bool assume_constant_matrix = false;
    double nl_crit = 0.0;
    int max_nl = 50;
    int last_nl_solves = 0;       // solve_system calls of the last solve() / iteration()
    double *d_old = nullptr;      // previous iterate of the non-linear loop
    femgpu_solve_opts dopts{};    // default_iterative_criterium and the rest of the solve options
    // get_precond(): preconditioner kept for solve_system_using_precond, even across assemblies
    bool precond_pinned = false, precond_stale = false;
    int pinned_kind = -1;
    femgpu_problem() { femgpu_default_solve_opts(&dopts); }
};
// ProblemAncestor (problem_ancestor.h:21-140) over formulations the caller owns.
struct femgpu_stepper {
    std::vector<femgpu_problem *> forms;
    double time = 0.0;
    bool initialised = false;
};
struct femgpu_batch {
    femgpu_problem H;
};
