// The C ABI's opaque handles (include/femgpu.h): a problem (one Problem and the host state of the
// reference's Formulation members around it), the time-stepping driver, a batch.  Shared by the
// entry-point files (capi.hip, noalloc.hip).
#pragma once
#include <string>
#include <vector>

#include "../../include/femgpu.h"
#include "constraint_expr.hpp"
#include "femgpu_internal.hpp"

namespace femgpu {
std::string &last_error();
}

// Formulation::nb_vectors = Carac::nb_vectors + 16 (formulation.h:112; 4 for the generated
// elasticity formulations, definition_probleme_elasticity_dep_static.h:587)
constexpr int NB_VECTORS = 20;

struct femgpu_problem {
    femgpu::Problem P;
    std::vector<std::vector<uint32_t>> sub_faces;  // per sub-problem, local node indices
    std::vector<std::vector<double>> sub_trac;
    bool faces_dirty = false;
    // time stepping (formulation.h:124-130, 424-445, 2624-2627): vectors[0] is P.d_u, vectors[k]
    // (k >= 1) d_hist[k], allocated when a shift first fills it (null: the zero vector)
    double time = 0.0;
    double time_steps[NB_VECTORS];
    double *d_hist[NB_VECTORS] = {};
    // constraints given as text (add_constraint(txt, penalty)): the parsed expression, its
    // penalty and its entry in P.cons, re-evaluated at each assembly when it depends on time
    struct TextConstraint { femgpu::ConstraintExpr expr; double penalty; size_t slot; uint64_t dof; };
    std::vector<TextConstraint> text_cons;
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
    femgpu_problem() {
        femgpu_default_solve_opts(&dopts);
        for (double &t : time_steps) t = 1e40;
    }
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
