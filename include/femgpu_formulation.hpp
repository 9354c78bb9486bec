// femgpu_formulation.hpp -- C++ host facade over the femgpu C ABI with the call surface of
// LMT++'s FormulationAncestor<double> (include/formulation/formulation_ancestor.h) for the
// elasticity_iso formulation:
//
//     allocate_matrices()            formulation_ancestor.h:51   (pattern built at construction)
//     shift(nb)                      formulation_ancestor.h:57; formulation.h:424-433
//     assemble(mat, vec)             formulation_ancestor.h:63
//     solve_system(crit) -> bool     formulation_ancestor.h:64
//     solve(crit) -> bool            formulation_ancestor.h:49
//     update_variables()             formulation_ancestor.h:179
//     add_constraint(dof, penalty)   formulation_ancestor.h:197  (penalty form)
//     add_constraint(txt, penalty)   formulation_ancestor.h:197; constraint.h:46-93 (one unknown)
//     unshift(nb)                    formulation_ancestor.h:58; formulation.h:435-445
//     get_sollicitation()            formulation_ancestor.h:236 (set_sollicitation for writes)
//     get_result(num)                formulation_ancestor.h:237
//     get_indice_noda(node)          formulation_ancestor.h:188
//     get_nodal_forces()             formulation_ancestor.h:65; formulation.h:447-458
//     call_after_solve()             formulation.h:2026-2050 (elem.epsilon / sigma / tr_epsilon)
//     default_iterative_criterium    formulation_ancestor.h:244-257
//     assume_constant_matrix, non_linear_iterative_criterium, max_non_linear_iteration
//                                    formulation_ancestor.h:31-35,244
//     get_precond(), solve_system_using_precond(crit)   formulation.h:1655-1656
//     set_initial_time_step / get_next_time_step / get_time / set_time   formulation.h:2624-2627
//
// and ProblemGpu, the ProblemAncestor (problem_ancestor.h:21-140) time-stepping driver:
//     iteration(), call_after_solve(), time
//
// The preconditioner is options().precond: FEMGPU_JACOBI, FEMGPU_BLOCK_JACOBI, FEMGPU_BLOCK_IC0
// (the reference's block incomplete Cholesky, MatWithTinyBlocks.h:426-709) or
// FEMGPU_BLOCK_IC0_COLOURED.
//
// Like the reference, solve()/solve_system() return false when the system could not be solved
// (formulation.h:1394,1530-1531); the femgpu status of the last call is in last_status().
// Bad arguments and HIP failures throw femgpu::Error.
#pragma once
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "femgpu.h"

namespace femgpu {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(int rc, const char *what) {
    if (rc != FEMGPU_OK) throw Error(rc, std::string(what) + ": " + femgpu_last_error());
}

class FormulationGpu {
public:
    // Mesh in the reference layout: node coordinates (3 per node) and element connectivity in
    // the reference local node order (make_rect output or any mesh reader).
    FormulationGpu(femgpu_elem_kind kind, const std::vector<double> &coords, const std::vector<uint32_t> &conn)
        : n_nodes_(uint32_t(coords.size() / 3)) {
        const int npe = kind == FEMGPU_TETRA ? 4 : kind == FEMGPU_TETRA10 ? 10 : 8;
        n_elem_ = uint32_t(conn.size() / npe);
        femgpu_mesh m{kind, n_nodes_, coords.data(), n_elem_, conn.data()};
        check(femgpu_create(&m, &h_), "femgpu_create");
        femgpu_default_solve_opts(&opts_);
    }
    FormulationGpu(const FormulationGpu &) = delete;
    FormulationGpu &operator=(const FormulationGpu &) = delete;
    ~FormulationGpu() { femgpu_destroy(h_); }

    // public knobs
    double default_iterative_criterium = 1e-10;
    femgpu_solve_opts &options() { return opts_; }
    void set_assume_constant_matrix(bool v) { check(femgpu_set_assume_constant_matrix(h_, v), "assume_constant_matrix"); }
    // non_linear_iterative_criterium (0: linear, one solve) and max_non_linear_iteration
    void set_non_linear(double crit, int max_iter = 50) { check(femgpu_set_nonlinear(h_, crit, max_iter), "set_nonlinear"); }
    int nonlinear_solves() const {
        int n = 0;
        check(femgpu_get_nonlinear_solves(h_, &n), "nonlinear_solves");
        return n;
    }

    // material: elastic_modulus (global or per element), poisson_ratio, density * f_vol
    void set_material(double E, double nu, const double body_force[3] = nullptr) {
        femgpu_material mat{nu, E, nullptr, {0, 0, 0}};
        if (body_force) for (int d = 0; d < 3; ++d) mat.body_force[d] = body_force[d];
        check(femgpu_set_material(h_, &mat), "set_material");
    }
    void set_material(const std::vector<double> &E_elem, double nu, const double body_force[3] = nullptr) {
        femgpu_material mat{nu, 0.0, E_elem.data(), {0, 0, 0}};
        if (body_force) for (int d = 0; d < 3; ++d) mat.body_force[d] = body_force[d];
        check(femgpu_set_material(h_, &mat), "set_material");
    }
    void set_traction(const std::vector<uint32_t> &faces, const std::vector<double> &trac) {
        check(femgpu_set_traction(h_, uint32_t(trac.size() / 3), faces.data(), trac.data()), "set_traction");
    }
    void add_constraint(uint32_t dof, double value, double penalty) {
        check(femgpu_add_constraint(h_, dof, value, penalty), "add_constraint");
    }
    // e.g. add_constraint("node[12].dep[2] + 0.001*time", 1e5); returns the constraint's number
    unsigned add_constraint(const std::string &txt, double penalty_value) {
        uint32_t num = 0;
        check(femgpu_add_constraint_text(h_, txt.c_str(), penalty_value, &num), "add_constraint");
        return num;
    }

    void allocate_matrices(bool = true, bool = true) {}
    // time += time step; the previous solution moves to vectors[1]; vectors[0] stays the next
    // initial guess
    void shift(int nb = 1) { check(femgpu_shift(h_, nb), "shift"); }
    void unshift(int nb = 1) { check(femgpu_unshift(h_, nb), "unshift"); }
    void set_initial_time_step(double ts) { check(femgpu_set_initial_time_step(h_, ts), "set_initial_time_step"); }
    double get_next_time_step() const {
        double v = 0;
        check(femgpu_get_next_time_step(h_, &v), "get_next_time_step");
        return v;
    }
    double get_time() const {
        double v = 0;
        check(femgpu_get_time(h_, &v), "get_time");
        return v;
    }
    void set_time(double t) { check(femgpu_set_time(h_, t), "set_time"); }
    void assemble(bool assemble_mat = true, bool assemble_vec = true) {
        check(femgpu_assemble(h_, assemble_mat, assemble_vec), "assemble");
    }
    // crit > 0: stop when max_i |r_i| <= crit (MatWithTinyBlocks.h:1115,1136); crit == 0 uses
    // default_iterative_criterium.  rel_tol > 0 selects ||r|| <= rel_tol ||b|| instead.
    bool solve_system(double crit = 0, double rel_tol = 0) {
        const femgpu_solve_opts o = solve_opts(crit, rel_tol);
        last_status_ = femgpu_solve_system(h_, &o, &stats_);
        if (last_status_ == FEMGPU_BAD_INPUT || last_status_ == FEMGPU_HIP_ERROR || last_status_ == FEMGPU_NOT_READY)
            check(last_status_, "solve_system");
        return last_status_ == FEMGPU_OK;
    }
    // shift, then assemble + solve_system (repeated by the non-linear loop); throws when the
    // loop reaches max_non_linear_iteration, as the reference throws SolveException
    bool solve(double crit = 0, double rel_tol = 0) {
        const femgpu_solve_opts o = solve_opts(crit, rel_tol);
        last_status_ = femgpu_solve(h_, &o, &stats_);
        if (last_status_ == FEMGPU_BAD_INPUT || last_status_ == FEMGPU_HIP_ERROR || last_status_ == FEMGPU_NOT_READY ||
            last_status_ == FEMGPU_NEG_JACOBIAN || (last_status_ == FEMGPU_MAX_ITER && stats_.status == FEMGPU_OK))
            check(last_status_, "solve");
        if (last_status_ == FEMGPU_OK) update_variables();
        return last_status_ == FEMGPU_OK;
    }
    // preconditioner of the current K (options().precond), kept through later assemblies
    void get_precond(double crit = 0, double rel_tol = 0) {
        const femgpu_solve_opts o = solve_opts(crit, rel_tol);
        check(femgpu_get_precond(h_, &o), "get_precond");
    }
    bool solve_system_using_precond(double crit = 0, double rel_tol = 0) {
        const femgpu_solve_opts o = solve_opts(crit, rel_tol);
        last_status_ = femgpu_solve_system_using_precond(h_, &o, &stats_);
        if (last_status_ == FEMGPU_BAD_INPUT || last_status_ == FEMGPU_HIP_ERROR || last_status_ == FEMGPU_NOT_READY)
            check(last_status_, "solve_system_using_precond");
        return last_status_ == FEMGPU_OK;
    }
    // options of a solve: crit > 0 max-norm test, else default_iterative_criterium; rel_tol > 0
    // the relative L2 test
    femgpu_solve_opts solve_opts(double crit = 0, double rel_tol = 0) const {
        femgpu_solve_opts o = opts_;
        if (rel_tol > 0) { o.crit_kind = FEMGPU_REL_L2; o.tol = rel_tol; }
        else { o.crit_kind = FEMGPU_ABS_MAXNORM; o.tol = crit > 0 ? crit : default_iterative_criterium; }
        return o;
    }
    // node.dep[d] = vectors[0][3 i + d] (formulation.h:2571-2576)
    const std::vector<double> &update_variables() {
        dep_.resize(3 * size_t(n_nodes_));
        check(femgpu_get_result(h_, dep_.data(), dep_.size()), "get_result");
        return dep_;
    }
    std::vector<double> get_result(unsigned num = 0) const {
        std::vector<double> u(3 * size_t(n_nodes_));
        check(femgpu_get_result_history(h_, num, u.data(), u.size()), "get_result");
        return u;
    }
    std::vector<double> get_sollicitation() const {
        std::vector<double> F(3 * size_t(n_nodes_));
        check(femgpu_get_sollicitation(h_, F.data(), F.size()), "get_sollicitation");
        return F;
    }
    void set_sollicitation(const std::vector<double> &F) {
        check(femgpu_set_sollicitation(h_, F.data(), F.size()), "set_sollicitation");
    }
    unsigned get_indice_noda(unsigned num_node) const {
        uint32_t i = 0;
        check(femgpu_get_indice_noda(h_, num_node, &i), "get_indice_noda");
        return i;
    }
    std::vector<double> get_nodal_forces() const {
        std::vector<double> F(3 * size_t(n_nodes_));
        check(femgpu_get_nodal_forces(h_, F.data(), F.size()), "get_nodal_forces");
        return F;
    }
    // Per-element strain and stress at xi = (1/2,1/2,1/2) from the current result: 6 values per
    // element each (Voigt xx, yy, zz, xy, xz, yz, tensor shear) and the trace.
    struct AfterSolve { std::vector<double> epsilon, sigma, tr_epsilon; };
    AfterSolve call_after_solve() const {
        AfterSolve r;
        r.epsilon.resize(6 * size_t(n_elem_));
        r.sigma.resize(6 * size_t(n_elem_));
        r.tr_epsilon.resize(n_elem_);
        check(femgpu_after_solve(h_, r.epsilon.data(), r.sigma.data(), r.tr_epsilon.data()), "after_solve");
        return r;
    }
    int last_status() const { return last_status_; }
    const femgpu_solve_stats &last_stats() const { return stats_; }
    void set_last(int status, const femgpu_solve_stats &st) { last_status_ = status; stats_ = st; }
    femgpu_problem *handle() const { return h_; }
    uint32_t nb_nodes() const { return n_nodes_; }
    uint32_t nb_elems() const { return n_elem_; }

private:
    femgpu_problem *h_ = nullptr;
    uint32_t n_nodes_ = 0, n_elem_ = 0;
    femgpu_solve_opts opts_{};
    femgpu_solve_stats stats_{};
    int last_status_ = FEMGPU_OK;
    std::vector<double> dep_;
};

// ProblemAncestor (problem_ancestor.h:21-140) over formulations the caller owns; each solve uses
// that formulation's solve_opts() (rel_tol > 0 selects the relative L2 test for all).
class ProblemGpu {
public:
    explicit ProblemGpu(std::vector<FormulationGpu *> forms, double rel_tol = 0) : forms_(std::move(forms)), rel_tol_(rel_tol) {
        std::vector<femgpu_problem *> h;
        for (FormulationGpu *f : forms_) h.push_back(f->handle());
        check(femgpu_stepper_create(uint32_t(h.size()), h.data(), &s_), "stepper_create");
    }
    ProblemGpu(const ProblemGpu &) = delete;
    ProblemGpu &operator=(const ProblemGpu &) = delete;
    ~ProblemGpu() { femgpu_stepper_destroy(s_); }

    // problem_ancestor.h:41-129; returns which formulations were updated
    std::vector<int> iteration() {
        const size_t n = forms_.size();
        for (FormulationGpu *f : forms_) {
            const femgpu_solve_opts o = f->solve_opts(0, rel_tol_);
            check(femgpu_set_default_solve_opts(f->handle(), &o), "set_default_solve_opts");
        }
        std::vector<femgpu_solve_stats> st(n);
        std::vector<int> upd(n, 0);
        const int rc = femgpu_stepper_iteration(s_, st.data(), upd.data());
        for (size_t i = 0; i < n; ++i)
            if (upd[i]) forms_[i]->set_last(st[i].status, st[i]);
        check(rc, "iteration");
        for (size_t i = 0; i < n; ++i)
            if (upd[i]) forms_[i]->update_variables();
        return upd;
    }
    double time() const {
        double t = 0;
        check(femgpu_stepper_get_time(s_, &t), "stepper_get_time");
        return t;
    }
    std::vector<FormulationGpu::AfterSolve> call_after_solve() const {
        std::vector<FormulationGpu::AfterSolve> r;
        for (FormulationGpu *f : forms_) r.push_back(f->call_after_solve());
        return r;
    }

private:
    std::vector<FormulationGpu *> forms_;
    double rel_tol_ = 0;
    femgpu_stepper *s_ = nullptr;
};

}  // namespace femgpu
