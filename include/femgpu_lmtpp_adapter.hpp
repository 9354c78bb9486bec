// femgpu_lmtpp_adapter.hpp -- the maintainer-side adapter that runs an LMT++ elasticity_iso solve
// on the MI355X path: it reads the mesh the reference already owns (m.node_list, m.elem_list,
// node.pos) and forwards the hot-path calls of Formulation (assemble, solve(crit),
// update_variables, call_after_solve) to the femgpu C ABI through femgpu_formulation.hpp, then
// writes node.dep and the elements' epsilon / sigma / tr_epsilon back.  It is included from an
// LMT++ translation unit, after the reference's mesh headers (it uses LMT::apply and the Mesh
// members of include/mesh/mesh.h), and needs -I<repo>/include and -lfemgpu.
#pragma once
#include <vector>

#include "femgpu_formulation.hpp"

namespace LMT {

// after-solve fields back into the elements, in element order
struct GpuElemScatter {
    const femgpu::FormulationGpu::AfterSolve *as;
    mutable size_t e;
    template<class TE> void operator()(TE &el) const {
        for (int i = 0; i < 6; ++i) {
            el.epsilon[0][i] = as->epsilon[6 * e + i];
            el.sigma[0][i] = as->sigma[6 * e + i];
        }
        el.tr_epsilon[0] = as->tr_epsilon[e];
        ++e;
    }
};

// connectivity in the reference local node order
template<class TM> struct GpuElemGather {
    TM *m;
    std::vector<uint32_t> *conn;
    template<class TE> void operator()(const TE &e) const {
        for (unsigned i = 0; i < TE::nb_nodes; ++i) conn->push_back(m->node_list.number(*e.node(i)));
    }
};

// Replaces f.assemble(); f.solve(crit); f.update_variables(); f.call_after_solve() for
// elasticity_iso on one element type (Tetra, Tetra_10 or Hexa); constraints in the penalty form of
// add_constraint (formulation.h:1028-1070), tractions on skin faces (formulation.h:584-594).
template<class TM>
bool solve_on_mi355x(TM &m, femgpu_elem_kind kind, double E, double nu, const double f_vol_times_rho[3],
                     const std::vector<uint32_t> &clamped_dofs, double penalty,
                     const std::vector<uint32_t> &traction_faces, const std::vector<double> &traction,
                     double crit) {
    std::vector<double> coords;
    for (unsigned i = 0; i < m.node_list.size(); ++i)
        for (int d = 0; d < 3; ++d) coords.push_back(m.node_list[i].pos[d]);
    std::vector<uint32_t> conn;
    GpuElemGather<TM> g;
    g.m = &m;
    g.conn = &conn;
    apply(m.elem_list, g);

    femgpu::FormulationGpu f(kind, coords, conn);       // allocate_matrices (formulation.h:338-412)
    f.options().precond = FEMGPU_BLOCK_IC0;             // the reference's block IC(0) PCG (or JACOBI, ...)
    f.set_material(E, nu, f_vol_times_rho);             // formulation_elasticity_iso.py:4-7
    f.set_traction(traction_faces, traction);           // f_surf on skin faces
    for (size_t k = 0; k < clamped_dofs.size(); ++k) f.add_constraint(clamped_dofs[k], 0.0, penalty);
    if (!f.solve(crit)) return false;                   // solve(crit) (formulation.h:1517-1568)
    const std::vector<double> &u = f.update_variables();
    for (unsigned i = 0; i < m.node_list.size(); ++i)   // update_variables (formulation.h:2571-2576)
        for (int d = 0; d < 3; ++d) m.node_list[i].dep[d] = u[3 * i + d];
    // call_after_solve (formulation.h:2026-2050): elem.epsilon / sigma / tr_epsilon
    const femgpu::FormulationGpu::AfterSolve as = f.call_after_solve();
    GpuElemScatter w;
    w.as = &as;
    w.e = 0;
    apply(m.elem_list, w);
    return true;
}

}  // namespace LMT
