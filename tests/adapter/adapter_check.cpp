// Instantiates include/femgpu_lmtpp_adapter.hpp against the reference's own mesh types (the
// generated elasticity_dep_static mesh, which lists Hexa and carries dep, epsilon, sigma and
// tr_epsilon): compiled, not run, by tests/test_integration_adapter.py where the LMT++ sources are.
#include "std_formulations/definition_probleme_elasticity_dep_static.h"
#include "mesh/make_rect.h"
#include "femgpu_lmtpp_adapter.hpp"

using namespace LMT;
typedef Mesh<Mesh_carac_pb_elasticity_dep_static<double, 3> > TM;

bool adapter_check(TM &m, const std::vector<uint32_t> &dofs, const std::vector<uint32_t> &faces,
                   const std::vector<double> &trac) {
    make_rect(m, Hexa(), TM::Pvec(0, 0, 0), TM::Pvec(1, 1, 1), TM::Pvec(5, 5, 5));
    const double bf[3] = {0, 0, 0};
    return solve_on_mi355x(m, FEMGPU_HEXA, 210000.0, 0.3, bf, dofs, 1e5, faces, trac, 1e-10);
}
