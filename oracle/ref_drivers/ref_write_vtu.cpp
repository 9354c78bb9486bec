// Writes a result file with the reference's own VTU writer (mesh/write_mesh_vtk.h:434-656, the
// binary/appended form write_mesh_vtk(filename, m) uses, :827-831) for a make_rect Hexa mesh of
// the elasticity_dep_static problem: node positions and displacements from the caller, strain and
// stress from the generated after_solve (std_formulations/definition_probleme_elasticity_dep_static.h:721-789),
// E and nu from the caller, every other field at its default.  Builds against the LMT++ headers
// where they lie; this file only drives them.
// usage: ref_write_vtu nx ny nz in.bin out.vtu fields.bin
// in.bin    : uint32 n_nodes, double coords[n_nodes*3], double u[n_nodes*3], double E, double nu
// fields.bin: uint32 n_elem, then per element double eps[6], sigma[6], tr_epsilon (the values
//             the writer was given)
#include "std_formulations/definition_probleme_elasticity_dep_static.h"
#include "mesh/make_rect.h"
#include "mesh/write_mesh_vtk.h"
#include <cstdio>
#include <cstdlib>
#include <vector>
using namespace LMT;

typedef Mesh<Mesh_carac_pb_elasticity_dep_static<double, 3> > TM;

struct FormulationView { TM *m; };

struct RunAfterSolve {
    TM *m; std::vector<std::vector<double> > *vec; std::vector<double> *out;
    template<class TE> void operator()(TE &e) const {
        unsigned ind[TE::nb_nodes];
        for (unsigned i = 0; i < TE::nb_nodes; ++i) ind[i] = 3 * m->node_list.number(*e.node(i));
        FormulationView f; f.m = m;
        CaracFormulationForElement<elasticity_dep_static, TE, TM, double>::after_solve(e, f, *vec, ind);
        for (int i = 0; i < 6; ++i) out->push_back(e.epsilon[0][i]);
        for (int i = 0; i < 6; ++i) out->push_back(e.sigma[0][i]);
        out->push_back(e.tr_epsilon[0]);
    }
};

int main(int argc, char **argv) {
    if (argc != 7) { fprintf(stderr, "usage: ref_write_vtu nx ny nz in.bin out.vtu fields.bin\n"); return 1; }
    TM m;
    make_rect(m, Hexa(), TM::Pvec(0, 0, 0), TM::Pvec(1, 1, 1), TM::Pvec(atoi(argv[1]), atoi(argv[2]), atoi(argv[3])));
    FILE *fi = fopen(argv[4], "rb");
    if (!fi) return 2;
    unsigned nn = 0;
    if (fread(&nn, 4, 1, fi) != 1 || nn != m.node_list.size()) return 3;
    std::vector<double> c(3 * (size_t)nn), u(3 * (size_t)nn);
    double E = 0, nu = 0;
    if (fread(c.data(), 8, c.size(), fi) != c.size() || fread(u.data(), 8, u.size(), fi) != u.size() ||
        fread(&E, 8, 1, fi) != 1 || fread(&nu, 8, 1, fi) != 1) return 4;
    fclose(fi);
    for (unsigned i = 0; i < nn; ++i)
        for (int d = 0; d < 3; ++d) {
            m.node_list[i].pos[d] = c[3 * i + d];
            m.node_list[i].dep[d] = u[3 * i + d];
        }
    m.elastic_modulus = E;
    m.poisson_ratio = nu;
    std::vector<std::vector<double> > vec(1, u);
    std::vector<double> out;
    RunAfterSolve r; r.m = &m; r.vec = &vec; r.out = &out;
    apply(m.elem_list, r);
    write_mesh_vtk(std::string(argv[5]), m);
    unsigned ne = m.elem_list.size();
    FILE *fo = fopen(argv[6], "wb");
    if (!fo) return 5;
    fwrite(&ne, 4, 1, fo);
    fwrite(out.data(), 8, out.size(), fo);
    fclose(fo);
    printf("%u nodes %u elements\n", nn, ne);
    return 0;
}
