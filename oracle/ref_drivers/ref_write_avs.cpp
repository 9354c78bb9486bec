// Writes a result file with the reference's own ASCII AVS UCD writer (mesh/write_avs.h:389-530,
// write_avs(mesh, name, display_fields, Ascii())) for a make_rect Hexa mesh of the
// elasticity_dep_static problem (node positions and displacements from the caller, strain and
// stress from the generated after_solve of std_formulations/definition_probleme_elasticity_dep_static.h,
// E and nu from the caller, every other field at its default), or for a make_rect Tetra mesh of
// the field-less MeshCaracStd<3,3,0,double,1> (mesh/meshcaracstd.h; the elasticity carac holds
// Hexa only) or a Tetra_10 one of degree 2, whose files are the geometry alone.  Builds against the LMT++ headers where they lie;
// this file only drives them.
// usage: ref_write_avs hexa|tetra|tetra10 nx ny nz in.bin out.inp fields.bin field...   ("all": every field)
// in.bin    : uint32 n_nodes, double coords[n_nodes*3], double u[n_nodes*3], double E, double nu
// fields.bin: uint32 n_elem, then per element double eps[6], sigma[6], tr_epsilon
#include "std_formulations/definition_probleme_elasticity_dep_static.h"
#include "mesh/meshcaracstd.h"
#include "mesh/make_rect.h"
#include "mesh/write_avs.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

// A field-less make_rect mesh with the caller's node positions.
template<class TT, class NE>
int geometry(NE ne, const TM::Pvec &nb, const char *in, const char *out, const Vec<std::string> &display) {
    TT t;
    make_rect(t, ne, typename TT::Pvec(0, 0, 0), typename TT::Pvec(1, 1, 1), typename TT::Pvec(nb[0], nb[1], nb[2]));
    FILE *fi = fopen(in, "rb");
    unsigned nn = 0;
    if (!fi || fread(&nn, 4, 1, fi) != 1 || nn != t.node_list.size()) return 3;
    std::vector<double> c(3 * (size_t)nn);
    if (fread(c.data(), 8, c.size(), fi) != c.size()) return 4;
    fclose(fi);
    for (unsigned i = 0; i < nn; ++i)
        for (int d = 0; d < 3; ++d) t.node_list[i].pos[d] = c[3 * i + d];
    write_avs(t, std::string(out), display, Ascii());
    printf("%u nodes %u elements\n", nn, (unsigned)t.elem_list.size());
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 9) { fprintf(stderr, "usage: ref_write_avs hexa|tetra nx ny nz in.bin out.inp fields.bin field...\n"); return 1; }
    const TM::Pvec lo(0, 0, 0), hi(1, 1, 1), nb(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
    Vec<std::string> display;
    for (int i = 8; i < argc; ++i) display.push_back(argv[i]);
    if (!strcmp(argv[1], "tetra"))
        return geometry<Mesh<MeshCaracStd<3, 3, 0, double, 1> > >(Tetra(), nb, argv[5], argv[6], display);
    if (!strcmp(argv[1], "tetra10"))
        return geometry<Mesh<MeshCaracStd<3, 3, 0, double, 2> > >(Tetra_10(), nb, argv[5], argv[6], display);
    if (strcmp(argv[1], "hexa")) return 1;
    TM m;
    make_rect(m, Hexa(), lo, hi, nb);
    FILE *fi = fopen(argv[5], "rb");
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
    write_avs(m, std::string(argv[6]), display, Ascii());
    unsigned ne = m.elem_list.size();
    FILE *fo = fopen(argv[7], "wb");
    if (!fo) return 5;
    fwrite(&ne, 4, 1, fo);
    fwrite(out.data(), 8, out.size(), fo);
    fclose(fo);
    printf("%u nodes %u elements\n", nn, ne);
    return 0;
}
