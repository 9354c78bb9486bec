// Dumps the reference make_rect() meshes (node coordinates and connectivity)
// for golden fixtures.  Builds against the LMT++ headers where they lie; this
// file only drives them (mesh/make_rect.h, util/rectilinear_iterator.h).
// usage: ref_mesh_dump <tetra|tetra10|hexa> nx ny nz out.bin
// out.bin: uint32 n_nodes, uint32 n_elem, uint32 npe, double[n_nodes*3], uint32[n_elem*npe]
#include "mesh/mesh.h"
#include "mesh/meshcaracstd.h"
#include "mesh/make_rect.h"
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <vector>
using namespace LMT;

template<class TM> struct GetConn {
    TM *m; std::vector<unsigned> *conn;
    template<class TE> void operator()(const TE &el) const {
        for (unsigned i = 0; i < TE::nb_nodes; ++i) conn->push_back(m->node_list.number(*el.node(i)));
    }
};

template<class TM, class E> int dump(const E &e, unsigned nx, unsigned ny, unsigned nz, const char *out, unsigned npe) {
    TM m;
    typedef typename TM::Pvec Pvec;
    make_rect(m, e, Pvec(0, 0, 0), Pvec(1, 1, 1), Pvec(nx, ny, nz));
    unsigned nn = m.node_list.size(), ne = m.elem_list.size();
    std::vector<double> c(3 * (size_t)nn);
    for (unsigned i = 0; i < nn; ++i) for (int d = 0; d < 3; ++d) c[3 * i + d] = m.node_list[i].pos[d];
    std::vector<unsigned> conn;
    GetConn<TM> g; g.m = &m; g.conn = &conn;
    apply(m.elem_list, g);
    FILE *f = fopen(out, "wb");
    if (!f) return 1;
    fwrite(&nn, 4, 1, f); fwrite(&ne, 4, 1, f); fwrite(&npe, 4, 1, f);
    fwrite(c.data(), 8, c.size(), f); fwrite(conn.data(), 4, conn.size(), f);
    fclose(f);
    printf("%s %u nodes %u elems\n", out, nn, ne);
    return conn.size() == (size_t)ne * npe ? 0 : 2;
}

int main(int argc, char **argv) {
    if (argc != 6) { fprintf(stderr, "usage\n"); return 1; }
    unsigned nx = atoi(argv[2]), ny = atoi(argv[3]), nz = atoi(argv[4]);
    if (!strcmp(argv[1], "tetra")) return dump<Mesh<MeshCaracStd<3, 3, 0, double, 1> > >(Tetra(), nx, ny, nz, argv[5], 4);
    if (!strcmp(argv[1], "tetra10")) return dump<Mesh<MeshCaracStd<3, 3, 0, double, 2> > >(Tetra_10(), nx, ny, nz, argv[5], 10);
    if (!strcmp(argv[1], "hexa")) return dump<Mesh<MeshCaracStd<3, 3, 1, double, 1> > >(Hexa(), nx, ny, nz, argv[5], 8);
    return 1;
}
