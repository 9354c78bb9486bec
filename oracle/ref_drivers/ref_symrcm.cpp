// The reference's node ordering symrcm(m) (util/symrcm.h:30-35, util/symrcm.cpp:38-101 with
// util/GetLinkForOrdering.h) on a make_rect mesh.  Builds against the LMT++ sources where they
// lie; this file only drives them.
// usage: ref_symrcm <tetra|hexa> nx ny nz out.bin
// out.bin: uint32 n_nodes, uint32 n_order, uint32 order[n_order]
#include "mesh/mesh.h"
#include "mesh/meshcaracstd.h"
#include "mesh/make_rect.h"
#include "util/symrcm.h"
#include "util/symrcm.cpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
using namespace LMT;

template<class TM, class E> int run(const E &e, unsigned nx, unsigned ny, unsigned nz, const char *out) {
    TM m;
    typedef typename TM::Pvec Pvec;
    make_rect(m, e, Pvec(0, 0, 0), Pvec(1, 1, 1), Pvec(nx, ny, nz));
    Vec<unsigned> o = symrcm(m);
    FILE *f = fopen(out, "wb");
    if (!f) return 2;
    unsigned nn = m.node_list.size(), no = o.size();
    fwrite(&nn, 4, 1, f);
    fwrite(&no, 4, 1, f);
    for (unsigned i = 0; i < no; ++i) { unsigned v = o[i]; fwrite(&v, 4, 1, f); }
    fclose(f);
    printf("%u nodes, %u ordered\n", nn, no);
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 6) { fprintf(stderr, "usage: ref_symrcm <tetra|hexa> nx ny nz out.bin\n"); return 1; }
    unsigned nx = atoi(argv[2]), ny = atoi(argv[3]), nz = atoi(argv[4]);
    if (!strcmp(argv[1], "hexa")) return run<Mesh<MeshCaracStd<3, 3, 1, double, 1> > >(Hexa(), nx, ny, nz, argv[5]);
    return run<Mesh<MeshCaracStd<3, 3, 0, double, 1> > >(Tetra(), nx, ny, nz, argv[5]);
}
