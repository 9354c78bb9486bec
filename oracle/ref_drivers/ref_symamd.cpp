// The reference's node ordering symamd(m) (util/symamd.h:27-36, util/symamd.cpp:16-36 with
// util/GetLinkForOrdering.h and the vendored AMD 1.1 of amd/, built with -DDINT as its
// SConscript does) on a make_rect mesh: the position of each node in the AMD order, as
// formulation.h:314-318 multiplies it into indice_noda.  Builds against the LMT++ sources where
// they lie; this file only drives them.
// usage: ref_symamd <tetra|tetra10|hexa> nx ny nz out.bin
//        ref_symamd graph in.bin out.bin      symamd(ind) (util/symamd.cpp:16-36) on a node graph
// in.bin : uint32 n, uint32 row_ptr[n+1], uint32 col[row_ptr[n]] (rows sorted, each holding its node)
// out.bin: uint32 n_nodes, uint32 n_order, uint32 position[n_order]
#include "mesh/mesh.h"
#include "mesh/meshcaracstd.h"
#include "mesh/make_rect.h"
#include "util/symamd.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
using namespace LMT;

template<class TM, class E> int run(const E &e, unsigned nx, unsigned ny, unsigned nz, const char *out) {
    TM m;
    typedef typename TM::Pvec Pvec;
    make_rect(m, e, Pvec(0, 0, 0), Pvec(1, 1, 1), Pvec(nx, ny, nz));
    Vec<unsigned> o = symamd(m);
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

int write_positions(const Vec<unsigned> &perm, const char *out) {
    unsigned n = perm.size();
    std::vector<unsigned> pos(n);
    for (unsigned i = 0; i < n; ++i) pos[perm[i]] = i;
    FILE *f = fopen(out, "wb");
    if (!f) return 2;
    fwrite(&n, 4, 1, f);
    fwrite(&n, 4, 1, f);
    fwrite(pos.data(), 4, n, f);
    fclose(f);
    printf("%u nodes ordered\n", n);
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 4 && !strcmp(argv[1], "graph")) {
        FILE *f = fopen(argv[2], "rb");
        unsigned n = 0;
        if (!f || fread(&n, 4, 1, f) != 1) return 3;
        std::vector<unsigned> rp(n + 1);
        if (fread(rp.data(), 4, n + 1, f) != n + 1) return 4;
        std::vector<unsigned> col(rp[n]);
        if (fread(col.data(), 4, col.size(), f) != col.size()) return 4;
        fclose(f);
        Vec<Vec<unsigned> > ind;
        ind.resize(n);
        for (unsigned i = 0; i < n; ++i)
            for (unsigned p = rp[i]; p < rp[i + 1]; ++p) ind[i].push_back(col[p]);
        return write_positions(symamd(ind), argv[3]);
    }
    if (argc != 6) {
        fprintf(stderr, "usage: ref_symamd <tetra|tetra10|hexa> nx ny nz out.bin | graph in.bin out.bin\n");
        return 1;
    }
    unsigned nx = atoi(argv[2]), ny = atoi(argv[3]), nz = atoi(argv[4]);
    if (!strcmp(argv[1], "hexa")) return run<Mesh<MeshCaracStd<3, 3, 1, double, 1> > >(Hexa(), nx, ny, nz, argv[5]);
    if (!strcmp(argv[1], "tetra10"))
        return run<Mesh<MeshCaracStd<3, 3, 0, double, 2> > >(Tetra_10(), nx, ny, nz, argv[5]);
    return run<Mesh<MeshCaracStd<3, 3, 0, double, 1> > >(Tetra(), nx, ny, nz, argv[5]);
}
