// Runs the reference's own block IC(0) + PCG (containers/MatWithTinyBlocks.h:
// conversion :224-269, chol_incomp :426-523, solve_using_incomplete_chol_factorize
// :1101-1151) on a matrix and right-hand side read from a file, for golden
// fixtures and for timing the reference's solve on this host (it prints the iteration count and
// the seconds of chol_incomp and of the PCG).  Builds against the LMT++ headers where they lie.
// in.bin : uint32 n_nodes, uint32 nnzb, uint32 row_ptr[n_nodes+1], uint32 col[nnzb],
//          double vals[nnzb*9] (full block CSR, block row-major), double b[3n],
//          double x0[3n], double crit
// out.bin: uint32 iters, double x[3n], then the factor: per block row r the
//          strictly-lower blocks in column order (9 doubles each, row-major) and
//          the lower-triangular diagonal block (9 doubles, upper part zero),
//          each block preceded by its uint32 block column.
#include "containers/mat.h"
#include "containers/matsparse.h"
#include "containers/MatWithTinyBlocks.h"
#include <cstdio>
#include <time.h>
#include <vector>
#include <sstream>
#include <iostream>
using namespace LMT;

static double now_s() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(int argc, char **argv) {
    if (argc != 3) return 1;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 1;
    unsigned nn, nnzb;
    if (fread(&nn, 4, 1, f) != 1 || fread(&nnzb, 4, 1, f) != 1) return 1;
    std::vector<unsigned> rp(nn + 1), col(nnzb);
    std::vector<double> vals(9 * (size_t)nnzb), b(3 * (size_t)nn), x0(3 * (size_t)nn);
    double crit;
    size_t ok = fread(rp.data(), 4, nn + 1, f) + fread(col.data(), 4, nnzb, f) + fread(vals.data(), 8, vals.size(), f)
              + fread(b.data(), 8, b.size(), f) + fread(x0.data(), 8, x0.size(), f) + fread(&crit, 8, 1, f);
    fclose(f);
    if (ok != nn + 1 + nnzb + vals.size() + b.size() + x0.size() + 1) return 2;
    const unsigned n = 3 * nn;
    Mat<double, Sym<>, SparseLine<> > K(n, n);
    for (unsigned r = 0; r < nn; ++r)
        for (unsigned p = rp[r]; p < rp[r + 1]; ++p)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    unsigned R = 3 * r + i, C = 3 * col[p] + j;
                    double v = vals[9 * (size_t)p + 3 * i + j];
                    if (R >= C && v != 0) K(R, C) = v;
                }
    MatWithTinyBlocks<double, Sym<3> > M(K);
    MatWithTinyBlocks<double, Sym<3> > F = M;
    const double t0 = now_s();
    F.chol_incomp();
    const double t1 = now_s();
    Vec<double> bb; bb.resize(n); for (unsigned i = 0; i < n; ++i) bb[i] = b[i];
    Vec<double> x; x.resize(n); for (unsigned i = 0; i < n; ++i) x[i] = x0[i];
    // count the residual prints (one before the loop, one per iteration)
    std::stringstream ss; std::streambuf *old = std::cout.rdbuf(ss.rdbuf());
    solve_using_incomplete_chol_factorize(F, M, bb, x, crit, true);
    const double t2 = now_s();
    std::cout.rdbuf(old);
    unsigned lines = 0; std::string line; while (std::getline(ss, line)) ++lines;
    unsigned iters = lines ? lines - 1 : 0;
    FILE *o = fopen(argv[2], "wb");
    fwrite(&iters, 4, 1, o);
    for (unsigned i = 0; i < n; ++i) { double v = x[i]; fwrite(&v, 8, 1, o); }
    for (unsigned r = 0; r < nn; ++r) {
        for (unsigned k = 0; k < F.rows[r].indices.size(); ++k) {
            typename MatWithTinyBlocks<double, Sym<3> >::RB::Block &B =
                *reinterpret_cast<typename MatWithTinyBlocks<double, Sym<3> >::RB::Block *>(F.rows[r].data.ptr() + k * MatWithTinyBlocks<double, Sym<3> >::RB::tiny_block_size);
            unsigned c = F.rows[r].indices[k] / 3; fwrite(&c, 4, 1, o);
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) { double v = B(i, j); fwrite(&v, 8, 1, o); }
        }
        fwrite(&r, 4, 1, o);
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) { double v = j <= i ? F.diags[r](i, j) : 0.0; fwrite(&v, 8, 1, o); }
    }
    fclose(o);
    printf("iters %u factor_s %.6f solve_s %.6f\n", iters, t1 - t0, t2 - t1);
    return 0;
}
