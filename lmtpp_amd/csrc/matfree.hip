// Matrix-free block product y = K x for Hexa and Tetra_10 problems on gfx950.
//
// The assembled FP32 product reads about 1 KB of matrix per node (27 blocks of 36 bytes) and runs
// at the HBM read bound.  The same product evaluated element by element reads a few tens of bytes
// per element (block-local corner indices and either its nodes and modulus or, for a box, one
// record), and spends vector instructions instead: K_e x_e is the element loop's quadrature of
// the elasticity weak form (formulation.h:626-647 with
// formulations/formulation_elasticity_iso.py:45-60, 2x2x2 Gauss points of
// formulation.h:2917-2933), applied without forming K_e:
//   per Gauss point q: G = grad u (3x3), sigma = H eps(G), f_a += w_q detJ_q sigma grad N_a,
// and, the Hexa shape functions being products phi(xi0) phi(xi1) phi(xi2), every reference
// derivative at the 8 Gauss points comes from 1D interpolations of lattice differences (sum
// factorisation), the transposed projection the same way.  With adj = detJ J^-1, the Gauss point
// term is (w_q / detJ) sigma(Gr adj^T) adj: one division per point.  Elements that are
// axis-aligned boxes (every lattice mesh) have a constant diagonal J: adj, w / detJ and the
// Hooke coefficients come from a per-element record, and the sums over the Gauss points take a
// closed form (about a third of the general element's arithmetic).
//
// Scatter without atomics: elements sorted along a Morton curve of their centres form blocks of
// MF_BLOCK, one workgroup each (compact bricks on lattices).  The workgroup stages the input at its
// block's nodes in LDS, adds its elements' nodal forces into LDS sub-colour by sub-colour (no two
// elements of a sub-colour share a node), then stores the nodes whose elements all lie in the
// block and, for a node shared with other blocks, its partial in a slot of its own; a
// node-parallel pass sums each shared node's partials in block order.  Every sum has a fixed
// order: the product is deterministic.  The penalty diagonal of the constraints is added per
// block, and nodes of no element get it in a pass of their own.  In the CG iteration the direction
// update p = z + beta p is fused into these passes.
//
// Tetra_10 elements (10 nodes, 4 Gauss points, no product structure to factor) take a per-slot
// record built on the device after each material snapshot: J^-1 and w detJ at each Gauss point
// and the Hooke coefficients, 44 values (176 B FP32), so the element loop reads neither the
// node coordinates nor the modulus; the reference gradients at the fixed Gauss points are
// compile-time constants.  A block then reads ~210 B per element (FP32 record and block-local
// indices) plus its staged nodes, against ~28 stored 3x3 blocks (~1 KB FP32) per node for the
// assembled product.
#include "device_common.hpp"
#include "generated/elasticity_iso.inc"

#include <algorithm>
#include <cstring>
#include <tuple>
#include <type_traits>

// Multiply-adds fuse only within one source expression (the C FP_CONTRACT rule), never across
// statements at the compiler's choice: every instantiation of an element routine (the box-only
// kernel, the mixed box/general one, batches) then rounds the same way, bit for bit.
#pragma clang fp contract(on)

namespace femgpu {

namespace {

// Node of the Hexa at corner (x, y, z) of the reference cube: the generated shape function N_a of
// formulations/element_Hexa.py is phi_x(xi0) phi_y(xi1) phi_z(xi2) with phi_0(t) = 1 - t,
// phi_1(t) = t, so N_a = 1 at that corner.
constexpr int HN[2][2][2] = {{{0, 4}, {3, 7}}, {{1, 5}, {2, 6}}};

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}


}  // namespace

// 1D abscissae of the element loop's 2x2x2 Gauss rule, from its generated table (every weight 1/8).
__device__ inline void hexa_gauss(double &g0, double &g1, double &w) {
    double xi[3];
    gen::vol<HEXA>::gauss(0, xi, w);
    g0 = xi[0];
    gen::vol<HEXA>::gauss(6, xi, w);
    g1 = xi[0];
}

// V[s][t] = sum_{p,q} D[p][q] phi_p(g_s) phi_q(g_t): values at the 2x2 Gauss points of a bilinear
// field given at the 2x2 lattice corners.
template <typename T>
__device__ inline void interp2(const T (&D)[2][2], const T (&g)[2], T (&V)[2][2]) {
    T e[2][2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const T dd = D[1][q] - D[0][q];
#pragma unroll
        for (int s = 0; s < 2; ++s) e[s][q] = D[0][q] + dd * g[s];
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const T dd = e[s][1] - e[s][0];
#pragma unroll
        for (int t = 0; t < 2; ++t) V[s][t] = e[s][0] + dd * g[t];
    }
}

// Transpose of interp2: R[p][q] = sum_{s,t} phi_p(g_s) phi_q(g_t) P[s][t], phi_p(g_s) = w1[p][s].
template <typename T>
__device__ inline void project2(const T (&P)[2][2], const T (&w1)[2][2], T (&R)[2][2]) {
    T r[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int q = 0; q < 2; ++q) r[s][q] = P[s][0] * w1[q][0] + P[s][1] * w1[q][1];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) R[p][q] = r[0][q] * w1[p][0] + r[1][q] * w1[p][1];
}

// det J and adj = det J * J^-1, in jac_inverse's convention (J[j][i] = dx_i/dxi_j).
template <typename T>
__device__ inline T adjugate(const T (&J)[3][3], T (&A)[3][3]) {
    A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    A[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    A[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    A[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    A[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    return J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
}

// w / det: exact division in FP64; in FP32 the hardware reciprocal (1 ulp, against the ~1e-7
// relative accuracy of the FP32 product) instead of the ten-instruction IEEE division.
__device__ inline double wq_over(double det, double w) { return w / det; }
__device__ inline float wq_over(float det, float w) { return w * __builtin_amdgcn_rcpf(det); }

// Ud[d][s][t][c] = d u_c / d xi_d at the Gauss points (s, t) of the two other axes (d = 0:
// (y, z), 1: (x, z), 2: (x, y)) from the corner values u (sum factorisation: 1D interpolations
// of lattice differences).
template <typename T>
__device__ inline void ref_gradients(const T (&u)[8][3], const T (&g)[2], T (&Ud)[3][2][2][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T DU[3][2][2];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                DU[0][p][q] = u[HN[1][p][q]][c] - u[HN[0][p][q]][c];
                DU[1][p][q] = u[HN[p][1][q]][c] - u[HN[p][0][q]][c];
                DU[2][p][q] = u[HN[p][q][1]][c] - u[HN[p][q][0]][c];
            }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            T W[2][2];
            interp2(DU[d], g, W);
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) Ud[d][s][t][c] = W[s][t];
        }
    }
}

// f_a = sum_d phi'_{a_d} (projection of Pd[d] onto the two other lattice axes)
template <typename T>
__device__ inline void project_out(const T (&Pd)[3][2][2][3], const T (&w1)[2][2], T (&f)[8][3]) {
    T Rd[3][2][2][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T Pin[2][2], Rout[2][2];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) Pin[s][t] = Pd[d][s][t][c];
            project2(Pin, w1, Rout);
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) Rd[d][s][t][c] = Rout[s][t];
        }
#pragma unroll
    for (int ax = 0; ax < 2; ++ax)
#pragma unroll
        for (int ay = 0; ay < 2; ++ay)
#pragma unroll
            for (int az = 0; az < 2; ++az)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    f[HN[ax][ay][az]][c] = (ax ? Rd[0][ay][az][c] : -Rd[0][ay][az][c]) +
                                           (ay ? Rd[1][ax][az][c] : -Rd[1][ax][az][c]) +
                                           (az ? Rd[2][ax][ay][c] : -Rd[2][ax][ay][c]);
}

// f[a][c] = (K_e u_e) of one Hexa with node coordinates X and displacements u, H as
// (H11 - H12, H12, H44 / 2).
template <typename T>
__device__ inline void hexa_apply(const double (&X)[8][3], const T (&u)[8][3], T h01, T h1, T h2, T (&f)[8][3]) {
    double g0d, g1d, wq;
    hexa_gauss(g0d, g1d, wq);
    const T g[2] = {(T)g0d, (T)g1d};
    const T w1[2][2] = {{(T)(1.0 - g0d), (T)(1.0 - g1d)}, {(T)g0d, (T)g1d}};
    // Jd[d][s][t][c] = d x_c / d xi_d at the Gauss points (s, t) of the two other axes; the
    // coordinate differences are taken in FP64 before rounding to T.
    T Jd[3][2][2][3], Ud[3][2][2][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T DX[3][2][2];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                DX[0][p][q] = (T)(X[HN[1][p][q]][c] - X[HN[0][p][q]][c]);
                DX[1][p][q] = (T)(X[HN[p][1][q]][c] - X[HN[p][0][q]][c]);
                DX[2][p][q] = (T)(X[HN[p][q][1]][c] - X[HN[p][q][0]][c]);
            }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            T V[2][2];
            interp2(DX[d], g, V);
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) Jd[d][s][t][c] = V[s][t];
        }
    }
    ref_gradients(u, g, Ud);
    // Gauss point (i, j, k): S = (w / detJ) sigma(Gr adj^T) adj, summed into Pd over the axis each
    // reference derivative does not depend on.
    T Pd[3][2][2][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int c = 0; c < 3; ++c) Pd[d][s][t][c] = 0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                T J[3][3], Gr[3][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    J[0][c] = Jd[0][j][k][c];
                    J[1][c] = Jd[1][i][k][c];
                    J[2][c] = Jd[2][i][j][c];
                    Gr[c][0] = Ud[0][j][k][c];
                    Gr[c][1] = Ud[1][i][k][c];
                    Gr[c][2] = Ud[2][i][j][c];
                }
                T Aj[3][3];
                const T kq = wq_over(adjugate(J, Aj), (T)wq);
                T G[3][3];  // detJ * d u_c / d x_m
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int m = 0; m < 3; ++m) G[c][m] = Gr[c][0] * Aj[m][0] + Gr[c][1] * Aj[m][1] + Gr[c][2] * Aj[m][2];
                const T tr = G[0][0] + G[1][1] + G[2][2];
                const T ks = kq * h2;
                T sg[3][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) sg[c][c] = kq * (h01 * G[c][c] + h1 * tr);
                sg[0][1] = sg[1][0] = ks * (G[0][1] + G[1][0]);
                sg[0][2] = sg[2][0] = ks * (G[0][2] + G[2][0]);
                sg[1][2] = sg[2][1] = ks * (G[1][2] + G[2][1]);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    Pd[0][j][k][c] += sg[c][0] * Aj[0][0] + sg[c][1] * Aj[1][0] + sg[c][2] * Aj[2][0];
                    Pd[1][i][k][c] += sg[c][0] * Aj[0][1] + sg[c][1] * Aj[1][1] + sg[c][2] * Aj[2][1];
                    Pd[2][i][j][c] += sg[c][0] * Aj[0][2] + sg[c][1] * Aj[1][2] + sg[c][2] * Aj[2][2];
                }
            }
    project_out(Pd, w1, f);
}

// One axis of the box product: Pd[D][s][t][.] = b sum over the two Gauss points along xi_D of
// sigma(G)[.][D], G[c][m] = Ud[m][.][c] ad_m, in closed form.  Column m of G does not vary along
// xi_m (the reference derivative along xi_m of a Q1 field is constant along xi_m), so the sum
// is twice the terms of column D plus the sums along xi_D of columns E and F (the other axes;
// (s, t) are the Gauss point indices along E < F).
template <int D, typename T>
__device__ inline void box_axis(const T (&Ud)[3][2][2][3], const T (&ad)[3], T b, T cs, T c1, T c2,
                                T (&Pd)[3][2][2][3]) {
    constexpr int E = D == 0 ? 1 : 0, F = D == 2 ? 1 : 2;
    // Ud[E] at (n_D, n_F) and Ud[F] at (n_D, n_E), each indexed by its own two axes in order
    auto udE = [&](int nD, int nF, int c) -> T { return D < F ? Ud[E][nD][nF][c] : Ud[E][nF][nD][c]; };
    auto udF = [&](int nD, int nE, int c) -> T { return D < E ? Ud[F][nD][nE][c] : Ud[F][nE][nD][c]; };
    const T bc1 = b * c1, bc2 = b * c2;
    const T A = b * cs * ad[D], BE = bc1 * ad[E], BF = bc1 * ad[F];
    const T C = bc2 * ((T)2 * ad[D]), DE = bc2 * ad[E], DF = bc2 * ad[F];
    T seE[2], seD[2], sfF[2], sfD[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        seE[t] = BE * (udE(0, t, E) + udE(1, t, E));
        seD[t] = DE * (udE(0, t, D) + udE(1, t, D));
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        sfF[q] = BF * (udF(0, q, F) + udF(1, q, F));
        sfD[q] = DF * (udF(0, q, D) + udF(1, q, D));
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            Pd[D][q][t][D] = A * Ud[D][q][t][D] + (seE[t] + sfF[q]);
            Pd[D][q][t][E] = C * Ud[D][q][t][E] + seD[t];
            Pd[D][q][t][F] = C * Ud[D][q][t][F] + sfD[q];
        }
}

// The same for an axis-aligned box (every edge along xi_d is h_d e_d): J = diag(h) at every
// Gauss point, so adj = diag(h1 h2, h0 h2, h0 h1) and w / detJ are constant, and the point terms
// (w / detJ) sigma(Gr adj^T) adj need the diagonal only; their sums over the points along each
// axis take the closed form of box_axis (no geometry: the "Cartesian" cell of matrix-free FE
// codes).  c0, c1, c2: H11 - H12, H12, H44 / 2.
template <typename T>
__device__ inline void hexa_apply_box(const T (&h)[3], const T (&u)[8][3], T c0, T c1, T c2, T (&f)[8][3]) {
    double g0d, g1d, wq;
    hexa_gauss(g0d, g1d, wq);
    const T g[2] = {(T)g0d, (T)g1d};
    const T w1[2][2] = {{(T)(1.0 - g0d), (T)(1.0 - g1d)}, {(T)g0d, (T)g1d}};
    T Ud[3][2][2][3];
    ref_gradients(u, g, Ud);
    const T ad[3] = {h[1] * h[2], h[0] * h[2], h[0] * h[1]};
    const T kq = wq_over(h[0] * ad[0], (T)wq);
    const T cs = (T)2 * (c0 + c1);
    T Pd[3][2][2][3];
    box_axis<0>(Ud, ad, kq * ad[0], cs, c1, c2, Pd);
    box_axis<1>(Ud, ad, kq * ad[1], cs, c1, c2, Pd);
    box_axis<2>(Ud, ad, kq * ad[2], cs, c1, c2, Pd);
    project_out(Pd, w1, f);
}

// Per-slot record of a box element for the product in precision T: its edge lengths and the Hooke
// coefficients (H11 - H12, H12, H44 / 2) of its modulus at the last matrix assembly, rounded to T
// exactly as the general path rounds them.  24 B in FP32 instead of the 44 B of nodes, element
// index and modulus plus the corner gathers.
template <typename T>
struct BoxRec {
    T h[3], c[3];
};

template <typename T>
__device__ inline void box_coefs(const double (&X0)[3], const double (&X6)[3], double E, double nu, BoxRec<T> &r) {
    double Hd[3];
    gen::hooke(E, nu, Hd);
#pragma unroll
    for (int d = 0; d < 3; ++d) r.h[d] = (T)(X6[d] - X0[d]);
    r.c[0] = (T)(Hd[0] - Hd[1]);
    r.c[1] = (T)Hd[1];
    r.c[2] = (T)(0.5 * Hd[2]);
}

// The records of every box slot, both precisions (after each snapshot of the material).
__global__ __launch_bounds__(256) void k_mf_boxrec(size_t n, const uint8_t *__restrict__ color,
                                                   const uint32_t *__restrict__ conn, const uint32_t *__restrict__ elem,
                                                   const uint32_t *__restrict__ bsub,
                                                   const double *__restrict__ coords, const double *__restrict__ E,
                                                   const SubMaterial *__restrict__ mat, BoxRec<double> *__restrict__ r64,
                                                   BoxRec<float> *__restrict__ r32) {
    const size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const uint32_t cb = color[s];
    if (cb == MF_EMPTY || !(cb & MF_BOX)) return;
    double X0[3], X6[3];
    for (int d = 0; d < 3; ++d) {
        X0[d] = coords[3 * (size_t)conn[8 * s] + d];
        X6[d] = coords[3 * (size_t)conn[8 * s + 6] + d];
    }
    const double nu = mat[bsub[s / MF_BLOCK]].nu;
    box_coefs<double>(X0, X6, E[elem[s]], nu, r64[s]);
    box_coefs<float>(X0, X6, E[elem[s]], nu, r32[s]);
}

// Tetra_10 slot record: per Gauss point q of the element loop's 4-point rule
// (element_Tetra_10.h:12-15), ji[q][3 m + j] = d xi_j / d x_m (J^-1 from the adjugate, J_ji =
// d x_i / d xi_j from all 10 nodes, formal_lf/element.py:195-205) and c[q] = w_q detJ_q (signed,
// formal_lf/formulation.py:553), then the Hooke coefficients (H11 - H12, H12, H44 / 2).
template <typename T>
struct TetRec {
    T ji[4][9];
    T c[4];
    T h[3];
    T pad;
};

template <typename T>
__device__ inline void tet_coefs(const double (&X)[10][3], double E, double nu, TetRec<T> &r) {
    double Hd[3];
    gen::hooke(E, nu, Hd);
    r.h[0] = (T)(Hd[0] - Hd[1]);
    r.h[1] = (T)Hd[1];
    r.h[2] = (T)(0.5 * Hd[2]);
    r.pad = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double xi[3], w;
        gen::vol<TETRA10>::gauss(q, xi, w);
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int a = 0; a < 10; ++a) {
            double d[3];
            (void)gen::vol<TETRA10>::shape(a, xi, d);
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) J[j][i] += d[j] * X[a][i];
        }
        double A[3][3];
        const double det = adjugate(J, A);  // A[m][j] = detJ d xi_j / d x_m
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int j = 0; j < 3; ++j) r.ji[q][3 * m + j] = (T)(A[m][j] / det);
        r.c[q] = (T)(w * det);
    }
}

// The records of every Tetra_10 slot, both precisions (after each snapshot of the material).
__global__ __launch_bounds__(256) void k_mf_tetrec(size_t n, const uint8_t *__restrict__ color,
                                                   const uint32_t *__restrict__ conn, const uint32_t *__restrict__ elem,
                                                   const uint32_t *__restrict__ bsub,
                                                   const double *__restrict__ coords, const double *__restrict__ E,
                                                   const SubMaterial *__restrict__ mat, TetRec<double> *__restrict__ r64,
                                                   TetRec<float> *__restrict__ r32) {
    const size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n || color[s] == MF_EMPTY) return;
    double X[10][3];
    for (int a = 0; a < 10; ++a)
        for (int d = 0; d < 3; ++d) X[a][d] = coords[3 * (size_t)conn[10 * s + a] + d];
    const double nu = mat[bsub[s / MF_BLOCK]].nu;
    tet_coefs<double>(X, E[elem[s]], nu, r64[s]);
    tet_coefs<float>(X, E[elem[s]], nu, r32[s]);
}

// f[a][c] = (K_e u_e) of one Tetra_10 from its record: per Gauss point, G = grad u (reference
// gradients of the generated shape functions, constants at the fixed points, times J^-1),
// sigma = H eps(G) scaled by w detJ, f_a += sigma grad N_a.
template <typename T>
__device__ inline void tet10_apply(const TetRec<T> &r, const T (&u)[10][3], T (&f)[10][3]) {
#pragma unroll
    for (int a = 0; a < 10; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) f[a][c] = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double xi[3], w;
        gen::vol<TETRA10>::gauss(q, xi, w);
        T dr[10][3];
#pragma unroll
        for (int a = 0; a < 10; ++a) {
            double d[3];
            (void)gen::vol<TETRA10>::shape(a, xi, d);
#pragma unroll
            for (int j = 0; j < 3; ++j) dr[a][j] = (T)d[j];
        }
        T Gr[3][3];  // d u_c / d xi_j
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                T v = 0;
#pragma unroll
                for (int a = 0; a < 10; ++a) v += u[a][c] * dr[a][j];
                Gr[c][j] = v;
            }
        T G[3][3];  // d u_c / d x_m
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int m = 0; m < 3; ++m)
                G[c][m] = Gr[c][0] * r.ji[q][3 * m] + Gr[c][1] * r.ji[q][3 * m + 1] + Gr[c][2] * r.ji[q][3 * m + 2];
        const T kq = r.c[q], tr = G[0][0] + G[1][1] + G[2][2], ks = kq * r.h[2];
        T sg[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) sg[c][c] = kq * (r.h[0] * G[c][c] + r.h[1] * tr);
        sg[0][1] = sg[1][0] = ks * (G[0][1] + G[1][0]);
        sg[0][2] = sg[2][0] = ks * (G[0][2] + G[2][0]);
        sg[1][2] = sg[2][1] = ks * (G[1][2] + G[2][1]);
        T Sx[3][3];  // sum_m sigma_cm d xi_j / d x_m
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Sx[c][j] = sg[c][0] * r.ji[q][j] + sg[c][1] * r.ji[q][3 + j] + sg[c][2] * r.ji[q][6 + j];
#pragma unroll
        for (int a = 0; a < 10; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) f[a][c] += Sx[c][0] * dr[a][0] + Sx[c][1] * dr[a][1] + Sx[c][2] * dr[a][2];
    }
}

// r.z partials per lane the folded element-block kernel loads with its node lists (a single
// problem's vector update leaves one per 1024 nodes: 4 per lane cover 1M nodes)
constexpr uint32_t MF_RZ_PER_LANE = 4;
// Waves per SIMD the box-only element-block kernels are held to (the second __launch_bounds__
// argument): 6 in FP32 (80 VGPRs) and 4 in FP64 (128), what the kernels without the folded
// vector-update end reach by themselves.
#ifndef MF_WG_F32
#define MF_WG_F32 6
#endif
#ifndef MF_WG_F64
#define MF_WG_F64 4
#endif
// LDS of k_mf_block: per local node its 3 input values and 3 accumulators, each triple at a
// stride of MF_NS values.  (A stride of 4, one 16-byte access per node, was measured: 29.7 us
// against 29.3 us per FP32 element-block launch at Hexa8 100^3, 49.0 against 46.0 us in FP64.)
constexpr int MF_NS = 3;
// Accumulator copies of the Hexa element blocks: sub-colour c adds into copy c % n in pass c / n,
// so the 8 sub-colours of a brick of boxes take 8 / n passes (and barriers), and the write-out
// sums the copies in order.  n is the block's own: the most copies, up to MF_ACC_F32 (FP32) or
// MF_ACC_F64 (FP64), whose LDS (input and copies, 3 values of T per local node) stays within the
// share of one of the MF_WG_F32 / MF_WG_F64 workgroups per CU (FP32: 4 copies up to 455 local
// nodes, a Hexa8 brick has 405; FP64: 2 up to 568).  A function of the block alone, so a
// sub-problem sums alike alone and in a batch; the launch's LDS is the largest block's.  (A fixed
// 4 copies sized by the largest block of a batch of uneven cubes left 3-4 workgroups per CU:
// 153 against 138 ms for the batched solve.)  Tetra_10: 1.
#ifndef MF_ACC_F32
#define MF_ACC_F32 4
#endif
#ifndef MF_ACC_F64
#define MF_ACC_F64 2
#endif
constexpr uint32_t mf_ilog2(int v) { return v > 1 ? 1 + mf_ilog2(v / 2) : 0; }
// log2 of the copies of a block whose LDS fits (every block of a lattice): UNI launches
template <typename T, int K>
constexpr uint32_t mf_acc_log2_max() { return K == TETRA10 ? 0 : mf_ilog2(sizeof(T) == 4 ? MF_ACC_F32 : MF_ACC_F64); }
// log2 of block copies for a block of m local nodes (T's size vsize)
inline uint32_t mf_acc_log2(uint32_t m, size_t vsize, int kind) {
    if (kind == TETRA10) return 0;
    const size_t budget = (size_t)160 * 1024 / (vsize == 4 ? MF_WG_F32 : MF_WG_F64);
    const int cap = vsize == 4 ? MF_ACC_F32 : MF_ACC_F64;
    uint32_t lg = 0;
    while ((2 << lg) <= cap && (size_t)std::max<uint32_t>(1, m) * (1 + (2 << lg)) * MF_NS * vsize <= budget) ++lg;
    return lg;
}
// LDS of a block: its input values and 2^lg accumulator copies, at a stride of m nodes
inline size_t mf_lds_bytes(uint32_t m, size_t vsize, uint32_t lg) {
    return (size_t)std::max<uint32_t>(1, m) * (1 + (1u << lg)) * MF_NS * vsize;
}

// The 3 values of local node i at LDS triple p (stride MF_NS), as vector accesses when the stride is 4.
template <typename T>
__device__ inline void lds_get3(const T *p, T &a, T &b, T &c) {
    if constexpr (MF_NS == 4 && sizeof(T) == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        a = v.x; b = v.y; c = v.z;
    } else if constexpr (MF_NS == 4 && sizeof(T) == 8) {
        const double2 v = *reinterpret_cast<const double2 *>(p);
        a = v.x; b = v.y; c = p[2];
    } else {
        a = p[0]; b = p[1]; c = p[2];
    }
}
template <typename T>
__device__ inline void lds_put3(T *p, T a, T b, T c) {
    if constexpr (MF_NS == 4 && sizeof(T) == 4) {
        *reinterpret_cast<float4 *>(p) = make_float4(a, b, c, 0.0f);
    } else if constexpr (MF_NS == 4 && sizeof(T) == 8) {
        *reinterpret_cast<double2 *>(p) = make_double2(a, b);
        p[2] = c;
    } else {
        p[0] = a; p[1] = b; p[2] = c;
    }
}

// One block of MF_BLOCK element slots per workgroup (logical block b = xcd_slab(), so each XCD
// takes a contiguous run of blocks): y of the block's own nodes, partials of its shared ones.
// The workgroup first stages the product's input at its block's local nodes in LDS (one
// coalesced pass over the node list), and each element lane reads its 8 corners' values there by
// their block-local indices; the corner coordinates (FP64, read-only) are gathered from memory.
// PQ: partial of sum_e x_e . (K_e x_e) = sum over the block's local nodes of x . (their summed
// element forces), taken at the write-out; with the penalty pass's partials it is x . K x.
// DIR: the CG direction update is fused in.  The direction lives in two buffers, x and xb, used in
// turn by the parity of the iteration count (device-side, so captured iterations, pauses and
// restarts agree): the product is applied to p = z + beta s (beta of the scalars, s the stored
// direction, in x for even iterations and xb for odd ones), and the block stores p into the other
// buffer at the nodes it owns and at the shared nodes it holds first (lslot bit 31), so every node
// is written once and no block reads a value another block has already replaced.
// LDS: per local node 3 input values and the block's copies of 3 accumulators (mf_acc_log2).
// (FP64 Tetra_10: capped at 3 waves per SIMD, 168 VGPRs and 40 B of spills, against 184 VGPRs and
// 2 waves uncapped: 80 against 84 us per product at 384k elements)
// FIN (with PQ and DIR: the two-kernel CG iteration): the end of the previous vector update is
// folded in.  When its partials are pending (fin_pending), every workgroup combines the r.z
// partials rpart itself (in the fixed order of wg_combine256, so all hold the same bits) and forms
// beta = r.z / rho; the sub-problem's first block also combines r.r and max|r|, applies the stop
// test (fin_update) and writes the advanced scalars into S1, the copy the vector update reads (no
// workgroup reads a value another one writes in the same kernel); without pending partials it
// copies S into S1 unchanged.  A stop or pause decided here takes effect from the vector update
// on: this launch still forms the next direction into the other buffer and its product, which
// nothing reads, and leaves the stored direction as it was.
// UNI: every block of the launch takes the most copies (the arithmetic of the per-block choice,
// with the copy count a compile-time constant; Hexa8 100^3 MIXED 85.8-86.4 against 87.1-87.8 ms
// read from the block's word).
template <typename T, bool PQ, bool DIR, bool BOXES, int K = HEXA, bool FIN = false, bool UNI = false>
__global__ __launch_bounds__(256, (K == TETRA10 && sizeof(T) == 8) ? 3 : (BOXES ? (sizeof(T) == 4 ? MF_WG_F32 : MF_WG_F64) : 1)) void k_mf_block(const uint32_t *__restrict__ conn, const uint16_t *__restrict__ loc,
                                                  const uint8_t *__restrict__ color, const uint32_t *__restrict__ elem,
                                                  const uint32_t *__restrict__ bsub,
                                                  uint32_t n_colors, uint32_t ls, const uint32_t *__restrict__ bls,
                                                  uint32_t lsl_all, const uint32_t *__restrict__ lnode,
                                                  const uint32_t *__restrict__ lslot,
                                                  const uint32_t *__restrict__ pptr, const uint32_t *__restrict__ pidx,
                                                  const uint32_t *__restrict__ psrc, const double *__restrict__ fpen,
                                                  const double *__restrict__ coords, const double *__restrict__ E,
                                                  const SubMaterial *__restrict__ mat,
                                                  const void *__restrict__ recv, const T *x, T *xb,
                                                  const T *__restrict__ z, T *__restrict__ y, T *__restrict__ halo,
                                                  const ProbScalars *S, double *partials, ProbScalars *S1,
                                                  const double *__restrict__ rpart, unsigned rstride,
                                                  const uint32_t *__restrict__ rw0, const uint32_t *__restrict__ rnw,
                                                  const uint32_t *__restrict__ b0) {
    extern __shared__ double mf_lds[];
    constexpr int NPE = K == TETRA10 ? 10 : 8;
    const unsigned b = xcd_slab();
    // the block's local nodes at a fixed stride ls (MF_NO_NODE past its own count), so their
    // loads need no other load first
    const size_t l0 = (size_t)b * ls;
    // everything the later phases read from memory at the block's own index: the penalty range,
    // the global node and partial slot of the lane's first MF_NODE_REGS local nodes (a brick's
    // ~400; larger blocks reload the rest), and the element's block-local node indices (8 u16 per
    // Hexa slot, 16 with 10 used per Tetra_10 slot) and a box's record.  Every index is in bounds
    // for every block, padding blocks and empty slots included.  The plain and p.q products load
    // them ahead of the sub-problem's scalars (a load after a branch or a barrier cannot be
    // hoisted above it: 24.5 against 25.2 us per FP32 launch at Hexa8 100^3, 38.8 against 39.6 in
    // FP64); the iteration's FIN kernel after its scalars and r.z partials (loaded ahead, its
    // iteration took 57.6 against 57.0 us in FP32).
    uint32_t pk0, pk1, gnode[MF_NODE_REGS], gslot[MF_NODE_REGS], cbits, lo[NPE];
    const size_t slot = (size_t)b * MF_BLOCK + threadIdx.x;
    BoxRec<T> br;    // a box: edge lengths and Hooke coefficients
    auto load_block = [&]() {
        pk0 = pptr[b];
        pk1 = pptr[b + 1];
#pragma unroll
        for (int r = 0; r < (int)MF_NODE_REGS; ++r) {
            const uint32_t i = threadIdx.x + r * MF_BLOCK;
            gnode[r] = i < ls ? lnode[l0 + i] : MF_NO_NODE;
            gslot[r] = i < ls ? lslot[l0 + i] : 0u;
        }
        cbits = color[slot];
        const uint4 lq = reinterpret_cast<const uint4 *>(loc)[K == TETRA10 ? 2 * slot : slot];
        lo[0] = lq.x & 0xffffu; lo[1] = lq.x >> 16; lo[2] = lq.y & 0xffffu; lo[3] = lq.y >> 16;
        lo[4] = lq.z & 0xffffu; lo[5] = lq.z >> 16; lo[6] = lq.w & 0xffffu; lo[7] = lq.w >> 16;
        if constexpr (K == TETRA10) {
            const uint32_t l89 = reinterpret_cast<const uint32_t *>(loc)[8 * slot + 4];
            lo[NPE - 2] = l89 & 0xffffu;
            lo[NPE - 1] = l89 >> 16;
        } else if (BOXES || (cbits != MF_EMPTY && (cbits & MF_BOX))) {
            br = reinterpret_cast<const BoxRec<T> *>(recv)[slot];
        }
    };
    if constexpr (!FIN) load_block();
    const uint32_t sb = bsub[b];  // the block's sub-problem (NO_SUB: an empty padding block)
    if (sb == NO_SUB) {
        // a single problem's p.q combine spans the padding blocks too: they contribute zero
        if (PQ && threadIdx.x == 0) partials[b] = 0.0;
        return;
    }
    if (PQ && !selected(S[sb], SEL_ACTIVE)) {
        if (FIN && b == b0[sb]) copy_scalars(S1 + sb, S + sb);
        return;
    }
    // FIN: the pending update's iteration count is the one this product belongs to
    const bool fin = FIN && S[sb].fin_pending;
    // FIN: the r.z partials of the pending update, loaded here with the node lists (their memory
    // round trip overlaps that of the node lists, ahead of the loads of z and the direction) when
    // they fit one batch of MF_RZ_PER_LANE per lane; summed in the order of lane_partials
    double rzt[MF_RZ_PER_LANE];
    bool rz_pre = false;
    if constexpr (FIN) {
        const unsigned nw = rnw[sb], w0 = rw0[sb];
        rz_pre = fin && nw <= MF_RZ_PER_LANE * MF_BLOCK;
#pragma unroll
        for (int j = 0; j < (int)MF_RZ_PER_LANE; ++j) {
            const unsigned i = threadIdx.x + j * MF_BLOCK;
            rzt[j] = rz_pre && i < nw ? rpart[w0 + i] : 0.0;
        }
    }
    // DIR: stored direction in, updated direction out (the two buffers by iteration parity)
    const bool odd = DIR && ((S[sb].iter + (fin ? 1 : 0)) & 1);
    const T *xs = odd ? xb : x;
    T *xo = odd ? const_cast<T *>(x) : xb;
    // the block's local node count and accumulator copies (2^lg: bits 24-27 FP32, 28-31 FP64)
    uint32_t lsl, lg;
    if constexpr (UNI) {
        lsl = lsl_all;
        lg = mf_acc_log2_max<T, K>();
    } else {
        const uint32_t bw = bls[b];
        lsl = bw & 0xffffffu;
        lg = (bw >> (sizeof(T) == 4 ? 24 : 28)) & 0xfu;
    }
    const int nacc = 1 << lg;
    T *Ps = reinterpret_cast<T *>(mf_lds);  // [lsl][MF_NS] the product's input
    T *acc = Ps + MF_NS * (size_t)lsl;      // nacc x [lsl][MF_NS] summed nodal forces
    const size_t astr = MF_NS * (size_t)lsl;
    T beta = DIR ? (T)S[sb].beta : (T)0;
    if constexpr (FIN) load_block();
    const uint32_t col = cbits == MF_EMPTY ? MF_EMPTY : cbits & MF_COLOR_BITS;
    const bool box = BOXES || (cbits != MF_EMPTY && (cbits & MF_BOX));
    auto stage = [&](uint32_t i, uint32_t gn) {
        const size_t g = 3 * (size_t)gn;
        T v[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] = DIR ? dir_value(z[g + d], beta, xs[g + d]) : x[g + d];
        lds_put3(Ps + MF_NS * i, v[0], v[1], v[2]);
        for (int k = 0; k < nacc; ++k) lds_put3(acc + k * astr + MF_NS * i, (T)0, (T)0, (T)0);
    };
    if constexpr (FIN) {
        // the operands of the first nodes' direction update are loaded before the combine of the
        // partials that beta needs, so their latencies overlap
        T zr[MF_NODE_REGS][3], sr[MF_NODE_REGS][3];
#pragma unroll
        for (int r = 0; r < (int)MF_NODE_REGS; ++r)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const size_t g = gnode[r] != MF_NO_NODE ? 3 * (size_t)gnode[r] + d : 0;
                zr[r][d] = gnode[r] != MF_NO_NODE ? z[g] : (T)0;
                sr[r][d] = gnode[r] != MF_NO_NODE ? xs[g] : (T)0;
            }
        const bool lead = b == b0[sb];
        if (lead) copy_scalars(S1 + sb, S + sb);
        if (fin) {
            // r.z in every block (its barrier orders the lead's copy before the record below)
            double a[1], rz[1];
            if (rz_pre) {
                a[0] = 0.0;
#pragma unroll
                for (int j = 0; j < (int)MF_RZ_PER_LANE; ++j) a[0] = a[0] + rzt[j];
            } else {
                lane_partials<1, 0u, MF_RZ_PER_LANE>(rpart, rstride, rw0[sb], rnw[sb], a);
            }
            wg_combine_lanes<1, 0u>(a, rz);
            beta = (T)(rz[0] / S[sb].rho);
            if (lead) {
                // r.r and max|r| for the stop test
                double rm[2];
                wg_combine256<2, 2u, 2>(rpart + rstride, rstride, rw0[sb], rnw[sb], rm);
                if (threadIdx.x == 0) {
                    const double o[3] = {rz[0], rm[0], rm[1]};
                    fin_update(S[sb], S1[sb], o, true);
                    S1[sb].fin_pending = 0;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < (int)MF_NODE_REGS; ++r)
            if (gnode[r] != MF_NO_NODE) {
                const uint32_t i = threadIdx.x + r * MF_BLOCK;
                lds_put3(Ps + MF_NS * i, dir_value(zr[r][0], beta, sr[r][0]), dir_value(zr[r][1], beta, sr[r][1]),
                         dir_value(zr[r][2], beta, sr[r][2]));
                for (int k = 0; k < nacc; ++k) lds_put3(acc + k * astr + MF_NS * i, (T)0, (T)0, (T)0);
            }
    } else {
#pragma unroll
        for (int r = 0; r < (int)MF_NODE_REGS; ++r)
            if (gnode[r] != MF_NO_NODE) stage(threadIdx.x + r * MF_BLOCK, gnode[r]);
    }
    for (uint32_t i = threadIdx.x + MF_NODE_REGS * MF_BLOCK; i < ls; i += MF_BLOCK)
        if (lnode[l0 + i] != MF_NO_NODE) stage(i, lnode[l0 + i]);
    double pq[1] = {0.0};
    T f[NPE][3];
    double X[8][3];  // corners of a general element
    double Hd[3] = {0, 0, 0};
    if (col != MF_EMPTY) {
        if constexpr (K != TETRA10) if (!box) {
            const uint4 ca = reinterpret_cast<const uint4 *>(conn)[2 * slot];
            const uint4 cb = reinterpret_cast<const uint4 *>(conn)[2 * slot + 1];
            const uint32_t nd[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
            gen::hooke(E[elem[slot]], mat[sb].nu, Hd);
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) X[a][c] = coords[3 * (size_t)nd[a] + c];
        }
    }
    __syncthreads();
    if (col != MF_EMPTY) {
        T u[NPE][3];
#pragma unroll
        for (int a = 0; a < NPE; ++a) lds_get3(Ps + MF_NS * lo[a], u[a][0], u[a][1], u[a][2]);
        if constexpr (K == TETRA10) {
            tet10_apply<T>(reinterpret_cast<const TetRec<T> *>(recv)[slot], u, f);
        } else {
            if (box) hexa_apply_box<T>(br.h, u, br.c[0], br.c[1], br.c[2], f);
            else hexa_apply<T>(X, u, (T)(Hd[0] - Hd[1]), (T)Hd[1], (T)(0.5 * Hd[2]), f);
        }
    }
    // (no barrier here: the element loop reads only the inputs, the passes only the accumulators,
    // which the staging zeroed before the barrier above)
    // no two lanes of a sub-colour touch one entry, and sub-colour c adds into copy c % nacc in
    // pass c / nacc, so the order of every sum is fixed: each copy in sub-colour order, the copies
    // summed in order at the write-out; plain read-add-write (LDS float atomics measured 2.6x
    // slower here in FP32)
    // (an element's nodes are distinct, so all its entries are read before any is written: one
    // LDS round trip per pass instead of one per node)
    const uint32_t n_pass = (n_colors + nacc - 1) >> lg;
    for (uint32_t c = 0; c < n_pass; ++c) {
        if (col != MF_EMPTY && (col >> lg) == c) {
            T *ac = acc + (col & (nacc - 1)) * astr;
            T cur[NPE][3];
#pragma unroll
            for (int a = 0; a < NPE; ++a) lds_get3(ac + MF_NS * lo[a], cur[a][0], cur[a][1], cur[a][2]);
#pragma unroll
            for (int a = 0; a < NPE; ++a)
                lds_put3(ac + MF_NS * lo[a], cur[a][0] + f[a][0], cur[a][1] + f[a][1], cur[a][2] + f[a][2]);
        }
        __syncthreads();
    }
    // penalty terms of the constrained DOFs assigned to this block (distinct local entries); the
    // range is the block's, so a block without any skips the barrier
    if (pk1 > pk0) {
        for (uint32_t k = pk0 + threadIdx.x; k < pk1; k += MF_BLOCK) acc[pidx[k]] += (T)fpen[psrc[k]] * Ps[pidx[k]];
        __syncthreads();
    }
    auto write_out = [&](uint32_t i, uint32_t s, uint32_t g) {
        const bool own = s == 0xffffffffu;
        T *dst = own ? y + 3 * (size_t)g : halo + 3 * (size_t)(s & MF_SLOT_BITS);
        T a0, a1, a2;
        lds_get3(acc + MF_NS * i, a0, a1, a2);
        for (int k = 1; k < nacc; ++k) {
            T b0, b1, b2;
            lds_get3(acc + k * astr + MF_NS * i, b0, b1, b2);
            a0 += b0;
            a1 += b1;
            a2 += b2;
        }
        dst[0] = a0;
        dst[1] = a1;
        dst[2] = a2;
        if (PQ || DIR) {
            T x0, x1, x2;
            lds_get3(Ps + MF_NS * i, x0, x1, x2);
            if (PQ) pq[0] += (double)x0 * (double)a0 + (double)x1 * (double)a1 + (double)x2 * (double)a2;
            if (DIR && (own || (s & MF_FIRST_HOLDER))) {
                xo[3 * (size_t)g] = x0;
                xo[3 * (size_t)g + 1] = x1;
                xo[3 * (size_t)g + 2] = x2;
            }
        }
    };
#pragma unroll
    for (int r = 0; r < (int)MF_NODE_REGS; ++r)
        if (gnode[r] != MF_NO_NODE) write_out(threadIdx.x + r * MF_BLOCK, gslot[r], gnode[r]);
    for (uint32_t i = threadIdx.x + MF_NODE_REGS * MF_BLOCK; i < ls; i += MF_BLOCK)
        if (lnode[l0 + i] != MF_NO_NODE) write_out(i, lslot[l0 + i], lnode[l0 + i]);
    if (PQ) wg_partial<1, 0u>(pq, partials, 0u, b);
}

// Shared nodes: y = sum of the blocks' partials, in block order (a node of a finished
// sub-problem sums stale partials into a q nobody reads).
template <typename T>
__global__ __launch_bounds__(256) void k_mf_halo(uint32_t n, const uint32_t *__restrict__ bnode,
                                                 const uint32_t *__restrict__ bptr, const T *__restrict__ halo,
                                                 T *__restrict__ y) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    T v0 = 0, v1 = 0, v2 = 0;
    for (uint32_t k = bptr[i]; k < bptr[i + 1]; ++k) {
        v0 += halo[3 * (size_t)k];
        v1 += halo[3 * (size_t)k + 1];
        v2 += halo[3 * (size_t)k + 2];
    }
    T *py = y + 3 * (size_t)bnode[i];
    py[0] = v0;
    py[1] = v1;
    py[2] = v2;
}

// Nodes of no element: their rows hold the penalty alone, y = M P x on their constrained DOFs and
// 0 on the others.
// DIR: applied to the updated direction, stored into the other buffer, as k_mf_block.
template <typename T, bool PQ, bool DIR>
__global__ __launch_bounds__(256) void k_mf_iso(uint32_t n, const uint32_t *__restrict__ dof,
                                                const uint32_t *__restrict__ src, const double *__restrict__ fpen,
                                                const T *x, T *xb, const T *__restrict__ z, T *__restrict__ y,
                                                const ProbScalars *S, double *partials, uint32_t slot0) {
    if (PQ && !selected(S[0], SEL_ACTIVE)) return;
    const bool odd = DIR && (S[0].iter & 1);
    const T *xs = odd ? xb : x;
    T *xo = odd ? const_cast<T *>(x) : xb;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    double acc[1] = {0.0};
    if (i < n) {
        const uint32_t d = dof[i], k = src[i];
        const T xv = DIR ? dir_value(z[d], (T)S[0].beta, xs[d]) : x[d];
        if (DIR) xo[d] = xv;
        const T a = k == 0xffffffffu ? (T)0 : (T)fpen[k] * xv;
        y[d] = a;
        acc[0] = (double)a * (double)xv;
    }
    if (PQ) wg_partial<1, 0u>(acc, partials, 0u, slot0 + blockIdx.x);
}

// Summed penalty per constrained DOF as k_constraints adds it to K (M = max |diag K| of the
// element loop, times each penalty in insertion order).
__global__ void k_mf_pen(uint32_t n_cdof, const uint32_t *__restrict__ cptr, const uint32_t *__restrict__ csub,
                         const double *__restrict__ cpen, const ProbScalars *S, double *__restrict__ fpen) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_cdof) return;
    const double M = S[csub[i]].max_diag;
    double s = 0.0;
    for (uint32_t k = cptr[i]; k < cptr[i + 1]; ++k) s += 1.0 * M * cpen[k];
    fpen[i] = s;
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------

bool mf_eligible(const Problem &P) {
    if (P.from_bsr || P.n_elem == 0) return false;
    return (P.kind == HEXA && P.conn_h.size() == 8 * (size_t)P.n_elem) ||
           (P.kind == TETRA10 && P.conn_h.size() == 10 * (size_t)P.n_elem);
}

void mf_free(MatFree &M) {
    dfree(M.d_conn); dfree(M.d_loc); dfree(M.d_color); dfree(M.d_elem); dfree(M.d_lptr); dfree(M.d_bls);
    dfree(M.d_lnode_pad); dfree(M.d_lslot_pad); dfree(M.d_bnode); dfree(M.d_bptr); dfree(M.d_halo); dfree(M.d_idof);
    dfree(M.d_isrc);
    dfree(M.d_pptr); dfree(M.d_pidx); dfree(M.d_psrc); dfree(M.d_E); dfree(M.d_mat);
    dfree(M.d_fpen); dfree(M.d_partials); dfree(M.d_rec64); dfree(M.d_rec32); dfree(M.d_pb); dfree(M.d_pb32); dfree(M.d_hbeg); dfree(M.d_bsub); dfree(M.d_wg0); dfree(M.d_nwg); dfree(M.d_b0);
    M = MatFree();
}

// An axis-aligned box: every edge along xi_d is (h_d > 0) e_d, with one h_d per axis, compared
// exactly (so the box product sees the coordinate differences of the general one).
bool hexa_is_box(const double *X, const uint32_t *nd) {
    for (int d = 0; d < 3; ++d) {
        double hd = 0;
        for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q) {
                const int a = d == 0 ? HN[0][p][q] : d == 1 ? HN[p][0][q] : HN[p][q][0];
                const int b = d == 0 ? HN[1][p][q] : d == 1 ? HN[p][1][q] : HN[p][q][1];
                for (int c = 0; c < 3; ++c) {
                    const double v = X[3 * (size_t)nd[b] + c] - X[3 * (size_t)nd[a] + c];
                    if (c != d && v != 0.0) return false;
                    if (c == d && (p || q) && v != hd) return false;
                    if (c == d) hd = v;
                }
            }
        if (!(hd > 0)) return false;
    }
    return true;
}

namespace {

// Renumbers the local nodes of block b (lnode[L0..]) by the parity class of their lattice
// position -- per axis the rank of the node's coordinate among the block's distinct ones -- then
// by (z, y, x) rank, and updates the block's corner indices (sloc of slots i0..i0+nk), its
// per-node element counts and the local index of the nodes the block holds first.
void renumber_by_parity(const double *X, uint32_t L0, std::vector<uint32_t> &lnode, std::vector<uint32_t> &in_blk,
                        std::vector<uint16_t> &sloc, size_t i0, size_t nk, uint32_t lst, uint32_t b, MatFree &M) {
    const uint32_t nl = (uint32_t)lnode.size() - L0;
    std::vector<uint32_t> rk(3 * (size_t)nl);
    std::vector<double> v(nl), u;
    for (int d = 0; d < 3; ++d) {
        for (uint32_t i = 0; i < nl; ++i) v[i] = X[3 * (size_t)lnode[L0 + i] + d];
        u = v;
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        for (uint32_t i = 0; i < nl; ++i)
            rk[3 * (size_t)i + d] = (uint32_t)(std::lower_bound(u.begin(), u.end(), v[i]) - u.begin());
    }
    auto key = [&](uint32_t i) {
        const uint32_t *r = &rk[3 * (size_t)i];
        const uint32_t cls = (r[0] & 1u) | (r[1] & 1u) << 1 | (r[2] & 1u) << 2;
        return std::make_tuple(cls, r[2], r[1], r[0]);
    };
    std::vector<uint32_t> ord(nl), nid(nl);
    for (uint32_t i = 0; i < nl; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](uint32_t p, uint32_t q) { return key(p) < key(q); });
    for (uint32_t t = 0; t < nl; ++t) nid[ord[t]] = t;
    const std::vector<uint32_t> ln(lnode.begin() + L0, lnode.end()), ib(in_blk.begin() + L0, in_blk.end());
    for (uint32_t t = 0; t < nl; ++t) {
        lnode[L0 + t] = ln[ord[t]];
        in_blk[L0 + t] = ib[ord[t]];
    }
    for (size_t k = 0; k < nk; ++k)
        for (uint32_t a = 0; a < 8; ++a) {
            uint16_t &l = sloc[lst * (i0 + k) + a];
            l = (uint16_t)nid[l];
        }
    for (uint32_t i = 0; i < nl; ++i)
        if (M.node_blk[ln[i]] == b) M.node_loc[ln[i]] = nid[i];
}

// 21-bit-per-axis Morton code (bit 3k + d holds bit k of axis d).
uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t m = 0;
    for (int k = 0; k < 21; ++k)
        m |= ((uint64_t)(x >> k & 1u) << (3 * k)) | ((uint64_t)(y >> k & 1u) << (3 * k + 1)) |
             ((uint64_t)(z >> k & 1u) << (3 * k + 2));
    return m;
}

template <class V>
bool upload_vec(V **d, const std::vector<V> &h) {
    if (hipMalloc(reinterpret_cast<void **>(d), std::max<size_t>(1, h.size()) * sizeof(V)) != hipSuccess) return false;
    return h.empty() || hipMemcpy(*d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice) == hipSuccess;
}

}  // namespace

int mf_build(Problem &P) {
    MatFree &M = P.mf;
    if (M.built) return 0;
    if (!mf_eligible(P)) return 1;
    const uint32_t ne = P.n_elem, nn = P.n_nodes;
    const uint32_t *conn = P.conn_h.data();
    // nodes per element, and the stride of the block-local index lists (16 u16 per Tetra_10 slot,
    // so a slot's list is two aligned 16-byte loads)
    const int npe = P.kind == TETRA10 ? 10 : 8, lst = P.kind == TETRA10 ? 16 : 8;
    std::vector<double> X(3 * (size_t)nn);
    if (hipMemcpy(X.data(), P.d_coords, X.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) return 5;
    // element order, per sub-problem (a block never straddles two): Morton code of the element
    // centre on a grid of the sub-problem's typical element size (on a lattice, one cell per
    // element, so aligned blocks are 8 x 8 x 4 bricks), ties by element id
    std::vector<std::pair<uint64_t, uint32_t>> key(ne);
    std::vector<uint32_t> bstart, bcount, bsub;  // per block: first sorted element, count, sub-problem
    for (uint32_t sp = 0; sp < P.n_sub; ++sp) {
        const uint32_t e0 = P.elem_off[sp], e1 = P.elem_off[sp + 1], n0 = P.node_off[sp], n1 = P.node_off[sp + 1];
        if (e1 == e0) continue;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (uint32_t n = n0; n < n1; ++n)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], X[3 * (size_t)n + d]);
                hi[d] = std::max(hi[d], X[3 * (size_t)n + d]);
            }
        double vol = 1.0, ext = 0.0;
        for (int d = 0; d < 3; ++d) {
            ext = std::max(ext, hi[d] - lo[d]);
            vol *= std::max(hi[d] - lo[d], 0.0);
        }
        double h = vol > 0 ? std::cbrt(vol / (e1 - e0)) : ext / std::cbrt((double)(e1 - e0));
        if (!(h > 0)) h = 1.0;
        for (uint32_t e = e0; e < e1; ++e) {
            uint32_t q[3];
            for (int d = 0; d < 3; ++d) {
                double c = 0;
                for (int a = 0; a < npe; ++a) c += X[3 * (size_t)conn[npe * (size_t)e + a] + d];
                q[d] = (uint32_t)std::min(2097151.0, std::max(0.0, std::floor((c / npe - lo[d]) / h)));
            }
            key[e] = {morton3(q[0], q[1], q[2]), e};
        }
        std::sort(key.begin() + e0, key.begin() + e1);
        for (uint32_t k = e0; k < e1; k += MF_BLOCK) {
            bstart.push_back(k);
            bcount.push_back(std::min(MF_BLOCK, e1 - k));
            bsub.push_back(sp);
        }
    }
    // blocks of up to MF_BLOCK elements of one sub-problem, padded to a multiple of 8 blocks (the
    // XCD mapping) with empty blocks of no sub-problem
    const uint32_t nb = ((uint32_t)bstart.size() + 7) / 8 * 8;
    const size_t ns = (size_t)nb * MF_BLOCK;
    while (bstart.size() < nb) {
        bstart.push_back(0);
        bcount.push_back(0);
        bsub.push_back(NO_SUB);
    }
    std::vector<uint32_t> sconn(npe * ns, 0), selem(ns, 0), lptr(nb + 1, 0), lnode, deg(nn, 0), stamp(nn, 0xffffffffu),
        lid(nn, 0), in_blk;
    M.node_blk.assign(nn, 0xffffffffu);
    M.node_loc.assign(nn, 0);
    std::vector<uint16_t> sloc(lst * ns, 0);
    std::vector<uint8_t> scol(ns, 0xffu);
    for (size_t s = 0; s < npe * (size_t)ne; ++s) ++deg[conn[s]];
    uint32_t n_colors = 0, max_local = 0;
    // element centres, for the sweep order of the sub-colouring
    std::vector<double> cen(3 * (size_t)ne, 0.0);
    for (uint32_t e = 0; e < ne; ++e)
        for (int a = 0; a < npe; ++a)
            for (int d = 0; d < 3; ++d) cen[3 * (size_t)e + d] += X[3 * (size_t)conn[npe * (size_t)e + a] + d] / npe;
    std::vector<uint32_t> blk;
    for (uint32_t b = 0; b < nb; ++b) {
        lptr[b] = (uint32_t)lnode.size();
        std::vector<uint32_t> mask;
        // the block's elements swept by (z, y, x) of their centres: greedy first-fit then finds the
        // 8-colour parity pattern on lattices (the Morton order alone needs more sub-colours)
        blk.clear();
        for (uint32_t k = 0; k < bcount[b]; ++k) blk.push_back(key[(size_t)bstart[b] + k].second);
        std::stable_sort(blk.begin(), blk.end(), [&](uint32_t p, uint32_t q) {
            for (int d = 2; d >= 0; --d)
                if (cen[3 * (size_t)p + d] != cen[3 * (size_t)q + d]) return cen[3 * (size_t)p + d] < cen[3 * (size_t)q + d];
            return p < q;
        });
        for (uint32_t k = 0; k < (uint32_t)blk.size(); ++k) {
            const size_t i = (size_t)b * MF_BLOCK + k;
            const uint32_t e = blk[k];
            uint32_t used = 0;
            for (int a = 0; a < npe; ++a) {
                const uint32_t n = conn[npe * (size_t)e + a];
                if (stamp[n] != b) {
                    stamp[n] = b;
                    lid[n] = (uint32_t)(lnode.size() - lptr[b]);
                    lnode.push_back(n);
                    in_blk.push_back(0);
                    mask.push_back(0);
                    if (M.node_blk[n] == 0xffffffffu) {
                        M.node_blk[n] = b;
                        M.node_loc[n] = lid[n];
                    }
                }
                sconn[npe * i + a] = n;
                sloc[lst * i + a] = (uint16_t)lid[n];
                ++in_blk[lptr[b] + lid[n]];
                used |= mask[lid[n]];
            }
            if (used == 0xffffffffu) return 1;  // more than 32 sub-colours: not served
            const uint32_t c = (uint32_t)__builtin_ctz(~used);
            for (int a = 0; a < npe; ++a) mask[lid[conn[npe * (size_t)e + a]]] |= 1u << c;
            scol[i] = (uint8_t)(c | (P.kind == HEXA && hexa_is_box(X.data(), conn + 8 * (size_t)e) ? MF_BOX : 0u));
            selem[i] = e;
            n_colors = std::max(n_colors, c + 1);
        }
        // slots by sub-colour (stable): a wave then holds 64 elements of one or two sub-colours
        // instead of a (z, y, x) layer spanning four, so it takes part in fewer accumulation phases
        // with more of its lanes; every entry still receives its adds in sub-colour order
        {
            const size_t i0 = (size_t)b * MF_BLOCK, nk = blk.size();
            std::vector<uint32_t> ord(nk);
            for (uint32_t k = 0; k < nk; ++k) ord[k] = k;
            std::stable_sort(ord.begin(), ord.end(), [&](uint32_t p, uint32_t q) {
                return (scol[i0 + p] & MF_COLOR_BITS) < (scol[i0 + q] & MF_COLOR_BITS);
            });
            const std::vector<uint32_t> c0(sconn.begin() + npe * i0, sconn.begin() + npe * (i0 + nk)),
                e0(selem.begin() + i0, selem.begin() + i0 + nk);
            const std::vector<uint16_t> l0(sloc.begin() + lst * i0, sloc.begin() + lst * (i0 + nk));
            const std::vector<uint8_t> s0(scol.begin() + i0, scol.begin() + i0 + nk);
            for (size_t k = 0; k < nk; ++k) {
                const size_t src = ord[k];
                for (int a = 0; a < npe; ++a) {
                    sconn[npe * (i0 + k) + a] = c0[npe * src + a];
                    sloc[lst * (i0 + k) + a] = l0[lst * src + a];
                }
                scol[i0 + k] = s0[src];
                selem[i0 + k] = e0[src];
            }
        }
        // a block of box elements (a brick of a lattice) numbers its local nodes by lattice parity
        // class, then (z, y, x): the lanes of one sub-colour reach, at one corner, nodes of one
        // class two lattice steps apart, which then lie at consecutive triples of the staged
        // input and the accumulator and so in distinct LDS banks (first-appearance numbering
        // put them 6 words apart, two to four lanes per bank)
        if (P.kind == HEXA && !blk.empty()) {
            const size_t i0 = (size_t)b * MF_BLOCK, nk = blk.size();
            bool boxes = true;
            for (size_t k = 0; k < nk && boxes; ++k) boxes = (scol[i0 + k] & MF_BOX) != 0;
            if (boxes) renumber_by_parity(X.data(), lptr[b], lnode, in_blk, sloc, i0, nk, lst, b, M);
        }
        max_local = std::max<uint32_t>(max_local, (uint32_t)lnode.size() - lptr[b]);
    }
    lptr[nb] = (uint32_t)lnode.size();
    const uint32_t ls = (max_local + 63) / 64 * 64;  // fixed stride of the blocks' node lists
    // per block its local node count and accumulator copies; the launches' LDS, the largest block's
    std::vector<uint32_t> bls(nb);
    size_t lds32 = 0, lds64 = 0;
    const uint32_t gmax32 = P.kind == TETRA10 ? mf_acc_log2_max<float, TETRA10>() : mf_acc_log2_max<float, HEXA>();
    const uint32_t gmax64 = P.kind == TETRA10 ? mf_acc_log2_max<double, TETRA10>() : mf_acc_log2_max<double, HEXA>();
    bool uni32 = true, uni64 = true;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t m = lptr[b + 1] - lptr[b], g32 = mf_acc_log2(m, 4, P.kind), g64 = mf_acc_log2(m, 8, P.kind);
        bls[b] = m | g32 << 24 | g64 << 28;
        lds32 = std::max(lds32, mf_lds_bytes(m, 4, g32));
        lds64 = std::max(lds64, mf_lds_bytes(m, 8, g64));
        uni32 &= g32 == gmax32;
        uni64 &= g64 == gmax64;
    }
    if (lds64 > 64 * 1024) return 1;  // LDS of the FP64 product's blocks
    // shared nodes: a partial slot per (node, block), grouped by node in block order
    std::vector<uint32_t> nshared(nn, 0);
    for (size_t l = 0; l < lnode.size(); ++l)
        if (in_blk[l] != deg[lnode[l]]) ++nshared[lnode[l]];
    std::vector<uint32_t> bnode, bptr(1, 0), first(nn, 0xffffffffu), idof;
    for (uint32_t n = 0; n < nn; ++n) {
        if (!deg[n])
            for (uint32_t d = 0; d < 3; ++d) idof.push_back(3 * n + d);
        if (!nshared[n]) continue;
        first[n] = bptr.back();
        bnode.push_back(n);
        bptr.push_back(bptr.back() + nshared[n]);
    }
    // nodes of no element take a pass of their own, whose partials follow the blocks': served for
    // a single problem only
    if (P.n_sub > 1 && !idof.empty()) return 1;
    std::vector<uint32_t> hbeg(first.begin(), first.end());  // first[n]: node n's first partial slot
    hbeg.push_back(bptr.back());
    for (uint32_t n = nn; n-- > 0;)
        if (hbeg[n] == 0xffffffffu) hbeg[n] = hbeg[n + 1];  // no partials: an empty range
    std::vector<uint32_t> lslot(lnode.size(), 0xffffffffu);
    for (uint32_t b = 0; b < nb; ++b)  // blocks in order, so slots follow block order
        for (uint32_t l = lptr[b]; l < lptr[b + 1]; ++l)
            if (in_blk[l] != deg[lnode[l]])
                lslot[l] = first[lnode[l]]++ | (M.node_blk[lnode[l]] == b ? MF_FIRST_HOLDER : 0u);
    if (bptr.back() > MF_SLOT_BITS) return 1;  // partial slots beyond the 31-bit slot field
    std::vector<uint32_t> lnode_pad((size_t)nb * ls, MF_NO_NODE), lslot_pad((size_t)nb * ls, 0xffffffffu);
    for (uint32_t b = 0; b < nb; ++b)
        for (uint32_t l = lptr[b]; l < lptr[b + 1]; ++l) {
            lnode_pad[(size_t)b * ls + (l - lptr[b])] = lnode[l];
            lslot_pad[(size_t)b * ls + (l - lptr[b])] = lslot[l];
        }
    const std::vector<uint32_t> zero(1, 0);
    if (!upload_vec(&M.d_conn, sconn) || !upload_vec(&M.d_loc, sloc) || !upload_vec(&M.d_color, scol) ||
        !upload_vec(&M.d_elem, selem) || !upload_vec(&M.d_lptr, lptr) || !upload_vec(&M.d_bls, bls) || !upload_vec(&M.d_lnode_pad, lnode_pad) ||
        !upload_vec(&M.d_lslot_pad, lslot_pad) || !upload_vec(&M.d_hbeg, hbeg) || !upload_vec(&M.d_bnode, bnode) || !upload_vec(&M.d_bptr, bptr) ||
        !upload_vec(&M.d_idof, idof) || !upload_vec(&M.d_isrc, std::vector<uint32_t>(idof.size(), 0xffffffffu)) ||
        !upload_vec(&M.d_bsub, bsub) || !upload_vec(&M.d_wg0, std::vector<uint32_t>(P.n_sub, 0)) ||
        !upload_vec(&M.d_nwg, std::vector<uint32_t>(P.n_sub, 0)) ||
        hipMalloc(&M.d_rec64, (P.kind == TETRA10 ? sizeof(TetRec<double>) : sizeof(BoxRec<double>)) * ns) != hipSuccess ||
        hipMalloc(&M.d_pb, sizeof(double) * 3 * (size_t)nn) != hipSuccess ||
        hipMalloc(&M.d_pb32, sizeof(float) * 3 * (size_t)nn) != hipSuccess ||
        hipMalloc(&M.d_rec32, (P.kind == TETRA10 ? sizeof(TetRec<float>) : sizeof(BoxRec<float>)) * ns) != hipSuccess ||
        hipMalloc(&M.d_halo, sizeof(double) * 3 * std::max<size_t>(1, bptr.back())) != hipSuccess) {
        mf_free(M);
        return 5;
    }
    M.n_blocks = nb;
    M.n_colors = n_colors;
    M.n_boxes = (uint32_t)std::count_if(scol.begin(), scol.end(), [](uint8_t c) { return c != MF_EMPTY && (c & MF_BOX); });
    M.all_boxes = M.n_boxes == ne;
    M.max_local = max_local;
    M.lds32 = lds32;
    M.lds64 = lds64;
    M.uni32 = uni32;
    M.uni64 = uni64;
    M.ls = ls;
    M.n_bnodes = (uint32_t)bnode.size();
    M.n_halo = bptr.back();
    M.n_iso = (uint32_t)idof.size() / 3;
    // each sub-problem's blocks are consecutive: their p.q partial slots
    M.sub_b0.assign(P.n_sub, 0);
    M.sub_nb.assign(P.n_sub, 0);
    for (uint32_t b = nb; b-- > 0;)
        if (bsub[b] != NO_SUB) {
            M.sub_b0[bsub[b]] = b;
            ++M.sub_nb[bsub[b]];
        }
    // the two-kernel iteration needs a block of every sub-problem to carry its scalars
    M.fold = std::find(M.sub_nb.begin(), M.sub_nb.end(), 0u) == M.sub_nb.end();
    if (!upload_vec(&M.d_b0, M.sub_b0)) {
        mf_free(M);
        return 5;
    }
    M.built = true;
    return 0;
}

hipError_t mf_snapshot(Problem &P) {
    if (!mf_eligible(P)) return hipSuccess;
    MatFree &M = P.mf;
    hipError_t e = hipSuccess;
    if (!M.d_E) e = hipMalloc(&M.d_E, sizeof(double) * P.n_elem);
    if (e == hipSuccess && !M.d_mat) e = hipMalloc(&M.d_mat, sizeof(SubMaterial) * P.n_sub);
    if (e == hipSuccess && P.n_cdof > M.cap_fix) {
        dfree(M.d_fpen);
        e = hipMalloc(&M.d_fpen, sizeof(double) * P.n_cdof);
        M.cap_fix = e == hipSuccess ? P.n_cdof : 0;
        destroy_graph(P);  // captured products read the penalties through this pointer
        M.pen_built = false;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(M.d_E, P.d_E, sizeof(double) * P.n_elem, hipMemcpyDeviceToDevice, P.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(M.d_mat, P.d_mat, sizeof(SubMaterial) * P.n_sub, hipMemcpyDeviceToDevice, P.stream);
    if (e == hipSuccess && P.n_cdof)
        hipLaunchKernelGGL(k_mf_pen, dim3((P.n_cdof + 255) / 256), dim3(256), 0, P.stream, P.n_cdof, P.d_cptr, P.d_csub,
                           P.d_cpen, P.d_ps, M.d_fpen);
    if (e != hipSuccess) return e;
    M.snap_dofs = P.cdof_h;
    M.n_fix = P.n_cdof;
    M.snap = true;
    M.rec_valid = false;
    return hipGetLastError();
}

// Penalty tables for the DOF list of the last snapshot: each constrained DOF goes to the first
// block holding its node (or to the nodes-of-no-element pass).
static int mf_penalties(Problem &P) {
    MatFree &M = P.mf;
    if (M.pen_built && M.pen_dofs == M.snap_dofs) return 0;
    std::vector<std::vector<uint32_t>> per(M.n_blocks);
    std::vector<uint32_t> isrc(3 * (size_t)M.n_iso, 0xffffffffu);
    std::vector<uint32_t> idof(3 * (size_t)M.n_iso);
    if (M.n_iso && hipMemcpy(idof.data(), M.d_idof, idof.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return 5;
    for (uint32_t i = 0; i < (uint32_t)M.snap_dofs.size(); ++i) {
        const uint32_t dof = M.snap_dofs[i], n = dof / 3;
        if (M.node_blk[n] == 0xffffffffu) {
            const size_t k = std::lower_bound(idof.begin(), idof.end(), dof) - idof.begin();
            if (k < idof.size() && idof[k] == dof) isrc[k] = i;
        } else {
            per[M.node_blk[n]].push_back(i);
        }
    }
    std::vector<uint32_t> pptr(1, 0), pidx, psrc;
    for (uint32_t b = 0; b < M.n_blocks; ++b) {
        for (uint32_t i : per[b]) {
            const uint32_t dof = M.snap_dofs[i], n = dof / 3;
            pidx.push_back(MF_NS * M.node_loc[n] + dof % 3);
            psrc.push_back(i);
        }
        pptr.push_back((uint32_t)pidx.size());
    }
    dfree(M.d_pptr); dfree(M.d_pidx); dfree(M.d_psrc);
    if (!upload_vec(&M.d_pptr, pptr) || !upload_vec(&M.d_pidx, pidx) ||
        !upload_vec(&M.d_psrc, psrc) ||
        (M.n_iso && hipMemcpy(M.d_isrc, isrc.data(), isrc.size() * 4, hipMemcpyHostToDevice) != hipSuccess))
        return 5;
    M.pen_dofs = M.snap_dofs;
    M.pen_built = true;
    destroy_graph(P);  // captured products hold the old table pointers
    return 0;
}

// Partial slots of the passes and the finalize's range, penalty tables; allocations happen here,
// never while a chunk of iterations is being captured.
int mf_prepare(Problem &P) {
    int rc = mf_build(P);
    if (rc) return rc;
    MatFree &M = P.mf;
    if (!M.snap) return 1;
    if (M.n_fix && !M.d_fpen) return 5;
    rc = mf_penalties(P);
    if (rc) return rc;
    if (!M.rec_valid && P.kind == TETRA10) {  // stream order: after the snapshot's copies
        const size_t ns = (size_t)M.n_blocks * MF_BLOCK;
        hipLaunchKernelGGL(k_mf_tetrec, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, P.stream, ns, M.d_color, M.d_conn,
                           M.d_elem, M.d_bsub, P.d_coords, M.d_E, M.d_mat, reinterpret_cast<TetRec<double> *>(M.d_rec64),
                           reinterpret_cast<TetRec<float> *>(M.d_rec32));
        if (hipGetLastError() != hipSuccess) return 5;
    } else if (!M.rec_valid && M.n_boxes) {  // stream order: after the snapshot's copies
        const size_t ns = (size_t)M.n_blocks * MF_BLOCK;
        hipLaunchKernelGGL(k_mf_boxrec, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, P.stream, ns, M.d_color, M.d_conn,
                           M.d_elem, M.d_bsub, P.d_coords, M.d_E, M.d_mat, reinterpret_cast<BoxRec<double> *>(M.d_rec64),
                           reinterpret_cast<BoxRec<float> *>(M.d_rec32));
        if (hipGetLastError() != hipSuccess) return 5;
    }
    M.rec_valid = true;
    const uint32_t s = M.n_blocks + (3 * M.n_iso + 255) / 256;
    if (s > M.cap_slots) {
        dfree(M.d_partials);
        if (hipMalloc(&M.d_partials, sizeof(double) * std::max<uint32_t>(1, s)) != hipSuccess) {
            M.cap_slots = 0;
            return 5;
        }
        M.cap_slots = s;
        destroy_graph(P);
    }
    if (s != M.n_slots) {
        // a single problem: every slot (blocks, then the nodes-of-no-element pass); a batch: each
        // sub-problem its blocks
        std::vector<uint32_t> w0(P.n_sub, 0), nw(P.n_sub, s);
        if (P.n_sub > 1) {
            w0 = M.sub_b0;
            nw = M.sub_nb;
        }
        if (hipMemcpy(M.d_wg0, w0.data(), 4 * w0.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(M.d_nwg, nw.data(), 4 * nw.size(), hipMemcpyHostToDevice) != hipSuccess)
            return 5;
        M.n_slots = s;
    }
    return 0;
}

// Element-block kernel of one variant (BOXES: every element of the mesh is an axis-aligned box;
// Tetra_10 problems take the Tetra_10 kernel).
template <typename T, bool PQ, bool DIR, bool BOXES, bool FIN = false>
static void mf_block_launch(Problem &P, const T *x, T *xb, const T *z, T *y, T *halo, const FinParts *fin = nullptr) {
    const MatFree &M = P.mf;
    const void *rec = sizeof(T) == 4 ? M.d_rec32 : M.d_rec64;
    const FinParts F = fin ? *fin : FinParts{nullptr, 0u, nullptr, nullptr};
#define FEMGPU_MF_BLOCK(KIND, UNI)                                                                                    \
    hipLaunchKernelGGL((k_mf_block<T, PQ, DIR, BOXES && KIND == HEXA, KIND, FIN, UNI>), dim3(M.n_blocks), dim3(MF_BLOCK), \
                       sizeof(T) == 4 ? M.lds32 : M.lds64, P.stream, M.d_conn, M.d_loc,                                 \
                       M.d_color, M.d_elem, M.d_bsub, M.n_colors, M.ls, M.d_bls, M.max_local, M.d_lnode_pad, M.d_lslot_pad, M.d_pptr, M.d_pidx, M.d_psrc, M.d_fpen,         \
                       P.d_coords, M.d_E, M.d_mat, rec, x, xb, z, y, halo, P.d_ps,                                      \
                       M.d_partials, P.d_ps + P.n_sub, F.part, F.stride, F.w0, F.nw, M.d_b0)
    if (P.kind == TETRA10) FEMGPU_MF_BLOCK(TETRA10, true);
    else if (sizeof(T) == 4 ? M.uni32 : M.uni64) FEMGPU_MF_BLOCK(HEXA, true);
    else FEMGPU_MF_BLOCK(HEXA, false);
#undef FEMGPU_MF_BLOCK
}

// halo: the shared-node partial slots (3 values of T each), the problem's own buffer or a caller's
// FIN: the two-kernel CG iteration (the nodes-of-no-element pass then reads the scalars the
// element blocks advanced)
template <typename T, bool PQ, bool DIR, bool FIN = false>
static void mf_passes(Problem &P, const T *x, T *xb, const T *z, T *y, T *halo, bool halo_pass = true,
                      const FinParts *fin = nullptr) {
    const MatFree &M = P.mf;
    if (M.all_boxes) mf_block_launch<T, PQ, DIR, true, FIN>(P, x, xb, z, y, halo, fin);
    else mf_block_launch<T, PQ, DIR, false, FIN>(P, x, xb, z, y, halo, fin);
    if (M.n_iso)
        hipLaunchKernelGGL((k_mf_iso<T, PQ, DIR>), dim3((3 * M.n_iso + 255) / 256), dim3(256), 0, P.stream, 3 * M.n_iso,
                           M.d_idof, M.d_isrc, M.d_fpen, x, xb, z, y, FIN ? P.d_ps + P.n_sub : P.d_ps, M.d_partials,
                           M.n_blocks);
    // DIR (or !halo_pass): the caller's vector kernel sums the shared nodes' partials itself
    // (k_update_mf, k_cheb)
    if (M.n_bnodes && !DIR && halo_pass)
        hipLaunchKernelGGL(k_mf_halo<T>, dim3((M.n_bnodes + 255) / 256), dim3(256), 0, P.stream, M.n_bnodes, M.d_bnode,
                           M.d_bptr, halo, y);
}

// y = K x; with z (PQ only), y = K p for the updated direction p = z + beta s, s the stored
// direction in x or in the problem's second direction buffer by the parity of the iteration
// count, p stored into the other one (the direction update of the CG step, fused into the
// product's passes); then y holds the nodes of no shared block only, and the shared nodes'
// partials are left to the caller's vector update (k_update_mf).
// fin (with z): the two-kernel CG iteration, the previous vector update's end folded in
// (k_mf_block FIN, combining the partials fin describes); the problem must have P.mf.fold.
template <typename T>
static hipError_t mf_launch_t(Problem &P, const T *x, T *y, bool pq, const T *z, void *halo,
                              const FinParts *fin = nullptr) {
    T *xb = reinterpret_cast<T *>(sizeof(T) == 4 ? (void *)P.mf.d_pb32 : (void *)P.mf.d_pb);
    T *h = reinterpret_cast<T *>(halo ? halo : (void *)P.mf.d_halo);
    if (pq && z && fin) mf_passes<T, true, true, true>(P, x, xb, z, y, h, true, fin);
    else if (pq && z) mf_passes<T, true, true>(P, x, xb, z, y, h);
    else if (pq) mf_passes<T, true, false>(P, x, nullptr, nullptr, y, h);
    else mf_passes<T, false, false>(P, x, nullptr, nullptr, y, h);
    // peek, not get: a failed launch stays pending for the caller's own check (the solver
    // tests hipGetLastError after each chunk and at the end of the solve)
    return hipPeekAtLastError();
}

// Average device time of the element-block kernel alone and of the whole product over reps launches
// each (after one untimed product), on the problem's stream.
int mf_time(Problem &P, int precision, int reps, double *block_ms, double *product_ms) {
    const int rc = mf_prepare(P);
    if (rc) return rc;
    const MatFree &M = P.mf;
    auto block = [&]() {
        if (precision) {
            float *h = reinterpret_cast<float *>(M.d_halo);
            if (M.all_boxes) mf_block_launch<float, false, false, true>(P, P.d_p32, nullptr, nullptr, P.d_q32, h);
            else mf_block_launch<float, false, false, false>(P, P.d_p32, nullptr, nullptr, P.d_q32, h);
        } else {
            double *h = M.d_halo;
            if (M.all_boxes) mf_block_launch<double, false, false, true>(P, P.d_p, nullptr, nullptr, P.d_q, h);
            else mf_block_launch<double, false, false, false>(P, P.d_p, nullptr, nullptr, P.d_q, h);
        }
    };
    auto product = [&]() {
        if (precision) (void)mf_launch_t<float>(P, P.d_p32, P.d_q32, false, nullptr, nullptr);
        else (void)mf_launch_t<double>(P, P.d_p, P.d_q, false, nullptr, nullptr);
    };
    float ms[2] = {0, 0};
    product();
    for (int k = 0; k < 2; ++k) {
        if (hipEventRecord(P.ev[2], P.stream) != hipSuccess) return 5;
        for (int i = 0; i < reps; ++i) k ? product() : block();
        if (hipEventRecord(P.ev[3], P.stream) != hipSuccess || hipEventSynchronize(P.ev[3]) != hipSuccess ||
            hipEventElapsedTime(&ms[k], P.ev[2], P.ev[3]) != hipSuccess)
            return 5;
    }
    if (hipGetLastError() != hipSuccess) return 5;
    if (block_ms) *block_ms = (double)ms[0] / reps;
    if (product_ms) *product_ms = (double)ms[1] / reps;
    return 0;
}

// y = K x leaving the shared nodes' partials in the halo buffer for the caller's vector kernel
// (pq: on the active sub-problems, with the p.q partials; else on every sub-problem).
template <typename T>
static hipError_t mf_launch_partial_t(Problem &P, const T *x, T *y, bool pq) {
    T *h = reinterpret_cast<T *>(P.mf.d_halo);
    if (pq) mf_passes<T, true, false>(P, x, nullptr, nullptr, y, h, false);
    else mf_passes<T, false, false>(P, x, nullptr, nullptr, y, h, false);
    return hipPeekAtLastError();
}
hipError_t mf_launch_partial(Problem &P, const float *x, float *y, bool pq) { return mf_launch_partial_t<float>(P, x, y, pq); }
hipError_t mf_launch_partial(Problem &P, const double *x, double *y, bool pq) { return mf_launch_partial_t<double>(P, x, y, pq); }

hipError_t mf_launch(Problem &P, const float *x, float *y, bool pq, const float *z, void *halo, const FinParts *fin) {
    return mf_launch_t<float>(P, x, y, pq, z, halo, fin);
}
hipError_t mf_launch(Problem &P, const double *x, double *y, bool pq, const double *z, void *halo,
                     const FinParts *fin) {
    return mf_launch_t<double>(P, x, y, pq, z, halo, fin);
}

bool mf_ready(const Problem &P) {
    const MatFree &M = P.mf;
    return M.built && M.snap && M.rec_valid && M.pen_built && M.pen_dofs == M.snap_dofs;
}

size_t mf_halo_values(const Problem &P) { return 3 * (size_t)P.mf.n_halo; }

}  // namespace femgpu
