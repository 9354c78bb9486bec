/*
 * fem_oracle.c -- CPU restatement of the LMT++ assemble + CG path (test infrastructure).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this code, and only as the checker or the timed CPU baseline.  See
 * fem_oracle.h for the per-function pinning status.
 *
 * Reference citations are relative to the LMT++ source tree.
 */
#include "fem_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

/* Host threads of the parallel paths (or_set_threads); 1 = the serial restatement. */
static int g_threads = 1;
int or_set_threads(int n) { g_threads = n > 0 ? n : omp_get_max_threads(); return g_threads; }

/* ------------------------------------------------------------------------- */
/* Element tables                                                             */
/* ------------------------------------------------------------------------- */

int or_nodes_per_elem(int kind) { return kind == OR_TETRA ? 4 : kind == OR_TETRA10 ? 10 : kind == OR_HEXA ? 8 : 0; }
int or_nodes_per_face(int fkind) { return fkind == OR_TRI3 ? 3 : fkind == OR_TRI6 ? 6 : fkind == OR_QUAD4 ? 4 : 0; }
int or_face_kind_of(int kind) { return kind == OR_TETRA ? OR_TRI3 : kind == OR_TETRA10 ? OR_TRI6 : OR_QUAD4; }

/* Volume quadrature used by the runtime Gauss loop (formulation.h:2917-2933)
 * with order_integration = 2*degree-2 (formal_lf/formulation.py:30-32,121-123):
 *   Tetra    order 0: include/mesh/element_Tetra.h:8          (1 point)
 *   Tetra_10 order 2: include/mesh/element_Tetra.h:12-15 table (4 points)
 *   Hexa     order 2: include/mesh/element_Hexa.h:12-19      (2x2x2, 12 significant digits) */
static const double GA_HEXA_LO = 0.211324865405, GA_HEXA_HI = 0.788675134595;
static const double GA_T10_A = 0.1381966011250105, GA_T10_B = 0.5854101966249685;

static int gauss_points(int kind, double *w, double (*xi)[3])
{
    if (kind == OR_TETRA) {
        w[0] = 0.16666666666666666667; xi[0][0] = xi[0][1] = xi[0][2] = 0.25;
        return 1;
    }
    if (kind == OR_TETRA10) {
        const double a = GA_T10_A, b = GA_T10_B;
        const double p[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
        for (int g = 0; g < 4; ++g) { w[g] = 0.041666666666666666667; for (int d = 0; d < 3; ++d) xi[g][d] = p[g][d]; }
        return 4;
    }
    if (kind == OR_HEXA) {
        const double a = GA_HEXA_LO, b = GA_HEXA_HI;
        const double p[8][3] = {{a, a, a}, {b, a, a}, {b, b, a}, {a, b, a}, {a, a, b}, {b, a, b}, {b, b, b}, {a, b, b}};
        for (int g = 0; g < 8; ++g) { w[g] = 0.125; for (int d = 0; d < 3; ++d) xi[g][d] = p[g][d]; }
        return 8;
    }
    return 0;
}

/* Shape functions and reference-space derivatives.
 *   Tetra    formulations/element_Tetra.py:27-30
 *   Tetra_10 formulations/element_Tetra_10.py:27-41 (nodes 4..9 on edges 01,12,02,03,13,23)
 *   Hexa     formulations/element_Hexa.py:61-68 */
static void shape(int kind, const double *p, double *N, double (*dN)[3])
{
    const double x = p[0], y = p[1], z = p[2];
    if (kind == OR_TETRA) {
        N[0] = 1 - x - y - z; N[1] = x; N[2] = y; N[3] = z;
        const double d[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        memcpy(dN, d, sizeof d);
    } else if (kind == OR_TETRA10) {
        const double L0 = 1 - x - y - z, L1 = x, L2 = y, L3 = z;
        const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        const double L[4] = {L0, L1, L2, L3};
        for (int a = 0; a < 4; ++a) {
            N[a] = L[a] * (2 * L[a] - 1);
            for (int d = 0; d < 3; ++d) dN[a][d] = (4 * L[a] - 1) * dL[a][d];
        }
        const int e[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
        for (int k = 0; k < 6; ++k) {
            const int i = e[k][0], j = e[k][1];
            N[4 + k] = 4 * L[i] * L[j];
            for (int d = 0; d < 3; ++d) dN[4 + k][d] = 4 * (dL[i][d] * L[j] + L[i] * dL[j][d]);
        }
    } else {
        const double X[2] = {1 - x, x}, Y[2] = {1 - y, y}, Z[2] = {1 - z, z};
        const double dX[2] = {-1, 1};
        const int ix[8] = {0, 1, 1, 0, 0, 1, 1, 0}, iy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, iz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
        for (int a = 0; a < 8; ++a) {
            N[a] = X[ix[a]] * Y[iy[a]] * Z[iz[a]];
            dN[a][0] = dX[ix[a]] * Y[iy[a]] * Z[iz[a]];
            dN[a][1] = X[ix[a]] * dX[iy[a]] * Z[iz[a]];
            dN[a][2] = X[ix[a]] * Y[iy[a]] * dX[iz[a]];
        }
    }
}

/* J[j][i] = dx_i/dxi_j (formal_lf/element.py:195-205), signed determinant, inverse. */
static double jacobian(int npe, const double *xe, double (*dN)[3], double J[3][3], double Ji[3][3])
{
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            double s = 0;
            for (int a = 0; a < npe; ++a) s += dN[a][j] * xe[3 * a + i];
            J[j][i] = s;
        }
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    if (Ji) {
        const double id = 1.0 / det;
        Ji[0][0] = c00 * id;
        Ji[1][0] = c01 * id;
        Ji[2][0] = c02 * id;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    }
    return det;
}

/* permutation_if_jac_neg: det J evaluated at xi = (1/2,1/2,1/2) (make_elem.py:28-45);
 * node permutation from formulations/element_<E>.py "permutation". */
static void permutation_if_jac_neg(int kind, const double *coords, uint32_t *nodes)
{
    const int npe = or_nodes_per_elem(kind);
    double xe[30], N[10], dN[10][3], J[3][3];
    for (int a = 0; a < npe; ++a)
        for (int d = 0; d < 3; ++d) xe[3 * a + d] = coords[3 * (size_t)nodes[a] + d];
    const double half[3] = {0.5, 0.5, 0.5};
    shape(kind, half, N, dN);
    if (jacobian(npe, xe, dN, J, NULL) >= 0) return;
    static const int p_tet[4] = {0, 2, 1, 3};
    static const int p_t10[10] = {0, 1, 3, 2, 4, 8, 7, 6, 5, 9};
    static const int p_hex[8] = {0, 3, 2, 1, 4, 7, 6, 5};
    const int *p = kind == OR_TETRA ? p_tet : kind == OR_TETRA10 ? p_t10 : p_hex;
    uint32_t tmp[10];
    for (int a = 0; a < npe; ++a) tmp[a] = nodes[p[a]];
    memcpy(nodes, tmp, sizeof(uint32_t) * npe);
}

/* ------------------------------------------------------------------------- */
/* make_rect                                                                  */
/* ------------------------------------------------------------------------- */

int or_make_rect_counts(int kind, const unsigned nb[3], uint32_t *n_nodes, uint32_t *n_elem)
{
    if (nb[0] < 2 || nb[1] < 2 || nb[2] < 2) return 1;
    const uint64_t cells = (uint64_t)(nb[0] - 1) * (nb[1] - 1) * (nb[2] - 1);
    if (kind == OR_TETRA) { *n_nodes = nb[0] * nb[1] * nb[2]; *n_elem = (uint32_t)(6 * cells); }
    else if (kind == OR_HEXA) { *n_nodes = nb[0] * nb[1] * nb[2]; *n_elem = (uint32_t)cells; }
    else if (kind == OR_TETRA10) { *n_nodes = (2 * nb[0] - 1) * (2 * nb[1] - 1) * (2 * nb[2] - 1); *n_elem = (uint32_t)(6 * cells); }
    else return 1;
    return 0;
}

/* Node lattice of make_rect(Tetra/Hexa) (make_rect.h:398-399,663-664): the
 * Rectilinear_iterator adds inc to pos[0] and carries into pos[1], pos[2]
 * (util/rectilinear_iterator.h:40-51), so coordinates are running sums, the
 * loop runs while pos[2] < X1[2] + incr[2]/2. */
static void rect_nodes(const double X0[3], const double X1[3], const unsigned nb[3], double *coords)
{
    double incr[3], upper[3], pos[3];
    for (int d = 0; d < 3; ++d) {
        incr[d] = (X1[d] - X0[d]) / (double)(nb[d] - 1);
        upper[d] = X1[d] + incr[d] * 0.5;
        pos[d] = X0[d];
    }
    size_t n = 0;
    while (pos[2] < upper[2]) {
        coords[3 * n + 0] = pos[0]; coords[3 * n + 1] = pos[1]; coords[3 * n + 2] = pos[2];
        ++n;
        pos[0] += incr[0];
        for (int d = 0; d < 2; ++d) {
            if (pos[d] >= upper[d]) { pos[d] = X0[d]; pos[d + 1] += incr[d + 1]; }
            else break;
        }
    }
}

/* Tetra_10 node position O + tr + incr*(chgt*rela[i]) (make_rect.h:477-542). */
static const double T10_RELA[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0.5, 0, 0},
                                       {0.5, 0.5, 0}, {0, 0.5, 0}, {0, 0, 0.5}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
/* per tetra: translation (in units of incr) and the three basis columns */
static const double T10_TR[6][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
static const double T10_BASIS[6][3][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{-1, 0, 0}, {0, -1, 0}, {0, -1, 1}},
    {{-1, 0, 1}, {-1, 1, 0}, {0, 0, 1}},
    {{-1, 0, 0}, {-1, 1, 0}, {-1, 1, -1}},
    {{-1, 0, 0}, {-1, 0, 1}, {0, -1, 1}},
    {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}}};

static void t10_pos(const double X0[3], const double incr[3], unsigned k, unsigned j, unsigned i, int t, int r, double out[3])
{
    const double O[3] = {X0[0] + (double)k * incr[0], X0[1] + (double)j * incr[1], X0[2] + (double)i * incr[2]};
    for (int d = 0; d < 3; ++d) {
        double v = 0; /* chgt * rela: exact small multiples of 1/2 */
        for (int c = 0; c < 3; ++c) v += T10_BASIS[t][c][d] * T10_RELA[r][c];
        const double tr = T10_TR[t][d] * incr[d];
        out[d] = (t == 0 ? O[d] : O[d] + tr) + incr[d] * v;
    }
}

typedef struct { int64_t key; uint32_t val; } hslot;

static uint64_t mix64(uint64_t x)
{
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb3f97a4fe83bULL; x ^= x >> 33;
    return x;
}

/* Tetra_10 make_rect.  The reference deduplicates nodes with a linear search in
 * node_list under tolerance 16*|X1-X0|*eps (make_rect.h:26-39,465), so the node
 * order is order of first appearance.  Positions that coincide within that
 * tolerance are exactly the points of the half-step lattice, so a hash keyed on
 * the rounded half-step index gives the same first-appearance order in O(N). */
static int make_rect_t10(const double X0[3], const double X1[3], const unsigned nb[3], double *coords, uint32_t *conn,
                         uint32_t *n_nodes_out)
{
    const unsigned ne[3] = {nb[0] - 1, nb[1] - 1, nb[2] - 1};
    double incr[3];
    for (int d = 0; d < 3; ++d) incr[d] = (X1[d] - X0[d]) / (double)ne[d];
    const uint64_t L[3] = {2ull * ne[0] + 1, 2ull * ne[1] + 1, 2ull * ne[2] + 1};
    const uint64_t nmax = L[0] * L[1] * L[2];
    size_t cap = 1; while (cap < 2 * nmax) cap <<= 1;
    hslot *tab = (hslot *)malloc(cap * sizeof(hslot));
    if (!tab) return 2;
    for (size_t s = 0; s < cap; ++s) tab[s].key = -1;
    uint32_t nn = 0; size_t ec = 0;
    for (unsigned i = 0; i < ne[2]; ++i)
        for (unsigned j = 0; j < ne[1]; ++j)
            for (unsigned k = 0; k < ne[0]; ++k)
                for (int t = 0; t < 6; ++t) {
                    uint32_t vn[10];
                    for (int r = 0; r < 10; ++r) {
                        double p[3]; t10_pos(X0, incr, k, j, i, t, r, p);
                        int64_t q[3];
                        for (int d = 0; d < 3; ++d) q[d] = (int64_t)llround((p[d] - X0[d]) / (0.5 * incr[d]));
                        const int64_t key = (q[2] * (int64_t)L[1] + q[1]) * (int64_t)L[0] + q[0];
                        size_t s = mix64((uint64_t)key) & (cap - 1);
                        while (tab[s].key != -1 && tab[s].key != key) s = (s + 1) & (cap - 1);
                        if (tab[s].key == -1) {
                            tab[s].key = key; tab[s].val = nn;
                            for (int d = 0; d < 3; ++d) coords[3 * (size_t)nn + d] = p[d];
                            ++nn;
                        }
                        vn[r] = tab[s].val;
                    }
                    permutation_if_jac_neg(OR_TETRA10, coords, vn);
                    memcpy(conn + 10 * ec, vn, sizeof vn); ++ec;
                }
    free(tab);
    if (n_nodes_out) *n_nodes_out = nn;
    return 0;
}

int or_make_rect_tetra10_linear_search(const double X0[3], const double X1[3], const unsigned nb[3], double *coords,
                                       uint32_t *conn, uint32_t *n_nodes_out)
{
    const unsigned ne[3] = {nb[0] - 1, nb[1] - 1, nb[2] - 1};
    double incr[3], dist = 0;
    for (int d = 0; d < 3; ++d) { incr[d] = (X1[d] - X0[d]) / (double)ne[d]; dist += (X1[d] - X0[d]) * (X1[d] - X0[d]); }
    const double prec = 16 * sqrt(dist) * DBL_EPSILON;
    uint32_t nn = 0; size_t ec = 0;
    for (unsigned i = 0; i < ne[2]; ++i)
        for (unsigned j = 0; j < ne[1]; ++j)
            for (unsigned k = 0; k < ne[0]; ++k)
                for (int t = 0; t < 6; ++t) {
                    uint32_t vn[10];
                    for (int r = 0; r < 10; ++r) {
                        double p[3]; t10_pos(X0, incr, k, j, i, t, r, p);
                        uint32_t found = nn;
                        for (uint32_t m = 0; m < nn; ++m) {
                            const double dx = coords[3 * m] - p[0], dy = coords[3 * m + 1] - p[1], dz = coords[3 * m + 2] - p[2];
                            if (sqrt(dx * dx + dy * dy + dz * dz) < prec) { found = m; break; }
                        }
                        if (found == nn) { for (int d = 0; d < 3; ++d) coords[3 * (size_t)nn + d] = p[d]; ++nn; }
                        vn[r] = found;
                    }
                    permutation_if_jac_neg(OR_TETRA10, coords, vn);
                    memcpy(conn + 10 * ec, vn, sizeof vn); ++ec;
                }
    *n_nodes_out = nn;
    return 0;
}

int or_make_rect(int kind, const double X0[3], const double X1[3], const unsigned nb[3], double *coords, uint32_t *conn)
{
    uint32_t nn, ne;
    if (or_make_rect_counts(kind, nb, &nn, &ne)) return 1;
    if (kind == OR_TETRA10) {
        uint32_t got = 0;
        int rc = make_rect_t10(X0, X1, nb, coords, conn, &got);
        return rc ? rc : (got == nn ? 0 : 5);
    }
    rect_nodes(X0, X1, nb, coords);
    const unsigned nx = nb[0], ny = nb[1], nz = nb[2];
    size_t e = 0;
    if (kind == OR_HEXA) { /* make_rect.h:665-678 */
        for (unsigned i = 0; i + 1 < nz; ++i)
            for (unsigned j = 0; j + 1 < ny; ++j)
                for (unsigned k = 0; k + 1 < nx; ++k) {
                    const uint32_t b = i * nx * ny + j * nx + k, up = nx * ny;
                    const uint32_t v[8] = {b, b + 1, b + 1 + nx, b + nx, b + up, b + up + 1, b + up + 1 + nx, b + up + nx};
                    memcpy(conn + 8 * e, v, sizeof v); ++e;
                }
        return 0;
    }
    /* Tetra: make_rect.h:400-428, corner bits x=1, y=2, z=4 */
    static const int lst[24] = {0, 1, 2, 4, 1, 2, 3, 5, 1, 2, 4, 5, 2, 4, 5, 6, 2, 3, 5, 6, 3, 5, 6, 7};
    for (unsigned i = 0; i + 1 < nz; ++i)
        for (unsigned j = 0; j + 1 < ny; ++j)
            for (unsigned k = 0; k + 1 < nx; ++k)
                for (int n = 0; n < 6; ++n) {
                    uint32_t vn[4];
                    for (int p = 0; p < 4; ++p) {
                        const int c = lst[4 * n + p];
                        vn[p] = (i + ((c & 4) >> 2)) * nx * ny + (j + ((c & 2) >> 1)) * nx + (k + (c & 1));
                    }
                    permutation_if_jac_neg(OR_TETRA, coords, vn);
                    memcpy(conn + 4 * e, vn, sizeof vn); ++e;
                }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Skin                                                                       */
/* ------------------------------------------------------------------------- */

/* Faces of each element in the reference's child ordering:
 *   Tetra    formulations/element_Tetra.py:17-22
 *   Tetra_10 formulations/element_Tetra_10.py:17-22
 *   Hexa     formulations/element_Hexa.py:22-29 */
static const int FACES_TET[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
static const int FACES_T10[4][6] = {{0, 2, 1, 6, 5, 4}, {0, 1, 3, 4, 8, 7}, {0, 3, 2, 7, 9, 6}, {1, 2, 3, 5, 9, 8}};
static const int FACES_HEX[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 4, 7, 3}, {1, 2, 6, 5}};

typedef struct { uint32_t k[4]; uint32_t elem; uint8_t face; } facekey;

static int cmp_facekey(const void *a, const void *b)
{
    const facekey *x = (const facekey *)a, *y = (const facekey *)b;
    for (int i = 0; i < 4; ++i) if (x->k[i] != y->k[i]) return x->k[i] < y->k[i] ? -1 : 1;
    if (x->elem != y->elem) return x->elem < y->elem ? -1 : 1;
    return (int)x->face - (int)y->face;
}

static void sort_small(uint32_t *v, int n)
{
    for (int i = 1; i < n; ++i) { uint32_t t = v[i]; int j = i - 1; while (j >= 0 && v[j] > t) { v[j + 1] = v[j]; --j; } v[j + 1] = t; }
}

/* update_skin: faces shared by exactly one element (mesh.h:151-188).  Output
 * ordered by (element, local face). */
int or_skin_faces(int kind, uint32_t n_nodes, uint32_t n_elem, const uint32_t *conn, uint32_t *faces, uint32_t cap,
                  uint32_t *n_faces)
{
    (void)n_nodes;
    const int npe = or_nodes_per_elem(kind), nfe = kind == OR_HEXA ? 6 : 4;
    const int nvk = kind == OR_HEXA ? 4 : 3; /* vertices that identify a face */
    const int npf = or_nodes_per_face(or_face_kind_of(kind));
    const size_t nf = (size_t)n_elem * nfe;
    facekey *fk = (facekey *)malloc(nf * sizeof(facekey));
    if (!fk) return 2;
    for (uint32_t e = 0; e < n_elem; ++e)
        for (int f = 0; f < nfe; ++f) {
            facekey *q = &fk[(size_t)e * nfe + f];
            const int *loc = kind == OR_TETRA ? FACES_TET[f] : kind == OR_TETRA10 ? FACES_T10[f] : FACES_HEX[f];
            q->k[3] = 0;
            for (int v = 0; v < nvk; ++v) q->k[v] = conn[(size_t)e * npe + loc[v]];
            sort_small(q->k, nvk);
            q->elem = e; q->face = (uint8_t)f;
        }
    qsort(fk, nf, sizeof(facekey), cmp_facekey);
    /* boundary flags per (elem, face) */
    uint8_t *bnd = (uint8_t *)calloc(nf, 1);
    for (size_t s = 0; s < nf;) {
        size_t t = s + 1;
        while (t < nf && memcmp(fk[t].k, fk[s].k, sizeof fk[s].k) == 0) ++t;
        if (t - s == 1) bnd[(size_t)fk[s].elem * nfe + fk[s].face] = 1;
        s = t;
    }
    uint32_t cnt = 0;
    for (uint32_t e = 0; e < n_elem; ++e)
        for (int f = 0; f < nfe; ++f)
            if (bnd[(size_t)e * nfe + f]) {
                if (faces && cnt < cap) {
                    const int *loc = kind == OR_TETRA ? FACES_TET[f] : kind == OR_TETRA10 ? FACES_T10[f] : FACES_HEX[f];
                    for (int v = 0; v < npf; ++v) faces[(size_t)cnt * npf + v] = conn[(size_t)e * npe + loc[v]];
                }
                ++cnt;
            }
    free(bnd); free(fk);
    *n_faces = cnt;
    return (faces && cnt > cap) ? 1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Element kernels                                                            */
/* ------------------------------------------------------------------------- */

/* Isotropic Hooke law: D is the inverse of the compliance with s1=1/E,
 * s2=-nu/E, s3=(1+nu)/E on tensor shear (formal_lf/material_help.py:116-128,181-183);
 * equal to lambda = E nu/((1+nu)(1-2nu)), mu = E/(2(1+nu)).  The energy
 * trace_sym_col(sigma, eps_test) sums normal terms and twice the tensor shear
 * terms (include/codegen/__init__.py:23-42,82-87), so with B from grad_sym_col
 * (formal_lf/element.py:270-296) the 3x3 block of nodes a, b is
 *   K_ab(i,j) = lambda g_a,i g_b,j + mu g_a,j g_b,i + mu delta_ij (g_a . g_b). */
static void lame(double E, double nu, double *lam, double *mu)
{
    *lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    *mu = E / (2 * (1 + nu));
}

typedef struct { int ngp; double w[8]; double wd[8]; double N[8][10]; double g[8][10][3]; double detj_min; } gp_data;

static int element_gp(int kind, const double *xe, gp_data *G)
{
    const int npe = or_nodes_per_elem(kind);
    double xi[8][3];
    G->ngp = gauss_points(kind, G->w, xi);
    G->detj_min = INFINITY;
    for (int q = 0; q < G->ngp; ++q) {
        double dN[10][3], J[3][3], Ji[3][3];
        shape(kind, xi[q], G->N[q], dN);
        const double det = jacobian(npe, xe, dN, J, Ji);
        if (det < G->detj_min) G->detj_min = det;
        G->wd[q] = G->w[q] * det; /* signed detJ (formal_lf/formulation.py:553) */
        for (int a = 0; a < npe; ++a)
            for (int j = 0; j < 3; ++j)
                G->g[q][a][j] = Ji[j][0] * dN[a][0] + Ji[j][1] * dN[a][1] + Ji[j][2] * dN[a][2];
    }
    return G->ngp;
}

/* One Gauss point's contribution to local entry (3a+i, 3b+j). */
static inline double kentry(const gp_data *G, int q, int a, int i, int b, int j, double lam, double mu)
{
    const double *ga = G->g[q][a], *gb = G->g[q][b];
    double v = lam * (ga[i] * gb[j]) + mu * (ga[j] * gb[i]);
    if (i == j) v += mu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
    return G->wd[q] * v;
}

int or_elem_matrix(int kind, const double *xe, double E, double nu, const double bf[3], double *Ke, double *Fe,
                   double *detj_min)
{
    const int npe = or_nodes_per_elem(kind), nd = 3 * npe;
    if (!npe) return 1;
    gp_data G; element_gp(kind, xe, &G);
    double lam, mu; lame(E, nu, &lam, &mu);
    if (Ke) memset(Ke, 0, sizeof(double) * nd * nd);
    if (Fe) memset(Fe, 0, sizeof(double) * nd);
    for (int q = 0; q < G.ngp; ++q) {
        if (Ke)
            for (int r = 0; r < nd; ++r)
                for (int c = r; c < nd; ++c) {
                    const double v = kentry(&G, q, r / 3, r % 3, c / 3, c % 3, lam, mu);
                    Ke[r * nd + c] += v;
                    if (c != r) Ke[c * nd + r] += v;
                }
        if (Fe && bf)
            for (int a = 0; a < npe; ++a)
                for (int d = 0; d < 3; ++d) Fe[3 * a + d] += G.wd[q] * G.N[q][a] * bf[d];
    }
    if (detj_min) *detj_min = G.detj_min;
    return G.detj_min > 0 ? 0 : 2;
}

/* Face quadrature: the skin integrand is integrated at generation time with the
 * face element's Python Gauss table (formal_lf/formulation.py:400-442), key =
 * smallest >= the parent's order: Triangle 1 point, Triangle_6 3 mid-edge points,
 * Quad 2x2 points at (1 +- sqrt(1/3))/2 (formulations/element_Quad.py:70-75). */
int or_face_load(int fkind, const double *xf, const double t[3], double *Ff)
{
    const int npf = or_nodes_per_face(fkind);
    double w[4], p[4][2]; int ng;
    if (fkind == OR_TRI3) { ng = 1; w[0] = 0.5; p[0][0] = p[0][1] = 1.0 / 3.0; }
    else if (fkind == OR_TRI6) {
        ng = 3; w[0] = w[1] = w[2] = 1.0 / 6.0;
        p[0][0] = 0.5; p[0][1] = 0.0; p[1][0] = 0.5; p[1][1] = 0.5; p[2][0] = 0.0; p[2][1] = 0.5;
    } else if (fkind == OR_QUAD4) {
        ng = 4; const double a = (1.0 - sqrt(1.0 / 3.0)) / 2.0, b = (1.0 + sqrt(1.0 / 3.0)) / 2.0;
        const double pp[4][2] = {{b, b}, {a, b}, {b, a}, {a, a}};
        for (int g = 0; g < 4; ++g) { w[g] = 0.25; p[g][0] = pp[g][0]; p[g][1] = pp[g][1]; }
    } else return 1;
    memset(Ff, 0, sizeof(double) * 3 * npf);
    for (int g = 0; g < ng; ++g) {
        const double x = p[g][0], y = p[g][1];
        double N[6], d0[6], d1[6];
        if (fkind == OR_TRI3) {
            N[0] = 1 - x - y; N[1] = x; N[2] = y;
            d0[0] = -1; d0[1] = 1; d0[2] = 0; d1[0] = -1; d1[1] = 0; d1[2] = 1;
        } else if (fkind == OR_TRI6) { /* formulations/element_Triangle_6.py:38-48 */
            const double L0 = 1 - x - y;
            N[0] = (2 * L0 - 1) * L0; N[1] = (2 * x - 1) * x; N[2] = (2 * y - 1) * y;
            N[3] = 4 * L0 * x; N[4] = 4 * x * y; N[5] = 4 * y * L0;
            d0[0] = -(4 * L0 - 1); d0[1] = 4 * x - 1; d0[2] = 0; d0[3] = 4 * (L0 - x); d0[4] = 4 * y; d0[5] = -4 * y;
            d1[0] = -(4 * L0 - 1); d1[1] = 0; d1[2] = 4 * y - 1; d1[3] = -4 * x; d1[4] = 4 * x; d1[5] = 4 * (L0 - y);
        } else { /* formulations/element_Quad.py:22-25 */
            N[0] = (1 - x) * (1 - y); N[1] = x * (1 - y); N[2] = x * y; N[3] = (1 - x) * y;
            d0[0] = -(1 - y); d0[1] = 1 - y; d0[2] = y; d0[3] = -y;
            d1[0] = -(1 - x); d1[1] = -x; d1[2] = x; d1[3] = 1 - x;
        }
        double t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
        for (int a = 0; a < npf; ++a)
            for (int d = 0; d < 3; ++d) { t0[d] += d0[a] * xf[3 * a + d]; t1[d] += d1[a] * xf[3 * a + d]; }
        const double n0 = t0[1] * t1[2] - t0[2] * t1[1], n1 = t0[2] * t1[0] - t0[0] * t1[2], n2 = t0[0] * t1[1] - t0[1] * t1[0];
        const double ds = w[g] * sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        for (int a = 0; a < npf; ++a)
            for (int d = 0; d < 3; ++d) Ff[3 * a + d] += ds * N[a] * t[d];
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Pattern + assembly                                                         */
/* ------------------------------------------------------------------------- */

static int cmp_u32(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* ResizeMat::GetNZ (formulation.h:200-230): every ordered node pair of every
 * element contributes the 3x3 block (DOF = 3*node + comp, formulation.h:376);
 * empty rows get a diagonal; rows sorted.  The reference keeps the lower
 * triangle; this restatement stores both triangles (blocks (r,c) and (c,r)). */
int or_pattern(int kind, uint32_t n_nodes, uint32_t n_elem, const uint32_t *conn, uint32_t *row_ptr, uint32_t *col,
               size_t *nnzb)
{
    const int npe = or_nodes_per_elem(kind);
    /* node -> element lists */
    uint32_t *deg = (uint32_t *)calloc((size_t)n_nodes + 1, sizeof(uint32_t));
    for (size_t s = 0; s < (size_t)n_elem * npe; ++s) deg[conn[s] + 1]++;
    for (uint32_t n = 0; n < n_nodes; ++n) deg[n + 1] += deg[n];
    uint32_t *ne = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n_elem * npe), *fill = (uint32_t *)calloc(n_nodes, 4);
    for (uint32_t e = 0; e < n_elem; ++e)
        for (int a = 0; a < npe; ++a) { const uint32_t n = conn[(size_t)e * npe + a]; ne[deg[n] + fill[n]++] = e; }
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * 4096);
    size_t tot = 0;
    for (uint32_t n = 0; n < n_nodes; ++n) {
        int m = 0;
        for (uint32_t s = deg[n]; s < deg[n + 1]; ++s)
            for (int b = 0; b < npe; ++b) buf[m++] = conn[(size_t)ne[s] * npe + b];
        if (m == 0) buf[m++] = n;
        qsort(buf, m, 4, cmp_u32);
        int u = 0;
        for (int i = 0; i < m; ++i) if (i == 0 || buf[i] != buf[i - 1]) buf[u++] = buf[i];
        if (row_ptr) row_ptr[n] = (uint32_t)tot;
        if (col) memcpy(col + tot, buf, 4 * (size_t)u);
        tot += u;
    }
    if (row_ptr) row_ptr[n_nodes] = (uint32_t)tot;
    *nnzb = tot;
    free(buf); free(fill); free(ne); free(deg);
    return 0;
}

static inline size_t find_block(const uint32_t *row_ptr, const uint32_t *col, uint32_t r, uint32_t c)
{
    uint32_t lo = row_ptr[r], hi = row_ptr[r + 1];
    while (lo < hi) { const uint32_t m = (lo + hi) / 2; if (col[m] < c) lo = m + 1; else hi = m; }
    return lo;
}

/* One element of the loop below: its Gauss points' local upper triangles (i <= j) added
 * straight into the global matrix, mirrored to the symmetric position
 * (formal_lf/calculate_matrix.py:64-98), and its body load.  Returns 2 for det J <= 0. */
static int assemble_elem(int kind, const double *coords, const uint32_t *conn, uint32_t e, const double *E_e, double nu,
                         const double bf[3], const uint32_t *row_ptr, const uint32_t *col, double *vals, double *F)
{
    const int npe = or_nodes_per_elem(kind), nd = 3 * npe;
    double xe[30];
    const uint32_t *en = conn + (size_t)e * npe;
    for (int a = 0; a < npe; ++a) for (int d = 0; d < 3; ++d) xe[3 * a + d] = coords[3 * (size_t)en[a] + d];
    gp_data G; element_gp(kind, xe, &G);
    const int status = G.detj_min > 0 ? 0 : 2;
    double lam, mu; lame(E_e[e], nu, &lam, &mu);
    size_t slot[10][10];
    for (int a = 0; a < npe; ++a) for (int b = 0; b < npe; ++b) slot[a][b] = find_block(row_ptr, col, en[a], en[b]);
    for (int q = 0; q < G.ngp; ++q) {
        for (int r = 0; r < nd; ++r)
            for (int c = r; c < nd; ++c) {
                const int a = r / 3, i = r % 3, b = c / 3, j = c % 3;
                const double v = kentry(&G, q, a, i, b, j, lam, mu);
                vals[9 * slot[a][b] + 3 * i + j] += v;
                if (r != c) vals[9 * slot[b][a] + 3 * j + i] += v;
            }
        if (bf)
            for (int a = 0; a < npe; ++a)
                for (int d = 0; d < 3; ++d) F[3 * (size_t)en[a] + d] += G.wd[q] * G.N[q][a] * bf[d];
    }
    return status;
}

/* Element loop (formulation.h:626-647) with the default per-Gauss-point
 * integration (formulation.h:2917-2933), in element order; then the skin faces
 * (formulation.h:584-594).  With more than one thread (or_set_threads) the
 * elements are greedily coloured so that no two of a colour share a node, and
 * the colours run in turn, each on all threads: every entry then receives its
 * additions in colour order instead of element order (the multi-core CPU line). */
int or_assemble(int kind, uint32_t n_nodes, const double *coords, uint32_t n_elem, const uint32_t *conn,
                const double *E_e, double nu, const double bf[3], uint32_t n_faces, const uint32_t *faces,
                const double *face_trac, const uint32_t *row_ptr, const uint32_t *col, double *vals, double *F)
{
    const int npe = or_nodes_per_elem(kind);
    memset(vals, 0, sizeof(double) * 9 * (size_t)row_ptr[n_nodes]);
    memset(F, 0, sizeof(double) * 3 * (size_t)n_nodes);
    int status = 0;
    if (g_threads > 1) {
        uint32_t *color = (uint32_t *)malloc(4 * (size_t)(n_elem ? n_elem : 1));
        uint64_t *used = (uint64_t *)calloc(n_nodes ? n_nodes : 1, 8);
        uint32_t n_col = 0;
        for (uint32_t e = 0; e < n_elem && status == 0; ++e) {
            const uint32_t *en = conn + (size_t)e * npe;
            uint64_t m = 0;
            for (int a = 0; a < npe; ++a) m |= used[en[a]];
            uint32_t c = 0;
            while (c < 64 && (m >> c & 1)) ++c;
            if (c == 64) { status = 1; break; }
            color[e] = c;
            if (c + 1 > n_col) n_col = c + 1;
            for (int a = 0; a < npe; ++a) used[en[a]] |= 1ull << c;
        }
        uint32_t *start = (uint32_t *)calloc(n_col + 2, 4), *order = (uint32_t *)malloc(4 * (size_t)(n_elem ? n_elem : 1));
        if (status == 0) {
            for (uint32_t e = 0; e < n_elem; ++e) ++start[color[e] + 1];
            for (uint32_t c = 0; c < n_col; ++c) start[c + 1] += start[c];
            for (uint32_t e = 0; e < n_elem; ++e) order[start[color[e]]++] = e;   /* start[c] -> end of c */
            for (uint32_t c = n_col; c > 0; --c) start[c] = start[c - 1];
            start[0] = 0;
        }
        for (uint32_t c = 0; c < n_col && status == 0; ++c) {
            int st = 0;
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(max : st)
            for (uint32_t k = start[c]; k < start[c + 1]; ++k) {
                const int s_ = assemble_elem(kind, coords, conn, order[k], E_e, nu, bf, row_ptr, col, vals, F);
                if (s_ > st) st = s_;
            }
            if (st) status = st;
        }
        free(color); free(used); free(start); free(order);
        if (status == 1) return 1;
    } else {
        for (uint32_t e = 0; e < n_elem; ++e) {
            const int st = assemble_elem(kind, coords, conn, e, E_e, nu, bf, row_ptr, col, vals, F);
            if (st) status = st;
        }
    }
    if (n_faces && faces && face_trac) {
        const int fk = or_face_kind_of(kind), npf = or_nodes_per_face(fk);
        for (uint32_t f = 0; f < n_faces; ++f) {
            double xf[18], Ff[18];
            for (int a = 0; a < npf; ++a) for (int d = 0; d < 3; ++d) xf[3 * a + d] = coords[3 * (size_t)faces[(size_t)f * npf + a] + d];
            or_face_load(fk, xf, face_trac + 3 * (size_t)f, Ff);
            for (int a = 0; a < npf; ++a) for (int d = 0; d < 3; ++d) F[3 * (size_t)faces[(size_t)f * npf + a] + d] += Ff[3 * a + d];
        }
    }
    return status;
}

/* assemble_constraints penalty branch (formulation.h:1028-1070): M = max|diag K|
 * taken once after element assembly; for a constraint u_p = value with unit
 * coefficient: K(p,p) += M*P, F(p) += M*P*value. */
int or_penalty(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, double *vals, double *F, uint32_t n_cons,
               const uint32_t *cons_dof, const double *cons_val, double P, double *max_diag)
{
    double M = 0;
    for (uint32_t r = 0; r < n_nodes; ++r) {
        const size_t s = find_block(row_ptr, col, r, r);
        for (int d = 0; d < 3; ++d) { const double v = fabs(vals[9 * s + 4 * d]); if (v > M) M = v; }
    }
    for (uint32_t k = 0; k < n_cons; ++k) {
        const uint32_t p = cons_dof[k], r = p / 3, d = p % 3;
        const double C = 1.0 * M * P;
        if (F) F[p] += C * (cons_val ? cons_val[k] : 0.0);
        vals[9 * find_block(row_ptr, col, r, r) + 4 * d] += C * 1.0;
    }
    if (max_diag) *max_diag = M;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Solvers                                                                    */
/* ------------------------------------------------------------------------- */

void or_spmv(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals, const double *x, double *y)
{
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (uint32_t r = 0; r < n_nodes; ++r) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (uint32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
            const double *B = vals + 9 * (size_t)p, *xc = x + 3 * (size_t)col[p];
            s0 += B[0] * xc[0] + B[1] * xc[1] + B[2] * xc[2];
            s1 += B[3] * xc[0] + B[4] * xc[1] + B[5] * xc[2];
            s2 += B[6] * xc[0] + B[7] * xc[1] + B[8] * xc[2];
        }
        y[3 * (size_t)r] = s0; y[3 * (size_t)r + 1] = s1; y[3 * (size_t)r + 2] = s2;
    }
}

static double dotv(size_t n, const double *a, const double *b) { double s = 0; for (size_t i = 0; i < n; ++i) s += a[i] * b[i]; return s; }

/* Dot product of the multi-threaded PCG: DOT_CHUNKS fixed index ranges summed in order, the
 * range sums added in range order -- the same value for every thread count. */
#define DOT_CHUNKS 1024
static double dotv_mt(size_t n, const double *a, const double *b)
{
    double part[DOT_CHUNKS];
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int k = 0; k < DOT_CHUNKS; ++k) {
        const size_t i0 = n * (size_t)k / DOT_CHUNKS, i1 = n * (size_t)(k + 1) / DOT_CHUNKS;
        double t = 0;
        for (size_t i = i0; i < i1; ++i) t += a[i] * b[i];
        part[k] = t;
    }
    double s = 0;
    for (int k = 0; k < DOT_CHUNKS; ++k) s += part[k];
    return s;
}
#define DOTP(n_, a_, b_) (g_threads > 1 ? dotv_mt(n_, a_, b_) : dotv(n_, a_, b_))

static int inv3(const double *A, double *Ai)
{
    const double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8], c02 = A[3] * A[7] - A[4] * A[6];
    const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    if (!(det != 0) || !isfinite(det)) return 1;
    const double id = 1.0 / det;
    Ai[0] = c00 * id; Ai[3] = c01 * id; Ai[6] = c02 * id;
    Ai[1] = (A[2] * A[7] - A[1] * A[8]) * id; Ai[4] = (A[0] * A[8] - A[2] * A[6]) * id; Ai[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    Ai[2] = (A[1] * A[5] - A[2] * A[4]) * id; Ai[5] = (A[2] * A[3] - A[0] * A[5]) * id; Ai[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    return 0;
}

/* Textbook preconditioned CG (the loop of conjugate_gradient.h:27-68 with the
 * Jacobi preconditioner, stop ||r||_2 <= rel_tol ||b||_2). */
int or_pcg_jacobi(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals, const double *b,
                  double *x, double rel_tol, int max_iter, int block_jacobi, int *iters, double *rel_res)
{
    return or_pcg_jacobi_hist(n_nodes, row_ptr, col, vals, b, x, rel_tol, max_iter, block_jacobi, iters, rel_res, NULL, 0);
}

int or_pcg_jacobi_hist(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals,
                       const double *b, double *x, double rel_tol, int max_iter, int block_jacobi, int *iters,
                       double *rel_res, double *hist, int hist_cap)
{
    const double t_start = omp_get_wtime();
    const size_t n = 3 * (size_t)n_nodes;
    double *r = (double *)malloc(8 * n), *z = (double *)malloc(8 * n), *p = (double *)malloc(8 * n), *q = (double *)malloc(8 * n);
    double *Dinv = (double *)malloc(8 * 9 * (size_t)n_nodes);
    int status = 0;
    for (uint32_t i = 0; i < n_nodes; ++i) {
        const double *B = vals + 9 * find_block(row_ptr, col, i, i);
        double *Di = Dinv + 9 * (size_t)i;
        if (block_jacobi) { if (inv3(B, Di)) status = 3; }
        else {
            memset(Di, 0, 72);
            for (int d = 0; d < 3; ++d) { if (!(B[4 * d] > 0)) status = 3; Di[4 * d] = 1.0 / B[4 * d]; }
        }
    }
    if (status) goto done;
    or_spmv(n_nodes, row_ptr, col, vals, x, q);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
#define APPLY_M(dst, src) \
    _Pragma("omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)") \
    for (uint32_t k = 0; k < n_nodes; ++k) { const double *Di = Dinv + 9 * (size_t)k; const double *s_ = src + 3 * (size_t)k; \
        for (int d = 0; d < 3; ++d) dst[3 * (size_t)k + d] = Di[3 * d] * s_[0] + Di[3 * d + 1] * s_[1] + Di[3 * d + 2] * s_[2]; }
    const double bn = sqrt(DOTP(n, b, b));
    const double stop = rel_tol * (bn > 0 ? bn : 1.0);
    double rn = sqrt(DOTP(n, r, r));
    int it = 0;
    APPLY_M(z, r);
    memcpy(p, z, 8 * n);
    double rho = DOTP(n, r, z);
    while (rn > stop) {
        if (it >= max_iter) { status = 4; break; }
        or_spmv(n_nodes, row_ptr, col, vals, p, q);
        const double pq = DOTP(n, p, q);
        if (!(pq > 0)) { status = 3; break; }
        const double alpha = rho / pq;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        APPLY_M(z, r);
        const double rho1 = DOTP(n, r, z);
        const double beta = rho1 / rho;
        rho = rho1;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rn = sqrt(DOTP(n, r, r));
        if (hist && it < hist_cap) {
            hist[2 * it] = rn / (bn > 0 ? bn : 1.0);
            hist[2 * it + 1] = omp_get_wtime() - t_start;
        }
        ++it;
    }
#undef APPLY_M
    if (iters) *iters = it;
    if (rel_res) *rel_res = rn / (bn > 0 ? bn : 1.0);
done:
    free(r); free(z); free(p); free(q); free(Dinv);
    return status;
}

int or_cg_csr(uint32_t n, const uint32_t *row_ptr, const uint32_t *col, const double *vals, const double *b, double *x,
              double rel_tol, int max_iter, int *iters)
{
    double *r = (double *)malloc(8 * (size_t)n), *z = (double *)malloc(8 * (size_t)n), *p = (double *)malloc(8 * (size_t)n),
           *q = (double *)malloc(8 * (size_t)n), *dinv = (double *)malloc(8 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        double d = 0;
        for (uint32_t s = row_ptr[i]; s < row_ptr[i + 1]; ++s) if (col[s] == i) d = vals[s];
        dinv[i] = d > 0 ? 1.0 / d : 1.0;
    }
#define SPMV(y, v) for (uint32_t i = 0; i < n; ++i) { double s = 0; for (uint32_t t = row_ptr[i]; t < row_ptr[i + 1]; ++t) s += vals[t] * v[col[t]]; y[i] = s; }
    SPMV(q, x);
    for (uint32_t i = 0; i < n; ++i) { r[i] = b[i] - q[i]; z[i] = dinv[i] * r[i]; p[i] = z[i]; }
    const double bn = sqrt(dotv(n, b, b));
    double rho = dotv(n, r, z), rn = sqrt(dotv(n, r, r));
    int it = 0, status = 0;
    while (rn > rel_tol * bn) {
        if (it >= max_iter) { status = 4; break; }
        SPMV(q, p);
        const double alpha = rho / dotv(n, p, q);
        for (uint32_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; z[i] = dinv[i] * r[i]; }
        const double rho1 = dotv(n, r, z);
        for (uint32_t i = 0; i < n; ++i) p[i] = z[i] + (rho1 / rho) * p[i];
        rho = rho1; rn = sqrt(dotv(n, r, r)); ++it;
    }
#undef SPMV
    if (iters) *iters = it;
    free(r); free(z); free(p); free(q); free(dinv);
    return status;
}

/* Block IC(0) of MatWithTinyBlocks::chol_incomp (containers/MatWithTinyBlocks.h:426-523):
 * row by row over the lower block pattern, each strictly-lower block
 *   L_rc = (A_rc - sum_{k<c, k in both rows} L_rk L_ck^T) L_cc^{-T}
 * with the column solve of :468-487 (forward substitution against the lower
 * diagonal factor), and the diagonal block as the dense 3x3 Cholesky of
 *   A_rr - sum_{k<r} L_rk L_rk^T   (:497-522).
 * L is stored at the full-BSR positions with c <= r. */
int or_block_ic0(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals, double *L)
{
    memcpy(L, vals, sizeof(double) * 9 * (size_t)row_ptr[n_nodes]);
    for (uint32_t r = 0; r < n_nodes; ++r) {
        const uint32_t rb = row_ptr[r], re = row_ptr[r + 1];
        uint32_t pd = rb; while (pd < re && col[pd] < r) ++pd; /* diagonal slot */
        if (pd == re || col[pd] != r) return 1;
        for (uint32_t pc = rb; pc < pd; ++pc) {
            const uint32_t c = col[pc];
            const uint32_t cb = row_ptr[c];
            double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            /* merge the strictly-lower parts of rows r and c over columns k < c */
            uint32_t i1 = rb, i2 = cb;
            while (i1 < pc && i2 < row_ptr[c + 1] && col[i2] < c) {
                if (col[i1] < col[i2]) ++i1;
                else if (col[i1] > col[i2]) ++i2;
                else {
                    const double *A = L + 9 * (size_t)i1, *B = L + 9 * (size_t)i2;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) s[3 * a + b] += A[3 * a] * B[3 * b] + A[3 * a + 1] * B[3 * b + 1] + A[3 * a + 2] * B[3 * b + 2];
                    ++i1; ++i2;
                }
            }
            uint32_t pcd = cb; while (col[pcd] != c) ++pcd;
            const double *D = L + 9 * (size_t)pcd; /* lower-triangular factor of block c */
            double *O = L + 9 * (size_t)pc;
            for (int a = 0; a < 3; ++a) {
                O[3 * a + 0] = (O[3 * a + 0] - s[3 * a + 0]) / D[0];
                O[3 * a + 1] = (O[3 * a + 1] - s[3 * a + 1] - O[3 * a + 0] * D[3]) / D[4];
                O[3 * a + 2] = (O[3 * a + 2] - s[3 * a + 2] - O[3 * a + 0] * D[6] - O[3 * a + 1] * D[7]) / D[8];
            }
        }
        double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t pc = rb; pc < pd; ++pc) {
            const double *A = L + 9 * (size_t)pc;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b <= a; ++b) s[3 * a + b] += A[3 * a] * A[3 * b] + A[3 * a + 1] * A[3 * b + 1] + A[3 * a + 2] * A[3 * b + 2];
        }
        double *D = L + 9 * (size_t)pd;
        const double d00 = D[0] - s[0];
        if (!(d00 > 0)) return 3;
        D[0] = sqrt(d00);
        D[3] = (D[3] - s[3]) / D[0];
        D[6] = (D[6] - s[6]) / D[0];
        const double d11 = D[4] - s[4] - D[3] * D[3];
        if (!(d11 > 0)) return 3;
        D[4] = sqrt(d11);
        D[7] = (D[7] - s[7] - D[6] * D[3]) / D[4];
        const double d22 = D[8] - s[8] - D[6] * D[6] - D[7] * D[7];
        if (!(d22 > 0)) return 3;
        D[8] = sqrt(d22);
        D[1] = D[2] = D[5] = 0;
    }
    return 0;
}

/* z = (L L^T)^{-1} r: forward over block rows, then backward scattering the
 * transposed blocks (MatWithTinyBlocks.h:587-709). */
static void ic0_solve(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *L, const double *r, double *z)
{
    memcpy(z, r, 24 * (size_t)n_nodes);
    for (uint32_t i = 0; i < n_nodes; ++i) {
        double *zi = z + 3 * (size_t)i;
        uint32_t p = row_ptr[i];
        for (; col[p] < i; ++p) {
            const double *B = L + 9 * (size_t)p, *zc = z + 3 * (size_t)col[p];
            for (int a = 0; a < 3; ++a) zi[a] -= B[3 * a] * zc[0] + B[3 * a + 1] * zc[1] + B[3 * a + 2] * zc[2];
        }
        const double *D = L + 9 * (size_t)p;
        zi[0] = zi[0] / D[0];
        zi[1] = (zi[1] - D[3] * zi[0]) / D[4];
        zi[2] = (zi[2] - D[6] * zi[0] - D[7] * zi[1]) / D[8];
    }
    for (uint32_t i = n_nodes; i-- > 0;) {
        double *zi = z + 3 * (size_t)i;
        uint32_t p = row_ptr[i]; while (col[p] < i) ++p;
        const double *D = L + 9 * (size_t)p;
        zi[2] = zi[2] / D[8];
        zi[1] = (zi[1] - D[7] * zi[2]) / D[4];
        zi[0] = (zi[0] - D[3] * zi[1] - D[6] * zi[2]) / D[0];
        for (uint32_t q = row_ptr[i]; col[q] < i; ++q) {
            const double *B = L + 9 * (size_t)q;
            double *zc = z + 3 * (size_t)col[q];
            for (int b = 0; b < 3; ++b) zc[b] -= B[b] * zi[0] + B[3 + b] * zi[1] + B[6 + b] * zi[2];
        }
    }
}

/* solve_using_incomplete_chol_factorize (MatWithTinyBlocks.h:1101-1151): stop as
 * soon as every |r_i| <= crit (absolute max-norm), tested on the initial residual
 * and after every update. */
int or_pcg_ic0(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals, const double *L,
               const double *b, double *x, double crit, int max_iter, int *iters)
{
    const size_t n = 3 * (size_t)n_nodes;
    double *r = (double *)malloc(8 * n), *d = (double *)malloc(8 * n), *q = (double *)malloc(8 * n), *s = (double *)malloc(8 * n);
    int status = 0, it = 0;
    or_spmv(n_nodes, row_ptr, col, vals, x, q);
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
#define CONVERGED() ({ int ok_ = 1; for (size_t i_ = 0; i_ < n; ++i_) if (fabs(r[i_]) > crit) { ok_ = 0; break; } ok_; })
    if (!CONVERGED()) {
        ic0_solve(n_nodes, row_ptr, col, L, r, d);
        double deltn = dotv(n, r, d);
        for (;;) {
            if (it >= max_iter) { status = 4; break; }
            const double delto = deltn;
            or_spmv(n_nodes, row_ptr, col, vals, d, q);
            const double alpha = deltn / dotv(n, d, q);
            for (size_t i = 0; i < n; ++i) { x[i] += alpha * d[i]; r[i] -= alpha * q[i]; }
            ++it;
            if (CONVERGED()) break;
            ic0_solve(n_nodes, row_ptr, col, L, r, s);
            deltn = dotv(n, r, s);
            const double beta = deltn / delto;
            for (size_t i = 0; i < n; ++i) d[i] = s[i] + beta * d[i];
        }
    }
#undef CONVERGED
    if (iters) *iters = it;
    free(r); free(d); free(q); free(s);
    return status;
}

/* ------------------------------------------------------------------------- */
/* After-solve recovery                                                       */
/* ------------------------------------------------------------------------- */

/* apply_on_elements_after_solve (formulations/formulation_elasticity_iso.py:64-85), called per
 * element by call_after_solve (include/formulation/formulation.h:2026-2050): every var_inter is
 * substituted by 1/2, so the gradient is taken at xi = (1/2,1/2,1/2) whatever the element (for the
 * tetrahedra that point lies outside the reference simplex).  G = sum_a u_a (J^-1 dN_a)^T,
 * eps = grad_sym_col(u) = [G00, G11, G22, (G01+G10)/2, (G02+G20)/2, (G12+G21)/2], and sigma = H eps
 * written with the Lame constants (lambda tr(eps) + 2 mu eps_ii, 2 mu eps_ij). */
int or_after_solve(int kind, uint32_t n_elem, const uint32_t *conn, const double *coords, const double *E_e,
                   double nu, const double *u, double *eps, double *sig, double *tr)
{
    const int npe = or_nodes_per_elem(kind);
    const double xi[3] = {0.5, 0.5, 0.5};
    double N[10], dN[10][3];
    shape(kind, xi, N, dN);
    for (uint32_t e = 0; e < n_elem; ++e) {
        const uint32_t *en = conn + (size_t)e * npe;
        double xe[30], J[3][3], Ji[3][3], G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < npe; ++a)
            for (int d = 0; d < 3; ++d) xe[3 * a + d] = coords[3 * (size_t)en[a] + d];
        const double det = jacobian(npe, xe, dN, J, Ji);
        if (!(det != 0)) return 2;
        for (int a = 0; a < npe; ++a) {
            double g[3];
            for (int j = 0; j < 3; ++j) g[j] = Ji[j][0] * dN[a][0] + Ji[j][1] * dN[a][1] + Ji[j][2] * dN[a][2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) G[i][j] += u[3 * (size_t)en[a] + i] * g[j];
        }
        double ev[6] = {G[0][0], G[1][1], G[2][2], 0.5 * (G[0][1] + G[1][0]), 0.5 * (G[0][2] + G[2][0]),
                        0.5 * (G[1][2] + G[2][1])};
        double lam, mu;
        lame(E_e[e], nu, &lam, &mu);
        const double t = ev[0] + ev[1] + ev[2];
        for (int i = 0; i < 6; ++i) {
            if (eps) eps[6 * (size_t)e + i] = ev[i];
            if (sig) sig[6 * (size_t)e + i] = i < 3 ? lam * t + 2 * mu * ev[i] : 2 * mu * ev[i];
        }
        if (tr) tr[e] = t;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Adjoint gradient                                                           */
/* ------------------------------------------------------------------------- */

int or_elem_grad_E(int kind, uint32_t n_elem, const uint32_t *conn, const double *coords, const double *E_e, double nu,
                   const double *u, const double *lam, double *g)
{
    const int npe = or_nodes_per_elem(kind), nd = 3 * npe;
    double Ke[900];
    (void)E_e;
    for (uint32_t e = 0; e < n_elem; ++e) {
        double xe[30], ue[30], le[30];
        const uint32_t *en = conn + (size_t)e * npe;
        for (int a = 0; a < npe; ++a)
            for (int d = 0; d < 3; ++d) {
                xe[3 * a + d] = coords[3 * (size_t)en[a] + d];
                ue[3 * a + d] = u[3 * (size_t)en[a] + d];
                le[3 * a + d] = lam[3 * (size_t)en[a] + d];
            }
        or_elem_matrix(kind, xe, 1.0, nu, NULL, Ke, NULL, NULL); /* dK_e/dE_e = K_e(E=1) */
        double s = 0;
        for (int r = 0; r < nd; ++r) { double t = 0; for (int c = 0; c < nd; ++c) t += Ke[r * nd + c] * ue[c]; s += le[r] * t; }
        g[e] = -s;
    }
    return 0;
}
