/*
 * fem_oracle.h -- CPU restatement of the LMT++ assemble + CG path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (lmtpp_amd/) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Paths are relative to the LMT++ source tree.
 *
 * Pinning status of each part (what checks it, see tests/test_oracle_*.py):
 *   or_make_rect        pinned: golden meshes dumped by the reference's own
 *                       make_rect (oracle/_ref), exact coordinates and connectivity.
 *   or_pattern          pinned: exact block counts of the reference meshes
 *                       (17,191 / 9,261 Tetra 11^3; 118,121 lower blocks Hexa 21^3
 *                       => 1,035,306 lower scalars), brute-force adjacency.
 *   or_elem_matrix      pinned by closed forms and invariants: Tetra4 V*B^T D B,
 *                       rigid-body null space, rank, patch test, linearity in E,
 *                       sympy-exact Hexa brick with the 12-digit Gauss table,
 *                       mpmath quadrature of the Tetra10 matrix.
 *                       No bit-level pin exists: the reference's generated
 *                       element kernels need its Python-2 generator.
 *   or_face_load        pinned by closed forms (A/3, A/4, Tri6 0 and A/3).
 *   or_skin_faces       pinned by geometry (test_oracle_pins.py): every boundary face of
 *                       the unit cube once (total area 6), oriented outward by its first
 *                       three nodes; Tetra_10 faces carry their mid-edge nodes at the
 *                       midpoints of edges (0,1), (1,2), (2,0) (element_Triangle_6.py);
 *                       a uniform traction loads only the mid-edge nodes, t A/3 per face.
 *   or_penalty          pinned against a dense K summed from or_elem_matrix: exactly
 *                       P * max|diag K| added at the constrained diagonal entries and
 *                       P * max|diag K| * value to F, nothing else; the solution holds
 *                       the prescribed values to O(1/P) (100x smaller defect at 100x P).
 *   or_block_ic0 / or_pcg_ic0
 *                       pinned against the reference's own MatWithTinyBlocks
 *                       chol_incomp + solve_using_incomplete_chol_factorize
 *                       compiled from source (oracle/_ref) on the same K, F.
 *   or_pcg_jacobi       pinned by the reference CG known-answer tests
 *                       (3x3 -> (11,2,4); 10x10 Maple solution) and by a dense
 *                       LAPACK solve of the assembled system.
 *   or_after_solve      pinned: Hexa against the reference's generated
 *                       after_solve (std_formulations/definition_probleme_
 *                       elasticity_dep_static.h:721-789, the same DSL expression)
 *                       compiled from source (oracle/_ref, golden fixture);
 *                       all kinds by closed forms (affine fields give the exact
 *                       constant strain, sigma = lambda tr I + 2 mu eps) and the
 *                       analytic gradient of a quadratic field on a straight
 *                       Tetra_10 at the point xi = (1/2,1/2,1/2).
 *   or_elem_grad_E      pinned by -lambda_e^T K_e(E_e) u_e / E_e from or_elem_matrix
 *                       and by central differences (step 1e-4 E_e) of the objective
 *                       with dense LAPACK solves, penalty scale held fixed
 *                       (test_oracle_pins.py); no reference counterpart exists (the
 *                       reference has no adjoint).
 */
#ifndef FEM_ORACLE_H
#define FEM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_TETRA = 0, OR_TETRA10 = 1, OR_HEXA = 2 };

/* Host threads of or_assemble, or_spmv and or_pcg_jacobi (OpenMP; n <= 0: all).  1, the
 * default, is the serial restatement in the reference's order; more threads colour the element
 * loop and chunk the dot products (the multi-core CPU line).  Returns the count set. */
int or_set_threads(int n);
enum { OR_TRI3 = 0, OR_TRI6 = 1, OR_QUAD4 = 2 };

int or_nodes_per_elem(int kind);
int or_nodes_per_face(int fkind);
int or_face_kind_of(int kind);

/* make_rect: node and element counts for nb_points[3] lattice points. */
int or_make_rect_counts(int kind, const unsigned nb_points[3], uint32_t *n_nodes, uint32_t *n_elem);
/* make_rect: fills coords (n_nodes*3) and conn (n_elem*npe). */
int or_make_rect(int kind, const double X0[3], const double X1[3], const unsigned nb_points[3],
                 double *coords, uint32_t *conn);
/* Tetra_10 with the reference's O(N^2) linear node search (tiny meshes only). */
int or_make_rect_tetra10_linear_search(const double X0[3], const double X1[3], const unsigned nb_points[3],
                                       double *coords, uint32_t *conn, uint32_t *n_nodes_out);

/* Boundary faces (faces with exactly one parent element), reference child ordering. */
int or_skin_faces(int kind, uint32_t n_nodes, uint32_t n_elem, const uint32_t *conn,
                  uint32_t *faces, uint32_t cap_faces, uint32_t *n_faces);

/* Element stiffness K_e ((3npe)^2, row-major, dof 3a+i) and load F_e (3npe) for a body force bf. */
int or_elem_matrix(int kind, const double *xe, double E, double nu, const double bf[3],
                   double *Ke, double *Fe, double *detj_min);
/* Face load F_f (3npf) for a uniform traction t. */
int or_face_load(int fkind, const double *xf, const double t[3], double *Ff);

/* Block pattern (node adjacency), full symmetric, rows sorted. Returns nnzb via *nnzb.
 * Call with col == NULL to get the count only. */
int or_pattern(int kind, uint32_t n_nodes, uint32_t n_elem, const uint32_t *conn,
               uint32_t *row_ptr, uint32_t *col, size_t *nnzb);

/* Assembly into full BSR-3 (vals: nnzb*9, block row-major) + F (3*n_nodes). E_e per element.
 * Faces: n_faces x npf nodes with per-face traction (n_faces*3). */
int or_assemble(int kind, uint32_t n_nodes, const double *coords, uint32_t n_elem, const uint32_t *conn,
                const double *E_e, double nu, const double bf[3],
                uint32_t n_faces, const uint32_t *faces, const double *face_trac,
                const uint32_t *row_ptr, const uint32_t *col, double *vals, double *F);

/* Penalty Dirichlet: K(p,p) += M*P, F(p) += M*P*value, M = max|diag K| (before constraints). */
int or_penalty(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, double *vals, double *F,
               uint32_t n_cons, const uint32_t *cons_dof, const double *cons_val, double P, double *max_diag);

void or_spmv(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals,
             const double *x, double *y);

/* Jacobi (scalar or 3x3 block) PCG, relative L2 residual stop. Returns 0 converged, 4 max-iter. */
int or_pcg_jacobi(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals,
                  const double *b, double *x, double rel_tol, int max_iter, int block_jacobi,
                  int *iters, double *rel_res);

/* The same, recording after iteration i (i < hist_cap) the relative recurrence residual in
 * hist[2i] and the seconds since the call began (set-up included) in hist[2i+1]. */
int or_pcg_jacobi_hist(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals,
                       const double *b, double *x, double rel_tol, int max_iter, int block_jacobi, int *iters,
                       double *rel_res, double *hist, int hist_cap);

/* Generic scalar CSR CG with scalar-Jacobi or no preconditioner (KAT tests). */
int or_cg_csr(uint32_t n, const uint32_t *row_ptr, const uint32_t *col, const double *vals,
              const double *b, double *x, double rel_tol, int max_iter, int *iters);

/* Block IC(0) on the lower block pattern.  L is returned in full-BSR positions
 * (only c <= r used; diagonal block lower-triangular).  Returns 3 (NOT_SPD) on a
 * non-positive pivot. */
int or_block_ic0(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals,
                 double *L);
/* PCG with the block IC(0) factor and the reference's stop test max|r_i| <= crit. */
int or_pcg_ic0(uint32_t n_nodes, const uint32_t *row_ptr, const uint32_t *col, const double *vals,
               const double *L, const double *b, double *x, double crit, int max_iter, int *iters);

/* After-solve recovery (apply_on_elements_after_solve,
 * formulations/formulation_elasticity_iso.py:64-85): per element at xi = (1/2,1/2,1/2),
 * eps = grad_sym_col(u) (Voigt xx,yy,zz,xy,xz,yz, tensor shear), sigma = H eps, tr_eps.
 * eps, sig: 6 per element; tr: one per element. */
int or_after_solve(int kind, uint32_t n_elem, const uint32_t *conn, const double *coords, const double *E_e,
                   double nu, const double *u, double *eps, double *sig, double *tr);

/* Per-element gradient of J w.r.t. E_e given u and the adjoint lambda:
 * g_e = -lambda_e^T (dK_e/dE_e) u_e  (K_e linear in E_e). */
int or_elem_grad_E(int kind, uint32_t n_elem, const uint32_t *conn, const double *coords,
                   const double *E_e, double nu, const double *u, const double *lam, double *g);

#ifdef __cplusplus
}
#endif
#endif
