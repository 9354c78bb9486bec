/*
 * femgpu.h -- C ABI of the MI355X (gfx950) assemble + solve path for LMT++-style
 * linear elasticity (Tetra, Tetra_10, Hexa).
 *
 * The entry points replace the hot-path virtuals of LMT++'s
 * FormulationAncestor<double> (include/formulation/formulation_ancestor.h) as
 * implemented by Formulation<...> (include/formulation/formulation.h); each
 * declaration cites the reference interface it stands in for.  Paths are
 * relative to the LMT++ source tree.
 *
 * Conventions
 *   - every call returns an int status (FEMGPU_OK == 0, see femgpu_status);
 *   - host buffers are caller-owned and copied; "_device" variants take device
 *     pointers on the problem's GPU and are read/written on the problem's stream
 *     after the call has synchronised that stream;
 *   - a problem owns its device buffers and one HIP stream; a problem is not
 *     thread-safe, distinct problems are independent;
 *   - DOF numbering is DOF(node n, component d) = 3n + d (formulation.h:376).
 */
#ifndef FEMGPU_H
#define FEMGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEMGPU_API __attribute__((visibility("default")))

typedef enum {
    FEMGPU_OK = 0,
    FEMGPU_BAD_INPUT = 1,     /* sizes, indices or options out of range */
    FEMGPU_NEG_JACOBIAN = 2,  /* an element has det J <= 0 at a Gauss point */
    FEMGPU_NOT_SPD = 3,       /* non-positive pivot / curvature (p.Ap <= 0) */
    FEMGPU_MAX_ITER = 4,      /* iteration cap reached before the stop test */
    FEMGPU_HIP_ERROR = 5,     /* a HIP runtime call failed */
    FEMGPU_NAN = 6,           /* NaN or Inf met in a reduction */
    FEMGPU_NOT_READY = 7,     /* call order violated (e.g. solve before assemble) */
    FEMGPU_STAGNATED = 8      /* the FP64 residual b - K u stopped decreasing above the
                                 tolerance (the tolerance is below what FP64 attains here) */
} femgpu_status;

typedef enum { FEMGPU_TETRA = 0, FEMGPU_TETRA10 = 1, FEMGPU_HEXA = 2 } femgpu_elem_kind;
typedef enum { FEMGPU_TRI3 = 0, FEMGPU_TRI6 = 1, FEMGPU_QUAD4 = 2 } femgpu_face_kind;

typedef enum { FEMGPU_FP64 = 0, FEMGPU_MIXED = 1 } femgpu_precision;
/* FEMGPU_BLOCK_IC0: the reference's block incomplete Cholesky (chol_incomp,
 * containers/MatWithTinyBlocks.h:426-523) in natural node order -- the reference's factor,
 * level-scheduled; FEMGPU_BLOCK_IC0_COLOURED: the same factorisation in a multicolour node
 * order (a different, GPU-friendly preconditioner with a few wide levels).
 * FEMGPU_CHEBYSHEV_BLOCK_JACOBI (no counterpart in the reference): z = p(D K) D r with D the
 * block-Jacobi inverse and p the degree-1 polynomial whose residual 1 - t p(t) is the scaled
 * Chebyshev polynomial of degree 2 on [b/20, b], b = 1.15 x a power-iteration estimate of the
 * largest eigenvalue of D K (20 FP64 steps at set-up); one more product per iteration, about half
 * the iterations of block Jacobi on elasticity.  A sub-problem on which it is not positive
 * definite stops with FEMGPU_NOT_SPD. */
typedef enum {
    FEMGPU_JACOBI = 0,
    FEMGPU_BLOCK_JACOBI = 1,
    FEMGPU_BLOCK_IC0 = 2,
    FEMGPU_BLOCK_IC0_COLOURED = 3,
    FEMGPU_CHEBYSHEV_BLOCK_JACOBI = 4
} femgpu_precond;
/* stop tests: relative L2 residual, or the reference's max_i |r_i| <= crit
 * (containers/MatWithTinyBlocks.h:1115,1136) */
typedef enum { FEMGPU_REL_L2 = 0, FEMGPU_ABS_MAXNORM = 1 } femgpu_crit_kind;

typedef struct femgpu_problem femgpu_problem;

typedef struct {
    int elem_kind;           /* femgpu_elem_kind */
    uint32_t n_nodes;
    const double *coords;    /* n_nodes * 3 */
    uint32_t n_elem;
    const uint32_t *conn;    /* n_elem * nodes_per_elem, reference local node order */
} femgpu_mesh;

typedef struct {
    double nu;                /* poisson_ratio (formulations/formulation_elasticity_iso.py:6) */
    double E;                 /* elastic_modulus when E_elem == NULL (global interpolation) */
    const double *E_elem;     /* per-element modulus (interpolation='elementary'), or NULL */
    double body_force[3];     /* density * f_vol (formulation_elasticity_iso.py:52) */
} femgpu_material;

typedef struct {
    int precision;            /* femgpu_precision */
    int precond;              /* femgpu_precond */
    int crit_kind;            /* femgpu_crit_kind */
    double tol;               /* REL_L2: ||r|| <= tol ||b||;  ABS_MAXNORM: max|r_i| <= tol, with
                                 r = b - K u formed in FP64 at the end: a recurrence residual that
                                 meets the test only triggers that check; a solve whose true
                                 residual stops decreasing above the tolerance ends with
                                 FEMGPU_STAGNATED, one that runs out of iterations with
                                 FEMGPU_MAX_ITER */
    int max_iter;             /* inner iterations cap (the reference has none) */
    double reliable_delta;    /* MIXED: FP64 residual refresh when the FP32 residual fell by this factor */
    int chunk;                /* iterations enqueued between host checks; a single problem's
                                 chunk ends at its predicted next check or refresh instead, up to
                                 4 x chunk once its convergence rate is known */
} femgpu_solve_opts;

typedef struct {
    int status;
    int iters;                /* CG iterations (inner FP32 iterations in MIXED mode) */
    int refreshes;            /* FP64 true-residual checks that restarted or ended the
                                 recurrence (MIXED: the outer refinement steps) */
    double rel_res;           /* final ||b - K u||_2 / ||b||_2, FP64 */
    double max_abs_res;       /* final max_i |b - K u|_i, FP64 */
    double solve_ms;          /* device time of the solve (HIP events on the problem's stream) */
} femgpu_solve_stats;

/* ---- lifecycle ---------------------------------------------------------------------------- */

/* Stands in for `new Formulation<TM,elasticity_iso,...>(m)` made by the generated factory
 * Problem_<name>::new_formulation_<f> (__init__.py:167-168,205-209) plus
 * update_connectivity / allocate_matrices (formulation.h:283-412): copies the mesh, builds the
 * DOF numbering, the 3x3-block pattern (ResizeMat::GetNZ, formulation.h:200-230), the element
 * colouring and the scatter map, and uploads everything to the current HIP device. */
FEMGPU_API int femgpu_create(const femgpu_mesh *mesh, femgpu_problem **out);
/* ProblemAncestor deletes its formulations (problem_ancestor.h:35-38). */
FEMGPU_API int femgpu_destroy(femgpu_problem *p);

/* ---- inputs ------------------------------------------------------------------------------- */

/* Material fields: elastic_modulus / poisson_ratio / density*f_vol of
 * formulations/formulation_elasticity_iso.py:4-7 (global or elementary interpolation). */
FEMGPU_API int femgpu_set_material(femgpu_problem *p, const femgpu_material *mat);
/* Per-element modulus from a device array of n_elem doubles. */
FEMGPU_API int femgpu_set_elastic_modulus_device(femgpu_problem *p, const double *d_E);
/* Skin loads: f_surf (interpolation='skin_elementary', formulation_elasticity_iso.py:22) on
 * boundary faces, integrated by the skin loop AssembleElem::ass_skin_elem
 * (formulation.h:584-594).  faces: n_faces * nodes_per_face in the reference's child order. */
FEMGPU_API int femgpu_set_traction(femgpu_problem *p, uint32_t n_faces, const uint32_t *faces, const double *trac);
/* add_constraint(txt, penalty) (formulation_ancestor.h:197) for u[dof] = value with a unit
 * coefficient, penalty branch of assemble_constraints (formulation.h:1028-1070). */
FEMGPU_API int femgpu_add_constraint(femgpu_problem *p, uint32_t dof, double value, double penalty);
FEMGPU_API int femgpu_add_constraints(femgpu_problem *p, uint32_t n, const uint32_t *dofs, const double *values,
                                      double penalty);
FEMGPU_API int femgpu_clear_constraints(femgpu_problem *p);
/* add_constraint(txt, penalty_value) (formulation_ancestor.h:197, formulation.h:2618; parsed as
 * Constraint, constraint.h:46-93): txt is an expression whose zero is the constraint, linear in
 * the unknowns node[i].dep[d] (caller's node numbering), with numbers, `time`, pi, + - * /
 * ** (or ^), parentheses and sin cos tan exp log sqrt abs of unknown-free arguments, e.g.
 * "node[12].dep[2] + 0.001*time".  Its penalty terms are those of assemble_constraints
 * (formulation.h:1058-1070): with the expression c u + b, K += M P c^2 and F += M P c (-b),
 * M = max|diag K|; an expression depending on time is re-evaluated at each assembly at the
 * formulation's time.  Expressions with more than one unknown (couplings between DOFs) are
 * refused (FEMGPU_BAD_INPUT), as is penalty <= 0 (the Lagrange form).  *number: the constraint's
 * index among all constraints, as the reference returns it. */
FEMGPU_API int femgpu_add_constraint_text(femgpu_problem *p, const char *txt, double penalty, uint32_t *number);
/* The parsing of femgpu_add_constraint_text alone (host only, no problem): at time t the
 * expression is coeff[k] u[dof[k]] summed over k < *n_terms, plus *b; dof / coeff hold cap
 * entries (*n_terms may exceed cap: then only the count is written).  FEMGPU_BAD_INPUT with the
 * message in femgpu_last_error() for a malformed or non-linear expression. */
FEMGPU_API int femgpu_constraint_parse(const char *txt, double t, uint32_t cap, uint64_t *dof, double *coeff,
                                       uint32_t *n_terms, double *b);

/* ---- hot path ----------------------------------------------------------------------------- */

/* assemble(assemble_mat, assemble_vec) (formulation_ancestor.h:63; formulation.h:1112-1125):
 * element loop (formulation.h:626-647) + skin loads + penalty constraints. */
FEMGPU_API int femgpu_assemble(femgpu_problem *p, int assemble_mat, int assemble_vec);
/* solve_system(crit) (formulation_ancestor.h:64; formulation.h:1494-1504 and the CG branch
 * :1487-1492, containers/MatWithTinyBlocks.h:1101-1151).  The initial guess is the current
 * result vector (vectors[0] after shift(), formulation.h:424-433). */
FEMGPU_API int femgpu_solve_system(femgpu_problem *p, const femgpu_solve_opts *opts, femgpu_solve_stats *stats);
/* solve(crit) (formulation_ancestor.h:49; formulation.h:1517-1568): shift(), then assemble +
 * solve_system, repeated while a non-linear criterium is set (femgpu_set_nonlinear) and the max
 * norm of the change since the previous iterate (zero at first) exceeds it; FEMGPU_MAX_ITER after
 * max_non_linear_iteration + 1 solves (the reference throws SolveException).  With no criterium
 * (the default) one assemble + solve_system. */
FEMGPU_API int femgpu_solve(femgpu_problem *p, const femgpu_solve_opts *opts, femgpu_solve_stats *stats);
FEMGPU_API void femgpu_default_solve_opts(femgpu_solve_opts *opts);

/* ---- results ------------------------------------------------------------------------------ */

/* get_result(0) / update_variables (formulation_ancestor.h:237,179; formulation.h:2571-2576). */
FEMGPU_API int femgpu_get_result(femgpu_problem *p, double *u, size_t n);
FEMGPU_API int femgpu_get_result_device(femgpu_problem *p, double *d_u);
/* Writes vectors[0], the initial guess of the next solve (shift(), formulation.h:424-433). */
FEMGPU_API int femgpu_set_result(femgpu_problem *p, const double *u, size_t n);
/* get_result(num) (formulation_ancestor.h:237, formulation.h:2648): vectors[num], num <
 * nb_vectors = 20 (formulation.h:112): num = 0 the current solution, num = k the solution k shifts
 * ago (zeros where no shift has filled the slot yet). */
FEMGPU_API int femgpu_get_result_history(femgpu_problem *p, uint32_t num, double *u, size_t n);
/* get_indice_noda(num_node) (formulation_ancestor.h:188, formulation.h:2655): the index of the
 * node's first unknown in every vector of this API (3 num_node; with FEMGPU_NODE_ORDER_RCM the
 * internal storage order is private and converted at the API, so the index stays 3 num_node). */
FEMGPU_API int femgpu_get_indice_noda(femgpu_problem *p, uint32_t node, uint32_t *index);
/* get_nodal_forces() (formulation_ancestor.h:65, formulation.h:447-458): the sollicitation of the
 * element and skin loops alone (assemble_clean_mat(false, true): body and surface loads, no
 * constraint terms), the assembled sollicitation left unchanged. */
FEMGPU_API int femgpu_get_nodal_forces(femgpu_problem *p, double *F, size_t n);
/* get_sollicitation() (formulation_ancestor.h:236). */
FEMGPU_API int femgpu_get_sollicitation(femgpu_problem *p, double *F, size_t n);
/* Writes the sollicitation that the next solve_system solves for: the reference's
 * get_sollicitation() returns Vec<T>& (formulation.h:2647), which callers modify in place
 * between assemble() and solve_system().  A later vector assembly overwrites it. */
FEMGPU_API int femgpu_set_sollicitation(femgpu_problem *p, const double *F, size_t n);
/* get_mat() (formulation_ancestor.h:230): the assembled K as full 3x3-block CSR, block row-major
 * values (both triangles; the reference stores the lower one).  Any pointer may be NULL. */
FEMGPU_API int femgpu_get_matrix(femgpu_problem *p, uint32_t *row_ptr, uint32_t *col, double *vals);
FEMGPU_API int femgpu_get_sizes(femgpu_problem *p, uint32_t *n_nodes, uint32_t *n_elem, uint64_t *nnzb,
                                uint32_t *n_colors);
/* Block IC(0) factor of the assembled K (chol_incomp, containers/MatWithTinyBlocks.h:426-523;
 * factorised on the GPU if not current) for precond FEMGPU_BLOCK_IC0 or _COLOURED, in
 * elimination numbering: order[t] = node eliminated t-th (n_nodes), row_ptr (n_nodes + 1) and
 * col (nnzb) = the full pattern renumbered by t, rows sorted; L (9 nnzb, block row-major) holds
 * L_tc for c < t, the lower-triangular diagonal factor for c = t and zeros for c > t.  Any
 * output may be NULL; *n_levels receives the number of level sets. */
FEMGPU_API int femgpu_get_ic0_factor(femgpu_problem *p, int precond, uint32_t *order, uint32_t *row_ptr,
                                     uint32_t *col, double *L, uint32_t *n_levels);
/* max|diag K| taken before the constraints (formulation.h:1029-1030). */
FEMGPU_API int femgpu_get_max_diag(femgpu_problem *p, double *max_diag);

/* ---- after solve ------------------------------------------------------------------------- */

/* call_after_solve() (formulation.h:2026-2050) running the generated after_solve of
 * apply_on_elements_after_solve (formulations/formulation_elasticity_iso.py:64-85; the Hexa
 * instance is generated in std_formulations/definition_probleme_elasticity_dep_static.h:721-789)
 * on the current result: per element, at xi = (1/2, 1/2, 1/2), eps = grad_sym_col(u) (Voigt xx,
 * yy, zz, xy, xz, yz, tensor shear), sigma = H(E_e, nu) eps and tr_epsilon.  eps and sigma hold
 * 6 values per element (element-major), tr_eps one; any output may be NULL. */
FEMGPU_API int femgpu_after_solve(femgpu_problem *p, double *eps, double *sigma, double *tr_eps);
FEMGPU_API int femgpu_after_solve_device(femgpu_problem *p, double *d_eps, double *d_sigma, double *d_tr_eps);
/* Device time (ms) of the last after_solve kernel. */
FEMGPU_API int femgpu_get_after_solve_ms(femgpu_problem *p, double *ms);

/* ---- differentiation ---------------------------------------------------------------------- */

/* Adjoint of the assembled system: solves K lambda = dJ/du and returns
 * dJ/dE_e = -lambda_e^T (dK_e/dE_e) u_e for every element.  Replaces the forward-sensitivity
 * loop of FormulationFit::exec / solve_and_get_derivatives (FormulationFit.h:19-85,
 * formulation.h:1588-1652), which costs one solve per parameter.  The penalty scale
 * max|diag K| is held fixed. */
FEMGPU_API int femgpu_adjoint(femgpu_problem *p, const double *dJdu, double *dJdE, const femgpu_solve_opts *opts,
                              femgpu_solve_stats *stats);
FEMGPU_API int femgpu_adjoint_device(femgpu_problem *p, const double *d_dJdu, double *d_dJdE,
                                     const femgpu_solve_opts *opts, femgpu_solve_stats *stats);
/* Element-gradient kernel alone for given u and lambda (host arrays, 3*n_nodes each). */
FEMGPU_API int femgpu_elem_grad(femgpu_problem *p, const double *u, const double *lam, double *dJdE);

/* ---- generic block systems ---------------------------------------------------------------- */

/* MatWithTinyBlocks<double,Sym<3>> built from a caller's matrix (MatWithTinyBlocks.h:224-269)
 * and solved by the same PCG: full 3x3-block CSR (both triangles, rows sorted). */
FEMGPU_API int femgpu_create_from_bsr(uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col,
                                      const double *vals, const double *rhs, femgpu_problem **out);

/* ---- mesh utilities (host, no GPU) -------------------------------------------------------- */

/* make_rect(m, Elem(), X0, X1, nb_points) (mesh/make_rect.h:389-430 Tetra, :443-544 Tetra_10,
 * :653-679 Hexa).  Call with coords == conn == NULL to get the counts. */
FEMGPU_API int femgpu_make_rect(int elem_kind, const double X0[3], const double X1[3], const uint32_t nb_points[3],
                                double *coords, uint32_t *conn, uint32_t *n_nodes, uint32_t *n_elem);
/* Skin faces (faces with one parent element, mesh.h:151-188) in the reference child order;
 * call with faces == NULL to get the count. */
FEMGPU_API int femgpu_skin_faces(int elem_kind, uint32_t n_nodes, uint32_t n_elem, const uint32_t *conn,
                                 uint32_t *faces, uint32_t *n_faces);

/* Elimination order and level sets of the block IC(0) for a full 3x3-block CSR pattern (both
 * triangles, rows sorted, diagonal present): order (n_brows), level_ptr (n_levels + 1; may be
 * NULL: *n_levels then only receives the count; else *n_levels gives its capacity in levels).
 * Rows of one level share no block.  No reference counterpart: chol_incomp runs row by row. */
FEMGPU_API int femgpu_ic0_schedule(uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col, int precond,
                                   uint32_t *order, uint32_t *level_ptr, uint32_t *n_levels);

/* ---- result files ----------------------------------------------------------------------- */

/* One named field of a result file: n_comp doubles per node or element (one item for a global
 * field). */
typedef struct {
    const char *name;
    uint32_t n_comp;
    const double *data;
} femgpu_vtu_field;
/* write_mesh_vtk(filename, m) (mesh/write_mesh_vtk.h:827-831 -> write_mesh_vtk<true>,
 * :434-656): a VTK XML unstructured grid with the reference writer's layout and bytes -- point
 * fields in the given order (2-component fields padded to 3), cell fields in name order, the
 * element group in ASCII (group NULL: 0), global fields repeated for every cell, the points, the
 * cells (VTK types 10 / 24 / 12) in ASCII, and every float array base64-encoded in one appended
 * block.  Host only. */
FEMGPU_API int femgpu_write_vtu(const char *path, int elem_kind, uint32_t n_nodes, const double *coords,
                                uint32_t n_elem, const uint32_t *conn, uint32_t n_point,
                                const femgpu_vtu_field *point, uint32_t n_cell, const femgpu_vtu_field *cell,
                                const uint32_t *group, uint32_t n_global, const femgpu_vtu_field *global);
/* The problem's mesh and current result in that layout: dep; elastic_modulus, epsilon, sigma,
 * tr_epsilon (call_after_solve of the current result); poisson_ratio. */
FEMGPU_API int femgpu_write_result_vtu(femgpu_problem *p, const char *path);

/* One named field of an AVS file.  nest is the shape of the value in the reference's mesh data,
 * which fixes its spacing in the file: 0 a scalar (n_comp 1), 1 a Vec<T,n_comp>, 2 a
 * Vec<Vec<T,n_comp>,1> (the element fields epsilon and sigma). */
typedef struct {
    const char *name;
    uint32_t n_comp;
    uint32_t nest;
    const double *data;
} femgpu_avs_field;
/* write_avs(mesh, name, display_fields, Ascii()) (mesh/write_avs.h:389-530): an AVS UCD ASCII
 * file with the reference writer's bytes -- the header (nodes, elements, then the component sums
 * of the node, cell and global fields written, 2-component node and global fields counted as 3),
 * the nodes, the cells ("hex" / "tet"; a Tetra_10 cell, which that writer has no type for, as its
 * number and material 0 alone, with no line end), then the node fields in the given order, the
 * cell fields in name order and the global fields in the given order, each written when its name
 * is in display[0..n_display) or display[0] is "all".  Host only. */
FEMGPU_API int femgpu_write_avs(const char *path, int elem_kind, uint32_t n_nodes, const double *coords,
                                uint32_t n_elem, const uint32_t *conn, uint32_t n_point,
                                const femgpu_avs_field *point, uint32_t n_cell, const femgpu_avs_field *cell,
                                uint32_t n_global, const femgpu_avs_field *global, uint32_t n_display,
                                const char *const *display);
/* The problem's mesh and current result in that layout (write_avs(mesh, num, structure, Ascii()),
 * write_avs.h:548-562, without its path and numbering): the node fields dep, pos; the cell fields
 * elastic_modulus, epsilon, sigma, tr_epsilon (call_after_solve of the current result); the
 * global fields poisson_ratio, f_vol; display as for femgpu_write_avs (NULL or none: "all"). */
FEMGPU_API int femgpu_write_result_avs(femgpu_problem *p, const char *path, uint32_t n_display,
                                       const char *const *display);

/* ---- time steps, non-linear loop, preconditioner reuse ------------------------------------- */

/* shift(nb) (formulation.h:424-433): time += time_steps[0], time_steps and the solution history
 * vectors[0..19] move up one slot (vectors[i] = vectors[i-1]); vectors[0] stays, the initial guess
 * of the next solve. */
FEMGPU_API int femgpu_shift(femgpu_problem *p, int nb);
/* unshift(nb) (formulation_ancestor.h:58, formulation.h:435-445): time -= time_steps[1],
 * time_steps and vectors move down one slot (vectors[0] takes vectors[1]), the last slot kept. */
FEMGPU_API int femgpu_unshift(femgpu_problem *p, int nb);
/* set_initial_time_step(ts) (formulation.h:2624): time step ts, time = -2 ts.  The time step
 * starts at 1e40 and time at 0 (formulation.h:128-129). */
FEMGPU_API int femgpu_set_initial_time_step(femgpu_problem *p, double ts);
/* time_steps[0] = ts (the public member the reference's callers set between steps). */
FEMGPU_API int femgpu_set_next_time_step(femgpu_problem *p, double ts);
/* get_next_time_step / get_time / set_time (formulation.h:2625-2627). */
FEMGPU_API int femgpu_get_next_time_step(femgpu_problem *p, double *ts);
FEMGPU_API int femgpu_get_time(femgpu_problem *p, double *time);
FEMGPU_API int femgpu_set_time(femgpu_problem *p, double time);
/* vectors[1]: the solution before the last shift (node-major, 3 n_nodes doubles; zeros before the
 * first shift), femgpu_get_result_history(p, 1, ...). */
FEMGPU_API int femgpu_get_previous_result(femgpu_problem *p, double *u, size_t n);
/* assume_constant_matrix (formulation_ancestor.h:31,244): ProblemAncestor::iteration then
 * assembles K once, builds the preconditioner once (get_precond) and afterwards assembles only the
 * right-hand side. */
FEMGPU_API int femgpu_set_assume_constant_matrix(femgpu_problem *p, int flag);
/* non_linear_iterative_criterium and max_non_linear_iteration (formulation_ancestor.h:33-35;
 * defaults 0 and 50). */
FEMGPU_API int femgpu_set_nonlinear(femgpu_problem *p, double crit, int max_iter);
/* solve_system calls made by the last femgpu_solve or stepper iteration. */
FEMGPU_API int femgpu_get_nonlinear_solves(femgpu_problem *p, int *n);
/* default_iterative_criterium and the other solve options ProblemAncestor::iteration uses for
 * this formulation (formulation_ancestor.h:32; femgpu_default_solve_opts until set). */
FEMGPU_API int femgpu_set_default_solve_opts(femgpu_problem *p, const femgpu_solve_opts *opts);
/* get_precond() (formulation.h:1655,1785-1795): builds opts->precond (FP64 and FP32 copies) for
 * the current K and keeps it, through later assemblies, for solve_system_using_precond.  The
 * reference factorises incomplete Cholesky there: FEMGPU_BLOCK_IC0 gives the same factor. */
FEMGPU_API int femgpu_get_precond(femgpu_problem *p, const femgpu_solve_opts *opts);
/* solve_system_using_precond(crit) (formulation.h:1656,1796-1805): PCG with get_precond's
 * preconditioner (opts->precond must be the same kind). */
FEMGPU_API int femgpu_solve_system_using_precond(femgpu_problem *p, const femgpu_solve_opts *opts,
                                                 femgpu_solve_stats *stats);

/* ProblemAncestor (problem_ancestor.h:21-140) over n formulations, which stay owned by the
 * caller.  iteration() (problem_ancestor.h:41-129): the first call shifts and assembles every
 * formulation and builds the preconditioner of those assuming a constant matrix; later calls
 * update only the formulations whose time equals the problem's time, shifting them and
 * reassembling F (and K unless constant).  Constant-matrix formulations then solve with the kept
 * preconditioner; the others run the non-linear loop (2-norm of the change, reassembling between
 * solves).  The problem's time becomes the smallest formulation time.  stats and updated receive
 * one entry per formulation (either may be NULL). */
typedef struct femgpu_stepper femgpu_stepper;
FEMGPU_API int femgpu_stepper_create(uint32_t n, femgpu_problem *const *forms, femgpu_stepper **out);
FEMGPU_API int femgpu_stepper_destroy(femgpu_stepper *s);
FEMGPU_API int femgpu_stepper_iteration(femgpu_stepper *s, femgpu_solve_stats *stats, int *updated);
FEMGPU_API int femgpu_stepper_get_time(femgpu_stepper *s, double *time);

/* ---- batched uneven meshes ---------------------------------------------------------------- */

typedef struct femgpu_batch femgpu_batch;
/* n independent problems assembled and solved together (no reference counterpart: the
 * reference solves one formulation at a time).  Each problem keeps its own scalars. */
FEMGPU_API int femgpu_batch_create(uint32_t n, const femgpu_mesh *meshes, femgpu_batch **out);
FEMGPU_API int femgpu_batch_set_material(femgpu_batch *b, uint32_t i, const femgpu_material *mat);
FEMGPU_API int femgpu_batch_set_traction(femgpu_batch *b, uint32_t i, uint32_t n_faces, const uint32_t *faces,
                                         const double *trac);
FEMGPU_API int femgpu_batch_add_constraints(femgpu_batch *b, uint32_t i, uint32_t n, const uint32_t *dofs,
                                            const double *values, double penalty);
FEMGPU_API int femgpu_batch_assemble(femgpu_batch *b);
FEMGPU_API int femgpu_batch_solve(femgpu_batch *b, const femgpu_solve_opts *opts, femgpu_solve_stats *per_problem,
                                  double *total_ms);
FEMGPU_API int femgpu_batch_get_result(femgpu_batch *b, uint32_t i, double *u, size_t n);
/* Zero initial guess for every problem of the batch (each solve otherwise starts from the previous
 * result, as shift() does for one formulation). */
FEMGPU_API int femgpu_batch_clear_result(femgpu_batch *b);
/* after_solve of problem i of the batch (outputs sized for that problem's elements). */
FEMGPU_API int femgpu_batch_after_solve(femgpu_batch *b, uint32_t i, double *eps, double *sigma, double *tr_eps);
/* femgpu_write_result_vtu for problem i of a batch (its own mesh, group = i). */
FEMGPU_API int femgpu_batch_write_result_vtu(femgpu_batch *b, uint32_t i, const char *path);
/* femgpu_write_result_avs for problem i of a batch (its own mesh and material). */
FEMGPU_API int femgpu_batch_write_result_avs(femgpu_batch *b, uint32_t i, const char *path, uint32_t n_display,
                                             const char *const *display);
/* Device time (ms) of the last batched assembly. */
FEMGPU_API int femgpu_batch_get_assemble_ms(femgpu_batch *b, double *ms);
/* The element-loop variant of a batch (femgpu_set_assembly_rows' values); every variant gives
 * each sub-problem bitwise the K and F it gets alone. */
FEMGPU_API int femgpu_batch_set_assembly_rows(femgpu_batch *b, uint32_t rows);
FEMGPU_API int femgpu_batch_get_assembly_rows(femgpu_batch *b, uint32_t *rows, uint32_t *max_rows);
FEMGPU_API int femgpu_batch_destroy(femgpu_batch *b);

/* ---- diagnostics -------------------------------------------------------------------------- */

FEMGPU_API const char *femgpu_status_string(int status);
FEMGPU_API const char *femgpu_last_error(void);
/* Device time (ms) of the last assembly (element + skin + constraint kernels). */
FEMGPU_API int femgpu_get_assemble_ms(femgpu_problem *p, double *ms);
/* Device time (ms) of the preconditioner set-up of the last solve (Jacobi factors, the block
 * IC(0) factorisation and its FP32 copy, or the block-Jacobi factors and the power iteration of
 * the Chebyshev preconditioner); 0 when the previous set-up was reused. */
FEMGPU_API int femgpu_get_precond_setup_ms(femgpu_problem *p, double *ms);
/* Average device time (ms) of one block-SpMV launch (precision FP64 or the FP32 copy), measured
 * with HIP events around reps back-to-back launches on the problem's stream, after one untimed
 * launch. */
FEMGPU_API int femgpu_time_spmv(femgpu_problem *p, int precision, int reps, double *avg_ms);

/* Average device time (ms) of a plain copy of n doubles (n even; 16 B per lane) over reps
 * launches: the measured-bandwidth calibration (moves 16 n bytes per launch). */
FEMGPU_API int femgpu_stream_copy(size_t n, int reps, double *avg_ms);
/* The same for a streaming read of n doubles (n even; moves 8 n bytes per launch): the ceiling
 * for read-dominated kernels such as the block SpMV. */
FEMGPU_API int femgpu_stream_read(size_t n, int reps, double *avg_ms);

/* ---- tuning (no reference counterpart: the reference has no device code) ------------------- */

/* Block-SpMV storage formats: FEMGPU_SPMV_FULL stores both triangles (each row all its blocks);
 * FEMGPU_SPMV_SYM stores each row's upper blocks and reads the lower ones transposed from their
 * owner row, as MatWithTinyBlocks::partial_mul uses each stored block twice
 * (containers/MatWithTinyBlocks.h:303-378); FEMGPU_SPMV_MATRIX_FREE (Hexa and Tetra_10 problems)
 * evaluates K x element by element from the mesh with the element loop's quadrature
 * (formulation.h:626-647), the material and penalties of the last matrix assembly, colour by colour
 * without atomics; residual products of that precision keep the full storage with the configured
 * unroll and loads. */
typedef enum { FEMGPU_SPMV_FULL = 0, FEMGPU_SPMV_SYM = 1, FEMGPU_SPMV_MATRIX_FREE = 2 } femgpu_spmv_format;
/* y = K x (n = 3 n_nodes, caller's node order) with the current product variant of precision
 * FEMGPU_FP64 or the FP32 one (x rounded to float, y widened): the operator the solver iterates
 * with, as Mat * Vec computes the stored matrix's product (containers/MatWithTinyBlocks.h:303-378). */
FEMGPU_API int femgpu_apply(femgpu_problem *p, int precision, const double *x, double *y, size_t n);
/* No counterpart in the reference: femgpu_apply without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  d_y = K d_x on device pointers of the problem's GPU, n = 3 * n_nodes values each,
 * in the caller's node order, with the product variant the solver uses for precision (FP64, or
 * MIXED for the FP32 one: x rounded to float, the result widened).  Every temporary lives in a
 * static 64 MiB device array per GPU: the call only enqueues kernels on the problem's stream
 * (synchronise it before reading d_y) and never synchronises or allocates.  FEMGPU_NOT_READY until
 * the product variant was built for the current matrix (by femgpu_apply or a solve of that
 * precision, which allocate); FEMGPU_BAD_INPUT when the product needs more scratch than the array
 * holds (the FP32 copies of x and y, 24 n bytes in all, the matrix-free product's shared-node
 * partials, and 16 n bytes more with an internal node order).  Calls on one GPU share the array:
 * calls that overlap on one GPU race. */
FEMGPU_API int femgpu_apply_noalloc(femgpu_problem *p, int precision, const double *d_x, double *d_y);
/* The matrix-free product's layout: info[0] element blocks (256 element slots each, sorted along a
 * Morton curve), [1] sub-colours, [2] block-local node entries, [3] nodes shared by several blocks,
 * [4] their partial slots, [5] nodes of no element, [6] elements that are axis-aligned boxes (their
 * constant diagonal Jacobian takes the reduced product; when every element is one, a kernel
 * without the general path serves). */
FEMGPU_API int femgpu_matrix_free_info(femgpu_problem *p, uint64_t info[7]);
/* Average device times (ms) of the matrix-free element-block kernel alone and of the whole product
 * (element blocks, shared nodes, penalties) of precision FEMGPU_FP64 or the FP32 one, over reps
 * launches each on the problem's stream (needs an assembled matrix). */
FEMGPU_API int femgpu_time_matrix_free(femgpu_problem *p, int precision, int reps, double *block_ms,
                                       double *product_ms);
/* Average device times (ms) of the two kernels of the CG iteration on this problem (block Jacobi;
 * the matrix-free product of precision FEMGPU_FP64 or the FP32 one with the previous vector
 * update's end folded in, then the vector update), on the problem's stream with graph replay off:
 * a solve of iters + 4 iterations with HIP events around each launch gives the two kernels'
 * shares (the first 4 iterations left out), a second one with events around its chunks only the
 * iteration's time, which is split in those shares.  The solves run into the problem's scratch
 * vector from a zero initial guess; the result is left alone, and a preconditioner pinned by
 * get_precond other than block Jacobi is released.  No counterpart in the reference: a
 * measurement hook for the roofline of the iteration's dominant kernel. */
FEMGPU_API int femgpu_time_iteration(femgpu_problem *p, int precision, int iters, double *block_ms,
                                     double *update_ms);
/* Times every block-SpMV variant (format x load kind x SELL slots per loop trip 1, 2, 3) on this assembled
 * matrix, for FP64 and for the FP32 copy, and keeps the fastest of each; *unroll receives the
 * FP64 pick (may be NULL).  Then times the element-loop variants (femgpu_set_assembly_rows) and
 * keeps the fastest, re-assembling K and F with it. */
FEMGPU_API int femgpu_autotune(femgpu_problem *p, int *unroll);
/* How the full format's matrix is read: plain loads; non-temporal loads (streamed once, L2 left to
 * the x gathers); non-temporal loads with 16-bit column offsets c - r (available when every
 * offset fits, e.g. lattice numberings). */
typedef enum { FEMGPU_LOADS_PLAIN = 0, FEMGPU_LOADS_NT = 1, FEMGPU_LOADS_NT_COL16 = 2 } femgpu_spmv_loads;
/* Current (or forced) variant per precision: index 0 FP64, 1 FP32; format femgpu_spmv_format,
 * unroll 1..3, loads femgpu_spmv_loads (not with the symmetric format; NULL = plain; with the
 * matrix-free format, unroll and loads are those of the residual products). */
FEMGPU_API int femgpu_get_spmv_config(femgpu_problem *p, int format[2], int unroll[2], int loads[2]);
FEMGPU_API int femgpu_set_spmv_config(femgpu_problem *p, const int format[2], const int unroll[2],
                                      const int loads[2]);
/* The same for a batch.  Non-temporal loads suit matrices streamed once per product; when most
 * sub-problems have converged, the few still iterating may fit the 256 MB Infinity Cache, where
 * plain loads keep them resident across iterations. */
FEMGPU_API int femgpu_batch_set_spmv_config(femgpu_batch *b, const int format[2], const int unroll[2],
                                            const int loads[2]);
/* Batch product variants femgpu_batch_autotune chooses among. */
typedef enum {
    FEMGPU_BATCH_STORED_PLAIN = 0,
    FEMGPU_BATCH_STORED_NT = 1,
    FEMGPU_BATCH_STORED_NT_COL16 = 2,
    FEMGPU_BATCH_MATRIX_FREE = -1
} femgpu_batch_product;
/* Times whole batched solves with the options o (NULL: defaults) from a zero initial guess, two
 * per product variant (the stored matrix with non-temporal loads and 16-bit column offsets when
 * every offset fits, with non-temporal loads, with plain loads -- the few sub-problems still
 * iterating late may stay in the 256 MB Infinity Cache --, and the matrix-free product for Hexa
 * and Tetra_10 batches), keeps the variant of the fastest second solve in both precisions
 * (*product, may be NULL), then times the element-loop variants as femgpu_autotune does and
 * re-assembles K and F with the fastest.  The results of the batch are left as they were.
 * Needs an assembled batch. */
FEMGPU_API int femgpu_batch_autotune(femgpu_batch *b, const femgpu_solve_opts *o, int *product);
/* Iteration chunks of the last solve: how many, the device time inside them, and the time the
 * device waited between consecutive chunks with no refresh between them (the host's read of the
 * per-problem scalars and the next launch) over n_gaps such transitions.  Any pointer may be
 * NULL. */
FEMGPU_API int femgpu_get_chunk_stats(femgpu_problem *p, int *n_chunks, double *chunk_ms, int *n_gaps,
                                      double *gap_ms);
FEMGPU_API int femgpu_batch_get_chunk_stats(femgpu_batch *b, int *n_chunks, double *chunk_ms, int *n_gaps,
                                            double *gap_ms);
/* Element-loop variant.  rows = 0: one pass per element colour, each element adding its blocks
 * into K (read-modify-write).  rows = FEMGPU_ASSEMBLY_STAGED: the same passes into a
 * block-contiguous copy of K (each block's 9 values adjacent; an extra K-sized buffer), then one
 * streaming pass into the SpMV's layout.  rows = FEMGPU_ASSEMBLY_GATHER: one pass stores every
 * element's blocks and loads (plain writes), a second builds every stored position of K and every
 * node's load from its list of contributions, written once (buffers of the element blocks and the
 * lists, about 2.6 GB and 0.4 GB for a million Hexa8; element blocks are indexed in 32 bits).
 * rows = FEMGPU_ASSEMBLY_GATHER_BOX (Hexa meshes whose elements are all axis-aligned boxes, of at
 * most 4096 distinct edge-length triples per sub-problem, as every make_rect lattice): the gather
 * with no stored element blocks -- the element matrix of a box is linear in its modulus, so each
 * class of equal edge lengths gets its unit-modulus matrix from the generated per-point blocks
 * once per assembly, and every position sums E_e times its class's blocks (K agrees with the
 * other variants to rounding, 1e-15 of the matrix scale, not bitwise).
 * rows = R (a power of two, at most *max_rows): each workgroup owns R consecutive block rows,
 * accumulates their blocks from the incident elements colour by colour in LDS and writes the rows
 * once.  Every variant sums each entry in the same (colour) order: the colour passes, the staged
 * passes and the gather give bitwise the same K and F.  The default is FEMGPU_ASSEMBLY_STAGED;
 * femgpu_autotune times them all. */
#define FEMGPU_ASSEMBLY_STAGED 0xffffffffu
#define FEMGPU_ASSEMBLY_GATHER 0xfffffffeu
#define FEMGPU_ASSEMBLY_GATHER_BOX 0xfffffffdu
FEMGPU_API int femgpu_set_assembly_rows(femgpu_problem *p, uint32_t rows);
FEMGPU_API int femgpu_get_assembly_rows(femgpu_problem *p, uint32_t *rows, uint32_t *max_rows);
/* Replays each chunk of CG iterations from a captured hipGraph (on by default for batches). */
FEMGPU_API int femgpu_set_graph_replay(femgpu_problem *p, int enable);

/* Storage order of the block rows in the SELL-64 slices (64 rows per slice, each slice as wide as
 * its longest row).  FEMGPU_ROW_ORDER_BY_LENGTH stores the rows of every run of 1024 block rows
 * (four workgroups of one sub-problem) by decreasing length, so a slice holds rows of similar
 * length and pads less (meshes mixing node kinds, such as Tetra_10 vertex and edge nodes);
 * FEMGPU_ROW_ORDER_AUTO does so when it saves more than 3% of the positions.  The order is
 * internal: every vector and the CSR view keep the reference's row order. */
typedef enum {
    FEMGPU_ROW_ORDER_AUTO = 0,
    FEMGPU_ROW_ORDER_NATURAL = 1,
    FEMGPU_ROW_ORDER_BY_LENGTH = 2
} femgpu_row_order;
/* Process-wide order for problems and batches created afterwards (default AUTO). */
FEMGPU_API int femgpu_set_row_order(int order);
/* Order a problem uses (NATURAL or BY_LENGTH) and its SELL positions (blocks + padding). */
FEMGPU_API int femgpu_get_row_order(femgpu_problem *p, int *order, uint64_t *n_positions);
/* want_rcm / want_amd (formulation.h:314-323, util/symrcm.h:30-35, util/symamd.h:27-36):
 * problems and batches created afterwards store their nodes in the Cuthill-McKee order of
 * femgpu_node_order (RCM), so that the block SpMV gathers neighbouring values from neighbouring
 * memory whatever the caller's node numbering, or in the approximate minimum degree order of
 * femgpu_node_order_amd (AMD: a fill-reducing order for direct factorisations, which scatters
 * neighbours in memory and slows the iterative path), per batch member.  The order is internal:
 * every API array (results, loads, constraints, faces, the CSR view, result files) keeps the
 * caller's numbering, and FEMGPU_BLOCK_IC0 still eliminates in the caller's node order (the
 * reference's factor).  Process-wide; default NATURAL. */
typedef enum {
    FEMGPU_NODE_ORDER_NATURAL = 0,
    FEMGPU_NODE_ORDER_RCM = 1,
    FEMGPU_NODE_ORDER_AMD = 2
} femgpu_node_order_kind;
FEMGPU_API int femgpu_set_node_order(int order);
/* Order a problem uses; perm (n_nodes, may be NULL) receives the caller's node held at each
 * internal position. */
FEMGPU_API int femgpu_get_node_order(femgpu_problem *p, int *order, uint32_t *perm);
/* Host only: the Cuthill-McKee visiting order symrcm builds (util/symrcm.cpp:38-101): breadth
 * first from the lowest-degree node, unvisited neighbours by increasing degree (ties in node
 * order), unreached components restarted from their lowest-degree node.  row_ptr / col: node
 * adjacency, every row holding its own node (the 3x3-block pattern).  order: n_brows nodes. */
FEMGPU_API int femgpu_node_order(uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col, uint32_t *order);
/* Host only: the approximate minimum degree order symamd computes (util/symamd.cpp:16-36 with the
 * AMD 1.1 package the reference vendors, default controls: rows of degree above
 * max(16, 10 sqrt(n)) ordered last, aggressive absorption, assembly-tree postorder).  row_ptr /
 * col: a symmetric node adjacency, rows strictly increasing, every row holding its own node.
 * order: n_brows nodes, the node at each position (symamd(m) returns the inverse, the position
 * of each node, which want_amd multiplies into indice_noda). */
FEMGPU_API int femgpu_node_order_amd(uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col,
                                     uint32_t *order);
/* Host only: the storage order the library would choose for a block-row length profile
 * (row_ptr, n_brows + 1 entries; one sub-problem).  perm[t] = block row stored at t (may be
 * NULL); *n_natural and *n_positions receive the SELL positions of the natural and the chosen
 * order (either may be NULL). */
FEMGPU_API int femgpu_sell_order(uint32_t n_brows, const uint32_t *row_ptr, int order, uint32_t *perm,
                                 uint64_t *n_natural, uint64_t *n_positions);

#ifdef __cplusplus
}
#endif
#endif
