// Internal structures shared by the C ABI (capi.hip) and the kernels (assemble.hip, solve.hip,
// ic0.hip).
//
// A Problem holds n_sub >= 1 independent elasticity systems in one device system: their nodes,
// elements, faces and constraints are concatenated (node indices offset), the block pattern is
// block diagonal across sub-problems, and every workgroup of the row-parallel kernels covers
// rows of exactly one sub-problem.  Each sub-problem keeps its own CG scalars and stop test, so
// one launch advances every unfinished system.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "host_mesh.hpp"

namespace femgpu {

// Device-resident scalars of one sub-problem's CG.  Written by the finalize kernel that combines
// the per-workgroup partials of a reducing kernel, read by the next kernels and, between chunks
// of iterations, by the host.
struct ProbScalars {
    double rho, rho_prev, pq, alpha, beta, rr, rmax;
    double bb, stop2, crit;       // stop test: rr <= stop2 (REL_L2) or rmax <= crit (ABS_MAXNORM)
    double rr_ref, delta2;        // mixed: pause for an FP64 refresh when rr <= delta2 * rr_ref
    double max_diag;              // assembly: max |diag K| of this sub-problem
    double res_best;              // smallest true residual measure of the checks so far
    int iter, max_iter, done, pause, status, crit_kind, mixed, refreshes;
    int checks, stalls;           // true-residual checks; consecutive ones without progress
    int iter_best;                // iteration of the check that set res_best
    double cheb_c0, cheb_c1;      // Chebyshev preconditioner: z = c0 D r + c1 D K D r
    double pw_norm, pw_rq;        // power iteration: norm of its vector, Rayleigh quotient
    // folded iteration (two kernels): the last vector update's partials of r.z, r.r, max|r| are
    // not yet applied (iter, rho, beta, the stop test); the next element-block kernel or the
    // chunk's closing pass applies them
    int fin_pending;
};

// pause states: 0 running, 1 FP64 refresh requested, 2 refreshed (direction update pending)
enum { PAUSE_NONE = 0, PAUSE_REQ = 1, PAUSE_DONE = 2 };
// sub-problem selectors of the refresh kernels
enum { SEL_ACTIVE = 0, SEL_PAUSE_REQ = 1, SEL_PAUSE_DONE = 2, SEL_ALL = 3 };

struct SubMaterial { double nu, bf0, bf1, bf2; };

constexpr uint32_t ROWS_PER_WG = 256;
// Rows per workgroup of the vector update of the two-kernel CG iteration (1024 lanes, one row
// each): the fewer its partials, the fewer bytes every element block of the next product reads
// to form beta.
constexpr uint32_t FOLD_ROWS_PER_WG = 1024;
constexpr uint32_t ASM_STAGED = 0xffffffffu;
constexpr uint32_t ASM_GATHER = 0xfffffffeu;
constexpr uint32_t ASM_GATHER_BOX = 0xfffffffdu;
constexpr uint32_t MAX_BOX_CLASSES = 4096;
constexpr uint32_t NO_SUB = 0xffffffffu;
// Storage order of the block rows, chosen when a problem is created: AUTO sorts rows by length
// within windows of SORT_WINDOW_WGS workgroups when that saves more than 3% of the SELL
// positions; NATURAL keeps t = r; BY_LENGTH always sorts.
enum { ROW_ORDER_AUTO = 0, ROW_ORDER_NATURAL = 1, ROW_ORDER_BY_LENGTH = 2 };
constexpr uint32_t SORT_WINDOW_WGS = 4;
int row_order_default();
// Node order chosen when a mesh problem is created: NATURAL keeps the caller's numbering, CM
// stores nodes in the Cuthill-McKee order of symrcm (cm_order) for locality, AMD in the
// approximate minimum degree order of symamd (amd_order).
enum { NODE_ORDER_NATURAL = 0, NODE_ORDER_CM = 1, NODE_ORDER_AMD = 2 };
int node_order_default();

// Block-SpMV variant per precision (index 0 = FP64, 1 = FP32): storage format, SELL slots per loop
// trip (1, 2 or 3) and how the full format's matrix is loaded: plain loads, non-temporal loads,
// or non-temporal loads with 16-bit column offsets (c - r; only when every offset fits).
// FMT_MF: the product is evaluated element by element from the mesh (Hexa and Tetra_10; see
// MatFree); the residual products of that precision keep the full storage with its unroll and
// loads.
enum { FMT_FULL = 0, FMT_SYM = 1, FMT_MF = 2 };
enum { LOAD_PLAIN = 0, LOAD_NT = 1, LOAD_NT16 = 2 };
struct SpmvConfig {
    int unroll[2] = {2, 2};
    int format[2] = {FMT_FULL, FMT_FULL};
    int loads[2] = {LOAD_NT, LOAD_NT};
    bool tuned = false;
};

// Device copy of the symmetric SELL storage (host_mesh.hpp SymSell) and its values.
struct SymDev {
    uint32_t *d_sptrU = nullptr, *d_lenU = nullptr, *d_colU = nullptr, *d_srcU = nullptr;
    uint32_t *d_sptrL = nullptr, *d_lenL = nullptr, *d_colL = nullptr, *d_tposL = nullptr;
    uint64_t psU = 0, nL = 0;
    double *d_K = nullptr;
    float *d_K32 = nullptr;
    bool built = false, valid[2] = {false, false};
};

// Matrix-free product y = K x of a Hexa problem with one sub-problem: K_e x_e evaluated per element
// with the element loop's quadrature.  Elements are sorted along a Morton curve of their centres
// and cut into blocks of MF_BLOCK; a workgroup stages the input at its block's local nodes in
// LDS, reads its elements' corners there by block-local index, sums their nodal forces in LDS (sub-colour by
// sub-colour, so every sum has a fixed order), adds the penalty terms of the constrained DOFs it
// is assigned (each to the first block holding its node), and stores the nodes all of whose
// elements are in the block; a node shared by several blocks gets one partial per block, summed in
// block order by a second pass.  Nodes of no element get the penalty alone.  Mesh arrays are
// built once; the material and penalties are those of the last matrix assembly, so the operator
// is the assembled K up to rounding even when the material changed later
// (assume_constant_matrix).
constexpr uint32_t MF_BLOCK = 256;
constexpr uint32_t MF_NODE_REGS = 2;
constexpr uint32_t MF_NO_NODE = 0xffffffffu;
constexpr uint32_t MF_FIRST_HOLDER = 0x80000000u, MF_SLOT_BITS = 0x7fffffffu;  // shared-node slot: flag | slot  // local nodes per lane whose node and slot k_mf_block keeps in registers
constexpr uint32_t MF_EMPTY = 0xffu, MF_BOX = 0x40u, MF_COLOR_BITS = 0x3fu;
struct MatFree {
    bool built = false, snap = false;
    bool all_boxes = false;          // every element an axis-aligned box (the box-only kernel serves)
    uint32_t n_boxes = 0;            // elements that are axis-aligned boxes
    void *d_rec64 = nullptr, *d_rec32 = nullptr;  // per slot of a box: edge lengths, Hooke coefficients
    bool rec_valid = false;          // the records hold the last snapshot's material
    uint32_t n_blocks = 0, n_colors = 0, max_local = 0, n_bnodes = 0, n_halo = 0, n_iso = 0;
    uint32_t n_slots = 0, n_fix = 0, cap_fix = 0, cap_slots = 0;
    uint32_t *d_conn = nullptr;      // 8 nodes per block slot (MF_BLOCK slots per block)
    uint16_t *d_loc = nullptr;       // 8 block-local node indices per slot
    uint8_t *d_color = nullptr;      // sub-colour per slot | MF_BOX for axis-aligned boxes (MF_EMPTY: empty slot)
    uint32_t *d_elem = nullptr;      // element per slot (0 for empty slots)
    uint32_t *d_lptr = nullptr;      // block b: local nodes [lptr[b], lptr[b + 1])
    uint32_t *d_bls = nullptr;       // block b: local node count | log2 accumulator copies << 24 (FP32), << 28 (FP64)
    size_t lds32 = 0, lds64 = 0;     // LDS of the element-block launches (the largest block's)
    bool uni32 = false, uni64 = false;  // every block takes the most accumulator copies
    uint32_t ls = 0;                 // stride of the per-block node lists (max_local rounded up to 64)
    // block b, local node l < its count (lptr[b + 1] - lptr[b]): entry b * ls + l holds the global
    // node (MF_NO_NODE past the count) and its partial slot when shared with another block (~0 if not)
    // (bit 31 of a shared node's slot: this block is the first holding the node and stores its
    // updated CG direction)
    uint32_t *d_lnode_pad = nullptr, *d_lslot_pad = nullptr;
    double *d_pb = nullptr;          // second CG direction buffers (with the problem's d_p / d_p32,
    float *d_pb32 = nullptr;         // used in turn by iteration parity when the update is fused)
    uint32_t *d_bnode = nullptr, *d_bptr = nullptr;  // shared node k: partials [bptr[k], bptr[k + 1])
    uint32_t *d_hbeg = nullptr;      // node n: partials [hbeg[n], hbeg[n + 1]) (empty unless shared)
    uint32_t *d_bsub = nullptr;      // sub-problem of each block (NO_SUB: padding); a block never straddles two
    std::vector<uint32_t> sub_b0, sub_nb;  // each sub-problem's consecutive blocks
    double *d_halo = nullptr;        // 3 values per partial slot (FP64, or FP32 in its first half)
    std::vector<uint32_t> node_blk, node_loc;  // first block holding each node, its local index (~0: none)
    uint32_t *d_idof = nullptr;      // the 3 DOFs of each node of no element ...
    uint32_t *d_isrc = nullptr;      // ... and the index of their penalty in d_fpen (~0: none)
    // penalty terms per block: block b adds fpen[psrc[k]] x_l to local entry l = pidx[k] (x_l its
    // staged input), k in [pptr[b], pptr[b + 1]); built for the DOF list pen_dofs
    uint32_t *d_pptr = nullptr, *d_pidx = nullptr, *d_psrc = nullptr;
    std::vector<uint32_t> snap_dofs, pen_dofs;
    bool pen_built = false;
    double *d_E = nullptr;           // E per element at the last matrix assembly
    SubMaterial *d_mat = nullptr;    // material at the last matrix assembly
    double *d_fpen = nullptr;        // summed penalty M * P of each DOF of snap_dofs at that assembly
    double *d_partials = nullptr;    // one p.q partial per workgroup (element blocks, then isolated nodes)
    uint32_t *d_wg0 = nullptr, *d_nwg = nullptr;  // per sub-problem: its p.q partial slots (first, count)
    uint32_t *d_b0 = nullptr;        // per sub-problem: its first block
    bool fold = false;               // every sub-problem has a block: the CG iteration runs in two kernels
};

// Device side of the block IC(0) preconditioner (schedule in host_mesh.hpp).
struct Ic0Args {
    const uint32_t *nat, *sub, *len, *nlow, *sptr, *colN, *colT, *kpos, *tpos;
    uint64_t ps;  // plane stride of the factor values
};
struct Ic0 {
    Ic0Schedule sched;
    uint32_t *d_nat = nullptr, *d_sub = nullptr, *d_len = nullptr, *d_nlow = nullptr, *d_sptr = nullptr,
             *d_colN = nullptr, *d_colT = nullptr, *d_kpos = nullptr, *d_tpos = nullptr;
    double *d_L = nullptr, *d_z = nullptr;   // factor (9 planes), preconditioned residual
    float *d_L32 = nullptr, *d_z32 = nullptr;
    uint64_t ps = 0;
    bool valid = false, valid32 = false;     // factor of the current K (FP64, FP32 copy)
    Ic0Args args() const { return Ic0Args{d_nat, d_sub, d_len, d_nlow, d_sptr, d_colN, d_colT, d_kpos, d_tpos, ps}; }
};

// Gather element loop (ASM_GATHER): pass 1 stores every element's upper blocks (d_em: 9 values
// per block, NPE (NPE + 1) / 2 blocks per element) and nodal loads (d_ef: 3 per (element, node));
// pass 2 sums, for every stored position of K, its entries d_kent[d_kptr[p] .. d_kptr[p + 1])
// ((element block index << 1) | transposed), and for every node its loads d_fent[d_fptr[n] ..]
// (element * NPE + local node); both lists in the colour order of the colour passes.
struct GatherAsm {
    double *d_em = nullptr, *d_ef = nullptr;
    uint32_t *d_kptr = nullptr, *d_kent = nullptr, *d_fptr = nullptr, *d_fent = nullptr;
    // ASM_GATHER_BOX, when every element is an axis-aligned Hexa box: elements fall into classes
    // of equal edge lengths and sub-problem; a class's unit-modulus element matrix K_c(1) and body
    // loads are formed once per matrix assembly, and pass 2 sums E_e K_c(1) (K_e is linear in E_e)
    // instead of reading stored element blocks.
    bool box = false;
    uint32_t n_classes = 0;
    uint16_t *d_cls = nullptr;        // class of each element
    double *d_ch = nullptr;           // edge lengths of each class (3)
    uint32_t *d_csub = nullptr;       // sub-problem of each class
    double *d_kref = nullptr;         // per class: the NUP upper 3x3 blocks of K_c(1)
    double *d_fref = nullptr;         // per class: the loads of its 8 nodes (3 each)
};

struct Problem {
    int kind = 0, npe = 0, device = 0;
    uint32_t n_sub = 1;
    uint32_t n_nodes = 0, n_elem = 0, n_faces = 0;          // totals over sub-problems
    std::vector<uint32_t> node_off, elem_off;                // n_sub + 1 offsets
    uint64_t nnzb = 0, ps = 0;  // stored blocks, plane stride (SELL positions, padded to 64)
    hipStream_t stream = nullptr;
    hipEvent_t ev[7] = {};  // 0-3 solve and kernel timing, 4-6 the ring timing iteration chunks

    Pattern pat;
    Coloring ecol, fcol;

    // workgroup tables of the row-parallel kernels (indexed by logical workgroup)
    uint32_t n_wg = 0;                                       // grid size, multiple of 8
    uint32_t *d_wg_row0 = nullptr, *d_wg_row1 = nullptr, *d_wg_sub = nullptr;
    uint32_t *d_sub_wg0 = nullptr, *d_sub_nwg = nullptr;     // per sub-problem workgroup range
    // the same with workgroups of 4 x 256 rows (the vector update of the two-kernel CG iteration)
    uint32_t n_fwg = 0;
    uint32_t *d_fwg_row0 = nullptr, *d_fwg_row1 = nullptr, *d_fwg_sub = nullptr;
    uint32_t *d_fsub_wg0 = nullptr, *d_fsub_nwg = nullptr;

    // device: mesh, pattern, scatter map
    double *d_coords = nullptr, *d_E = nullptr;
    int16_t *d_col16 = nullptr;  // SELL column offsets c - r, when all fit in 16 bits (else null)
    // storage order of the block rows (null: identity): storage row t holds block row d_perm[t]
    uint32_t *d_perm = nullptr, *d_iperm = nullptr;
    // node order (empty: the caller's): internal node i is the caller's node nperm[i], ninv the
    // inverse; every vector, the pattern and the device mesh use the internal numbering, the API
    // the caller's (block-diagonal over sub-problems)
    std::vector<uint32_t> nperm, ninv;
    uint32_t *d_nperm = nullptr, *d_ninv = nullptr;
    int node_order = NODE_ORDER_NATURAL;  // the order nperm follows (NATURAL when empty)
    uint32_t *d_conn = nullptr, *d_rp = nullptr, *d_col = nullptr, *d_diag = nullptr, *d_slot = nullptr,
             *d_order = nullptr, *d_sptr = nullptr, *d_elem_sub = nullptr;
    SubMaterial *d_mat = nullptr;
    std::vector<SubMaterial> mat_h;
    std::vector<uint8_t> material_set;
    // device: system
    double *d_K = nullptr, *d_F = nullptr, *d_u = nullptr;
    float *d_K32 = nullptr;
    bool k32_valid = false, assembled_mat = false, assembled_vec = false;
    // skin loads
    uint32_t *d_faces = nullptr, *d_forder = nullptr;
    double *d_trac = nullptr;
    std::vector<uint32_t> faces_h;
    std::vector<double> trac_h;
    // constraints, grouped by dof in insertion order
    struct Cons { uint32_t dof; double val, pen; };
    std::vector<Cons> cons;
    bool cons_dirty = true;
    uint32_t n_cdof = 0;
    std::vector<uint32_t> cdof_h;  // the constrained DOFs of d_cdof, in its order
    uint32_t *d_cdof = nullptr, *d_cptr = nullptr, *d_csub = nullptr;
    double *d_cval = nullptr, *d_cpen = nullptr;
    // solver work vectors
    double *d_r = nullptr, *d_p = nullptr, *d_q = nullptr, *d_Minv = nullptr;
    double *d_aux = nullptr;  // adjoint state / scratch vector
    double *d_red = nullptr;  // diff_norm partials
    float *d_r32 = nullptr, *d_p32 = nullptr, *d_q32 = nullptr, *d_x32 = nullptr;
    uint16_t *d_Minv32 = nullptr;  // the FP32 iterations' preconditioner values, bfloat16 bits
    double *d_z = nullptr;    // z = M^-1 r of the Jacobi preconditioners (written by the update)
    float *d_z32 = nullptr;
    double *d_partials = nullptr;
    // d_ps: the scalars of every sub-problem (n_sub), then a second copy (n_sub) that the
    // two-kernel CG iteration passes between its kernels (the element-block kernel writes it from
    // the first, the vector update writes the first back from it)
    ProbScalars *d_ps = nullptr, *h_ps = nullptr;
    int *d_flags = nullptr, *h_flags = nullptr;  // [0]: element with det J <= 0; [1]: IC(0) pivot <= 0
    int minv_kind = -1;  // preconditioner currently stored in d_Minv
    // Chebyshev preconditioner: per sub-problem coefficients (c0, c1) of its polynomial, from the
    // power iteration run with the block-Jacobi factors it was built with (minv_kind == 3)
    std::vector<double> cheb_c;
    SpmvConfig spmv;
    std::vector<double> max_diag;  // per sub-problem, of the last matrix assembly
    double last_asm_ms = 0;
    double last_precond_ms = 0;
    double last_after_solve_ms = 0;    // device time of the last preconditioner set-up (0: reused)
    bool from_bsr = false;
    // iteration chunks of the last solve: count, device time inside them, and the time the device
    // waited between consecutive chunks with no refresh between them (the host's poll of the
    // scalars and the next launch), over n_gaps such transitions
    int last_chunks = 0, last_gaps = 0;
    double last_chunk_ms = 0, last_gap_ms = 0;
    bool use_graph = false;        // replay chunks of iterations from a captured hipGraph
    // kernel timing of the two-kernel iteration (femgpu_time_iteration, graphs off): events before
    // the element blocks, between them and the vector update, after it, for the first kt_cap
    // iterations
    hipEvent_t *kt = nullptr;
    int kt_cap = 0, kt_used = 0;
    // captured chunks: slot 0 holds a chunk of the solve's maximum length, slot k (1..5) one of
    // 2^k iterations (the shorter chunks that end at a predicted pause are sums of these)
    static constexpr int N_GRAPHS = 6;
    hipGraphExec_t graph_exec[N_GRAPHS] = {};
    uint64_t graph_key[N_GRAPHS] = {};
    const double *graph_x[N_GRAPHS] = {};  // solution vector each captured chunk updates
    Ic0 ic0;
    SymDev sym;
    MatFree mf;
    // element loop variant: 0 = colour passes over all elements (read-modify-write of K),
    // R > 0 = row groups of R rows accumulated in LDS (work lists in rg / d_gptr, d_gel),
    // ASM_STAGED = colour passes into the block-contiguous staging copy d_Kb, then the planes
    uint32_t asm_rows = 0;
    double *d_Kb = nullptr;
    GatherAsm ga;                    // ASM_GATHER: element blocks and the per-position gather lists
    std::vector<uint32_t> conn_h;  // host copy of the (offset) connectivity, for the work lists
    RowGroups rg;
    uint32_t *d_gptr = nullptr, *d_gel = nullptr;
};

// ---- launchers (assemble.hip) ----
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s);
hipError_t launch_copy(double *dst, const double *src, size_t n, hipStream_t s);  // n even
// out: 4 doubles per 1024 input pairs (one per wave)
hipError_t launch_read(const double *src, size_t n, double *out, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
// dst[3 i + d] = src[3 map[i] + d] for i < n (node-major vectors)
hipError_t launch_permute_nodes(double *dst, const double *src, const uint32_t *map, uint32_t n, hipStream_t s);
hipError_t launch_assemble(Problem &P, int assemble_mat, int assemble_vec);
// Largest row-group size (power of two <= 64) whose accumulator fits the LDS budget; 0 if none.
uint32_t max_assembly_rows(const Problem &P);
// An axis-aligned Hexa box: every edge along xi_d is (h_d > 0) e_d, one h_d per axis, compared
// exactly; X the node coordinates, nd the element's 8 nodes.
bool hexa_is_box(const double *X, const uint32_t *nd);
hipError_t launch_constraints(Problem &P, int assemble_mat, int assemble_vec);
hipError_t launch_to_fp32(Problem &P);
hipError_t launch_elem_grad(Problem &P, const double *d_u, const double *d_lam, double *d_g);
// eps (6 per element), sigma (6 per element), tr_epsilon (1 per element); any may be null
hipError_t launch_after_solve(Problem &P, const double *d_u, double *d_eps, double *d_sig, double *d_tr);

// ---- solver (solve.hip) ----
struct SolveOpts { int precision, precond, crit_kind; double tol; int max_iter; double delta; int chunk; };
struct SolveOut { int status, iters, refreshes; double rel_res, max_abs_res, ms; };
// Solves every sub-problem; out[s] per sub-problem; returns the first non-zero status.
int solve_pcg(Problem &P, const double *d_rhs, double *d_x, const SolveOpts &o, SolveOut *out);
// Average device times (ms) of the two kernels of the CG iteration (block Jacobi, the matrix-free
// product of precision 0 (FP64) or 1 (FP32) folded with the previous update's end, then the vector
// update) over iters iterations of a solve of K d = F into the scratch vector d_aux, after
// warm-up ones; returns 0, 1 (the product is not the matrix-free one for that precision, or the
// iteration has three kernels) or 5.
int time_iteration(Problem &P, int precision, int iters, double *block_ms, double *update_ms);
double time_spmv(Problem &P, int precision, int reps);
// d_y = K d_x (device, internal order) with the product variant of precision 0 (FP64) or 1 (FP32);
// returns 0, 1 (the variant cannot serve this problem) or 5.
int apply_spmv(Problem &P, int precision, const double *d_x, double *d_y);
// Bytes of scratch apply_spmv_scratch needs for precision (FP32 copies of x and y, the matrix-free
// product's shared-node partials), or -1 while the product variant's data is not built for the
// current K (building it allocates: apply_spmv or a solve builds it).
int64_t apply_scratch_bytes(const Problem &P, int precision);
// apply_spmv with every temporary in the caller's scratch (apply_scratch_bytes of it, 256-byte
// aligned); allocates nothing and does not synchronise.  Returns 0 or 5.
int apply_spmv_scratch(Problem &P, int precision, const double *d_x, double *d_y, char *scratch);
// Preconditioner precond (femgpu_precond) of the current K with its FP32 copy, as get_precond
// does before the solves that reuse it; returns 0, 1 (bad input), 3 (non-positive pivot) or 5.
int precond_build(Problem &P, int precond);
// ||a - b||_2 and max |a - b| over n device values (deterministic); returns 0 or 5.
int diff_norm(Problem &P, const double *a, const double *b, size_t n, double *norm2, double *norm_inf);
// Picks format and unroll per precision by timing every variant on this matrix; returns 0 or 5.
int autotune_spmv(Problem &P);
// Symmetric storage and its values for precision (0 FP64, 1 FP32) from the current K.
int sym_ensure(Problem &P, int precision);
void sym_free(SymDev &S);
void destroy_graph(Problem &P);

// ---- matrix-free product (matfree.hip) ----
bool mf_eligible(const Problem &P);
// Colour-order mesh arrays (once per problem); returns 0, 1 (not a Hexa problem with one
// sub-problem) or 5.
int mf_build(Problem &P);
// Material and penalties of the matrix assembly just launched (stream order); eligible problems only.
hipError_t mf_snapshot(Problem &P);
// Builds what the product needs outside any graph capture: returns 0, 1 (not eligible, or no
// matrix assembled yet) or 5.
int mf_prepare(Problem &P);
// y = K x (pq: on the active sub-problem, with the p.q partials in mf.d_partials).  With z
// (pq only): y = K p for the CG direction p = z + beta x (beta of the scalars), stored into x.
// halo: where the shared nodes' partials go (3 values of T per slot); null: the problem's buffer.
// The partials of r.z, r.r, max|r| a two-kernel CG iteration's element blocks combine: values
// k * stride + w0[s] + i, i < nw[s], of sub-problem s.
struct FinParts {
    const double *part;
    unsigned stride;
    const uint32_t *w0, *nw;
};
// fin (with z): the two-kernel CG iteration, with the previous vector update's pending partials
// (fin) applied by the element blocks (P.mf.fold problems only).
hipError_t mf_launch(Problem &P, const float *x, float *y, bool pq, const float *z = nullptr, void *halo = nullptr,
                     const FinParts *fin = nullptr);
hipError_t mf_launch(Problem &P, const double *x, double *y, bool pq, const double *z = nullptr, void *halo = nullptr,
                     const FinParts *fin = nullptr);
// y = K x with the shared nodes' partials left in the halo buffer (summed by the caller's vector
// kernel); pq: active sub-problems only, with the p.q partials.
hipError_t mf_launch_partial(Problem &P, const float *x, float *y, bool pq);
hipError_t mf_launch_partial(Problem &P, const double *x, double *y, bool pq);
// The product's arrays are built for the current matrix assembly (mf_prepare has nothing left to
// allocate or upload for a plain product).
bool mf_ready(const Problem &P);
// Values in the shared-node partial slots (3 per slot).
size_t mf_halo_values(const Problem &P);
void mf_free(MatFree &M);
// Device time (ms) of the element-block kernel alone and of the whole product; returns 0, 1 or 5.
int mf_time(Problem &P, int precision, int reps, double *block_ms, double *product_ms);

// ---- block IC(0) (ic0.hip) ----
// Schedule for `ordering` (built and uploaded when it changes) and the numeric factor of the
// current K (FP64, plus the FP32 copy when fp32).  Returns 0, 3 (non-positive pivot: the
// sub-problem's status is set to 3) or 5.
int ic0_ensure(Problem &P, int ordering, bool fp32);
// z = (L L^T)^-1 r on the sub-problems selected by sel: forward sweep level by level, backward
// sweep in reverse level order (T = double or float).
void ic0_apply(Problem &P, const double *r, double *z, int sel);
void ic0_apply(Problem &P, const float *r, float *z, int sel);
// Factor in elimination numbering: order[t] = node, full-pattern CSR over t (rows sorted), values
// of the lower blocks and of the lower-triangular diagonal blocks, zeros in the upper blocks.
int ic0_download(Problem &P, uint32_t *order, uint32_t *row_ptr, uint32_t *col, double *L);
void ic0_free(Ic0 &I);

std::string &last_error();

}  // namespace femgpu
