// C ABI of the MI355X assemble + solve path (declarations and reference citations in
// include/femgpu.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>

#include "../../include/femgpu.h"
#include "femgpu_internal.hpp"

namespace femgpu {
std::string &last_error() {
    static thread_local std::string s;
    return s;
}
}  // namespace femgpu

using namespace femgpu;

struct femgpu_problem {
    Problem P;
};

namespace {

int fail(int status, const std::string &msg) {
    last_error() = msg;
    return status;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(FEMGPU_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t dalloc(T **p, size_t n) {
    return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(1, n) * sizeof(T));
}
template <class T>
hipError_t upload(T **p, const T *h, size_t n) {
    hipError_t e = dalloc(p, n);
    if (e == hipSuccess && n) e = hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice);
    return e;
}
template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_problem(Problem &P) {
    dfree(P.d_coords); dfree(P.d_E); dfree(P.d_conn); dfree(P.d_rp); dfree(P.d_col); dfree(P.d_diag);
    dfree(P.d_slot); dfree(P.d_order); dfree(P.d_K); dfree(P.d_F); dfree(P.d_u); dfree(P.d_K32);
    dfree(P.d_faces); dfree(P.d_forder); dfree(P.d_trac); dfree(P.d_cdof); dfree(P.d_cptr); dfree(P.d_cval);
    dfree(P.d_cpen); dfree(P.d_r); dfree(P.d_p); dfree(P.d_q); dfree(P.d_Minv); dfree(P.d_r32); dfree(P.d_p32);
    dfree(P.d_q32); dfree(P.d_x32); dfree(P.d_Minv32); dfree(P.d_partials); dfree(P.d_sc); dfree(P.d_wgrows);
    dfree(P.d_aux);
    if (P.h_sc) (void)hipHostFree(P.h_sc);
    P.h_sc = nullptr;
    for (auto &e : P.ev)
        if (e) (void)hipEventDestroy(e);
    if (P.stream) (void)hipStreamDestroy(P.stream);
    P.stream = nullptr;
}

// Device buffers common to mesh problems and plain block systems.
int alloc_system(Problem &P) {
    const size_t n = 3 * (size_t)P.n_nodes;
    P.ps = (P.nnzb + 63) / 64 * 64;
    HIPCHK(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    for (auto &e : P.ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(upload(&P.d_rp, P.pat.row_ptr.data(), P.pat.row_ptr.size()));
    HIPCHK(upload(&P.d_col, P.pat.col.data(), P.pat.col.size()));
    HIPCHK(upload(&P.d_diag, P.pat.diag.data(), P.pat.diag.size()));
    HIPCHK(dalloc(&P.d_K, 9 * P.ps));
    HIPCHK(dalloc(&P.d_F, n));
    HIPCHK(dalloc(&P.d_u, n));
    HIPCHK(dalloc(&P.d_r, n));
    HIPCHK(dalloc(&P.d_p, n));
    HIPCHK(dalloc(&P.d_q, n));
    HIPCHK(dalloc(&P.d_aux, n));
    HIPCHK(dalloc(&P.d_Minv, 9 * (size_t)P.n_nodes));
    HIPCHK(dalloc(&P.d_r32, n));
    HIPCHK(dalloc(&P.d_p32, n));
    HIPCHK(dalloc(&P.d_q32, n));
    HIPCHK(dalloc(&P.d_x32, n));
    HIPCHK(dalloc(&P.d_partials, 8 * 4096));
    HIPCHK(dalloc(&P.d_sc, 1));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&P.h_sc), sizeof(Scalars), hipHostMallocDefault));
    std::memset(P.h_sc, 0, sizeof(Scalars));
    HIPCHK(hipMemcpy(P.d_sc, P.h_sc, sizeof(Scalars), hipMemcpyHostToDevice));
    HIPCHK(launch_fill(P.d_u, n, 0.0, P.stream));
    HIPCHK(launch_fill(P.d_p, n, 0.0, P.stream));
    HIPCHK(launch_fill(P.d_K, 9 * P.ps, 0.0, P.stream));
    HIPCHK(launch_fill(P.d_F, n, 0.0, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    return FEMGPU_OK;
}

int upload_constraints(Problem &P) {
    if (!P.cons_dirty) return FEMGPU_OK;
    dfree(P.d_cdof); dfree(P.d_cptr); dfree(P.d_cval); dfree(P.d_cpen);
    // group by DOF, keeping insertion order inside a group
    std::vector<size_t> idx(P.cons.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return P.cons[a].dof < P.cons[b].dof; });
    std::vector<uint32_t> dof, ptr(1, 0);
    std::vector<double> val, pen;
    for (size_t k = 0; k < idx.size(); ++k) {
        const auto &c = P.cons[idx[k]];
        if (dof.empty() || dof.back() != c.dof) {
            if (!dof.empty()) ptr.push_back((uint32_t)val.size());
            dof.push_back(c.dof);
        }
        val.push_back(c.val);
        pen.push_back(c.pen);
    }
    if (!dof.empty()) ptr.push_back((uint32_t)val.size());
    P.n_cdof = (uint32_t)dof.size();
    HIPCHK(upload(&P.d_cdof, dof.data(), dof.size()));
    HIPCHK(upload(&P.d_cptr, ptr.data(), ptr.size()));
    HIPCHK(upload(&P.d_cval, val.data(), val.size()));
    HIPCHK(upload(&P.d_cpen, pen.data(), pen.size()));
    P.cons_dirty = false;
    return FEMGPU_OK;
}

SolveOpts to_internal(const femgpu_solve_opts *o) {
    femgpu_solve_opts d;
    femgpu_default_solve_opts(&d);
    if (!o) o = &d;
    return SolveOpts{o->precision, o->precond, o->crit_kind, o->tol, o->max_iter, o->reliable_delta, o->chunk};
}

void to_stats(const SolveOut &s, femgpu_solve_stats *st) {
    if (!st) return;
    st->status = s.status;
    st->iters = s.iters;
    st->refreshes = s.refreshes;
    st->rel_res = s.rel_res;
    st->max_abs_res = s.max_abs_res;
    st->solve_ms = s.ms;
}

int check_opts(const femgpu_solve_opts *o) {
    if (!o) return FEMGPU_OK;
    if (o->precision < 0 || o->precision > 1 || o->precond < 0 || o->precond > 1 || o->crit_kind < 0 ||
        o->crit_kind > 1 || !(o->tol >= 0) || o->max_iter < 0 || o->chunk < 1 || !(o->reliable_delta > 0) ||
        !(o->reliable_delta < 1))
        return fail(FEMGPU_BAD_INPUT, "invalid solve options");
    return FEMGPU_OK;
}

}  // namespace

extern "C" {

void femgpu_default_solve_opts(femgpu_solve_opts *o) {
    if (!o) return;
    o->precision = FEMGPU_FP64;
    o->precond = FEMGPU_JACOBI;
    o->crit_kind = FEMGPU_REL_L2;
    o->tol = 1e-10;
    o->max_iter = 100000;
    o->reliable_delta = 0.1;
    o->chunk = 32;
}

const char *femgpu_status_string(int s) {
    switch (s) {
        case FEMGPU_OK: return "ok";
        case FEMGPU_BAD_INPUT: return "bad input";
        case FEMGPU_NEG_JACOBIAN: return "negative jacobian";
        case FEMGPU_NOT_SPD: return "matrix not SPD";
        case FEMGPU_MAX_ITER: return "iteration cap reached";
        case FEMGPU_HIP_ERROR: return "HIP error";
        case FEMGPU_NAN: return "NaN or Inf";
        case FEMGPU_NOT_READY: return "not ready";
        default: return "unknown status";
    }
}

const char *femgpu_last_error(void) { return last_error().c_str(); }

int femgpu_make_rect(int kind, const double X0[3], const double X1[3], const uint32_t nb[3], double *coords,
                     uint32_t *conn, uint32_t *n_nodes, uint32_t *n_elem) {
    uint32_t nn, ne;
    if (!nb || kind < 0 || kind > 2 || !rect_counts(kind, nb, nn, ne)) return fail(FEMGPU_BAD_INPUT, "make_rect: bad sizes");
    if (n_nodes) *n_nodes = nn;
    if (n_elem) *n_elem = ne;
    if (!coords && !conn) return FEMGPU_OK;
    if (!coords || !conn || !X0 || !X1) return fail(FEMGPU_BAD_INPUT, "make_rect: null output");
    const int rc = make_rect(kind, X0, X1, nb, coords, conn);
    return rc ? fail(FEMGPU_BAD_INPUT, "make_rect failed") : FEMGPU_OK;
}

int femgpu_skin_faces(int kind, uint32_t n_nodes, uint32_t n_elem, const uint32_t *conn, uint32_t *faces,
                      uint32_t *n_faces) {
    if (kind < 0 || kind > 2 || !conn || !n_faces) return fail(FEMGPU_BAD_INPUT, "skin_faces: bad input");
    const int npe = nodes_per_elem(kind);
    for (size_t s = 0; s < (size_t)n_elem * npe; ++s)
        if (conn[s] >= n_nodes) return fail(FEMGPU_BAD_INPUT, "skin_faces: node index out of range");
    std::vector<uint32_t> f = skin_faces(kind, n_nodes, n_elem, conn);
    const uint32_t npf = nodes_per_face(face_kind_of(kind));
    const uint32_t cnt = (uint32_t)(f.size() / npf);
    if (faces) {
        if (*n_faces < cnt) return fail(FEMGPU_BAD_INPUT, "skin_faces: output too small");
        std::memcpy(faces, f.data(), f.size() * 4);
    }
    *n_faces = cnt;
    return FEMGPU_OK;
}

int femgpu_create(const femgpu_mesh *m, femgpu_problem **out) {
    if (!m || !out || !m->coords || !m->conn || m->elem_kind < 0 || m->elem_kind > 2 || !m->n_nodes || !m->n_elem)
        return fail(FEMGPU_BAD_INPUT, "create: bad mesh");
    *out = nullptr;
    const int npe = nodes_per_elem(m->elem_kind);
    for (size_t s = 0; s < (size_t)m->n_elem * npe; ++s)
        if (m->conn[s] >= m->n_nodes) return fail(FEMGPU_BAD_INPUT, "create: node index out of range");
    auto *h = new (std::nothrow) femgpu_problem();
    if (!h) return fail(FEMGPU_BAD_INPUT, "create: out of host memory");
    Problem &P = h->P;
    P.kind = m->elem_kind;
    P.npe = npe;
    P.n_nodes = m->n_nodes;
    P.n_elem = m->n_elem;
    (void)hipGetDevice(&P.device);
    P.coords_h.assign(m->coords, m->coords + 3 * (size_t)m->n_nodes);
    P.conn_h.assign(m->conn, m->conn + (size_t)m->n_elem * npe);
    P.pat = build_pattern(P.n_nodes, P.n_elem, npe, m->conn);
    P.nnzb = P.pat.col.size();
    P.ecol = color_elements(P.n_nodes, P.n_elem, npe, m->conn);
    int rc = alloc_system(P);
    if (rc == FEMGPU_OK) {
        std::vector<uint32_t> slot = build_slot_map(P.pat, P.n_elem, npe, m->conn);
        hipError_t e = upload(&P.d_slot, slot.data(), slot.size());
        if (e == hipSuccess) e = upload(&P.d_coords, m->coords, 3 * (size_t)P.n_nodes);
        if (e == hipSuccess) e = upload(&P.d_conn, m->conn, (size_t)P.n_elem * npe);
        if (e == hipSuccess) e = upload(&P.d_order, P.ecol.order.data(), P.ecol.order.size());
        if (e == hipSuccess) e = dalloc(&P.d_E, P.n_elem);
        if (e != hipSuccess) rc = fail(FEMGPU_HIP_ERROR, std::string("create: ") + hipGetErrorString(e));
    }
    if (rc != FEMGPU_OK) {
        free_problem(P);
        delete h;
        return rc;
    }
    *out = h;
    return FEMGPU_OK;
}

int femgpu_create_from_bsr(uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col, const double *vals,
                           const double *rhs, femgpu_problem **out) {
    if (!out || !n_brows || !row_ptr || !col || !vals) return fail(FEMGPU_BAD_INPUT, "create_from_bsr: bad input");
    *out = nullptr;
    if (row_ptr[0] != 0) return fail(FEMGPU_BAD_INPUT, "create_from_bsr: row_ptr[0] != 0");
    Pattern pat;
    pat.n_rows = n_brows;
    pat.row_ptr.assign(row_ptr, row_ptr + n_brows + 1);
    pat.col.assign(col, col + row_ptr[n_brows]);
    pat.diag.assign(n_brows, 0xffffffffu);
    for (uint32_t r = 0; r < n_brows; ++r) {
        if (row_ptr[r + 1] < row_ptr[r]) return fail(FEMGPU_BAD_INPUT, "create_from_bsr: row_ptr not monotone");
        for (uint32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
            if (col[p] >= n_brows || (p > row_ptr[r] && col[p] <= col[p - 1]))
                return fail(FEMGPU_BAD_INPUT, "create_from_bsr: columns out of range or unsorted");
            if (col[p] == r) pat.diag[r] = p;
        }
        if (pat.diag[r] == 0xffffffffu) return fail(FEMGPU_BAD_INPUT, "create_from_bsr: missing diagonal block");
    }
    auto *h = new (std::nothrow) femgpu_problem();
    if (!h) return fail(FEMGPU_BAD_INPUT, "create_from_bsr: out of host memory");
    Problem &P = h->P;
    P.from_bsr = true;
    P.n_nodes = n_brows;
    P.pat = std::move(pat);
    P.nnzb = P.pat.col.size();
    (void)hipGetDevice(&P.device);
    int rc = alloc_system(P);
    if (rc == FEMGPU_OK) {
        // block row-major input -> nine planes
        std::vector<double> planes(9 * P.ps, 0.0);
        for (size_t p = 0; p < P.nnzb; ++p)
            for (int c = 0; c < 9; ++c) planes[c * P.ps + p] = vals[9 * p + c];
        hipError_t e = hipMemcpy(P.d_K, planes.data(), planes.size() * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess && rhs) e = hipMemcpy(P.d_F, rhs, 3 * (size_t)n_brows * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(FEMGPU_HIP_ERROR, hipGetErrorString(e));
        P.assembled_mat = P.assembled_vec = true;
    }
    if (rc != FEMGPU_OK) {
        free_problem(P);
        delete h;
        return rc;
    }
    *out = h;
    return FEMGPU_OK;
}

int femgpu_destroy(femgpu_problem *h) {
    if (!h) return FEMGPU_OK;
    free_problem(h->P);
    delete h;
    return FEMGPU_OK;
}

int femgpu_set_material(femgpu_problem *h, const femgpu_material *m) {
    if (!h || !m || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "set_material: bad input");
    Problem &P = h->P;
    if (!(m->nu > -1.0 && m->nu < 0.5)) return fail(FEMGPU_BAD_INPUT, "set_material: nu out of (-1, 0.5)");
    P.nu = m->nu;
    for (int d = 0; d < 3; ++d) P.bf[d] = m->body_force[d];
    if (m->E_elem) {
        for (uint32_t e = 0; e < P.n_elem; ++e)
            if (!std::isfinite(m->E_elem[e])) return fail(FEMGPU_NAN, "set_material: non-finite E");
        HIPCHK(hipMemcpy(P.d_E, m->E_elem, (size_t)P.n_elem * 8, hipMemcpyHostToDevice));
    } else {
        if (!std::isfinite(m->E)) return fail(FEMGPU_NAN, "set_material: non-finite E");
        HIPCHK(launch_fill(P.d_E, P.n_elem, m->E, P.stream));
        HIPCHK(hipStreamSynchronize(P.stream));
    }
    P.material_set = true;
    return FEMGPU_OK;
}

int femgpu_set_elastic_modulus_device(femgpu_problem *h, const double *d_E) {
    if (!h || !d_E || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "set_elastic_modulus_device: bad input");
    Problem &P = h->P;
    HIPCHK(hipMemcpyAsync(P.d_E, d_E, (size_t)P.n_elem * 8, hipMemcpyDeviceToDevice, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    return FEMGPU_OK;
}

int femgpu_set_traction(femgpu_problem *h, uint32_t n_faces, const uint32_t *faces, const double *trac) {
    if (!h || h->P.from_bsr || (n_faces && (!faces || !trac))) return fail(FEMGPU_BAD_INPUT, "set_traction: bad input");
    Problem &P = h->P;
    const int npf = nodes_per_face(face_kind_of(P.kind));
    for (size_t s = 0; s < (size_t)n_faces * npf; ++s)
        if (faces[s] >= P.n_nodes) return fail(FEMGPU_BAD_INPUT, "set_traction: node index out of range");
    dfree(P.d_faces); dfree(P.d_forder); dfree(P.d_trac);
    P.n_faces = n_faces;
    P.fcol = Coloring{};
    if (!n_faces) return FEMGPU_OK;
    P.fcol = color_elements(P.n_nodes, n_faces, npf, faces);
    HIPCHK(upload(&P.d_faces, faces, (size_t)n_faces * npf));
    HIPCHK(upload(&P.d_forder, P.fcol.order.data(), P.fcol.order.size()));
    HIPCHK(upload(&P.d_trac, trac, 3 * (size_t)n_faces));
    return FEMGPU_OK;
}

int femgpu_add_constraint(femgpu_problem *h, uint32_t dof, double value, double penalty) {
    return femgpu_add_constraints(h, 1, &dof, &value, penalty);
}

int femgpu_add_constraints(femgpu_problem *h, uint32_t n, const uint32_t *dofs, const double *values, double penalty) {
    if (!h || (n && !dofs)) return fail(FEMGPU_BAD_INPUT, "add_constraints: bad input");
    Problem &P = h->P;
    if (!(penalty > 0)) return fail(FEMGPU_BAD_INPUT, "add_constraints: only the penalty form (penalty > 0) keeps K SPD");
    for (uint32_t k = 0; k < n; ++k) {
        if (dofs[k] >= 3 * P.n_nodes) return fail(FEMGPU_BAD_INPUT, "add_constraints: dof out of range");
        P.cons.push_back({dofs[k], values ? values[k] : 0.0, penalty});
    }
    P.cons_dirty = true;
    return FEMGPU_OK;
}

int femgpu_clear_constraints(femgpu_problem *h) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "clear_constraints: null");
    h->P.cons.clear();
    h->P.cons_dirty = true;
    return FEMGPU_OK;
}

int femgpu_assemble(femgpu_problem *h, int do_mat, int do_vec) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "assemble: null");
    Problem &P = h->P;
    if (P.from_bsr) return fail(FEMGPU_BAD_INPUT, "assemble: problem built from a matrix");
    if (!P.material_set) return fail(FEMGPU_NOT_READY, "assemble: material not set");
    int rc = upload_constraints(P);
    if (rc) return rc;
    HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t *>(&P.d_sc->pad), 1, 0u, P.stream));
    HIPCHK(hipEventRecord(P.ev[0], P.stream));
    HIPCHK(launch_assemble(P, do_mat, do_vec, &P.d_sc->pad));
    HIPCHK(launch_constraints(P, do_mat, do_vec));
    HIPCHK(hipEventRecord(P.ev[1], P.stream));
    HIPCHK(hipMemcpyAsync(P.h_sc, P.d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, P.ev[0], P.ev[1]));
    P.last_asm_ms = ms;
    if (do_mat) { P.assembled_mat = true; P.k32_valid = false; P.minv_kind = -1; P.max_diag = P.h_sc->max_diag; }
    if (do_vec) P.assembled_vec = true;
    if (P.h_sc->pad) return fail(FEMGPU_NEG_JACOBIAN, "assemble: element with det J <= 0");
    return FEMGPU_OK;
}

int femgpu_solve_system(femgpu_problem *h, const femgpu_solve_opts *o, femgpu_solve_stats *st) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "solve_system: null");
    int rc = check_opts(o);
    if (rc) return rc;
    Problem &P = h->P;
    if (!P.assembled_mat || !P.assembled_vec) return fail(FEMGPU_NOT_READY, "solve_system: assemble first");
    SolveOut s;
    rc = solve_pcg(P, P.d_F, P.d_u, to_internal(o), s);
    to_stats(s, st);
    if (st) st->status = rc;
    if (rc) return fail(rc, std::string("solve_system: ") + femgpu_status_string(rc));
    return FEMGPU_OK;
}

int femgpu_solve(femgpu_problem *h, const femgpu_solve_opts *o, femgpu_solve_stats *st) {
    int rc = femgpu_assemble(h, 1, 1);
    return rc ? rc : femgpu_solve_system(h, o, st);
}

int femgpu_get_result(femgpu_problem *h, double *u, size_t n) {
    if (!h || !u || n != 3 * (size_t)h->P.n_nodes) return fail(FEMGPU_BAD_INPUT, "get_result: size");
    HIPCHK(hipStreamSynchronize(h->P.stream));
    HIPCHK(hipMemcpy(u, h->P.d_u, n * 8, hipMemcpyDeviceToHost));
    return FEMGPU_OK;
}

int femgpu_get_result_device(femgpu_problem *h, double *d_u) {
    if (!h || !d_u) return fail(FEMGPU_BAD_INPUT, "get_result_device: null");
    HIPCHK(hipMemcpyAsync(d_u, h->P.d_u, 3 * (size_t)h->P.n_nodes * 8, hipMemcpyDeviceToDevice, h->P.stream));
    HIPCHK(hipStreamSynchronize(h->P.stream));
    return FEMGPU_OK;
}

int femgpu_set_result(femgpu_problem *h, const double *u, size_t n) {
    if (!h || !u || n != 3 * (size_t)h->P.n_nodes) return fail(FEMGPU_BAD_INPUT, "set_result: size");
    HIPCHK(hipStreamSynchronize(h->P.stream));
    HIPCHK(hipMemcpy(h->P.d_u, u, n * 8, hipMemcpyHostToDevice));
    return FEMGPU_OK;
}

int femgpu_get_sollicitation(femgpu_problem *h, double *F, size_t n) {
    if (!h || !F || n != 3 * (size_t)h->P.n_nodes) return fail(FEMGPU_BAD_INPUT, "get_sollicitation: size");
    HIPCHK(hipStreamSynchronize(h->P.stream));
    HIPCHK(hipMemcpy(F, h->P.d_F, n * 8, hipMemcpyDeviceToHost));
    return FEMGPU_OK;
}

int femgpu_get_matrix(femgpu_problem *h, uint32_t *row_ptr, uint32_t *col, double *vals) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_matrix: null");
    Problem &P = h->P;
    if (row_ptr) std::memcpy(row_ptr, P.pat.row_ptr.data(), P.pat.row_ptr.size() * 4);
    if (col) std::memcpy(col, P.pat.col.data(), P.pat.col.size() * 4);
    if (vals) {
        std::vector<double> planes(9 * P.ps);
        HIPCHK(hipStreamSynchronize(P.stream));
        HIPCHK(hipMemcpy(planes.data(), P.d_K, planes.size() * 8, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < P.nnzb; ++p)
            for (int c = 0; c < 9; ++c) vals[9 * p + c] = planes[c * P.ps + p];
    }
    return FEMGPU_OK;
}

int femgpu_get_sizes(femgpu_problem *h, uint32_t *n_nodes, uint32_t *n_elem, uint64_t *nnzb, uint32_t *n_colors) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_sizes: null");
    if (n_nodes) *n_nodes = h->P.n_nodes;
    if (n_elem) *n_elem = h->P.n_elem;
    if (nnzb) *nnzb = h->P.nnzb;
    if (n_colors) *n_colors = h->P.ecol.n_colors();
    return FEMGPU_OK;
}

int femgpu_get_max_diag(femgpu_problem *h, double *md) {
    if (!h || !md) return fail(FEMGPU_BAD_INPUT, "get_max_diag: null");
    *md = h->P.max_diag;
    return FEMGPU_OK;
}

int femgpu_elem_grad(femgpu_problem *h, const double *u, const double *lam, double *g) {
    if (!h || !u || !lam || !g || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "elem_grad: bad input");
    Problem &P = h->P;
    const size_t n = 3 * (size_t)P.n_nodes;
    double *d_g = nullptr;
    HIPCHK(dalloc(&d_g, P.n_elem));
    hipError_t e = hipMemcpy(P.d_aux, u, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P.d_q, lam, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_elem_grad(P, P.d_aux, P.d_q, d_g);
    if (e == hipSuccess) e = hipStreamSynchronize(P.stream);
    if (e == hipSuccess) e = hipMemcpy(g, d_g, (size_t)P.n_elem * 8, hipMemcpyDeviceToHost);
    (void)hipFree(d_g);
    if (e != hipSuccess) return fail(FEMGPU_HIP_ERROR, hipGetErrorString(e));
    return FEMGPU_OK;
}

int femgpu_adjoint_device(femgpu_problem *h, const double *d_dJdu, double *d_dJdE, const femgpu_solve_opts *o,
                          femgpu_solve_stats *st) {
    if (!h || !d_dJdu || !d_dJdE || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "adjoint: bad input");
    int rc = check_opts(o);
    if (rc) return rc;
    Problem &P = h->P;
    if (!P.assembled_mat) return fail(FEMGPU_NOT_READY, "adjoint: assemble first");
    const size_t n = 3 * (size_t)P.n_nodes;
    // lambda in d_aux: K lambda = dJ/du (K symmetric), zero initial guess
    double *d_rhs = nullptr;
    HIPCHK(dalloc(&d_rhs, n));
    hipError_t e = hipMemcpyAsync(d_rhs, d_dJdu, n * 8, hipMemcpyDeviceToDevice, P.stream);
    if (e == hipSuccess) e = launch_fill(P.d_aux, n, 0.0, P.stream);
    if (e != hipSuccess) { (void)hipFree(d_rhs); return fail(FEMGPU_HIP_ERROR, hipGetErrorString(e)); }
    SolveOut s;
    rc = solve_pcg(P, d_rhs, P.d_aux, to_internal(o), s);
    to_stats(s, st);
    if (rc == FEMGPU_OK) {
        e = launch_elem_grad(P, P.d_u, P.d_aux, d_dJdE);
        if (e == hipSuccess) e = hipStreamSynchronize(P.stream);
        if (e != hipSuccess) rc = fail(FEMGPU_HIP_ERROR, hipGetErrorString(e));
    }
    (void)hipFree(d_rhs);
    if (rc) return fail(rc, std::string("adjoint: ") + femgpu_status_string(rc));
    return FEMGPU_OK;
}

int femgpu_adjoint(femgpu_problem *h, const double *dJdu, double *dJdE, const femgpu_solve_opts *o,
                   femgpu_solve_stats *st) {
    if (!h || !dJdu || !dJdE) return fail(FEMGPU_BAD_INPUT, "adjoint: null");
    Problem &P = h->P;
    const size_t n = 3 * (size_t)P.n_nodes;
    double *d_in = nullptr, *d_out = nullptr;
    HIPCHK(dalloc(&d_in, n));
    if (dalloc(&d_out, P.n_elem) != hipSuccess) { (void)hipFree(d_in); return fail(FEMGPU_HIP_ERROR, "adjoint: alloc"); }
    int rc = FEMGPU_OK;
    if (hipMemcpy(d_in, dJdu, n * 8, hipMemcpyHostToDevice) != hipSuccess) rc = fail(FEMGPU_HIP_ERROR, "adjoint: copy");
    if (!rc) rc = femgpu_adjoint_device(h, d_in, d_out, o, st);
    if (!rc && hipMemcpy(dJdE, d_out, (size_t)P.n_elem * 8, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(FEMGPU_HIP_ERROR, "adjoint: copy back");
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return rc;
}

int femgpu_get_assemble_ms(femgpu_problem *h, double *ms) {
    if (!h || !ms) return fail(FEMGPU_BAD_INPUT, "get_assemble_ms: null");
    *ms = h->P.last_asm_ms;
    return FEMGPU_OK;
}

int femgpu_time_spmv(femgpu_problem *h, int precision, int reps, double *avg_ms) {
    if (!h || !avg_ms || reps < 1 || precision < 0 || precision > 1) return fail(FEMGPU_BAD_INPUT, "time_spmv: bad input");
    Problem &P = h->P;
    if (precision == 1 && !P.k32_valid) {
        if (!P.d_K32) HIPCHK(dalloc(&P.d_K32, 9 * P.ps));
        HIPCHK(launch_to_fp32(P));
        P.k32_valid = true;
    }
    const double t = time_spmv(P, precision, reps);
    if (t < 0) return fail(FEMGPU_HIP_ERROR, "time_spmv failed");
    *avg_ms = t;
    return FEMGPU_OK;
}

}  // extern "C"
