// C ABI of the MI355X assemble + solve path (declarations and reference citations in
// include/femgpu.h).  Single problems and batches share one implementation: a Problem with
// n_sub sub-problems (n_sub = 1 for femgpu_create).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <tuple>
#include <new>

#include "../../include/femgpu.h"
#include "capi_handles.hpp"
#include "femgpu_internal.hpp"
#include "io_avs.hpp"
#include "io_vtu.hpp"

namespace femgpu {
std::string &last_error() {
    static thread_local std::string s;
    return s;
}
}  // namespace femgpu

using namespace femgpu;

namespace {

int fail(int status, const std::string &msg) {
    last_error() = msg;
    return status;
}

#define HIPCHK(expr)                                                                                              \
    do {                                                                                                          \
        hipError_t e_ = (expr);                                                                                   \
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

void free_gather(GatherAsm &G) {
    dfree(G.d_em); dfree(G.d_ef); dfree(G.d_kptr); dfree(G.d_kent); dfree(G.d_fptr); dfree(G.d_fent);
    dfree(G.d_cls); dfree(G.d_ch); dfree(G.d_csub); dfree(G.d_kref); dfree(G.d_fref);
    G = GatherAsm();
}

void free_problem(Problem &P) {
    destroy_graph(P);
    ic0_free(P.ic0);
    sym_free(P.sym);
    mf_free(P.mf);
    dfree(P.d_coords); dfree(P.d_E); dfree(P.d_conn); dfree(P.d_rp); dfree(P.d_col); dfree(P.d_diag); dfree(P.d_col16);
    dfree(P.d_slot); dfree(P.d_order); dfree(P.d_sptr); dfree(P.d_elem_sub); dfree(P.d_mat);
    dfree(P.d_K); dfree(P.d_F); dfree(P.d_u); dfree(P.d_K32);
    dfree(P.d_faces); dfree(P.d_forder); dfree(P.d_trac);
    dfree(P.d_cdof); dfree(P.d_cptr); dfree(P.d_csub); dfree(P.d_cval); dfree(P.d_cpen);
    dfree(P.d_r); dfree(P.d_p); dfree(P.d_q); dfree(P.d_Minv); dfree(P.d_aux); dfree(P.d_red);
    dfree(P.d_r32); dfree(P.d_p32); dfree(P.d_q32); dfree(P.d_x32); dfree(P.d_Minv32); dfree(P.d_z); dfree(P.d_z32);
    dfree(P.d_partials); dfree(P.d_ps); dfree(P.d_flags);
    dfree(P.d_wg_row0); dfree(P.d_wg_row1); dfree(P.d_wg_sub); dfree(P.d_sub_wg0); dfree(P.d_sub_nwg);
    dfree(P.d_fwg_row0); dfree(P.d_fwg_row1); dfree(P.d_fwg_sub); dfree(P.d_fsub_wg0); dfree(P.d_fsub_nwg);
    dfree(P.d_gptr); dfree(P.d_gel); dfree(P.d_perm); dfree(P.d_iperm); dfree(P.d_nperm); dfree(P.d_ninv); dfree(P.d_Kb);
    free_gather(P.ga);
    if (P.h_ps) (void)hipHostFree(P.h_ps);
    if (P.h_flags) (void)hipHostFree(P.h_flags);
    P.h_ps = nullptr;
    P.h_flags = nullptr;
    for (auto &e : P.ev)
        if (e) (void)hipEventDestroy(e);
    if (P.stream) (void)hipStreamDestroy(P.stream);
    P.stream = nullptr;
}

// Workgroup tables: every sub-problem's rows split into workgroups of up to `rows` rows; the grid
// is rounded up to a multiple of 8 with idle workgroups (sub = NO_SUB).  Returns the grid size.
uint32_t build_wg_tables(Problem &P, uint32_t rows, std::vector<uint32_t> &row0, std::vector<uint32_t> &row1,
                         std::vector<uint32_t> &sub, std::vector<uint32_t> &swg0, std::vector<uint32_t> &snwg) {
    swg0.assign(P.n_sub, 0);
    snwg.assign(P.n_sub, 0);
    for (uint32_t s = 0; s < P.n_sub; ++s) {
        swg0[s] = (uint32_t)row0.size();
        for (uint32_t r = P.node_off[s]; r < P.node_off[s + 1]; r += rows) {
            row0.push_back(r);
            row1.push_back(std::min(P.node_off[s + 1], r + rows));
            sub.push_back(s);
        }
        snwg[s] = (uint32_t)row0.size() - swg0[s];
    }
    while (row0.size() % 8 || row0.empty()) { row0.push_back(0); row1.push_back(0); sub.push_back(NO_SUB); }
    return (uint32_t)row0.size();
}

// Classes of ASM_GATHER_BOX: every element an axis-aligned Hexa box (else 1), classed by its exact
// edge lengths and its sub-problem, at most MAX_BOX_CLASSES classes (else 1).
static int box_classes(const Problem &P, std::vector<uint16_t> &cls, std::vector<double> &ch, std::vector<uint32_t> &csub) {
    if (P.kind != HEXA || P.n_elem == 0) return 1;
    std::vector<double> X(3 * (size_t)P.n_nodes);
    if (hipMemcpy(X.data(), P.d_coords, X.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) return 5;
    std::map<std::tuple<double, double, double, uint32_t>, uint16_t> ids;
    cls.resize(P.n_elem);
    uint32_t s = 0;
    for (uint32_t e = 0; e < P.n_elem; ++e) {
        while (e >= P.elem_off[s + 1]) ++s;
        const uint32_t *nd = P.conn_h.data() + 8 * (size_t)e;
        if (!hexa_is_box(X.data(), nd)) return 1;
        const double *x0 = &X[3 * (size_t)nd[0]], *x6 = &X[3 * (size_t)nd[6]];
        const auto key = std::make_tuple(x6[0] - x0[0], x6[1] - x0[1], x6[2] - x0[2], s);
        auto it = ids.find(key);
        if (it == ids.end()) {
            if (ids.size() >= MAX_BOX_CLASSES) return 1;
            it = ids.emplace(key, (uint16_t)ids.size()).first;
            ch.insert(ch.end(), {std::get<0>(key), std::get<1>(key), std::get<2>(key)});
            csub.push_back(s);
        }
        cls[e] = it->second;
    }
    return 0;
}

// The gather lists of ASM_GATHER (host, once per problem): every (element, a, b) block goes to
// the list of its position, every (element, a) load to the list of its node, walking the elements
// in colour order, so each list is in the order the colour passes add.  box: the lists of
// ASM_GATHER_BOX, with the element classes instead of the stored element blocks.
static int build_gather(Problem &P, bool box) {
    if (box ? P.ga.box : (P.ga.d_em != nullptr)) return FEMGPU_OK;
    std::vector<uint16_t> cls;
    std::vector<double> ch;
    std::vector<uint32_t> csub;
    if (box) {
        const int rc = box_classes(P, cls, ch, csub);
        if (rc == 1) return fail(FEMGPU_BAD_INPUT, "assembly: the box gather variant needs Hexa elements that are all "
                                                   "axis-aligned boxes, of at most 4096 shapes");
        if (rc) return fail(FEMGPU_HIP_ERROR, "assembly: box classes");
    }
    const uint32_t npe = (uint32_t)P.npe, nup = npe * (npe + 1) / 2;
    if ((uint64_t)P.n_elem * nup * 2 >= (1ull << 32) || (uint64_t)P.n_elem * npe * npe >= (1ull << 32) ||
        P.ps >= (1ull << 32) || P.ecol.order.size() != P.n_elem)
        return fail(FEMGPU_BAD_INPUT, "assembly: the gather variant indexes at most 2^32 element blocks");
    const uint32_t *conn = P.conn_h.data();
    const std::vector<uint32_t> slot = build_slot_map(P.pat, P.n_elem, (int)npe, conn);
    std::vector<uint32_t> kptr(P.ps + 1, 0), fptr(P.n_nodes + 1, 0);
    for (size_t i = 0; i < slot.size(); ++i) ++kptr[(size_t)slot[i] + 1];
    for (size_t i = 0; i < (size_t)P.n_elem * npe; ++i) ++fptr[(size_t)conn[i] + 1];
    for (size_t p = 0; p < P.ps; ++p) kptr[p + 1] += kptr[p];
    for (uint32_t n = 0; n < P.n_nodes; ++n) fptr[n + 1] += fptr[n];
    std::vector<uint32_t> kent(kptr.back()), fent(fptr.back());
    std::vector<uint32_t> kf(kptr.begin(), kptr.end() - 1), ff(fptr.begin(), fptr.end() - 1);
    for (const uint32_t e : P.ecol.order) {
        for (uint32_t a = 0; a < npe; ++a) {
            for (uint32_t b = 0; b < npe; ++b) {
                const uint32_t lo = std::min(a, b), hi = std::max(a, b);
                const uint32_t k = lo * npe - lo * (lo - 1) / 2 + (hi - lo);  // lo (lo - 1) is even
                kent[kf[slot[((size_t)e * npe + a) * npe + b]]++] = ((e * nup + k) << 1) | (a > b ? 1u : 0u);
            }
            fent[ff[conn[(size_t)e * npe + a]]++] = e * npe + a;
        }
    }
    free_gather(P.ga);
    GatherAsm G;
    bool ok = upload(&G.d_kptr, kptr.data(), kptr.size()) == hipSuccess && upload(&G.d_kent, kent.data(), kent.size()) == hipSuccess &&
              upload(&G.d_fptr, fptr.data(), fptr.size()) == hipSuccess && upload(&G.d_fent, fent.data(), fent.size()) == hipSuccess;
    if (ok && box) {
        G.box = true;
        G.n_classes = (uint32_t)csub.size();
        ok = upload(&G.d_cls, cls.data(), cls.size()) == hipSuccess && upload(&G.d_ch, ch.data(), ch.size()) == hipSuccess &&
             upload(&G.d_csub, csub.data(), csub.size()) == hipSuccess &&
             dalloc(&G.d_kref, (size_t)G.n_classes * nup * 9) == hipSuccess &&
             dalloc(&G.d_fref, (size_t)G.n_classes * npe * 3) == hipSuccess;
    } else if (ok) {
        ok = dalloc(&G.d_em, (size_t)P.n_elem * nup * 9) == hipSuccess && dalloc(&G.d_ef, (size_t)P.n_elem * npe * 3) == hipSuccess;
    }
    if (!ok) {
        free_gather(G);
        return fail(FEMGPU_HIP_ERROR, "assembly: gather buffers");
    }
    P.ga = G;
    return FEMGPU_OK;
}

// Element-loop variant: 0 = colour passes, ASM_STAGED = colour passes through the block-contiguous
// copy, ASM_GATHER = element blocks then per-position gathers, R = row groups of R rows (a power
// of two <= 64 whose LDS accumulator fits, see max_assembly_rows).
int set_assembly_rows(Problem &P, uint32_t R) {
    if (R == ASM_GATHER || R == ASM_GATHER_BOX) {
        const int rc = build_gather(P, R == ASM_GATHER_BOX);
        if (rc) return rc;
        dfree(P.d_Kb);
        P.asm_rows = R;
        return FEMGPU_OK;
    }
    free_gather(P.ga);
    if (R == ASM_STAGED) {
        if (!P.d_Kb) HIPCHK(dalloc(&P.d_Kb, 9 * P.ps));
        P.asm_rows = ASM_STAGED;
        return FEMGPU_OK;
    }
    dfree(P.d_Kb);
    if (R == 0) {
        P.asm_rows = 0;
        return FEMGPU_OK;
    }
    if (R > 64 || (R & (R - 1)) || R > max_assembly_rows(P))
        return fail(FEMGPU_BAD_INPUT, "assembly rows: a power of two, at most " + std::to_string(max_assembly_rows(P)));
    if (P.rg.R != R) {
        P.rg = build_row_groups(P.n_nodes, R, P.n_elem, P.npe, P.conn_h.data(), P.ecol,
                                P.pat.iperm.empty() ? nullptr : P.pat.iperm.data());
        dfree(P.d_gptr);
        dfree(P.d_gel);
        HIPCHK(upload(&P.d_gptr, P.rg.ptr.data(), P.rg.ptr.size()));
        HIPCHK(upload(&P.d_gel, P.rg.elems.data(), P.rg.elems.size()));
    }
    P.asm_rows = R;
    return FEMGPU_OK;
}

namespace {
std::atomic<int> g_row_order{ROW_ORDER_AUTO};
std::atomic<int> g_node_order{NODE_ORDER_NATURAL};
}
int row_order_default() { return g_row_order.load(); }
int node_order_default() { return g_node_order.load(); }

// Device buffers common to mesh problems and plain block systems.
int alloc_system(Problem &P) {
    const size_t n = 3 * (size_t)P.n_nodes;
    HIPCHK(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    for (auto &e : P.ev) HIPCHK(hipEventCreate(&e));
    const int order = row_order_default();
    if (order != ROW_ORDER_NATURAL)
        sort_rows_by_length(P.pat, P.node_off, SORT_WINDOW_WGS * ROWS_PER_WG, order == ROW_ORDER_BY_LENGTH ? -1.0 : 0.03);
    P.ps = (P.pat.n_sell + 63) / 64 * 64;
    if (!P.pat.perm.empty()) {
        HIPCHK(upload(&P.d_perm, P.pat.perm.data(), P.pat.perm.size()));
        HIPCHK(upload(&P.d_iperm, P.pat.iperm.data(), P.pat.iperm.size()));
    }
    HIPCHK(upload(&P.d_rp, P.pat.row_ptr.data(), P.pat.row_ptr.size()));
    {
        std::vector<uint32_t> cs = sell_columns(P.pat);
        HIPCHK(upload(&P.d_col, cs.data(), cs.size()));
        // 16-bit column offsets c - r, r the block row held by the position's storage row
        std::vector<int16_t> c16(cs.size(), 0);
        bool fits = true;
        for (uint32_t t = 0; t < P.pat.n_rows && fits; ++t) {
            const uint32_t r = P.pat.row_of(t), s = t >> 6, w = (P.pat.sell_ptr[s + 1] - P.pat.sell_ptr[s]) / 64;
            for (uint32_t k = 0; k < w; ++k) {
                const int64_t d = (int64_t)cs[P.pat.pos(r, k)] - (int64_t)r;
                if (d < -32768 || d > 32767) { fits = false; break; }
                c16[P.pat.pos(r, k)] = (int16_t)d;
            }
        }
        if (fits) HIPCHK(upload(&P.d_col16, c16.data(), c16.size()));
        else P.spmv.loads[0] = P.spmv.loads[1] = LOAD_NT;
    }
    HIPCHK(upload(&P.d_sptr, P.pat.sell_ptr.data(), P.pat.sell_ptr.size()));
    HIPCHK(upload(&P.d_diag, P.pat.diag.data(), P.pat.diag.size()));
    {
        std::vector<uint32_t> row0, row1, sub, swg0, snwg;
        P.n_wg = build_wg_tables(P, ROWS_PER_WG, row0, row1, sub, swg0, snwg);
        HIPCHK(upload(&P.d_wg_row0, row0.data(), row0.size()));
        HIPCHK(upload(&P.d_wg_row1, row1.data(), row1.size()));
        HIPCHK(upload(&P.d_wg_sub, sub.data(), sub.size()));
        HIPCHK(upload(&P.d_sub_wg0, swg0.data(), swg0.size()));
        HIPCHK(upload(&P.d_sub_nwg, snwg.data(), snwg.size()));
    }
    {
        std::vector<uint32_t> row0, row1, sub, swg0, snwg;
        P.n_fwg = build_wg_tables(P, FOLD_ROWS_PER_WG, row0, row1, sub, swg0, snwg);
        HIPCHK(upload(&P.d_fwg_row0, row0.data(), row0.size()));
        HIPCHK(upload(&P.d_fwg_row1, row1.data(), row1.size()));
        HIPCHK(upload(&P.d_fwg_sub, sub.data(), sub.size()));
        HIPCHK(upload(&P.d_fsub_wg0, swg0.data(), swg0.size()));
        HIPCHK(upload(&P.d_fsub_nwg, snwg.data(), snwg.size()));
    }
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
    HIPCHK(dalloc(&P.d_z, n));
    HIPCHK(dalloc(&P.d_z32, n));
    HIPCHK(dalloc(&P.d_partials, 3 * (size_t)P.n_wg));
    HIPCHK(dalloc(&P.d_ps, 2 * (size_t)P.n_sub));
    HIPCHK(dalloc(&P.d_flags, 4));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&P.h_ps), sizeof(ProbScalars) * P.n_sub, hipHostMallocDefault));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&P.h_flags), sizeof(int) * 4, hipHostMallocDefault));
    std::memset(P.h_ps, 0, sizeof(ProbScalars) * P.n_sub);
    HIPCHK(hipMemcpy(P.d_ps, P.h_ps, sizeof(ProbScalars) * P.n_sub, hipMemcpyHostToDevice));
    P.max_diag.assign(P.n_sub, 0.0);
    HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t *>(P.d_flags), 4, 0u, P.stream));
    HIPCHK(launch_fill(P.d_u, n, 0.0, P.stream));
    HIPCHK(launch_fill(P.d_p, n, 0.0, P.stream));
    HIPCHK(launch_fill(P.d_K, 9 * P.ps, 0.0, P.stream));
    HIPCHK(launch_fill(P.d_F, n, 0.0, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    return FEMGPU_OK;
}

// n meshes of one element kind concatenated into one device system.
int build_problem(femgpu_problem &H, uint32_t n_sub, const femgpu_mesh *meshes) {
    Problem &P = H.P;
    if (!n_sub || !meshes) return fail(FEMGPU_BAD_INPUT, "create: no mesh");
    const int kind = meshes[0].elem_kind;
    if (kind < 0 || kind > 2) return fail(FEMGPU_BAD_INPUT, "create: unknown element kind");
    const int npe = nodes_per_elem(kind);
    P.kind = kind;
    P.npe = npe;
    P.n_sub = n_sub;
    P.node_off.assign(n_sub + 1, 0);
    P.elem_off.assign(n_sub + 1, 0);
    for (uint32_t s = 0; s < n_sub; ++s) {
        const femgpu_mesh &m = meshes[s];
        if (m.elem_kind != kind || !m.coords || !m.conn || !m.n_nodes || !m.n_elem)
            return fail(FEMGPU_BAD_INPUT, "create: bad mesh " + std::to_string(s));
        for (size_t k = 0; k < (size_t)m.n_elem * npe; ++k)
            if (m.conn[k] >= m.n_nodes) return fail(FEMGPU_BAD_INPUT, "create: node index out of range");
        for (size_t k = 0; k < 3 * (size_t)m.n_nodes; ++k)
            if (!std::isfinite(m.coords[k])) return fail(FEMGPU_NAN, "create: non-finite coordinate");
        if ((uint64_t)P.node_off[s] + m.n_nodes > 0x7fffffffull || (uint64_t)P.elem_off[s] + m.n_elem > 0x7fffffffull)
            return fail(FEMGPU_BAD_INPUT, "create: mesh too large");
        P.node_off[s + 1] = P.node_off[s] + m.n_nodes;
        P.elem_off[s + 1] = P.elem_off[s] + m.n_elem;
    }
    P.n_nodes = P.node_off[n_sub];
    P.n_elem = P.elem_off[n_sub];
    std::vector<double> coords(3 * (size_t)P.n_nodes);
    std::vector<uint32_t> conn((size_t)P.n_elem * npe), esub(P.n_elem);
    for (uint32_t s = 0; s < n_sub; ++s) {
        const femgpu_mesh &m = meshes[s];
        std::memcpy(coords.data() + 3 * (size_t)P.node_off[s], m.coords, 24 * (size_t)m.n_nodes);
        uint32_t *c = conn.data() + (size_t)P.elem_off[s] * npe;
        for (size_t k = 0; k < (size_t)m.n_elem * npe; ++k) c[k] = m.conn[k] + P.node_off[s];
        std::fill(esub.begin() + P.elem_off[s], esub.begin() + P.elem_off[s + 1], s);
    }
    (void)hipGetDevice(&P.device);
    const int node_order = node_order_default();
    if (node_order != NODE_ORDER_NATURAL) {
        // want_rcm / want_amd renumbering (formulation.h:314-323), per sub-problem
        const Pattern api = build_pattern(P.n_nodes, P.n_elem, npe, conn.data());
        std::vector<uint32_t> perm;
        perm.reserve(P.n_nodes);
        for (uint32_t s = 0; s < n_sub; ++s) {
            const std::vector<uint32_t> o = node_order == NODE_ORDER_AMD
                                                ? amd_order(api, P.node_off[s], P.node_off[s + 1])
                                                : cm_order(api, P.node_off[s], P.node_off[s + 1]);
            perm.insert(perm.end(), o.begin(), o.end());
        }
        bool identity = true;
        for (uint32_t i = 0; i < P.n_nodes && identity; ++i) identity = perm[i] == i;
        if (!identity) {
            P.ninv.assign(P.n_nodes, 0);
            for (uint32_t i = 0; i < P.n_nodes; ++i) P.ninv[perm[i]] = i;
            P.nperm = std::move(perm);
            P.node_order = node_order;
            for (uint32_t &v : conn) v = P.ninv[v];
            std::vector<double> c2(coords.size());
            for (uint32_t i = 0; i < P.n_nodes; ++i)
                for (int d = 0; d < 3; ++d) c2[3 * (size_t)i + d] = coords[3 * (size_t)P.nperm[i] + d];
            coords.swap(c2);
        }
    }
    P.pat = build_pattern(P.n_nodes, P.n_elem, npe, conn.data());
    P.nnzb = P.pat.col.size();
    P.ecol = color_elements(P.n_nodes, P.n_elem, npe, conn.data());
    int rc = alloc_system(P);
    if (rc) return rc;
    std::vector<uint32_t> slot = build_slot_map(P.pat, P.n_elem, npe, conn.data());
    HIPCHK(upload(&P.d_slot, slot.data(), slot.size()));
    if (!P.nperm.empty()) {
        HIPCHK(upload(&P.d_nperm, P.nperm.data(), P.nperm.size()));
        HIPCHK(upload(&P.d_ninv, P.ninv.data(), P.ninv.size()));
    }
    HIPCHK(upload(&P.d_coords, coords.data(), coords.size()));
    HIPCHK(upload(&P.d_conn, conn.data(), conn.size()));
    HIPCHK(upload(&P.d_order, P.ecol.order.data(), P.ecol.order.size()));
    HIPCHK(upload(&P.d_elem_sub, esub.data(), esub.size()));
    HIPCHK(dalloc(&P.d_E, P.n_elem));
    P.mat_h.assign(n_sub, SubMaterial{0.3, 0, 0, 0});
    P.material_set.assign(n_sub, 0);
    HIPCHK(upload(&P.d_mat, P.mat_h.data(), P.mat_h.size()));
    H.sub_faces.assign(n_sub, {});
    H.sub_trac.assign(n_sub, {});
    P.conn_h = std::move(conn);
    // staged colour passes: the fastest variant on the lattices measured (Hexa8 100^3 7.1 ms,
    // Tetra_10 40^3 4.7 ms, against 11.0 / 11.7 ms for the best row groups)
    return set_assembly_rows(P, ASM_STAGED);
}

int set_material(femgpu_problem &H, uint32_t s, const femgpu_material *m) {
    Problem &P = H.P;
    if (!m || P.from_bsr || s >= P.n_sub) return fail(FEMGPU_BAD_INPUT, "set_material: bad input");
    if (!(m->nu > -1.0 && m->nu < 0.5)) return fail(FEMGPU_BAD_INPUT, "set_material: nu out of (-1, 0.5)");
    const uint32_t ne = P.elem_off[s + 1] - P.elem_off[s];
    for (int d = 0; d < 3; ++d)
        if (!std::isfinite(m->body_force[d])) return fail(FEMGPU_NAN, "set_material: non-finite body force");
    if (m->E_elem) {
        for (uint32_t e = 0; e < ne; ++e)
            if (!std::isfinite(m->E_elem[e])) return fail(FEMGPU_NAN, "set_material: non-finite E");
        HIPCHK(hipMemcpy(P.d_E + P.elem_off[s], m->E_elem, (size_t)ne * 8, hipMemcpyHostToDevice));
    } else {
        if (!std::isfinite(m->E)) return fail(FEMGPU_NAN, "set_material: non-finite E");
        HIPCHK(launch_fill(P.d_E + P.elem_off[s], ne, m->E, P.stream));
        HIPCHK(hipStreamSynchronize(P.stream));
    }
    P.mat_h[s] = SubMaterial{m->nu, m->body_force[0], m->body_force[1], m->body_force[2]};
    HIPCHK(hipMemcpy(P.d_mat, P.mat_h.data(), P.mat_h.size() * sizeof(SubMaterial), hipMemcpyHostToDevice));
    P.material_set[s] = 1;
    return FEMGPU_OK;
}

int set_traction(femgpu_problem &H, uint32_t s, uint32_t n_faces, const uint32_t *faces, const double *trac) {
    Problem &P = H.P;
    if (P.from_bsr || s >= P.n_sub || (n_faces && (!faces || !trac))) return fail(FEMGPU_BAD_INPUT, "set_traction: bad input");
    const int npf = nodes_per_face(face_kind_of(P.kind));
    const uint32_t nn = P.node_off[s + 1] - P.node_off[s];
    for (size_t k = 0; k < (size_t)n_faces * npf; ++k)
        if (faces[k] >= nn) return fail(FEMGPU_BAD_INPUT, "set_traction: node index out of range");
    for (size_t k = 0; k < 3 * (size_t)n_faces; ++k)
        if (!std::isfinite(trac[k])) return fail(FEMGPU_NAN, "set_traction: non-finite traction");
    H.sub_faces[s].assign(faces, faces + (size_t)n_faces * npf);
    H.sub_trac[s].assign(trac, trac + 3 * (size_t)n_faces);
    H.faces_dirty = true;
    return FEMGPU_OK;
}

int upload_faces(femgpu_problem &H) {
    Problem &P = H.P;
    if (!H.faces_dirty) return FEMGPU_OK;
    const int npf = nodes_per_face(face_kind_of(P.kind));
    std::vector<uint32_t> faces;
    std::vector<double> trac;
    for (uint32_t s = 0; s < P.n_sub; ++s) {
        for (uint32_t v : H.sub_faces[s]) faces.push_back(P.ninv.empty() ? v + P.node_off[s] : P.ninv[v + P.node_off[s]]);
        trac.insert(trac.end(), H.sub_trac[s].begin(), H.sub_trac[s].end());
    }
    dfree(P.d_faces); dfree(P.d_forder); dfree(P.d_trac);
    P.n_faces = (uint32_t)(trac.size() / 3);
    P.fcol = Coloring{};
    if (P.n_faces) {
        P.fcol = color_elements(P.n_nodes, P.n_faces, npf, faces.data());
        HIPCHK(upload(&P.d_faces, faces.data(), faces.size()));
        HIPCHK(upload(&P.d_forder, P.fcol.order.data(), P.fcol.order.size()));
        HIPCHK(upload(&P.d_trac, trac.data(), trac.size()));
    }
    H.faces_dirty = false;
    return FEMGPU_OK;
}

int add_constraints(femgpu_problem &H, uint32_t s, uint32_t n, const uint32_t *dofs, const double *values,
                    double penalty) {
    Problem &P = H.P;
    if (s >= P.n_sub || (n && !dofs)) return fail(FEMGPU_BAD_INPUT, "add_constraints: bad input");
    if (!(penalty > 0) || !std::isfinite(penalty))
        return fail(FEMGPU_BAD_INPUT, "add_constraints: only the penalty form (penalty > 0) keeps K SPD");
    const uint32_t nd = 3 * (P.node_off[s + 1] - P.node_off[s]);
    for (uint32_t k = 0; k < n; ++k) {
        if (dofs[k] >= nd) return fail(FEMGPU_BAD_INPUT, "add_constraints: dof out of range");
        if (values && !std::isfinite(values[k])) return fail(FEMGPU_NAN, "add_constraints: non-finite value");
    }
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t g = dofs[k] + 3 * P.node_off[s];
        P.cons.push_back({P.ninv.empty() ? g : 3 * P.ninv[g / 3] + g % 3, values ? values[k] : 0.0, penalty});
    }
    P.cons_dirty = true;
    return FEMGPU_OK;
}

int upload_constraints(Problem &P) {
    if (!P.cons_dirty) return FEMGPU_OK;
    dfree(P.d_cdof); dfree(P.d_cptr); dfree(P.d_csub); dfree(P.d_cval); dfree(P.d_cpen);
    // group by DOF, keeping insertion order inside a group
    std::vector<size_t> idx(P.cons.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return P.cons[a].dof < P.cons[b].dof; });
    std::vector<uint32_t> dof, ptr(1, 0), sub;
    std::vector<double> val, pen;
    for (size_t k = 0; k < idx.size(); ++k) {
        const auto &c = P.cons[idx[k]];
        if (dof.empty() || dof.back() != c.dof) {
            if (!dof.empty()) ptr.push_back((uint32_t)val.size());
            dof.push_back(c.dof);
            const uint32_t node = c.dof / 3;
            sub.push_back(uint32_t(std::upper_bound(P.node_off.begin(), P.node_off.end(), node) - P.node_off.begin()) - 1);
        }
        val.push_back(c.val);
        pen.push_back(c.pen);
    }
    if (!dof.empty()) ptr.push_back((uint32_t)val.size());
    P.n_cdof = (uint32_t)dof.size();
    P.cdof_h = dof;
    HIPCHK(upload(&P.d_cdof, dof.data(), dof.size()));
    HIPCHK(upload(&P.d_cptr, ptr.data(), ptr.size()));
    HIPCHK(upload(&P.d_csub, sub.data(), sub.size()));
    HIPCHK(upload(&P.d_cval, val.data(), val.size()));
    HIPCHK(upload(&P.d_cpen, pen.data(), pen.size()));
    P.cons_dirty = false;
    return FEMGPU_OK;
}

// Value and penalty of a text constraint at time t: its expression is c u_dof + b, so the penalty
// terms M P c^2 on the diagonal and M P c (-b) on the right-hand side of assemble_constraints
// (formulation.h:1058-1070) are those of u_dof = -b / c with penalty P c^2.
int eval_text_constraint(const femgpu_problem::TextConstraint &tc, double t, double &value, double &pen) {
    std::map<uint64_t, double> co;
    double b = 0;
    const std::string err = tc.expr.eval(t, co, b);
    if (!err.empty()) return fail(FEMGPU_BAD_INPUT, "constraint '" + tc.expr.text() + "': " + err);
    if (co.size() != 1 || co.begin()->first != tc.dof)
        return fail(FEMGPU_BAD_INPUT, "constraint '" + tc.expr.text() + "': no unknown left at time " + std::to_string(t));
    const double c = co.begin()->second;
    value = -b / c;
    pen = tc.penalty * c * c;
    if (!std::isfinite(value) || !std::isfinite(pen)) return fail(FEMGPU_NAN, "constraint '" + tc.expr.text() + "': not finite");
    return FEMGPU_OK;
}

int assemble(femgpu_problem &H, int do_mat, int do_vec) {
    Problem &P = H.P;
    if (P.from_bsr) return fail(FEMGPU_BAD_INPUT, "assemble: problem built from a matrix");
    for (uint32_t s = 0; s < P.n_sub; ++s)
        if (!P.material_set[s]) return fail(FEMGPU_NOT_READY, "assemble: material not set");
    // text constraints depending on time take their values at the current time
    for (const auto &tc : H.text_cons) {
        if (!tc.expr.uses_time()) continue;
        double v = 0, pen = 0;
        const int rc = eval_text_constraint(tc, H.time, v, pen);
        if (rc) return rc;
        Problem::Cons &c = P.cons[tc.slot];
        if (c.val != v || c.pen != pen) {
            c.val = v;
            c.pen = pen;
            P.cons_dirty = true;
        }
    }
    int rc = upload_constraints(P);
    if (!rc) rc = upload_faces(H);
    if (rc) return rc;
    HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t *>(P.d_flags), 4, 0u, P.stream));
    HIPCHK(hipEventRecord(P.ev[0], P.stream));
    HIPCHK(launch_assemble(P, do_mat, do_vec));
    HIPCHK(launch_constraints(P, do_mat, do_vec));
    HIPCHK(hipEventRecord(P.ev[1], P.stream));
    if (do_mat) HIPCHK(mf_snapshot(P));
    HIPCHK(hipMemcpyAsync(P.h_flags, P.d_flags, sizeof(int) * 4, hipMemcpyDeviceToHost, P.stream));
    HIPCHK(hipMemcpyAsync(P.h_ps, P.d_ps, sizeof(ProbScalars) * P.n_sub, hipMemcpyDeviceToHost, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, P.ev[0], P.ev[1]));
    P.last_asm_ms = ms;
    if (do_mat) {
        P.assembled_mat = true;
        P.k32_valid = false;
        if (H.precond_pinned) {
            H.precond_stale = true;  // get_precond's factors stay for solve_system_using_precond
        } else {
            P.minv_kind = -1;
            P.ic0.valid = P.ic0.valid32 = false;
        }
        P.sym.valid[0] = P.sym.valid[1] = false;
        for (uint32_t s = 0; s < P.n_sub; ++s) P.max_diag[s] = P.h_ps[s].max_diag;
    }
    if (do_vec) P.assembled_vec = true;
    if (P.h_flags[0]) return fail(FEMGPU_NEG_JACOBIAN, "assemble: element with det J <= 0");
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
    if (o->precision < 0 || o->precision > 1 || o->precond < 0 || o->precond > 4 || o->crit_kind < 0 ||
        o->crit_kind > 1 || !(o->tol >= 0) || o->max_iter < 0 || o->chunk < 1 || !(o->reliable_delta > 0) ||
        !(o->reliable_delta < 1))
        return fail(FEMGPU_BAD_INPUT, "invalid solve options");
    return FEMGPU_OK;
}

// A solve other than solve_system_using_precond builds the preconditioner of the current K: a
// preconditioner pinned by get_precond is released when K changed since (its factors no longer
// match K) or when the solve asks for another kind (building that one replaces it), so that
// solve_system_using_precond then reports NOT_READY instead of running with other factors.
void unpin_precond(femgpu_problem &H, const femgpu_solve_opts *o) {
    if (!H.precond_pinned) return;
    const int kind = o ? o->precond : H.dopts.precond;
    if (H.precond_stale || kind != H.pinned_kind) {
        if (H.precond_stale) {
            H.P.minv_kind = -1;
            H.P.ic0.valid = H.P.ic0.valid32 = false;
        }
        H.precond_pinned = H.precond_stale = false;
    }
}

int solve_system(femgpu_problem &H, const femgpu_solve_opts *o, femgpu_solve_stats *per, double *d_rhs, double *d_x) {
    int rc = check_opts(o);
    if (rc) return rc;
    Problem &P = H.P;
    if (!P.assembled_mat || !P.assembled_vec) return fail(FEMGPU_NOT_READY, "solve_system: assemble first");
    std::vector<SolveOut> out(P.n_sub);
    rc = solve_pcg(P, d_rhs ? d_rhs : P.d_F, d_x ? d_x : P.d_u, to_internal(o), out.data());
    if (per)
        for (uint32_t s = 0; s < P.n_sub; ++s) to_stats(out[s], &per[s]);
    if (rc) return fail(rc, std::string("solve_system: ") + femgpu_status_string(rc));
    return FEMGPU_OK;
}

// Node-major values of sub-problem s from the internal node order to the caller's, in place.
void to_api_order(const Problem &P, uint32_t s, double *v) {
    if (P.ninv.empty()) return;
    const uint32_t n0 = P.node_off[s], nn = P.node_off[s + 1] - n0;
    const std::vector<double> t(v, v + 3 * (size_t)nn);
    for (uint32_t a = 0; a < nn; ++a)
        for (int d = 0; d < 3; ++d) v[3 * (size_t)a + d] = t[3 * (size_t)(P.ninv[n0 + a] - n0) + d];
}

int copy_sub(femgpu_problem &H, uint32_t s, const double *src, double *dst_host, size_t n) {
    Problem &P = H.P;
    if (s >= P.n_sub || !dst_host || n != 3 * (size_t)(P.node_off[s + 1] - P.node_off[s]))
        return fail(FEMGPU_BAD_INPUT, "get: size");
    HIPCHK(hipStreamSynchronize(P.stream));
    HIPCHK(hipMemcpy(dst_host, src + 3 * (size_t)P.node_off[s], n * 8, hipMemcpyDeviceToHost));
    to_api_order(P, s, dst_host);
    return FEMGPU_OK;
}

// Formulation::solve's loop (formulation.h:1517-1568) and the non-constant branch of
// ProblemAncestor::iteration (problem_ancestor.h:85-106): solve_system, then stop when no
// non-linear criterium is set or when the change since the previous iterate (zero at first) is at
// most |criterium| -- max norm for solve(), 2-norm for iteration() -- else reassemble and solve
// again, at most max_non_linear_iteration + 1 times.  solve() assembles before every solve;
// iteration() has assembled already and reassembles after.
int nonlinear_solve(femgpu_problem &H, const femgpu_solve_opts *o, femgpu_solve_stats *st, bool norm2,
                    bool assemble_first) {
    Problem &P = H.P;
    const size_t n = 3 * (size_t)P.n_nodes;
    H.last_nl_solves = 0;
    bool have_old = false;
    for (int nb = 0;; ) {
        int rc = FEMGPU_OK;
        if (assemble_first) rc = assemble(H, 1, 1);
        if (!rc) rc = femgpu_solve_system(&H, o, st);
        if (rc) return rc;
        ++H.last_nl_solves;
        if (H.nl_crit == 0.0) return FEMGPU_OK;
        if (!have_old) {
            if (!H.d_old) HIPCHK(dalloc(&H.d_old, n));
            HIPCHK(launch_fill(H.d_old, n, 0.0, P.stream));
            have_old = true;
        }
        double d2 = 0, dinf = 0;
        if (diff_norm(P, H.d_old, P.d_u, n, &d2, &dinf)) return fail(FEMGPU_HIP_ERROR, "solve: difference norm");
        if ((norm2 ? d2 : dinf) <= std::fabs(H.nl_crit)) return FEMGPU_OK;
        if (nb++ >= H.max_nl) return fail(FEMGPU_MAX_ITER, "solve: max_non_linear_iteration reached");
        HIPCHK(hipMemcpyAsync(H.d_old, P.d_u, n * 8, hipMemcpyDeviceToDevice, P.stream));
        if (!assemble_first) {
            rc = assemble(H, H.assume_constant_matrix ? 0 : 1, 1);
            if (rc) return rc;
        }
    }
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
        case FEMGPU_STAGNATED: return "residual stagnated above the tolerance";
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
    if (!m || !out) return fail(FEMGPU_BAD_INPUT, "create: null");
    *out = nullptr;
    auto *h = new (std::nothrow) femgpu_problem();
    if (!h) return fail(FEMGPU_BAD_INPUT, "create: out of host memory");
    const int rc = build_problem(*h, 1, m);
    if (rc) {
        free_problem(h->P);
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
    for (uint32_t r = 0; r < n_brows; ++r)
        if (row_ptr[r + 1] < row_ptr[r]) return fail(FEMGPU_BAD_INPUT, "create_from_bsr: row_ptr not monotone");
    pat.col.assign(col, col + row_ptr[n_brows]);
    for (uint32_t r = 0; r < n_brows; ++r) {
        bool has_diag = false;
        for (uint32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
            if (col[p] >= n_brows || (p > row_ptr[r] && col[p] <= col[p - 1]))
                return fail(FEMGPU_BAD_INPUT, "create_from_bsr: columns out of range or unsorted");
            has_diag |= col[p] == r;
        }
        if (!has_diag) return fail(FEMGPU_BAD_INPUT, "create_from_bsr: missing diagonal block");
    }
    for (size_t k = 0; k < 9 * (size_t)row_ptr[n_brows]; ++k)
        if (!std::isfinite(vals[k])) return fail(FEMGPU_NAN, "create_from_bsr: non-finite value");
    finish_pattern(pat);
    auto *h = new (std::nothrow) femgpu_problem();
    if (!h) return fail(FEMGPU_BAD_INPUT, "create_from_bsr: out of host memory");
    Problem &P = h->P;
    P.from_bsr = true;
    P.n_sub = 1;
    P.n_nodes = n_brows;
    P.node_off = {0, n_brows};
    P.elem_off = {0, 0};
    P.pat = std::move(pat);
    P.nnzb = P.pat.col.size();
    (void)hipGetDevice(&P.device);
    int rc = alloc_system(P);
    if (rc == FEMGPU_OK) {
        std::vector<double> planes(9 * P.ps, 0.0);
        for (uint32_t r = 0; r < n_brows; ++r)
            for (uint32_t p = P.pat.row_ptr[r]; p < P.pat.row_ptr[r + 1]; ++p) {
                const uint64_t q = P.pat.pos_of_slot(r, p);
                for (int c = 0; c < 9; ++c) planes[c * P.ps + q] = vals[9 * (size_t)p + c];
            }
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
    for (double *&v : h->d_hist) dfree(v);
    dfree(h->d_old);
    delete h;
    return FEMGPU_OK;
}

int femgpu_set_material(femgpu_problem *h, const femgpu_material *m) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "set_material: null");
    return set_material(*h, 0, m);
}

int femgpu_set_elastic_modulus_device(femgpu_problem *h, const double *d_E) {
    if (!h || !d_E || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "set_elastic_modulus_device: bad input");
    Problem &P = h->P;
    HIPCHK(hipMemcpyAsync(P.d_E, d_E, (size_t)P.n_elem * 8, hipMemcpyDeviceToDevice, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    std::fill(P.material_set.begin(), P.material_set.end(), 1);
    return FEMGPU_OK;
}

int femgpu_set_traction(femgpu_problem *h, uint32_t n_faces, const uint32_t *faces, const double *trac) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "set_traction: null");
    return set_traction(*h, 0, n_faces, faces, trac);
}

int femgpu_add_constraint(femgpu_problem *h, uint32_t dof, double value, double penalty) {
    return femgpu_add_constraints(h, 1, &dof, &value, penalty);
}

int femgpu_add_constraints(femgpu_problem *h, uint32_t n, const uint32_t *dofs, const double *values, double penalty) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "add_constraints: null");
    return add_constraints(*h, 0, n, dofs, values, penalty);
}

int femgpu_constraint_parse(const char *txt, double t, uint32_t cap, uint64_t *dof, double *coeff, uint32_t *n_terms,
                            double *b) {
    if (!txt || !n_terms || !b || (cap && (!dof || !coeff))) return fail(FEMGPU_BAD_INPUT, "constraint_parse: null");
    ConstraintExpr ex;
    std::string err = ex.parse(txt);
    std::map<uint64_t, double> co;
    if (err.empty()) err = ex.eval(t, co, *b);
    if (!err.empty()) return fail(FEMGPU_BAD_INPUT, std::string("constraint '") + txt + "': " + err);
    *n_terms = (uint32_t)co.size();
    uint32_t k = 0;
    for (const auto &kv : co) {
        if (k >= cap) break;
        dof[k] = kv.first;
        coeff[k] = kv.second;
        ++k;
    }
    return FEMGPU_OK;
}

int femgpu_add_constraint_text(femgpu_problem *h, const char *txt, double penalty, uint32_t *number) {
    if (!h || !txt) return fail(FEMGPU_BAD_INPUT, "add_constraint: null");
    Problem &P = h->P;
    if (P.from_bsr) return fail(FEMGPU_BAD_INPUT, "add_constraint: problem built from a matrix");
    if (!(penalty > 0) || !std::isfinite(penalty))
        return fail(FEMGPU_BAD_INPUT, "add_constraint: only the penalty form (penalty > 0) keeps K SPD");
    femgpu_problem::TextConstraint tc;
    std::string err = tc.expr.parse(txt);
    if (!err.empty()) return fail(FEMGPU_BAD_INPUT, std::string("add_constraint '") + txt + "': " + err);
    std::map<uint64_t, double> co;
    double b = 0;
    err = tc.expr.eval(h->time, co, b);
    if (!err.empty()) return fail(FEMGPU_BAD_INPUT, std::string("add_constraint '") + txt + "': " + err);
    if (co.size() != 1)
        return fail(FEMGPU_BAD_INPUT, std::string("add_constraint '") + txt + "': " +
                                          (co.empty() ? "no unknown"
                                                      : "couples several unknowns; its penalty terms between DOFs lie "
                                                        "outside the node-adjacency pattern and the matrix-free product"));
    const uint64_t dof = co.begin()->first;
    if (dof >= 3 * (uint64_t)(P.node_off[1] - P.node_off[0]))
        return fail(FEMGPU_BAD_INPUT, std::string("add_constraint '") + txt + "': node beyond the mesh");
    tc.dof = dof;
    tc.penalty = penalty;
    double v = 0, pen = 0;
    const int rc = eval_text_constraint(tc, h->time, v, pen);
    if (rc) return rc;
    const uint32_t g = (uint32_t)dof;
    tc.slot = P.cons.size();
    P.cons.push_back({P.ninv.empty() ? g : 3 * P.ninv[g / 3] + g % 3, v, pen});
    P.cons_dirty = true;
    h->text_cons.push_back(tc);
    if (number) *number = (uint32_t)(P.cons.size() - 1);
    return FEMGPU_OK;
}

int femgpu_clear_constraints(femgpu_problem *h) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "clear_constraints: null");
    h->P.cons.clear();
    h->text_cons.clear();
    h->P.cons_dirty = true;
    return FEMGPU_OK;
}

int femgpu_assemble(femgpu_problem *h, int do_mat, int do_vec) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "assemble: null");
    return assemble(*h, do_mat, do_vec);
}

int femgpu_solve_system(femgpu_problem *h, const femgpu_solve_opts *o, femgpu_solve_stats *st) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "solve_system: null");
    unpin_precond(*h, o);
    femgpu_solve_stats tmp{};
    const int rc = solve_system(*h, o, st ? st : &tmp, nullptr, nullptr);
    if (st) st->status = rc;
    return rc;
}

int femgpu_solve(femgpu_problem *h, const femgpu_solve_opts *o, femgpu_solve_stats *st) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "solve: null");
    int rc = femgpu_shift(h, 1);
    if (rc) return rc;
    return nonlinear_solve(*h, o, st, false, true);
}

int femgpu_get_result(femgpu_problem *h, double *u, size_t n) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_result: null");
    return copy_sub(*h, 0, h->P.d_u, u, n);
}

int femgpu_get_result_device(femgpu_problem *h, double *d_u) {
    if (!h || !d_u) return fail(FEMGPU_BAD_INPUT, "get_result_device: null");
    if (h->P.d_ninv) HIPCHK(launch_permute_nodes(d_u, h->P.d_u, h->P.d_ninv, h->P.n_nodes, h->P.stream));
    else HIPCHK(hipMemcpyAsync(d_u, h->P.d_u, 3 * (size_t)h->P.n_nodes * 8, hipMemcpyDeviceToDevice, h->P.stream));
    HIPCHK(hipStreamSynchronize(h->P.stream));
    return FEMGPU_OK;
}

int femgpu_set_result(femgpu_problem *h, const double *u, size_t n) {
    if (!h || !u || n != 3 * (size_t)h->P.n_nodes) return fail(FEMGPU_BAD_INPUT, "set_result: size");
    const Problem &P = h->P;
    std::vector<double> t;
    if (!P.nperm.empty()) {
        t.resize(n);
        for (uint32_t i = 0; i < P.n_nodes; ++i)
            for (int d = 0; d < 3; ++d) t[3 * (size_t)i + d] = u[3 * (size_t)P.nperm[i] + d];
        u = t.data();
    }
    HIPCHK(hipStreamSynchronize(P.stream));
    HIPCHK(hipMemcpy(P.d_u, u, n * 8, hipMemcpyHostToDevice));
    return FEMGPU_OK;
}

int femgpu_get_sollicitation(femgpu_problem *h, double *F, size_t n) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_sollicitation: null");
    return copy_sub(*h, 0, h->P.d_F, F, n);
}

int femgpu_set_sollicitation(femgpu_problem *h, const double *F, size_t n) {
    if (!h || !F || n != 3 * (size_t)h->P.n_nodes) return fail(FEMGPU_BAD_INPUT, "set_sollicitation: size");
    Problem &P = h->P;
    std::vector<double> t;
    if (!P.nperm.empty()) {  // to the internal node order
        t.resize(n);
        for (uint32_t i = 0; i < P.n_nodes; ++i)
            for (int d = 0; d < 3; ++d) t[3 * (size_t)i + d] = F[3 * (size_t)P.nperm[i] + d];
        F = t.data();
    }
    HIPCHK(hipStreamSynchronize(P.stream));
    HIPCHK(hipMemcpy(P.d_F, F, n * 8, hipMemcpyHostToDevice));
    P.assembled_vec = true;
    return FEMGPU_OK;
}

int femgpu_get_matrix(femgpu_problem *h, uint32_t *row_ptr, uint32_t *col, double *vals) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_matrix: null");
    Problem &P = h->P;
    const Pattern &pat = P.pat;
    std::vector<double> planes;
    if (vals) {
        planes.resize(9 * P.ps);
        HIPCHK(hipStreamSynchronize(P.stream));
        HIPCHK(hipMemcpy(planes.data(), P.d_K, planes.size() * 8, hipMemcpyDeviceToHost));
    }
    // caller's row a = internal row ninv[a]; its blocks re-sorted by the caller's column
    std::vector<std::pair<uint32_t, uint64_t>> row;
    uint32_t p_out = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (uint32_t a = 0; a < P.n_nodes; ++a) {
        const uint32_t r = P.ninv.empty() ? a : P.ninv[a];
        row.clear();
        for (uint32_t p = pat.row_ptr[r]; p < pat.row_ptr[r + 1]; ++p)
            row.push_back({P.nperm.empty() ? pat.col[p] : P.nperm[pat.col[p]], pat.pos_of_slot(r, p)});
        if (!P.nperm.empty()) std::sort(row.begin(), row.end());
        for (const auto &e : row) {
            if (col) col[p_out] = e.first;
            if (vals)
                for (int c = 0; c < 9; ++c) vals[9 * (size_t)p_out + c] = planes[c * P.ps + e.second];
            ++p_out;
        }
        if (row_ptr) row_ptr[a + 1] = p_out;
    }
    return FEMGPU_OK;
}

int femgpu_get_ic0_factor(femgpu_problem *h, int precond, uint32_t *order, uint32_t *row_ptr, uint32_t *col, double *L,
                          uint32_t *n_levels) {
    if (!h || (precond != FEMGPU_BLOCK_IC0 && precond != FEMGPU_BLOCK_IC0_COLOURED))
        return fail(FEMGPU_BAD_INPUT, "get_ic0_factor: bad input");
    Problem &P = h->P;
    if (!P.assembled_mat) return fail(FEMGPU_NOT_READY, "get_ic0_factor: assemble first");
    const int rc = ic0_ensure(P, precond == FEMGPU_BLOCK_IC0 ? IC0_NATURAL : IC0_COLOURED, false);
    if (rc == 3) return fail(FEMGPU_NOT_SPD, "get_ic0_factor: non-positive pivot");
    if (rc) return fail(rc == 1 ? FEMGPU_BAD_INPUT : FEMGPU_HIP_ERROR,
                        rc == 1 ? "get_ic0_factor: pattern not symmetric" : "get_ic0_factor: factorisation failed");
    if (n_levels) *n_levels = P.ic0.sched.n_levels();
    if (ic0_download(P, order, row_ptr, col, L)) return fail(FEMGPU_HIP_ERROR, "get_ic0_factor: download");
    if (order && !P.nperm.empty())
        for (uint32_t t = 0; t < P.n_nodes; ++t) order[t] = P.nperm[order[t]];
    return FEMGPU_OK;
}

int femgpu_ic0_schedule(uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col, int precond, uint32_t *order,
                        uint32_t *level_ptr, uint32_t *n_levels) {
    if (!n_brows || !row_ptr || !col || !order || !n_levels ||
        (precond != FEMGPU_BLOCK_IC0 && precond != FEMGPU_BLOCK_IC0_COLOURED))
        return fail(FEMGPU_BAD_INPUT, "ic0_schedule: bad input");
    Pattern pat;
    pat.n_rows = n_brows;
    pat.row_ptr.assign(row_ptr, row_ptr + n_brows + 1);
    pat.col.assign(col, col + row_ptr[n_brows]);
    for (uint32_t r = 0; r < n_brows; ++r) {
        bool has_diag = false;
        for (uint32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
            if (col[p] >= n_brows || (p > row_ptr[r] && col[p] <= col[p - 1]))
                return fail(FEMGPU_BAD_INPUT, "ic0_schedule: columns out of range or unsorted");
            has_diag |= col[p] == r;
        }
        if (!has_diag) return fail(FEMGPU_BAD_INPUT, "ic0_schedule: missing diagonal block");
    }
    finish_pattern(pat);
    Ic0Schedule S;
    if (!build_ic0_schedule(pat, precond == FEMGPU_BLOCK_IC0 ? IC0_NATURAL : IC0_COLOURED, S))
        return fail(FEMGPU_BAD_INPUT, "ic0_schedule: pattern not symmetric");
    std::memcpy(order, S.order.data(), 4 * (size_t)n_brows);
    if (level_ptr) {
        if (*n_levels < S.n_levels()) return fail(FEMGPU_BAD_INPUT, "ic0_schedule: level_ptr too small");
        std::memcpy(level_ptr, S.lv_ptr.data(), 4 * S.lv_ptr.size());
    }
    *n_levels = S.n_levels();
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
    *md = h->P.max_diag.empty() ? 0.0 : h->P.max_diag[0];
    return FEMGPU_OK;
}

int femgpu_after_solve_device(femgpu_problem *h, double *d_eps, double *d_sigma, double *d_tr) {
    if (!h || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "after_solve: bad input");
    Problem &P = h->P;
    for (uint32_t s = 0; s < P.n_sub; ++s)
        if (!P.material_set[s]) return fail(FEMGPU_NOT_READY, "after_solve: material not set");
    HIPCHK(hipEventRecord(P.ev[2], P.stream));
    HIPCHK(launch_after_solve(P, P.d_u, d_eps, d_sigma, d_tr));
    HIPCHK(hipEventRecord(P.ev[3], P.stream));
    HIPCHK(hipEventSynchronize(P.ev[3]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, P.ev[2], P.ev[3]));
    P.last_after_solve_ms = ms;
    return FEMGPU_OK;
}

int femgpu_get_after_solve_ms(femgpu_problem *h, double *ms) {
    if (!h || !ms) return fail(FEMGPU_BAD_INPUT, "get_after_solve_ms: null");
    *ms = h->P.last_after_solve_ms;
    return FEMGPU_OK;
}

// Recovery over every element, then the rows of elements [e0, e0 + ne) copied to the host.
static int after_solve_host(femgpu_problem *h, uint32_t e0, uint32_t ne, double *eps, double *sig, double *tr) {
    Problem &P = h->P;
    double *d = nullptr;
    HIPCHK(dalloc(&d, 13 * (size_t)P.n_elem));
    double *d_eps = d, *d_sig = d + 6 * (size_t)P.n_elem, *d_tr = d + 12 * (size_t)P.n_elem;
    int rc = femgpu_after_solve_device(h, d_eps, d_sig, d_tr);
    hipError_t e = hipSuccess;
    if (!rc && eps) e = hipMemcpy(eps, d_eps + 6 * (size_t)e0, 48 * (size_t)ne, hipMemcpyDeviceToHost);
    if (!rc && sig && e == hipSuccess) e = hipMemcpy(sig, d_sig + 6 * (size_t)e0, 48 * (size_t)ne, hipMemcpyDeviceToHost);
    if (!rc && tr && e == hipSuccess) e = hipMemcpy(tr, d_tr + e0, 8 * (size_t)ne, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (rc) return rc;
    if (e != hipSuccess) return fail(FEMGPU_HIP_ERROR, hipGetErrorString(e));
    return FEMGPU_OK;
}

int femgpu_after_solve(femgpu_problem *h, double *eps, double *sigma, double *tr) {
    if (!h || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "after_solve: bad input");
    return after_solve_host(h, 0, h->P.n_elem, eps, sigma, tr);
}

int femgpu_batch_after_solve(femgpu_batch *b, uint32_t i, double *eps, double *sigma, double *tr) {
    if (!b || i >= b->H.P.n_sub) return fail(FEMGPU_BAD_INPUT, "batch_after_solve: bad input");
    const Problem &P = b->H.P;
    return after_solve_host(&b->H, P.elem_off[i], P.elem_off[i + 1] - P.elem_off[i], eps, sigma, tr);
}

int femgpu_write_vtu(const char *path, int elem_kind, uint32_t n_nodes, const double *coords, uint32_t n_elem,
                     const uint32_t *conn, uint32_t n_point, const femgpu_vtu_field *point, uint32_t n_cell,
                     const femgpu_vtu_field *cell, const uint32_t *group, uint32_t n_global, const femgpu_vtu_field *global) {
    int npe = 0;
    if (!path || !vtk_cell_type(elem_kind, &npe) || (n_nodes && !coords) || (n_elem && !conn) ||
        (n_point && !point) || (n_cell && !cell) || (n_global && !global))
        return fail(FEMGPU_BAD_INPUT, "write_vtu: bad input");
    for (size_t k = 0; k < (size_t)n_elem * npe; ++k)
        if (conn[k] >= n_nodes) return fail(FEMGPU_BAD_INPUT, "write_vtu: node index out of range");
    auto fields = [](uint32_t n, const femgpu_vtu_field *f, size_t items, std::vector<VtuField> &out) {
        for (uint32_t i = 0; i < n; ++i) {
            if (!f[i].name || (f[i].n_comp && items && !f[i].data)) return false;
            out.push_back(VtuField{f[i].name, f[i].n_comp, f[i].data});
        }
        return true;
    };
    std::vector<VtuField> pf, cf, gf;
    if (!fields(n_point, point, n_nodes, pf) || !fields(n_cell, cell, n_elem, cf) || !fields(n_global, global, 1, gf))
        return fail(FEMGPU_BAD_INPUT, "write_vtu: field without name or data");
    if (!write_vtu(path, elem_kind, n_nodes, coords, n_elem, conn, pf, cf, group, gf))
        return fail(FEMGPU_BAD_INPUT, std::string("write_vtu: cannot write ") + path);
    return FEMGPU_OK;
}

int femgpu_write_avs(const char *path, int elem_kind, uint32_t n_nodes, const double *coords, uint32_t n_elem,
                     const uint32_t *conn, uint32_t n_point, const femgpu_avs_field *point, uint32_t n_cell,
                     const femgpu_avs_field *cell, uint32_t n_global, const femgpu_avs_field *global, uint32_t n_display,
                     const char *const *display) {
    const int npe = nodes_per_elem(elem_kind);
    if (!path || !npe || (n_nodes && !coords) || (n_elem && !conn) || (n_point && !point) || (n_cell && !cell) ||
        (n_global && !global) || (n_display && !display))
        return fail(FEMGPU_BAD_INPUT, "write_avs: bad input");
    for (size_t k = 0; k < (size_t)n_elem * npe; ++k)
        if (conn[k] >= n_nodes) return fail(FEMGPU_BAD_INPUT, "write_avs: node index out of range");
    auto fields = [](uint32_t n, const femgpu_avs_field *f, size_t items, std::vector<AvsField> &out) {
        for (uint32_t i = 0; i < n; ++i) {
            if (!f[i].name || !f[i].n_comp || f[i].nest > 2 || (f[i].nest == 0 && f[i].n_comp != 1) ||
                (items && !f[i].data))
                return false;
            out.push_back(AvsField{f[i].name, f[i].n_comp, f[i].nest, f[i].data});
        }
        return true;
    };
    std::vector<AvsField> pf, cf, gf;
    if (!fields(n_point, point, n_nodes, pf) || !fields(n_cell, cell, n_elem, cf) || !fields(n_global, global, 1, gf))
        return fail(FEMGPU_BAD_INPUT, "write_avs: a field needs a name, data, n_comp >= 1 and nest 0 (n_comp 1), 1 or 2");
    std::vector<std::string> disp;
    for (uint32_t i = 0; i < n_display; ++i) {
        if (!display[i]) return fail(FEMGPU_BAD_INPUT, "write_avs: null display name");
        disp.emplace_back(display[i]);
    }
    if (!write_avs(path, elem_kind, n_nodes, coords, n_elem, conn, pf, cf, gf, disp))
        return fail(FEMGPU_BAD_INPUT, std::string("write_avs: cannot write ") + path);
    return FEMGPU_OK;
}

namespace {
// Sub-problem s of a mesh problem, gathered for a result file in the API's node order.
struct ResultData {
    uint32_t nn = 0, ne = 0;
    std::vector<double> coords, u, E, eps, sig, tr;
    std::vector<uint32_t> conn;
    double nu = 0, bf[3] = {0, 0, 0};
};

int gather_result(femgpu_problem &H, uint32_t s, const char *path, ResultData &R) {
    Problem &P = H.P;
    if (!path || P.from_bsr || s >= P.n_sub) return fail(FEMGPU_BAD_INPUT, "write_result: bad input");
    const uint32_t n0 = P.node_off[s], e0 = P.elem_off[s];
    R.nn = P.node_off[s + 1] - n0;
    R.ne = P.elem_off[s + 1] - e0;
    R.coords.resize(3 * (size_t)R.nn);
    R.u.resize(3 * (size_t)R.nn);
    R.E.resize(R.ne);
    R.eps.resize(6 * (size_t)R.ne);
    R.sig.resize(6 * (size_t)R.ne);
    R.tr.resize(R.ne);
    HIPCHK(hipStreamSynchronize(P.stream));
    HIPCHK(hipMemcpy(R.coords.data(), P.d_coords + 3 * (size_t)n0, R.coords.size() * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(R.u.data(), P.d_u + 3 * (size_t)n0, R.u.size() * 8, hipMemcpyDeviceToHost));
    to_api_order(P, s, R.coords.data());
    to_api_order(P, s, R.u.data());
    HIPCHK(hipMemcpy(R.E.data(), P.d_E + e0, R.E.size() * 8, hipMemcpyDeviceToHost));
    int rc = after_solve_host(&H, e0, R.ne, R.eps.data(), R.sig.data(), R.tr.data());
    if (rc) return rc;
    R.conn.assign(P.conn_h.begin() + (size_t)e0 * P.npe, P.conn_h.begin() + (size_t)(e0 + R.ne) * P.npe);
    for (uint32_t &v : R.conn) v = (P.nperm.empty() ? v : P.nperm[v]) - n0;
    const SubMaterial &m = P.mat_h[s];
    R.nu = m.nu;
    R.bf[0] = m.bf0;
    R.bf[1] = m.bf1;
    R.bf[2] = m.bf2;
    return FEMGPU_OK;
}

// Sub-problem s as a VTU file: dep; elastic_modulus, epsilon, sigma, tr_epsilon (the after-solve
// recovery of the current result); group = s; poisson_ratio.
int write_result(femgpu_problem &H, uint32_t s, const char *path) {
    ResultData R;
    int rc = gather_result(H, s, path, R);
    if (rc) return rc;
    const std::vector<uint32_t> group(R.ne, s);
    const femgpu_vtu_field pf[] = {{"dep", 3, R.u.data()}};
    const femgpu_vtu_field cf[] = {{"elastic_modulus", 1, R.E.data()}, {"epsilon", 6, R.eps.data()},
                                   {"sigma", 6, R.sig.data()}, {"tr_epsilon", 1, R.tr.data()}};
    const femgpu_vtu_field gf[] = {{"poisson_ratio", 1, &R.nu}};
    return femgpu_write_vtu(path, H.P.kind, R.nn, R.coords.data(), R.ne, R.conn.data(), 1, pf, 4, cf, group.data(), 1,
                            gf);
}

// Sub-problem s as an AVS file, with the shapes of the reference's elasticity mesh data: dep, pos
// (Vec<T,3>); elastic_modulus (a scalar per element), epsilon, sigma (Vec<Vec<T,6>,1>),
// tr_epsilon (Vec<T,1>); poisson_ratio (scalar), f_vol (Vec<T,3>).
int write_result_avs(femgpu_problem &H, uint32_t s, const char *path, uint32_t n_display, const char *const *display) {
    ResultData R;
    int rc = gather_result(H, s, path, R);
    if (rc) return rc;
    const femgpu_avs_field pf[] = {{"dep", 3, 1, R.u.data()}, {"pos", 3, 1, R.coords.data()}};
    const femgpu_avs_field cf[] = {{"elastic_modulus", 1, 0, R.E.data()}, {"epsilon", 6, 2, R.eps.data()},
                                   {"sigma", 6, 2, R.sig.data()}, {"tr_epsilon", 1, 1, R.tr.data()}};
    const femgpu_avs_field gf[] = {{"poisson_ratio", 1, 0, &R.nu}, {"f_vol", 3, 1, R.bf}};
    static const char *const all[] = {"all"};
    if (!n_display || !display) {
        n_display = 1;
        display = all;
    }
    return femgpu_write_avs(path, H.P.kind, R.nn, R.coords.data(), R.ne, R.conn.data(), 2, pf, 4, cf, 2, gf, n_display,
                            display);
}
}  // namespace

int femgpu_write_result_vtu(femgpu_problem *h, const char *path) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "write_result_vtu: null");
    return write_result(*h, 0, path);
}

int femgpu_batch_write_result_vtu(femgpu_batch *b, uint32_t i, const char *path) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_write_result_vtu: null");
    return write_result(b->H, i, path);
}

int femgpu_write_result_avs(femgpu_problem *h, const char *path, uint32_t n_display, const char *const *display) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "write_result_avs: null");
    return write_result_avs(*h, 0, path, n_display, display);
}

int femgpu_batch_write_result_avs(femgpu_batch *b, uint32_t i, const char *path, uint32_t n_display,
                                  const char *const *display) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_write_result_avs: null");
    return write_result_avs(b->H, i, path, n_display, display);
}

int femgpu_elem_grad(femgpu_problem *h, const double *u, const double *lam, double *g) {
    if (!h || !u || !lam || !g || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "elem_grad: bad input");
    Problem &P = h->P;
    const size_t n = 3 * (size_t)P.n_nodes;
    std::vector<double> ui, li;
    if (!P.nperm.empty()) {  // to the internal node order
        ui.resize(n);
        li.resize(n);
        for (uint32_t i = 0; i < P.n_nodes; ++i)
            for (int d = 0; d < 3; ++d) {
                ui[3 * (size_t)i + d] = u[3 * (size_t)P.nperm[i] + d];
                li[3 * (size_t)i + d] = lam[3 * (size_t)P.nperm[i] + d];
            }
        u = ui.data();
        lam = li.data();
    }
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
    Problem &P = h->P;
    if (!P.assembled_mat) return fail(FEMGPU_NOT_READY, "adjoint: assemble first");
    const size_t n = 3 * (size_t)P.n_nodes;
    // lambda in d_aux: K lambda = dJ/du (K symmetric), zero initial guess
    double *d_rhs = nullptr;
    HIPCHK(dalloc(&d_rhs, n));
    hipError_t e = P.d_nperm ? launch_permute_nodes(d_rhs, d_dJdu, P.d_nperm, P.n_nodes, P.stream)
                             : hipMemcpyAsync(d_rhs, d_dJdu, n * 8, hipMemcpyDeviceToDevice, P.stream);
    if (e == hipSuccess) e = launch_fill(P.d_aux, n, 0.0, P.stream);
    if (e != hipSuccess) { (void)hipFree(d_rhs); return fail(FEMGPU_HIP_ERROR, hipGetErrorString(e)); }
    femgpu_solve_stats tmp{};
    unpin_precond(*h, o);
    int rc = solve_system(*h, o, st ? st : &tmp, d_rhs, P.d_aux);
    if (rc == FEMGPU_OK) {
        e = launch_elem_grad(P, P.d_u, P.d_aux, d_dJdE);
        if (e == hipSuccess) e = hipStreamSynchronize(P.stream);
        if (e != hipSuccess) rc = fail(FEMGPU_HIP_ERROR, hipGetErrorString(e));
    }
    (void)hipFree(d_rhs);
    if (st) st->status = rc;
    return rc;
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

int femgpu_get_precond_setup_ms(femgpu_problem *h, double *ms) {
    if (!h || !ms) return fail(FEMGPU_BAD_INPUT, "get_precond_setup_ms: null");
    *ms = h->P.last_precond_ms;
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

int femgpu_apply(femgpu_problem *h, int precision, const double *x, double *y, size_t n) {
    if (!h || !x || !y || precision < 0 || precision > 1 || n != 3 * (size_t)h->P.n_nodes)
        return fail(FEMGPU_BAD_INPUT, "apply: bad input");
    Problem &P = h->P;
    if (!P.assembled_mat) return fail(FEMGPU_NOT_READY, "apply: matrix not assembled");
    std::vector<double> t;
    if (!P.nperm.empty()) {
        t.resize(n);
        for (uint32_t i = 0; i < P.n_nodes; ++i)
            for (int d = 0; d < 3; ++d) t[3 * (size_t)i + d] = x[3 * (size_t)P.nperm[i] + d];
        x = t.data();
    }
    HIPCHK(hipStreamSynchronize(P.stream));
    HIPCHK(hipMemcpy(P.d_aux, x, n * 8, hipMemcpyHostToDevice));
    const int rc = apply_spmv(P, precision, P.d_aux, P.d_q);
    if (rc) return fail(rc == 1 ? FEMGPU_BAD_INPUT : FEMGPU_HIP_ERROR, "apply: product variant");
    HIPCHK(hipStreamSynchronize(P.stream));
    HIPCHK(hipMemcpy(y, P.d_q, n * 8, hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < P.n_sub; ++s) to_api_order(P, s, y + 3 * (size_t)P.node_off[s]);
    return FEMGPU_OK;
}

int femgpu_time_matrix_free(femgpu_problem *h, int precision, int reps, double *block_ms, double *product_ms) {
    if (!h || reps < 1 || precision < 0 || precision > 1) return fail(FEMGPU_BAD_INPUT, "time_matrix_free: bad input");
    const int rc = mf_time(h->P, precision, reps, block_ms, product_ms);
    if (rc) return fail(rc == 1 ? FEMGPU_BAD_INPUT : FEMGPU_HIP_ERROR, "time_matrix_free: not a Hexa problem with one "
                                                                        "sub-problem and an assembled matrix");
    return FEMGPU_OK;
}

int femgpu_time_iteration(femgpu_problem *h, int precision, int iters, double *block_ms, double *update_ms) {
    if (!h || iters < 1 || precision < 0 || precision > 1) return fail(FEMGPU_BAD_INPUT, "time_iteration: bad input");
    Problem &P = h->P;
    if (P.n_sub != 1) return fail(FEMGPU_BAD_INPUT, "time_iteration: a single problem only");
    if (!P.assembled_mat || !P.assembled_vec) return fail(FEMGPU_NOT_READY, "time_iteration: assemble first");
    int rc = mf_prepare(P);
    if (rc) return fail(rc == 1 ? FEMGPU_BAD_INPUT : FEMGPU_HIP_ERROR, "time_iteration: no matrix-free product");
    femgpu_solve_opts bj;
    femgpu_default_solve_opts(&bj);
    bj.precond = FEMGPU_BLOCK_JACOBI;
    unpin_precond(*h, &bj);
    rc = time_iteration(P, precision, iters, block_ms, update_ms);
    if (rc == 1) return fail(FEMGPU_BAD_INPUT, "time_iteration: the iteration of that precision is not the two-kernel one "
                                               "(matrix-free product, one sub-problem)");
    if (rc) return fail(FEMGPU_HIP_ERROR, "time_iteration: HIP error");
    return FEMGPU_OK;
}

int femgpu_matrix_free_info(femgpu_problem *h, uint64_t info[7]) {
    if (!h || !info) return fail(FEMGPU_BAD_INPUT, "matrix_free_info: null");
    const int rc = mf_build(h->P);
    if (rc) return fail(rc == 1 ? FEMGPU_BAD_INPUT : FEMGPU_HIP_ERROR, "matrix_free_info: not a Hexa problem with one "
                                                                        "sub-problem");
    const MatFree &M = h->P.mf;
    uint32_t n_local = 0;
    HIPCHK(hipMemcpy(&n_local, M.d_lptr + M.n_blocks, sizeof(uint32_t), hipMemcpyDeviceToHost));
    info[0] = M.n_blocks;
    info[1] = M.n_colors;
    info[2] = n_local;
    info[3] = M.n_bnodes;
    info[4] = M.n_halo;
    info[5] = M.n_iso;
    info[6] = M.n_boxes;
    return FEMGPU_OK;
}

// Calibration kernels on a fresh buffer of n doubles: plain copy (16 n bytes moved per launch) or
// streaming read (8 n bytes).
static int stream_bench(size_t n, int reps, double *avg_ms, bool read_only) {
    if (!avg_ms || reps < 1 || n < 2 || (n & 1)) return fail(FEMGPU_BAD_INPUT, "stream bench: bad input");
    double *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FEMGPU_OK;
    const size_t nb = read_only ? 4 * (n / 2048 + 1) : n;
    auto run = [&]() { return read_only ? launch_read(a, n, b, s) : launch_copy(b, a, n, s); };
    if (dalloc(&a, n) != hipSuccess || dalloc(&b, nb) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(FEMGPU_HIP_ERROR, "stream bench: setup");
    if (!rc && (launch_fill(a, n, 1.0, s) != hipSuccess || run() != hipSuccess)) rc = fail(FEMGPU_HIP_ERROR, "stream bench: warmup");
    if (!rc) {
        (void)hipEventRecord(e0, s);
        for (int i = 0; i < reps; ++i) (void)run();
        (void)hipEventRecord(e1, s);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(FEMGPU_HIP_ERROR, "stream bench: timing");
        *avg_ms = (double)ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    dfree(a);
    dfree(b);
    return rc;
}

int femgpu_stream_copy(size_t n, int reps, double *avg_ms) { return stream_bench(n, reps, avg_ms, false); }
int femgpu_stream_read(size_t n, int reps, double *avg_ms) { return stream_bench(n, reps, avg_ms, true); }

}  // extern "C"

namespace {
// Element loop: colour passes (direct or staged), the gathers, or row groups of Rmax, Rmax/2,
// Rmax/4 rows (min of 3 timings); the fastest is kept and K and F re-assembled with it.
int tune_assembly(femgpu_problem *h) {
    Problem &P = h->P;
    const uint32_t Rmax = max_assembly_rows(P);
    uint32_t best_R = P.asm_rows;
    double best = 1e30;
    for (uint32_t R : {0u, ASM_STAGED, ASM_GATHER, ASM_GATHER_BOX, Rmax, Rmax / 2, Rmax / 4}) {
        if (R && R != ASM_STAGED && R != ASM_GATHER && R != ASM_GATHER_BOX && R < 8) continue;
        int rc = set_assembly_rows(P, R);
        // beyond the gather lists' index range, or not a mesh of box elements
        if (rc == FEMGPU_BAD_INPUT && (R == ASM_GATHER || R == ASM_GATHER_BOX)) continue;
        if (rc) return rc;
        double t = 1e30;
        for (int rep = 0; rep < 4; ++rep) {
            HIPCHK(hipEventRecord(P.ev[2], P.stream));
            HIPCHK(launch_assemble(P, 1, 1));
            HIPCHK(hipEventRecord(P.ev[3], P.stream));
            HIPCHK(hipEventSynchronize(P.ev[3]));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, P.ev[2], P.ev[3]));
            if (rep) t = std::min(t, (double)ms);
        }
        if (t < best) { best = t; best_R = R; }
    }
    int rc = set_assembly_rows(P, best_R);
    if (!rc) rc = assemble(*h, 1, P.assembled_vec ? 1 : 0);   // K and F as the chosen variant makes them
    return rc;
}
}  // namespace

extern "C" {

int femgpu_autotune(femgpu_problem *h, int *unroll) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "autotune: null");
    Problem &P = h->P;
    if (!P.assembled_mat) return fail(FEMGPU_NOT_READY, "autotune: assemble first");
    if (autotune_spmv(P)) return fail(FEMGPU_HIP_ERROR, "autotune failed");
    if (unroll) *unroll = P.spmv.unroll[0];
    if (P.from_bsr) return FEMGPU_OK;
    return tune_assembly(h);
}

int femgpu_batch_autotune(femgpu_batch *b, const femgpu_solve_opts *o, int *product) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_autotune: null");
    int rc = check_opts(o);
    if (rc) return rc;
    femgpu_problem &H = b->H;
    Problem &P = H.P;
    if (!P.assembled_mat) return fail(FEMGPU_NOT_READY, "batch_autotune: assemble first");
    // the caller's current results are kept
    const size_t n = 3 * (size_t)P.n_nodes;
    double *d_keep = nullptr;
    HIPCHK(dalloc(&d_keep, n));
    hipError_t e = hipMemcpyAsync(d_keep, P.d_u, n * 8, hipMemcpyDeviceToDevice, P.stream);
    // candidates: stored full matrix with non-temporal loads and 16-bit column offsets (when
    // every offset fits), with non-temporal loads, with plain loads; the matrix-free product
    struct Variant { int id, format, loads; };
    std::vector<Variant> vs;
    if (P.d_col16) vs.push_back({FEMGPU_BATCH_STORED_NT_COL16, FMT_FULL, LOAD_NT16});
    vs.push_back({FEMGPU_BATCH_STORED_NT, FMT_FULL, LOAD_NT});
    vs.push_back({FEMGPU_BATCH_STORED_PLAIN, FMT_FULL, LOAD_PLAIN});
    if (mf_eligible(P)) vs.push_back({FEMGPU_BATCH_MATRIX_FREE, FMT_MF, LOAD_NT});
    double best = 1e30;
    int best_i = -1;
    std::vector<femgpu_solve_stats> st(P.n_sub);
    for (size_t i = 0; i < vs.size() && e == hipSuccess && !rc; ++i) {
        const int fmt[2] = {vs[i].format, vs[i].format}, unroll[2] = {2, 2}, loads[2] = {vs[i].loads, vs[i].loads};
        if ((rc = femgpu_set_spmv_config(&H, fmt, unroll, loads))) break;
        double t = 1e30;
        for (int rep = 0; rep < 2 && !rc; ++rep) {  // the first solve also captures the chunk graph
            if ((e = launch_fill(P.d_u, n, 0.0, P.stream)) != hipSuccess) break;
            rc = solve_system(H, o, st.data(), nullptr, nullptr);
            if (rc == FEMGPU_MAX_ITER || rc == FEMGPU_STAGNATED) rc = FEMGPU_OK;  // timed all the same
            if (!rc && rep) t = st[0].solve_ms;
        }
        if (!rc && t < best) {
            best = t;
            best_i = (int)i;
        }
    }
    if (!rc && e == hipSuccess && best_i >= 0) {
        const int fmt[2] = {vs[best_i].format, vs[best_i].format}, unroll[2] = {2, 2},
                  loads[2] = {vs[best_i].loads, vs[best_i].loads};
        rc = femgpu_set_spmv_config(&H, fmt, unroll, loads);
        if (!rc && !P.from_bsr) rc = tune_assembly(&H);
        if (product) *product = vs[best_i].id;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(P.d_u, d_keep, n * 8, hipMemcpyDeviceToDevice, P.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(P.stream);
    (void)hipFree(d_keep);
    if (e != hipSuccess) return fail(FEMGPU_HIP_ERROR, hipGetErrorString(e));
    return rc;
}

int femgpu_get_chunk_stats(femgpu_problem *h, int *n_chunks, double *chunk_ms, int *n_gaps, double *gap_ms) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_chunk_stats: null");
    const Problem &P = h->P;
    if (n_chunks) *n_chunks = P.last_chunks;
    if (chunk_ms) *chunk_ms = P.last_chunk_ms;
    if (n_gaps) *n_gaps = P.last_gaps;
    if (gap_ms) *gap_ms = P.last_gap_ms;
    return FEMGPU_OK;
}

int femgpu_batch_get_chunk_stats(femgpu_batch *b, int *n_chunks, double *chunk_ms, int *n_gaps, double *gap_ms) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_get_chunk_stats: null");
    return femgpu_get_chunk_stats(&b->H, n_chunks, chunk_ms, n_gaps, gap_ms);
}

int femgpu_batch_set_spmv_config(femgpu_batch *b, const int format[2], const int unroll[2], const int loads[2]) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_set_spmv_config: null");
    return femgpu_set_spmv_config(&b->H, format, unroll, loads);
}

int femgpu_set_assembly_rows(femgpu_problem *h, uint32_t rows) {
    if (!h || h->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "set_assembly_rows: bad input");
    return set_assembly_rows(h->P, rows);
}

int femgpu_get_assembly_rows(femgpu_problem *h, uint32_t *rows, uint32_t *max_rows) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_assembly_rows: null");
    if (rows) *rows = h->P.asm_rows;
    if (max_rows) *max_rows = h->P.from_bsr ? 0 : max_assembly_rows(h->P);
    return FEMGPU_OK;
}

int femgpu_get_spmv_config(femgpu_problem *h, int format[2], int unroll[2], int loads[2]) {
    if (!h || !format || !unroll) return fail(FEMGPU_BAD_INPUT, "get_spmv_config: null");
    for (int i = 0; i < 2; ++i) {
        format[i] = h->P.spmv.format[i];
        unroll[i] = h->P.spmv.unroll[i];
        if (loads) loads[i] = h->P.spmv.loads[i];
    }
    return FEMGPU_OK;
}

int femgpu_set_spmv_config(femgpu_problem *h, const int format[2], const int unroll[2], const int loads[2]) {
    if (!h || !format || !unroll) return fail(FEMGPU_BAD_INPUT, "set_spmv_config: null");
    for (int i = 0; i < 2; ++i) {
        const int ld = loads ? loads[i] : LOAD_PLAIN;
        if (format[i] < 0 || format[i] > 2 || unroll[i] < 1 || unroll[i] > 3 || ld < 0 || ld > 2 ||
            (ld != LOAD_PLAIN && format[i] == FMT_SYM) || (ld == LOAD_NT16 && !h->P.d_col16) ||
            (format[i] == FMT_MF && !mf_eligible(h->P)))
            return fail(FEMGPU_BAD_INPUT, "set_spmv_config: format 0/1/2, unroll 1..3, loads 0..2 (1, 2 not with "
                                          "format 1; 2 needs every column offset in 16 bits; format 2 needs a Hexa "
                                          "or Tetra_10 problem)");
    }
    Problem &P = h->P;
    if (format[0] == FMT_MF || format[1] == FMT_MF) {
        const int rc = mf_build(P);  // the block layout, once; 1: a layout this product does not serve
        if (rc) return fail(rc == 1 ? FEMGPU_BAD_INPUT : FEMGPU_HIP_ERROR,
                            "set_spmv_config: matrix-free product not served (more than 32 sub-colours in a block, a "
                            "batch with a node of no element, or block lists beyond their index range)");
    }
    if ((format[0] == FMT_SYM || format[1] == FMT_SYM) && !P.sym.built) {
        const int rc = sym_ensure(P, format[0] == FMT_SYM ? 0 : 1);
        if (rc) return fail(rc == 1 ? FEMGPU_BAD_INPUT : FEMGPU_HIP_ERROR, "set_spmv_config: symmetric storage");
        P.sym.valid[0] = P.sym.valid[1] = false;  // values are (re)made from K before the next product
    }
    for (int i = 0; i < 2; ++i) {
        P.spmv.format[i] = format[i];
        P.spmv.unroll[i] = unroll[i];
        P.spmv.loads[i] = loads ? loads[i] : LOAD_PLAIN;
    }
    destroy_graph(P);
    return FEMGPU_OK;
}

int femgpu_set_row_order(int order) {
    if (order < ROW_ORDER_AUTO || order > ROW_ORDER_BY_LENGTH) return fail(FEMGPU_BAD_INPUT, "set_row_order: 0, 1 or 2");
    g_row_order.store(order);
    return FEMGPU_OK;
}

int femgpu_get_row_order(femgpu_problem *h, int *order, uint64_t *n_positions) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_row_order: null");
    if (order) *order = h->P.pat.perm.empty() ? ROW_ORDER_NATURAL : ROW_ORDER_BY_LENGTH;
    if (n_positions) *n_positions = h->P.pat.n_sell;
    return FEMGPU_OK;
}

int femgpu_set_node_order(int order) {
    if (order != NODE_ORDER_NATURAL && order != NODE_ORDER_CM && order != NODE_ORDER_AMD)
        return fail(FEMGPU_BAD_INPUT, "set_node_order: 0, 1 or 2");
    g_node_order.store(order);
    return FEMGPU_OK;
}

int femgpu_get_node_order(femgpu_problem *h, int *order, uint32_t *perm) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_node_order: null");
    const Problem &P = h->P;
    if (order) *order = P.nperm.empty() ? NODE_ORDER_NATURAL : P.node_order;
    if (perm)
        for (uint32_t i = 0; i < P.n_nodes; ++i) perm[i] = P.nperm.empty() ? i : P.nperm[i];
    return FEMGPU_OK;
}

namespace {
// A node graph given as CSR: every row holds its node; with sym, rows strictly increasing and
// the pattern symmetric.
int graph_pattern(const char *what, uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col, bool sym,
                  Pattern &pat) {
    if (!n_brows || !row_ptr || !col) return fail(FEMGPU_BAD_INPUT, std::string(what) + ": bad input");
    if (row_ptr[0] != 0) return fail(FEMGPU_BAD_INPUT, std::string(what) + ": row_ptr[0] must be 0");
    for (uint32_t r = 0; r < n_brows; ++r) {
        if (row_ptr[r + 1] < row_ptr[r]) return fail(FEMGPU_BAD_INPUT, std::string(what) + ": row_ptr decreasing");
        bool has_diag = false;
        for (uint32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
            if (col[p] >= n_brows) return fail(FEMGPU_BAD_INPUT, std::string(what) + ": column out of range");
            if (sym && p > row_ptr[r] && col[p] <= col[p - 1])
                return fail(FEMGPU_BAD_INPUT, std::string(what) + ": rows must be strictly increasing");
            has_diag |= col[p] == r;
        }
        if (!has_diag) return fail(FEMGPU_BAD_INPUT, std::string(what) + ": every row must hold its node");
    }
    if (sym)
        for (uint32_t r = 0; r < n_brows; ++r)
            for (uint32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
                if (!std::binary_search(col + row_ptr[col[p]], col + row_ptr[col[p] + 1], r))
                    return fail(FEMGPU_BAD_INPUT, std::string(what) + ": the pattern must be symmetric");
    pat.n_rows = n_brows;
    pat.row_ptr.assign(row_ptr, row_ptr + n_brows + 1);
    pat.col.assign(col, col + row_ptr[n_brows]);
    return FEMGPU_OK;
}
}  // namespace

int femgpu_node_order(uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col, uint32_t *order) {
    Pattern pat;
    if (!order) return fail(FEMGPU_BAD_INPUT, "node_order: bad input");
    int rc = graph_pattern("node_order", n_brows, row_ptr, col, false, pat);
    if (rc) return rc;
    const std::vector<uint32_t> o = cm_order(pat, 0, n_brows);
    std::memcpy(order, o.data(), 4 * (size_t)n_brows);
    return FEMGPU_OK;
}

int femgpu_node_order_amd(uint32_t n_brows, const uint32_t *row_ptr, const uint32_t *col, uint32_t *order) {
    Pattern pat;
    if (!order) return fail(FEMGPU_BAD_INPUT, "node_order_amd: bad input");
    int rc = graph_pattern("node_order_amd", n_brows, row_ptr, col, true, pat);
    if (rc) return rc;
    const std::vector<uint32_t> o = amd_order(pat, 0, n_brows);
    std::memcpy(order, o.data(), 4 * (size_t)n_brows);
    return FEMGPU_OK;
}

int femgpu_sell_order(uint32_t n_brows, const uint32_t *row_ptr, int order, uint32_t *perm, uint64_t *n_natural,
                      uint64_t *n_positions) {
    if (!n_brows || !row_ptr || order < ROW_ORDER_AUTO || order > ROW_ORDER_BY_LENGTH)
        return fail(FEMGPU_BAD_INPUT, "sell_order: bad input");
    for (uint32_t r = 0; r < n_brows; ++r)
        if (row_ptr[r + 1] < row_ptr[r]) return fail(FEMGPU_BAD_INPUT, "sell_order: row_ptr decreasing");
    Pattern pat;
    pat.n_rows = n_brows;
    pat.row_ptr.assign(row_ptr, row_ptr + n_brows + 1);
    if (order != ROW_ORDER_NATURAL)
        pat.perm = length_order(pat, {0, n_brows}, SORT_WINDOW_WGS * ROWS_PER_WG, order == ROW_ORDER_BY_LENGTH ? -1.0 : 0.03);
    if (perm)
        for (uint32_t t = 0; t < n_brows; ++t) perm[t] = pat.row_of(t);
    if (n_natural) *n_natural = sell_size(pat, {});
    if (n_positions) *n_positions = sell_size(pat, pat.perm);
    return FEMGPU_OK;
}

int femgpu_shift(femgpu_problem *h, int nb) {
    if (!h || nb < 0) return fail(FEMGPU_BAD_INPUT, "shift: bad input");
    Problem &P = h->P;
    const size_t n = 3 * (size_t)P.n_nodes;
    while (nb--) {
        h->time += h->time_steps[0];
        for (int i = NB_VECTORS - 1; i > 0; --i) h->time_steps[i] = h->time_steps[i - 1];
        h->time_steps[0] = h->time_steps[1];
        // vectors[i] = vectors[i - 1]: the buffers move up one slot, the last one's buffer takes
        // vectors[0] (which stays as the next solve's initial guess)
        double *last = h->d_hist[NB_VECTORS - 1];
        for (int i = NB_VECTORS - 1; i > 1; --i) h->d_hist[i] = h->d_hist[i - 1];
        h->d_hist[1] = nullptr;
        if (!last) HIPCHK(dalloc(&last, n));
        h->d_hist[1] = last;
        HIPCHK(hipMemcpyAsync(last, P.d_u, n * 8, hipMemcpyDeviceToDevice, P.stream));
    }
    HIPCHK(hipStreamSynchronize(P.stream));
    return FEMGPU_OK;
}

int femgpu_unshift(femgpu_problem *h, int nb) {
    if (!h || nb < 0) return fail(FEMGPU_BAD_INPUT, "unshift: bad input");
    Problem &P = h->P;
    const size_t n = 3 * (size_t)P.n_nodes;
    while (nb--) {
        h->time -= h->time_steps[1];
        for (int i = 0; i < NB_VECTORS - 2; ++i) h->time_steps[i] = h->time_steps[i + 1];
        h->time_steps[NB_VECTORS - 1] = h->time_steps[NB_VECTORS - 2];
        // vectors[i] = vectors[i + 1] for i < nb_vectors - 2, vectors[last] = vectors[last - 1]
        if (h->d_hist[1]) HIPCHK(hipMemcpyAsync(P.d_u, h->d_hist[1], n * 8, hipMemcpyDeviceToDevice, P.stream));
        else HIPCHK(launch_fill(P.d_u, n, 0.0, P.stream));
        double *spare = h->d_hist[1];
        for (int i = 1; i < NB_VECTORS - 1; ++i) h->d_hist[i] = h->d_hist[i + 1];
        h->d_hist[NB_VECTORS - 1] = nullptr;
        if (h->d_hist[NB_VECTORS - 2]) {
            if (!spare) HIPCHK(dalloc(&spare, n));
            HIPCHK(hipMemcpyAsync(spare, h->d_hist[NB_VECTORS - 2], n * 8, hipMemcpyDeviceToDevice, P.stream));
            h->d_hist[NB_VECTORS - 1] = spare;
        } else if (spare) {
            HIPCHK(hipStreamSynchronize(P.stream));
            dfree(spare);
        }
    }
    HIPCHK(hipStreamSynchronize(P.stream));
    return FEMGPU_OK;
}

int femgpu_set_initial_time_step(femgpu_problem *h, double ts) {
    if (!h || !std::isfinite(ts)) return fail(FEMGPU_BAD_INPUT, "set_initial_time_step: bad input");
    for (double &t : h->time_steps) t = ts;
    h->time = -2.0 * ts;
    return FEMGPU_OK;
}

int femgpu_set_next_time_step(femgpu_problem *h, double ts) {
    if (!h || !std::isfinite(ts)) return fail(FEMGPU_BAD_INPUT, "set_next_time_step: bad input");
    h->time_steps[0] = ts;
    return FEMGPU_OK;
}

int femgpu_get_next_time_step(femgpu_problem *h, double *ts) {
    if (!h || !ts) return fail(FEMGPU_BAD_INPUT, "get_next_time_step: null");
    *ts = h->time_steps[0];
    return FEMGPU_OK;
}

int femgpu_get_time(femgpu_problem *h, double *t) {
    if (!h || !t) return fail(FEMGPU_BAD_INPUT, "get_time: null");
    *t = h->time;
    return FEMGPU_OK;
}

int femgpu_set_time(femgpu_problem *h, double t) {
    if (!h || !std::isfinite(t)) return fail(FEMGPU_BAD_INPUT, "set_time: bad input");
    h->time = t;
    return FEMGPU_OK;
}

int femgpu_get_result_history(femgpu_problem *h, uint32_t num, double *u, size_t n) {
    if (!h || num >= (uint32_t)NB_VECTORS) return fail(FEMGPU_BAD_INPUT, "get_result: num beyond nb_vectors");
    if (num == 0) return copy_sub(*h, 0, h->P.d_u, u, n);
    if (!h->d_hist[num]) {
        if (!u || n != 3 * (size_t)(h->P.node_off[1] - h->P.node_off[0])) return fail(FEMGPU_BAD_INPUT, "get_result: size");
        std::fill(u, u + n, 0.0);
        return FEMGPU_OK;
    }
    return copy_sub(*h, 0, h->d_hist[num], u, n);
}

int femgpu_get_previous_result(femgpu_problem *h, double *u, size_t n) {
    return femgpu_get_result_history(h, 1, u, n);
}

int femgpu_get_indice_noda(femgpu_problem *h, uint32_t node, uint32_t *index) {
    if (!h || !index || node >= h->P.node_off[1]) return fail(FEMGPU_BAD_INPUT, "get_indice_noda: bad input");
    *index = 3 * node;
    return FEMGPU_OK;
}

int femgpu_get_nodal_forces(femgpu_problem *h, double *F, size_t n) {
    if (!h || !F || n != 3 * (size_t)(h->P.node_off[1] - h->P.node_off[0]))
        return fail(FEMGPU_BAD_INPUT, "get_nodal_forces: size");
    Problem &P = h->P;
    if (P.from_bsr) return fail(FEMGPU_BAD_INPUT, "get_nodal_forces: problem built from a matrix");
    for (uint32_t s = 0; s < P.n_sub; ++s)
        if (!P.material_set[s]) return fail(FEMGPU_NOT_READY, "get_nodal_forces: material not set");
    int rc = upload_faces(*h);
    if (rc) return rc;
    const size_t nv = 3 * (size_t)P.n_nodes;
    // the sollicitation of the element and skin loops alone (assemble_clean_mat(false, true),
    // formulation.h:447-458); the assembled one is kept aside and restored
    HIPCHK(hipMemcpyAsync(P.d_aux, P.d_F, nv * 8, hipMemcpyDeviceToDevice, P.stream));
    HIPCHK(launch_assemble(P, 0, 1));
    rc = copy_sub(*h, 0, P.d_F, F, n);
    HIPCHK(hipMemcpyAsync(P.d_F, P.d_aux, nv * 8, hipMemcpyDeviceToDevice, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    return rc;
}

int femgpu_set_assume_constant_matrix(femgpu_problem *h, int flag) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "set_assume_constant_matrix: null");
    h->assume_constant_matrix = flag != 0;
    return FEMGPU_OK;
}

int femgpu_set_nonlinear(femgpu_problem *h, double crit, int max_iter) {
    if (!h || !std::isfinite(crit) || max_iter < 0) return fail(FEMGPU_BAD_INPUT, "set_nonlinear: bad input");
    h->nl_crit = crit;
    h->max_nl = max_iter;
    return FEMGPU_OK;
}

int femgpu_get_nonlinear_solves(femgpu_problem *h, int *n) {
    if (!h || !n) return fail(FEMGPU_BAD_INPUT, "get_nonlinear_solves: null");
    *n = h->last_nl_solves;
    return FEMGPU_OK;
}

int femgpu_set_default_solve_opts(femgpu_problem *h, const femgpu_solve_opts *o) {
    if (!h || !o) return fail(FEMGPU_BAD_INPUT, "set_default_solve_opts: null");
    const int rc = check_opts(o);
    if (rc) return rc;
    h->dopts = *o;
    return FEMGPU_OK;
}

int femgpu_get_precond(femgpu_problem *h, const femgpu_solve_opts *o) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "get_precond: null");
    int rc = check_opts(o);
    if (rc) return rc;
    Problem &P = h->P;
    if (!P.assembled_mat) return fail(FEMGPU_NOT_READY, "get_precond: assemble first");
    const int kind = o ? o->precond : h->dopts.precond;
    P.minv_kind = -1;
    P.ic0.valid = P.ic0.valid32 = false;
    h->precond_pinned = h->precond_stale = false;
    rc = precond_build(P, kind);
    if (rc == 3) return fail(FEMGPU_NOT_SPD, "get_precond: non-positive pivot");
    if (rc) return fail(rc == 1 ? FEMGPU_BAD_INPUT : FEMGPU_HIP_ERROR, "get_precond failed");
    h->precond_pinned = true;
    h->pinned_kind = kind;
    return FEMGPU_OK;
}

int femgpu_solve_system_using_precond(femgpu_problem *h, const femgpu_solve_opts *o, femgpu_solve_stats *st) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "solve_system_using_precond: null");
    if (!h->precond_pinned) return fail(FEMGPU_NOT_READY, "solve_system_using_precond: call get_precond first");
    if ((o ? o->precond : h->dopts.precond) != h->pinned_kind)
        return fail(FEMGPU_BAD_INPUT, "solve_system_using_precond: preconditioner differs from get_precond's");
    femgpu_solve_stats tmp{};
    const int rc = solve_system(*h, o ? o : &h->dopts, st ? st : &tmp, nullptr, nullptr);
    if (st) st->status = rc;
    return rc;
}

int femgpu_stepper_create(uint32_t n, femgpu_problem *const *forms, femgpu_stepper **out) {
    if (!out || (n && !forms)) return fail(FEMGPU_BAD_INPUT, "stepper_create: null");
    for (uint32_t i = 0; i < n; ++i)
        if (!forms[i] || forms[i]->P.from_bsr) return fail(FEMGPU_BAD_INPUT, "stepper_create: formulation " + std::to_string(i));
    auto *s = new (std::nothrow) femgpu_stepper();
    if (!s) return fail(FEMGPU_BAD_INPUT, "stepper_create: out of host memory");
    s->forms.assign(forms, forms + n);
    *out = s;
    return FEMGPU_OK;
}

int femgpu_stepper_destroy(femgpu_stepper *s) {
    delete s;
    return FEMGPU_OK;
}

int femgpu_stepper_get_time(femgpu_stepper *s, double *t) {
    if (!s || !t) return fail(FEMGPU_BAD_INPUT, "stepper_get_time: null");
    *t = s->time;
    return FEMGPU_OK;
}

int femgpu_stepper_iteration(femgpu_stepper *s, femgpu_solve_stats *stats, int *updated) {
    if (!s) return fail(FEMGPU_BAD_INPUT, "stepper_iteration: null");
    const size_t nf = s->forms.size();
    std::vector<int> upd(nf, 1);
    int rc = FEMGPU_OK;
    if (!s->initialised) {
        for (size_t i = 0; i < nf && !rc; ++i) {
            femgpu_problem *f = s->forms[i];
            rc = femgpu_shift(f, 1);
            if (!rc) rc = femgpu_assemble(f, 1, 1);
            if (!rc && f->assume_constant_matrix) rc = femgpu_get_precond(f, &f->dopts);
        }
        if (rc) return rc;
        s->initialised = true;
    } else {
        for (size_t i = 0; i < nf; ++i) upd[i] = s->forms[i]->time == s->time;
        for (size_t i = 0; i < nf && !rc; ++i) {
            if (!upd[i]) continue;
            femgpu_problem *f = s->forms[i];
            rc = femgpu_shift(f, 1);
            if (!rc) rc = femgpu_assemble(f, f->assume_constant_matrix ? 0 : 1, 1);
        }
        if (rc) return rc;
    }
    for (size_t i = 0; i < nf; ++i) {
        femgpu_problem *f = s->forms[i];
        femgpu_solve_stats st{};
        if (upd[i]) {
            if (f->assume_constant_matrix) {
                rc = femgpu_solve_system_using_precond(f, &f->dopts, &st);
                f->last_nl_solves = 1;
            } else {
                rc = nonlinear_solve(*f, &f->dopts, &st, true, false);
            }
        }
        if (stats) stats[i] = st;
        if (updated) updated[i] = upd[i];
        if (rc) return rc;
    }
    double t = std::numeric_limits<double>::max();
    for (femgpu_problem *f : s->forms) t = std::min(t, f->time);
    s->time = t;
    return FEMGPU_OK;
}

int femgpu_set_graph_replay(femgpu_problem *h, int enable) {
    if (!h) return fail(FEMGPU_BAD_INPUT, "set_graph_replay: null");
    h->P.use_graph = enable != 0;
    if (!enable) destroy_graph(h->P);
    return FEMGPU_OK;
}

// ---- batched uneven problems -------------------------------------------------------------------

int femgpu_batch_create(uint32_t n, const femgpu_mesh *meshes, femgpu_batch **out) {
    if (!out || !meshes || !n) return fail(FEMGPU_BAD_INPUT, "batch_create: bad input");
    *out = nullptr;
    auto *b = new (std::nothrow) femgpu_batch();
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_create: out of host memory");
    const int rc = build_problem(b->H, n, meshes);
    if (rc) {
        free_problem(b->H.P);
        delete b;
        return rc;
    }
    b->H.P.use_graph = true;
    *out = b;
    return FEMGPU_OK;
}

int femgpu_batch_set_material(femgpu_batch *b, uint32_t i, const femgpu_material *mat) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_set_material: null");
    return set_material(b->H, i, mat);
}

int femgpu_batch_set_traction(femgpu_batch *b, uint32_t i, uint32_t n_faces, const uint32_t *faces, const double *trac) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_set_traction: null");
    return set_traction(b->H, i, n_faces, faces, trac);
}

int femgpu_batch_add_constraints(femgpu_batch *b, uint32_t i, uint32_t n, const uint32_t *dofs, const double *values,
                                 double penalty) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_add_constraints: null");
    return add_constraints(b->H, i, n, dofs, values, penalty);
}

int femgpu_batch_assemble(femgpu_batch *b) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_assemble: null");
    return assemble(b->H, 1, 1);
}

int femgpu_batch_solve(femgpu_batch *b, const femgpu_solve_opts *opts, femgpu_solve_stats *per, double *total_ms) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_solve: null");
    std::vector<femgpu_solve_stats> st(b->H.P.n_sub);
    const int rc = solve_system(b->H, opts, st.data(), nullptr, nullptr);
    if (per) std::copy(st.begin(), st.end(), per);
    if (total_ms) *total_ms = st.empty() ? 0.0 : st[0].solve_ms;
    return rc;
}

int femgpu_batch_get_result(femgpu_batch *b, uint32_t i, double *u, size_t n) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_get_result: null");
    return copy_sub(b->H, i, b->H.P.d_u, u, n);
}

int femgpu_batch_clear_result(femgpu_batch *b) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_clear_result: null");
    HIPCHK(launch_fill(b->H.P.d_u, 3 * (size_t)b->H.P.n_nodes, 0.0, b->H.P.stream));
    HIPCHK(hipStreamSynchronize(b->H.P.stream));
    return FEMGPU_OK;
}

int femgpu_batch_get_assemble_ms(femgpu_batch *b, double *ms) {
    if (!b || !ms) return fail(FEMGPU_BAD_INPUT, "batch_get_assemble_ms: null");
    *ms = b->H.P.last_asm_ms;
    return FEMGPU_OK;
}

int femgpu_batch_set_assembly_rows(femgpu_batch *b, uint32_t rows) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_set_assembly_rows: null");
    return femgpu_set_assembly_rows(&b->H, rows);
}

int femgpu_batch_get_assembly_rows(femgpu_batch *b, uint32_t *rows, uint32_t *max_rows) {
    if (!b) return fail(FEMGPU_BAD_INPUT, "batch_get_assembly_rows: null");
    return femgpu_get_assembly_rows(&b->H, rows, max_rows);
}

int femgpu_batch_destroy(femgpu_batch *b) {
    if (!b) return FEMGPU_OK;
    free_problem(b->H.P);
    delete b;
    return FEMGPU_OK;
}

}  // extern "C"
