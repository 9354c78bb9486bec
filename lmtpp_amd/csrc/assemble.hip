// Assembly kernels for gfx950: per-element stiffness and load of isotropic linear elasticity
// (Tetra, Tetra_10, Hexa), scattered colour by colour into the 3x3-block CSR without atomics;
// skin tractions; penalty constraints; FP32 copy; per-element adjoint gradient.
//
// Layout of K in HBM: 9 planes of nnzb doubles (plane q = block entry (q/3, q%3)), plane
// stride P.ps; a block's position p is its CSR slot.  Plane-major storage makes every
// SpMV load a contiguous run across the lanes of a row group.
#include "device_common.hpp"
#include "generated/elasticity_iso.inc"

namespace femgpu {

// ---------------------------------------------------------------------------------------------
// Element descriptions, Hooke law, per-Gauss-point blocks and loads come from the generator
// (lmtpp_amd/formal_hip -> generated/elasticity_iso.inc): Gauss tables of the runtime loop
// (formulation.h:2917-2933) at order 2*degree-2, shape functions of formulations/element_*.py,
// and the derivatives of the elasticity_iso weak form.
// ---------------------------------------------------------------------------------------------
template <int K> struct ET;
template <> struct ET<TETRA> { static constexpr int NPE = 4, NGP = 1, TEAM = 16; };
template <> struct ET<TETRA10> { static constexpr int NPE = 10, NGP = 4, TEAM = 128; };
template <> struct ET<HEXA> { static constexpr int NPE = 8, NGP = 8, TEAM = 64; };
static_assert(gen::vol<TETRA>::NPE == 4 && gen::vol<TETRA>::NGP == 1, "generated Tetra");
static_assert(gen::vol<TETRA10>::NPE == 10 && gen::vol<TETRA10>::NGP == 4, "generated Tetra_10");
static_assert(gen::vol<HEXA>::NPE == 8 && gen::vol<HEXA>::NGP == 8, "generated Hexa");

// Phases A (coordinates to LDS) and B (per Gauss point gradients) shared by the assembly and
// the adjoint-gradient kernels.  J[j][i] = dx_i/dxi_j (formal_lf/element.py:195-205), detJ
// signed (formal_lf/formulation.py:553), grad = J^-1 dN.
template <int K>
struct ElemLds {
    double x[ET<K>::NPE * 3];
    double g[ET<K>::NGP][ET<K>::NPE][3];
    double wd[ET<K>::NGP];
    double N[ET<K>::NGP][ET<K>::NPE];
    double ji[ET<K>::NGP][9];   // J^-1 at each Gauss point (ji[q][3 j + i] = Ji[j][i])
};

// Phase B alone, coordinates already in L.x (written before the barrier this function starts with).
template <int K>
__device__ inline void elem_gradients(ElemLds<K> &L, int t, bool valid, int *flags);

template <int K>
__device__ inline void elem_geometry(ElemLds<K> &L, int t, bool valid, uint32_t e, const uint32_t *__restrict__ conn,
                                     const double *__restrict__ coords, int *flags) {
    constexpr int NPE = ET<K>::NPE;
    if (valid && t < NPE * 3) L.x[t] = coords[3 * (size_t)conn[(size_t)e * NPE + t / 3] + t % 3];
    elem_gradients<K>(L, t, valid, flags);
}

// J, det J and J^-1 once per Gauss point (a lane each), then the gradient of every (point, node)
// pair (a lane each).
template <int K>
__device__ inline void elem_gradients(ElemLds<K> &L, int t, bool valid, int *flags) {
    constexpr int NPE = ET<K>::NPE, NGP = ET<K>::NGP;
    __syncthreads();
    if (valid && t < NGP) {
        double xi[3], w;
        gen::vol<K>::gauss(t, xi, w);
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int b = 0; b < NPE; ++b) {
            double d[3];
            gen::vol<K>::shape(b, xi, d);
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) J[j][i] += d[j] * L.x[3 * b + i];
        }
        double Ji[3][3];
        const double det = jac_inverse(J, Ji);
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) L.ji[t][3 * j + i] = Ji[j][i];
        L.wd[t] = w * det;
        if (!(det > 0)) atomicOr(flags, 1);
    }
    __syncthreads();
    if (valid && t < NGP * NPE) {
        const int q = t / NPE, a = t % NPE;
        double xi[3], w;
        gen::vol<K>::gauss(q, xi, w);
        double d[3];
        L.N[q][a] = gen::vol<K>::shape(a, xi, d);
        const double *Ji = L.ji[q];
        for (int j = 0; j < 3; ++j) L.g[q][a][j] = Ji[3 * j] * d[0] + Ji[3 * j + 1] * d[1] + Ji[3 * j + 2] * d[2];
    }
    __syncthreads();
}

// 3x3 block of element nodes (a,b): sum over Gauss points of the generated per-point block
//   w_q detJ_q [lambda g_a,i g_b,j + mu g_a,j g_b,i + mu delta_ij g_a.g_b]   (H = lambda+2mu, lambda, 2mu)
template <int K>
__device__ inline void elem_block(const ElemLds<K> &L, int a, int b, const double H[3], double k[9]) {
#pragma unroll
    for (int c = 0; c < 9; ++c) k[c] = 0;
    for (int q = 0; q < ET<K>::NGP; ++q) gen::block(L.g[q][a], L.g[q][b], H, L.wd[q], k);
}

// One colour of the element loop (formulation.h:626-647).  A team of ET<K>::TEAM lanes owns one
// element; no two elements of a colour share a node, so the read-modify-write scatter into K and
// F needs no atomics and its order is fixed.  BLOCKS: K is a block-contiguous staging copy (the
// 9 values of position p at 9 p .. 9 p + 8, one or two cache lines per block instead of nine
// plane-strided ones), turned into planes afterwards by k_blocks_to_planes.
template <int K, bool BLOCKS>
__global__ __launch_bounds__(256) void k_assemble(const uint32_t *__restrict__ order, uint32_t count,
                                                  const uint32_t *__restrict__ conn, const double *__restrict__ coords,
                                                  const double *__restrict__ E, const uint32_t *__restrict__ elem_sub,
                                                  const SubMaterial *__restrict__ mat, const uint32_t *__restrict__ slot,
                                                  double *__restrict__ Kv, uint64_t ps, double *__restrict__ F, int do_mat,
                                                  int do_vec, int *flags) {
    constexpr int NPE = ET<K>::NPE, NGP = ET<K>::NGP, TEAM = ET<K>::TEAM, TPW = 256 / TEAM;
    __shared__ ElemLds<K> lds[TPW];
    const int team = threadIdx.x / TEAM, t = threadIdx.x % TEAM;
    const uint32_t idx = blockIdx.x * TPW + team;
    const bool valid = idx < count;
    const uint32_t e = valid ? order[idx] : 0;
    ElemLds<K> &L = lds[team];
    elem_geometry<K>(L, t, valid, e, conn, coords, flags);
    if (!valid) return;
    const SubMaterial m = mat[elem_sub[e]];
    // Only the local upper triangle is evaluated (formal_lf/calculate_matrix.py:69-97); block (a,b),
    // a < b, is added at (a,b) and, transposed, at (b,a), so the stored matrix is exactly symmetric.
    if (do_mat && t < NPE * NPE && t / NPE <= t % NPE) {
        const int a = t / NPE, b = t % NPE;
        double H[3];
        gen::hooke(E[e], m.nu, H);
        double k[9];
        elem_block<K>(L, a, b, H, k);
        const uint64_t p = slot[(size_t)e * NPE * NPE + t];
        const uint64_t sp = BLOCKS ? 1 : ps, bp = BLOCKS ? 9 * p : p;
        if (a == b) {
            k[3] = k[1]; k[6] = k[2]; k[7] = k[5];
#pragma unroll
            for (int c = 0; c < 9; ++c) Kv[c * sp + bp] += k[c];
        } else {
            const uint64_t pt = slot[(size_t)e * NPE * NPE + b * NPE + a], bt = BLOCKS ? 9 * pt : pt;
#pragma unroll
            for (int c = 0; c < 9; ++c) Kv[c * sp + bp] += k[c];
#pragma unroll
            for (int c = 0; c < 9; ++c) Kv[c * sp + bt] += k[3 * (c % 3) + c / 3];
        }
    }
    if (do_vec && t < NPE * 3) {
        // F_a = sum_q w_q detJ_q N_a(q) (density f_vol) (formulation_elasticity_iso.py:52)
        const int a = t / 3, d = t % 3;
        double f3[3] = {0, 0, 0};
        for (int q = 0; q < NGP; ++q) gen::load(L.N[q][a], m.bf0, m.bf1, m.bf2, L.wd[q], f3);
        F[3 * (size_t)conn[(size_t)e * NPE + a] + d] += f3[d];
    }
}

// ---------------------------------------------------------------------------------------------
// Gather element loop, pass 1: every element's upper 3x3 blocks (summed over its Gauss points,
// exactly as k_assemble forms them) and nodal loads, stored with plain writes: block (a, b),
// a <= b, at EM[(e NUP + upper_index(a, b)) 9 ..], load of node a at EF[(e NPE + a) 3 ..].
// ---------------------------------------------------------------------------------------------
template <int K>
__device__ constexpr int upper_index(int a, int b) { return a * ET<K>::NPE - a * (a - 1) / 2 + (b - a); }

template <int K>
__global__ __launch_bounds__(256) void k_elem_mats(uint32_t n_elem, const uint32_t *__restrict__ conn,
                                                   const double *__restrict__ coords, const double *__restrict__ E,
                                                   const uint32_t *__restrict__ elem_sub,
                                                   const SubMaterial *__restrict__ mat, double *__restrict__ EM,
                                                   double *__restrict__ EF, int do_mat, int do_vec, int *flags) {
    constexpr int NPE = ET<K>::NPE, NGP = ET<K>::NGP, TEAM = ET<K>::TEAM, TPW = 256 / TEAM;
    constexpr int NUP = NPE * (NPE + 1) / 2;
    __shared__ ElemLds<K> lds[TPW];
    const int team = threadIdx.x / TEAM, t = threadIdx.x % TEAM;
    const uint32_t e = blockIdx.x * TPW + team;
    const bool valid = e < n_elem;
    ElemLds<K> &L = lds[team];
    elem_geometry<K>(L, t, valid, valid ? e : 0, conn, coords, flags);
    if (!valid) return;
    const SubMaterial m = mat[elem_sub[e]];
    if (do_mat && t < NPE * NPE && t / NPE <= t % NPE) {
        const int a = t / NPE, b = t % NPE;
        double H[3];
        gen::hooke(E[e], m.nu, H);
        double k[9];
        elem_block<K>(L, a, b, H, k);
        if (a == b) { k[3] = k[1]; k[6] = k[2]; k[7] = k[5]; }
        double *dst = EM + ((size_t)e * NUP + upper_index<K>(a, b)) * 9;
#pragma unroll
        for (int c = 0; c < 9; ++c) dst[c] = k[c];
    }
    if (do_vec && t < NPE * 3) {
        const int a = t / 3, d = t % 3;
        double f3[3] = {0, 0, 0};
        for (int q = 0; q < NGP; ++q) gen::load(L.N[q][a], m.bf0, m.bf1, m.bf2, L.wd[q], f3);
        EF[((size_t)e * NPE + a) * 3 + d] = f3[d];
    }
}

// Pass 2: every stored position p of K (a lane each, the planes written once and coalesced) sums
// its contributions ent[ptr[p] .. ptr[p + 1]) in the colour order of the colour passes; an entry is
// (element block index << 1) | transposed.  Same adds in the same order from zero as the colour
// passes, so K is bitwise theirs.
__global__ __launch_bounds__(256) void k_gather_blocks(uint64_t ps, const uint32_t *__restrict__ ptr,
                                                       const uint32_t *__restrict__ ent, const double *__restrict__ EM,
                                                       double *__restrict__ Kv) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < ps; p += (uint64_t)gridDim.x * blockDim.x) {
        double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t j = ptr[p]; j < ptr[p + 1]; ++j) {
            const uint32_t c = ent[j];
            const double *blk = EM + (size_t)(c >> 1) * 9;
            if (c & 1u) {
#pragma unroll
                for (int q = 0; q < 9; ++q) v[q] += blk[3 * (q % 3) + q / 3];
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) v[q] += blk[q];
            }
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) Kv[q * ps + p] = v[q];
    }
}

// Box classes (ASM_GATHER_BOX): the unit-modulus element matrix K_c(1) and the body loads of every
// class of axis-aligned Hexa boxes, a team of 64 lanes per class, from a box at the origin with the
// class's edge lengths through the same phases and generated per-point blocks as the element loop.
__global__ __launch_bounds__(64) void k_box_ref(const double *__restrict__ ch, const uint32_t *__restrict__ csub,
                                                const SubMaterial *__restrict__ mat, double *__restrict__ kref,
                                                double *__restrict__ fref, int *flags) {
    constexpr int NPE = 8, NGP = 8, NUP = NPE * (NPE + 1) / 2;
    __shared__ ElemLds<HEXA> L;
    const uint32_t c = blockIdx.x;
    const int t = threadIdx.x;
    if (t < NPE * 3) {
        // node a of element_Hexa.py at corner (x, y, z) = ((a + 1) / 2 % 2, a / 2 % 2, a / 4)
        const int a = t / 3, i = t % 3;
        const int bit = i == 0 ? ((a + 1) >> 1) & 1 : i == 1 ? (a >> 1) & 1 : a >> 2;
        L.x[t] = bit ? ch[3 * (size_t)c + i] : 0.0;
    }
    elem_gradients<HEXA>(L, t, true, flags);
    const SubMaterial m = mat[csub[c]];
    if (t / NPE <= t % NPE) {
        const int a = t / NPE, b = t % NPE;
        double H[3];
        gen::hooke(1.0, m.nu, H);
        double k[9];
        elem_block<HEXA>(L, a, b, H, k);
        if (a == b) { k[3] = k[1]; k[6] = k[2]; k[7] = k[5]; }
        double *dst = kref + ((size_t)c * NUP + upper_index<HEXA>(a, b)) * 9;
#pragma unroll
        for (int q = 0; q < 9; ++q) dst[q] = k[q];
    }
    if (t < NPE * 3) {
        const int a = t / 3, d = t % 3;
        double f3[3] = {0, 0, 0};
        for (int q = 0; q < NGP; ++q) gen::load(L.N[q][a], m.bf0, m.bf1, m.bf2, L.wd[q], f3);
        fref[((size_t)c * NPE + a) * 3 + d] = f3[d];
    }
}

// Pass 2 of ASM_GATHER_BOX: as k_gather_blocks, each entry's block being E_e times its class's
// unit-modulus block (entry (e NUP + pair) << 1 | transposed).
__global__ __launch_bounds__(256) void k_gather_blocks_box(uint64_t ps, const uint32_t *__restrict__ ptr,
                                                           const uint32_t *__restrict__ ent, const double *__restrict__ E,
                                                           const uint16_t *__restrict__ cls,
                                                           const double *__restrict__ kref, double *__restrict__ Kv) {
    constexpr uint32_t NUP = 36;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < ps; p += (uint64_t)gridDim.x * blockDim.x) {
        double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t j = ptr[p]; j < ptr[p + 1]; ++j) {
            const uint32_t c = ent[j], blk = c >> 1, e = blk / NUP, pair = blk - e * NUP;
            const double Ee = E[e];
            const double *r = kref + ((size_t)cls[e] * NUP + pair) * 9;
            if (c & 1u) {
#pragma unroll
                for (int q = 0; q < 9; ++q) v[q] += Ee * r[3 * (q % 3) + q / 3];
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) v[q] += Ee * r[q];
            }
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) Kv[q * ps + p] = v[q];
    }
}

// Loads of ASM_GATHER_BOX: node n sums its class loads (entry e NPE + a) in colour order.
__global__ __launch_bounds__(256) void k_gather_loads_box(uint32_t n_nodes, const uint32_t *__restrict__ ptr,
                                                          const uint32_t *__restrict__ ent,
                                                          const uint16_t *__restrict__ cls,
                                                          const double *__restrict__ fref, double *__restrict__ F) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= n_nodes) return;
    double v[3] = {0, 0, 0};
    for (uint32_t j = ptr[n]; j < ptr[n + 1]; ++j) {
        const uint32_t c = ent[j];
        const double *f = fref + ((size_t)cls[c >> 3] * 8 + (c & 7u)) * 3;
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] += f[d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) F[3 * (size_t)n + d] = v[d];
}

// The loads: node n sums the element loads ent[ptr[n] .. ptr[n + 1]) (entry e NPE + a), in colour order.
__global__ __launch_bounds__(256) void k_gather_loads(uint32_t n_nodes, const uint32_t *__restrict__ ptr,
                                                      const uint32_t *__restrict__ ent, const double *__restrict__ EF,
                                                      double *__restrict__ F) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= n_nodes) return;
    double v[3] = {0, 0, 0};
    for (uint32_t j = ptr[n]; j < ptr[n + 1]; ++j) {
        const double *f = EF + (size_t)ent[j] * 3;
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] += f[d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) F[3 * (size_t)n + d] = v[d];
}

// ---------------------------------------------------------------------------------------------
// Row-group element loop: workgroup G owns storage rows [G R, G R + R) (R divides 64, so the group
// lies in one SELL slice; storage row t holds block row perm[t], perm null for the identity).  It walks the elements incident to its rows colour by colour, the same
// order as the colour-pass kernel above, computes each element's geometry in a team of lanes and
// adds the blocks (and loads) of its own rows into an LDS accumulator; rows of one colour's
// elements are distinct, so no two lanes add to one entry at once and the summation order of
// every entry is the colour order.  The finished rows are then streamed out once, each (slot,
// plane) a contiguous run of R values: K and F are written, not read-modified-written, and need
// no clearing.  Elements touching several groups redo their geometry in each (about 4x on a Hexa
// lattice with R = 64); every block is evaluated once, by the group owning its row.
// ---------------------------------------------------------------------------------------------
constexpr uint32_t ASM_SLOTS = 1792;   // R x (slice width) block slots accumulated in LDS
constexpr int ASM_THREADS = 512;

template <int K>
__global__ __launch_bounds__(ASM_THREADS) void k_assemble_rows(
    uint32_t R, uint32_t n_rows, uint32_t nc, const uint32_t *__restrict__ gptr, const uint32_t *__restrict__ gel,
    const uint32_t *__restrict__ conn, const double *__restrict__ coords, const double *__restrict__ E,
    const uint32_t *__restrict__ elem_sub, const SubMaterial *__restrict__ mat, const uint32_t *__restrict__ slot,
    const uint32_t *__restrict__ sptr, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ iperm,
    double *__restrict__ Kv, uint64_t ps, double *__restrict__ F, int do_mat, int do_vec, int *flags) {
    constexpr int NPE = ET<K>::NPE, NGP = ET<K>::NGP, TEAM = ET<K>::TEAM, TPW = ASM_THREADS / TEAM;
    __shared__ double acc[ASM_SLOTS * 9];   // acc[(k * 9 + q) * R + i]: storage row r0 + i, slot k, plane q
    __shared__ double facc[64 * 3];
    __shared__ ElemLds<K> lds[TPW];
    const uint32_t G = blockIdx.x, r0 = G * R, r1 = min(r0 + R, n_rows);
    const uint32_t s = r0 >> 6, W = (sptr[s + 1] - sptr[s]) >> 6, base = sptr[s];
    for (uint32_t i = threadIdx.x; i < W * 9 * R; i += blockDim.x) acc[i] = 0.0;
    for (uint32_t i = threadIdx.x; i < R * 3; i += blockDim.x) facc[i] = 0.0;
    const int team = threadIdx.x / TEAM, t = threadIdx.x % TEAM;
    ElemLds<K> &L = lds[team];
    // Batches of TPW elements never straddle colours.  The coordinates of the next batch are
    // loaded into registers while the current one computes (the gather latency is otherwise
    // exposed: one workgroup per CU).
    const uint32_t *gp = gptr + (size_t)G * (nc + 1);
    uint32_t c = 0, bb = gp[0];
    while (c < nc && bb >= gp[c + 1]) bb = gp[++c];
    uint32_t e_nx = 0;
    bool v_nx = false;
    double x_nx = 0.0;
    auto fetch = [&](uint32_t cc, uint32_t b) {
        const uint32_t idx = b + team;
        v_nx = cc < nc && idx < gp[cc + 1];
        e_nx = v_nx ? gel[idx] : 0;
        if (v_nx && t < NPE * 3) x_nx = coords[3 * (size_t)conn[(size_t)e_nx * NPE + t / 3] + t % 3];
    };
    fetch(c, bb);
    while (c < nc) {
        const bool valid = v_nx;
        const uint32_t e = e_nx;
        if (valid && t < NPE * 3) L.x[t] = x_nx;
        // advance to the next batch (next colour when this one is exhausted) and prefetch it
        bb += TPW;
        while (c < nc && bb >= gp[c + 1]) bb = gp[++c];
        fetch(c, bb);
        elem_gradients<K>(L, t, valid, flags);   // its barriers order the batches
        if (!valid) continue;
        const SubMaterial m = mat[elem_sub[e]];
        if (do_mat && t < NPE * NPE) {
            const int a = t / NPE, b = t % NPE;
            const uint32_t row = conn[(size_t)e * NPE + a], tr = iperm ? iperm[row] : row;
            if (tr >= r0 && tr < r1) {
                double H[3];
                gen::hooke(E[e], m.nu, H);
                double k[9];
                elem_block<K>(L, a < b ? a : b, a < b ? b : a, H, k);
                const uint32_t kk = (slot[(size_t)e * NPE * NPE + t] - base - (tr & 63u)) >> 6;
                double *A = acc + (size_t)kk * 9 * R + (tr - r0);
                if (a == b) {
                    k[3] = k[1]; k[6] = k[2]; k[7] = k[5];
#pragma unroll
                    for (int q = 0; q < 9; ++q) A[q * R] += k[q];
                } else if (a < b) {
#pragma unroll
                    for (int q = 0; q < 9; ++q) A[q * R] += k[q];
                } else {  // block (a, b) = block (b, a)^T
#pragma unroll
                    for (int q = 0; q < 9; ++q) A[q * R] += k[3 * (q % 3) + q / 3];
                }
            }
        }
        if (do_vec && t < NPE * 3) {
            const int a = t / 3, d = t % 3;
            const uint32_t row = conn[(size_t)e * NPE + a], tr = iperm ? iperm[row] : row;
            if (tr >= r0 && tr < r1) {
                double f3[3] = {0, 0, 0};
                for (int q = 0; q < NGP; ++q) gen::load(L.N[q][a], m.bf0, m.bf1, m.bf2, L.wd[q], f3);
                facc[(tr - r0) * 3 + d] += f3[d];
            }
        }
    }
    __syncthreads();
    if (do_mat)
        for (uint32_t i = threadIdx.x; i < W * 9 * R; i += blockDim.x) {
            const uint32_t li = i % R, kq = i / R, q = kq % 9, kk = kq / 9;
            if (r0 + li < r1) Kv[q * ps + base + 64ull * kk + ((r0 + li) & 63u)] = acc[i];
        }
    if (do_vec)
        for (uint32_t i = threadIdx.x; i < R * 3; i += blockDim.x) {
            const uint32_t tr = r0 + i / 3;
            if (tr < r1) F[3 * (size_t)(perm ? perm[tr] : tr) + i % 3] = facc[i];
        }
}

// Skin loads, one thread per face, faces coloured like elements.  Gauss tables of the face
// elements at the generator's key (formal_lf/formulation.py:400-442): Triangle 1 point,
// Triangle_6 three mid-edge points, Quad 2x2 at (1 +- sqrt(1/3))/2 (element_Quad.py:70-75).
template <int FK>
__global__ __launch_bounds__(256) void k_face_load(const uint32_t *__restrict__ order, uint32_t count,
                                                   const uint32_t *__restrict__ faces, const double *__restrict__ trac,
                                                   const double *__restrict__ coords, double *__restrict__ F) {
    using FE = gen::face<FK>;
    constexpr int NPF = FE::NPE;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t f = order[i];
    double x[NPF][3];
    uint32_t nd[NPF];
    for (int a = 0; a < NPF; ++a) {
        nd[a] = faces[(size_t)f * NPF + a];
        for (int d = 0; d < 3; ++d) x[a][d] = coords[3 * (size_t)nd[a] + d];
    }
    const double t0 = trac[3 * (size_t)f], t1 = trac[3 * (size_t)f + 1], t2 = trac[3 * (size_t)f + 2];
    double acc[NPF] = {};
    for (int g = 0; g < FE::NGP; ++g) {
        double uv[2], w;
        FE::gauss(g, uv, w);
        double N[NPF], du[NPF], dv[NPF];
        for (int a = 0; a < NPF; ++a) {
            double d[2];
            N[a] = FE::shape(a, uv, d);
            du[a] = d[0];
            dv[a] = d[1];
        }
        double a0[3] = {0, 0, 0}, a1[3] = {0, 0, 0};
        for (int a = 0; a < NPF; ++a)
            for (int d = 0; d < 3; ++d) { a0[d] += du[a] * x[a][d]; a1[d] += dv[a] * x[a][d]; }
        const double n0 = a0[1] * a1[2] - a0[2] * a1[1], n1 = a0[2] * a1[0] - a0[0] * a1[2], n2 = a0[0] * a1[1] - a0[1] * a1[0];
        const double ds = w * sqrt(n0 * n0 + n1 * n1 + n2 * n2);   // |dS| of the skin element
        for (int a = 0; a < NPF; ++a) acc[a] += ds * N[a];
    }
    for (int a = 0; a < NPF; ++a) {
        F[3 * (size_t)nd[a]] += acc[a] * t0;
        F[3 * (size_t)nd[a] + 1] += acc[a] * t1;
        F[3 * (size_t)nd[a] + 2] += acc[a] * t2;
    }
}

__global__ void k_fill(double *p, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Plain copy, four 16-byte loads per lane issued before the four stores: the calibration kernel
// for measured bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters (it moves exactly
// 2 * 8 * n bytes).
__global__ __launch_bounds__(256) void k_copy(double2 *__restrict__ dst, const double2 *__restrict__ src, size_t n2) {
    const size_t base = blockIdx.x * (size_t)1024 + threadIdx.x;
    double2 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + 256 * k < n2) v[k] = src[base + 256 * k];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + 256 * k < n2) dst[base + 256 * k] = v[k];
}

// Streaming read (sum of the input, one partial per workgroup), four 16-byte loads per lane in
// flight: the read-bandwidth calibration for read-dominated kernels (reads exactly 8 n bytes).
__global__ __launch_bounds__(256) void k_read(const double2 *__restrict__ src, size_t n2, double *__restrict__ out) {
    const size_t base = blockIdx.x * (size_t)1024 + threadIdx.x;
    double2 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = base + 256 * k < n2 ? src[base + 256 * k] : double2{0.0, 0.0};
    double s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) s += v[k].x + v[k].y;
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) out[blockIdx.x * 4 + (threadIdx.x >> 6)] = s;
}

hipError_t launch_read(const double *src, size_t n, double *out, hipStream_t s) {
    const size_t n2 = n / 2;
    hipLaunchKernelGGL(k_read, dim3((unsigned)((n2 + 1023) / 1024)), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(src), n2, out);
    return hipGetLastError();
}

hipError_t launch_copy(double *dst, const double *src, size_t n, hipStream_t s) {
    const size_t n2 = n / 2;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n2 + 1023) / 1024)), dim3(256), 0, s, reinterpret_cast<double2 *>(dst),
                       reinterpret_cast<const double2 *>(src), n2);
    return hipGetLastError();
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s) {
    if (!n) return hipSuccess;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
    if (!n) return hipSuccess;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// Block-contiguous staging -> the 9 value planes: one lane per position.
__global__ __launch_bounds__(256) void k_blocks_to_planes(const double *__restrict__ Kb, double *__restrict__ Kv,
                                                          uint64_t ps) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < ps; p += (uint64_t)gridDim.x * blockDim.x) {
        double v[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) v[c] = Kb[9 * p + c];
#pragma unroll
        for (int c = 0; c < 9; ++c) Kv[c * ps + p] = v[c];
    }
}

template <int K, bool BLOCKS>
static hipError_t assemble_kind(Problem &P, int do_mat, int do_vec) {
    constexpr int TPW = 256 / ET<K>::TEAM;
    double *Kt = BLOCKS ? P.d_Kb : P.d_K;
    for (uint32_t c = 0; c < P.ecol.n_colors(); ++c) {
        const uint32_t b = P.ecol.color_ptr[c], n = P.ecol.color_ptr[c + 1] - b;
        if (!n) continue;
        hipLaunchKernelGGL((k_assemble<K, BLOCKS>), dim3((n + TPW - 1) / TPW), dim3(256), 0, P.stream, P.d_order + b, n,
                           P.d_conn, P.d_coords, P.d_E, P.d_elem_sub, P.d_mat, P.d_slot, Kt, P.ps, P.d_F, do_mat,
                           do_vec, P.d_flags);
    }
    if (BLOCKS && do_mat)
        hipLaunchKernelGGL(k_blocks_to_planes, dim3(8192), dim3(256), 0, P.stream, P.d_Kb, P.d_K, P.ps);
    return hipGetLastError();
}

template <int K>
static hipError_t assemble_gather_kind(Problem &P, int do_mat, int do_vec) {
    constexpr int TPW = 256 / ET<K>::TEAM;
    hipLaunchKernelGGL(k_elem_mats<K>, dim3((P.n_elem + TPW - 1) / TPW), dim3(256), 0, P.stream, P.n_elem, P.d_conn,
                       P.d_coords, P.d_E, P.d_elem_sub, P.d_mat, P.ga.d_em, P.ga.d_ef, do_mat, do_vec, P.d_flags);
    if (do_mat)
        hipLaunchKernelGGL(k_gather_blocks, dim3(8192), dim3(256), 0, P.stream, P.ps, P.ga.d_kptr, P.ga.d_kent, P.ga.d_em,
                           P.d_K);
    if (do_vec)
        hipLaunchKernelGGL(k_gather_loads, dim3((P.n_nodes + 255) / 256), dim3(256), 0, P.stream, P.n_nodes, P.ga.d_fptr,
                           P.ga.d_fent, P.ga.d_ef, P.d_F);
    return hipGetLastError();
}

static hipError_t assemble_gather_box(Problem &P, int do_mat, int do_vec) {
    const GatherAsm &G = P.ga;
    hipLaunchKernelGGL(k_box_ref, dim3(G.n_classes), dim3(64), 0, P.stream, G.d_ch, G.d_csub, P.d_mat, G.d_kref, G.d_fref,
                       P.d_flags);
    if (do_mat)
        hipLaunchKernelGGL(k_gather_blocks_box, dim3(8192), dim3(256), 0, P.stream, P.ps, G.d_kptr, G.d_kent, P.d_E,
                           G.d_cls, G.d_kref, P.d_K);
    if (do_vec)
        hipLaunchKernelGGL(k_gather_loads_box, dim3((P.n_nodes + 255) / 256), dim3(256), 0, P.stream, P.n_nodes, G.d_fptr,
                           G.d_fent, G.d_cls, G.d_fref, P.d_F);
    return hipGetLastError();
}

template <int K>
static void assemble_rows_kind(Problem &P, int do_mat, int do_vec) {
    hipLaunchKernelGGL(k_assemble_rows<K>, dim3(P.rg.n_groups), dim3(ASM_THREADS), 0, P.stream, P.rg.R, P.n_nodes,
                       P.rg.n_colors, P.d_gptr, P.d_gel, P.d_conn, P.d_coords, P.d_E, P.d_elem_sub, P.d_mat, P.d_slot,
                       P.d_sptr, P.d_perm, P.d_iperm, P.d_K, P.ps, P.d_F, do_mat, do_vec, P.d_flags);
}

uint32_t max_assembly_rows(const Problem &P) {
    uint32_t w = 0;
    for (uint32_t r = 0; r < P.n_nodes; ++r) w = std::max(w, P.pat.row_ptr[r + 1] - P.pat.row_ptr[r]);
    uint32_t R = 64;
    while (R > 0 && R * w > ASM_SLOTS) R /= 2;
    return R;
}

__global__ __launch_bounds__(256) void k_permute_nodes(double *__restrict__ dst, const double *__restrict__ src,
                                                       const uint32_t *__restrict__ map, uint32_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t m = map[i];
#pragma unroll
        for (int d = 0; d < 3; ++d) dst[3 * i + d] = src[3 * m + d];
    }
}

hipError_t launch_permute_nodes(double *dst, const double *src, const uint32_t *map, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_permute_nodes, dim3(std::min<uint32_t>(8192, (n + 255) / 256)), dim3(256), 0, s, dst, src, map, n);
    return hipGetLastError();
}

hipError_t launch_assemble(Problem &P, int do_mat, int do_vec) {
    hipError_t e = hipSuccess;
    bool body_force = false;
    for (const SubMaterial &m : P.mat_h) body_force |= m.bf0 != 0.0 || m.bf1 != 0.0 || m.bf2 != 0.0;
    if (!do_mat && !body_force) {
        // right-hand side only, no body force: every element contribution is exactly zero
        if (do_vec) e = launch_fill(P.d_F, 3 * (size_t)P.n_nodes, 0.0, P.stream);
    } else if (P.asm_rows == ASM_GATHER_BOX && P.ga.box) {
        e = assemble_gather_box(P, do_mat, do_vec);
    } else if (P.asm_rows == ASM_GATHER && P.ga.d_em) {
        if (P.kind == TETRA) e = assemble_gather_kind<TETRA>(P, do_mat, do_vec);
        else if (P.kind == TETRA10) e = assemble_gather_kind<TETRA10>(P, do_mat, do_vec);
        else e = assemble_gather_kind<HEXA>(P, do_mat, do_vec);
    } else if (P.asm_rows && P.rg.R == P.asm_rows) {
        if (P.kind == TETRA) assemble_rows_kind<TETRA>(P, do_mat, do_vec);
        else if (P.kind == TETRA10) assemble_rows_kind<TETRA10>(P, do_mat, do_vec);
        else assemble_rows_kind<HEXA>(P, do_mat, do_vec);
        e = hipGetLastError();
    } else {
        const bool blocks = P.asm_rows == ASM_STAGED && P.d_Kb;
        if (do_mat) e = launch_fill(blocks ? P.d_Kb : P.d_K, 9 * P.ps, 0.0, P.stream);
        if (e == hipSuccess && do_vec) e = launch_fill(P.d_F, 3 * (size_t)P.n_nodes, 0.0, P.stream);
        if (e != hipSuccess) return e;
        if (blocks) {
            if (P.kind == TETRA) e = assemble_kind<TETRA, true>(P, do_mat, do_vec);
            else if (P.kind == TETRA10) e = assemble_kind<TETRA10, true>(P, do_mat, do_vec);
            else e = assemble_kind<HEXA, true>(P, do_mat, do_vec);
        } else {
            if (P.kind == TETRA) e = assemble_kind<TETRA, false>(P, do_mat, do_vec);
            else if (P.kind == TETRA10) e = assemble_kind<TETRA10, false>(P, do_mat, do_vec);
            else e = assemble_kind<HEXA, false>(P, do_mat, do_vec);
        }
    }
    if (e != hipSuccess || !do_vec || !P.n_faces) return e;
    const int fk = face_kind_of(P.kind);
    for (uint32_t c = 0; c < P.fcol.n_colors(); ++c) {
        const uint32_t b = P.fcol.color_ptr[c], n = P.fcol.color_ptr[c + 1] - b;
        if (!n) continue;
        const dim3 g((n + 255) / 256);
        if (fk == TRI3) hipLaunchKernelGGL(k_face_load<TRI3>, g, dim3(256), 0, P.stream, P.d_forder + b, n, P.d_faces, P.d_trac, P.d_coords, P.d_F);
        else if (fk == TRI6) hipLaunchKernelGGL(k_face_load<TRI6>, g, dim3(256), 0, P.stream, P.d_forder + b, n, P.d_faces, P.d_trac, P.d_coords, P.d_F);
        else hipLaunchKernelGGL(k_face_load<QUAD4>, g, dim3(256), 0, P.stream, P.d_forder + b, n, P.d_faces, P.d_trac, P.d_coords, P.d_F);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Penalty constraints (formulation.h:1028-1070): M = max|diag K| over all scalar diagonal
// entries, taken after the element loop and before any constraint; then, per constrained DOF in
// insertion order, K(p,p) += M*P and F(p) += M*P*value.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_max_diag(const uint32_t *__restrict__ diag, const double *__restrict__ Kv,
                                                  uint64_t ps, WgTables wt, ProbScalars *S, double *partials) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB) return;
    double m[1] = {0.0};
    const uint32_t r = wt.row0[L] + threadIdx.x;
    if (r < wt.row1[L]) {
        const uint64_t p = diag[r];
        m[0] = fmax(fabs(Kv[p]), fmax(fabs(Kv[4 * ps + p]), fabs(Kv[8 * ps + p])));
    }
    wg_partial<1, 1u>(m, partials, gridDim.x, L);
}

__global__ __launch_bounds__(256) void k_fin_max_diag(ProbScalars *S, const double *__restrict__ partials, unsigned stride,
                                                      const uint32_t *__restrict__ sub_wg0,
                                                      const uint32_t *__restrict__ sub_nwg) {
    double o[1];
    wg_combine<1, 1u>(partials, stride, sub_wg0[blockIdx.x], sub_nwg[blockIdx.x], o);
    if (threadIdx.x == 0) S[blockIdx.x].max_diag = o[0];
}

__global__ void k_constraints(uint32_t n_cdof, const uint32_t *__restrict__ cdof, const uint32_t *__restrict__ cptr,
                              const uint32_t *__restrict__ csub, const double *__restrict__ cval,
                              const double *__restrict__ cpen, const uint32_t *__restrict__ diag, double *Kv, uint64_t ps,
                              double *F, const ProbScalars *S, int do_mat, int do_vec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_cdof) return;
    const double M = S[csub[i]].max_diag;
    const uint32_t dof = cdof[i], r = dof / 3, d = dof % 3;
    double *kp = &Kv[(uint64_t)(4 * d) * ps + diag[r]];
    for (uint32_t k = cptr[i]; k < cptr[i + 1]; ++k) {
        const double C = 1.0 * M * cpen[k];
        if (do_vec) F[dof] += C * cval[k];
        if (do_mat) *kp += C * 1.0;
    }
}

hipError_t launch_constraints(Problem &P, int do_mat, int do_vec) {
    const WgTables wt{P.d_wg_row0, P.d_wg_row1, P.d_wg_sub, P.d_sub_wg0, P.d_sub_nwg};
    if (do_mat) {
        hipLaunchKernelGGL(k_max_diag, dim3(P.n_wg), dim3(256), 0, P.stream, P.d_diag, P.d_K, P.ps, wt, P.d_ps, P.d_partials);
        hipLaunchKernelGGL(k_fin_max_diag, dim3(P.n_sub), dim3(256), 0, P.stream, P.d_ps, P.d_partials, P.n_wg, P.d_sub_wg0,
                           P.d_sub_nwg);
    }
    if (!P.n_cdof) return hipGetLastError();
    hipLaunchKernelGGL(k_constraints, dim3((P.n_cdof + 255) / 256), dim3(256), 0, P.stream, P.n_cdof, P.d_cdof, P.d_cptr,
                       P.d_csub, P.d_cval, P.d_cpen, P.d_diag, P.d_K, P.ps, P.d_F, P.d_ps, do_mat, do_vec);
    return hipGetLastError();
}

__global__ void k_to_fp32(const double *__restrict__ a, float *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = (float)a[i];
}

hipError_t launch_to_fp32(Problem &P) {
    hipLaunchKernelGGL(k_to_fp32, dim3(8192), dim3(256), 0, P.stream, P.d_K, P.d_K32, 9 * P.ps);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// dJ/dE_e = -lambda_e^T (dK_e/dE_e) u_e; K_e is linear in E_e, so dK_e/dE_e = K_e(E=1).
// Element-parallel, no scatter; the sum over the team's (a,b) blocks is a fixed-order LDS tree.
// ---------------------------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void k_elem_grad(uint32_t n_elem, const uint32_t *__restrict__ conn,
                                                   const double *__restrict__ coords,
                                                   const uint32_t *__restrict__ elem_sub,
                                                   const SubMaterial *__restrict__ mat, const double *__restrict__ u,
                                                   const double *__restrict__ lam, double *__restrict__ g, int *flags) {
    constexpr int NPE = ET<K>::NPE, TEAM = ET<K>::TEAM, TPW = 256 / TEAM;
    __shared__ ElemLds<K> lds[TPW];
    __shared__ double red[256];
    const int team = threadIdx.x / TEAM, t = threadIdx.x % TEAM;
    const uint32_t e = blockIdx.x * TPW + team;
    const bool valid = e < n_elem;
    ElemLds<K> &L = lds[team];
    elem_geometry<K>(L, t, valid, valid ? e : 0, conn, coords, flags);
    double s = 0;
    if (valid && t < NPE * NPE) {
        const int a = t / NPE, b = t % NPE;
        double H[3];
        gen::hooke(1.0, mat[elem_sub[e]].nu, H);
        double k[9];
        elem_block<K>(L, a, b, H, k);
        const uint32_t na = conn[(size_t)e * NPE + a], nb = conn[(size_t)e * NPE + b];
        for (int i = 0; i < 3; ++i) {
            const double ku = k[3 * i] * u[3 * (size_t)nb] + k[3 * i + 1] * u[3 * (size_t)nb + 1] + k[3 * i + 2] * u[3 * (size_t)nb + 2];
            s += lam[3 * (size_t)na + i] * ku;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = TEAM / 2; o > 0; o >>= 1) {
        if (t < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (valid && t == 0) g[e] = -red[threadIdx.x];
}

hipError_t launch_elem_grad(Problem &P, const double *d_u, const double *d_lam, double *d_g) {
    int *flags = P.d_flags;
    if (P.kind == TETRA) hipLaunchKernelGGL(k_elem_grad<TETRA>, dim3((P.n_elem + 15) / 16), dim3(256), 0, P.stream, P.n_elem, P.d_conn, P.d_coords, P.d_elem_sub, P.d_mat, d_u, d_lam, d_g, flags);
    else if (P.kind == TETRA10) hipLaunchKernelGGL(k_elem_grad<TETRA10>, dim3((P.n_elem + 1) / 2), dim3(256), 0, P.stream, P.n_elem, P.d_conn, P.d_coords, P.d_elem_sub, P.d_mat, d_u, d_lam, d_g, flags);
    else hipLaunchKernelGGL(k_elem_grad<HEXA>, dim3((P.n_elem + 3) / 4), dim3(256), 0, P.stream, P.n_elem, P.d_conn, P.d_coords, P.d_elem_sub, P.d_mat, d_u, d_lam, d_g, flags);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// After-solve recovery (call_after_solve, formulation.h:2026-2050, calling the generated
// after_solve of apply_on_elements_after_solve, formulations/formulation_elasticity_iso.py:64-85):
// eps = grad_sym_col(u), tr_epsilon = eps_xx + eps_yy + eps_zz and sigma = H(E_e, nu) eps, at
// xi = (1/2, 1/2, 1/2) for every element kind (each var_inter is set to 1/2; for Tetra and
// Tetra_10 that point lies outside the reference simplex, as in the reference).  One lane per
// element; element-major outputs eps[6e + i], sig[6e + i], tr[e] (elem.epsilon[0][i],
// elem.sigma[0][i], elem.tr_epsilon).
// ---------------------------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void k_after_solve(uint32_t n_elem, const uint32_t *__restrict__ conn,
                                                     const double *__restrict__ coords, const double *__restrict__ E,
                                                     const uint32_t *__restrict__ elem_sub,
                                                     const SubMaterial *__restrict__ mat, const double *__restrict__ u,
                                                     double *__restrict__ eps, double *__restrict__ sig,
                                                     double *__restrict__ tr) {
    constexpr int NPE = ET<K>::NPE;
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_elem) return;
    const double xi[3] = {0.5, 0.5, 0.5};
    uint32_t nd[NPE];
    double dN[NPE][3];
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        nd[a] = conn[(size_t)e * NPE + a];
        gen::vol<K>::shape(a, xi, dN[a]);
        const double x0 = coords[3 * (size_t)nd[a]], x1 = coords[3 * (size_t)nd[a] + 1], x2 = coords[3 * (size_t)nd[a] + 2];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            J[j][0] += dN[a][j] * x0;
            J[j][1] += dN[a][j] * x1;
            J[j][2] += dN[a][j] * x2;
        }
    }
    double Ji[3][3];
    jac_inverse(J, Ji);
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        double g[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) g[j] = Ji[j][0] * dN[a][0] + Ji[j][1] * dN[a][1] + Ji[j][2] * dN[a][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double ui = u[3 * (size_t)nd[a] + i];
#pragma unroll
            for (int j = 0; j < 3; ++j) G[3 * i + j] += ui * g[j];
        }
    }
    double ev[6], sv[6], H[3];
    gen::strain(G, ev);
    gen::hooke(E[e], mat[elem_sub[e]].nu, H);
    gen::stress(H, ev, sv);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        if (eps) eps[6 * (size_t)e + i] = ev[i];
        if (sig) sig[6 * (size_t)e + i] = sv[i];
    }
    if (tr) tr[e] = ev[0] + ev[1] + ev[2];
}

hipError_t launch_after_solve(Problem &P, const double *d_u, double *d_eps, double *d_sig, double *d_tr) {
    const dim3 g((P.n_elem + 255) / 256), b(256);
    if (P.kind == TETRA) hipLaunchKernelGGL(k_after_solve<TETRA>, g, b, 0, P.stream, P.n_elem, P.d_conn, P.d_coords, P.d_E, P.d_elem_sub, P.d_mat, d_u, d_eps, d_sig, d_tr);
    else if (P.kind == TETRA10) hipLaunchKernelGGL(k_after_solve<TETRA10>, g, b, 0, P.stream, P.n_elem, P.d_conn, P.d_coords, P.d_E, P.d_elem_sub, P.d_mat, d_u, d_eps, d_sig, d_tr);
    else hipLaunchKernelGGL(k_after_solve<HEXA>, g, b, 0, P.stream, P.n_elem, P.d_conn, P.d_coords, P.d_E, P.d_elem_sub, P.d_mat, d_u, d_eps, d_sig, d_tr);
    return hipGetLastError();
}

}  // namespace femgpu
