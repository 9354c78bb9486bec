// Preconditioned CG on the 3x3-block SELL-64 matrix, FP64 or FP32-inner with FP64 residual
// refreshes, for one or many independent sub-problems at once.
//
// The iteration is the PCG of containers/MatWithTinyBlocks.h:1101-1151 /
// containers/conjugate_gradient.h:27-68:  q = K d; alpha = rho / d.q; x += alpha d;
// r -= alpha q; z = M^-1 r; rho' = r.z; beta = rho'/rho; d = z + beta d, with M = scalar or 3x3
// block Jacobi (fused into the vector updates), the block IC(0) factor of ic0.hip (level
// sweeps between the update and the r.z reduction) or a degree-2 Chebyshev polynomial in the
// block-Jacobi operator (a second product per iteration, k_cheb).
//
// Every row-parallel kernel uses the same workgroup tables: logical workgroup L covers rows
// [wg_row0[L], wg_row1[L]) of sub-problem wg_sub[L].  Reductions are segmented by sub-problem and
// deterministic: a wave-shuffle tree and a fixed 4-wave sum give one partial per workgroup, and
// k_finalize (one workgroup per sub-problem, launched after the row kernel) combines that
// sub-problem's partials in a fixed order and advances its scalars -- bitwise reproducible, no
// fences or atomics.  Workgroups of finished or paused sub-problems return at once.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.hpp"

namespace femgpu {

enum { MODE_PLAIN = 0, MODE_PQ = 1, MODE_RES = 2 };

// ---------------------------------------------------------------------------------------------
// Block SpMV on SELL-64 storage: one lane per storage row; the 64 lanes of a wave are 64
// consecutive storage rows, so every value-plane and column load of slot k is one contiguous run.
// Rows shorter than their slice stop early (masked lanes fetch nothing).  FMT_FULL storage row t
// holds block row perm[t] (rows sorted by length within windows of whole workgroups, so a
// workgroup's storage rows are block rows of its own sub-problem); FMT_SYM keeps t = r.
// FMT_FULL: every block of the row is stored in the row (row lengths from row_ptr).
// FMT_SYM: the row stores its upper blocks (c >= r); each lower block K_rc = K_cr^T is read from
// row c's upper storage through tposL, as MatWithTinyBlocks::partial_mul uses every stored block
// for both rows (containers/MatWithTinyBlocks.h:303-378).
// MODE_PQ: y = K x and the partial of p.q.  MODE_RES: y = b - K x (FP64) with partials of ||y||^2
// and max|y|.  The matching k_finalize turns the partials into the sub-problem's scalars.
// ---------------------------------------------------------------------------------------------
struct SpmvArgs {
    const uint32_t *len;    // FMT_FULL: row_ptr; FMT_SYM: upper lengths
    const uint32_t *sptr, *col;
    uint64_t ps;
    const uint32_t *sptrL, *lenL, *colL, *tposL;  // FMT_SYM lower part
    const int16_t *col16;   // LOAD_NT16: column offsets c - r
    const uint32_t *perm;   // FMT_FULL: block row held by storage row t (null: t)
};

// NT: the matrix values and columns are read with non-temporal loads (streamed once; L2 is left
// to the x gathers).  C16: columns come as 16-bit offsets c - r.
template <bool NT, typename V>
__device__ inline V ld(const V *p) {
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <typename T, int UNROLL, bool NT, bool C16>
__device__ inline void spmv_rows(const uint32_t *__restrict__ colS, const int16_t *__restrict__ col16, uint32_t r,
                                 const T *__restrict__ Kv, uint64_t ps, uint64_t base, uint32_t len,
                                 const T *__restrict__ x, T &s0, T &s1, T &s2) {
    uint32_t k = 0;
    for (; k + UNROLL <= len; k += UNROLL) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const uint64_t p = base + 64ull * (k + u);
            const uint32_t c = C16 ? r + (int)ld<NT>(col16 + p) : ld<NT>(colS + p);
            const T a0 = ld<NT>(Kv + p), a1 = ld<NT>(Kv + ps + p), a2 = ld<NT>(Kv + 2 * ps + p);
            const T a3 = ld<NT>(Kv + 3 * ps + p), a4 = ld<NT>(Kv + 4 * ps + p), a5 = ld<NT>(Kv + 5 * ps + p);
            const T a6 = ld<NT>(Kv + 6 * ps + p), a7 = ld<NT>(Kv + 7 * ps + p), a8 = ld<NT>(Kv + 8 * ps + p);
            const T x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
            s0 += a0 * x0 + a1 * x1 + a2 * x2;
            s1 += a3 * x0 + a4 * x1 + a5 * x2;
            s2 += a6 * x0 + a7 * x1 + a8 * x2;
        }
    }
    for (; k < len; ++k) {
        const uint64_t p = base + 64ull * k;
        const uint32_t c = C16 ? r + (int)ld<NT>(col16 + p) : ld<NT>(colS + p);
        const T x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
        s0 += ld<NT>(Kv + p) * x0 + ld<NT>(Kv + ps + p) * x1 + ld<NT>(Kv + 2 * ps + p) * x2;
        s1 += ld<NT>(Kv + 3 * ps + p) * x0 + ld<NT>(Kv + 4 * ps + p) * x1 + ld<NT>(Kv + 5 * ps + p) * x2;
        s2 += ld<NT>(Kv + 6 * ps + p) * x0 + ld<NT>(Kv + 7 * ps + p) * x1 + ld<NT>(Kv + 8 * ps + p) * x2;
    }
}

// lower blocks of FMT_SYM: K_rc x_c = K_cr^T x_c with K_cr at position tposL in the upper part
template <typename T, int UNROLL>
__device__ inline void spmv_rows_t(const uint32_t *__restrict__ colL, const uint32_t *__restrict__ tposL,
                                   const T *__restrict__ Kv, uint64_t ps, uint64_t base, uint32_t len,
                                   const T *__restrict__ x, T &s0, T &s1, T &s2) {
    uint32_t k = 0;
    for (; k + UNROLL <= len; k += UNROLL) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const uint64_t pl = base + 64ull * (k + u);
            const uint32_t c = colL[pl];
            const uint64_t p = tposL[pl];
            const T a0 = Kv[p], a1 = Kv[ps + p], a2 = Kv[2 * ps + p];
            const T a3 = Kv[3 * ps + p], a4 = Kv[4 * ps + p], a5 = Kv[5 * ps + p];
            const T a6 = Kv[6 * ps + p], a7 = Kv[7 * ps + p], a8 = Kv[8 * ps + p];
            const T x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
            s0 += a0 * x0 + a3 * x1 + a6 * x2;
            s1 += a1 * x0 + a4 * x1 + a7 * x2;
            s2 += a2 * x0 + a5 * x1 + a8 * x2;
        }
    }
    for (; k < len; ++k) {
        const uint64_t pl = base + 64ull * k;
        const uint32_t c = colL[pl];
        const uint64_t p = tposL[pl];
        const T x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
        s0 += Kv[p] * x0 + Kv[3 * ps + p] * x1 + Kv[6 * ps + p] * x2;
        s1 += Kv[ps + p] * x0 + Kv[4 * ps + p] * x1 + Kv[7 * ps + p] * x2;
        s2 += Kv[2 * ps + p] * x0 + Kv[5 * ps + p] * x1 + Kv[8 * ps + p] * x2;
    }
}

template <typename T, int MODE, int UNROLL, int FMT, int LD>
__global__ __launch_bounds__(256) void k_spmv(SpmvArgs A, const T *__restrict__ Kv, const T *__restrict__ x,
                                              const double *b, T *__restrict__ y, WgTables wt, const ProbScalars *S,
                                              double *partials, int sel) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB) return;
    if (MODE != MODE_PLAIN && !selected(S[s], MODE == MODE_PQ ? (int)SEL_ACTIVE : sel)) return;
    const uint32_t t = wt.row0[L] + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (t < wt.row1[L]) {
        T s0 = 0, s1 = 0, s2 = 0;
        const uint64_t off = t & 63u;
        const uint32_t r = FMT == FMT_FULL && A.perm ? A.perm[t] : t;
        if (FMT == FMT_FULL) {
            spmv_rows<T, UNROLL, LD != LOAD_PLAIN, LD == LOAD_NT16>(A.col, A.col16, r, Kv, A.ps, A.sptr[t >> 6] + off,
                                                                    A.len[r + 1] - A.len[r], x, s0, s1, s2);
        } else {
            spmv_rows<T, UNROLL, false, false>(A.col, nullptr, r, Kv, A.ps, A.sptr[r >> 6] + off, A.len[r], x, s0, s1, s2);
            spmv_rows_t<T, UNROLL>(A.colL, A.tposL, Kv, A.ps, A.sptrL[r >> 6] + off, A.lenL[r], x, s0, s1, s2);
        }
        const T sv[3] = {s0, s1, s2};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t i = 3 * (size_t)r + d;
            if (MODE == MODE_RES) {
                const double v = b[i] - (double)sv[d];
                y[i] = (T)v;
                acc[0] += v * v;
                acc[1] = fmax(acc[1], fabs(v));
            } else {
                y[i] = sv[d];
                if (MODE == MODE_PQ) acc[0] += (double)sv[d] * (double)x[i];
            }
        }
    }
    if (MODE == MODE_PLAIN) return;
    if (MODE == MODE_PQ) {
        double v1[1] = {acc[0]};
        wg_partial<1, 0u>(v1, partials, gridDim.x, L);
    } else {
        wg_partial<2, 2u>(acc, partials, gridDim.x, L);
    }
}

// y = b - y for y holding K x (the residual of MODE_RES after a matrix-free product), with the
// partials of ||y||^2 and max |y| that k_spmv's MODE_RES writes
__global__ __launch_bounds__(256) void k_residual(const double *__restrict__ b, double *__restrict__ y, WgTables wt,
                                                  const ProbScalars *S, double *partials, int sel) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB || !selected(S[s], sel)) return;
    double acc[2] = {0.0, 0.0};
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n < wt.row1[L]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t i = 3 * (size_t)n + d;
            const double v = b[i] - y[i];
            y[i] = v;
            acc[0] += v * v;
            acc[1] = fmax(acc[1], fabs(v));
        }
    }
    wg_partial<2, 2u>(acc, partials, gridDim.x, L);
}

// Values of the symmetric storage from the full K: plane q of upper position p is K[q][src[p]]
// (zero at padding positions).
template <typename T>
__global__ void k_to_sym(const double *__restrict__ K, uint64_t ps, const uint32_t *__restrict__ src, T *__restrict__ Ku,
                         uint64_t psU) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < psU; p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = src[p];
#pragma unroll
        for (int j = 0; j < 9; ++j) Ku[j * psU + p] = q == 0xffffffffu ? (T)0 : (T)K[j * ps + q];
    }
}

// z = M^-1 r for one node; PRE 0: scalar Jacobi (3 values per node), 1: 3x3 block Jacobi, the
// inverse of a symmetric diagonal block stored as its 6 upper values (m00 m01 m02 m11 m12 m22),
// 2: block IC(0), whose z the level sweeps write (not applied here), 3: the Chebyshev polynomial
// z = c0 D r + c1 D K D r in the block-Jacobi D (D r here; k_cheb adds the second term).
enum { PRE_JACOBI = 0, PRE_BLOCK_JACOBI = 1, PRE_IC0 = 2, PRE_CHEB = 3 };
// FP32 vector update of the two-kernel CG iteration: lanes per 1024-row workgroup (1024 / UPD_NT32
// rows each) and p.q partials loaded per lane in one round of the combine (Hexa8 100^3, us per
// CG iteration: 256 lanes and 4 partials 55.6, 8 partials 55.6, 16 60.3; 512 lanes and 4
// partials 55.1, 8 partials 58.0; 1024 lanes 59.3.  512 lanes: MIXED solve 84.7-85.7 against
// 85.7-86.5 ms with 256, the batched 256-cube solve 141 against 148 ms)
#ifndef UPD_NT32
#define UPD_NT32 512
#endif
#ifndef UPD_CPL32
#define UPD_CPL32 4
#endif
// the block-Jacobi factors of D serve both
template <int PRE>
constexpr int jacobi_kind() { return PRE == PRE_CHEB ? PRE_BLOCK_JACOBI : PRE; }
// PRE_CHEB in the vector updates and the restart: z = D r (k_cheb forms the preconditioned
// residual and its partials afterwards, so these write none)
// Stored preconditioner values: FP64 solves keep them in FP64.  The FP32 inner iterations of a
// MIXED solve keep the block-Jacobi inverse blocks (the upper triangle, 6 values a node) and the
// Jacobi reciprocals rounded to bfloat16: 12 B a node instead of 24 in the vector update, which
// streams 11% fewer bytes (Hexa8 100^3: 110 -> 97 MB a launch).  bfloat16 keeps FP32's exponent
// range, so no value under- or overflows; a rounded block is a relative perturbation of at most
// 2^-9 of each entry, which keeps an SPD 3x3 block SPD while its condition number is below
// 2^9 / sqrt(3) = 295, and k_build_minv checks the rounded block's leading minors and stores the
// node's point-Jacobi reciprocals instead when one is not positive.  The preconditioner only
// steers the inner iterations; the FP64 refreshes set the solution's accuracy.
template <typename T>
struct MStore { using type = T; };
template <>
struct MStore<float> { using type = uint16_t; };
template <typename T>
using mstore_t = typename MStore<T>::type;

// bfloat16 bits of a finite float, rounded to nearest even
__device__ inline uint16_t bf16_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
template <typename T>
__device__ inline T mval(const mstore_t<T> *__restrict__ Minv, size_t i) {
    if constexpr (sizeof(T) == 4) return bf16_value(Minv[i]);
    else return Minv[i];
}

// The preconditioner's values of node n (3 for Jacobi, 6 for block Jacobi), loaded ahead of their use.
template <int PRE, typename T>
__device__ inline void load_m(const mstore_t<T> *__restrict__ Minv, size_t n, T (&m)[6]) {
    if (PRE == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) m[c] = mval<T>(Minv, 3 * n + c);
    } else if (PRE == PRE_BLOCK_JACOBI || PRE == PRE_CHEB) {
        if constexpr (sizeof(T) == 4) {
            // a node's 6 values are 12 bytes at a 4-byte aligned offset: 3 dword loads
            const uint32_t *w = reinterpret_cast<const uint32_t *>(Minv + 6 * n);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const uint32_t v = w[c];
                m[2 * c] = bf16_value((uint16_t)(v & 0xffffu));
                m[2 * c + 1] = bf16_value((uint16_t)(v >> 16));
            }
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) m[c] = Minv[6 * n + c];
        }
    }
}

// z = M^-1 r from the loaded values, the arithmetic of apply_m.
template <int PRE, typename T>
__device__ inline void apply_mv(const T (&m)[6], const T r[3], T z[3]) {
    if (PRE == 0) {
        z[0] = m[0] * r[0]; z[1] = m[1] * r[1]; z[2] = m[2] * r[2];
    } else if (PRE == PRE_BLOCK_JACOBI || PRE == PRE_CHEB) {
        z[0] = m[0] * r[0] + m[1] * r[1] + m[2] * r[2];
        z[1] = m[1] * r[0] + m[3] * r[1] + m[4] * r[2];
        z[2] = m[2] * r[0] + m[4] * r[1] + m[5] * r[2];
    }
}

template <int PRE, typename T>
__device__ inline void apply_m(const mstore_t<T> *__restrict__ Minv, size_t n, const T r[3], T z[3]) {
    if (PRE == 0) {
        z[0] = mval<T>(Minv, 3 * n) * r[0]; z[1] = mval<T>(Minv, 3 * n + 1) * r[1]; z[2] = mval<T>(Minv, 3 * n + 2) * r[2];
    } else if (PRE == PRE_BLOCK_JACOBI || PRE == PRE_CHEB) {
        T m[6];
        load_m<PRE, T>(Minv, n, m);
        z[0] = m[0] * r[0] + m[1] * r[1] + m[2] * r[2];
        z[1] = m[1] * r[0] + m[3] * r[1] + m[4] * r[2];
        z[2] = m[2] * r[0] + m[4] * r[1] + m[5] * r[2];
    }
}

// One node of the vector update: x += alpha p; r -= alpha q; z = M^-1 r; partials of r.z, r.r,
// max|r| accumulated into acc.
template <int PRE, typename T>
__device__ inline void update_node(uint32_t n, T alpha, T *__restrict__ x, T *__restrict__ r, const T *p,
                                   const T (&qn)[3], const mstore_t<T> *__restrict__ Minv, T *__restrict__ z, double (&acc)[3]) {
    T rv[3], zv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t i = 3 * (size_t)n + d;
        x[i] += alpha * p[i];
        rv[d] = r[i] - alpha * qn[d];
        r[i] = rv[d];
    }
    if (PRE != PRE_IC0) apply_m<PRE, T>(Minv, n, rv, zv);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (PRE != PRE_IC0) {
            z[3 * (size_t)n + d] = zv[d];
            acc[0] += (double)rv[d] * (double)zv[d];
        }
        acc[1] += (double)rv[d] * (double)rv[d];
        acc[2] = fmax(acc[2], fabs((double)rv[d]));
    }
}

// x += alpha p; r -= alpha q; z = M^-1 r (stored for k_update_p); partials of r.z, r.r, max|r|
// (k_finalize FIN_UPD advances the scalars and applies the stop test).  IC(0): z needs the
// sweeps, so r.z is left to k_dot_rz (partial 0 stays zero, FIN_UPD_IC).  p_alt (direction
// update fused into a matrix-free product): the product of an even iteration stored the
// direction in p_alt, of an odd one in p.
template <int PRE, typename T>
__global__ __launch_bounds__(256) void k_update(T *__restrict__ x, T *__restrict__ r, const T *p, const T *p_alt,
                                                const T *__restrict__ q, const mstore_t<T> *__restrict__ Minv, T *__restrict__ z,
                                                WgTables wt, const ProbScalars *S, double *partials) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB || !selected(S[s], SEL_ACTIVE)) return;
    if (p_alt && !(S[s].iter & 1)) p = p_alt;
    const T alpha = (T)S[s].alpha;
    double acc[3] = {0, 0, 0};
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n < wt.row1[L]) {
        const T qn[3] = {q[3 * (size_t)n], q[3 * (size_t)n + 1], q[3 * (size_t)n + 2]};
        update_node<PRE, T>(n, alpha, x, r, p, qn, Minv, z, acc);
    }
    if (PRE != PRE_CHEB) wg_partial<3, 4u>(acc, partials, gridDim.x, L);
}

// The same after a matrix-free product with the fused direction update, with the product's
// shared-node pass folded in: every workgroup combines its sub-problem's p.q partials itself (in
// the fixed order of wg_combine, so every workgroup of a sub-problem holds the same alpha; the
// sub-problem's first workgroup records it), and a node shared by several blocks takes q as the
// sum of its partials halo[hbeg[n] .. hbeg[n + 1]) in block order, any other node the product's y.
// FOLD (two-kernel iteration): the scalars are read from Sin, the copy the element blocks wrote,
// and the first workgroup writes them back into Sout with p.q recorded and the partials of this
// update marked pending (fin_pending: the next element-block kernel applies them); else Sin and
// Sout are the same array and only p.q is recorded.  FOLD workgroups take the tables of 1024-row
// workgroups: a quarter of the partials for every element block to combine, and a quarter of the
// workgroups combining the blocks' p.q partials.
template <int PRE, typename T, bool FOLD>
__global__ __launch_bounds__(FOLD ? (sizeof(T) == 8 ? 1024 : UPD_NT32) : 256, FOLD && sizeof(T) == 8 ? 8 : 1) void k_update_mf(T *__restrict__ x, T *__restrict__ r, const T *p, const T *p_alt,
                                                   const T *__restrict__ q, const T *__restrict__ halo,
                                                   const uint32_t *__restrict__ hbeg, const double *__restrict__ mfpart,
                                                   const uint32_t *__restrict__ mf_w0, const uint32_t *__restrict__ mf_nw,
                                                   const mstore_t<T> *__restrict__ Minv, T *__restrict__ z, WgTables wt,
                                                   const ProbScalars *Sin, ProbScalars *Sout, double *partials) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB) return;
    const ProbScalars *S = Sin;
    const bool lead = L == wt.sub_wg0[s];
    if (!selected(S[s], SEL_ACTIVE)) {
        if (FOLD && lead) copy_scalars(Sout + s, Sin + s);
        return;
    }
    if (FOLD && lead) copy_scalars(Sout + s, Sin + s);  // ordered before the record below by the combine's barrier
    if (!(S[s].iter & 1)) p = p_alt;
    // the node's operands: q (a shared node: the sum of its partials in block order), x, r, the
    // direction, the preconditioner's values
    auto load = [&](uint32_t n, T (&qn)[3], T (&xv)[3], T (&rv)[3], T (&pv)[3], T (&mv)[6]) {
        load_m<PRE, T>(Minv, n, mv);
        const uint32_t k0 = hbeg[n], k1 = hbeg[n + 1];
        if (k0 == k1) {
#pragma unroll
            for (int d = 0; d < 3; ++d) qn[d] = q[3 * (size_t)n + d];
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) qn[d] = 0;
            for (uint32_t k = k0; k < k1; ++k) {
                qn[0] += halo[3 * (size_t)k];
                qn[1] += halo[3 * (size_t)k + 1];
                qn[2] += halo[3 * (size_t)k + 2];
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xv[d] = x[3 * (size_t)n + d];
            rv[d] = r[3 * (size_t)n + d];
            pv[d] = p[3 * (size_t)n + d];
        }
    };
    auto update = [&](uint32_t n, T alpha, const T (&qn)[3], const T (&xv)[3], T (&rv)[3], const T (&pv)[3],
                      const T (&mv)[6], double (&acc)[3]) {
        T zv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x[3 * (size_t)n + d] = xv[d] + alpha * pv[d];
            rv[d] = rv[d] - alpha * qn[d];
            r[3 * (size_t)n + d] = rv[d];
        }
        apply_mv<PRE, T>(mv, rv, zv);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            z[3 * (size_t)n + d] = zv[d];
            acc[0] += (double)rv[d] * (double)zv[d];
            acc[1] += (double)rv[d] * (double)rv[d];
            acc[2] = fmax(acc[2], fabs((double)rv[d]));
        }
    };
    auto record = [&](double pq) {
        if (lead && threadIdx.x == 0) {
            finish_pq(Sin[s], Sout[s], pq);
            if (FOLD) Sout[s].fin_pending = Sout[s].done ? 0 : 1;
        }
    };
    double o[1];
    double acc[3] = {0, 0, 0};
    // FOLD: 1024 rows per workgroup, as 1024 lanes of one row (FP64) or UPD_NT32 lanes of
    // 1024 / UPD_NT32 rows, rows row0 + lane + UPD_NT32 j (FP32); measured at Hexa8 100^3 against
    // 256-row workgroups: FP64 45.9 against 51.4 us, FP32 (256 lanes) 24.1 against 25.5 us
    constexpr int NT = FOLD ? (sizeof(T) == 8 ? 1024 : UPD_NT32) : 256;
    constexpr int RPL = FOLD ? 1024 / NT : 1;
    // FP32: the p.q partials (one per element block) in one round of loads per lane
    constexpr int CPL = FOLD && sizeof(T) == 4 ? UPD_CPL32 : 4;
    if constexpr (RPL == 1) {
        // FP64: the node's operands are loaded before the combine of the p.q partials (which alpha
        // needs), so that their latency overlaps it (48.2 against 50.1 us per launch at Hexa8
        // 100^3); FP32 combines first (25.6 against 27.7 us: memory returns in order, and the
        // combine's L2 reads then wait behind the operand loads)
        const uint32_t n = wt.row0[L] + threadIdx.x;
        const bool valid = n < wt.row1[L];
        T qn[3] = {0, 0, 0}, xv[3] = {0, 0, 0}, rv[3] = {0, 0, 0}, pv[3] = {0, 0, 0}, mv[6] = {0, 0, 0, 0, 0, 0};
        if (sizeof(T) == 4) wg_combine<1, 0u>(mfpart, 0u, mf_w0[s], mf_nw[s], o);
        if (valid) load(n, qn, xv, rv, pv, mv);
        if (sizeof(T) == 8) wg_combine<1, 0u>(mfpart, 0u, mf_w0[s], mf_nw[s], o);
        record(o[0]);
        if (!finite(o[0]) || !(o[0] > 0)) return;  // recorded by workgroup 0: the solve stops
        const T alpha = (T)(S[s].rho / o[0]);
        if (valid) update(n, alpha, qn, xv, rv, pv, mv, acc);
    } else {
        // FP32: the operands of all RPL rows are loaded before the combine of the p.q partials
        // (56.5 against 57.8 us per CG iteration at Hexa8 100^3; 2 rows ahead, or 4 loaded right
        // after the combine, 57.1-57.8)
        T qn[RPL][3], xv[RPL][3], rv[RPL][3], pv[RPL][3], mv[RPL][6];
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
            const uint32_t n = wt.row0[L] + threadIdx.x + (uint32_t)NT * j;
            if (n < wt.row1[L]) load(n, qn[j], xv[j], rv[j], pv[j], mv[j]);
        }
        wg_combine<1, 0u, CPL>(mfpart, 0u, mf_w0[s], mf_nw[s], o);
        record(o[0]);
        if (!finite(o[0]) || !(o[0] > 0)) return;
        const T alpha = (T)(S[s].rho / o[0]);
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
            const uint32_t n = wt.row0[L] + threadIdx.x + (uint32_t)NT * j;
            if (n < wt.row1[L]) update(n, alpha, qn[j], xv[j], rv[j], pv[j], mv[j], acc);
        }
    }
    if (PRE != PRE_CHEB) {
        if constexpr (NT == 1024) wg_partial_n<3, 4u, 16>(acc, partials, gridDim.x, L);
        else if constexpr (NT == 512) wg_partial_n<3, 4u, 8>(acc, partials, gridDim.x, L);
        else wg_partial<3, 4u>(acc, partials, gridDim.x, L);
    }
}

// p = z + beta p, or p = z when first, with z = M^-1 r as the update, the restart or the IC(0)
// sweeps stored it (sel: SEL_ACTIVE in the iteration, SEL_PAUSE_DONE after a refresh, SEL_ALL
// for the first direction of IC(0))
template <typename T>
__global__ __launch_bounds__(256) void k_update_p(const T *__restrict__ z, T *__restrict__ p, WgTables wt,
                                                  const ProbScalars *S, int sel, int first) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB || !selected(S[s], sel)) return;
    const T beta = (T)S[s].beta;
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n >= wt.row1[L]) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t i = 3 * (size_t)n + d;
        p[i] = first ? z[i] : dir_value(z[i], beta, p[i]);
    }
}

// Start (first = 1, every unfinished sub-problem) or restart after an FP64 residual refresh
// (first = 0, sub-problems with a pending refresh) of the recurrence from the FP64 residual r64:
// r = (T) r64; z = M^-1 r; first: p = z, rho = r.z;  refresh: rho = r.z, beta = rho / rho_prev,
// pause -> PAUSE_DONE (the direction update follows).  IC(0): only r = (T) r64 here; the sweeps,
// k_dot_rz and the direction update follow.
template <int PRE, typename T>
__global__ __launch_bounds__(256) void k_restart(const double *r64, T *r, T *__restrict__ p,
                                                 const mstore_t<T> *__restrict__ Minv, T *__restrict__ z, WgTables wt,
                                                 const ProbScalars *S, double *partials, int first) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB || !selected(S[s], first ? (int)SEL_ALL : (int)SEL_PAUSE_REQ)) return;
    double acc[1] = {0};
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n < wt.row1[L]) {
        T rv[3], zv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) { rv[d] = (T)r64[3 * (size_t)n + d]; r[3 * (size_t)n + d] = rv[d]; }
        if (PRE != PRE_IC0) {
            apply_m<PRE, T>(Minv, n, rv, zv);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                z[3 * (size_t)n + d] = zv[d];
                if (first) p[3 * (size_t)n + d] = zv[d];
                acc[0] += (double)rv[d] * (double)zv[d];
            }
        }
    }
    if (PRE != PRE_IC0 && PRE != PRE_CHEB) wg_partial<1, 0u>(acc, partials, gridDim.x, L);
}

// Partials of r.z (IC(0): after the sweeps); finalized by FIN_RHO, FIN_RESTART_FIRST or
// FIN_RESTART.
template <typename T>
__global__ __launch_bounds__(256) void k_dot_rz(const T *__restrict__ r, const T *__restrict__ z, WgTables wt,
                                                const ProbScalars *S, double *partials, int sel) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB || !selected(S[s], sel)) return;
    double acc[1] = {0};
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n < wt.row1[L]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[0] += (double)r[3 * (size_t)n + d] * (double)z[3 * (size_t)n + d];
    }
    wg_partial<1, 0u>(acc, partials, gridDim.x, L);
}

// Second stage of the Chebyshev preconditioner: with v = K y for y = D r (y stored in z by the
// vector update or the restart; v in q, or for a node shared by several matrix-free blocks the sum
// of its partials halo[hbeg[n] .. hbeg[n + 1]) when hbeg is given), z = c0 y + c1 D v =
// D (c0 r + c1 v), written over y.  first: also p = z.  NV 3 (iteration): partials of r.z, r.r,
// max|r| for FIN_UPD; NV 1 (restart): of r.z.
template <typename T, int NV>
__global__ __launch_bounds__(256) void k_cheb(const T *__restrict__ r, T *__restrict__ z, const T *__restrict__ q,
                                              const T *__restrict__ halo, const uint32_t *__restrict__ hbeg,
                                              const mstore_t<T> *__restrict__ Minv, T *__restrict__ p, WgTables wt,
                                              const ProbScalars *S, double *partials, int sel, int first) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB || !selected(S[s], sel)) return;
    const T c0 = (T)S[s].cheb_c0, c1 = (T)S[s].cheb_c1;
    double acc[3] = {0, 0, 0};
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n < wt.row1[L]) {
        T m[6], v[3] = {0, 0, 0}, rv[3], w[3], zv[3];
        load_m<PRE_BLOCK_JACOBI, T>(Minv, n, m);
        const uint32_t k0 = hbeg ? hbeg[n] : 0u, k1 = hbeg ? hbeg[n + 1] : 0u;
        if (k0 == k1) {
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = q[3 * (size_t)n + d];
        } else {
            for (uint32_t k = k0; k < k1; ++k) {
                v[0] += halo[3 * (size_t)k];
                v[1] += halo[3 * (size_t)k + 1];
                v[2] += halo[3 * (size_t)k + 2];
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            rv[d] = r[3 * (size_t)n + d];
            w[d] = c0 * rv[d] + c1 * v[d];
        }
        apply_mv<PRE_BLOCK_JACOBI, T>(m, w, zv);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t i = 3 * (size_t)n + d;
            z[i] = zv[d];
            if (first) p[i] = zv[d];
            acc[0] += (double)rv[d] * (double)zv[d];
            acc[1] += (double)rv[d] * (double)rv[d];
            acc[2] = fmax(acc[2], fabs((double)rv[d]));
        }
    }
    if (NV == 3) {
        wg_partial<3, 4u>(acc, partials, gridDim.x, L);
    } else {
        double a1[1] = {acc[0]};
        wg_partial<1, 0u>(a1, partials, gridDim.x, L);
    }
}

// Start vector of the power iteration for the largest eigenvalue of D K: a fixed pseudo-random
// value in [-1, 1) per DOF (the MurmurHash3 32-bit finalizer of its index within its
// sub-problem, so a sub-problem of a batch starts as it would alone), with the partials
// (0, 0, ||v||^2) for FIN_POWER.
__device__ inline double unit_hash(uint32_t i) {
    uint32_t h = i * 0x9e3779b1u;
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return (double)h * (2.0 / 4294967296.0) - 1.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_power_init(T *__restrict__ v, WgTables wt, double *partials) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB) return;
    double acc[3] = {0, 0, 0};
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n < wt.row1[L]) {
        const uint32_t local = n - wt.row0[wt.sub_wg0[s]];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T h = (T)unit_hash(3 * local + d);
            v[3 * (size_t)n + d] = h;
            acc[2] += (double)h * (double)h;
        }
    }
    wg_partial<3, 0u>(acc, partials, gridDim.x, L);
}

// One power step from u = K v: partials of v.u and u.D u (their ratio, the Rayleigh quotient of
// K^1/2 D K^1/2 at K^1/2 v, bounds the largest eigenvalue of D K from below) and v <- D u / ||v||
// with the partial of its squared norm.
template <typename T>
__global__ __launch_bounds__(256) void k_power(T *__restrict__ v, const T *__restrict__ u, const mstore_t<T> *__restrict__ Minv,
                                               WgTables wt, const ProbScalars *S, double *partials) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB) return;
    const T inv = (T)(1.0 / S[s].pw_norm);
    double acc[3] = {0, 0, 0};
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n < wt.row1[L]) {
        T m[6], un[3], w[3];
        load_m<PRE_BLOCK_JACOBI, T>(Minv, n, m);
#pragma unroll
        for (int d = 0; d < 3; ++d) un[d] = u[3 * (size_t)n + d];
        apply_mv<PRE_BLOCK_JACOBI, T>(m, un, w);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t i = 3 * (size_t)n + d;
            acc[0] += (double)v[i] * (double)un[d];
            acc[1] += (double)un[d] * (double)w[d];
            const T vn = w[d] * inv;
            v[i] = vn;
            acc[2] += (double)vn * (double)vn;
        }
    }
    wg_partial<3, 0u>(acc, partials, gridDim.x, L);
}

// Finalize of a row kernel's partials, one workgroup per sub-problem (grid = n_sub); selects the
// same sub-problems as the row kernel it follows (their flags are unchanged in between).
// FIN_UPD_IC: as FIN_UPD without r.z (rho_prev = rho; FIN_RHO then sets rho and beta).
// FIN_POWER: a power step's Rayleigh quotient and vector norm (every sub-problem).
// Chebyshev preconditioner (cheb_c1 != 0): a negative r.z (FIN_UPD, FIN_RESTART*) stops the
// sub-problem as NOT_SPD -- the polynomial is not positive on the spectrum of D K (a bound below
// its largest eigenvalue).
enum { FIN_PQ = 0, FIN_UPD = 1, FIN_RES = 2, FIN_RES_CHECK = 3, FIN_RESTART_FIRST = 4, FIN_RESTART = 5,
       FIN_UPD_IC = 6, FIN_RHO = 7, FIN_POWER = 8 };

__global__ __launch_bounds__(1024) void k_finalize(int what, ProbScalars *S, const double *__restrict__ partials,
                                                  unsigned stride, const uint32_t *__restrict__ sub_wg0,
                                                  const uint32_t *__restrict__ sub_nwg, int sel) {
    const uint32_t s = blockIdx.x;
    ProbScalars &P = S[s];
    const int esel = what == FIN_PQ || what == FIN_UPD || what == FIN_UPD_IC || what == FIN_RHO ? (int)SEL_ACTIVE
                     : what == FIN_RESTART_FIRST       ? (int)SEL_ALL
                     : what == FIN_RESTART             ? (int)SEL_PAUSE_REQ
                                                       : sel;
    if (!selected(P, esel)) return;
    const unsigned w0 = sub_wg0[s], nw = sub_nwg[s];
    if (what == FIN_PQ) {
        double o[1];
        wg_combine<1, 0u>(partials, stride, w0, nw, o);
        if (threadIdx.x) return;
        finish_pq(P, o[0]);
    } else if (what == FIN_UPD || what == FIN_UPD_IC) {
        double o[3];
        wg_combine<3, 4u>(partials, stride, w0, nw, o);
        if (threadIdx.x) return;
        fin_update(P, o, what == FIN_UPD);
    } else if (what == FIN_RES || what == FIN_RES_CHECK) {
        double o[2];
        wg_combine<2, 2u>(partials, stride, w0, nw, o);
        if (threadIdx.x) return;
        P.rr = o[0];
        P.rmax = o[1];
        if (what == FIN_RES_CHECK) {
            P.rr_ref = o[0];
            // progress of the true residual: a check that does not halve the smallest one so far
            // (in the norm of the stop test) is a stall; STALL_CHECKS in a row end the solve --
            // the tolerance lies below what FP64 attains on this system
            const double v = P.crit_kind == 0 ? o[0] : o[1] * o[1];
            if (!finite(o[0])) { P.status = 6; P.done = 1; }
            else if (P.crit_kind == 0 ? o[0] <= P.stop2 : o[1] <= P.crit) P.done = 1;
            else if (P.checks > 0 && !(v < 0.25 * P.res_best)) {
                if (++P.stalls >= STALL_CHECKS) { P.status = 8; P.done = 1; }
            } else {
                P.res_best = v;
                P.stalls = 0;
                P.iter_best = P.iter;
            }
            P.checks += 1;
        }
    } else if (what == FIN_POWER) {
        double o[3];
        wg_combine<3, 0u>(partials, stride, w0, nw, o);
        if (threadIdx.x) return;
        if (o[0] > 0 && finite(o[1])) P.pw_rq = o[1] / o[0];
        P.pw_norm = o[2] > 0 && finite(o[2]) ? sqrt(o[2]) : 1.0;
    } else {
        double o[1];
        wg_combine<1, 0u>(partials, stride, w0, nw, o);
        if (threadIdx.x) return;
        P.rho = o[0];
        if (what == FIN_RESTART_FIRST) P.beta = 0;  // p = z: a fused direction update keeps it
        if (what == FIN_RHO) P.beta = o[0] / P.rho_prev;
        if (what == FIN_RESTART) {
            P.beta = o[0] / P.rho_prev;
            P.pause = PAUSE_DONE;
            P.refreshes += 1;
        }
        if (!finite(o[0])) { P.status = 6; P.done = 1; }
        else if (P.cheb_c1 != 0 && o[0] < 0) { P.status = 3; P.done = 1; }
    }
}

// Closing pass of a chunk of two-kernel iterations: applies the last vector update's pending
// partials (fin_update), combined exactly as the element-block kernel combines them (the same
// bits whether a chunk ends there or the next product folds them in).  One 256-lane workgroup per
// sub-problem.
__global__ __launch_bounds__(256) void k_fin_pending(ProbScalars *S, const double *__restrict__ partials, unsigned stride,
                                                     const uint32_t *__restrict__ sub_wg0,
                                                     const uint32_t *__restrict__ sub_nwg) {
    const uint32_t s = blockIdx.x;
    ProbScalars &P = S[s];
    if (!selected(P, SEL_ACTIVE) || !P.fin_pending) return;
    double o[3];
    wg_combine256<3, 4u, 4>(partials, stride, sub_wg0[s], sub_nwg[s], o);
    if (threadIdx.x) return;
    fin_update(P, o, true);
    P.fin_pending = 0;
}

__global__ void k_clear_pause(ProbScalars *S, uint32_t n_sub) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_sub && S[s].pause == PAUSE_DONE) S[s].pause = PAUSE_NONE;
}

// x64 += x32; x32 = 0 on the selected sub-problems
__global__ __launch_bounds__(256) void k_accumulate(double *__restrict__ x, float *__restrict__ x32, WgTables wt,
                                                    const ProbScalars *S, int sel) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB || !selected(S[s], sel)) return;
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n >= wt.row1[L]) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[3 * (size_t)n + d] += (double)x32[3 * (size_t)n + d];
        x32[3 * (size_t)n + d] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_zero_f32(size_t n, float *__restrict__ x) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = 0.0f;
}

// Jacobi factors from the diagonal blocks; a non-positive pivot marks the sub-problem NOT_SPD.
template <int PRE>
__global__ __launch_bounds__(256) void k_build_minv(const uint32_t *__restrict__ diag, const double *__restrict__ Kv,
                                                    uint64_t ps, double *__restrict__ M, uint16_t *__restrict__ M32,
                                                    WgTables wt, ProbScalars *S) {
    const unsigned L = xcd_slab();
    const uint32_t s = wt.sub[L];
    if (s == NO_SUB) return;
    const uint32_t n = wt.row0[L] + threadIdx.x;
    if (n >= wt.row1[L]) return;
    const uint64_t p = diag[n];
    double a[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) a[c] = Kv[c * ps + p];
    if (PRE == 0) {
        for (int d = 0; d < 3; ++d) {
            const double v = a[4 * d];
            if (!(v > 0)) S[s].status = 3;
            M[3 * (size_t)n + d] = 1.0 / v;
            if (M32) M32[3 * (size_t)n + d] = bf16_bits((float)(1.0 / v));
        }
    } else {
        const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
        const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
        if (!(det > 0) || !(a[0] > 0)) S[s].status = 3;
        const double id = 1.0 / det;
        double m[9];
        m[0] = c00 * id; m[3] = c01 * id; m[6] = c02 * id;
        m[1] = (a[2] * a[7] - a[1] * a[8]) * id; m[4] = (a[0] * a[8] - a[2] * a[6]) * id; m[7] = (a[1] * a[6] - a[0] * a[7]) * id;
        m[2] = (a[1] * a[5] - a[2] * a[4]) * id; m[5] = (a[2] * a[3] - a[0] * a[5]) * id; m[8] = (a[0] * a[4] - a[1] * a[3]) * id;
        const int up[6] = {0, 1, 2, 4, 5, 8};   // m is exactly symmetric (K's diagonal block is)
        uint16_t h[6];
        double v[6];
        for (int c = 0; c < 6; ++c) {
            M[6 * (size_t)n + c] = m[up[c]];
            h[c] = bf16_bits((float)m[up[c]]);
            v[c] = bf16_value(h[c]);
        }
        if (M32) {
            // the rounded block's leading minors (v = m00 m01 m02 m11 m12 m22); a block that is not
            // positive definite after rounding gives way to the node's point-Jacobi reciprocals
            const double d2 = v[0] * v[3] - v[1] * v[1];
            const double d3 = v[0] * (v[3] * v[5] - v[4] * v[4]) - v[1] * (v[1] * v[5] - v[4] * v[2]) +
                              v[2] * (v[1] * v[4] - v[3] * v[2]);
            if (!(v[0] > 0) || !(d2 > 0) || !(d3 > 0)) {
                h[0] = bf16_bits((float)(1.0 / a[0]));
                h[3] = bf16_bits((float)(1.0 / a[4]));
                h[5] = bf16_bits((float)(1.0 / a[8]));
                h[1] = h[2] = h[4] = 0;
            }
            for (int c = 0; c < 6; ++c) M32[6 * (size_t)n + c] = h[c];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------

static void finalize(Problem &P, int what, int sel = 0) {
    hipLaunchKernelGGL(k_finalize, dim3(P.n_sub), dim3(1024), 0, P.stream, what, P.d_ps, P.d_partials, P.n_wg, P.d_sub_wg0,
                       P.d_sub_nwg, sel);
}

// alpha = rho / p.q from the partials of the product just launched (the row-parallel formats
// write one per row workgroup, the matrix-free passes one per element block and per workgroup of
// the nodes-of-no-element pass, each sub-problem's in a range of its own).
static void finalize_pq(Problem &P, int precision) {
    if (P.spmv.format[precision] != FMT_MF)
        finalize(P, FIN_PQ);
    else
        hipLaunchKernelGGL(k_finalize, dim3(P.n_sub), dim3(1024), 0, P.stream, (int)FIN_PQ, P.d_ps, P.mf.d_partials,
                           P.mf.n_slots, P.mf.d_wg0, P.mf.d_nwg, 0);
}

static WgTables tables(const Problem &P) {
    return WgTables{P.d_wg_row0, P.d_wg_row1, P.d_wg_sub, P.d_sub_wg0, P.d_sub_nwg};
}
// workgroups of 1024 rows (4 per lane)
static WgTables tables_fat(const Problem &P) {
    return WgTables{P.d_fwg_row0, P.d_fwg_row1, P.d_fwg_sub, P.d_fsub_wg0, P.d_fsub_nwg};
}

// zdir (matrix-free MODE_PQ only): the direction update p = z + beta x is fused into the product.
// fin (with zdir): the two-kernel iteration (the previous update's end folded into the product).
template <typename T, int MODE>
static void launch_spmv_t(Problem &P, const T *K, const T *x, const double *b, T *y, int sel,
                          const T *zdir = nullptr, void *halo = nullptr, const FinParts *fin = nullptr) {
    const dim3 g(P.n_wg), blk(256);
    const WgTables wt = tables(P);
    const int pr = sizeof(T) == 4 ? 1 : 0;
    // matrix-free products; an FP64 residual is the product into y, then y = b - y with the
    // residual's partials
    if (P.spmv.format[pr] == FMT_MF) {
        // a launch error stays pending (mf_launch peeks) for run_chunk's and solve_impl's checks
        (void)mf_launch(P, x, y, MODE == MODE_PQ, MODE == MODE_PQ ? zdir : nullptr, halo, MODE == MODE_PQ ? fin : nullptr);
        if (MODE == MODE_RES)
            hipLaunchKernelGGL(k_residual, g, blk, 0, P.stream, b, reinterpret_cast<double *>(y), wt, P.d_ps, P.d_partials,
                               sel);
        return;
    }
    if (P.spmv.format[pr] == FMT_SYM) {
        const SymDev &Y = P.sym;
        const SpmvArgs A{Y.d_lenU, Y.d_sptrU, Y.d_colU, Y.psU, Y.d_sptrL, Y.d_lenL, Y.d_colL, Y.d_tposL, nullptr, nullptr};
        const T *Ks = pr ? reinterpret_cast<const T *>(Y.d_K32) : reinterpret_cast<const T *>(Y.d_K);
#define FEMGPU_SPMV_CASE(U) \
    hipLaunchKernelGGL((k_spmv<T, MODE, U, FMT_SYM, LOAD_PLAIN>), g, blk, 0, P.stream, A, Ks, x, b, y, wt, P.d_ps, P.d_partials, sel)
        switch (P.spmv.unroll[pr]) {
            case 1: FEMGPU_SPMV_CASE(1); break;
            case 3: FEMGPU_SPMV_CASE(3); break;
            default: FEMGPU_SPMV_CASE(2); break;
        }
#undef FEMGPU_SPMV_CASE
        return;
    }
    const SpmvArgs A{P.d_rp, P.d_sptr, P.d_col, P.ps, nullptr, nullptr, nullptr, nullptr, P.d_col16, P.d_perm};
#define FEMGPU_SPMV_CASE(U, LD) \
    hipLaunchKernelGGL((k_spmv<T, MODE, U, FMT_FULL, LD>), g, blk, 0, P.stream, A, K, x, b, y, wt, P.d_ps, P.d_partials, sel)
#define FEMGPU_SPMV_LOADS(U)                                                   \
    switch (P.spmv.loads[pr]) {                                                \
        case LOAD_PLAIN: FEMGPU_SPMV_CASE(U, LOAD_PLAIN); break;               \
        case LOAD_NT16: FEMGPU_SPMV_CASE(U, LOAD_NT16); break;                 \
        default: FEMGPU_SPMV_CASE(U, LOAD_NT); break;                          \
    }
    switch (P.spmv.unroll[pr]) {
        case 1: FEMGPU_SPMV_LOADS(1); break;
        case 3: FEMGPU_SPMV_LOADS(3); break;
        default: FEMGPU_SPMV_LOADS(2); break;
    }
#undef FEMGPU_SPMV_LOADS
#undef FEMGPU_SPMV_CASE
}

namespace {
template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <class T>
hipError_t upload(T **p, const std::vector<T> &h, size_t n_alloc) {
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(1, n_alloc) * sizeof(T));
    if (e == hipSuccess && !h.empty()) e = hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}
}  // namespace

void sym_free(SymDev &S) {
    dfree(S.d_sptrU); dfree(S.d_lenU); dfree(S.d_colU); dfree(S.d_srcU);
    dfree(S.d_sptrL); dfree(S.d_lenL); dfree(S.d_colL); dfree(S.d_tposL);
    dfree(S.d_K); dfree(S.d_K32);
    S = SymDev();
}

int sym_ensure(Problem &P, int precision) {
    SymDev &Y = P.sym;
    if (!Y.built) {
        SymSell H;
        if (!build_sym_sell(P.pat, H)) return 1;
        Y.psU = (H.nU + 63) / 64 * 64;
        Y.nL = H.nL;
        std::vector<uint32_t> src = H.srcU;
        src.resize(std::max<uint64_t>(Y.psU, 1), 0xffffffffu);
        if (upload(&Y.d_sptrU, H.sptrU, H.sptrU.size()) || upload(&Y.d_lenU, H.lenU, H.lenU.size()) ||
            upload(&Y.d_colU, H.colU, H.colU.size()) || upload(&Y.d_srcU, src, src.size()) ||
            upload(&Y.d_sptrL, H.sptrL, H.sptrL.size()) || upload(&Y.d_lenL, H.lenL, H.lenL.size()) ||
            upload(&Y.d_colL, H.colL, H.colL.size()) || upload(&Y.d_tposL, H.tposL, H.tposL.size()))
            return 5;
        Y.built = true;
        destroy_graph(P);
    }
    if (Y.valid[precision]) return 0;
    const dim3 g(8192), blk(256);
    if (precision == 1) {
        if (!Y.d_K32 && hipMalloc(&Y.d_K32, sizeof(float) * 9 * std::max<uint64_t>(Y.psU, 1)) != hipSuccess) return 5;
        hipLaunchKernelGGL(k_to_sym<float>, g, blk, 0, P.stream, P.d_K, P.ps, Y.d_srcU, Y.d_K32, Y.psU);
    } else {
        if (!Y.d_K && hipMalloc(&Y.d_K, sizeof(double) * 9 * std::max<uint64_t>(Y.psU, 1)) != hipSuccess) return 5;
        hipLaunchKernelGGL(k_to_sym<double>, g, blk, 0, P.stream, P.d_K, P.ps, Y.d_srcU, Y.d_K, Y.psU);
    }
    if (hipGetLastError() != hipSuccess) return 5;
    Y.valid[precision] = true;
    return 0;
}

// Values the SpMV of this precision reads (FP32 full copy, symmetric storage, or the matrix-free
// arrays).
static int spmv_ensure(Problem &P, int precision) {
    if (P.spmv.format[precision] == FMT_SYM) return sym_ensure(P, precision);
    if (P.spmv.format[precision] == FMT_MF) return mf_prepare(P);
    if (precision == 1 && !P.k32_valid) {
        if (!P.d_K32 && hipMalloc(&P.d_K32, sizeof(float) * 9 * P.ps) != hipSuccess) return 5;
        if (launch_to_fp32(P) != hipSuccess) return 5;
        P.k32_valid = true;
    }
    return 0;
}

// Average device time of one product over reps launches, after one untimed launch (the first
// launch of a kernel in a process also loads its code object).
double time_spmv(Problem &P, int precision, int reps) {
    if (spmv_ensure(P, precision)) return -1;
    if (precision == 1) launch_spmv_t<float, MODE_PLAIN>(P, P.d_K32, P.d_p32, nullptr, P.d_q32, 0);
    else launch_spmv_t<double, MODE_PLAIN>(P, P.d_K, P.d_p, nullptr, P.d_q, 0);
    (void)hipEventRecord(P.ev[2], P.stream);
    for (int i = 0; i < reps; ++i) {
        if (precision == 1) launch_spmv_t<float, MODE_PLAIN>(P, P.d_K32, P.d_p32, nullptr, P.d_q32, 0);
        else launch_spmv_t<double, MODE_PLAIN>(P, P.d_K, P.d_p, nullptr, P.d_q, 0);
    }
    (void)hipEventRecord(P.ev[3], P.stream);
    if (hipEventSynchronize(P.ev[3]) != hipSuccess) return -1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, P.ev[2], P.ev[3]);
    return (double)ms / reps;
}

// Times every SpMV variant (storage format x SELL slots per loop trip) on this matrix, for each
// precision (10 launches after 2 warm-up launches), and keeps the fastest.  The symmetric storage
// holds a second copy of the values, so it is kept only when it beats the full storage by more
// than 3%; storage no precision picked is released.  The matrix-free product (Hexa and Tetra_10
// problems) is kept when it beats the stored formats; the full storage's fastest variant then
// serves the residual products.
int autotune_spmv(Problem &P) {
    for (int pr = 0; pr < 2; ++pr) {
        double best[2] = {1e30, 1e30};
        int bu[2] = {P.spmv.unroll[pr], P.spmv.unroll[pr]};
        int bld = LOAD_NT;
        for (int f : {FMT_FULL, FMT_SYM})
            for (int ld = LOAD_PLAIN; ld <= (f == FMT_FULL ? (P.d_col16 ? LOAD_NT16 : LOAD_NT) : LOAD_PLAIN); ++ld)
                for (int u : {1, 2, 3}) {
                    P.spmv.format[pr] = f;
                    P.spmv.unroll[pr] = u;
                    P.spmv.loads[pr] = ld;
                    if (time_spmv(P, pr, 2) < 0) continue;
                    const double t = time_spmv(P, pr, 10);
                    if (t > 0 && t < best[f]) {
                        best[f] = t;
                        bu[f] = u;
                        if (f == FMT_FULL) bld = ld;
                    }
                }
        const int f = best[FMT_SYM] < 0.97 * best[FMT_FULL] ? FMT_SYM : FMT_FULL;
        double best_mf = 1e30;
        if (mf_eligible(P)) {
            P.spmv.format[pr] = FMT_MF;
            if (time_spmv(P, pr, 2) >= 0) best_mf = time_spmv(P, pr, 10);
            if (!(best_mf > 0)) best_mf = 1e30;
        }
        if (best_mf < best[f]) {
            P.spmv.format[pr] = FMT_MF;
            P.spmv.unroll[pr] = bu[FMT_FULL];
            P.spmv.loads[pr] = bld;
        } else {
            P.spmv.format[pr] = f;
            P.spmv.unroll[pr] = bu[f];
            P.spmv.loads[pr] = f == FMT_FULL ? bld : LOAD_PLAIN;
        }
    }
    if (P.spmv.format[0] != FMT_SYM && P.spmv.format[1] != FMT_SYM) sym_free(P.sym);
    else {
        if (P.spmv.format[0] != FMT_SYM) { dfree(P.sym.d_K); P.sym.valid[0] = false; }
        if (P.spmv.format[1] != FMT_SYM) { dfree(P.sym.d_K32); P.sym.valid[1] = false; }
    }
    P.spmv.tuned = true;
    destroy_graph(P);
    return hipGetLastError() == hipSuccess ? 0 : 5;
}

__global__ __launch_bounds__(256) void k_vec_to_f32(const double *__restrict__ a, float *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = (float)a[i];
}
__global__ __launch_bounds__(256) void k_vec_to_f64(const float *__restrict__ a, double *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = (double)a[i];
}

// y = K x with the product variant of this precision (FP32: x rounded to float, y widened back).
int apply_spmv(Problem &P, int precision, const double *d_x, double *d_y) {
    const int rc = spmv_ensure(P, precision);
    if (rc) return rc;
    const size_t n = 3 * (size_t)P.n_nodes;
    const dim3 g((unsigned)std::max<size_t>(1, std::min<size_t>(8192, (n + 255) / 256))), blk(256);
    if (precision == 1) {
        hipLaunchKernelGGL(k_vec_to_f32, g, blk, 0, P.stream, d_x, P.d_p32, n);
        launch_spmv_t<float, MODE_PLAIN>(P, P.d_K32, P.d_p32, nullptr, P.d_q32, 0);
        hipLaunchKernelGGL(k_vec_to_f64, g, blk, 0, P.stream, P.d_q32, d_y, n);
    } else {
        launch_spmv_t<double, MODE_PLAIN>(P, P.d_K, d_x, nullptr, d_y, 0);
    }
    return hipGetLastError() == hipSuccess ? 0 : 5;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

int64_t apply_scratch_bytes(const Problem &P, int precision) {
    const size_t n = 3 * (size_t)P.n_nodes;
    size_t b = precision == 1 ? align256(2 * n * sizeof(float)) : 0;
    switch (P.spmv.format[precision]) {
        case FMT_MF:
            if (!mf_ready(P)) return -1;
            return (int64_t)(b + align256(mf_halo_values(P) * (precision == 1 ? sizeof(float) : sizeof(double))));
        case FMT_SYM: return P.sym.valid[precision] ? (int64_t)b : -1;
        default: return precision == 1 && !P.k32_valid ? -1 : (int64_t)b;
    }
}

int apply_spmv_scratch(Problem &P, int precision, const double *d_x, double *d_y, char *scratch) {
    const size_t n = 3 * (size_t)P.n_nodes;
    const dim3 g((unsigned)std::max<size_t>(1, std::min<size_t>(8192, (n + 255) / 256))), blk(256);
    if (precision == 1) {
        float *x32 = reinterpret_cast<float *>(scratch), *y32 = x32 + n;
        char *halo = scratch + align256(2 * n * sizeof(float));
        hipLaunchKernelGGL(k_vec_to_f32, g, blk, 0, P.stream, d_x, x32, n);
        launch_spmv_t<float, MODE_PLAIN>(P, P.d_K32, x32, nullptr, y32, 0, nullptr, halo);
        hipLaunchKernelGGL(k_vec_to_f64, g, blk, 0, P.stream, y32, d_y, n);
    } else {
        launch_spmv_t<double, MODE_PLAIN>(P, P.d_K, d_x, nullptr, d_y, 0, nullptr, scratch);
    }
    return hipGetLastError() == hipSuccess ? 0 : 5;
}

void destroy_graph(Problem &P) {
    for (int k = 0; k < Problem::N_GRAPHS; ++k) {
        if (P.graph_exec[k]) (void)hipGraphExecDestroy(P.graph_exec[k]);
        P.graph_exec[k] = nullptr;
        P.graph_key[k] = 0;
        P.graph_x[k] = nullptr;
    }
}

static int read_scalars(Problem &P) {
    if (hipMemcpyAsync(P.h_ps, P.d_ps, sizeof(ProbScalars) * P.n_sub, hipMemcpyDeviceToHost, P.stream) != hipSuccess)
        return 5;
    if (hipStreamSynchronize(P.stream) != hipSuccess) return 5;
    return 0;
}

static int write_scalars(Problem &P) {
    if (hipMemcpyAsync(P.d_ps, P.h_ps, sizeof(ProbScalars) * P.n_sub, hipMemcpyHostToDevice, P.stream) != hipSuccess)
        return 5;
    return 0;
}

// z = M^-1 r of precision T: the IC(0) sweep output, or the vector the Jacobi updates store.
template <int PRE, typename T>
static T *zbuf(Problem &P) {
    if (sizeof(T) == 4) return reinterpret_cast<T *>(PRE == PRE_IC0 ? P.ic0.d_z32 : P.d_z32);
    return reinterpret_cast<T *>(PRE == PRE_IC0 ? P.ic0.d_z : P.d_z);
}

// One PCG iteration on precision T (K, vectors and the Jacobi factors or IC(0) buffers of T).
// Jacobi preconditioners with the matrix-free product: the direction update p = z + beta p that
// ends an iteration is fused into the next iteration's product (beta = 0 after the first restart,
// where p = z already; a refresh leaves its beta for the next product).
template <typename T>
static bool fused_direction(const Problem &P, int pre) {
    return pre != PRE_IC0 && P.spmv.format[sizeof(T) == 4 ? 1 : 0] == FMT_MF;
}

// Two kernels per iteration (fused direction, every sub-problem with an element block): the end
// of each vector update (k_finalize FIN_UPD) is folded into the next product's element blocks,
// and k_fin_pending closes each chunk.
template <typename T>
static bool folded_iteration(const Problem &P, int pre) {
    return fused_direction<T>(P, pre) && P.mf.fold;
}

// The partials a two-kernel iteration's element blocks (and the chunk's closing pass) combine:
// the vector update's, in the 1024-row workgroup tables, or for the Chebyshev preconditioner
// k_cheb's, in the 256-row tables.
static FinParts fin_parts(const Problem &P, int pre) {
    if (pre == PRE_CHEB) return FinParts{P.d_partials, P.n_wg, P.d_sub_wg0, P.d_sub_nwg};
    return FinParts{P.d_partials, P.n_fwg, P.d_fsub_wg0, P.d_fsub_nwg};
}

// Chebyshev preconditioner, after y = D r was stored in z: v = K y into q (matrix-free: the
// shared nodes' partials stay in the halo buffer for k_cheb), then z = c0 y + c1 D v with the
// partials of r.z, r.r, max|r| (iteration, NV 3: the active sub-problems) or of r.z (restart, NV
// 1: sel, with p = z when first).
template <typename T, int NV>
static void launch_cheb(Problem &P, const T *K, const T *r, T *q, const mstore_t<T> *Minv, T *z, T *p, int sel, int first) {
    const WgTables wt = tables(P);
    const dim3 g(P.n_wg), blk(256);
    const bool mf = P.spmv.format[sizeof(T) == 4 ? 1 : 0] == FMT_MF;
    // the product runs on the iteration's active sub-problems (MODE_PQ; its p.q partials are not
    // used) or on every sub-problem
    if (mf) (void)mf_launch_partial(P, z, q, NV == 3);
    else if (NV == 3) launch_spmv_t<T, MODE_PQ>(P, K, z, nullptr, q, SEL_ACTIVE);
    else launch_spmv_t<T, MODE_PLAIN>(P, K, z, nullptr, q, SEL_ALL);
    hipLaunchKernelGGL((k_cheb<T, NV>), g, blk, 0, P.stream, r, z, q, reinterpret_cast<const T *>(P.mf.d_halo),
                       mf ? (const uint32_t *)P.mf.d_hbeg : nullptr, Minv, p, wt, P.d_ps, P.d_partials, sel, first);
}

template <int PRE, typename T>
static void launch_iteration_t(Problem &P, const T *K, T *x, T *r, T *p, T *q, const mstore_t<T> *Minv, T *z) {
    const WgTables wt = tables(P);
    const dim3 g(P.n_wg), blk(256);
    const bool fused = fused_direction<T>(P, PRE);
    const bool fold = folded_iteration<T>(P, PRE);
    const FinParts fp = fin_parts(P, PRE);
    hipEvent_t *kt = fold && P.kt_used < P.kt_cap ? P.kt + 3 * (size_t)P.kt_used++ : nullptr;
    if (kt) (void)hipEventRecord(kt[0], P.stream);
    launch_spmv_t<T, MODE_PQ>(P, K, p, nullptr, q, SEL_ACTIVE, fused ? z : nullptr, nullptr, fold ? &fp : nullptr);
    if (kt) (void)hipEventRecord(kt[1], P.stream);
    if constexpr (PRE != PRE_IC0) {
        if (fused) {
            // the product's shared-node pass and alpha are folded into the update
            const T *p_alt = reinterpret_cast<const T *>(sizeof(T) == 4 ? (void *)P.mf.d_pb32 : (void *)P.mf.d_pb);
            if (fold)
                hipLaunchKernelGGL((k_update_mf<PRE, T, true>), dim3(P.n_fwg), dim3(sizeof(T) == 8 ? 1024 : UPD_NT32), 0, P.stream, x, r, p, p_alt, q,
                                   reinterpret_cast<const T *>(P.mf.d_halo), P.mf.d_hbeg, P.mf.d_partials, P.mf.d_wg0,
                                   P.mf.d_nwg, Minv, z, tables_fat(P), P.d_ps + P.n_sub, P.d_ps, P.d_partials);
            else
                hipLaunchKernelGGL((k_update_mf<PRE, T, false>), g, blk, 0, P.stream, x, r, p, p_alt, q,
                                   reinterpret_cast<const T *>(P.mf.d_halo), P.mf.d_hbeg, P.mf.d_partials, P.mf.d_wg0,
                                   P.mf.d_nwg, Minv, z, wt, P.d_ps, P.d_ps, P.d_partials);
            if (kt) (void)hipEventRecord(kt[2], P.stream);
        }
    }
    if (!fused) {
        finalize_pq(P, sizeof(T) == 4 ? 1 : 0);
        hipLaunchKernelGGL((k_update<PRE, T>), g, blk, 0, P.stream, x, r, p, nullptr, q, Minv, z, wt, P.d_ps, P.d_partials);
    }
    // Chebyshev: the update stored y = D r in z (its partials are replaced by k_cheb's)
    if (PRE == PRE_CHEB) launch_cheb<T, 3>(P, K, r, q, Minv, z, nullptr, (int)SEL_ACTIVE, 0);
    if (PRE == PRE_IC0) {
        finalize(P, FIN_UPD_IC);
        ic0_apply(P, r, z, SEL_ACTIVE);
        hipLaunchKernelGGL(k_dot_rz<T>, g, blk, 0, P.stream, r, z, wt, P.d_ps, P.d_partials, (int)SEL_ACTIVE);
        finalize(P, FIN_RHO);
    } else if (!fold) {
        finalize(P, FIN_UPD);
    }
    if (!fused) hipLaunchKernelGGL(k_update_p<T>, g, blk, 0, P.stream, z, p, wt, P.d_ps, (int)SEL_ACTIVE, 0);
}

template <int PRE>
static void launch_iteration(Problem &P, bool fp32, double *d_x) {
    if (fp32)
        launch_iteration_t<PRE, float>(P, P.d_K32, P.d_x32, P.d_r32, P.d_p32, P.d_q32, P.d_Minv32, zbuf<PRE, float>(P));
    else
        launch_iteration_t<PRE, double>(P, P.d_K, d_x, P.d_r, P.d_p, P.d_q, P.d_Minv, zbuf<PRE, double>(P));
}

// (Re)start of the recurrence on precision T from the FP64 residual r64: first = 1 for every
// unfinished sub-problem (p = z, rho = r.z), else for those with a pending refresh (rho = r.z,
// beta = rho / rho_prev, p = z + beta p).
template <int PRE, typename T>
static void launch_restart(Problem &P, const double *r64, T *r, T *p, const mstore_t<T> *Minv, T *z, int first) {
    const WgTables wt = tables(P);
    const dim3 g(P.n_wg), blk(256);
    const int sel = first ? (int)SEL_ALL : (int)SEL_PAUSE_REQ;
    if (PRE == PRE_IC0) {
        if ((const void *)r64 != (const void *)r)
            hipLaunchKernelGGL((k_restart<PRE, T>), g, blk, 0, P.stream, r64, r, p, Minv, z, wt, P.d_ps, P.d_partials, first);
        ic0_apply(P, r, z, sel);
        hipLaunchKernelGGL(k_dot_rz<T>, g, blk, 0, P.stream, r, z, wt, P.d_ps, P.d_partials, sel);
        finalize(P, first ? FIN_RESTART_FIRST : FIN_RESTART);
        hipLaunchKernelGGL(k_update_p<T>, g, blk, 0, P.stream, z, p, wt, P.d_ps, first ? (int)SEL_ALL : (int)SEL_PAUSE_DONE,
                           first);
    } else {
        hipLaunchKernelGGL((k_restart<PRE, T>), g, blk, 0, P.stream, r64, r, p, Minv, z, wt, P.d_ps,
                           P.d_partials, first);
        // Chebyshev: k_restart stored y = D r in z; the product of y goes to the dead q buffer (in
        // FP64 r64 itself, read by k_restart already)
        if (PRE == PRE_CHEB) {
            T *q = reinterpret_cast<T *>(sizeof(T) == 4 ? (void *)P.d_q32 : (void *)P.d_q);
            const T *K = reinterpret_cast<const T *>(sizeof(T) == 4 ? (const void *)P.d_K32 : (const void *)P.d_K);
            launch_cheb<T, 1>(P, K, r, q, Minv, z, p, sel, first);
        }
        finalize(P, first ? FIN_RESTART_FIRST : FIN_RESTART);
        if (!first && !fused_direction<T>(P, PRE))
            hipLaunchKernelGGL(k_update_p<T>, g, blk, 0, P.stream, z, p, wt, P.d_ps, (int)SEL_PAUSE_DONE, 0);
    }
}

// n iterations replayed from the captured graph of slot (captured on first use or when the
// configuration or the solution vector changed; the kernels read all state from device memory, so
// one graph serves every chunk of every solve on d_x).
template <int PRE>
static int run_graph(Problem &P, bool fp32, double *d_x, int n, int slot) {
    const int pr = fp32 ? 1 : 0;
    // every field in its own bits; the solution pointer is compared separately
    const uint64_t key = (uint64_t)fp32 | ((uint64_t)PRE << 1) | ((uint64_t)P.spmv.format[pr] << 3) |
                         ((uint64_t)P.spmv.unroll[pr] << 5) | ((uint64_t)P.spmv.loads[pr] << 7) | ((uint64_t)n << 9);
    hipGraphExec_t &ge = P.graph_exec[slot];
    if (!ge || P.graph_key[slot] != key || P.graph_x[slot] != d_x) {
        if (ge) (void)hipGraphExecDestroy(ge);
        ge = nullptr;
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return 5;
        for (int k = 0; k < n; ++k) launch_iteration<PRE>(P, fp32, d_x);
        if (hipStreamEndCapture(P.stream, &g) != hipSuccess) return 5;
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { ge = nullptr; return 5; }
        P.graph_key[slot] = key;
        P.graph_x[slot] = d_x;
    }
    return hipGraphLaunch(ge, P.stream) == hipSuccess ? 0 : 5;
}

// One chunk of n iterations (n even): replayed from captured hipGraphs when enabled -- graphs of
// the full length (the solve's chunk option), the rest as a sum of graphs of 2^k iterations,
// launched back to back.
template <int PRE>
static int run_chunk(Problem &P, bool fp32, double *d_x, int n, int full) {
    // two-kernel iterations: the chunk's last update is applied by a closing pass, so the host
    // reads the scalars of a whole iteration
    const bool fold = fp32 ? folded_iteration<float>(P, PRE) : folded_iteration<double>(P, PRE);
    auto close = [&]() {
        if (fold) {
            const FinParts fp = fin_parts(P, PRE);
            hipLaunchKernelGGL(k_fin_pending, dim3(P.n_sub), dim3(256), 0, P.stream, P.d_ps, fp.part, fp.stride, fp.w0,
                               fp.nw);
        }
        return hipGetLastError() == hipSuccess ? 0 : 5;
    };
    // level-scheduled IC(0) in natural order has hundreds of sweep launches per iteration: no graph
    if (!P.use_graph || (PRE == PRE_IC0 && P.ic0.sched.n_levels() > 64)) {
        for (int k = 0; k < n; ++k) launch_iteration<PRE>(P, fp32, d_x);
        return close();
    }
    for (; n >= full; n -= full)
        if (run_graph<PRE>(P, fp32, d_x, full, 0)) return 5;
    while (n >= 2) {
        int k = 1;
        while (k + 1 < Problem::N_GRAPHS && (2 << k) <= n) ++k;
        if (run_graph<PRE>(P, fp32, d_x, 1 << k, k)) return 5;
        n -= 1 << k;
    }
    return close();
}

// Iterations until the next pause of a single running problem (a check or an FP64 refresh: its
// recurrence residual below the stop test's bound, or in MIXED below delta^2 of the last FP64
// residual), from its squared residual rr now, the last FP64 one rr_ref and the geometric rate
// of rr per iteration (log, < 0) measured over the chunks so far, kept across refreshes (the
// recurrence restarts on the same operator); rounded up to even, within [2, cap]; dflt without a
// rate.  A chunk that ends at the pause saves the iterations that would run after it as no-ops
// (every kernel of a paused problem returns at once) until the chunk's end, and a long one the
// host's polls in between.  Up to 4 chunk lengths: a longer prediction is the likelier to
// overshoot (Hexa8 100^3 MIXED 85.8-86.4 ms against 86.7-87.4 with up to 8).
constexpr int PREDICTED_CHUNKS = 4;
static int predict_chunk(const ProbScalars &S, double rr, double rr_ref, double rate, int dflt, int cap) {
    if (!(rate < 0) || !(rr > 0)) return dflt;
    double thr = S.crit_kind == 0 ? S.stop2 : 0.0;
    if (S.mixed) thr = std::max(thr, S.delta2 * rr_ref);
    if (!(thr > 0)) return dflt;
    const double k = std::log(thr / rr) / rate;
    if (!(k < (double)cap)) return cap;
    const int n = std::max(2, ((int)std::ceil(k) + 1) & ~1);
    return std::min(n, cap);
}

// Chebyshev preconditioner: z = p(D K) D r with p(t) = c0 + c1 t, the degree-2 residual
// polynomial 1 - t p(t) = T2((b + a - 2t) / (b - a)) / T2((b + a) / (b - a)) (T2(x) = 2x^2 - 1)
// smallest in magnitude on [a, b], a = CHEB_LOWER b.  b = CHEB_SAFETY x the Rayleigh quotient
// after CHEB_POWER_STEPS power steps from a fixed start (a lower bound of the largest eigenvalue
// of D K).  The preconditioner is positive definite while every eigenvalue lies below
// a + b = 1.05 b, so b may undershoot the largest eigenvalue by about 5% (the quotient by about
// 20%) before r.z can turn negative.
constexpr int CHEB_POWER_STEPS = 20;
constexpr double CHEB_SAFETY = 1.15, CHEB_LOWER = 0.05;

// Power iteration on D K in the solve's precision (FP32: the FP32 copies of K or the FP32
// matrix-free product and of the block-Jacobi factors; vectors in the direction and product
// buffers, dead before a solve), every sub-problem at once; then (c0, c1) per sub-problem into
// P.cheb_c and the host scalars (written back).  Returns 3 when a sub-problem's quotient is not
// positive.
template <typename T>
static void cheb_power(Problem &P, const T *K, T *v, T *u, const mstore_t<T> *Minv) {
    const WgTables wt = tables(P);
    const dim3 g(P.n_wg), blk(256);
    hipLaunchKernelGGL(k_power_init<T>, g, blk, 0, P.stream, v, wt, P.d_partials);
    finalize(P, FIN_POWER, SEL_EVERY);
    for (int k = 0; k < CHEB_POWER_STEPS; ++k) {
        launch_spmv_t<T, MODE_PLAIN>(P, K, v, nullptr, u, SEL_EVERY);
        hipLaunchKernelGGL(k_power<T>, g, blk, 0, P.stream, v, u, Minv, wt, P.d_ps, P.d_partials);
        finalize(P, FIN_POWER, SEL_EVERY);
    }
}

static int cheb_setup(Problem &P, bool fp32) {
    if (fp32) cheb_power<float>(P, P.d_K32, P.d_p32, P.d_q32, P.d_Minv32);
    else cheb_power<double>(P, P.d_K, P.d_p, P.d_q, P.d_Minv);
    if (hipGetLastError() != hipSuccess || read_scalars(P)) return 5;
    P.cheb_c.assign(2 * (size_t)P.n_sub, 0.0);
    int rc = 0;
    for (uint32_t s = 0; s < P.n_sub; ++s) {
        ProbScalars &S = P.h_ps[s];
        const double lam = S.pw_rq;
        if (!(lam > 0) || !std::isfinite(lam)) {
            S.status = 3;
            rc = 3;
            continue;
        }
        const double b = CHEB_SAFETY * lam, a = CHEB_LOWER * b;
        const double th = 0.5 * (b + a), de = 0.5 * (b - a), t2 = 2 * (th / de) * (th / de) - 1;
        S.cheb_c0 = P.cheb_c[2 * s] = 4 * th / (de * de) / t2;
        S.cheb_c1 = P.cheb_c[2 * s + 1] = -2 / (de * de) / t2;
    }
    if (write_scalars(P)) return 5;
    return rc;
}

template <int PRE>
static int solve_impl(Problem &P, const double *d_rhs, double *d_x, const SolveOpts &o, SolveOut *out, int ordering) {
    const bool mixed = o.precision == 1;
    const size_t n = 3 * (size_t)P.n_nodes;
    const WgTables wt = tables(P);
    const dim3 g(P.n_wg), blk(256);
    std::memset(P.h_ps, 0, sizeof(ProbScalars) * P.n_sub);
    for (uint32_t s = 0; s < P.n_sub; ++s) {
        ProbScalars &S = P.h_ps[s];
        S.crit_kind = o.crit_kind;
        S.crit = o.tol;
        S.max_iter = o.max_iter;
        S.mixed = mixed ? 1 : 0;
        S.delta2 = o.delta * o.delta;
        S.max_diag = s < P.max_diag.size() ? P.max_diag[s] : 0.0;
    }
    if (write_scalars(P)) return 5;
    // ||b||^2 per sub-problem: the residual kernel with x = 0
    if (launch_fill(P.d_q, n, 0.0, P.stream) != hipSuccess) return 5;
    launch_spmv_t<double, MODE_RES>(P, P.d_K, P.d_q, d_rhs, P.d_r, SEL_EVERY);
    finalize(P, FIN_RES, SEL_EVERY);
    if (read_scalars(P)) return 5;
    std::vector<double> bb(P.n_sub);
    for (uint32_t s = 0; s < P.n_sub; ++s) {
        ProbScalars &S = P.h_ps[s];
        bb[s] = S.rr;
        S.bb = S.rr;
        S.stop2 = o.tol * o.tol * (S.rr > 0 ? S.rr : 1.0);
        S.rr = S.rmax = 0;
    }
    if (write_scalars(P)) return 5;
    bool built_minv = false;
    if (PRE == PRE_IC0) {
        const int rc = ic0_ensure(P, ordering, mixed);
        if (rc != 0 && rc != 3) return rc == 1 ? 1 : 5;  // 3: statuses are read below
    } else if (P.minv_kind != PRE || (mixed && !P.d_Minv32)) {
        if (mixed && !P.d_Minv32 && hipMalloc(&P.d_Minv32, sizeof(uint16_t) * 6 * std::max<uint32_t>(1, P.n_nodes)) != hipSuccess)
            return 5;
        (void)hipEventRecord(P.ev[2], P.stream);
        hipLaunchKernelGGL(k_build_minv<jacobi_kind<PRE>()>, g, blk, 0, P.stream, P.d_diag, P.d_K, P.ps, P.d_Minv,
                           P.d_Minv32, wt, P.d_ps);
        (void)hipEventRecord(P.ev[3], P.stream);
        P.minv_kind = PRE;
        built_minv = true;
    } else {
        P.last_precond_ms = 0;
    }
    if (read_scalars(P)) return 5;
    // Chebyshev: its coefficients from a power iteration with the factors just built, else those
    // kept from the build
    if (PRE == PRE_CHEB) {
        bool ok = true;
        for (uint32_t s = 0; s < P.n_sub; ++s) ok &= !P.h_ps[s].status;
        if (ok && (built_minv || P.cheb_c.size() != 2 * (size_t)P.n_sub)) {
            const int rc = cheb_setup(P, mixed);
            if (rc == 5) return 5;
            (void)hipEventRecord(P.ev[3], P.stream);
            if (hipEventSynchronize(P.ev[3]) != hipSuccess) return 5;
        } else if (ok) {
            for (uint32_t s = 0; s < P.n_sub; ++s) {
                P.h_ps[s].cheb_c0 = P.cheb_c[2 * s];
                P.h_ps[s].cheb_c1 = P.cheb_c[2 * s + 1];
            }
            if (write_scalars(P)) return 5;
        }
    }
    if (built_minv) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, P.ev[2], P.ev[3]) == hipSuccess) P.last_precond_ms = ms;
    }
    for (uint32_t s = 0; s < P.n_sub; ++s)
        if (P.h_ps[s].status) {
            P.minv_kind = -1;
            for (uint32_t t = 0; t < P.n_sub; ++t) out[t] = SolveOut{P.h_ps[t].status, 0, 0, 0, 0, 0};
            return P.h_ps[s].status;
        }

    (void)hipEventRecord(P.ev[0], P.stream);
    launch_spmv_t<double, MODE_RES>(P, P.d_K, d_x, d_rhs, P.d_r, SEL_ALL);
    finalize(P, FIN_RES_CHECK, SEL_ALL);
    if (mixed) {
        hipLaunchKernelGGL(k_zero_f32, dim3(std::max<size_t>(1, std::min<size_t>(8192, (n + 255) / 256))), blk, 0, P.stream, n, P.d_x32);
        launch_restart<PRE, float>(P, P.d_r, P.d_r32, P.d_p32, P.d_Minv32, zbuf<PRE, float>(P), 1);
    } else {
        launch_restart<PRE, double>(P, P.d_r, P.d_r, P.d_p, P.d_Minv, zbuf<PRE, double>(P), 1);
    }
    const int full = std::max(2, o.chunk);
    int chunk = P.use_graph ? full : 2;
    // single problems: chunks end at the predicted pause, up to PREDICTED_CHUNKS chunk lengths
    // once a rate is known (a batch keeps full chunks: its paused sub-problems idle while the
    // others run); without graphs the first chunks double from 2
    const bool predict = P.n_sub == 1;
    const int cap = PREDICTED_CHUNKS * full;
    double rr_prev = 0, rate = 0, rr_cyc = 0;
    long it_prev = -1, it_cyc = -1;
    const long max_chunks = (long)o.max_iter / 2 + 1000;
    // chunk timing: events 4-6 as a ring (the previous chunk's end stays for the gap)
    P.last_chunks = P.last_gaps = 0;
    P.last_chunk_ms = P.last_gap_ms = 0;
    int prev_end = -1;
    bool prev_refresh = false;
    for (long c = 0;; ++c) {
        if (c > max_chunks) {  // cannot happen while every sub-problem honours max_iter
            for (uint32_t s = 0; s < P.n_sub; ++s) if (!P.h_ps[s].done) P.h_ps[s].status = 4;
            break;
        }
        int ev_s = 4, ev_e = 5;  // the two events other than the previous chunk's end
        if (prev_end == 4) ev_s = 6;
        else if (prev_end == 5) ev_e = 6;
        (void)hipEventRecord(P.ev[ev_s], P.stream);
        if (run_chunk<PRE>(P, mixed, d_x, chunk, full)) return 5;
        (void)hipEventRecord(P.ev[ev_e], P.stream);
        if (read_scalars(P)) return 5;
        float t = 0;
        if (hipEventElapsedTime(&t, P.ev[ev_s], P.ev[ev_e]) == hipSuccess) P.last_chunk_ms += t;
        if (prev_end >= 0 && !prev_refresh && hipEventElapsedTime(&t, P.ev[prev_end], P.ev[ev_s]) == hipSuccess) {
            P.last_gap_ms += t;
            ++P.last_gaps;
        }
        ++P.last_chunks;
        prev_end = ev_e;
        const int grown = P.use_graph ? full : std::min(chunk * 2, full);
        if (predict) {
            const ProbScalars &S = P.h_ps[0];
            if (it_prev >= 0 && S.iter > it_prev && S.rr > 0 && rr_prev > S.rr)
                rate = std::log(S.rr / rr_prev) / (double)(S.iter - it_prev);
            // the residual of CG falls in steps: a chunk that saw only a plateau predicts a long
            // way to the pause, so the faster of its rate and the mean rate since the cycle's
            // first chunk (the solve's start or the last refresh) sets the prediction (Hexa8
            // 100^3 MIXED: 433 launches past a pause with the chunk's rate alone)
            if (it_cyc < 0) {
                it_cyc = S.iter;
                rr_cyc = S.rr;
            } else if (S.iter - it_cyc >= 8 && S.rr > 0 && rr_cyc > S.rr) {
                const double mean = std::log(S.rr / rr_cyc) / (double)(S.iter - it_cyc);
                if (!(rate < mean)) rate = mean;
            }
            chunk = predict_chunk(S, S.rr, S.rr_ref, rate, grown, cap);
            rr_prev = S.rr;
            it_prev = S.iter;
        } else {
            chunk = grown;
        }
        bool all_done = true, refresh = false;
        for (uint32_t s = 0; s < P.n_sub; ++s) {
            all_done &= P.h_ps[s].done != 0;
            refresh |= !P.h_ps[s].done && P.h_ps[s].pause == PAUSE_REQ;
        }
        if (all_done) break;
        prev_refresh = refresh;
        if (refresh) {
            // Check of the paused sub-problems on the true residual: MIXED first adds the FP32
            // correction (x += x32); r = b - K x in FP64 with the stop test on it; those that
            // fail it restart the recurrence from r (FP32 or FP64), keeping the direction.
            // FP64 forms that residual in q (dead between iterations), not in its recurrence
            // residual r: the matrix-free product writes every sub-problem's rows, and the
            // sub-problems still running keep their r; the restart copies r of the checked ones.
            if (mixed) hipLaunchKernelGGL(k_accumulate, g, blk, 0, P.stream, d_x, P.d_x32, wt, P.d_ps, (int)SEL_PAUSE_REQ);
            launch_spmv_t<double, MODE_RES>(P, P.d_K, d_x, d_rhs, mixed ? P.d_r : P.d_q, SEL_PAUSE_REQ);
            finalize(P, FIN_RES_CHECK, SEL_PAUSE_REQ);
            if (mixed) launch_restart<PRE, float>(P, P.d_r, P.d_r32, P.d_p32, P.d_Minv32, zbuf<PRE, float>(P), 0);
            else launch_restart<PRE, double>(P, P.d_q, P.d_r, P.d_p, P.d_Minv, zbuf<PRE, double>(P), 0);
            hipLaunchKernelGGL(k_clear_pause, dim3((P.n_sub + 255) / 256), blk, 0, P.stream, P.d_ps, P.n_sub);
            it_prev = -1;  // the recurrence restarts from the FP64 residual
            it_cyc = -1;
            chunk = P.use_graph ? full : std::max(chunk, 2);
            // the restarted residual is the true one: the next pause comes after about the
            // iterations the last cycle took; half of that, so that a faster start wastes nothing
            if (predict) chunk = std::max(2, (predict_chunk(P.h_ps[0], P.h_ps[0].rr, P.h_ps[0].rr, rate, 2 * chunk, 2 * cap) / 2) & ~1);
        }
    }
    if (mixed) hipLaunchKernelGGL(k_accumulate, g, blk, 0, P.stream, d_x, P.d_x32, wt, P.d_ps, (int)SEL_EVERY);
    (void)hipEventRecord(P.ev[1], P.stream);
    // final true residual of every sub-problem, for the report
    launch_spmv_t<double, MODE_RES>(P, P.d_K, d_x, d_rhs, P.d_q, SEL_EVERY);
    finalize(P, FIN_RES, SEL_EVERY);
    if (hipGetLastError() != hipSuccess) return 5;
    if (read_scalars(P)) return 5;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, P.ev[0], P.ev[1]);
    int first_bad = 0;
    for (uint32_t s = 0; s < P.n_sub; ++s) {
        const ProbScalars &S = P.h_ps[s];
        out[s] = SolveOut{S.status, S.iter, S.refreshes, std::sqrt(S.rr / (bb[s] > 0 ? bb[s] : 1.0)), S.rmax, (double)ms};
        if (S.status && !first_bad) first_bad = S.status;
    }
    return first_bad;
}

int precond_build(Problem &P, int precond) {
    if (precond == 2 || precond == 3) return ic0_ensure(P, precond == 2 ? IC0_NATURAL : IC0_COLOURED, true);
    if (precond < 0 || precond > 4) return 1;
    if (precond == 4) {
        const int rc = spmv_ensure(P, 0);  // the power iteration's FP64 product
        if (rc) return rc;
    }
    if (!P.d_Minv32 && hipMalloc(&P.d_Minv32, sizeof(uint16_t) * 6 * std::max<uint32_t>(1, P.n_nodes)) != hipSuccess)
        return 5;
    std::memset(P.h_ps, 0, sizeof(ProbScalars) * P.n_sub);
    // max |diag K| stays: an F-only assembly after this call scales the penalty terms with it
    // (formulation.h:1029-1030,1059-1062)
    for (uint32_t s = 0; s < P.n_sub; ++s) P.h_ps[s].max_diag = s < P.max_diag.size() ? P.max_diag[s] : 0.0;
    if (write_scalars(P)) return 5;
    const WgTables wt = tables(P);
    const dim3 g(P.n_wg), blk(256);
    (void)hipEventRecord(P.ev[2], P.stream);
    if (precond != 0)
        hipLaunchKernelGGL(k_build_minv<PRE_BLOCK_JACOBI>, g, blk, 0, P.stream, P.d_diag, P.d_K, P.ps, P.d_Minv, P.d_Minv32, wt, P.d_ps);
    else
        hipLaunchKernelGGL(k_build_minv<PRE_JACOBI>, g, blk, 0, P.stream, P.d_diag, P.d_K, P.ps, P.d_Minv, P.d_Minv32, wt, P.d_ps);
    (void)hipEventRecord(P.ev[3], P.stream);
    if (hipGetLastError() != hipSuccess || read_scalars(P)) return 5;
    bool ok = true;
    for (uint32_t s = 0; s < P.n_sub; ++s) ok &= !P.h_ps[s].status;
    if (ok && precond == 4) {
        const int rc = cheb_setup(P, false);
        if (rc == 5) return 5;
        (void)hipEventRecord(P.ev[3], P.stream);
        if (hipEventSynchronize(P.ev[3]) != hipSuccess) return 5;
    }
    for (uint32_t s = 0; s < P.n_sub; ++s)
        if (P.h_ps[s].status) {
            P.minv_kind = -1;
            return P.h_ps[s].status;
        }
    P.minv_kind = precond == 4 ? PRE_CHEB : precond == 1 ? PRE_BLOCK_JACOBI : PRE_JACOBI;
    float ms = 0;
    if (hipEventElapsedTime(&ms, P.ev[2], P.ev[3]) == hipSuccess) P.last_precond_ms = ms;
    return 0;
}

// sum (a - b)^2 and max |a - b| over n values: one partial pair per workgroup (fixed grid, fixed
// tree order), combined on the host in workgroup order -- deterministic.
constexpr int DIFF_WGS = 512;
__global__ __launch_bounds__(256) void k_diff_norm(const double *__restrict__ a, const double *__restrict__ b, size_t n,
                                                   double *__restrict__ out) {
    __shared__ double s2[256], sm[256];
    double acc = 0, mx = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double d = a[i] - b[i];
        acc += d * d;
        mx = fmax(mx, fabs(d));
    }
    s2[threadIdx.x] = acc;
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s2[threadIdx.x] += s2[threadIdx.x + w];
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = s2[0];
        out[2 * blockIdx.x + 1] = sm[0];
    }
}

int diff_norm(Problem &P, const double *a, const double *b, size_t n, double *norm2, double *norm_inf) {
    if (!P.d_red && hipMalloc(&P.d_red, sizeof(double) * 2 * DIFF_WGS) != hipSuccess) return 5;
    hipLaunchKernelGGL(k_diff_norm, dim3(DIFF_WGS), dim3(256), 0, P.stream, a, b, n, P.d_red);
    double h[2 * DIFF_WGS];
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h, P.d_red, sizeof(h), hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
        hipStreamSynchronize(P.stream) != hipSuccess)
        return 5;
    double s = 0, m = 0;
    for (int w = 0; w < DIFF_WGS; ++w) {
        s += h[2 * w];
        m = std::max(m, h[2 * w + 1]);
    }
    if (norm2) *norm2 = std::sqrt(s);
    if (norm_inf) *norm_inf = m;
    return 0;
}

int time_iteration(Problem &P, int precision, int iters, double *block_ms, double *update_ms) {
    const int warm = 4, n = iters + warm;
    if (iters < 1 || P.spmv.format[precision ? 1 : 0] != FMT_MF || !P.mf.fold || !P.assembled_mat || !P.assembled_vec)
        return 1;
    std::vector<hipEvent_t> ev(3 * (size_t)n, nullptr);
    int rc = 0;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = 5;
    const bool graph = P.use_graph;
    if (!rc) {
        P.use_graph = false;  // one launch at a time, the events between them
        P.kt = ev.data();
        P.kt_cap = n;
        P.kt_used = 0;
        // every iteration runs: a relative residual far below any the solve reaches
        const SolveOpts o{precision ? 1 : 0, 1, 0, 1e-300, n, 1e-300, 32};
        std::vector<SolveOut> out(P.n_sub);
        if (launch_fill(P.d_aux, 3 * (size_t)P.n_nodes, 0.0, P.stream) != hipSuccess) rc = 5;
        if (!rc) rc = solve_impl<PRE_BLOCK_JACOBI>(P, P.d_F, P.d_aux, o, out.data(), 0);
        if (rc == 4 || rc == 8) rc = 0;  // the iteration cap (or a stall) ends this solve
        if (!rc && hipStreamSynchronize(P.stream) != hipSuccess) rc = 5;
        double tb = 0, tu = 0;
        int m = 0;
        for (int k = warm; !rc && k < P.kt_used; ++k) {
            float a = 0, b = 0;
            if (hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]) != hipSuccess ||
                hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]) != hipSuccess) {
                rc = 5;
                break;
            }
            tb += a;
            tu += b;
            ++m;
        }
        if (!rc && m == 0) rc = 1;
        // an event between two dependent kernels adds the gap of its own record to the second
        // one: the iteration's time comes from a second solve with events around its chunks only
        // (the closing pass of each chunk included), split in the proportion just measured
        P.kt = nullptr;
        P.kt_cap = P.kt_used = 0;
        if (!rc && launch_fill(P.d_aux, 3 * (size_t)P.n_nodes, 0.0, P.stream) != hipSuccess) rc = 5;
        if (!rc) {
            rc = solve_impl<PRE_BLOCK_JACOBI>(P, P.d_F, P.d_aux, o, out.data(), 0);
            if (rc == 4 || rc == 8) rc = 0;
        }
        if (!rc && !(out[0].iters > 0 && P.last_chunk_ms > 0)) rc = 1;
        if (!rc) {
            const double t_it = P.last_chunk_ms / out[0].iters, f_b = tb / (tb + tu);
            if (block_ms) *block_ms = t_it * f_b;
            if (update_ms) *update_ms = t_it * (1.0 - f_b);
        }
    }
    P.kt = nullptr;
    P.kt_cap = P.kt_used = 0;
    P.use_graph = graph;
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int solve_pcg(Problem &P, const double *d_rhs, double *d_x, const SolveOpts &o, SolveOut *out) {
    for (uint32_t s = 0; s < P.n_sub; ++s) out[s] = SolveOut{};
    int rc = spmv_ensure(P, 0);
    if (!rc && o.precision == 1) rc = spmv_ensure(P, 1);
    if (rc) return rc;
    switch (o.precond) {
        case 1: return solve_impl<PRE_BLOCK_JACOBI>(P, d_rhs, d_x, o, out, 0);
        case 2: return solve_impl<PRE_IC0>(P, d_rhs, d_x, o, out, IC0_NATURAL);
        case 3: return solve_impl<PRE_IC0>(P, d_rhs, d_x, o, out, IC0_COLOURED);
        case 4: return solve_impl<PRE_CHEB>(P, d_rhs, d_x, o, out, 0);
        default: return solve_impl<PRE_JACOBI>(P, d_rhs, d_x, o, out, 0);
    }
}

}  // namespace femgpu
