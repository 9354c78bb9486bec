// Block IC(0) preconditioner on the GPU: the factorisation of MatWithTinyBlocks::chol_incomp
// (containers/MatWithTinyBlocks.h:426-523) and the triangular solves of :587-709, run level by
// level over the schedule of host_mesh.cpp (build_ic0_schedule).
//
// Rows of one level share no block, so each level is one launch with one lane per block row.
// With the natural elimination order the factor is the reference's (levels grow with the
// lattice diagonal: i + 2j + 4k on a Hexa lattice); with the coloured order the levels are the
// colour classes (8 for Hexa, the same count for the Tetra splits) and every sweep is a handful
// of wide launches.  The factor lives in SELL-64 over the level-sorted rows t, 9 value planes like
// K: a wave reads 64 consecutive positions per slot.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.hpp"

namespace femgpu {

__device__ inline uint64_t ic0_base(const Ic0Args &a, uint32_t t) { return a.sptr[t >> 6] + (t & 63u); }

template <typename T>
__device__ inline void ld9(const T *__restrict__ V, uint64_t ps, uint64_t p, double (&b)[9]) {
#pragma unroll
    for (int c = 0; c < 9; ++c) b[c] = (double)V[c * ps + p];
}

// One level of the factorisation, rows [t0, t1).  Row t, for each strictly lower block (t, c) in
// increasing c:  L_tc = (A_tc - sum_{k<c} L_tk L_ck^T) L_cc^-T, the sum over the blocks k present
// in both rows (a merge of the two sorted rows), the product by L_cc^-T as the forward
// substitution of :468-487; L_tc is also stored in row c's upper slot for the backward sweep.
// Diagonal: 3x3 Cholesky of A_tt - sum_k L_tk L_tk^T (:497-522); a non-positive pivot marks the
// sub-problem NOT_SPD (the reference takes the square root regardless).
__global__ __launch_bounds__(256) void k_ic0_factor(uint32_t t0, uint32_t t1, Ic0Args a, const double *__restrict__ Kv,
                                                    uint64_t kps, double *L, ProbScalars *S, int *flags) {
    const uint32_t t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    const uint64_t ps = a.ps, base = ic0_base(a, t);
    const uint32_t nl = a.nlow[t];
    for (uint32_t k = 0; k < nl; ++k) {
        const uint64_t p = base + 64ull * k;
        const uint32_t c = a.colT[p];
        const uint64_t cb = ic0_base(a, c);
        const uint32_t cl = a.nlow[c];
        double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (k && cl) {
            uint32_t i1 = 0, i2 = 0, e1 = a.colT[base], e2 = a.colT[cb];
            for (;;) {
                if (e1 < e2) {
                    if (++i1 == k) break;
                    e1 = a.colT[base + 64ull * i1];
                } else if (e1 > e2) {
                    if (++i2 == cl) break;
                    e2 = a.colT[cb + 64ull * i2];
                } else {
                    double A[9], B[9];
                    ld9(L, ps, base + 64ull * i1, A);
                    ld9(L, ps, cb + 64ull * i2, B);
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int y = 0; y < 3; ++y)
                            s[3 * x + y] += A[3 * x] * B[3 * y] + A[3 * x + 1] * B[3 * y + 1] + A[3 * x + 2] * B[3 * y + 2];
                    ++i1;
                    ++i2;
                    if (i1 == k || i2 == cl) break;
                    e1 = a.colT[base + 64ull * i1];
                    e2 = a.colT[cb + 64ull * i2];
                }
            }
        }
        double O[9], D[9];
        ld9(Kv, kps, a.kpos[p], O);
        ld9(L, ps, cb + 64ull * cl, D);
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            O[3 * x + 0] = (O[3 * x + 0] - s[3 * x + 0]) / D[0];
            O[3 * x + 1] = (O[3 * x + 1] - s[3 * x + 1] - O[3 * x + 0] * D[3]) / D[4];
            O[3 * x + 2] = (O[3 * x + 2] - s[3 * x + 2] - O[3 * x + 0] * D[6] - O[3 * x + 1] * D[7]) / D[8];
        }
        const uint64_t q = a.tpos[p];
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            L[m * ps + p] = O[m];
            L[m * ps + q] = O[m];
        }
    }
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t k = 0; k < nl; ++k) {
        double A[9];
        ld9(L, ps, base + 64ull * k, A);
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y <= x; ++y) s[3 * x + y] += A[3 * x] * A[3 * y] + A[3 * x + 1] * A[3 * y + 1] + A[3 * x + 2] * A[3 * y + 2];
    }
    const uint64_t pd = base + 64ull * nl;
    double D[9];
    ld9(Kv, kps, a.kpos[pd], D);
    const double d00 = D[0] - s[0];
    D[0] = sqrt(d00);
    D[3] = (D[3] - s[3]) / D[0];
    D[6] = (D[6] - s[6]) / D[0];
    const double d11 = D[4] - s[4] - D[3] * D[3];
    D[4] = sqrt(d11);
    D[7] = (D[7] - s[7] - D[6] * D[3]) / D[4];
    const double d22 = D[8] - s[8] - D[6] * D[6] - D[7] * D[7];
    D[8] = sqrt(d22);
    D[1] = D[2] = D[5] = 0;
    if (!(d00 > 0) || !(d11 > 0) || !(d22 > 0)) {
        S[a.sub[t]].status = 3;
        flags[1] = 1;
    }
#pragma unroll
    for (int m = 0; m < 9; ++m) L[m * ps + pd] = D[m];
}

// The sweeps stream the factor once each: its values and columns are read non-temporally.
template <typename V>
__device__ inline V nt(const V *p) { return __builtin_nontemporal_load(p); }

// Forward sweep of one level: z_t = L_tt^-1 (r_t - sum_{c<t} L_tc z_c).
template <typename T>
__global__ __launch_bounds__(256) void k_ic0_fwd(uint32_t t0, uint32_t t1, Ic0Args a, const T *__restrict__ L,
                                                 const T *__restrict__ r, T *z, const ProbScalars *S, int sel) {
    const uint32_t t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1 || !selected(S[a.sub[t]], sel)) return;
    const uint64_t ps = a.ps, base = ic0_base(a, t);
    const uint32_t nl = a.nlow[t], n = a.nat[t];
    T s0 = r[3 * (size_t)n], s1 = r[3 * (size_t)n + 1], s2 = r[3 * (size_t)n + 2];
    for (uint32_t k = 0; k < nl; ++k) {
        const uint64_t p = base + 64ull * k;
        const uint32_t c = nt(a.colN + p);
        const T x0 = z[3 * (size_t)c], x1 = z[3 * (size_t)c + 1], x2 = z[3 * (size_t)c + 2];
        s0 -= nt(L + p) * x0 + nt(L + ps + p) * x1 + nt(L + 2 * ps + p) * x2;
        s1 -= nt(L + 3 * ps + p) * x0 + nt(L + 4 * ps + p) * x1 + nt(L + 5 * ps + p) * x2;
        s2 -= nt(L + 6 * ps + p) * x0 + nt(L + 7 * ps + p) * x1 + nt(L + 8 * ps + p) * x2;
    }
    const uint64_t pd = base + 64ull * nl;
    const T z0 = s0 / L[pd];
    const T z1 = (s1 - L[3 * ps + pd] * z0) / L[4 * ps + pd];
    const T z2 = (s2 - L[6 * ps + pd] * z0 - L[7 * ps + pd] * z1) / L[8 * ps + pd];
    z[3 * (size_t)n] = z0;
    z[3 * (size_t)n + 1] = z1;
    z[3 * (size_t)n + 2] = z2;
}

// Backward sweep of one level, in place: z_t = L_tt^-T (z_t - sum_{c>t} L_ct^T z_c); the upper
// slot (t, c) holds L_ct.
template <typename T>
__global__ __launch_bounds__(256) void k_ic0_bwd(uint32_t t0, uint32_t t1, Ic0Args a, const T *__restrict__ L, T *z,
                                                 const ProbScalars *S, int sel) {
    const uint32_t t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1 || !selected(S[a.sub[t]], sel)) return;
    const uint64_t ps = a.ps, base = ic0_base(a, t);
    const uint32_t nl = a.nlow[t], len = a.len[t], n = a.nat[t];
    T s0 = z[3 * (size_t)n], s1 = z[3 * (size_t)n + 1], s2 = z[3 * (size_t)n + 2];
    for (uint32_t k = nl + 1; k < len; ++k) {
        const uint64_t p = base + 64ull * k;
        const uint32_t c = nt(a.colN + p);
        const T y0 = z[3 * (size_t)c], y1 = z[3 * (size_t)c + 1], y2 = z[3 * (size_t)c + 2];
        s0 -= nt(L + p) * y0 + nt(L + 3 * ps + p) * y1 + nt(L + 6 * ps + p) * y2;
        s1 -= nt(L + ps + p) * y0 + nt(L + 4 * ps + p) * y1 + nt(L + 7 * ps + p) * y2;
        s2 -= nt(L + 2 * ps + p) * y0 + nt(L + 5 * ps + p) * y1 + nt(L + 8 * ps + p) * y2;
    }
    const uint64_t pd = base + 64ull * nl;
    const T z2 = s2 / L[8 * ps + pd];
    const T z1 = (s1 - L[7 * ps + pd] * z2) / L[4 * ps + pd];
    const T z0 = (s0 - L[3 * ps + pd] * z1 - L[6 * ps + pd] * z2) / L[pd];
    z[3 * (size_t)n] = z0;
    z[3 * (size_t)n + 1] = z1;
    z[3 * (size_t)n + 2] = z2;
}

__global__ void k_ic0_to_fp32(const double *__restrict__ a, float *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------

namespace {

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <class T>
hipError_t upload(T **p, const std::vector<T> &h) {
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(1, h.size()) * sizeof(T));
    if (e == hipSuccess && !h.empty()) e = hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

dim3 level_grid(uint32_t rows) { return dim3((rows + 255) / 256); }

}  // namespace

void ic0_free(Ic0 &I) {
    dfree(I.d_nat); dfree(I.d_sub); dfree(I.d_len); dfree(I.d_nlow); dfree(I.d_sptr);
    dfree(I.d_colN); dfree(I.d_colT); dfree(I.d_kpos); dfree(I.d_tpos);
    dfree(I.d_L); dfree(I.d_z); dfree(I.d_L32); dfree(I.d_z32);
    I.sched = Ic0Schedule();
    I.ps = 0;
    I.valid = I.valid32 = false;
}

static int ic0_prepare(Problem &P, int ordering) {
    Ic0 &I = P.ic0;
    if (I.sched.ordering == ordering && I.d_L) return 0;
    destroy_graph(P);  // captured sweeps refer to the old buffers
    ic0_free(I);
    // natural order = the caller's node order (internal node of each API node)
    if (!build_ic0_schedule(P.pat, ordering, I.sched, P.ninv.empty() ? nullptr : &P.ninv) ||
        I.sched.n_sell >= (1ull << 32)) {
        I.sched = Ic0Schedule();
        return 1;
    }
    const Ic0Schedule &S = I.sched;
    I.ps = (S.n_sell + 63) / 64 * 64;
    std::vector<uint32_t> sub(P.n_nodes);
    for (uint32_t t = 0; t < P.n_nodes; ++t)
        sub[t] = (uint32_t)(std::upper_bound(P.node_off.begin(), P.node_off.end(), S.order[t]) - P.node_off.begin()) - 1;
    const size_t n = 3 * (size_t)std::max<uint32_t>(1, P.n_nodes);
    if (upload(&I.d_nat, S.order) || upload(&I.d_sub, sub) || upload(&I.d_len, S.len) || upload(&I.d_nlow, S.nlow) ||
        upload(&I.d_sptr, S.sptr) || upload(&I.d_colN, S.colN) || upload(&I.d_colT, S.colT) ||
        upload(&I.d_kpos, S.kpos) || upload(&I.d_tpos, S.tpos) ||
        hipMalloc(&I.d_L, sizeof(double) * 9 * I.ps) || hipMalloc(&I.d_z, sizeof(double) * n))
        return 5;
    // padding positions are never read; clear them once so downloads are deterministic
    if (hipMemsetAsync(I.d_L, 0, sizeof(double) * 9 * I.ps, P.stream) != hipSuccess) return 5;
    return 0;
}

int ic0_ensure(Problem &P, int ordering, bool fp32) {
    int rc = ic0_prepare(P, ordering);
    if (rc) return rc;
    Ic0 &I = P.ic0;
    P.last_precond_ms = 0;
    if (I.valid && (!fp32 || I.valid32)) return 0;
    (void)hipEventRecord(P.ev[2], P.stream);
    if (!I.valid) {
        I.valid32 = false;
        if (hipMemsetAsync(P.d_flags + 1, 0, sizeof(int), P.stream) != hipSuccess) return 5;
        const Ic0Args a = I.args();
        const auto &lv = I.sched.lv_ptr;
        for (uint32_t l = 0; l + 1 < lv.size(); ++l)
            hipLaunchKernelGGL(k_ic0_factor, level_grid(lv[l + 1] - lv[l]), dim3(256), 0, P.stream, lv[l], lv[l + 1], a,
                               P.d_K, P.ps, I.d_L, P.d_ps, P.d_flags);
        if (hipGetLastError() != hipSuccess) return 5;
        if (hipMemcpyAsync(P.h_flags + 1, P.d_flags + 1, sizeof(int), hipMemcpyDeviceToHost, P.stream) != hipSuccess ||
            hipStreamSynchronize(P.stream) != hipSuccess)
            return 5;
        if (P.h_flags[1]) return 3;
        I.valid = true;
    }
    if (fp32 && !I.valid32) {
        const size_t n = 3 * (size_t)std::max<uint32_t>(1, P.n_nodes);
        if (!I.d_L32 && (hipMalloc(&I.d_L32, sizeof(float) * 9 * I.ps) != hipSuccess ||
                         hipMalloc(&I.d_z32, sizeof(float) * n) != hipSuccess))
            return 5;
        hipLaunchKernelGGL(k_ic0_to_fp32, dim3(8192), dim3(256), 0, P.stream, I.d_L, I.d_L32, 9 * I.ps);
        if (hipGetLastError() != hipSuccess) return 5;
        I.valid32 = true;
    }
    (void)hipEventRecord(P.ev[3], P.stream);
    if (hipEventSynchronize(P.ev[3]) != hipSuccess) return 5;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, P.ev[2], P.ev[3]);
    P.last_precond_ms = ms;
    return 0;
}

template <typename T>
static void apply_t(Problem &P, const T *L, const T *r, T *z, int sel) {
    const Ic0Args a = P.ic0.args();
    const auto &lv = P.ic0.sched.lv_ptr;
    const uint32_t nl = (uint32_t)lv.size() - 1;
    for (uint32_t l = 0; l < nl; ++l)
        hipLaunchKernelGGL((k_ic0_fwd<T>), level_grid(lv[l + 1] - lv[l]), dim3(256), 0, P.stream, lv[l], lv[l + 1], a, L, r, z,
                           P.d_ps, sel);
    for (uint32_t l = nl; l-- > 0;)
        hipLaunchKernelGGL((k_ic0_bwd<T>), level_grid(lv[l + 1] - lv[l]), dim3(256), 0, P.stream, lv[l], lv[l + 1], a, L, z,
                           P.d_ps, sel);
}

void ic0_apply(Problem &P, const double *r, double *z, int sel) { apply_t<double>(P, P.ic0.d_L, r, z, sel); }
void ic0_apply(Problem &P, const float *r, float *z, int sel) { apply_t<float>(P, P.ic0.d_L32, r, z, sel); }

int ic0_download(Problem &P, uint32_t *order, uint32_t *row_ptr, uint32_t *col, double *L) {
    const Ic0 &I = P.ic0;
    const Ic0Schedule &S = I.sched;
    const uint32_t n = S.n_rows;
    std::vector<double> planes;
    if (L) {
        planes.resize(9 * I.ps);
        if (hipStreamSynchronize(P.stream) != hipSuccess ||
            hipMemcpy(planes.data(), I.d_L, sizeof(double) * planes.size(), hipMemcpyDeviceToHost) != hipSuccess)
            return 5;
    }
    if (order) std::memcpy(order, S.order.data(), sizeof(uint32_t) * n);
    uint32_t o = 0;
    for (uint32_t t = 0; t < n; ++t) {
        if (row_ptr) row_ptr[t] = o;
        for (uint32_t k = 0; k < S.len[t]; ++k, ++o) {
            const uint64_t p = S.pos(t, k);
            if (col) col[o] = S.colT[p];
            if (L)
                for (int m = 0; m < 9; ++m) L[9 * (size_t)o + m] = k <= S.nlow[t] ? planes[m * I.ps + p] : 0.0;
        }
    }
    if (row_ptr) row_ptr[n] = o;
    return 0;
}

}  // namespace femgpu
