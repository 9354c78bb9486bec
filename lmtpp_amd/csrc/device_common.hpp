// Device helpers shared by the assembly and solver kernels: wave/workgroup reductions segmented
// by sub-problem, the XCD-aware workgroup mapping and the sub-problem selectors.
#pragma once
#include "femgpu_internal.hpp"

namespace femgpu {

enum { SEL_EVERY = 4 };

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// XCD-aware workgroup index: hardware hands workgroup i to XCD i % 8, so logical index
// (i % 8) * (n / 8) + i / 8 gives each XCD one contiguous range of rows and keeps its
// x-vector neighbourhood in that XCD's L2.  Grids are multiples of 8.
__device__ inline unsigned xcd_slab() { return (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3); }

__device__ inline bool selected(const ProbScalars &S, int sel) {
    switch (sel) {
        case SEL_ACTIVE: return !S.done && S.pause == PAUSE_NONE;
        case SEL_PAUSE_REQ: return !S.done && S.pause == PAUSE_REQ;
        case SEL_PAUSE_DONE: return !S.done && S.pause == PAUSE_DONE;
        case SEL_ALL: return !S.done;
        default: return true;
    }
}

// Per-workgroup partial of NV values (bit k of MAXMASK set = max, else sum): a wave-shuffle tree
// then a fixed 4-wave combination, stored at partials[k * stride + L].  No fences: the partials
// are combined by a finalize kernel after the kernel boundary.
template <int NV, unsigned MAXMASK>
__device__ void wg_partial(double (&v)[NV], double *partials, unsigned stride, unsigned L) {
    __shared__ double sm[NV][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double t = (MAXMASK >> k & 1) ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) sm[k][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k)
            partials[(size_t)k * stride + L] = (MAXMASK >> k & 1) ? fmax(fmax(sm[k][0], sm[k][1]), fmax(sm[k][2], sm[k][3]))
                                                                  : ((sm[k][0] + sm[k][1]) + sm[k][2]) + sm[k][3];
    }
}

// The same for a workgroup of NW waves (the waves combined in order).
template <int NV, unsigned MAXMASK, int NW>
__device__ void wg_partial_n(double (&v)[NV], double *partials, unsigned stride, unsigned L) {
    __shared__ double sm[NV][NW];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double t = (MAXMASK >> k & 1) ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) sm[k][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        double o = sm[k][0];
        for (int w = 1; w < NW; ++w) o = (MAXMASK >> k & 1) ? fmax(o, sm[k][w]) : o + sm[k][w];
        partials[(size_t)k * stride + L] = o;
    }
}

// One workgroup (up to 1024 lanes) combines the partials of workgroups [w0, w0 + nw) in a fixed
// order: each round issues the loads of all NV values for PER strided elements per lane at once
// (one memory round trip per round: a finalize after a row kernel of 1024 x 4 workgroups needs
// one), then a pairwise tree over the PER accumulators, a wave tree and a fixed combination of
// the waves.  out[] is valid in every thread.
template <int NV, unsigned MAXMASK, int PER = 4>
__device__ void wg_combine(const double *partials, unsigned stride, unsigned w0, unsigned nw, double (&out)[NV]) {
    static_assert(PER >= 1 && (PER & (PER - 1)) == 0, "PER: a power of two");
    __shared__ double sm[NV][16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, n_waves = (blockDim.x + 63) >> 6;
    const unsigned B = blockDim.x;
    double a[NV][PER];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int j = 0; j < PER; ++j) a[k][j] = 0.0;
    for (unsigned i0 = threadIdx.x; i0 < nw; i0 += PER * B) {
        double t[NV][PER];
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const unsigned i = i0 + j * B;
                t[k][j] = i < nw ? partials[(size_t)k * stride + w0 + i] : 0.0;
            }
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int j = 0; j < PER; ++j) a[k][j] = (MAXMASK >> k & 1) ? fmax(a[k][j], t[k][j]) : a[k][j] + t[k][j];
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
#pragma unroll
        for (int h = 1; h < PER; h *= 2)
#pragma unroll
            for (int j = 0; j + h < PER; j += 2 * h)
                a[k][j] = (MAXMASK >> k & 1) ? fmax(a[k][j], a[k][j + h]) : a[k][j] + a[k][j + h];
        const double t = (MAXMASK >> k & 1) ? wave_max(a[k][0]) : wave_sum(a[k][0]);
        if (lane == 0) sm[k][wave] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double o = sm[k][0];
        for (int w = 1; w < n_waves; ++w) o = (MAXMASK >> k & 1) ? fmax(o, sm[k][w]) : o + sm[k][w];
        out[k] = o;
    }
}

// A 256-lane workgroup combines the partials of workgroups [w0, w0 + nw): lane l takes elements
// l, l + 256, l + 512, ... in that order (lane_partials, loads issued BATCH at a time ahead of
// their adds), then a wave tree and the four waves in order (wg_combine_lanes).  The value of
// index k depends only on the partials of index k, so every workgroup that combines the same
// partials gets the same bits whatever NV it asks for and however its loads are batched (the
// two-kernel CG iteration relies on it: every element block forms beta itself).
template <int NV, unsigned MAXMASK, int BATCH>
__device__ void lane_partials(const double *partials, unsigned stride, unsigned w0, unsigned nw, double (&a)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) a[k] = 0.0;
    for (unsigned i0 = threadIdx.x; i0 < nw; i0 += BATCH * 256u) {
        double t[NV][BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; ++j)
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const unsigned i = i0 + j * 256u;
                t[k][j] = i < nw ? partials[(size_t)k * stride + w0 + i] : 0.0;
            }
#pragma unroll
        for (int j = 0; j < BATCH; ++j)
#pragma unroll
            for (int k = 0; k < NV; ++k) a[k] = (MAXMASK >> k & 1) ? fmax(a[k], t[k][j]) : a[k] + t[k][j];
    }
}

template <int NV, unsigned MAXMASK>
__device__ void wg_combine_lanes(const double (&a)[NV], double (&out)[NV]) {
    __shared__ double sm[NV][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double t = (MAXMASK >> k & 1) ? wave_max(a[k]) : wave_sum(a[k]);
        if (lane == 0) sm[k][wave] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k)
        out[k] = (MAXMASK >> k & 1) ? fmax(fmax(sm[k][0], sm[k][1]), fmax(sm[k][2], sm[k][3]))
                                    : ((sm[k][0] + sm[k][1]) + sm[k][2]) + sm[k][3];
}

template <int NV, unsigned MAXMASK, int BATCH>
__device__ void wg_combine256(const double *partials, unsigned stride, unsigned w0, unsigned nw, double (&out)[NV]) {
    double a[NV];
    lane_partials<NV, MAXMASK, BATCH>(partials, stride, w0, nw, a);
    wg_combine_lanes<NV, MAXMASK>(a, out);
}

__device__ inline bool finite(double v) { return v == v && fabs(v) <= 1.7976931348623157e308; }

constexpr int STALL_CHECKS = 8, STALL_ITERS = 1000;

// One sub-problem's scalars copied by the first lanes of a workgroup, one 8-byte word each (no
// lane holds the whole record).
__device__ inline void copy_scalars(ProbScalars *dst, const ProbScalars *src) {
    static_assert(sizeof(ProbScalars) % 8 == 0 && sizeof(ProbScalars) / 8 <= 64, "scalar record layout");
    if (threadIdx.x < sizeof(ProbScalars) / 8)
        reinterpret_cast<uint64_t *>(dst)[threadIdx.x] = reinterpret_cast<const uint64_t *>(src)[threadIdx.x];
}

// End of a CG iteration's vector update from its combined partials o = (r.z, r.r, max|r|) (with_rz
// false: IC(0), whose r.z follows the sweeps): advance the iteration count and rho, form beta and
// apply the stop test, reading the scalars from in and writing the changed ones into out (the same
// record, or the copy the next kernel reads: in is then read-only in the kernel, its fields load
// ahead of the partials' combine).  The recurrence residual only requests a check: the solve ends
// on the true residual b - K x (FIN_RES_CHECK), in FP64 as in MIXED, where an FP32 recurrence that
// has fallen by delta since the last FP64 residual also pauses for a refresh.
__device__ inline void fin_update(const ProbScalars &in, ProbScalars &out, const double (&o)[3], bool with_rz) {
    const int iter = in.iter + 1, checks = in.checks, iter_best = in.iter_best, max_iter = in.max_iter;
    const double rho = in.rho, stop2 = in.stop2, crit = in.crit, delta2 = in.delta2, rr_ref = in.rr_ref;
    const bool mixed = in.mixed != 0, rel = in.crit_kind == 0, cheb = in.cheb_c1 != 0;
    int done = in.done, status = in.status, pause = in.pause;
    if (!finite(o[0]) || !finite(o[1])) {
        status = 6;
        done = 1;
    } else if (with_rz && cheb && o[0] < 0) {
        status = 3;
        done = 1;
    } else {
        const bool conv = rel ? o[1] <= stop2 : o[2] <= crit;
        if (conv || (mixed && o[1] <= delta2 * rr_ref)) pause = PAUSE_REQ;
        // no progress of the true residual since the last check that improved it, for as many
        // iterations as all progress took (at least STALL_ITERS): stagnated above the tolerance
        if (!done && checks >= 2 && iter - iter_best > max(iter_best, STALL_ITERS)) { status = 8; done = 1; }
        if (!done && iter >= max_iter) { status = 4; done = 1; }
    }
    out.iter = iter;
    out.rho_prev = rho;
    if (with_rz) {
        out.rho = o[0];
        out.beta = o[0] / rho;
    }
    out.rr = o[1];
    out.rmax = o[2];
    out.status = status;
    out.done = done;
    out.pause = pause;
}
__device__ inline void fin_update(ProbScalars &P, const double (&o)[3], bool with_rz) { fin_update(P, P, o, with_rz); }

// det J and J^-1 by cofactors (exmatrix.cpp:98-125), J[j][i] = dx_i/dxi_j, so that
// Ji[j][i] = dxi_i/dx_j and the physical gradient is grad_j = sum_i Ji[j][i] d/dxi_i.
template <typename T>
__device__ inline T jac_inverse(const T (&J)[3][3], T (&Ji)[3][3]) {
    const T c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const T c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const T c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const T det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const T id = T(1) / det;
    Ji[0][0] = c00 * id; Ji[1][0] = c01 * id; Ji[2][0] = c02 * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    return det;
}

// alpha = rho / p.q once p.q of the iteration is known; a non-finite or non-positive p.q ends the
// sub-problem (status 6 or 3).  Reads in, writes out (the same record or a copy, as fin_update).
__device__ inline void finish_pq(const ProbScalars &in, ProbScalars &out, double pq) {
    const double rho = in.rho;
    out.pq = pq;
    if (!finite(pq)) { out.status = 6; out.done = 1; }
    else if (!(pq > 0)) { out.status = 3; out.done = 1; }
    else out.alpha = rho / pq;
}
__device__ inline void finish_pq(ProbScalars &P, double pq) { finish_pq(P, P, pq); }

// Direction update p = z + beta p of one value, with one rounding (every kernel that forms the
// direction uses this, so the separate update and the one fused into the matrix-free product
// agree bit for bit).
__device__ inline float dir_value(float z, float beta, float p) { return __builtin_fmaf(beta, p, z); }
__device__ inline double dir_value(double z, double beta, double p) { return __builtin_fma(beta, p, z); }

struct WgTables {
    const uint32_t *row0, *row1, *sub, *sub_wg0, *sub_nwg;
};

}  // namespace femgpu
