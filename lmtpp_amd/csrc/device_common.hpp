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

// One workgroup combines the partials of workgroups [w0, w0 + nw) in a fixed order (strided
// per thread, wave tree, 4-wave combination).  out[] is valid in every thread.
template <int NV, unsigned MAXMASK>
__device__ void wg_combine(const double *partials, unsigned stride, unsigned w0, unsigned nw, double (&out)[NV]) {
    __shared__ double sm[NV][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        // four independent chains per thread (i, i+B, i+2B, i+3B), joined in a fixed order
        const double *src = partials + (size_t)k * stride + w0;
        const unsigned B = blockDim.x;
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        unsigned i = threadIdx.x;
        for (; i + 3 * B < nw; i += 4 * B) {
            const double t0 = src[i], t1 = src[i + B], t2 = src[i + 2 * B], t3 = src[i + 3 * B];
            if (MAXMASK >> k & 1) { a0 = fmax(a0, t0); a1 = fmax(a1, t1); a2 = fmax(a2, t2); a3 = fmax(a3, t3); }
            else { a0 += t0; a1 += t1; a2 += t2; a3 += t3; }
        }
        for (; i < nw; i += B) a0 = (MAXMASK >> k & 1) ? fmax(a0, src[i]) : a0 + src[i];
        v[k] = (MAXMASK >> k & 1) ? fmax(fmax(a0, a1), fmax(a2, a3)) : (a0 + a1) + (a2 + a3);
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double t = (MAXMASK >> k & 1) ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) sm[k][wave] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k)
        out[k] = (MAXMASK >> k & 1) ? fmax(fmax(sm[k][0], sm[k][1]), fmax(sm[k][2], sm[k][3]))
                                    : ((sm[k][0] + sm[k][1]) + sm[k][2]) + sm[k][3];
}

__device__ inline bool finite(double v) { return v == v && fabs(v) <= 1.7976931348623157e308; }

struct WgTables {
    const uint32_t *row0, *row1, *sub, *sub_wg0, *sub_nwg;
};

}  // namespace femgpu
