// Jacobi-preconditioned CG on the 3x3-block CSR, FP64 or FP32-inner with FP64 residual
// refreshes.  Block SpMV fused with the p.q reduction; vector updates fused with r.z, r.r and
// max|r|; every reduction is a fixed-order two-level sum (waves, then workgroup partials summed
// by the last workgroup), so results are bitwise reproducible run to run.
//
// The iteration is the PCG of containers/MatWithTinyBlocks.h:1101-1151 /
// containers/conjugate_gradient.h:27-68 with the block IC(0) preconditioner replaced by
// (block-)Jacobi:  q = K d; alpha = rho / d.q; x += alpha d; r -= alpha q; z = M^-1 r;
// rho' = r.z; beta = rho'/rho; d = z + beta d.
#include <algorithm>
#include <cmath>

#include "femgpu_internal.hpp"

namespace femgpu {

enum { MODE_PLAIN = 0, MODE_PQ = 1, MODE_RES = 2 };

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

// Two-level reduction of NV values (bit k of MAXMASK set = max, else sum).  Returns true in the
// last workgroup to arrive; there, out[] is valid in every thread.
template <int NV, unsigned MAXMASK>
__device__ bool grid_reduce(double (&v)[NV], double *partials, unsigned *counter, double (&out)[NV]) {
    __shared__ double sm[NV][4];
    __shared__ bool last;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double t = (MAXMASK >> k & 1) ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) sm[k][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const double t = (MAXMASK >> k & 1) ? fmax(fmax(sm[k][0], sm[k][1]), fmax(sm[k][2], sm[k][3]))
                                                : ((sm[k][0] + sm[k][1]) + sm[k][2]) + sm[k][3];
            partials[(size_t)k * gridDim.x + blockIdx.x] = t;
        }
        __threadfence();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    const volatile double *vp = partials;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double a = 0;
        for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
            const double t = vp[(size_t)k * gridDim.x + i];
            a = (MAXMASK >> k & 1) ? fmax(a, t) : a + t;
        }
        v[k] = a;
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
    if (threadIdx.x == 0) *counter = 0;
    return true;
}

__device__ inline bool finite(double v) { return v == v && fabs(v) <= 1.7976931348623157e308; }

// XCD-aware workgroup index: hardware hands workgroup i to XCD i % 8, so logical slab
// (i % 8) * (nwg / 8) + i / 8 gives each XCD one contiguous range of rows and keeps its
// x-vector neighbourhood in that XCD's L2.
__device__ inline unsigned xcd_slab(unsigned nwg) { return (blockIdx.x & 7u) * (nwg >> 3) + (blockIdx.x >> 3); }

// ---------------------------------------------------------------------------------------------
// Block SpMV, G lanes per block row, lane k of a row handles blocks k, k+G, ...; the nine value
// planes are read as contiguous runs.  MODE_PQ: y = K x and p.q -> alpha.  MODE_RES:
// y = b - K x (FP64) with ||y||^2 and max|y|.
// ---------------------------------------------------------------------------------------------
template <int G, typename T, int MODE>
__global__ __launch_bounds__(256) void k_spmv(const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                                              const T *__restrict__ Kv, uint64_t ps, const T *__restrict__ x,
                                              const double *__restrict__ b, T *__restrict__ y,
                                              const uint32_t *__restrict__ wgrows, Scalars *sc, double *partials,
                                              int check) {
    if (MODE == MODE_PQ && (sc->done | sc->pause)) return;
    if (MODE == MODE_RES && check && sc->done) return;
    constexpr int RPT = 256 / G;
    const unsigned w = xcd_slab(gridDim.x);
    const uint32_t r0 = wgrows[w], r1 = wgrows[w + 1];
    const int lg = threadIdx.x % G, grp = threadIdx.x / G;
    double acc[2] = {0.0, 0.0};
    for (uint32_t base = r0; base < r1; base += RPT) {
        const uint32_t r = base + grp;
        T s0 = 0, s1 = 0, s2 = 0;
        if (r < r1) {
            const uint32_t pe = rp[r + 1];
            for (uint32_t p = rp[r] + lg; p < pe; p += G) {
                const uint32_t c = col[p];
                const T a0 = Kv[p], a1 = Kv[ps + p], a2 = Kv[2 * ps + p];
                const T a3 = Kv[3 * ps + p], a4 = Kv[4 * ps + p], a5 = Kv[5 * ps + p];
                const T a6 = Kv[6 * ps + p], a7 = Kv[7 * ps + p], a8 = Kv[8 * ps + p];
                const T x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
                s0 += a0 * x0 + a1 * x1 + a2 * x2;
                s1 += a3 * x0 + a4 * x1 + a5 * x2;
                s2 += a6 * x0 + a7 * x1 + a8 * x2;
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            s0 += __shfl_xor(s0, o);
            s1 += __shfl_xor(s1, o);
            s2 += __shfl_xor(s2, o);
        }
        if (r < r1 && lg < 3) {
            const T s = lg == 0 ? s0 : lg == 1 ? s1 : s2;
            const size_t i = 3 * (size_t)r + lg;
            if (MODE == MODE_RES) {
                const double v = b[i] - (double)s;
                y[i] = (T)v;
                acc[0] += v * v;
                acc[1] = fmax(acc[1], fabs(v));
            } else {
                y[i] = s;
                if (MODE == MODE_PQ) acc[0] += (double)s * (double)x[i];
            }
        }
    }
    if (MODE == MODE_PLAIN) return;
    double out[2];
    if (MODE == MODE_PQ) {
        double v1[1] = {acc[0]}, o1[1];
        if (!grid_reduce<1, 0u>(v1, partials, &sc->counter[CTR_SPMV], o1)) return;
        out[0] = o1[0];
        if (threadIdx.x == 0) {
            sc->pq = out[0];
            if (!finite(out[0])) { sc->status = 6; sc->done = 1; }
            else if (!(out[0] > 0)) { sc->status = 3; sc->done = 1; }
            else sc->alpha = sc->rho / out[0];
        }
    } else {
        if (!grid_reduce<2, 2u>(acc, partials, &sc->counter[CTR_RES], out)) return;
        if (threadIdx.x == 0) {
            sc->rr = out[0];
            sc->rmax = out[1];
            if (check) {
                sc->rr_ref = out[0];
                if (!finite(out[0])) { sc->status = 6; sc->done = 1; }
                else if (sc->crit_kind == 0 ? out[0] <= sc->stop2 : out[1] <= sc->crit) sc->done = 1;
            }
        }
    }
}

// z = M^-1 r for one node; PRE 0: scalar Jacobi (3 values per node), 1: 3x3 block Jacobi (9).
template <int PRE, typename T>
__device__ inline void apply_m(const T *__restrict__ Minv, size_t n, const T r[3], T z[3]) {
    if (PRE == 0) {
        z[0] = Minv[3 * n] * r[0]; z[1] = Minv[3 * n + 1] * r[1]; z[2] = Minv[3 * n + 2] * r[2];
    } else {
        const T *m = Minv + 9 * n;
        z[0] = m[0] * r[0] + m[1] * r[1] + m[2] * r[2];
        z[1] = m[3] * r[0] + m[4] * r[1] + m[5] * r[2];
        z[2] = m[6] * r[0] + m[7] * r[1] + m[8] * r[2];
    }
}

// x += alpha p; r -= alpha q; z = M^-1 r; reductions r.z, r.r, max|r|; the last workgroup
// advances the scalars and applies the stop test.
template <int PRE, typename T>
__global__ __launch_bounds__(256) void k_update(uint32_t n_nodes, T *__restrict__ x, T *__restrict__ r,
                                                const T *__restrict__ p, const T *__restrict__ q,
                                                const T *__restrict__ Minv, Scalars *sc, double *partials) {
    if (sc->done | sc->pause) return;
    const T alpha = (T)sc->alpha;
    double acc[3] = {0, 0, 0};
    for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < n_nodes; n += (size_t)gridDim.x * blockDim.x) {
        T rv[3], zv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t i = 3 * n + d;
            x[i] += alpha * p[i];
            rv[d] = r[i] - alpha * q[i];
            r[i] = rv[d];
        }
        apply_m<PRE, T>(Minv, n, rv, zv);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            acc[0] += (double)rv[d] * (double)zv[d];
            acc[1] += (double)rv[d] * (double)rv[d];
            acc[2] = fmax(acc[2], fabs((double)rv[d]));
        }
    }
    double out[3];
    if (!grid_reduce<3, 4u>(acc, partials, &sc->counter[CTR_UPD], out)) return;
    if (threadIdx.x == 0) {
        sc->iter += 1;
        sc->rho_prev = sc->rho;
        sc->rho = out[0];
        sc->beta = out[0] / sc->rho_prev;
        sc->rr = out[1];
        sc->rmax = out[2];
        if (!finite(out[0]) || !finite(out[1])) { sc->status = 6; sc->done = 1; return; }
        const bool conv = sc->crit_kind == 0 ? out[1] <= sc->stop2 : out[2] <= sc->crit;
        if (sc->mixed) {
            if (conv || out[1] <= sc->delta2 * sc->rr_ref) sc->pause = 1;
        } else if (conv) {
            sc->done = 1;
        }
        if (!sc->done && sc->iter >= sc->max_iter) { sc->status = 4; sc->done = 1; }
    }
}

// p = M^-1 r + beta p
template <int PRE, typename T>
__global__ __launch_bounds__(256) void k_update_p(uint32_t n_nodes, const T *__restrict__ r, T *__restrict__ p,
                                                  const T *__restrict__ Minv, const Scalars *sc) {
    if (sc->done | sc->pause) return;
    const T beta = (T)sc->beta;
    for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < n_nodes; n += (size_t)gridDim.x * blockDim.x) {
        T rv[3] = {r[3 * n], r[3 * n + 1], r[3 * n + 2]}, zv[3];
        apply_m<PRE, T>(Minv, n, rv, zv);
#pragma unroll
        for (int d = 0; d < 3; ++d) p[3 * n + d] = zv[d] + beta * p[3 * n + d];
    }
}

// Start (or restart after an FP64 residual refresh) of the search direction from an FP64
// residual r64:  r = (T) r64; z = M^-1 r; first: p = z, rho = r.z;  refresh: rho = r.z,
// beta = rho / rho_prev and the pause is lifted (p is then updated by k_update_p).
template <int PRE, typename T>
__global__ __launch_bounds__(256) void k_restart(uint32_t n_nodes, const double *r64, T *r,
                                                 T *__restrict__ p, const T *__restrict__ Minv, Scalars *sc,
                                                 double *partials, int first) {
    if (sc->done) return;
    double acc[1] = {0};
    for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < n_nodes; n += (size_t)gridDim.x * blockDim.x) {
        T rv[3], zv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) { rv[d] = (T)r64[3 * n + d]; r[3 * n + d] = rv[d]; }
        apply_m<PRE, T>(Minv, n, rv, zv);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (first) p[3 * n + d] = zv[d];
            acc[0] += (double)rv[d] * (double)zv[d];
        }
    }
    double out[1];
    if (!grid_reduce<1, 0u>(acc, partials, &sc->counter[CTR_REFRESH], out)) return;
    if (threadIdx.x == 0) {
        if (first) {
            sc->rho = out[0];
        } else {
            sc->rho = out[0];
            sc->beta = out[0] / sc->rho_prev;
        }
        if (!finite(out[0])) { sc->status = 6; sc->done = 1; }
        sc->pause = 0;
    }
}

__global__ __launch_bounds__(256) void k_zero_f32(size_t n, float *__restrict__ x) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = 0.0f;
}

// x64 += x32; x32 = 0
__global__ __launch_bounds__(256) void k_accumulate(size_t n, double *__restrict__ x, float *__restrict__ x32) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        x[i] += (double)x32[i];
        x32[i] = 0.0f;
    }
}

// Jacobi factors from the diagonal blocks; a non-positive pivot marks the matrix NOT_SPD.
template <int PRE>
__global__ __launch_bounds__(256) void k_build_minv(uint32_t n_nodes, const uint32_t *__restrict__ diag,
                                                    const double *__restrict__ Kv, uint64_t ps, double *__restrict__ M,
                                                    float *__restrict__ M32, Scalars *sc) {
    for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < n_nodes; n += (size_t)gridDim.x * blockDim.x) {
        const uint64_t p = diag[n];
        double a[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) a[c] = Kv[c * ps + p];
        if (PRE == 0) {
            for (int d = 0; d < 3; ++d) {
                const double v = a[4 * d];
                if (!(v > 0)) sc->status = 3;
                M[3 * n + d] = 1.0 / v;
                if (M32) M32[3 * n + d] = (float)(1.0 / v);
            }
        } else {
            const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
            const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
            if (!(det > 0) || !(a[0] > 0)) sc->status = 3;
            const double id = 1.0 / det;
            double m[9];
            m[0] = c00 * id; m[3] = c01 * id; m[6] = c02 * id;
            m[1] = (a[2] * a[7] - a[1] * a[8]) * id; m[4] = (a[0] * a[8] - a[2] * a[6]) * id; m[7] = (a[1] * a[6] - a[0] * a[7]) * id;
            m[2] = (a[1] * a[5] - a[2] * a[4]) * id; m[5] = (a[2] * a[3] - a[0] * a[5]) * id; m[8] = (a[0] * a[4] - a[1] * a[3]) * id;
            for (int c = 0; c < 9; ++c) {
                M[9 * n + c] = m[c];
                if (M32) M32[9 * n + c] = (float)m[c];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------

hipError_t ensure_wgrows(Problem &P) {
    if (P.d_wgrows && P.wgrows_nwg == P.spmv.nwg) return hipSuccess;
    if (P.d_wgrows) { (void)hipFree(P.d_wgrows); P.d_wgrows = nullptr; }
    std::vector<uint32_t> cut = balance_rows(P.pat.row_ptr, (uint32_t)P.spmv.nwg);
    hipError_t e = hipMalloc(&P.d_wgrows, cut.size() * 4);
    if (e != hipSuccess) return e;
    e = hipMemcpy(P.d_wgrows, cut.data(), cut.size() * 4, hipMemcpyHostToDevice);
    P.wgrows_nwg = P.spmv.nwg;
    return e;
}

template <typename T, int MODE>
static void launch_spmv_t(Problem &P, const T *K, const T *x, const double *b, T *y, int check) {
    const dim3 g(P.spmv.nwg), blk(256);
#define FEMGPU_SPMV_CASE(GG) \
    hipLaunchKernelGGL((k_spmv<GG, T, MODE>), g, blk, 0, P.stream, P.d_rp, P.d_col, K, P.ps, x, b, y, P.d_wgrows, P.d_sc, P.d_partials, check)
    switch (P.spmv.group) {
        case 8: FEMGPU_SPMV_CASE(8); break;
        case 16: FEMGPU_SPMV_CASE(16); break;
        case 64: FEMGPU_SPMV_CASE(64); break;
        default: FEMGPU_SPMV_CASE(32); break;
    }
#undef FEMGPU_SPMV_CASE
}

static unsigned vec_grid(uint32_t n_nodes) {
    return (unsigned)std::max<size_t>(8, std::min<size_t>(2048, (n_nodes + 255) / 256));
}

hipError_t spmv_apply(Problem &P, const double *d_x, double *d_y) {
    hipError_t e = ensure_wgrows(P);
    if (e != hipSuccess) return e;
    launch_spmv_t<double, MODE_PLAIN>(P, P.d_K, d_x, nullptr, d_y, 0);
    return hipGetLastError();
}

double time_spmv(Problem &P, int precision, int reps) {
    if (ensure_wgrows(P) != hipSuccess) return -1;
    if (precision == 1 && !P.k32_valid) return -1;
    hipEventRecord(P.ev[2], P.stream);
    for (int i = 0; i < reps; ++i) {
        if (precision == 1) launch_spmv_t<float, MODE_PLAIN>(P, P.d_K32, P.d_p32, nullptr, P.d_q32, 0);
        else launch_spmv_t<double, MODE_PLAIN>(P, P.d_K, P.d_p, nullptr, P.d_q, 0);
    }
    hipEventRecord(P.ev[3], P.stream);
    hipEventSynchronize(P.ev[3]);
    float ms = 0;
    hipEventElapsedTime(&ms, P.ev[2], P.ev[3]);
    return (double)ms / reps;
}

static int read_scalars(Problem &P) {
    if (hipMemcpyAsync(P.h_sc, P.d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, P.stream) != hipSuccess) return 5;
    if (hipStreamSynchronize(P.stream) != hipSuccess) return 5;
    return 0;
}

template <int PRE>
static void launch_iteration(Problem &P, bool fp32, unsigned vg) {
    if (fp32) {
        launch_spmv_t<float, MODE_PQ>(P, P.d_K32, P.d_p32, nullptr, P.d_q32, 0);
        hipLaunchKernelGGL((k_update<PRE, float>), dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_x32, P.d_r32, P.d_p32, P.d_q32, P.d_Minv32, P.d_sc, P.d_partials);
        hipLaunchKernelGGL((k_update_p<PRE, float>), dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_r32, P.d_p32, P.d_Minv32, P.d_sc);
    } else {
        launch_spmv_t<double, MODE_PQ>(P, P.d_K, P.d_p, nullptr, P.d_q, 0);
        hipLaunchKernelGGL((k_update<PRE, double>), dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_x, P.d_r, P.d_p, P.d_q, P.d_Minv, P.d_sc, P.d_partials);
        hipLaunchKernelGGL((k_update_p<PRE, double>), dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_r, P.d_p, P.d_Minv, P.d_sc);
    }
}

template <int PRE>
static int solve_impl(Problem &P, const double *d_rhs, double *d_x, const SolveOpts &o, SolveOut &out) {
    const bool mixed = o.precision == 1;
    const unsigned vg = vec_grid(P.n_nodes);
    const size_t n = 3 * (size_t)P.n_nodes;
    P.d_b = const_cast<double *>(d_rhs);
    P.d_x = d_x;
    // b.b on the host side of the stop test: one pass
    Scalars s0 = {};
    s0.crit_kind = o.crit_kind;
    s0.crit = o.tol;
    s0.max_iter = o.max_iter;
    s0.mixed = mixed ? 1 : 0;
    s0.delta2 = o.delta * o.delta;
    if (hipMemcpyAsync(P.d_sc, &s0, sizeof(Scalars), hipMemcpyHostToDevice, P.stream) != hipSuccess) return 5;
    // ||b||^2 with the residual kernel on x = 0 (r = b)
    launch_fill(P.d_q, n, 0.0, P.stream);
    launch_spmv_t<double, MODE_RES>(P, P.d_K, P.d_q, d_rhs, P.d_r, 0);
    if (read_scalars(P)) return 5;
    const double bb = P.h_sc->rr;
    const double stop2 = o.tol * o.tol * (bb > 0 ? bb : 1.0);
    P.h_sc->bb = bb;
    P.h_sc->stop2 = stop2;
    P.h_sc->rr = 0;
    P.h_sc->rmax = 0;
    if (hipMemcpyAsync(P.d_sc, P.h_sc, sizeof(Scalars), hipMemcpyHostToDevice, P.stream) != hipSuccess) return 5;
    // preconditioner
    if (P.minv_kind != PRE || (mixed && !P.d_Minv32)) {
        if (mixed && !P.d_Minv32 && hipMalloc(&P.d_Minv32, sizeof(float) * 9 * P.n_nodes) != hipSuccess) return 5;
        hipLaunchKernelGGL(k_build_minv<PRE>, dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_diag, P.d_K, P.ps, P.d_Minv,
                           P.d_Minv32, P.d_sc);
        P.minv_kind = PRE;
    }
    if (read_scalars(P)) return 5;
    if (P.h_sc->status) { out.status = P.h_sc->status; return out.status; }

    hipEventRecord(P.ev[0], P.stream);
    // r = b - K x; stop if already converged
    launch_spmv_t<double, MODE_RES>(P, P.d_K, d_x, d_rhs, P.d_r, 1);
    if (mixed) {
        hipLaunchKernelGGL(k_zero_f32, dim3(vg), dim3(256), 0, P.stream, n, P.d_x32);
        hipLaunchKernelGGL((k_restart<PRE, float>), dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_r, P.d_r32, P.d_p32, P.d_Minv32, P.d_sc, P.d_partials, 1);
    } else {
        hipLaunchKernelGGL((k_restart<PRE, double>), dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_r, P.d_r, P.d_p, P.d_Minv, P.d_sc, P.d_partials, 1);
    }
    int chunk = 2, refreshes = 0;
    for (;;) {
        for (int k = 0; k < chunk; ++k) launch_iteration<PRE>(P, mixed, vg);
        if (hipGetLastError() != hipSuccess) return 5;
        if (read_scalars(P)) return 5;
        chunk = std::min(chunk * 2, std::max(2, o.chunk));
        if (P.h_sc->done) break;
        if (mixed && P.h_sc->pause) {
            // FP64 refresh: x += x32; r = b - K x (FP64); restart the FP32 recurrence from it
            hipLaunchKernelGGL(k_accumulate, dim3(vg), dim3(256), 0, P.stream, n, d_x, P.d_x32);
            launch_spmv_t<double, MODE_RES>(P, P.d_K, d_x, d_rhs, P.d_r, 1);
            hipLaunchKernelGGL((k_restart<PRE, float>), dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_r, P.d_r32, P.d_p32, P.d_Minv32, P.d_sc, P.d_partials, 0);
            hipLaunchKernelGGL((k_update_p<PRE, float>), dim3(vg), dim3(256), 0, P.stream, P.n_nodes, P.d_r32, P.d_p32, P.d_Minv32, P.d_sc);
            ++refreshes;
        }
    }
    if (mixed) hipLaunchKernelGGL(k_accumulate, dim3(vg), dim3(256), 0, P.stream, n, d_x, P.d_x32);
    hipEventRecord(P.ev[1], P.stream);
    // final true residual for the report
    launch_spmv_t<double, MODE_RES>(P, P.d_K, d_x, d_rhs, P.d_q, 0);
    if (read_scalars(P)) return 5;
    float ms = 0;
    hipEventElapsedTime(&ms, P.ev[0], P.ev[1]);
    out.ms = ms;
    out.iters = P.h_sc->iter;
    out.refreshes = refreshes;
    out.rel_res = std::sqrt(P.h_sc->rr / (bb > 0 ? bb : 1.0));
    out.max_abs_res = P.h_sc->rmax;
    out.status = P.h_sc->status;
    return out.status;
}

int solve_pcg(Problem &P, const double *d_rhs, double *d_x, const SolveOpts &o, SolveOut &out) {
    out = SolveOut{};
    if (ensure_wgrows(P) != hipSuccess) return 5;
    if (o.precision == 1 && !P.k32_valid) {
        if (!P.d_K32 && hipMalloc(&P.d_K32, sizeof(float) * 9 * P.ps) != hipSuccess) return 5;
        if (launch_to_fp32(P) != hipSuccess) return 5;
        P.k32_valid = true;
    }
    return o.precond == 1 ? solve_impl<1>(P, d_rhs, d_x, o, out) : solve_impl<0>(P, d_rhs, d_x, o, out);
}

}  // namespace femgpu
