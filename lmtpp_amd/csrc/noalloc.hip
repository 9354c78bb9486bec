// femgpu_apply_noalloc: y = K x on device pointers without allocating device memory, for callers
// that cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid it).
//
// Every temporary of the product lives in `scratch`, a static device array of SCRATCH_BYTES held
// by each GPU the library runs on until the process ends: the FP32 copies of x and y, the
// matrix-free product's shared-node partials, and the node-order permutations of x and y when the
// problem stores its nodes in an internal order.  Its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and passed to the kernels as a pointer; host
// code never hands the array's name to a kernel or a copy.  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU race.  Nothing needs to start at zero: each call writes
// every scratch value it reads (each shared-node partial slot is written by the one element block
// holding the node before the shared-node pass sums it).
#include <mutex>

#include "capi_handles.hpp"

namespace femgpu {
namespace {

constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void *g_scratch[MAX_DEVICES] = {};

// Device address of `scratch` on device dev (looked up once per device, with dev current).
hipError_t scratch_on(int dev, char **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = static_cast<char *>(g_scratch[dev]);
    return hipSuccess;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace
}  // namespace femgpu

using namespace femgpu;

extern "C" int femgpu_apply_noalloc(femgpu_problem *h, int precision, const double *d_x, double *d_y) {
    if (!h || !d_x || !d_y || precision < 0 || precision > 1) {
        last_error() = "apply_noalloc: bad input";
        return FEMGPU_BAD_INPUT;
    }
    Problem &P = h->P;
    if (!P.assembled_mat) {
        last_error() = "apply_noalloc: matrix not assembled";
        return FEMGPU_NOT_READY;
    }
    const int64_t prod = apply_scratch_bytes(P, precision);
    if (prod < 0) {
        last_error() = "apply_noalloc: the product variant of this precision is not built for the current matrix "
                       "(femgpu_apply or a solve of that precision builds it, which allocates)";
        return FEMGPU_NOT_READY;
    }
    const size_t n = 3 * (size_t)P.n_nodes;
    const size_t vec = P.d_nperm ? align256(n * sizeof(double)) : 0;
    const size_t need = 2 * vec + (size_t)prod;
    if (need > SCRATCH_BYTES) {
        last_error() = "apply_noalloc: the product needs " + std::to_string(need) + " bytes of scratch, more than the " +
                       std::to_string(SCRATCH_BYTES) + " bytes of the static scratch array";
        return FEMGPU_BAD_INPUT;
    }
    char *base = nullptr;
    hipError_t e = scratch_on(P.device, &base);
    if (e != hipSuccess) {
        last_error() = std::string("apply_noalloc: scratch address: ") + hipGetErrorString(e);
        return FEMGPU_HIP_ERROR;
    }
    // caller's node order -> internal order in scratch, the product, and back into d_y
    const double *x = d_x;
    double *y = d_y;
    if (P.d_nperm) {
        double *xi = reinterpret_cast<double *>(base), *yi = reinterpret_cast<double *>(base + vec);
        if ((e = launch_permute_nodes(xi, d_x, P.d_nperm, P.n_nodes, P.stream)) != hipSuccess) {
            last_error() = std::string("apply_noalloc: ") + hipGetErrorString(e);
            return FEMGPU_HIP_ERROR;
        }
        x = xi;
        y = yi;
    }
    if (apply_spmv_scratch(P, precision, x, y, base + 2 * vec)) {
        last_error() = "apply_noalloc: product launch failed";
        return FEMGPU_HIP_ERROR;
    }
    if (P.d_nperm && (e = launch_permute_nodes(d_y, y, P.d_ninv, P.n_nodes, P.stream)) != hipSuccess) {
        last_error() = std::string("apply_noalloc: ") + hipGetErrorString(e);
        return FEMGPU_HIP_ERROR;
    }
    return FEMGPU_OK;
}
