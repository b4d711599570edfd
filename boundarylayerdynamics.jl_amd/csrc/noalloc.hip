// Allocation-free batched LDLt solve (bld_ldlt_solve_noalloc): the same
// elimination as bld_ldlt_create + bld_ldlt_solve on device-resident inputs,
// with the factor and the float32 staging in a fixed device array instead of
// memory the library allocates.  For callers that cannot allocate: code under
// HIP graph capture, serving loops.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "bld.h"
#include "kernels.hpp"
#include "ldlt.hpp"

namespace bld {

// The fixed work array of every GPU the library runs on (held until the
// process ends).  Host code never uses its name as an address: each device's
// copy is reached through the pointer hipGetSymbolAddress returns on it.
static __device__ __align__(256) unsigned char scratch[BLD_LDLT_NOALLOC_SCRATCH];

namespace {

std::mutex g_scratch_mutex;
std::vector<void *> g_scratch_ptr;  // per device id, nullptr until set up
thread_local std::string g_noalloc_error;

// The device address of `scratch` on the current device (set up on the first
// call made with that device current).
hipError_t scratch_of_current_device(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (int(g_scratch_ptr.size()) <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

// γ_1 = 1/d_1, γ_i = 1/(d_i − γ_{i−1}·β_{i−1}²) (pressure.jl:86-95), one thread
// per system, float64 with the operations in the order of the handle's host
// factorisation (no contraction), stored as A.
template <typename A>
__global__ void __launch_bounds__(256) k_ldlt_factor(const double *__restrict__ dv, const double *__restrict__ ev,
                                                     A *__restrict__ gamma, int nk, int nz) {
#pragma clang fp contract(off)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nk) return;
    double g = 1 / dv[m];
    gamma[m] = A(g);
    for (int i = 1; i < nz; ++i) {
        const double b = ev[size_t(i - 1) * nk + m];
        g = 1 / (dv[size_t(i) * nk + m] - g * (b * b));
        gamma[size_t(i) * nk + m] = A(g);
    }
}

// complex float64 -> complex float32 (the handle's upload rounding) and back
__global__ void __launch_bounds__(256) k_to_f32(const cplx<double> *__restrict__ a, cplx<float> *__restrict__ o,
                                               size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const cplx<double> v = a[i];
        o[i] = mk(float(v.x), float(v.y));
    }
}
__global__ void __launch_bounds__(256) k_to_f64(const cplx<float> *__restrict__ a, cplx<double> *__restrict__ o,
                                               size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const cplx<float> v = a[i];
        o[i] = mk(double(v.x), double(v.y));
    }
}

}  // namespace bld

using namespace bld;

extern "C" {

const char *bld_ldlt_noalloc_last_error(void) { return g_noalloc_error.c_str(); }

int bld_ldlt_solve_noalloc(int64_t nk, int64_t nz, const double *dvs, const double *evs, double *b, int precision,
                           void *stream) {
    g_noalloc_error.clear();
    if (nk < 1 || nz < 1 || !dvs || (nz > 1 && !evs) || !b) {
        g_noalloc_error = "invalid LDLt arguments";
        return BLD_EARG;
    }
    if (precision != BLD_FP64 && precision != BLD_FP32_MIXED) {
        g_noalloc_error = "unknown precision";
        return BLD_EARG;
    }
    if (nz > kLdltChunkedMaxLevels || nk > (int64_t(1) << 30)) {
        g_noalloc_error = "allocation-free LDLt: at most 4096 levels per system";
        return BLD_EARG;
    }
    const size_t n = size_t(nk) * size_t(nz);
    const size_t need = precision == BLD_FP64 ? n * sizeof(double)
                                              : n * (sizeof(float) + 2 * sizeof(cplx<float>));
    if (need > size_t(BLD_LDLT_NOALLOC_SCRATCH)) {
        g_noalloc_error = "allocation-free LDLt: " + std::to_string(n) + " unknowns need " + std::to_string(need) +
                          " bytes of work space, more than the fixed " +
                          std::to_string(size_t(BLD_LDLT_NOALLOC_SCRATCH)) + " bytes";
        return BLD_EARG;
    }
    void *ptr = nullptr;
    hipError_t e = scratch_of_current_device(&ptr);
    if (e != hipSuccess) {
        g_noalloc_error = std::string("HIP error: ") + hipGetErrorString(e) + " in hipGetSymbolAddress(scratch)";
        return BLD_EHIP;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const double *beta = nz > 1 ? evs : dvs;  // nz = 1 reads no off-diagonal
    const int fb = int((nk + 255) / 256);
    const int cb = int(std::min<size_t>((n + 255) / 256, 4096));
    if (precision == BLD_FP64) {
        double *gamma = static_cast<double *>(ptr);
        k_ldlt_factor<double><<<fb, 256, 0, st>>>(dvs, evs, gamma, int(nk), int(nz));
        ldlt_chunked_launch(false, gamma, beta, true, nullptr, b, b, int(nk), int(nz), st);
    } else {
        // [γ float n][b32 complex float n][x32 complex float n], each 256-B aligned
        auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
        unsigned char *base = static_cast<unsigned char *>(ptr);
        float *gamma = reinterpret_cast<float *>(base);
        cplx<float> *b32 = reinterpret_cast<cplx<float> *>(base + align(n * sizeof(float)));
        cplx<float> *x32 = reinterpret_cast<cplx<float> *>(base + align(n * sizeof(float)) + align(n * sizeof(cplx<float>)));
        if (align(n * sizeof(float)) + 2 * align(n * sizeof(cplx<float>)) > size_t(BLD_LDLT_NOALLOC_SCRATCH)) {
            g_noalloc_error = "allocation-free LDLt: work space exceeds the fixed array";
            return BLD_EARG;
        }
        k_ldlt_factor<float><<<fb, 256, 0, st>>>(dvs, evs, gamma, int(nk), int(nz));
        k_to_f32<<<cb, 256, 0, st>>>(reinterpret_cast<const cplx<double> *>(b), b32, n);
        ldlt_chunked_launch(true, gamma, beta, true, dvs, b32, x32, int(nk), int(nz), st);
        k_to_f64<<<cb, 256, 0, st>>>(x32, reinterpret_cast<cplx<double> *>(b), n);
    }
    e = hipGetLastError();
    if (e != hipSuccess) {
        g_noalloc_error = std::string("HIP error: ") + hipGetErrorString(e) + " launching the LDLt kernels";
        return BLD_EHIP;
    }
    return BLD_OK;
}

}  // extern "C"
