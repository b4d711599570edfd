// Shared device types, FFT plans and descriptors of the time-step engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bld {

template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T a, T b) {
    cplx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return mk(s * a.x, s * a.y); }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {  // a * conj(b)
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// (i*d) * a for real d: spectral derivative factor i·k·2π/L times a
template <typename T>
__device__ __forceinline__ cplx<T> imul(T d, cplx<T> a) { return mk(-d * a.y, d * a.x); }

// Loads and stores through the global address space.  Pointers read from the
// device descriptors are generic to the compiler, whose flat instructions also
// wait on the LDS counter; global instructions let the loads of a thread stay
// in flight across LDS traffic.
template <typename T>
__device__ __forceinline__ T gload(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ cplx<T> gload(const cplx<T> *p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = *(const __attribute__((address_space(1))) v2 *)p;
    return mk(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void gstore(T *p, T v) {
    *(__attribute__((address_space(1))) T *)p = v;
}
template <typename T>
__device__ __forceinline__ void gstore(cplx<T> *p, cplx<T> v) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 w;
    w.x = v.x;
    w.y = v.y;
    *(__attribute__((address_space(1))) v2 *)p = w;
}

constexpr int kMaxStages = 12;

// Mixed-radix Stockham plan for one transform length.
struct FftPlan {
    int n;
    int nstages;
    int radix[kMaxStages];
    int tw;  // offset of this length's twiddle table exp(-2πi t/n), t = 0..n-1
};

// Recipes for the spectral input of a c2r transform (what is transformed).
enum Recipe : int {
    R_U = 0, R_V, R_W,           // state components
    R_D1U, R_D2U, R_D1V, R_D2V,  // horizontal derivatives (C-nodes)
    R_D1W, R_D2W,                // horizontal derivatives of w (I-nodes)
    R_VORT1, R_VORT2, R_VORT3,   // spectral vorticity (derivatives.jl:566-568)
    R_FIELD                      // a plain field given by pointer (set/get helpers)
};

constexpr int kMaxTerms = 10;
constexpr int kMaxTargets = 3;

// Physical-space process attached to a space.
enum SpaceKind : int { SK_ADVECTION = 0, SK_SGS = 1, SK_TRANSFER = 2 };

// One physical space (one size N1 x N2) of one model instance: the c2r terms,
// the physical rates that are transformed back, and the accumulation recipe.
template <typename T>
struct Space {
    int kind;
    int N1, N2, K1, K2, k1m, k2m, nk1, n3;
    FftPlan p1, p2;
    const cplx<T> *tw;  // twiddle table base
    // c2r terms
    int nin;
    int in_recipe[kMaxTerms];
    int in_layers[kMaxTerms];
    int in_job0[kMaxTerms + 1];
    T *in_phys[kMaxTerms];
    const cplx<T> *in_field[kMaxTerms];  // for R_FIELD
    // r2c sources
    int nsrc;
    int src_layers[kMaxTerms];
    int src_job0[kMaxTerms + 1];
    T *src_phys[kMaxTerms];
    // r2c targets: out[f] += FFT(Σ_G pre·src) + D2·FFT(Σ_H src)
    int ntgt;
    int tgt_layers[kMaxTargets];
    int tgt_job0[kMaxTargets + 1];
    int tgt_nG[kMaxTargets], tgt_G[kMaxTargets][3], tgt_Gd1[kMaxTargets][3];
    int tgt_nH[kMaxTargets], tgt_H[kMaxTargets][3];
    cplx<T> *tgt_out[kMaxTargets];
    // shared intermediate buffer [job][N2][nk1]
    cplx<T> *inter;
    // spectral input state (set per launch through `sel`: 0 = u, 1 = utmp)
    const cplx<T> *state[2][3];
    // derivative factors (double): D1[k1] = k1·2π/L1, D2[i2] = k2·2π/L2 (imaginary parts)
    const double *D1, *D2;
    const double *alpha_c, *alpha_i;  // [n3], [n3+1]
    // physical-process data
    T *visc_c;                        // SGS eddy viscosity at the C-nodes
    int store_visc;                   // the SGS kernel stores visc_c (read by the τ33 mean profile only)
    const double *lc, *li;            // SGS length scales
    int wall_rough[2];
    double fstrain[2], fstress[2];
    double wbc[2];                    // vel3 Dirichlet values (physical)
    int xfer_centered;                // transfer space: values at the cell centres (physical_space.jl:322-336)
    // learned closure (SGS spaces whose stresses come from the network)
    const unsigned short *mlp_w;      // bfloat16 W1, W2, W3
    const float *mlp_b;               // b1, b2, b3
    const double *mlp_xw;             // wall feature per C-layer [n3]
    double mlp_delta2;                // Δ² of the stress scale Δ²|∇u|²
    int mlp_guard;                    // Smagorinsky where the learned stress is not dissipative
    T *tau[6];                        // C-node stresses τ11, τ22, τ33, τ12, τ13, τ23
};

// Per-launch block bookkeeping of a multi-space transform launch.
struct LaunchEntry {
    int space;        // index into the space array
    int block_begin;  // first block of this entry
    int blocks_per_job;
    int njobs;
    int batch;        // 1D transforms per block
};

// Block ranges of one multi-entry launch; the entries live in device memory so
// one launch can cover every space (or instance) of an ensemble.
struct LaunchTable {
    int n;
    const LaunchEntry *e;
};

// entry of block b (entries sorted by block_begin)
__device__ __forceinline__ int find_entry(const LaunchTable &lt, int b) {
    int lo = 0, hi = lt.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b >= lt.e[mid].block_begin)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

}  // namespace bld
