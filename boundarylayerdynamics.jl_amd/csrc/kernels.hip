// Device kernels of the SSPRK33 time-step engine (gfx950).
//
// Transforms.  A 2D c2r of one layer (src/physical_space.jl:277-295) is done in
// two passes: along x2 (complex, zero-padded from the K2 resolved modes) into an
// intermediate [x2][k1] array, then along x1, where two real rows are produced
// by one complex inverse FFT (z = a + i·b).  The forward r2c with truncation,
// scaling by 1/(N1·N2), prefactor and accumulation (physical_space.jl:208-246)
// runs the same passes in the opposite order.  Each block transforms `batch`
// columns/row-pairs of one (term, layer) job in LDS.
//
// Physics.  Pointwise kernels for the rotation-form advection products
// (momentum_advection.jl:179-206), the static Smagorinsky model with its wall
// model (static_smagorinsky.jl:129-249), the spectral linear terms
// (molecular_diffusion.jl:57-72, sources.jl:32-35), and one fused kernel per
// SSPRK33 stage that combines the stage, builds the pressure right-hand side,
// runs the batched symmetric Thomas sweep per (k1,k2) column and subtracts the
// gradient (ODEMethods.jl:147-166, pressure.jl:334-426).
#include "kernels.hpp"

#ifndef BLD_KERNEL_PART
#define BLD_KERNEL_PART 0
#endif

#include "fft.hpp"
#include "terms.hpp"
#include "transforms_2d.hpp"
#include "transforms_reg.hpp"

namespace bld {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

template <typename T, bool INV, int FN>
__device__ __forceinline__ cplx<T> *run_fft(cplx<T> *a, cplx<T> *b, const FftPlan &p, int batch,
                                           const cplx<T> *__restrict__ tw) {
    if constexpr (FN > 0)
        return fft_block_ct<T, INV, FN>(a, b, batch, tw);
    else
        return fft_block<T, INV>(a, b, p, batch, tw);
}

// Lengths with compile-time FFT bodies; anything else takes the generic path.
#define BLD_FFT_SWITCH(N, CALL)      \
    switch (N) {                     \
        case 16: CALL(16); break;    \
        case 24: CALL(24); break;    \
        case 32: CALL(32); break;    \
        case 48: CALL(48); break;    \
        case 64: CALL(64); break;    \
        case 72: CALL(72); break;    \
        case 96: CALL(96); break;    \
        case 128: CALL(128); break;  \
        case 144: CALL(144); break;  \
        case 192: CALL(192); break;  \
        case 256: CALL(256); break;  \
        case 384: CALL(384); break;  \
        default: CALL(0); break;     \
    }

// ---------------------------------------------------------------------------
// c2r, pass 1: inverse FFT along x2 of `batch` k1-columns
// ---------------------------------------------------------------------------
template <typename T, int FN>
__device__ __forceinline__ void c2r_x2_body(const Space<T> &S, const LaunchEntry &E, int sel, cplx<T> *smem) {
    const int local = blockIdx.x - E.block_begin;
    const int job = local / E.blocks_per_job;
    const int chunk = local - job * E.blocks_per_job;
    int term = 0;
    while (job >= S.in_job0[term + 1]) ++term;
    const int layer = job - S.in_job0[term];
    const int recipe = S.in_recipe[term];
    const int B = E.batch;
    const int N2 = FN > 0 ? FN : S.N2;
    cplx<T> *bufA = smem, *bufB = smem + B * N2;
    const int k1_0 = chunk * B;
    const int ncols = min(B, S.nk1 - k1_0);
    const int lb = __ffs(B) - 1;  // B is a power of two
    for (int idx = threadIdx.x; idx < B * N2; idx += blockDim.x) {
        const int pos = idx >> lb, c = idx & (B - 1);  // c fastest: coalesced over k1
        cplx<T> val = czero<T>();
        if (c < ncols) {
            int i2 = -1;
            if (pos <= S.k2m)
                i2 = pos;
            else if (pos >= N2 - S.k2m)
                i2 = S.K2 - (N2 - pos);
            if (i2 >= 0) val = term_value(S, term, recipe, sel, layer, k1_0 + c, i2);
        }
        bufA[c * N2 + pos] = val;
    }
    __syncthreads();
    cplx<T> *res = run_fft<T, true, FN>(bufA, bufB, S.p2, B, S.tw + S.p2.tw);
    cplx<T> *out = S.inter + size_t(job) * N2 * S.nk1;
    for (int idx = threadIdx.x; idx < B * N2; idx += blockDim.x) {
        const int x2 = idx >> lb, c = idx & (B - 1);
        if (c >= ncols) continue;
        cplx<T> v = res[c * N2 + x2];
        if (k1_0 + c == 0) v.y = T(0);  // c2r uses the real part of the k1 = 0 column
        out[size_t(x2) * S.nk1 + k1_0 + c] = v;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_c2r_x2(const Space<T> *__restrict__ spaces, LaunchTable lt, int sel) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *smem = reinterpret_cast<cplx<T> *>(smem_raw);
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
#define BLD_CALL(n)                                   \
    if constexpr (reg_fft_a(n) > 0)                   \
        c2r_x2_reg<T, n>(S, E, sel, smem);                 \
    else                                              \
        c2r_x2_body<T, n>(S, E, sel, smem)
    BLD_FFT_SWITCH(S.N2, BLD_CALL)
#undef BLD_CALL
}

// fused c2r of whole (term, layer) planes of square grids, one plane per
// block of NT threads
template <typename T, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) k_c2r_2d(const Space<T> *__restrict__ spaces, LaunchTable lt, int sel) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *smem = reinterpret_cast<cplx<T> *>(smem_raw);
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
#define BLD_CALL(n)                                 \
    if constexpr (c2r2d_supported<T>(n))           \
        c2r_2d_reg<T, n, NT>(S, E, sel, smem)
    BLD_FFT_SWITCH(S.N2, BLD_CALL)
#undef BLD_CALL
}

// half-plane c2r: two blocks per plane, one for its even and one for its odd
// rows (placed 8 blocks apart, on one XCD)
template <typename T>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
k_c2r_2h(const Space<T> *__restrict__ spaces, LaunchTable lt, int sel) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *smem = reinterpret_cast<cplx<T> *>(smem_raw);
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
#define BLD_CALL(n)                           \
    if constexpr (c2r2h_supported<T>(n))     \
        c2r_2h_reg<T, n, 512>(S, E, sel, smem)
    BLD_FFT_SWITCH(S.N2, BLD_CALL)
#undef BLD_CALL
}

// fused r2c of whole (target, layer) planes of square grids, one plane per block
template <typename T>
__global__ void __launch_bounds__(kC2r2dThreads) k_r2c_2d(const Space<T> *__restrict__ spaces, LaunchTable lt) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *smem = reinterpret_cast<cplx<T> *>(smem_raw);
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
#define BLD_CALL(n)                       \
    if constexpr (c2r2d_supported<T>(n)) \
        r2c_2d_any<T, n>(S, E, smem)
    BLD_FFT_SWITCH(S.N2, BLD_CALL)
#undef BLD_CALL
}

// ---------------------------------------------------------------------------
// c2r, pass 2: Hermitian completion along x1, two real rows per complex FFT
// ---------------------------------------------------------------------------
template <typename T, int FN>
__device__ __forceinline__ void c2r_x1_body(const Space<T> &S, const LaunchEntry &E, cplx<T> *smem) {
    const int local = blockIdx.x - E.block_begin;
    const int job = local / E.blocks_per_job;
    const int chunk = local - job * E.blocks_per_job;
    int term = 0;
    while (job >= S.in_job0[term + 1]) ++term;
    const int layer = job - S.in_job0[term];
    const int B = E.batch, N2 = S.N2, k1m = S.k1m, nk1 = S.nk1;
    const int N1 = FN > 0 ? FN : S.N1;
    const Div divN(N1);
    const int npairs_tot = (N2 + 1) / 2;
    const int pair0 = chunk * B;
    const int np = min(B, npairs_tot - pair0);
    cplx<T> *bufA = smem, *bufB = smem + B * N1;
    const cplx<T> *in = S.inter + size_t(job) * N2 * nk1;
    for (int idx = threadIdx.x; idx < B * N1; idx += blockDim.x) {
        const int p = divN.q(idx), k = idx - p * N1;
        cplx<T> z = czero<T>();
        if (p < np) {
            const int xa = 2 * (pair0 + p), xb = xa + 1;
            const bool hb = xb < N2;
            if (k == 0) {
                cplx<T> a = in[size_t(xa) * nk1];
                z = mk(a.x, hb ? in[size_t(xb) * nk1].x : T(0));
            } else if (2 * k == N1 && k <= k1m) {  // Nyquist: real parts only
                cplx<T> a = in[size_t(xa) * nk1 + k];
                z = mk(a.x, hb ? in[size_t(xb) * nk1 + k].x : T(0));
            } else if (k <= k1m) {
                cplx<T> a = in[size_t(xa) * nk1 + k];
                cplx<T> b = hb ? in[size_t(xb) * nk1 + k] : czero<T>();
                z = mk(a.x - b.y, a.y + b.x);
            } else if (N1 - k <= k1m) {
                const int kk = N1 - k;
                cplx<T> a = in[size_t(xa) * nk1 + kk];
                cplx<T> b = hb ? in[size_t(xb) * nk1 + kk] : czero<T>();
                z = mk(a.x + b.y, b.x - a.y);
            }
        }
        bufA[idx] = z;
    }
    __syncthreads();
    cplx<T> *res = run_fft<T, true, FN>(bufA, bufB, S.p1, B, S.tw + S.p1.tw);
    T *out = S.in_phys[term] + size_t(layer) * N1 * N2;
    for (int idx = threadIdx.x; idx < np * N1; idx += blockDim.x) {
        const int p = divN.q(idx), x = idx - p * N1;
        const int xa = 2 * (pair0 + p);
        cplx<T> z = res[p * N1 + x];
        out[size_t(xa) * N1 + x] = z.x;
        if (xa + 1 < N2) out[size_t(xa + 1) * N1 + x] = z.y;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_c2r_x1(const Space<T> *__restrict__ spaces, LaunchTable lt) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *smem = reinterpret_cast<cplx<T> *>(smem_raw);
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
#define BLD_CALL(n)                                   \
    if constexpr (reg_fft_a(n) > 0)                   \
        c2r_x1_reg<T, n>(S, E, smem);                 \
    else                                              \
        c2r_x1_body<T, n>(S, E, smem)
    BLD_FFT_SWITCH(S.N1, BLD_CALL)
#undef BLD_CALL
}

// ---------------------------------------------------------------------------
// r2c, pass 1: forward FFT along x1 of two real rows at once, truncated to k1m
// ---------------------------------------------------------------------------
template <typename T, int FN>
__device__ __forceinline__ void r2c_x1_body(const Space<T> &S, const LaunchEntry &E, cplx<T> *smem) {
    const int local = blockIdx.x - E.block_begin;
    const int job = local / E.blocks_per_job;
    const int chunk = local - job * E.blocks_per_job;
    int src = 0;
    while (job >= S.src_job0[src + 1]) ++src;
    const int layer = job - S.src_job0[src];
    const int B = E.batch, N2 = S.N2, nk1 = S.nk1;
    const int N1 = FN > 0 ? FN : S.N1;
    const Div divN(N1), divK(nk1);
    const int npairs_tot = (N2 + 1) / 2;
    const int pair0 = chunk * B;
    const int np = min(B, npairs_tot - pair0);
    cplx<T> *bufA = smem, *bufB = smem + B * N1;
    const T *in = S.src_phys[src] + size_t(layer) * N1 * N2;
    for (int idx = threadIdx.x; idx < B * N1; idx += blockDim.x) {
        const int p = divN.q(idx), x = idx - p * N1;
        cplx<T> z = czero<T>();
        if (p < np) {
            const int xa = 2 * (pair0 + p);
            z.x = in[size_t(xa) * N1 + x];
            if (xa + 1 < N2) z.y = in[size_t(xa + 1) * N1 + x];
        }
        bufA[idx] = z;
    }
    __syncthreads();
    cplx<T> *res = run_fft<T, false, FN>(bufA, bufB, S.p1, B, S.tw + S.p1.tw);
    cplx<T> *out = S.inter + size_t(job) * N2 * nk1;
    for (int idx = threadIdx.x; idx < np * nk1; idx += blockDim.x) {
        const int p = divK.q(idx), k = idx - p * nk1;
        const int xa = 2 * (pair0 + p);
        const cplx<T> zk = res[p * N1 + k];
        const cplx<T> zn = res[p * N1 + (k == 0 ? 0 : N1 - k)];
        out[size_t(xa) * nk1 + k] = mk(T(0.5) * (zk.x + zn.x), T(0.5) * (zk.y - zn.y));
        if (xa + 1 < N2) out[size_t(xa + 1) * nk1 + k] = mk(T(0.5) * (zk.y + zn.y), T(-0.5) * (zk.x - zn.x));
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_r2c_x1(const Space<T> *__restrict__ spaces, LaunchTable lt) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *smem = reinterpret_cast<cplx<T> *>(smem_raw);
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
#define BLD_CALL(n)                                   \
    if constexpr (reg_fft_a(n) > 0)                   \
        r2c_x1_reg<T, n>(S, E, smem);                 \
    else                                              \
        r2c_x1_body<T, n>(S, E, smem)
    BLD_FFT_SWITCH(S.N1, BLD_CALL)
#undef BLD_CALL
}

// ---------------------------------------------------------------------------
// r2c, pass 2: forward FFT along x2, truncation, 1/(N1·N2), prefactors,
// accumulation into the spectral rate (one block owns each output element)
// ---------------------------------------------------------------------------
template <typename T, int FN>
__device__ __forceinline__ void r2c_x2_body(const Space<T> &S, const LaunchEntry &E, cplx<T> *smem) {
    const int local = blockIdx.x - E.block_begin;
    const int job = local / E.blocks_per_job;
    const int chunk = local - job * E.blocks_per_job;
    int tg = 0;
    while (job >= S.tgt_job0[tg + 1]) ++tg;
    const int layer = job - S.tgt_job0[tg];
    const int B = E.batch, nk1 = S.nk1;
    const int N2 = FN > 0 ? FN : S.N2;
    const int nG = S.tgt_nG[tg], nH = S.tgt_nH[tg];
    const int nT = nH > 0 ? 2 * B : B;  // G columns, then H columns
    cplx<T> *bufA = smem, *bufB = smem + nT * N2;
    const int k1_0 = chunk * B;
    const int ncols = min(B, nk1 - k1_0);
    const int lb = __ffs(B) - 1;  // B is a power of two
    for (int idx = threadIdx.x; idx < B * N2; idx += blockDim.x) {
        const int x2 = idx >> lb, c = idx & (B - 1);  // c fastest: coalesced over k1
        cplx<T> g = czero<T>(), h = czero<T>();
        if (c < ncols) {
            const int k1 = k1_0 + c;
            const T d1 = T(S.D1[k1]);
            for (int s = 0; s < nG; ++s) {
                const int sj = S.src_job0[S.tgt_G[tg][s]] + layer;
                cplx<T> v = S.inter[(size_t(sj) * N2 + x2) * nk1 + k1];
                g = g + (S.tgt_Gd1[tg][s] ? imul(d1, v) : v);
            }
            for (int s = 0; s < nH; ++s) {
                const int sj = S.src_job0[S.tgt_H[tg][s]] + layer;
                h = h + S.inter[(size_t(sj) * N2 + x2) * nk1 + k1];
            }
        }
        bufA[c * N2 + x2] = g;
        if (nH > 0) bufA[(B + c) * N2 + x2] = h;
    }
    __syncthreads();
    cplx<T> *res = run_fft<T, false, FN>(bufA, bufB, S.p2, nT, S.tw + S.p2.tw);
    const T scale = T(1.0 / (double(S.N1) * double(N2)));
    const int nres = 2 * S.k2m + 1;
    cplx<T> *out = S.tgt_out[tg] + size_t(layer) * S.K1 * S.K2;
    for (int idx = threadIdx.x; idx < B * nres; idx += blockDim.x) {
        const int r = idx >> lb, c = idx & (B - 1);
        if (c >= ncols) continue;
        const int i2 = r <= S.k2m ? r : S.K2 - (nres - r);
        const int pos = r <= S.k2m ? r : N2 - (nres - r);
        cplx<T> v = res[c * N2 + pos];
        if (nH > 0) v = v + imul(T(S.D2[i2]), res[(B + c) * N2 + pos]);
        const size_t o = size_t(i2) * S.K1 + k1_0 + c;
        if (S.xfer_centered) v = cmul(v, centre_shift<T>(S, k1_0 + c, i2, -1.0));
        out[o] = out[o] + scale * v;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_r2c_x2(const Space<T> *__restrict__ spaces, LaunchTable lt) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *smem = reinterpret_cast<cplx<T> *>(smem_raw);
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
#define BLD_CALL(n)                                   \
    if constexpr (reg_fft_a(n) > 0)                   \
        r2c_x2_reg<T, n>(S, E, smem);                 \
    else                                              \
        r2c_x2_body<T, n>(S, E, smem)
    BLD_FFT_SWITCH(S.N2, BLD_CALL)
#undef BLD_CALL
}

// ---------------------------------------------------------------------------
// advection products (momentum_advection.jl:179-206); w = 0 at the walls
// ---------------------------------------------------------------------------
// A thread walks kColLevels consecutive C-levels of one column x (columns
// fastest over the threads): the I-level values above a level and the C-level
// values of the next level are carried to the next level in registers.
template <typename T>
__global__ void __launch_bounds__(256) k_adv_products(const Space<T> *__restrict__ spaces, LaunchTable lt) {
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
    const uint32_t P = uint32_t(S.N1) * S.N2;  // point counts < 2^31 (checked at set-up)
    const uint32_t gid = uint32_t(blockIdx.x - E.block_begin) * blockDim.x + threadIdx.x;
    const int n3 = S.n3;
    const uint32_t nch = uint32_t(n3 + kColLevels - 1) / kColLevels;
    if (gid >= P * nch) return;
    const int c = int(gid / P);
    const uint32_t x = gid - uint32_t(c) * P;
    const int l0 = c * kColLevels, l1 = min(l0 + kColLevels, n3);
    const T *u = S.in_phys[0], *v = S.in_phys[1], *w = S.in_phys[2];
    const T *o1 = S.in_phys[3], *o2 = S.in_phys[4], *o3 = S.in_phys[5];
    T *a1 = S.src_phys[0], *a2 = S.src_phys[1], *a3 = S.src_phys[2];
    // products on the I-level below the first level (zero at the wall) and the
    // C-values of the first level
    T p2b = T(0), p1b = T(0);
    if (l0 > 0) {
        const uint32_t jb = uint32_t(l0 - 1) * P + x;
        const T wb = gload(w + jb);
        p2b = gload(o2 + jb) * wb;
        p1b = gload(o1 + jb) * wb;
    }
    T ul = gload(u + uint32_t(l0) * P + x), vl = gload(v + uint32_t(l0) * P + x);
    for (int l = l0; l < l1; ++l) {
        const uint32_t i = uint32_t(l) * P + x;
        const uint32_t ja = uint32_t(min(l, n3 - 2)) * P + x;
        const uint32_t in = uint32_t(min(l + 1, n3 - 1)) * P + x;
        const T o3l = gload(o3 + i);
        const T o2a = gload(o2 + ja), o1a = gload(o1 + ja), wa = gload(w + ja);
        const T un = gload(u + in), vn = gload(v + in);
        T p2a = T(0), p1a = T(0);
        if (l < n3 - 1) {
            p2a = o2a * wa;
            p1a = o1a * wa;
            gstore(a3 + ja, o2a * T(0.5) * (ul + un) - o1a * T(0.5) * (vl + vn));
        }
        gstore(a1 + i, o3l * vl - T(0.5) * (p2b + p2a));
        gstore(a2 + i, T(0.5) * (p1b + p1a) - o3l * ul);
        p2b = p2a;
        p1b = p1a;
        ul = un;
        vl = vn;
    }
}

// ---------------------------------------------------------------------------
// static Smagorinsky model (static_smagorinsky.jl:129-249)
// terms: 0 u, 1 v, 2 w, 3 u_1, 4 v_2, 5 u_2, 6 v_1, 7 w_1, 8 w_2
// sources: 0 R1 (vel1), 1 R2 (vel2), 2 R3 (vel3), 3 T11 (vel1_1), 4 T22 (vel2_2),
//          5 T12 (vel1_2 & vel2_1), 6 T13 (vel3_1), 7 T23 (vel3_2)
// ---------------------------------------------------------------------------
// The vertical stencil of one point (C-level l, I-levels l-1 and l), loaded
// together at clamped levels before any value is used; the strain-rate
// components are then formed from registers.
template <typename T>
struct SgsPoint {
    const Space<T> *S;
    int l, n3;
    T u[3], v[3];           // C-levels l-1, l, l+1
    T w[3];                 // I-levels l-1, l, l+1
    T u1[2], v2[2], u2[2], v1[2];  // C-levels l, l+1
    T w1[2], w2[2];         // I-levels l-1, l
    T ac[2], ai[2];         // alpha_c[l], alpha_c[l+1]; alpha_i[l], alpha_i[l+1] (set by the walk)
    T lc, li;               // filter widths of C-layer l and I-layer min(l, n3-2) (set by the walk)
    SgsPoint() = default;
    __device__ SgsPoint(const Space<T> &s, uint32_t P, uint32_t x, int lev) : S(&s), l(lev), n3(s.n3) {
        const uint32_t cm = uint32_t(max(l - 1, 0)) * P + x, c0 = uint32_t(l) * P + x;
        const uint32_t cp = uint32_t(min(l + 1, n3 - 1)) * P + x;
        const uint32_t i0 = uint32_t(min(l, n3 - 2)) * P + x, ip = uint32_t(min(l + 1, n3 - 2)) * P + x;
        const T *const *f = s.in_phys;
        u[0] = gload(f[0] + cm), u[1] = gload(f[0] + c0), u[2] = gload(f[0] + cp);
        v[0] = gload(f[1] + cm), v[1] = gload(f[1] + c0), v[2] = gload(f[1] + cp);
        w[0] = gload(f[2] + cm), w[1] = gload(f[2] + i0), w[2] = gload(f[2] + ip);
        u1[0] = gload(f[3] + c0), u1[1] = gload(f[3] + cp);
        v2[0] = gload(f[4] + c0), v2[1] = gload(f[4] + cp);
        u2[0] = gload(f[5] + c0), u2[1] = gload(f[5] + cp);
        v1[0] = gload(f[6] + c0), v1[1] = gload(f[6] + cp);
        w1[0] = gload(f[7] + cm), w1[1] = gload(f[7] + i0);
        w2[0] = gload(f[8] + cm), w2[1] = gload(f[8] + i0);
    }
    // The field values that the point of C-layer m takes in on top of those
    // it shares with the point of layer m - 1 (m clamped to the top layer),
    // loaded ahead of their use so that they are in flight while the layer
    // below is being formed.
    struct Next {
        T u, v, w, u1, v2, u2, v1, w1, w2;
    };
    __device__ Next fetch(uint32_t P, uint32_t x, int m) const {
        m = min(m, n3 - 1);
        const uint32_t cp = uint32_t(min(m + 1, n3 - 1)) * P + x, ip = uint32_t(min(m + 1, n3 - 2)) * P + x;
        const uint32_t i0 = uint32_t(min(m, n3 - 2)) * P + x;
        const T *const *f = S->in_phys;
        Next n;
        n.u = gload(f[0] + cp), n.v = gload(f[1] + cp), n.w = gload(f[2] + ip);
        n.u1 = gload(f[3] + cp), n.v2 = gload(f[4] + cp), n.u2 = gload(f[5] + cp), n.v1 = gload(f[6] + cp);
        n.w1 = gload(f[7] + i0), n.w2 = gload(f[8] + i0);
        return n;
    }
    // the point of the C-layer above (l + 1 <= n3 - 1) from the values fetched
    // for it: the same values as a point constructed there
    __device__ void advance(const Next &n) {
        ++l;
        u[0] = u[1], u[1] = u[2], u[2] = n.u;
        v[0] = v[1], v[1] = v[2], v[2] = n.v;
        w[0] = w[1], w[1] = w[2], w[2] = n.w;
        u1[0] = u1[1], u1[1] = n.u1;
        v2[0] = v2[1], v2[1] = n.v2;
        u2[0] = u2[1], u2[1] = n.u2;
        v1[0] = v1[1], v1[1] = n.v1;
        w1[0] = w1[1], w1[1] = n.w1;
        w2[0] = w2[1], w2[1] = n.w2;
    }
    // S33 on C-layer l (k = 0) or l+1 (k = 1) (vel3_3 via dx3_i2c!, physical Dirichlet values)
    __device__ T s33(int k) const {
        const int lk = l + k;
        const T wb = lk > 0 ? w[k] : T(S->wbc[0]);
        const T wa = lk < n3 - 1 ? w[k + 1] : T(S->wbc[1]);
        return (-ac[k]) * wb + ac[k] * wa;
    }
    __device__ T s12(int k) const { return (u2[k] + v1[k]) / T(2); }
    // S13, S23 on I-layer l-1 (k = 0) or l (k = 1) (vel1_3 via dx3_c2i!)
    __device__ T s13(int k) const { return ((-ai[k]) * u[k] + ai[k] * u[k + 1] + w1[k]) / T(2); }
    __device__ T s23(int k) const { return ((-ai[k]) * v[k] + ai[k] * v[k + 1] + w2[k]) / T(2); }
};

// Eddy viscosities of point q: at its C-node (evc) and at the I-node above it
// (evi; 0 on the top C-layer, which has none).
template <typename T>
__device__ __forceinline__ void sgs_visc_point(const Space<T> &S, const SgsPoint<T> &q, T &evc, T &evi) {
    const int l = q.l, n3 = S.n3;
    const T lcl = q.lc, lil = q.li;
    // C-nodes: total_strain! (:208, :225-229)
    const T S11 = q.u1[0], S22 = q.v2[0], S33 = q.s33(0), S12 = q.s12(0);
    T sc = S11 * S11 + S22 * S22 + S33 * S33 + T(2) * S12 * S12;
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        // lower neighbour (I-layer l-1 or wall), upper neighbour (I-layer l or wall)
        const T sb = l > 0 ? (comp == 0 ? q.s13(0) : q.s23(0)) : T(0);
        const T sa = l < n3 - 1 ? (comp == 0 ? q.s13(1) : q.s23(1)) : T(0);
        const T ul = comp == 0 ? q.u[1] : q.v[1];
        if (l == 0 && S.wall_rough[0]) {
            const T val = ul * T(S.fstrain[0]);
            sc += T(2) * (val * val);
        } else if (l == n3 - 1 && S.wall_rough[1]) {
            const T val = ul * T(S.fstrain[1]);
            sc += T(2) * (val * val);
        } else {
            const T s = sb + sa;
            sc += T(2) * (s * s) / T(4);
        }
    }
    sc = sqrt(T(2) * sc);
    evc = sc * (lcl * lcl);
    evi = T(0);
    // I-nodes (:209-220)
    if (l < n3 - 1) {
        const T a13 = q.s13(1), a23 = q.s23(1);
        T si = T(2) * a13 * a13 + T(2) * a23 * a23;
        T s;
        s = q.u1[0] + q.u1[1];
        si += s * s / T(4);
        s = q.v2[0] + q.v2[1];
        si += s * s / T(4);
        s = S33 + q.s33(1);
        si += s * s / T(4);
        s = S12 + q.s12(1);
        si += T(2) * (s * s) / T(4);
        si = sqrt(T(2) * si);
        evi = si * (lil * lil);
    }
}

// The stress terms of point q (index i of its C-layer) from the eddy
// viscosities at its C-node (evc), the C-node above (evn), and the I-nodes
// below (evb) and above (eva), at clamped levels.
template <typename T>
__device__ __forceinline__ void sgs_stress_point(const Space<T> &S, const SgsPoint<T> &q, uint32_t i, T evc, T evn,
                                                 T evb, T eva) {
    const int l = q.l, n3 = S.n3;
    // τ11, τ22, τ12 on C-nodes: rates vel1_1 -= τ11, vel2_2 -= τ22, vel1_2/vel2_1 -= τ12
    gstore(S.src_phys[3] + i, -(T(-2) * evc * q.u1[0]));
    gstore(S.src_phys[4] + i, -(T(-2) * evc * q.v2[0]));
    gstore(S.src_phys[5] + i, -(T(-2) * evc * q.s12(0)));
    // vel3 -= d/dx3 τ33 (dx3_c2i!) on I-layer l
    if (l < n3 - 1) {
        const T a = q.ai[1];
        const T t0 = T(-2) * evc * q.s33(0);
        const T t1 = T(-2) * evn * q.s33(1);
        gstore(S.src_phys[2] + i, -((-a) * t0 + a * t1));
        gstore(S.src_phys[6] + i, -(T(-2) * eva * q.s13(1)));
        gstore(S.src_phys[7] + i, -(T(-2) * eva * q.s23(1)));
    }
    // vel1 -= d/dx3 τ13, vel2 -= d/dx3 τ23 on C-layer l (dx3_i2c! with wall stress)
    const T ac = q.ac[0];
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        T tb, ta;
        const T ul = comp == 0 ? q.u[1] : q.v[1];
        if (l > 0) {
            tb = T(-2) * evb * (comp == 0 ? q.s13(0) : q.s23(0));
        } else if (S.wall_rough[0]) {
            const T u0 = q.u[1], v0 = q.v[1];
            tb = T(S.fstress[0]) * ul * sqrt(u0 * u0 + v0 * v0);
        } else {
            tb = T(0);
        }
        if (l < n3 - 1) {
            ta = T(-2) * eva * (comp == 0 ? q.s13(1) : q.s23(1));
        } else if (S.wall_rough[1]) {
            const T u1 = q.u[1], v1 = q.v[1];
            ta = T(S.fstress[1]) * ul * sqrt(u1 * u1 + v1 * v1);
        } else {
            ta = T(0);
        }
        gstore(S.src_phys[comp] + i, -((-ac) * tb + ac * ta));
    }
}

// Eddy viscosity and stresses in one pass: a thread walks kSgsLevels
// consecutive C-layers of one column x (columns fastest over the threads),
// computing the viscosities of each layer and of the layer below the first
// in registers; the stresses of a layer are formed once the viscosity of the
// layer above it is known.  The C-node viscosity is also stored while the τ33
// mean profile, which reads it, is collected.
template <typename T>
__global__ void __launch_bounds__(256) k_sgs_stress(const Space<T> *__restrict__ spaces, LaunchTable lt) {
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
    const uint32_t P = uint32_t(S.N1) * S.N2;  // point counts < 2^31 (checked at set-up)
    const uint32_t gid = uint32_t(blockIdx.x - E.block_begin) * blockDim.x + threadIdx.x;
    const int n3 = S.n3;
    const uint32_t nch = uint32_t(n3 + kSgsLevels - 1) / kSgsLevels;
    if (gid >= P * nch) return;
    const int c = int(gid / P);
    const uint32_t x = gid - uint32_t(c) * P;
    const int l0 = c * kSgsLevels, l1 = min(l0 + kSgsLevels, n3);
    const int js = max(l0 - 1, 0), je = min(l1, n3 - 1);
    // evc_p, evi_p: viscosities of layer j-1 (C-node, I-node above); evi_pp of layer j-2
    T evc_p = T(0), evi_p = T(0), evi_pp = T(0);
    // layer factors of the walk (at clamped levels), loaded before the fields
    // so that converting them waits for them alone
    constexpr int NW = kSgsLevels + 2;  // layers of the walk at most
    T fac[NW + 1], fai[NW + 1], flc[NW], fli[NW];
#pragma unroll
    for (int t = 0; t <= NW; ++t) {
        fac[t] = T(gload(S.alpha_c + min(js + t, n3 - 1)));
        fai[t] = T(gload(S.alpha_i + min(js + t, n3)));
        if (t < NW) {
            flc[t] = T(gload(S.lc + min(js + t, n3 - 1)));
            fli[t] = T(gload(S.li + min(js + t, n3 - 2)));
        }
    }
    // the walk is unrolled over the chunk and the loads of layer j+1 are issued
    // before layer j is formed (unconditionally, at a clamped layer), so one
    // layer's loads are always in flight, without waits at joins or on copies
    SgsPoint<T> prev, q(S, P, x, js);
    typename SgsPoint<T>::Next nx[2];
#pragma unroll
    for (int t = 0; t < NW; ++t) {
        const int j = js + t;
        if (j > je) break;
        nx[t & 1] = q.fetch(P, x, j + 1);
        if (t > 0) q.advance(nx[(t - 1) & 1]);
        q.ac[0] = fac[t], q.ac[1] = fac[t + 1], q.ai[0] = fai[t], q.ai[1] = fai[t + 1];
        q.lc = flc[t], q.li = fli[t];
        T evc, evi;
        sgs_visc_point(S, q, evc, evi);
        if (S.store_visc && j >= l0 && j < l1) gstore(S.visc_c + uint32_t(j) * P + x, evc);
        const int l = j - 1;
        if (l >= l0)  // layer l < n3-1: evn = evc(l+1), evb = evi(max(l-1, 0)), eva = evi(l)
            sgs_stress_point(S, prev, uint32_t(l) * P + x, evc_p, evc, l > 0 ? evi_pp : evi_p, evi_p);
        evi_pp = evi_p;
        evi_p = evi;
        evc_p = evc;
        prev = q;
    }
    // the top layer: evn = its own C-node viscosity, evb = eva = evi(n3-2)
    if (l1 == n3) sgs_stress_point(S, prev, uint32_t(n3 - 1) * P + x, evc_p, evc_p, evi_pp, evi_pp);
}

// ---------------------------------------------------------------------------
// linear terms: du = ν·Laplacian(u) + S (molecular_diffusion.jl:57-72, sources.jl:32-35)
// ---------------------------------------------------------------------------
template <typename T, typename A>
__device__ __forceinline__ cplx<A> laplacian_point(const LinDesc<T> &L, int f, int l, int m, const cplx<A> &q,
                                                   const cplx<A> &lo, const cplx<A> &hi) {
    const int k1 = m % L.K1, i2 = m / L.K1;
    const A dd = A(L.DD1[k1]) + A(L.DD2[i2]);
    A a0, a1, a2;
    if (f < 2) {
        a0 = A(L.alpha_i[l]);
        a1 = A(L.alpha_c[l]);
        a2 = A(L.alpha_i[l + 1]);
    } else {
        a0 = A(L.alpha_c[l]);
        a1 = A(L.alpha_i[l + 1]);
        a2 = A(L.alpha_c[l + 1]);
    }
    const A nu = A(L.nu[f]);
    const A c0 = a0 * a1, c2 = a1 * a2;
    const A c1 = dd - a0 * a1 - a1 * a2;
    return mk(nu * (c0 * lo.x + c1 * q.x + c2 * hi.x), nu * (c0 * lo.y + c1 * q.y + c2 * hi.y));
}

// ghost layer of a C-node field from 4th-order extrapolation (boundary_conditions.jl:286-356)
template <typename A>
__device__ __forceinline__ cplx<A> ghost(int kind, double value, double difference, bool mean, bool upper,
                                        cplx<A> l1, cplx<A> l2, cplx<A> l3) {
    cplx<A> g;
    if (kind == 0) {
        g = mk((A(-15) * l1.x + A(5) * l2.x - l3.x) / A(5), (A(-15) * l1.y + A(5) * l2.y - l3.y) / A(5));
        if (mean) g.x += A(16) * A(value) / A(5);
    } else {
        g = mk((A(21) * l1.x + A(3) * l2.x - l3.x) / A(23), (A(21) * l1.y + A(3) * l2.y - l3.y) / A(23));
        const A l0 = upper ? -A(difference) : A(difference);
        if (mean) g.x -= A(24) * l0 / A(23);
    }
    return g;
}

template <typename T>
__global__ void __launch_bounds__(256) k_linear(const LinDesc<T> *__restrict__ descs, LaunchTable lt, int sel,
                                                int *step_count) {
    // the trailing rate evaluation of a step counts the step (read by the
    // divergence records of the kernels ordered after this one)
    if (step_count && blockIdx.x == 0 && threadIdx.x == 0) step_count[0] += 1;
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const LinDesc<T> &L = descs[E.space];
    const int nk = L.nk, n3 = L.n3;
    // one job per field: the field index is uniform over the block, so the
    // descriptor fields it selects are scalar loads; a thread takes kLinLevels
    // consecutive levels of one mode (modes fastest over the threads)
    const int bj = blockIdx.x - E.block_begin;
    const int f = bj / E.blocks_per_job;
    const int nl = f < 2 ? n3 : n3 - 1;
    const uint32_t nch = uint32_t(nl + kLinLevels - 1) / kLinLevels;
    const uint32_t r = uint32_t(bj - f * E.blocks_per_job) * blockDim.x + threadIdx.x;
    if (r >= nch * uint32_t(nk)) return;
    const int c = int(r / uint32_t(nk)), m = int(r - uint32_t(c) * nk);
    const int l0 = c * kLinLevels;
    const cplx<T> *q = L.state[sel][f];
    cplx<T> *du = L.du[f];
    const cplx<double> *qs = L.shadow[sel][f];
    if (qs != nullptr && m < kLinLevels) {
        // (0,0) mode of the mixed form, float64 from the float64 shadow: one
        // level per thread of the chunk's first modes (mode m takes levels
        // l0 + m, l0 + m + nk, ...), so that its levels run side by side
        for (int j = m; j < kLinLevels; j += nk) {
            const int l = l0 + j;
            if (l >= nl) break;
            cplx<double> res = mk(0.0, 0.0);
            if (L.has_diff[f]) {
                cplx<double> lo, hi, cc = qs[l];
                if (f < 2) {
                    lo = l > 0 ? qs[l - 1] : ghost<double>(L.bc_kind[f][0], L.bc_val[f][0], L.bc_diff[f][0], true, false, qs[0], qs[1], qs[2]);
                    hi = l < nl - 1 ? qs[l + 1]
                                    : ghost<double>(L.bc_kind[f][1], L.bc_val[f][1], L.bc_diff[f][1], true, true, qs[nl - 1], qs[nl - 2], qs[nl - 3]);
                } else {
                    lo = l > 0 ? qs[l - 1] : mk(L.bc_val[2][0], 0.0);
                    hi = l < nl - 1 ? qs[l + 1] : mk(L.bc_val[2][1], 0.0);
                }
                res = laplacian_point<T, double>(L, f, l, 0, cc, lo, hi);
            }
            res.x += L.src[f];
            L.du00[f][l] = res;
            du[size_t(l) * nk] = mk(T(0), T(0));
        }
    }
    if (m == 0 && qs != nullptr) return;
    const bool diff = L.has_diff[f] != 0;
    const T src = T(L.src[f]);
    // the column window: levels l0-1 .. l0+kLinLevels at clamped indices, and
    // the second level from a boundary for the ghost layers
    cplx<T> w[kLinLevels + 2], qx0 = mk(T(0), T(0)), qx1 = mk(T(0), T(0));
    T dd = T(0);
    if (diff) {
#pragma unroll
        for (int i = 0; i < kLinLevels + 2; ++i) w[i] = gload(q + size_t(min(max(l0 - 1 + i, 0), nl - 1)) * nk + m);
        qx0 = gload(q + size_t(min(2, nl - 1)) * nk + m);
        qx1 = gload(q + size_t(max(nl - 3, 0)) * nk + m);
        const int k1 = m % L.K1, i2 = m / L.K1;
        dd = T(gload(L.DD1 + k1)) + T(gload(L.DD2 + i2));
    }
    const double *pa = f < 2 ? L.alpha_i : L.alpha_c, *pb = f < 2 ? L.alpha_c : L.alpha_i;
    const T nu = T(L.nu[f]);
#pragma unroll
    for (int j = 0; j < kLinLevels; ++j) {
        const int l = l0 + j;
        if (l >= nl) break;
        cplx<T> res = mk(T(0), T(0));
        if (diff) {
            const cplx<T> cc = w[j + 1], qlo = w[j], qhi = w[j + 2];
            const T a0 = T(gload(pa + l)), a1 = T(gload(pb + l + (f < 2 ? 0 : 1))), a2 = T(gload(pa + l + 1));
            cplx<T> lo, hi;
            if (f < 2) {
                lo = l > 0 ? qlo : ghost<T>(L.bc_kind[f][0], L.bc_val[f][0], L.bc_diff[f][0], m == 0, false, cc, qhi, qx0);
                hi = l < nl - 1 ? qhi : ghost<T>(L.bc_kind[f][1], L.bc_val[f][1], L.bc_diff[f][1], m == 0, true, cc, qlo, qx1);
            } else {
                // I-nodes: Dirichlet values enter every mode (boundary_conditions.jl:263-274)
                lo = l > 0 ? qlo : mk(T(L.bc_val[2][0]), T(0));
                hi = l < nl - 1 ? qhi : mk(T(L.bc_val[2][1]), T(0));
            }
            const T c0 = a0 * a1, c2 = a1 * a2;
            const T c1 = dd - a0 * a1 - a1 * a2;
            res = mk(nu * (c0 * lo.x + c1 * cc.x + c2 * hi.x), nu * (c0 * lo.y + c1 * cc.y + c2 * hi.y));
        }
        if (m == 0) res.x += src;
        gstore(du + size_t(l) * nk + m, res);
    }
}

// Divergence record (the isfinite check of BoundaryLayerDynamics.jl:155-157):
// flag[1] counts the steps whose trailing rate evaluation has started (bumped
// by the linear-term kernel of that evaluation), so a non-finite rate found by
// the kernel that consumes it is recorded as the number of time advances made
// before the reference would raise; flag[0] keeps the smallest such count.
__device__ __forceinline__ void record_diverged(int *flag) { atomicMin(flag, flag[1]); }

// ---------------------------------------------------------------------------
// SSPRK33 stage combination + pressure projection, one thread per (k1,k2)
// ---------------------------------------------------------------------------
template <typename T, typename A>
__device__ __forceinline__ cplx<A> combine(int stage, cplx<A> u, cplx<A> ut, cplx<A> du, A dt) {
    switch (stage) {
        case STAGE_RK_1: return mk(u.x + dt * du.x, u.y + dt * du.y);  // utmp = u + dt·du
        case STAGE_RK3_2:
            return mk((A(3) * u.x + ut.x + dt * du.x) / A(4), (A(3) * u.y + ut.y + dt * du.y) / A(4));
        case STAGE_RK3_3:
            return mk((u.x + A(2) * ut.x + A(2) * dt * du.x) / A(3), (u.y + A(2) * ut.y + A(2) * dt * du.y) / A(3));
        case STAGE_EULER:
        case STAGE_AB2_FIRST_1: return mk(u.x + dt * du.x, u.y + dt * du.y);  // u += dt·du
        case STAGE_RK2_2:  // u = (u + utmp + dt·du)/2
            return mk((u.x + ut.x + dt * du.x) / A(2), (u.y + ut.y + dt * du.y) / A(2));
        case STAGE_AB2:  // u += dt/2·(3·du − du_prev), du_prev in utmp
            return mk(u.x + dt / A(2) * (A(3) * du.x - ut.x), u.y + dt / A(2) * (A(3) * du.y - ut.y));
        case STAGE_AB2_FIRST_2:  // u += dt/2·(du − du_prev)
            return mk(u.x + dt / A(2) * (du.x - ut.x), u.y + dt / A(2) * (du.y - ut.y));
        default: return u;  // projection of the current state only
    }
}

template <typename T>
__global__ void __launch_bounds__(64) k_stage_project(ProjDesc<T> D, int stage, double dt, int mode_limit) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= D.nk || m >= mode_limit) return;
    const int nk = D.nk, n3 = D.n3;
    const bool mixed00 = (m == 0) && (D.u00[0] != nullptr);
    // input/output selection: stage 0 projects u in place, the first stages of
    // the Runge-Kutta methods write utmp, the others u
    cplx<T> *const *outp = stage_writes_ut(stage) ? D.ut : D.u;
    cplx<double> **outs = stage_writes_ut(stage) ? D.ut00 : D.u00;
    bool bad = false;
    auto fin = [&](cplx<T> v) {
        if (!(isfinite(v.x) && isfinite(v.y))) bad = true;
    };
    auto comb = [&](int f, int l) -> cplx<double> {
        const size_t i = size_t(l) * nk + m;
        if (mixed00) {
            cplx<double> du = mk(0.0, 0.0);
            if (stage > 0) {
                const cplx<T> dn = D.du[f][i];
                fin(dn);
                du = mk(D.du00[f][l].x + double(dn.x), D.du00[f][l].y + double(dn.y));
                if (!(isfinite(du.x) && isfinite(du.y))) bad = true;
            }
            const cplx<double> c = combine<T, double>(stage, D.u00[f][l], D.ut00[f][l], du, dt);
            if (stage_saves_du(stage)) D.ut00[f][l] = du;
            return c;
        }
        const cplx<T> u = D.u[f][i];
        cplx<T> ut = mk(T(0), T(0)), du = mk(T(0), T(0));
        if (stage_reads_ut(stage)) ut = D.ut[f][i];
        if (stage > 0) {
            du = D.du[f][i];
            fin(du);
        }
        const cplx<T> c = combine<T, T>(stage, u, ut, du, T(dt));
        if (stage_saves_du(stage)) D.ut[f][i] = du;
        return mk(double(c.x), double(c.y));
    };
    auto store = [&](int f, int l, cplx<double> v) {
        const size_t i = size_t(l) * nk + m;
        outp[f][i] = mk(T(v.x), T(v.y));
        if (mixed00) outs[f][l] = v;
    };
    auto load_out = [&](int f, int l) -> cplx<double> {
        if (mixed00) return outs[f][l];
        const cplx<T> v = outp[f][size_t(l) * nk + m];
        return mk(double(v.x), double(v.y));
    };

    if (!D.has_pressure) {
        if (stage > 0) {
            for (int l = 0; l < n3; ++l) {
                store(0, l, comb(0, l));
                store(1, l, comb(1, l));
                if (l < n3 - 1) store(2, l, comb(2, l));
            }
        }
        if (bad) record_diverged(D.flag);
        return;
    }
    const int k1 = m % D.K1, i2 = m / D.K1;
    const double d1 = D.D1[k1], d2 = D.D2[i2];
    // forward sweep: combine, divergence (div!, pressure.jl:368-378), elimination (:126-132)
    cplx<double> wprev = mk(0.0, 0.0), yprev = mk(0.0, 0.0);
    for (int l = 0; l < n3; ++l) {
        const cplx<double> c1 = comb(0, l), c2 = comb(1, l);
        store(0, l, c1);
        store(1, l, c2);
        cplx<double> c3 = mk(0.0, 0.0);
        if (l < n3 - 1) {
            c3 = comb(2, l);
            store(2, l, c3);
        }
        const double hf = 1.0 / D.alpha_c[l];
        // u1·D1·hf + u2·D2·hf with D = i·d
        cplx<double> r = mk((-d1 * c1.y) * hf + (-d2 * c2.y) * hf, (d1 * c1.x) * hf + (d2 * c2.x) * hf);
        if (l > 0) {
            r = mk(r.x - wprev.x, r.y - wprev.y);
        }
        if (l < n3 - 1) {
            r = mk(r.x + c3.x, r.y + c3.y);
        }
        if (m == 0) {
            if (l == 0) r.x -= D.wbc[0];
            if (l == n3 - 1) r.x += D.wbc[1];
            if (l == n3 - 1) r = mk(0.0, 0.0);  // adjust_pressure_rhs! (:396)
        }
        const double g = D.gamma[size_t(l) * nk + m];
        cplx<double> y;
        if (l == 0) {
            y = mk(g * r.x, g * r.y);
        } else {
            const double b = D.alpha_i[l];
            y = mk(g * (r.x - b * yprev.x), g * (r.y - b * yprev.y));
        }
        D.pbuf[size_t(l) * nk + m] = y;
        yprev = y;
        wprev = c3;
    }
    // backward sweep (:155-160) and gradient (add_gradient!, :405-426)
    cplx<double> pnext = mk(0.0, 0.0);
    for (int l = n3 - 1; l >= 0; --l) {
        cplx<double> p = D.pbuf[size_t(l) * nk + m];
        if (l < n3 - 1) {
            const double gb = D.gamma[size_t(l) * nk + m] * D.alpha_i[l + 1];
            p = mk(p.x - gb * pnext.x, p.y - gb * pnext.y);
        }
        cplx<double> v1 = load_out(0, l), v2 = load_out(1, l);
        // f += (-1·D)·p with D = i·d: (-i d)·p = (d·p.y, -d·p.x)
        v1 = mk(v1.x + d1 * p.y, v1.y - d1 * p.x);
        v2 = mk(v2.x + d2 * p.y, v2.y - d2 * p.x);
        store(0, l, v1);
        store(1, l, v2);
        if (l < n3 - 1) {
            const double a = D.alpha_i[l + 1];
            cplx<double> v3 = load_out(2, l);
            const cplx<double> g = mk((-a) * p.x + a * pnext.x, (-a) * p.y + a * pnext.y);
            v3 = mk(v3.x - g.x, v3.y - g.y);
            store(2, l, v3);
        }
        pnext = p;
    }
    if (bad) record_diverged(D.flag);
}


// Same stage + projection with each (k1,k2) column split into chunks of CL
// levels held in registers by different threads: lanes are consecutive modes
// (coalesced), threadIdx / MB is the chunk.  Both triangular sweeps of the
// LDLt solve are affine recurrences y_l = a_l·y_{l-1} + b_l; each chunk runs
// its sweep from a zero incoming value while tracking the product of the a_l,
// and the true values follow from one combine over the chunks through LDS.
// In the first (forward) / last (backward) chunk this is the sequential
// recurrence of pressure.jl:126-160 operation for operation.
template <typename T, int CL>
__device__ __forceinline__ void stage_chunked_body(const ProjDesc<T> &D, int blk, int stage, double dt,
                                                   unsigned char *smem_raw) {
    const int MB = D.MB, skip0 = D.skip0;
    const int NC = blockDim.x / MB;
    cplx<double> *s_y = reinterpret_cast<cplx<double> *>(smem_raw);  // [NC][MB]
    double *s_p = reinterpret_cast<double *>(s_y + NC * MB);         // [NC][MB]
    cplx<T> *s_w = reinterpret_cast<cplx<T> *>(s_p + NC * MB);      // [NC][MB]
    const int tx = threadIdx.x % MB, ty = threadIdx.x / MB;
    const int m = blk * MB + tx;
    const int nk = D.nk, n3 = D.n3;
    const bool active = m < nk && !(skip0 && m == 0);
    const int l0 = ty * CL;
    const int mm = active ? m : 0;
    cplx<T> *const *outp = stage_writes_ut(stage) ? D.ut : D.u;
    const bool rd_ut = stage_reads_ut(stage), rd_du = stage > 0, sv_du = stage_saves_du(stage);
    // Every load of the chunk is issued before any value is used: indices are
    // clamped into the arrays (vel3 has n3-1 levels) and values outside the
    // column are replaced by zero afterwards, so the loads carry no branches.
    cplx<T> c1[CL], c2[CL], c3[CL];
    double gam[CL], ac[CL], ai[CL + 1];
    {
        cplx<T> lu[3][CL], lt[3][CL], ld[3][CL];
        cplx<float> le[3][CL];
        const cplx<T> *const *tsrc = rd_ut ? D.ut : D.u;
        const cplx<T> *const *dsrc = rd_du ? D.du : D.u;
#pragma unroll
        for (int j = 0; j < CL; ++j) {
            const int lc = min(l0 + j, n3 - 1), li = min(l0 + j, n3 - 2);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const size_t i = size_t(f == 2 ? li : lc) * nk + mm;
                lu[f][j] = gload(D.u[f] + i);
                lt[f][j] = gload(tsrc[f] + i);
                ld[f][j] = gload(dsrc[f] + i);
                if (D.ext_stride) le[f][j] = gload(D.du_add[f] + size_t(f == 2 ? li : lc) * D.ext_stride);
            }
            gam[j] = D.has_pressure ? gload(D.gamma + size_t(lc) * nk + mm) : 1.0;
            ac[j] = gload(D.alpha_c + lc);
            ai[j] = gload(D.alpha_i + lc);
        }
        ai[CL] = gload(D.alpha_i + min(l0 + CL, n3));
        bool bad = false;
#pragma unroll
        for (int j = 0; j < CL; ++j) {
            const int l = l0 + j;
            const bool ok = active && l < n3;
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const bool okf = ok && (f < 2 || l < n3 - 1);
                cplx<T> du = mk(T(0), T(0)), ut = mk(T(0), T(0));
                if (rd_ut) ut = lt[f][j];
                if (rd_du) {
                    du = ld[f][j];
                    if (D.ext_stride) du = mk(du.x + T(le[f][j].x), du.y + T(le[f][j].y));
                    if (okf && !(isfinite(du.x) && isfinite(du.y))) bad = true;
                }
                cplx<T> c = combine<T, T>(stage, lu[f][j], ut, du, T(dt));
                if (sv_du && okf) {  // Adams–Bashforth: keep f(u) for the next step
                    const size_t i = size_t(l) * nk + m;
                    gstore(D.ut[f] + i, du);
                }
                if (!okf) c = mk(T(0), T(0));
                if (f == 0) c1[j] = c;
                else if (f == 1) c2[j] = c;
                else c3[j] = c;
            }
        }
        if (bad) record_diverged(D.flag);
    }
    if (!D.has_pressure) {
        if (stage > 0 && active) {
            cplx<float> *const *mir = stage_writes_ut(stage) ? D.mirror_ut : D.mirror_u;
#pragma unroll
            for (int j = 0; j < CL; ++j) {
                const int l = l0 + j;
                if (l < n3) {
                    const size_t i = size_t(l) * nk + m;
                    gstore(outp[0] + i, c1[j]);
                    gstore(outp[1] + i, c2[j]);
                    if (l < n3 - 1) gstore(outp[2] + i, c3[j]);
                    if (D.ext_stride) {
                        const size_t e = size_t(l) * D.ext_stride;
                        gstore(mir[0] + e, mk(float(c1[j].x), float(c1[j].y)));
                        gstore(mir[1] + e, mk(float(c2[j].x), float(c2[j].y)));
                        if (l < n3 - 1) gstore(mir[2] + e, mk(float(c3[j].x), float(c3[j].y)));
                    }
                }
            }
        }
        return;
    }
    const int k1 = mm % D.K1, i2 = mm / D.K1;
    const double d1 = gload(D.D1 + k1), d2 = gload(D.D2 + i2);
    // w at the I-level below this chunk (last level of the previous chunk)
    {
        const int jl = min(CL, n3 - l0) - 1;
        cplx<T> last = mk(T(0), T(0));
#pragma unroll
        for (int j = 0; j < CL; ++j)
            if (j == jl) last = c3[j];
        s_w[ty * MB + tx] = last;
    }
    __syncthreads();
    cplx<double> wb = mk(0.0, 0.0);
    if (ty > 0) {
        const cplx<T> v = s_w[(ty - 1) * MB + tx];
        wb = mk(double(v.x), double(v.y));
    }
    // forward sweep from a zero incoming value
    cplx<double> y[CL];
    double P[CL], g[CL];
    cplx<double> yprev = mk(0.0, 0.0);
    double Pprev = 1.0;
#pragma unroll
    for (int j = 0; j < CL; ++j) {
        const int l = l0 + j;
        y[j] = mk(0.0, 0.0);
        P[j] = 0.0;
        g[j] = 0.0;
        if (l < n3) {
            const cplx<double> a1 = mk(double(c1[j].x), double(c1[j].y));
            const cplx<double> a2 = mk(double(c2[j].x), double(c2[j].y));
            const cplx<double> a3 = mk(double(c3[j].x), double(c3[j].y));
            const double hf = 1.0 / ac[j];
            cplx<double> r = mk((-d1 * a1.y) * hf + (-d2 * a2.y) * hf, (d1 * a1.x) * hf + (d2 * a2.x) * hf);
            if (l > 0) r = mk(r.x - wb.x, r.y - wb.y);
            if (l < n3 - 1) r = mk(r.x + a3.x, r.y + a3.y);
            if (m == 0) {
                if (l == 0) r.x -= D.wbc[0];
                if (l == n3 - 1) r.x += D.wbc[1];
                if (l == n3 - 1) r = mk(0.0, 0.0);
            }
            const double gl = active ? gam[j] : 1.0;
            g[j] = gl;
            if (l == 0) {
                y[j] = mk(gl * r.x, gl * r.y);
                P[j] = 0.0;
            } else {
                const double b = ai[j];
                y[j] = mk(gl * (r.x - b * yprev.x), gl * (r.y - b * yprev.y));
                P[j] = -(gl * b) * Pprev;
            }
            yprev = y[j];
            Pprev = P[j];
            wb = a3;
        }
    }
    s_y[ty * MB + tx] = yprev;
    s_p[ty * MB + tx] = Pprev;
    __syncthreads();
    cplx<double> yin = mk(0.0, 0.0);
#pragma unroll 8
    for (int c = 0; c < ty; ++c) {
        const cplx<double> v = s_y[c * MB + tx];
        const double pp = s_p[c * MB + tx];
        yin = mk(v.x + pp * yin.x, v.y + pp * yin.y);
    }
    if (ty > 0) {
#pragma unroll
        for (int j = 0; j < CL; ++j) y[j] = mk(y[j].x + P[j] * yin.x, y[j].y + P[j] * yin.y);
    }
    __syncthreads();
    // backward sweep from a zero value above the chunk
    cplx<double> pnext = mk(0.0, 0.0);
    double Qnext = 1.0;
#pragma unroll
    for (int j = CL - 1; j >= 0; --j) {
        const int l = l0 + j;
        if (l < n3) {
            if (l == n3 - 1) {
                P[j] = 0.0;  // Q
            } else {
                const double gb = g[j] * ai[j + 1];
                y[j] = mk(y[j].x - gb * pnext.x, y[j].y - gb * pnext.y);
                P[j] = -gb * Qnext;
            }
            pnext = y[j];
            Qnext = P[j];
        }
    }
    s_y[ty * MB + tx] = pnext;  // p̃ at l0
    s_p[ty * MB + tx] = Qnext;
    __syncthreads();
    cplx<double> pin = mk(0.0, 0.0);  // p at the level just above this chunk
#pragma unroll 8
    for (int c = NC - 1; c > ty; --c) {
        const cplx<double> v = s_y[c * MB + tx];
        const double q = s_p[c * MB + tx];
        pin = mk(v.x + q * pin.x, v.y + q * pin.y);
    }
    if (ty < NC - 1) {
#pragma unroll
        for (int j = 0; j < CL; ++j) y[j] = mk(y[j].x + P[j] * pin.x, y[j].y + P[j] * pin.y);
    }
    if (!active) return;
    // gradient (add_gradient!, pressure.jl:405-426) and store
    cplx<float> *const *mir = stage_writes_ut(stage) ? D.mirror_ut : D.mirror_u;
#pragma unroll
    for (int j = 0; j < CL; ++j) {
        const int l = l0 + j;
        if (l < n3) {
            const size_t i = size_t(l) * nk + m;
            const cplx<double> p = y[j];
            cplx<double> v1 = mk(double(c1[j].x), double(c1[j].y));
            cplx<double> v2 = mk(double(c2[j].x), double(c2[j].y));
            v1 = mk(v1.x + d1 * p.y, v1.y - d1 * p.x);
            v2 = mk(v2.x + d2 * p.y, v2.y - d2 * p.x);
            gstore(outp[0] + i, mk(T(v1.x), T(v1.y)));
            gstore(outp[1] + i, mk(T(v2.x), T(v2.y)));
            const size_t e = size_t(l) * D.ext_stride;
            if (D.ext_stride) {
                gstore(mir[0] + e, mk(float(v1.x), float(v1.y)));
                gstore(mir[1] + e, mk(float(v2.x), float(v2.y)));
            }
            if (l < n3 - 1) {
                const cplx<double> pn = (j + 1 < CL) ? y[j + 1] : pin;
                const double a = ai[j + 1];
                const cplx<double> gr = mk((-a) * p.x + a * pn.x, (-a) * p.y + a * pn.y);
                cplx<double> v3 = mk(double(c3[j].x), double(c3[j].y));
                v3 = mk(v3.x - gr.x, v3.y - gr.y);
                gstore(outp[2] + i, mk(T(v3.x), T(v3.y)));
                if (D.ext_stride) gstore(mir[2] + e, mk(float(v3.x), float(v3.y)));
            }
        }
    }
}

// One launch of the stage kernel: blocks below nmain cover the columns of
// every model (launch table); the blocks after them take the float64 (0,0)
// column of each mixed-form model, so that it needs no second stream.
template <typename T, int CL>
__global__ void __launch_bounds__(256) k_stage_chunked(const ProjDesc<T> *__restrict__ descs, LaunchTable lt, int stage,
                                                       double dt, const ProjDesc<double> *__restrict__ descs00,
                                                       int nmain) {
    extern __shared__ unsigned char smem_raw[];
    if (int(blockIdx.x) >= nmain) {
        if constexpr (sizeof(T) != sizeof(double))
            stage_chunked_body<double, stage00_cl(CL)>(descs00[blockIdx.x - nmain], 0, stage, dt, smem_raw);
        return;
    }
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    stage_chunked_body<T, CL>(descs[E.space], blockIdx.x - E.block_begin, stage, dt, smem_raw);
}

// ConstantMean projection (sources.jl:63-77): shift the (0,0) column of each
// field with a prescribed mean so that Σ_l w_l·Re q(0,0,l) equals it; one block
// per instance, the float64 shadow first where the mixed form keeps one
template <typename T>
__global__ void __launch_bounds__(256) k_constant_mean(CMeanDesc<T> D, int out_ut) {
    __shared__ double red[256];
    for (int f = 0; f < 2; ++f) {
        if (!D.on[f]) continue;
        cplx<T> *q = out_ut ? D.ut[f] : D.u[f];
        cplx<double> *q00 = out_ut ? D.ut00[f] : D.u00[f];
        double s = 0.0;
        for (int l = threadIdx.x; l < D.n3; l += blockDim.x)
            s += D.w[l] * (q00 ? q00[l].x : double(q[size_t(l) * D.nk].x));
        red[threadIdx.x] = s;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        const double shift = D.value[f] - red[0];
        __syncthreads();
        for (int l = threadIdx.x; l < D.n3; l += blockDim.x) {
            if (q00) {
                q00[l].x += shift;
                q[size_t(l) * D.nk].x = T(q00[l].x);
            } else {
                q[size_t(l) * D.nk].x = T(double(q[size_t(l) * D.nk].x) + shift);
            }
        }
    }
}

// finite check of the rates (BoundaryLayerDynamics.jl:155-157)
template <typename T>
__global__ void __launch_bounds__(256) k_check_finite(const cplx<T> *a, size_t n, int *flag) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const cplx<T> v = a[i];
        if (!(isfinite(v.x) && isfinite(v.y))) record_diverged(flag);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(cplx<T> *a, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        a[i] = mk(T(0), T(0));
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// One transform kernel kind (PassKind) over the blocks of a launch table.
template <typename T, int KIND>
void launch_xf(const Space<T> *spaces_dev, const LaunchTable &lt, int nblocks, size_t smem, int sel, hipStream_t st) {
    if constexpr (KIND == PASS_C2R_X2) k_c2r_x2<T><<<nblocks, 256, smem, st>>>(spaces_dev, lt, sel);
    else if constexpr (KIND == PASS_C2R_X1) k_c2r_x1<T><<<nblocks, 256, smem, st>>>(spaces_dev, lt);
    else if constexpr (KIND == PASS_R2C_X1) k_r2c_x1<T><<<nblocks, 256, smem, st>>>(spaces_dev, lt);
    else if constexpr (KIND == PASS_R2C_X2) k_r2c_x2<T><<<nblocks, 256, smem, st>>>(spaces_dev, lt);
    else if constexpr (KIND == PASS_C2R_2D) k_c2r_2d<T, 1024><<<nblocks, 1024, smem, st>>>(spaces_dev, lt, sel);
    else if constexpr (KIND == PASS_C2R_2D_HALF) k_c2r_2d<T, 512><<<nblocks, 512, smem, st>>>(spaces_dev, lt, sel);
    else if constexpr (KIND == PASS_C2R_2H) k_c2r_2h<T><<<nblocks, 512, smem, st>>>(spaces_dev, lt, sel);
    else k_r2c_2d<T><<<nblocks, kC2r2dThreads, smem, st>>>(spaces_dev, lt);
}

// the fused transforms keep a whole plane in LDS: allow up to 160 KiB per block
template <typename T, int KIND>
hipError_t xf_allow_large_lds() {
    const void *f = nullptr;
    if constexpr (KIND == PASS_C2R_2D) f = reinterpret_cast<const void *>(&k_c2r_2d<T, 1024>);
    else if constexpr (KIND == PASS_C2R_2D_HALF) f = reinterpret_cast<const void *>(&k_c2r_2d<T, 512>);
    else if constexpr (KIND == PASS_C2R_2H) f = reinterpret_cast<const void *>(&k_c2r_2h<T>);
    else if constexpr (KIND == PASS_R2C_2D) f = reinterpret_cast<const void *>(&k_r2c_2d<T>);
    if (!f) return hipSuccess;
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

// Each transform kind is compiled in a translation unit of its own (the
// BLD_KERNEL_PART builds of this file), so the unrolled FFT bodies of all
// lengths build in parallel.
#define BLD_XF_DECL(T, K)                                                                                      \
    extern template void launch_xf<T, K>(const Space<T> *, const LaunchTable &, int, size_t, int, hipStream_t); \
    extern template hipError_t xf_allow_large_lds<T, K>();
#define BLD_XF_INST(T, K)                                                                               \
    template void launch_xf<T, K>(const Space<T> *, const LaunchTable &, int, size_t, int, hipStream_t); \
    template hipError_t xf_allow_large_lds<T, K>();

#if BLD_KERNEL_PART == 0
#define BLD_XF_ALL(M, T) \
    M(T, PASS_C2R_X2) M(T, PASS_C2R_X1) M(T, PASS_R2C_X1) M(T, PASS_R2C_X2) M(T, PASS_C2R_2D) M(T, PASS_C2R_2D_HALF) \
    M(T, PASS_C2R_2H) M(T, PASS_R2C_2D)
BLD_XF_ALL(BLD_XF_DECL, float)
BLD_XF_ALL(BLD_XF_DECL, double)

template <typename T>
void launch_transform(int kind, const Space<T> *spaces_dev, const LaunchTable &lt, int nblocks, size_t smem, int sel,
                      hipStream_t st) {
    if (nblocks <= 0) return;
    switch (kind) {
        case PASS_C2R_X2: launch_xf<T, PASS_C2R_X2>(spaces_dev, lt, nblocks, smem, sel, st); break;
        case PASS_C2R_X1: launch_xf<T, PASS_C2R_X1>(spaces_dev, lt, nblocks, smem, sel, st); break;
        case PASS_R2C_X1: launch_xf<T, PASS_R2C_X1>(spaces_dev, lt, nblocks, smem, sel, st); break;
        case PASS_R2C_X2: launch_xf<T, PASS_R2C_X2>(spaces_dev, lt, nblocks, smem, sel, st); break;
        case PASS_C2R_2D: launch_xf<T, PASS_C2R_2D>(spaces_dev, lt, nblocks, smem, sel, st); break;
        case PASS_C2R_2D_HALF: launch_xf<T, PASS_C2R_2D_HALF>(spaces_dev, lt, nblocks, smem, sel, st); break;
        case PASS_C2R_2H: launch_xf<T, PASS_C2R_2H>(spaces_dev, lt, nblocks, smem, sel, st); break;
        case PASS_R2C_2D: launch_xf<T, PASS_R2C_2D>(spaces_dev, lt, nblocks, smem, sel, st); break;
    }
}

template <typename T>
hipError_t allow_large_lds() {
    for (hipError_t e : {xf_allow_large_lds<T, PASS_C2R_2D>(), xf_allow_large_lds<T, PASS_C2R_2D_HALF>(),
                         xf_allow_large_lds<T, PASS_C2R_2H>(), xf_allow_large_lds<T, PASS_R2C_2D>()})
        if (e != hipSuccess) return e;
    return hipSuccess;
}

template <typename T>
void launch_physics(int kind, const Space<T> *spaces_dev, const LaunchTable &lt, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return;
    switch (kind) {
        case PHYS_ADV: k_adv_products<T><<<nblocks, 256, 0, st>>>(spaces_dev, lt); break;
        case PHYS_SGS_STRESS: k_sgs_stress<T><<<nblocks, 256, 0, st>>>(spaces_dev, lt); break;
    }
}

template <typename T>
void launch_constant_mean(const CMeanDesc<T> &D, int stage, hipStream_t st) {
    k_constant_mean<T><<<1, 256, 0, st>>>(D, stage_writes_ut(stage) ? 1 : 0);
}

template <typename T>
void launch_linear(const LinDesc<T> *descs_dev, const LaunchTable &lt, int nblocks, int sel, int *step_count,
                   hipStream_t st) {
    if (nblocks > 0) k_linear<T><<<nblocks, 256, 0, st>>>(descs_dev, lt, sel, step_count);
}

// One launch of the chunked stage kernel over all descriptors of the table;
// CL is the chunk length shared by the launch (block = 256 threads = MB·NC).
template <typename T>
void launch_stage_chunked(int CL, const ProjDesc<T> *descs_dev, const LaunchTable &lt, int nblocks, int stage,
                          double dt, hipStream_t st, const ProjDesc<double> *descs00_dev, int n00) {
    if (nblocks + n00 <= 0) return;
    const int grid = nblocks + n00;
    const size_t smem = 256 * (sizeof(cplx<double>) + sizeof(double) + sizeof(cplx<double>));
    if (CL == 2)
        k_stage_chunked<T, 2><<<grid, 256, smem, st>>>(descs_dev, lt, stage, dt, descs00_dev, nblocks);
    else if (CL == 4)
        k_stage_chunked<T, 4><<<grid, 256, smem, st>>>(descs_dev, lt, stage, dt, descs00_dev, nblocks);
    else if (CL == 8)
        k_stage_chunked<T, 8><<<grid, 256, smem, st>>>(descs_dev, lt, stage, dt, descs00_dev, nblocks);
    else
        k_stage_chunked<T, 16><<<grid, 256, smem, st>>>(descs_dev, lt, stage, dt, descs00_dev, nblocks);
}

// one thread per column (very tall columns, n3 > 16·16)
template <typename T>
void launch_stage_sequential(const ProjDesc<T> &D, int stage, double dt, hipStream_t st) {
    k_stage_project<T><<<(D.nk + 63) / 64, 64, 0, st>>>(D, stage, dt, D.nk);
}

template <typename T>
void launch_check_finite(const void *a, size_t n, int *flag, hipStream_t st) {
    int nb = int(std::min<size_t>((n + 255) / 256, 4096));
    if (nb > 0) k_check_finite<T><<<nb, 256, 0, st>>>(reinterpret_cast<const cplx<T> *>(a), n, flag);
}

template <typename T>
void launch_fill(void *a, size_t n, hipStream_t st) {
    int nb = int(std::min<size_t>((n + 255) / 256, 4096));
    if (nb > 0) k_fill<T><<<nb, 256, 0, st>>>(reinterpret_cast<cplx<T> *>(a), n);
}

#define BLD_INSTANTIATE(T)                                                                                         \
    template void launch_transform<T>(int, const Space<T> *, const LaunchTable &, int, size_t, int, hipStream_t); \
    template void launch_physics<T>(int, const Space<T> *, const LaunchTable &, int, hipStream_t);                \
    template void launch_linear<T>(const LinDesc<T> *, const LaunchTable &, int, int, int *, hipStream_t);               \
    template void launch_constant_mean<T>(const CMeanDesc<T> &, int, hipStream_t);                                \
    template void launch_stage_chunked<T>(int, const ProjDesc<T> *, const LaunchTable &, int, int, double, hipStream_t, \
                                          const ProjDesc<double> *, int);                                        \
    template void launch_stage_sequential<T>(const ProjDesc<T> &, int, double, hipStream_t);                      \
    template void launch_check_finite<T>(const void *, size_t, int *, hipStream_t);                               \
    template void launch_fill<T>(void *, size_t, hipStream_t);                                                    \
    template hipError_t allow_large_lds<T>();

BLD_INSTANTIATE(float)
BLD_INSTANTIATE(double)

#else
// parts 1..16: one transform kind of one type each (kind = (part − 1) % 8 in
// PassKind order, float for parts 1..8, double for 9..16)
#define BLD_PART_T std::conditional_t<(BLD_KERNEL_PART <= 8), float, double>
BLD_XF_INST(BLD_PART_T, ((BLD_KERNEL_PART - 1) % 8))
#endif

}  // namespace bld
