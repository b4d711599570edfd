// Learned SGS closure on the engine's SGS spaces, and the filtered-DNS sampler
// that produces its training data (no counterpart in the reference; the
// closure takes the place of the stresses of static_smagorinsky.jl:129-249
// and keeps that model's wall treatment and its eddy viscosity as fallback).
//
// Features of a C-node (shared by the closure and the sampler).  From the SGS
// space's physical terms (u, v at C; w at I; ∂u/∂x1, ∂u/∂x2, ∂v/∂x1, ∂v/∂x2 at
// C; ∂w/∂x1, ∂w/∂x2 at I) the velocity-gradient tensor g_ij = ∂u_i/∂x_j is
//   g11, g12, g21, g22   spectral derivatives at the C-node,
//   g33 = α_c (w⁺ − w⁻)   from the bounding I-nodes or the wall values,
//   g13, g23              mean of the I-node differences α_i (u_{l+1} − u_l)
//                         on the I-layers below and above (one-sided at the
//                         first and last C-layer),
//   g31, g32              mean of the I-node values below and above, with 0 at
//                         a boundary (w is constant there).
// Network input x = (g/|g| in row-major order, wall feature of the layer);
// output y = τ^dev / (Δ²|g|²) for (11, 22, 33, 12, 13, 23).  The stress is
// τ = Δ²|g|² y with the trace removed; where y is not finite or |g| vanishes
// (or, with the dissipation guard, where −τ:S < 0), the Smagorinsky stress
// −2 l_c² |S| S of the same C-node is used.
#include "kernels.hpp"
#include "mlp.hpp"

namespace bld {

using namespace mlp;

template <typename T, typename A>
__device__ __forceinline__ void velocity_gradient(const Space<T> &S, size_t P, int l, size_t x, A g[9]) {
    auto f = [&](int t, int j) { return A(S.in_phys[t][size_t(j) * P + x]); };
    const int n3 = S.n3;
    A su = 0, sv = 0, w1 = 0, w2 = 0;
    int n = 0;
    if (l > 0) {
        const A a = A(S.alpha_i[l]);
        su += a * (f(0, l) - f(0, l - 1));
        sv += a * (f(1, l) - f(1, l - 1));
        w1 += f(7, l - 1);
        w2 += f(8, l - 1);
        ++n;
    }
    if (l < n3 - 1) {
        const A a = A(S.alpha_i[l + 1]);
        su += a * (f(0, l + 1) - f(0, l));
        sv += a * (f(1, l + 1) - f(1, l));
        w1 += f(7, l);
        w2 += f(8, l);
        ++n;
    }
    const A wb = l > 0 ? f(2, l - 1) : A(S.wbc[0]);
    const A wa = l < n3 - 1 ? f(2, l) : A(S.wbc[1]);
    g[0] = f(3, l);
    g[1] = f(5, l);
    g[2] = n ? su / A(n) : A(0);
    g[3] = f(6, l);
    g[4] = f(4, l);
    g[5] = n ? sv / A(n) : A(0);
    g[6] = w1 / A(2);
    g[7] = w2 / A(2);
    g[8] = A(S.alpha_c[l]) * (wa - wb);
}

// ---------------------------------------------------------------------------
// learned stresses at the C-nodes: one thread per point for the features and
// the stress, one wave per 64 points on MFMA for the network (weights and
// activations in registers, forward_regs)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_sgs_mlp(const Space<T> *__restrict__ spaces, LaunchTable lt) {
    __shared__ __align__(16) unsigned short sX[256 * PAD];  // vector reads of 4 bf16
    __shared__ __align__(16) float sY[256 * 8];             // vector writes of 4 floats
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
    const size_t P = size_t(S.N1) * S.N2, npts = P * S.n3;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, gq = lane >> 4;
    NetRegs net;  // weights and biases stay in registers for all tiles of the wave
    load_net(S.mlp_w, S.mlp_b, lane, net);
    for (int c = 0; c < kMlpChunks; ++c) {
        const size_t gid = (size_t(blockIdx.x - E.block_begin) * kMlpChunks + c) * 256 + threadIdx.x;
        const bool live = gid < npts;
        float g[9], gn2 = 0.f;
        int l = 0;
        if (live) {
            l = int(uint32_t(gid) / uint32_t(P));
            velocity_gradient<T, float>(S, P, l, gid - size_t(l) * P, g);
#pragma unroll
            for (int k = 0; k < 9; ++k) gn2 += g[k] * g[k];
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) g[k] = 0.f;
        }
        const float inv = gn2 > 0.f ? rsqrtf(gn2) : 0.f;
        unsigned short *xr = sX + threadIdx.x * PAD;
#pragma unroll
        for (int k = 0; k < 9; ++k) xr[k] = to_bf16(g[k] * inv);
        xr[9] = to_bf16(live ? float(S.mlp_xw[l]) : 0.f);
#pragma unroll
        for (int k = 10; k < PAD; ++k) xr[k] = 0;
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            // lane holds x[point r][feature 4g + j]: one 8-byte read per lane
            const bf16x4 xb = *reinterpret_cast<const bf16x4 *>(sX + (64 * wave + 16 * s + r) * PAD + 4 * gq);
            const floatx4 y = forward_regs(net, xb);  // lane: y[point r][output 4g + j]
            if (gq < 2) *reinterpret_cast<floatx4 *>(sY + (64 * wave + 16 * s + r) * 8 + 4 * gq) = y;
        }
        __syncthreads();
        if (live) {
            const float *y = sY + threadIdx.x * 8;
            const float sc = float(S.mlp_delta2) * gn2;
            float t[6];
            bool ok = gn2 > 0.f;
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                t[m] = sc * y[m];
                ok = ok && isfinite(t[m]);
            }
            const float s11 = g[0], s22 = g[4], s33 = g[8];
            const float s12 = 0.5f * (g[1] + g[3]), s13 = 0.5f * (g[2] + g[6]), s23 = 0.5f * (g[5] + g[7]);
            if (ok) {
                const float tr = (t[0] + t[1] + t[2]) / 3.f;
                t[0] -= tr;
                t[1] -= tr;
                t[2] -= tr;
                // dissipation guard: where the learned stress would backscatter (−τ:S < 0),
                // the Smagorinsky stress is used instead
                if (S.mlp_guard) {
                    const float eps = -(t[0] * s11 + t[1] * s22 + t[2] * s33 +
                                        2.f * (t[3] * s12 + t[4] * s13 + t[5] * s23));
                    ok = eps >= 0.f;
                }
            }
            if (!ok) {  // Smagorinsky stress of this C-node
                const float ss = s11 * s11 + s22 * s22 + s33 * s33 + 2.f * (s12 * s12 + s13 * s13 + s23 * s23);
                const float lcl = float(S.lc[l]);
                const float nu = lcl * lcl * sqrtf(2.f * ss);
                t[0] = -2.f * nu * s11;
                t[1] = -2.f * nu * s22;
                t[2] = -2.f * nu * s33;
                t[3] = -2.f * nu * s12;
                t[4] = -2.f * nu * s13;
                t[5] = -2.f * nu * s23;
            }
#pragma unroll
            for (int m = 0; m < 6; ++m) S.tau[m][gid] = T(t[m]);
        }
        __syncthreads();
    }
}

// Rates of the learned stresses, assembled as static_smagorinsky.jl:145-186
// does with the eddy-viscosity stresses: τ11, τ22, τ12 at C; τ33 differenced
// to the I-nodes; τ13, τ23 interpolated to the I-nodes and differenced back to
// C with the rough-wall stress (or zero at a free-slip boundary) at the walls.
template <typename T>
__global__ void __launch_bounds__(256) k_sgs_mlp_stress(const Space<T> *__restrict__ spaces, LaunchTable lt) {
    const LaunchEntry E = lt.e[find_entry(lt, blockIdx.x)];
    const Space<T> &S = spaces[E.space];
    const size_t P = size_t(S.N1) * S.N2;
    const size_t gid = size_t(blockIdx.x - E.block_begin) * blockDim.x + threadIdx.x;
    if (gid >= P * S.n3) return;
    const int l = int(uint32_t(gid) / uint32_t(P)), n3 = S.n3;
    const size_t x = gid - size_t(l) * P;
    S.src_phys[3][gid] = -S.tau[0][gid];
    S.src_phys[4][gid] = -S.tau[1][gid];
    S.src_phys[5][gid] = -S.tau[3][gid];
    if (l < n3 - 1) {
        const T a = T(S.alpha_i[l + 1]);
        S.src_phys[2][gid] = -(a * (S.tau[2][gid + P] - S.tau[2][gid]));
        S.src_phys[6][gid] = -(S.tau[4][gid] + S.tau[4][gid + P]) / T(2);
        S.src_phys[7][gid] = -(S.tau[5][gid] + S.tau[5][gid + P]) / T(2);
    }
    const T ac = T(S.alpha_c[l]);
    auto u = [&](int comp, int j) { return S.in_phys[comp][size_t(j) * P + x]; };
    for (int comp = 0; comp < 2; ++comp) {
        const T *tc = S.tau[4 + comp];
        T tb, ta;
        if (l > 0) {
            tb = (tc[gid - P] + tc[gid]) / T(2);
        } else if (S.wall_rough[0]) {
            const T u0 = u(0, 0), v0 = u(1, 0);
            tb = T(S.fstress[0]) * u(comp, 0) * sqrt(u0 * u0 + v0 * v0);
        } else {
            tb = T(0);
        }
        if (l < n3 - 1) {
            ta = (tc[gid] + tc[gid + P]) / T(2);
        } else if (S.wall_rough[1]) {
            const T u1 = u(0, n3 - 1), v1 = u(1, n3 - 1);
            ta = T(S.fstress[1]) * u(comp, n3 - 1) * sqrt(u1 * u1 + v1 * v1);
        } else {
            ta = T(0);
        }
        S.src_phys[comp][gid] = -(ac * (ta - tb));
    }
}

// ---------------------------------------------------------------------------
// filtered-DNS sampler: box filter of width F grid cells in x1 and x2 on each
// C-layer; one thread per coarse point.  X [n][16]: features as above from the
// filtered gradient; Y [n][16]: Y[0..5] = τ^dev/(Δ²|ḡ|²) with
// τ_ij = avg(u_i u_j) − avg(u_i) avg(u_j), Y[15] = Δ²|ḡ|²; w is interpolated to
// the C-nodes (boundary values at the walls).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_closure_sample(const Space<T> *__restrict__ sp, int F, double delta2,
                                                       const double *__restrict__ xw, float *__restrict__ X,
                                                       float *__restrict__ Y) {
    const Space<T> &S = *sp;
    const int N1c = S.N1 / F, N2c = S.N2 / F, n3 = S.n3;
    const size_t P = size_t(S.N1) * S.N2, Pc = size_t(N1c) * N2c;
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= Pc * n3) return;
    const int l = int(c / Pc);
    const size_t rem = c - size_t(l) * Pc;
    const int i2c = int(rem / N1c), i1c = int(rem - size_t(i2c) * N1c);
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, U[3] = {0, 0, 0}, UU[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < F; ++b)
        for (int a = 0; a < F; ++a) {
            const size_t x = size_t(F * i1c + a) + size_t(S.N1) * size_t(F * i2c + b);
            double g[9];
            velocity_gradient<T, double>(S, P, l, x, g);
#pragma unroll
            for (int k = 0; k < 9; ++k) G[k] += g[k];
            const double u = double(S.in_phys[0][size_t(l) * P + x]);
            const double v = double(S.in_phys[1][size_t(l) * P + x]);
            const double wb = l > 0 ? double(S.in_phys[2][size_t(l - 1) * P + x]) : S.wbc[0];
            const double wa = l < n3 - 1 ? double(S.in_phys[2][size_t(l) * P + x]) : S.wbc[1];
            const double w = 0.5 * (wb + wa);
            U[0] += u;
            U[1] += v;
            U[2] += w;
            UU[0] += u * u;
            UU[1] += v * v;
            UU[2] += w * w;
            UU[3] += u * v;
            UU[4] += u * w;
            UU[5] += v * w;
        }
    const double inv = 1.0 / double(F * F);
    double gn2 = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        G[k] *= inv;
        gn2 += G[k] * G[k];
    }
    for (int k = 0; k < 3; ++k) U[k] *= inv;
    for (int k = 0; k < 6; ++k) UU[k] *= inv;
    double t[6] = {UU[0] - U[0] * U[0], UU[1] - U[1] * U[1], UU[2] - U[2] * U[2],
                   UU[3] - U[0] * U[1], UU[4] - U[0] * U[2], UU[5] - U[1] * U[2]};
    const double tr = (t[0] + t[1] + t[2]) / 3.0;
    t[0] -= tr;
    t[1] -= tr;
    t[2] -= tr;
    const double sc = delta2 * gn2;
    const double ig = gn2 > 0 ? 1.0 / sqrt(gn2) : 0.0;
    float *xo = X + c * PAD, *yo = Y + c * PAD;
    for (int k = 0; k < 9; ++k) xo[k] = float(G[k] * ig);
    xo[9] = float(xw[l]);
    for (int k = 10; k < PAD; ++k) xo[k] = 0.f;
    for (int m = 0; m < 6; ++m) yo[m] = sc > 0 ? float(t[m] / sc) : 0.f;
    for (int k = 6; k < PAD - 1; ++k) yo[k] = 0.f;
    yo[PAD - 1] = float(sc);
}

template <typename T>
void launch_closure(int kind, const Space<T> *spaces_dev, const LaunchTable &lt, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return;
    if (kind == PHYS_SGS_MLP)
        k_sgs_mlp<T><<<nblocks, 256, 0, st>>>(spaces_dev, lt);
    else
        k_sgs_mlp_stress<T><<<nblocks, 256, 0, st>>>(spaces_dev, lt);
}

template <typename T>
void launch_closure_sample(const Space<T> *space_dev, int N1, int N2, int n3, int F, double delta2, const double *xw,
                           float *X, float *Y, hipStream_t st) {
    const size_t n = size_t(N1 / F) * (N2 / F) * n3;
    if (n == 0) return;
    k_closure_sample<T><<<unsigned((n + 255) / 256), 256, 0, st>>>(space_dev, F, delta2, xw, X, Y);
}

#define BLD_INSTANTIATE_CLOSURE(T)                                                                        \
    template void launch_closure<T>(int, const Space<T> *, const LaunchTable &, int, hipStream_t);       \
    template void launch_closure_sample<T>(const Space<T> *, int, int, int, int, double, const double *, \
                                           float *, float *, hipStream_t);

BLD_INSTANTIATE_CLOSURE(float)
BLD_INSTANTIATE_CLOSURE(double)

}  // namespace bld
