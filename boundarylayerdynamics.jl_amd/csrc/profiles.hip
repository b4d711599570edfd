// Mean profiles on the device (MeanProfiles, logging.jl:310-525): horizontal
// means of the state and of the physical-space terms at the rate evaluation
// that ends each step, integrated in time with the trapezoidal rule
// (prepare_samples!/process_log!, logging.jl:370-451) without host copies.
//
// Samples (logging.jl:380-413, hmean): vel1, vel2 — real part of the (0,0)
// mode at the C-nodes; vel3 — the same at the interior I-nodes, 0 at the two
// boundary entries; adv1, adv2, adv3 — the plane means of the advection terms
// that enter the rates (momentum_advection.jl:58-71), adv3 with 0 boundary
// entries; sgs11, sgs12, sgs22, sgs33 — plane means of the SGS stresses at the
// C-nodes; sgs13, sgs23 — at the I-nodes with the wall-stress means as
// boundary entries (static_smagorinsky.jl:145-186, stress_bc).
#include "kernels.hpp"

namespace bld {

template <typename T>
__device__ __forceinline__ double plane_mean(const T *a, size_t P, double *red) {
    double s = 0.0;
    for (size_t x = threadIdx.x; x < P; x += blockDim.x) s += double(a[x]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    const double r = red[0] / double(P);
    __syncthreads();
    return r;
}

// one block per (profile, slot); slots 0..n3 (C profiles use 0..n3-1)
template <typename T>
__global__ void __launch_bounds__(256) k_profile_sample(ProfDesc<T> D, double weight) {
    __shared__ double red[256];
    const int kind = blockIdx.x / (D.n3 + 1), j = blockIdx.x - kind * (D.n3 + 1);
    if (!(D.mask & (1u << kind))) return;
    const int n3 = D.n3;
    const bool inode = kind == PROF_VEL3 || kind == PROF_ADV3 || kind == PROF_SGS13 || kind == PROF_SGS23;
    if (!inode && j == n3) return;
    const Space<T> *A = D.adv, *S = D.sgs;
    double s = 0.0;
    switch (kind) {
        case PROF_VEL1:
        case PROF_VEL2: {
            const int f = kind - PROF_VEL1;
            s = D.u00[f] ? D.u00[f][j].x : double(D.u[f][size_t(j) * D.nk].x);
            break;
        }
        case PROF_VEL3:
            if (j > 0 && j < n3) s = D.u00[2] ? D.u00[2][j - 1].x : double(D.u[2][size_t(j - 1) * D.nk].x);
            break;
        case PROF_ADV1:
        case PROF_ADV2:
            if (A) {
                const size_t P = size_t(A->N1) * A->N2;
                s = plane_mean(A->src_phys[kind - PROF_ADV1] + size_t(j) * P, P, red);
            }
            break;
        case PROF_ADV3:
            if (A && j > 0 && j < n3) {
                const size_t P = size_t(A->N1) * A->N2;
                s = plane_mean(A->src_phys[2] + size_t(j - 1) * P, P, red);
            }
            break;
        case PROF_SGS11:
        case PROF_SGS22:
        case PROF_SGS12:
            if (S) {  // the stress kernels store −τ as the rate source
                const int src = kind == PROF_SGS11 ? 3 : (kind == PROF_SGS22 ? 4 : 5);
                const size_t P = size_t(S->N1) * S->N2;
                s = -plane_mean(S->src_phys[src] + size_t(j) * P, P, red);
            }
            break;
        case PROF_SGS33:
            if (S) {
                const size_t P = size_t(S->N1) * S->N2;
                double acc = 0.0;
                if (S->mlp_w) {  // learned closure: τ33 stored at the C-nodes
                    acc = plane_mean(S->tau[2] + size_t(j) * P, P, red);
                } else {         // τ33 = −2 ν_T ∂w/∂x3 at the C-node
                    const double a = S->alpha_c[j];
                    double t = 0.0;
                    for (size_t x = threadIdx.x; x < P; x += blockDim.x) {
                        const double wb = j > 0 ? double(S->in_phys[2][size_t(j - 1) * P + x]) : S->wbc[0];
                        const double wa = j < n3 - 1 ? double(S->in_phys[2][size_t(j) * P + x]) : S->wbc[1];
                        t += -2.0 * double(S->visc_c[size_t(j) * P + x]) * a * (wa - wb);
                    }
                    red[threadIdx.x] = t;
                    __syncthreads();
                    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
                        if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
                        __syncthreads();
                    }
                    acc = red[0] / double(P);
                }
                s = acc;
            }
            break;
        case PROF_SGS13:
        case PROF_SGS23:
            if (S) {
                const int comp = kind == PROF_SGS13 ? 0 : 1;
                const size_t P = size_t(S->N1) * S->N2;
                if (j > 0 && j < n3) {
                    s = -plane_mean(S->src_phys[6 + comp] + size_t(j - 1) * P, P, red);
                } else {
                    const int side = j == 0 ? 0 : 1, l = j == 0 ? 0 : n3 - 1;
                    if (S->wall_rough[side]) {  // rough-wall stress (stress_bc)
                        double t = 0.0;
                        for (size_t x = threadIdx.x; x < P; x += blockDim.x) {
                            const double u = S->in_phys[0][size_t(l) * P + x], v = S->in_phys[1][size_t(l) * P + x];
                            t += S->fstress[side] * (comp == 0 ? u : v) * sqrt(u * u + v * v);
                        }
                        red[threadIdx.x] = t;
                        __syncthreads();
                        for (int h = blockDim.x / 2; h > 0; h >>= 1) {
                            if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
                            __syncthreads();
                        }
                        s = red[0] / double(P);
                    }
                }
            }
            break;
    }
    if (threadIdx.x == 0) {
        const int o = kind * (n3 + 1) + j;
        D.acc[o] += weight * (D.prev[o] + s);
        D.prev[o] = s;
    }
}

template <typename T>
void launch_profile_sample(const ProfDesc<T> &D, double weight, hipStream_t st) {
    k_profile_sample<T><<<PROF_COUNT * (D.n3 + 1), 256, 0, st>>>(D, weight);
}

template void launch_profile_sample<float>(const ProfDesc<float> &, double, hipStream_t);
template void launch_profile_sample<double>(const ProfDesc<double> &, double, hipStream_t);

}  // namespace bld
