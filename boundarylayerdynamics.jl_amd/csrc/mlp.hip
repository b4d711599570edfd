// Learned SGS closure: bfloat16 MFMA forward, forward+backward (per-wave
// partial gradients) and a reduction + Adam kernel; C ABI (bld_mlp_*).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bld.h"
#include "mlp.hpp"

namespace bld {
namespace mlp {

constexpr int WAVES = 4;  // waves per block

// weights (bf16) and biases (float) into LDS
__device__ __forceinline__ void load_params(const unsigned short *__restrict__ wb, const float *__restrict__ p,
                                            unsigned short *sW, float *sB) {
    for (int i = threadIdx.x; i < NWB; i += blockDim.x) sW[i] = wb[i];
    for (int i = threadIdx.x; i < 2 * NH + PAD; i += blockDim.x) sB[i] = p[OB1 + i];
    __syncthreads();
}

// A operand of a 16-point input tile: lane holds x[point r][feature 4g+j]
__device__ __forceinline__ bf16x4 load_x(const float *__restrict__ X, int64_t p0, int64_t n, int lane) {
    const int r = lane & 15, g = lane >> 4;
    bf16x4 a;
    const int64_t p = p0 + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = short(p < n ? to_bf16(X[p * PAD + 4 * g + j]) : 0);
    return a;
}

// inference: Y[n][16] = net(X[n][16])
__global__ void __launch_bounds__(256) k_forward(const unsigned short *__restrict__ wb, const float *__restrict__ p,
                                                 const float *__restrict__ X, float *__restrict__ Y, int64_t n) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    NetRegs net;
    load_net(wb, p + OB1, lane, net);
    const int64_t ntiles = (n + 15) / 16;
    for (int64_t t = int64_t(blockIdx.x) * WAVES + wave; t < ntiles; t += int64_t(gridDim.x) * WAVES) {
        const bf16x4 xa = load_x(X, t * 16, n, lane);
        const floatx4 y = forward_regs(net, xa);
        const int64_t pt = t * 16 + r;
        if (pt < n) *reinterpret_cast<floatx4 *>(Y + pt * PAD + 4 * g) = y;
    }
}

// forward + backward of the mean squared error over the 6 outputs; the waves
// of a block sum their gradients in LDS and the block writes one partial
// gradient (NPARAM floats) and one partial loss
__global__ void __launch_bounds__(256) k_gradient(const unsigned short *__restrict__ wb, const float *__restrict__ p,
                                                  const float *__restrict__ X, const float *__restrict__ Tg,
                                                  int64_t n, float scale, float *__restrict__ part,
                                                  float *__restrict__ lpart) {
    __shared__ unsigned short sW[NWB];
    __shared__ float sB[2 * NH + PAD];
    __shared__ unsigned short sT[WAVES][16 * NH];
    __shared__ float sAcc[NPARAM + 1];  // block sums of the gradient and (last entry) the loss
    for (int i = threadIdx.x; i <= NPARAM; i += blockDim.x) sAcc[i] = 0.f;
    load_params(wb, p, sW, sB);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
    floatx4 gW1[4], gW2[4][4], gW3[4];
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0, lsum = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gW1[a] = zero;
        gW3[a] = zero;
#pragma unroll
        for (int b = 0; b < 4; ++b) gW2[a][b] = zero;
    }
    unsigned short *tile = sT[wave];
    const int64_t ntiles = (n + 15) / 16;
    for (int64_t t = int64_t(blockIdx.x) * WAVES + wave; t < ntiles; t += int64_t(gridDim.x) * WAVES) {
        const int64_t p0 = t * 16;
        const bf16x4 xa = load_x(X, p0, n, lane);
        Fwd o;
        forward_tile(xa, sW, sB, tile, lane, o);
        // dy (D layout: lane holds point 4g+j, output r)
        floatx4 dy;
        bf16x4 dyb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t pt = p0 + 4 * g + j;
            float d = 0.f;
            if (pt < n && r < NOUT) {
                const float e = o.y[j] - Tg[pt * PAD + r];
                lsum += e * e;
                d = scale * e;
            }
            dy[j] = d;
            dyb[j] = short(to_bf16(d));
            gb3 += d;
        }
        // x in D-layout arrangement (lane: x[point 4g+j][feature r]) for dW1 = xᵀ·dz1
        bf16x4 xd;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t pt = p0 + 4 * g + j;
            xd[j] = short(pt < n ? to_bf16(X[pt * PAD + r]) : 0);
        }
        // dW3 += h2ᵀ·dy
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) gW3[mt] = mfma(o.h2d[mt], dyb, gW3[mt]);
        // dh2 = dy·W3ᵀ (A operand of dy needs the point-major arrangement)
        store_d(tile, NH, 0, dy, lane);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        const bf16x4 dya = a_frag(tile, NH, 0, lane);
        __builtin_amdgcn_wave_barrier();
        floatx4 dz2[4];
        bf16x4 dz2b[4];
#pragma unroll
        for (int t2 = 0; t2 < 4; ++t2) {
            floatx4 acc = mfma(dya, bt_frag(sW + OW3, PAD, 0, 16 * t2, lane), zero);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[j] *= gelu_grad(o.z2[t2][j]);
                dz2b[t2][j] = short(to_bf16(acc[j]));
                gb2[t2] += acc[j];
            }
            dz2[t2] = acc;
        }
        // dW2 += h1ᵀ·dz2
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) gW2[mt][nt] = mfma(o.h1d[mt], dz2b[nt], gW2[mt][nt]);
        // dh1 = dz2·W2ᵀ
#pragma unroll
        for (int t2 = 0; t2 < 4; ++t2) store_d(tile, NH, 16 * t2, dz2[t2], lane);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        bf16x4 a[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = a_frag(tile, NH, 16 * c, lane);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t1 = 0; t1 < 4; ++t1) {
            floatx4 acc = zero;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc = mfma(a[c], bt_frag(sW + OW2, NH, 16 * c, 16 * t1, lane), acc);
            bf16x4 dz1b;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[j] *= gelu_grad(o.z1[t1][j]);
                dz1b[j] = short(to_bf16(acc[j]));
                gb1[t1] += acc[j];
            }
            // dW1 += xᵀ·dz1
            gW1[t1] = mfma(xd, dz1b, gW1[t1]);
        }
    }
    // add this wave's gradient into the block sums (D layouts → parameter offsets)
    float *out = sAcc;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) atomicAdd(&out[OW1 + (4 * g + j) * NH + 16 * nt + r], gW1[nt][j]);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) atomicAdd(&out[OW2 + (16 * mt + 4 * g + j) * NH + 16 * nt + r], gW2[mt][nt][j]);
            atomicAdd(&out[OW3 + (16 * mt + 4 * g + j) * PAD + r], gW3[mt][j]);
        }
    }
    // bias gradients: sum the four lane groups
#pragma unroll
    for (int t2 = 0; t2 < 4; ++t2) {
        gb1[t2] += __shfl_xor(gb1[t2], 16);
        gb1[t2] += __shfl_xor(gb1[t2], 32);
        gb2[t2] += __shfl_xor(gb2[t2], 16);
        gb2[t2] += __shfl_xor(gb2[t2], 32);
    }
    gb3 += __shfl_xor(gb3, 16);
    gb3 += __shfl_xor(gb3, 32);
    for (int off = 1; off < 64; off <<= 1) lsum += __shfl_xor(lsum, off);
    if (g == 0) {
#pragma unroll
        for (int t2 = 0; t2 < 4; ++t2) {
            atomicAdd(&out[OB1 + 16 * t2 + r], gb1[t2]);
            atomicAdd(&out[OB2 + 16 * t2 + r], gb2[t2]);
        }
        atomicAdd(&out[OB3 + r], gb3);
    }
    if (lane == 0) atomicAdd(&sAcc[NPARAM], lsum);
    __syncthreads();
    float *dst = part + int64_t(blockIdx.x) * NPARAM;
    for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) dst[i] = sAcc[i];
    if (threadIdx.x == 0) lpart[blockIdx.x] = sAcc[NPARAM];
}

// sum of the per-block partials; optional Adam step on the float32 master
// parameters (Kingma & Ba 2015) and refresh of the bfloat16 weight copy
__global__ void __launch_bounds__(256) k_reduce_adam(const float *__restrict__ part, int nw, float *__restrict__ grad,
                                                     float *__restrict__ p, float *__restrict__ m,
                                                     float *__restrict__ v, unsigned short *__restrict__ wb, int adam,
                                                     float lr, float beta1, float beta2, float eps, float bc1,
                                                     float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    float s = 0.f;
    for (int w = 0; w < nw; ++w) s += part[int64_t(w) * NPARAM + i];
    grad[i] = s;
    if (!adam) return;
    const float mi = beta1 * m[i] + (1.f - beta1) * s;
    const float vi = beta2 * v[i] + (1.f - beta2) * s * s;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    if (i < NWB) wb[i] = to_bf16(pi);
}

__global__ void k_refresh_bf16(const float *__restrict__ p, unsigned short *__restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < NWB) wb[i] = to_bf16(p[i]);
}

// one block of 256 threads: out = Σ a[0..n)
__global__ void __launch_bounds__(256) k_sum(const float *__restrict__ a, int n, float *__restrict__ out) {
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += a[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace mlp
}  // namespace bld

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace bld::mlp;

namespace {
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    bool alloc(size_t b) {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = b;
        return b == 0 || hipMalloc(&p, b) == hipSuccess;
    }
    template <typename U>
    U *as() const {
        return reinterpret_cast<U *>(p);
    }
};

constexpr int GRID_MAX = 1024;  // blocks of the gradient kernel (one partial gradient each)

// blocks for n points: about two 16-point tiles per wave, at most GRID_MAX
int gradient_grid(int64_t n) {
    const int64_t tiles = (n + 15) / 16;
    const int64_t blocks = (tiles + 2 * WAVES - 1) / (2 * WAVES);
    return int(std::max<int64_t>(1, std::min<int64_t>(blocks, GRID_MAX)));
}
}  // namespace

struct bld_mlp {
    Buf params, m, v, wb, grad, part, lpart, loss, X, T, Y;
    int64_t ndata = 0;
    int64_t step = 0;
    std::string err;
};

static thread_local std::string g_mlp_err;

#define MLP_CHECK(expr)                                                   \
    do {                                                                  \
        hipError_t _e = (expr);                                           \
        if (_e != hipSuccess) {                                           \
            h->err = std::string("HIP error: ") + hipGetErrorString(_e);  \
            return BLD_EHIP;                                              \
        }                                                                 \
    } while (0)

extern "C" {

int bld_mlp_create(const float *params, bld_mlp **out) {
    if (!out || !params) return BLD_EARG;
    *out = nullptr;
    std::unique_ptr<bld_mlp> h(new bld_mlp());
    if (!h->params.alloc(NPARAM * 4) || !h->m.alloc(NPARAM * 4) || !h->v.alloc(NPARAM * 4) ||
        !h->wb.alloc(NWB * 2) || !h->grad.alloc(NPARAM * 4) || !h->part.alloc(size_t(GRID_MAX) * NPARAM * 4) ||
        !h->lpart.alloc(GRID_MAX * 4) || !h->loss.alloc(4)) {
        g_mlp_err = "out of device memory";
        return BLD_ENOMEM;
    }
    if (hipMemcpy(h->params.p, params, NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->m.p, 0, NPARAM * 4) != hipSuccess || hipMemset(h->v.p, 0, NPARAM * 4) != hipSuccess) {
        g_mlp_err = "HIP error during setup";
        return BLD_EHIP;
    }
    k_refresh_bf16<<<(NWB + 255) / 256, 256>>>(h->params.as<float>(), h->wb.as<unsigned short>());
    if (hipDeviceSynchronize() != hipSuccess) {
        g_mlp_err = "HIP error during setup";
        return BLD_EHIP;
    }
    *out = h.release();
    return BLD_OK;
}

void bld_mlp_destroy(bld_mlp *h) { delete h; }

const char *bld_mlp_last_error(const bld_mlp *h) { return h ? h->err.c_str() : g_mlp_err.c_str(); }

int bld_mlp_nparams(void) { return NPARAM; }

int bld_mlp_get_params(bld_mlp *h, float *params) {
    if (!h || !params) return BLD_EARG;
    MLP_CHECK(hipMemcpy(params, h->params.p, NPARAM * 4, hipMemcpyDeviceToHost));
    return BLD_OK;
}

int bld_mlp_forward(bld_mlp *h, const float *x, int64_t n, float *y) {
    if (!h || !x || !y || n < 1) return BLD_EARG;
    Buf X, Y;
    if (!X.alloc(size_t(n) * PAD * 4) || !Y.alloc(size_t(n) * PAD * 4)) return BLD_ENOMEM;
    MLP_CHECK(hipMemcpy(X.p, x, size_t(n) * PAD * 4, hipMemcpyHostToDevice));
    MLP_CHECK(hipMemset(Y.p, 0, size_t(n) * PAD * 4));
    const int64_t tiles = (n + 15) / 16;
    const int grid = int(std::min<int64_t>((tiles + WAVES - 1) / WAVES, 4096));
    k_forward<<<grid, 256>>>(h->wb.as<unsigned short>(), h->params.as<float>(), X.as<float>(), Y.as<float>(), n);
    MLP_CHECK(hipGetLastError());
    MLP_CHECK(hipDeviceSynchronize());
    MLP_CHECK(hipMemcpy(y, Y.p, size_t(n) * PAD * 4, hipMemcpyDeviceToHost));
    return BLD_OK;
}

// gradient of the batch into the per-block partials, its loss sum into *loss_dst
static int run_gradient(bld_mlp *h, const float *X, const float *T, int64_t n, float *loss_dst, int *grid_out) {
    const float scale = 2.0f / float(double(n) * NOUT);
    const int grid = gradient_grid(n);
    k_gradient<<<grid, 256>>>(h->wb.as<unsigned short>(), h->params.as<float>(), X, T, n, scale, h->part.as<float>(),
                             h->lpart.as<float>());
    k_sum<<<1, 256>>>(h->lpart.as<float>(), grid, loss_dst);
    MLP_CHECK(hipGetLastError());
    *grid_out = grid;
    return BLD_OK;
}

int bld_mlp_gradient(bld_mlp *h, const float *x, const float *t, int64_t n, float *grad, double *loss) {
    if (!h || !x || !t || !grad || n < 1) return BLD_EARG;
    Buf X, T;
    if (!X.alloc(size_t(n) * PAD * 4) || !T.alloc(size_t(n) * PAD * 4)) return BLD_ENOMEM;
    MLP_CHECK(hipMemcpy(X.p, x, size_t(n) * PAD * 4, hipMemcpyHostToDevice));
    MLP_CHECK(hipMemcpy(T.p, t, size_t(n) * PAD * 4, hipMemcpyHostToDevice));
    int grid = 0;
    int rc = run_gradient(h, X.as<float>(), T.as<float>(), n, h->loss.as<float>(), &grid);
    if (rc) return rc;
    k_reduce_adam<<<(NPARAM + 255) / 256, 256>>>(h->part.as<float>(), grid, h->grad.as<float>(), nullptr,
                                                  nullptr, nullptr, nullptr, 0, 0, 0, 0, 0, 1, 1);
    MLP_CHECK(hipDeviceSynchronize());
    MLP_CHECK(hipMemcpy(grad, h->grad.p, NPARAM * 4, hipMemcpyDeviceToHost));
    float l = 0;
    MLP_CHECK(hipMemcpy(&l, h->loss.p, 4, hipMemcpyDeviceToHost));
    if (loss) *loss = double(l) / (double(n) * NOUT);
    return BLD_OK;
}

int bld_mlp_set_data(bld_mlp *h, const float *x, const float *t, int64_t n) {
    if (!h || !x || !t || n < 16) return BLD_EARG;
    if (!h->X.alloc(size_t(n) * PAD * 4) || !h->T.alloc(size_t(n) * PAD * 4)) return BLD_ENOMEM;
    MLP_CHECK(hipMemcpy(h->X.p, x, size_t(n) * PAD * 4, hipMemcpyHostToDevice));
    MLP_CHECK(hipMemcpy(h->T.p, t, size_t(n) * PAD * 4, hipMemcpyHostToDevice));
    h->ndata = n;
    return BLD_OK;
}

int bld_mlp_train(bld_mlp *h, int64_t steps, int64_t batch, float lr, double *losses) {
    if (!h || steps < 0 || batch < 16 || h->ndata < batch) return BLD_EARG;
    const int64_t nb = h->ndata / batch;
    Buf L;  // batch losses stay on the device until the end
    if (!L.alloc(size_t(std::max<int64_t>(steps, 1)) * 4)) return BLD_ENOMEM;
    for (int64_t s = 0; s < steps; ++s) {
        const int64_t off = (h->step % nb) * batch;
        int grid = 0;
        int rc = run_gradient(h, h->X.as<float>() + off * PAD, h->T.as<float>() + off * PAD, batch,
                              L.as<float>() + s, &grid);
        if (rc) return rc;
        h->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.f - std::pow(b1, float(h->step)), bc2 = 1.f - std::pow(b2, float(h->step));
        k_reduce_adam<<<(NPARAM + 255) / 256, 256>>>(h->part.as<float>(), grid, h->grad.as<float>(),
                                                      h->params.as<float>(), h->m.as<float>(), h->v.as<float>(),
                                                      h->wb.as<unsigned short>(), 1, lr, b1, b2, 1e-8f, bc1, bc2);
        MLP_CHECK(hipGetLastError());
    }
    MLP_CHECK(hipDeviceSynchronize());
    if (losses && steps > 0) {
        std::vector<float> l(steps);
        MLP_CHECK(hipMemcpy(l.data(), L.p, size_t(steps) * 4, hipMemcpyDeviceToHost));
        for (int64_t s = 0; s < steps; ++s) losses[s] = double(l[s]) / (double(batch) * NOUT);
    }
    return BLD_OK;
}

int bld_mlp_time(bld_mlp *h, int64_t n, int reps, double *ms_forward, double *ms_gradient) {
    if (!h || n < 16 || reps < 1) return BLD_EARG;
    Buf X, Y;
    if (!X.alloc(size_t(n) * PAD * 4) || !Y.alloc(size_t(n) * PAD * 4)) return BLD_ENOMEM;
    MLP_CHECK(hipMemset(X.p, 0, size_t(n) * PAD * 4));
    MLP_CHECK(hipMemset(Y.p, 0, size_t(n) * PAD * 4));
    const int64_t tiles = (n + 15) / 16;
    const int grid = int(std::min<int64_t>((tiles + WAVES - 1) / WAVES, 4096));
    hipEvent_t a, b;
    MLP_CHECK(hipEventCreate(&a));
    MLP_CHECK(hipEventCreate(&b));
    k_forward<<<grid, 256>>>(h->wb.as<unsigned short>(), h->params.as<float>(), X.as<float>(), Y.as<float>(), n);
    MLP_CHECK(hipEventRecord(a, nullptr));
    for (int r = 0; r < reps; ++r)
        k_forward<<<grid, 256>>>(h->wb.as<unsigned short>(), h->params.as<float>(), X.as<float>(), Y.as<float>(), n);
    MLP_CHECK(hipEventRecord(b, nullptr));
    MLP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    MLP_CHECK(hipEventElapsedTime(&ms, a, b));
    if (ms_forward) *ms_forward = ms / reps;
    MLP_CHECK(hipEventRecord(a, nullptr));
    for (int r = 0; r < reps; ++r) {
        int grid = 0;
        int rc = run_gradient(h, X.as<float>(), Y.as<float>(), n, h->loss.as<float>(), &grid);
        if (rc) return rc;
        k_reduce_adam<<<(NPARAM + 255) / 256, 256>>>(h->part.as<float>(), grid, h->grad.as<float>(), nullptr,
                                                      nullptr, nullptr, nullptr, 0, 0, 0, 0, 0, 1, 1);
    }
    MLP_CHECK(hipEventRecord(b, nullptr));
    MLP_CHECK(hipEventSynchronize(b));
    MLP_CHECK(hipEventElapsedTime(&ms, a, b));
    if (ms_gradient) *ms_gradient = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return BLD_OK;
}

}  // extern "C"
