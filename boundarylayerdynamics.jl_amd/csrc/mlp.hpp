// Learned SGS closure network: 16 (10 used) → 64 → 64 → 16 (6 used), GELU,
// bfloat16 operands on MFMA with float32 accumulation: the inference network
// runs its hidden layers on v_mfma_f32_16x16x32_bf16 (gfx950), the first layer
// (16 padded features) and the training kernels on v_mfma_f32_16x16x16_bf16.
//
// MFMA 16x16x16 operand layout (lane l, r = l % 16, g = l / 16, element j = 0..3):
//   A[m][k]: lane holds A[r][4g + j]     B[k][n]: lane holds B[4g + j][r]
//   D[m][n]: lane holds D[4g + j][r]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bld {
namespace mlp {

constexpr int NIN = 10, NH = 64, NOUT = 6, PAD = 16;
// parameter offsets (float32 master copy, row-major matrices [in][out])
constexpr int OW1 = 0, OW2 = OW1 + PAD * NH, OW3 = OW2 + NH * NH, OB1 = OW3 + NH * PAD, OB2 = OB1 + NH,
              OB3 = OB2 + NH, NPARAM = OB3 + PAD;
// bfloat16 weight copy: W1, W2, W3 in the same row-major layout
constexpr int NWB = OB1;

typedef float floatx4 __attribute__((__vector_size__(4 * sizeof(float))));

// round to nearest even, as bits (one v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ unsigned short to_bf16(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
typedef float floatv2 __attribute__((ext_vector_type(2)));
typedef short bf16x4 __attribute__((__vector_size__(4 * sizeof(short))));
// four values rounded to nearest even: two v_cvt_pk_bf16_f32
__device__ __forceinline__ bf16x4 to_bf16x4(float a, float b, float c, float d) {
    const bf16v2 lo = __builtin_convertvector((floatv2){a, b}, bf16v2);
    const bf16v2 hi = __builtin_convertvector((floatv2){c, d}, bf16v2);
    typedef unsigned u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned))));
    return __builtin_bit_cast(bf16x4, (u32x2){__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)});
}
__device__ __forceinline__ float from_bf16(unsigned short b) { return __uint_as_float(uint32_t(b) << 16); }

// GELU, tanh form: z/2·(1 + tanh(u)), u = √(2/π)(z + 0.044715 z³).  With
// 1 + tanh(u) = 2σ(2u), σ the logistic function, one exp and one reciprocal
// give it (exp overflow for very negative z gives σ = 0, as it should).
// The exponential is taken as a power of two, 2u·log2(e) = z·(k + k·0.044715·z²)
// with k = 2·sqrt(2/π)·log2(e).
__device__ __forceinline__ float gelu_sigma(float z) {
    constexpr float k = 2.0f * 0.7978845608028654f * 1.4426950408889634f;
    const float m = __builtin_fmaf(-k * 0.044715f, z * z, -k);  // -(2u log2 e) / z
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * m));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sigma(z); }
// d/dz: σ + z·σ(1−σ)·2√(2/π)(1 + 3·0.044715 z²)
__device__ __forceinline__ float gelu_grad(float z) {
    const float s2 = 2.0f * 0.7978845608028654f;
    const float sg = gelu_sigma(z);
    return sg + z * sg * (1.0f - sg) * s2 * (1.0f + 3.0f * 0.044715f * z * z);
}

__device__ __forceinline__ floatx4 mfma(bf16x4 a, bf16x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// gfx950 v_mfma_f32_16x16x32_bf16: A[m][k] lane holds A[r][8g + j], B[k][n]
// lane holds B[8g + j][r] (j = 0..7), D as above
typedef short bf16x8 __attribute__((__vector_size__(8 * sizeof(short))));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float floatv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ floatx4 mfma32(bf16x8 a, bf16x8 b, floatx4 c) {
    const floatv4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b),
                                                              __builtin_bit_cast(floatv4, c), 0, 0, 0);
    return __builtin_bit_cast(floatx4, d);
}
// Hidden-neuron order of the K = 32 blocks of the register-resident network:
// element j of lane group g in K block c is neuron 16·(2c + j/4) + 4g + j%4,
// i.e. the values the lane already holds in the D layouts of output blocks
// 2c and 2c+1, so a layer's output feeds the next layer's 16x16x32 MFMA
// without moving across lanes.
__device__ __forceinline__ constexpr int k32_neuron(int c, int g, int j) { return 16 * (2 * c + j / 4) + 4 * g + j % 4; }

// Forward pass of one 16-point tile by one wave.  `xa` is the A operand of the
// input (lane holds x[point r][feature 4g+j] as bf16); weights in LDS (bf16,
// row-major), biases in LDS (float).  `tile` is this wave's 16x64 bf16 LDS
// scratch.  On return y holds D-layout outputs (lane: y[point 4g+j][out r]),
// z1/z2 the D-layout pre-activations and h1d/h2d the D-layout bf16 activations.
struct Fwd {
    floatx4 z1[4], z2[4], y;
    bf16x4 h1d[4], h2d[4];
};

__device__ __forceinline__ bf16x4 b_frag(const unsigned short *W, int ld, int k0, int n0, int lane) {
    // B operand of rows k0.., columns n0..: lane holds W[k0 + 4g + j][n0 + r]
    const int r = lane & 15, g = lane >> 4;
    bf16x4 b;
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = short(W[(k0 + 4 * g + j) * ld + n0 + r]);
    return b;
}

__device__ __forceinline__ bf16x4 bt_frag(const unsigned short *W, int ld, int k0, int n0, int lane) {
    // B operand of the transpose: lane holds Wᵀ[k0 + 4g + j][n0 + r] = W[n0 + r][k0 + 4g + j]
    const int r = lane & 15, g = lane >> 4;
    bf16x4 b;
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = short(W[(n0 + r) * ld + k0 + 4 * g + j]);
    return b;
}

// D-layout tile (16 x 16 block at column c0 of a 16-row LDS tile of width ld)
// stored as bf16, and read back as the A operand (lane: T[r][k0 + 4g + j]).
__device__ __forceinline__ void store_d(unsigned short *T, int ld, int c0, const floatx4 &v, int lane) {
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) T[(4 * g + j) * ld + c0 + r] = to_bf16(v[j]);
}
__device__ __forceinline__ bf16x4 a_frag(const unsigned short *T, int ld, int k0, int lane) {
    const int r = lane & 15, g = lane >> 4;
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = short(T[r * ld + k0 + 4 * g + j]);
    return a;
}

__device__ __forceinline__ void forward_tile(const bf16x4 &xa, const unsigned short *W, const float *bias,
                                             unsigned short *tile, int lane, Fwd &o) {
    const int r = lane & 15, g = lane >> 4;
    const unsigned short *W1 = W + OW1, *W2 = W + OW2, *W3 = W + OW3;
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        floatx4 acc = mfma(xa, b_frag(W1, NH, 0, 16 * t, lane), zero);
        floatx4 h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc[j] += bias[OB1 - OB1 + 16 * t + r];
            h[j] = gelu(acc[j]);
            o.h1d[t][j] = short(to_bf16(h[j]));
        }
        o.z1[t] = acc;
        store_d(tile, NH, 16 * t, h, lane);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    bf16x4 a[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = a_frag(tile, NH, 16 * c, lane);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        floatx4 acc = zero;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc = mfma(a[c], b_frag(W2, NH, 16 * c, 16 * t, lane), acc);
        floatx4 h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc[j] += bias[NH + 16 * t + r];
            h[j] = gelu(acc[j]);
            o.h2d[t][j] = short(to_bf16(h[j]));
        }
        o.z2[t] = acc;
        store_d(tile, NH, 16 * t, h, lane);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = a_frag(tile, NH, 16 * c, lane);
    __builtin_amdgcn_wave_barrier();
    floatx4 acc = zero;
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = mfma(a[c], b_frag(W3, PAD, 16 * c, 0, lane), acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] += bias[2 * NH + r];
    o.y = acc;
    (void)g;
}

// Register-resident forward pass (inference).  Each layer is computed
// transposed, Zᵀ = Wᵀ·Xᵀ, with the 16 points of a tile along N: the A operand
// is a 16x16 block of Wᵀ (held in registers for the whole kernel), the B
// operand the activations, and the D layout of a layer's output (lane holds
// Z[point r][neuron 16t + 4g + j]) is exactly the B layout the next layer
// needs, so activations never leave the registers.
struct NetRegs {
    bf16x4 w1[4];                     // Wᵀ1 blocks [out block] (K = the 16 padded features)
    bf16x8 w2[4][2], w3[2];           // Wᵀ2 [out block][K block of 32 neurons], Wᵀ3 [K block]
    float b1[4][4], b2[4][4], b3[4];  // biases of the D rows 16t + 4g + j
};

// W: bfloat16 weights (W1, W2, W3 row-major [in][out]); bias: b1, b2, b3 (float)
__device__ __forceinline__ void load_net(const unsigned short *__restrict__ W, const float *__restrict__ bias,
                                         int lane, NetRegs &n) {
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            n.w1[t][j] = short(W[OW1 + (4 * g + j) * NH + 16 * t + r]);
            n.b1[t][j] = bias[16 * t + 4 * g + j];
            n.b2[t][j] = bias[NH + 16 * t + 4 * g + j];
        }
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int j = 0; j < 8; ++j) n.w2[t][c][j] = short(W[OW2 + k32_neuron(c, g, j) * NH + 16 * t + r]);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int j = 0; j < 8; ++j) n.w3[c][j] = short(W[OW3 + k32_neuron(c, g, j) * PAD + r]);
#pragma unroll
    for (int j = 0; j < 4; ++j) n.b3[j] = bias[2 * NH + 4 * g + j];
}

// xb: lane holds x[point r][feature 4g + j] (bf16); returns y with lane holding
// y[point r][output 4g + j]
__device__ __forceinline__ floatx4 forward_regs(const NetRegs &n, const bf16x4 &xb) {
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
    bf16x4 h1[4], h2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        floatx4 acc = mfma(n.w1[t], xb, zero);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = gelu(acc[j] + n.b1[t][j]);
        h1[t] = to_bf16x4(acc[0], acc[1], acc[2], acc[3]);
    }
    // K blocks of 32 neurons from the D-layout registers of two output blocks
    auto kblock = [](const bf16x4 &lo, const bf16x4 &hi) {
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            b[j] = lo[j];
            b[4 + j] = hi[j];
        }
        return b;
    };
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        floatx4 acc = zero;
#pragma unroll
        for (int c = 0; c < 2; ++c) acc = mfma32(n.w2[t][c], kblock(h1[2 * c], h1[2 * c + 1]), acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = gelu(acc[j] + n.b2[t][j]);
        h2[t] = to_bf16x4(acc[0], acc[1], acc[2], acc[3]);
    }
    floatx4 y = zero;
#pragma unroll
    for (int c = 0; c < 2; ++c) y = mfma32(n.w3[c], kblock(h2[2 * c], h2[2 * c + 1]), y);
#pragma unroll
    for (int j = 0; j < 4; ++j) y[j] += n.b3[j];
    return y;
}

}  // namespace mlp
}  // namespace bld
