// Spectral values of the c2r terms and block-uniform integer division, shared
// by the LDS and the register transform kernels.
#pragma once

#include "common.hpp"

namespace bld {

template <typename T>
__device__ __forceinline__ cplx<T> czero() {
    return mk(T(0), T(0));
}

// exp(±iπ(k1/N1 + k2/N2)): moves physical samples by half a cell in x1 and x2
// (shift!/unshift!, physical_space.jl:322-336)
template <typename T>
__device__ __forceinline__ cplx<T> centre_shift(const Space<T> &S, int k1, int i2, double sign) {
    const int k2 = i2 <= (S.K2 - 1) / 2 ? i2 : i2 - S.K2;
    double sn, cs;
    sincospi(sign * (double(k1) / S.N1 + double(k2) / S.N2), &sn, &cs);
    return mk(T(cs), T(sn));
}

// Spectral value of a c2r term at (k1, i2) of `layer` (derivatives.jl:404-418,566-568).
template <typename T>
__device__ __forceinline__ cplx<T> term_value(const Space<T> &S, int term, int recipe, int sel, int layer, int k1,
                                              int i2) {
    const size_t P = size_t(S.K1) * S.K2;
    const size_t idx = size_t(layer) * P + k1 + size_t(S.K1) * i2;
    const cplx<T> *u = S.state[sel][0], *v = S.state[sel][1], *w = S.state[sel][2];
    // derivative factors are loaded only by the recipes that use them
    auto d1 = [&] { return T(gload(S.D1 + k1)); };
    auto d2 = [&] { return T(gload(S.D2 + i2)); };
    switch (recipe) {
        case R_U: return gload(u + idx);
        case R_V: return gload(v + idx);
        case R_W: return gload(w + idx);
        case R_D1U: return imul(d1(), gload(u + idx));
        case R_D2U: return imul(d2(), gload(u + idx));
        case R_D1V: return imul(d1(), gload(v + idx));
        case R_D2V: return imul(d2(), gload(v + idx));
        case R_D1W: return imul(d1(), gload(w + idx));
        case R_D2W: return imul(d2(), gload(w + idx));
        case R_VORT1: {  // D2·w - (-α·v¯ + α·v⁺)
            const T a = T(gload(S.alpha_i + layer + 1));
            cplx<T> dz = (-a) * gload(v + idx) + a * gload(v + idx + P);
            return imul(d2(), gload(w + idx)) - dz;
        }
        case R_VORT2: {  // (-α·u¯ + α·u⁺) - D1·w
            const T a = T(gload(S.alpha_i + layer + 1));
            cplx<T> dz = (-a) * gload(u + idx) + a * gload(u + idx + P);
            return dz - imul(d1(), gload(w + idx));
        }
        case R_VORT3: return imul(d1(), gload(v + idx)) - imul(d2(), gload(u + idx));
        default: {
            const cplx<T> val = gload(S.in_field[term] + idx);
            return S.xfer_centered ? cmul(val, centre_shift<T>(S, k1, i2, 1.0)) : val;
        }
    }
}

// Division by a block-uniform runtime divisor d (n·d < 2^32): q = ⌊n·m/2^32⌋
// with m = ⌈2^32/d⌉.
struct Div {
    uint32_t d, m;
    __device__ explicit Div(int dd) : d(uint32_t(dd)), m(uint32_t(0xFFFFFFFFu / uint32_t(dd)) + 1u) {}
    __device__ __forceinline__ int q(int n) const { return d == 1 ? n : int(__umulhi(uint32_t(n), m)); }
};

}  // namespace bld
