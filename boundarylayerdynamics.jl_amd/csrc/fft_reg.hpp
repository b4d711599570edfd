// Register-resident FFTs for the transform passes (gfx950).
//
// A length N = A·B transform of one column runs in two register steps with a
// single exchange through LDS:
//   n = B·n1 + n2, k = k1 + A·k2:
//   X[k1 + A·k2] = Σ_n2 w_B^{n2·k2} · ( w_N^{n2·k1} · Σ_n1 x[B·n1 + n2] · w_A^{n1·k1} )
// step 1: one task per n2 holds the A values x[B·n1 + n2] in registers, takes
//         their length-A DFT, multiplies by w_N^{n2·k1} and writes the A results
//         to LDS;
// step 2: one task per k1 reads the B values back, takes their length-B DFT and
//         holds X[k1 + A·k2], k2 = 0..B-1.
// The short DFTs are unrolled at compile time (mixed radix 4/3/2 with
// compile-time twiddles), so the only table lookups are the A·B inter-step
// twiddles w_N^{n2·k1} (from the host's extended-precision table).
// Sign convention as the LDS FFT: forward exp(-2πi nk/N), inverse exp(+2πi nk/N),
// both unnormalised.
#pragma once

#include <type_traits>

#include "fft.hpp"

namespace bld {

constexpr double kPiD = 3.14159265358979323846264338327950288;

constexpr double ct_sin_poly(double x) {
    double t = x, s = x;
    const double x2 = x * x;
    for (int n = 1; n < 12; ++n) {
        t *= -x2 / double((2 * n) * (2 * n + 1));
        s += t;
    }
    return s;
}
constexpr double ct_cos_poly(double x) {
    double t = 1.0, s = 1.0;
    const double x2 = x * x;
    for (int n = 1; n < 12; ++n) {
        t *= -x2 / double((2 * n - 1) * (2 * n));
        s += t;
    }
    return s;
}
// cos and sin of the angle 2π·m/R, reduced to |φ| <= π/4 by quadrant
constexpr double ct_cos_turn(int m, int R) {
    m = ((m % R) + R) % R;
    const int q = (8 * m + R) / (2 * R);
    const double phi = 2.0 * kPiD * double(4 * m - q * R) / double(4 * R);
    const double c = ct_cos_poly(phi), s = ct_sin_poly(phi);
    switch (q & 3) {
        case 0: return c;
        case 1: return -s;
        case 2: return -c;
        default: return s;
    }
}
constexpr double ct_sin_turn(int m, int R) {
    m = ((m % R) + R) % R;
    const int q = (8 * m + R) / (2 * R);
    const double phi = 2.0 * kPiD * double(4 * m - q * R) / double(4 * R);
    const double c = ct_cos_poly(phi), s = ct_sin_poly(phi);
    switch (q & 3) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// a · w_R^M (forward: exp(-2πi M/R), inverse: conjugate), M known at compile time
template <typename T, bool INV, int R, int M>
__device__ __forceinline__ cplx<T> ct_twiddle(const cplx<T> &a) {
    constexpr int m = M % R;
    if constexpr (m == 0) {
        return a;
    } else if constexpr (2 * m == R) {
        return mk(-a.x, -a.y);
    } else if constexpr (4 * m == R) {
        return rot90<T, INV>(a);
    } else if constexpr (4 * m == 3 * R) {
        return rot90<T, !INV>(a);
    } else {
        constexpr double c = ct_cos_turn(m, R);
        constexpr double s = INV ? ct_sin_turn(m, R) : -ct_sin_turn(m, R);
        return mk(a.x * T(c) - a.y * T(s), a.x * T(s) + a.y * T(c));
    }
}

constexpr int rdft_first_factor(int R) { return R % 4 == 0 ? 4 : (R % 3 == 0 ? 3 : 2); }

// In-register DFT of length R (R = 2^a·3^b), natural order in and out.
template <typename T, bool INV, int R>
__device__ __forceinline__ void rdft(cplx<T> (&v)[R]) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        const cplx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (R == 3) {
        const T h = T(0.86602540378443864676);  // sin(2π/3)
        const cplx<T> a0 = v[0], s = v[1] + v[2], d = v[1] - v[2];
        const cplx<T> m = a0 - T(0.5) * s;
        const cplx<T> r = rot90<T, INV>(h * d);
        v[0] = a0 + s;
        v[1] = m + r;
        v[2] = m - r;
    } else if constexpr (R == 4) {
        const cplx<T> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3], t3 = rot90<T, INV>(v[1] - v[3]);
        v[0] = t0 + t2;
        v[1] = t1 + t3;
        v[2] = t0 - t2;
        v[3] = t1 - t3;
    } else {
        constexpr int P = rdft_first_factor(R), Q = R / P;
        static_assert(P * Q == R, "length must factor into 2, 3 and 4");
        cplx<T> y[R];  // y[n2·P + k1]
        static_for<0, Q>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            cplx<T> a[P];
            static_for<0, P>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                a[n1] = v[Q * n1 + n2];
            });
            rdft<T, INV, P>(a);
            static_for<0, P>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                y[n2 * P + k1] = ct_twiddle<T, INV, R, n2 * k1>(a[k1]);
            });
        });
        static_for<0, P>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            cplx<T> b[Q];
            static_for<0, Q>([&](auto n2c) {
                constexpr int n2 = decltype(n2c)::value;
                b[n2] = y[n2 * P + k1];
            });
            rdft<T, INV, Q>(b);
            static_for<0, Q>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                v[k1 + P * k2] = b[k2];
            });
        });
    }
}

// Register plan N = A·B of the lengths the transform passes specialise
// (0: no register plan, the LDS Stockham FFT is used).
constexpr int reg_fft_a(int N) {
    switch (N) {
        case 16: return 4;
        case 24: return 6;
        case 32: return 8;
        case 48: return 8;
        case 64: return 8;
        case 72: return 9;
        case 96: return 12;
        case 128: return 16;
        case 144: return 12;
        case 192: return 16;
        case 256: return 16;
        case 384: return 24;
        default: return 0;
    }
}

}  // namespace bld
