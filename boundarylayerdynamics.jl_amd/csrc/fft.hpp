// Batched mixed-radix Stockham FFT on LDS buffers, executed by a whole block.
//
// For each stage of radix R (span Ns = product of the previous radices):
//   y[(j/Ns)·Ns·R + j%Ns + q·Ns] = Σ_r x[j + r·N/R] · w^{r·(j%Ns)·N/(Ns·R)} · ω_R^{r·q}
// (self-sorting, natural order in and out).  Radices 2, 3, 4, 5 have closed
// butterflies; any other prime factor uses a direct DFT of that length.
// Twiddles come from one table exp(-2πi t/N) per length (computed on the host
// in extended precision); the inverse transform uses their conjugates.
#pragma once

#include "common.hpp"

namespace bld {

template <typename T, bool INV>
__device__ __forceinline__ cplx<T> twiddle(const cplx<T> *__restrict__ tw, int t) {
    cplx<T> w = tw[t];
    if (INV) w.y = -w.y;
    return w;
}

// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> rot90(cplx<T> a) {
    return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

template <typename T, bool INV>
__device__ __forceinline__ void stage_radix(const cplx<T> *__restrict__ in, cplx<T> *__restrict__ out, int N, int R,
                                            int Ns, int jj, const cplx<T> *__restrict__ tw) {
    const int M = N / R;
    const int k = jj % Ns;
    const int idxD = (jj / Ns) * Ns * R + k;
    const int tstep = k * (N / (Ns * R));
    if (R == 4) {
        cplx<T> a0 = in[jj], a1 = in[jj + M], a2 = in[jj + 2 * M], a3 = in[jj + 3 * M];
        if (tstep) {
            a1 = cmul(a1, twiddle<T, INV>(tw, tstep));
            a2 = cmul(a2, twiddle<T, INV>(tw, 2 * tstep));
            a3 = cmul(a3, twiddle<T, INV>(tw, 3 * tstep));
        }
        cplx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = rot90<T, INV>(a1 - a3);
        out[idxD] = t0 + t2;
        out[idxD + Ns] = t1 + t3;
        out[idxD + 2 * Ns] = t0 - t2;
        out[idxD + 3 * Ns] = t1 - t3;
    } else if (R == 2) {
        cplx<T> a0 = in[jj], a1 = in[jj + M];
        if (tstep) a1 = cmul(a1, twiddle<T, INV>(tw, tstep));
        out[idxD] = a0 + a1;
        out[idxD + Ns] = a0 - a1;
    } else if (R == 3) {
        const T h = T(0.86602540378443864676);  // sin(2π/3)
        cplx<T> a0 = in[jj], a1 = in[jj + M], a2 = in[jj + 2 * M];
        if (tstep) {
            a1 = cmul(a1, twiddle<T, INV>(tw, tstep));
            a2 = cmul(a2, twiddle<T, INV>(tw, 2 * tstep));
        }
        cplx<T> s = a1 + a2, d = a1 - a2;
        cplx<T> m = a0 - T(0.5) * s;
        cplx<T> r = rot90<T, INV>(h * d);  // ∓i·sin(2π/3)·d
        out[idxD] = a0 + s;
        out[idxD + Ns] = m + r;
        out[idxD + 2 * Ns] = m - r;
    } else if (R == 5) {
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
        cplx<T> a0 = in[jj], a1 = in[jj + M], a2 = in[jj + 2 * M], a3 = in[jj + 3 * M], a4 = in[jj + 4 * M];
        if (tstep) {
            a1 = cmul(a1, twiddle<T, INV>(tw, tstep));
            a2 = cmul(a2, twiddle<T, INV>(tw, 2 * tstep));
            a3 = cmul(a3, twiddle<T, INV>(tw, 3 * tstep));
            a4 = cmul(a4, twiddle<T, INV>(tw, 4 * tstep));
        }
        cplx<T> b1 = a1 + a4, b2 = a2 + a3, d1 = a1 - a4, d2 = a2 - a3;
        cplx<T> m1 = a0 + c1 * b1 + c2 * b2, m2 = a0 + c2 * b1 + c1 * b2;
        cplx<T> r1 = rot90<T, INV>(s1 * d1 + s2 * d2), r2 = rot90<T, INV>(s2 * d1 - s1 * d2);
        out[idxD] = a0 + b1 + b2;
        out[idxD + Ns] = m1 + r1;
        out[idxD + 4 * Ns] = m1 - r1;
        out[idxD + 2 * Ns] = m2 + r2;
        out[idxD + 3 * Ns] = m2 - r2;
    } else {  // direct DFT of a prime length R
        const int wr = N / R;  // ω_R^m = w^{m·N/R}
        for (int q = 0; q < R; ++q) {
            cplx<T> acc = mk(T(0), T(0));
            for (int r = 0; r < R; ++r) {
                cplx<T> a = in[jj + r * M];
                if (tstep) a = cmul(a, twiddle<T, INV>(tw, r * tstep));
                acc = acc + cmul(a, twiddle<T, INV>(tw, ((r * q) % R) * wr));
            }
            out[idxD + q * Ns] = acc;
        }
    }
}

// Runs `batch` transforms of length p.n stored contiguously in `a` (stride n);
// `b` is scratch of the same size.  Returns the buffer that holds the result.
// Must be called by all threads of the block; ends with a barrier.
template <typename T, bool INV>
__device__ cplx<T> *fft_block(cplx<T> *a, cplx<T> *b, const FftPlan &p, int batch, const cplx<T> *__restrict__ tw) {
    const int N = p.n;
    int Ns = 1;
    for (int s = 0; s < p.nstages; ++s) {
        const int R = p.radix[s];
        const int M = N / R;
        for (int j = threadIdx.x; j < batch * M; j += blockDim.x) {
            const int t = j / M;
            const int jj = j - t * M;
            stage_radix<T, INV>(a + t * N, b + t * N, N, R, Ns, jj, tw);
        }
        __syncthreads();
        cplx<T> *tmp = a;
        a = b;
        b = tmp;
        Ns *= R;
    }
    return a;
}

// ---------------------------------------------------------------------------
// Compile-time lengths: the radix sequence, spans and strides are constants, so
// the per-butterfly index arithmetic reduces to multiplies and shifts and the
// stage loop unrolls.  Same factorisation and arithmetic as fft_block.
// ---------------------------------------------------------------------------
struct Radices {
    int n;
    int r[kMaxStages];
};

constexpr Radices factor_ct(int N) {
    Radices f{0, {}};
    int n = N;
    while (n % 4 == 0) {
        f.r[f.n++] = 4;
        n /= 4;
    }
    for (int p = 2; p <= 5; ++p) {
        if (p == 4) continue;
        while (n % p == 0) {
            f.r[f.n++] = p;
            n /= p;
        }
    }
    for (int p = 7; n > 1; p += 2) {
        while (n % p == 0) {
            f.r[f.n++] = p;
            n /= p;
        }
    }
    return f;
}

template <typename T, bool INV, int N, int S, int NS>
__device__ __forceinline__ void fft_stages_ct(cplx<T> *&a, cplx<T> *&b, int batch, const cplx<T> *__restrict__ tw) {
    constexpr Radices F = factor_ct(N);
    if constexpr (S < F.n) {
        constexpr int R = F.r[S];
        constexpr int M = N / R;
        for (int j = threadIdx.x; j < batch * M; j += blockDim.x) {
            const int t = j / M;
            const int jj = j - t * M;
            stage_radix<T, INV>(a + t * N, b + t * N, N, R, NS, jj, tw);
        }
        __syncthreads();
        cplx<T> *tmp = a;
        a = b;
        b = tmp;
        fft_stages_ct<T, INV, N, S + 1, NS * R>(a, b, batch, tw);
    }
}

template <typename T, bool INV, int N>
__device__ __forceinline__ cplx<T> *fft_block_ct(cplx<T> *a, cplx<T> *b, int batch, const cplx<T> *__restrict__ tw) {
    fft_stages_ct<T, INV, N, 0, 1>(a, b, batch, tw);
    return a;
}

}  // namespace bld
