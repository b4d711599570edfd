// Transform passes on the register FFT (lengths with a register plan N = A·B).
//
// Same dataflow as the LDS passes (physical_space.jl:191-312): c2r along x2
// into the intermediate [job][N2][k1m+1], c2r along x1 with two real rows per
// complex FFT, r2c along x1, r2c along x2 with truncation, 1/(N1·N2) scaling,
// derivative prefactors and accumulation into the rates.  Here the values
// move from global memory straight into registers, cross LDS once between the
// two register steps, and go back to global memory from registers:
//  * x2 passes: a block takes `batch` = C consecutive k1 columns of one job;
//    tasks are (column c, index t) with c fastest, so loads and stores are
//    contiguous over k1;
//  * x1 passes: a block takes C row pairs; tasks are (pair p, index t) with t
//    fastest, so loads and stores are contiguous along the row.  The LDS rows
//    of an x1 pass are padded to A+1 entries to spread the stride-A accesses
//    of step 1 over the banks.
#pragma once

#include "fft_reg.hpp"
#include "terms.hpp"

namespace bld {

// twiddle exp(∓2πi t/N) from the table of the length (conjugated for INV)
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> gtwiddle(const cplx<T> *__restrict__ tw, int t) {
    cplx<T> w = gload(tw + t);
    if (INV) w.y = -w.y;
    return w;
}

// LDS entries (complex) per batch item of a transform pass with a register
// plan: x2 passes one column per region (r2c_x2: G and H sums), x1 passes the
// padded rows of a pair (r2c_x1: and the pair's spectrum)
__host__ __device__ constexpr int reg_item_lds(int kind, int N, int regions) {
    return kind == 0 ? N
         : kind == 3 ? regions * N
         : kind == 1 ? (N / reg_fft_a(N)) * (reg_fft_a(N) + 1)
                     : (N / reg_fft_a(N)) * (reg_fft_a(N) + 1) + N;
}

// c2r, pass 1: inverse FFT along x2 of C k1-columns of one (term, layer) job
template <typename T, int N>
__device__ __forceinline__ void c2r_x2_reg(const Space<T> &S, const LaunchEntry &E, int sel, cplx<T> *Y) {
    constexpr int A = reg_fft_a(N), Bq = N / A;
    const int local = blockIdx.x - E.block_begin;
    const int job = local / E.blocks_per_job;
    const int chunk = local - job * E.blocks_per_job;
    int term = 0;
    while (job >= S.in_job0[term + 1]) ++term;
    const int layer = job - S.in_job0[term];
    const int recipe = S.in_recipe[term];
    const int C = E.batch, lc = __ffs(C) - 1;  // C is a power of two
    const int k1_0 = chunk * C;
    const int ncols = min(C, S.nk1 - k1_0);
    const cplx<T> *__restrict__ tw = S.tw + S.p2.tw;
    const size_t P = size_t(S.K1) * S.K2;
    for (int task = threadIdx.x; task < C * Bq; task += blockDim.x) {
        const int c = task & (C - 1), t = task >> lc;
        const int k1 = k1_0 + min(c, ncols - 1);  // clamped: every load below is in bounds
        cplx<T> v[A];
        // indices of the A values (resolved modes only; the padding is zero):
        // all loads are issued before any is used, no branches around them
        int i2s[A];
        bool ok[A];
        static_for<0, A>([&](auto n1c) {
            constexpr int n1 = decltype(n1c)::value;
            const int pos = Bq * n1 + t;
            const bool lo = pos <= S.k2m, hi = pos >= N - S.k2m;
            ok[n1] = (lo || hi) && c < ncols;
            i2s[n1] = lo ? pos : (hi ? S.K2 - (N - pos) : 0);
        });
        const size_t base = size_t(layer) * P + k1;
        auto gather = [&](auto value) {
            static_for<0, A>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                const cplx<T> x = value(base + size_t(S.K1) * i2s[n1], i2s[n1]);
                v[n1] = ok[n1] ? x : czero<T>();
            });
        };
        const cplx<T> *u = S.state[sel][0], *vv = S.state[sel][1], *w = S.state[sel][2];
        const T d1 = T(gload(S.D1 + k1));
        switch (recipe) {  // uniform per block
            case R_U: gather([&](size_t i, int) { return gload(u + i); }); break;
            case R_V: gather([&](size_t i, int) { return gload(vv + i); }); break;
            case R_W: gather([&](size_t i, int) { return gload(w + i); }); break;
            case R_D1U: gather([&](size_t i, int) { return imul(d1, gload(u + i)); }); break;
            case R_D1V: gather([&](size_t i, int) { return imul(d1, gload(vv + i)); }); break;
            case R_D1W: gather([&](size_t i, int) { return imul(d1, gload(w + i)); }); break;
            case R_D2U: gather([&](size_t i, int i2) { return imul(T(gload(S.D2 + i2)), gload(u + i)); }); break;
            case R_D2V: gather([&](size_t i, int i2) { return imul(T(gload(S.D2 + i2)), gload(vv + i)); }); break;
            case R_D2W: gather([&](size_t i, int i2) { return imul(T(gload(S.D2 + i2)), gload(w + i)); }); break;
            default: gather([&](size_t i, int i2) { return term_value(S, term, recipe, sel, layer, k1, i2); });
        }
        rdft<T, true, A>(v);
        static_for<0, A>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            const cplx<T> y = k1 == 0 ? v[0] : cmul(v[k1], gtwiddle<T, true>(tw, t * k1));
            Y[(t * A + k1) * C + c] = y;
        });
    }
    __syncthreads();
    cplx<T> *out = S.inter + size_t(job) * N * S.nk1;
    for (int task = threadIdx.x; task < C * A; task += blockDim.x) {
        const int c = task & (C - 1), t = task >> lc;
        cplx<T> v[Bq];
        static_for<0, Bq>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            v[n2] = Y[(n2 * A + t) * C + c];
        });
        rdft<T, true, Bq>(v);
        if (c < ncols) {
            static_for<0, Bq>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                cplx<T> val = v[k2];
                if (k1_0 + c == 0) val.y = T(0);  // c2r uses the real part of the k1 = 0 column
                gstore(out + size_t(t + A * k2) * S.nk1 + k1_0 + c, val);
            });
        }
    }
}

// c2r, pass 2: Hermitian completion along x1, two real rows per complex FFT
template <typename T, int N>
__device__ __forceinline__ void c2r_x1_reg(const Space<T> &S, const LaunchEntry &E, cplx<T> *Y) {
    constexpr int A = reg_fft_a(N), Bq = N / A, RS = A + 1, PS = Bq * RS;
    const int local = blockIdx.x - E.block_begin;
    const int job = local / E.blocks_per_job;
    const int chunk = local - job * E.blocks_per_job;
    int term = 0;
    while (job >= S.in_job0[term + 1]) ++term;
    const int layer = job - S.in_job0[term];
    const int C = E.batch, N2 = S.N2, k1m = S.k1m, nk1 = S.nk1;
    const int npairs_tot = (N2 + 1) / 2;
    const int pair0 = chunk * C;
    const int np = min(C, npairs_tot - pair0);
    const cplx<T> *in = S.inter + size_t(job) * N2 * nk1;
    const cplx<T> *__restrict__ tw = S.tw + S.p1.tw;
    for (int task = threadIdx.x; task < C * Bq; task += blockDim.x) {
        const int p = task / Bq, t = task - p * Bq;
        const int xa = 2 * (pair0 + min(p, np - 1));  // clamped: loads stay in bounds
        const bool live = p < np, hb = xa + 1 < N2;
        const cplx<T> *ra = in + size_t(xa) * nk1, *rb = in + size_t(hb ? xa + 1 : xa) * nk1;
        cplx<T> v[A];
        // Hermitian completion of the two rows (a + i·b): branch-free selects
        // after unconditional loads, so all loads are in flight together
        static_for<0, A>([&](auto n1c) {
            constexpr int n1 = decltype(n1c)::value;
            const int k = Bq * n1 + t;
            const bool direct = k <= k1m, mirror = !direct && N - k <= k1m;
            const bool real_only = k == 0 || (2 * k == N && direct);
            const int kk = direct ? k : (mirror ? N - k : 0);
            const cplx<T> a = gload(ra + kk);
            cplx<T> b = gload(rb + kk);
            if (!hb) b = czero<T>();
            cplx<T> z = real_only ? mk(a.x, b.x)
                      : direct   ? mk(a.x - b.y, a.y + b.x)
                                 : mk(a.x + b.y, b.x - a.y);
            if (!(live && (direct || mirror))) z = czero<T>();
            v[n1] = z;
        });
        rdft<T, true, A>(v);
        static_for<0, A>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            Y[p * PS + t * RS + k1] = k1 == 0 ? v[0] : cmul(v[k1], gtwiddle<T, true>(tw, t * k1));
        });
    }
    __syncthreads();
    T *out = S.in_phys[term] + size_t(layer) * N * N2;
    for (int task = threadIdx.x; task < C * A; task += blockDim.x) {
        const int p = task / A, t = task - p * A;
        cplx<T> v[Bq];
        static_for<0, Bq>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            v[n2] = Y[p * PS + n2 * RS + t];
        });
        rdft<T, true, Bq>(v);
        if (p < np) {
            const int xa = 2 * (pair0 + p);
            const bool hb = xa + 1 < N2;
            static_for<0, Bq>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                const int x = t + A * k2;
                gstore(out + size_t(xa) * N + x, v[k2].x);
                if (hb) gstore(out + size_t(xa + 1) * N + x, v[k2].y);
            });
        }
    }
}

// r2c, pass 1: forward FFT along x1 of two real rows at a time, split into the
// two half spectra (k1 = 0..k1m)
template <typename T, int N>
__device__ __forceinline__ void r2c_x1_reg(const Space<T> &S, const LaunchEntry &E, cplx<T> *Y) {
    constexpr int A = reg_fft_a(N), Bq = N / A, RS = A + 1, PS = Bq * RS;
    const int local = blockIdx.x - E.block_begin;
    const int job = local / E.blocks_per_job;
    const int chunk = local - job * E.blocks_per_job;
    int src = 0;
    while (job >= S.src_job0[src + 1]) ++src;
    const int layer = job - S.src_job0[src];
    const int C = E.batch, N2 = S.N2, nk1 = S.nk1;
    const int npairs_tot = (N2 + 1) / 2;
    const int pair0 = chunk * C;
    const int np = min(C, npairs_tot - pair0);
    const T *in = S.src_phys[src] + size_t(layer) * N * N2;
    const cplx<T> *__restrict__ tw = S.tw + S.p1.tw;
    for (int task = threadIdx.x; task < C * Bq; task += blockDim.x) {
        const int p = task / Bq, t = task - p * Bq;
        const int xa = 2 * (pair0 + min(p, np - 1));  // clamped: loads stay in bounds
        const bool live = p < np, hb = xa + 1 < N2;
        const T *ra = in + size_t(xa) * N, *rb = in + size_t(hb ? xa + 1 : xa) * N;
        cplx<T> v[A];
        static_for<0, A>([&](auto n1c) {
            constexpr int n1 = decltype(n1c)::value;
            const int x = Bq * n1 + t;
            const T a = gload(ra + x), b = gload(rb + x);
            v[n1] = mk(live ? a : T(0), live && hb ? b : T(0));
        });
        rdft<T, false, A>(v);
        static_for<0, A>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            Y[p * PS + t * RS + k1] = k1 == 0 ? v[0] : cmul(v[k1], gtwiddle<T, false>(tw, t * k1));
        });
    }
    __syncthreads();
    // spectrum of pair p in natural order at Z[p·N + k]
    cplx<T> *Z = Y + C * PS;
    for (int task = threadIdx.x; task < C * A; task += blockDim.x) {
        const int p = task / A, t = task - p * A;
        cplx<T> v[Bq];
        static_for<0, Bq>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            v[n2] = Y[p * PS + n2 * RS + t];
        });
        rdft<T, false, Bq>(v);
        static_for<0, Bq>([&](auto k2c) {
            constexpr int k2 = decltype(k2c)::value;
            Z[p * N + t + A * k2] = v[k2];
        });
    }
    __syncthreads();
    cplx<T> *out = S.inter + size_t(job) * N2 * nk1;
    for (int idx = threadIdx.x; idx < np * nk1; idx += blockDim.x) {
        const int p = idx / nk1, k = idx - p * nk1;
        const int xa = 2 * (pair0 + p);
        const cplx<T> zk = Z[p * N + k];
        const cplx<T> zn = Z[p * N + (k == 0 ? 0 : N - k)];
        gstore(out + size_t(xa) * nk1 + k, mk(T(0.5) * (zk.x + zn.x), T(0.5) * (zk.y - zn.y)));
        if (xa + 1 < N2) gstore(out + size_t(xa + 1) * nk1 + k, mk(T(0.5) * (zk.y + zn.y), T(-0.5) * (zk.x - zn.x)));
    }
}

// r2c, pass 2: forward FFT along x2 of C k1-columns of the summed sources,
// truncation to the resolved k2, scaling and accumulation into the rates
template <typename T, int N>
__device__ __forceinline__ void r2c_x2_reg(const Space<T> &S, const LaunchEntry &E, cplx<T> *Y) {
    constexpr int A = reg_fft_a(N), Bq = N / A;
    const int local = blockIdx.x - E.block_begin;
    const int job = local / E.blocks_per_job;
    const int chunk = local - job * E.blocks_per_job;
    int tg = 0;
    while (job >= S.tgt_job0[tg + 1]) ++tg;
    const int layer = job - S.tgt_job0[tg];
    const int C = E.batch, lc = __ffs(C) - 1, nk1 = S.nk1;
    const int nG = S.tgt_nG[tg], nH = S.tgt_nH[tg];
    const int k1_0 = chunk * C;
    const int ncols = min(C, nk1 - k1_0);
    const cplx<T> *__restrict__ tw = S.tw + S.p2.tw;
    cplx<T> *YH = Y + C * N;
    for (int task = threadIdx.x; task < C * Bq; task += blockDim.x) {
        const int c = task & (C - 1), t = task >> lc;
        const int k1 = k1_0 + min(c, ncols - 1);  // clamped: loads stay in bounds
        const bool live = c < ncols;
        const T d1 = T(gload(S.D1 + k1));
        cplx<T> g[A], h[A];
        static_for<0, A>([&](auto n1c) {
            constexpr int n1 = decltype(n1c)::value;
            g[n1] = czero<T>();
            h[n1] = czero<T>();
        });
        // one source at a time (uniform loop), its A loads issued together
        for (int s = 0; s < nG; ++s) {
            const cplx<T> *src = S.inter + size_t(S.src_job0[S.tgt_G[tg][s]] + layer) * N * nk1 + k1;
            const bool dx = S.tgt_Gd1[tg][s];
            static_for<0, A>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                const cplx<T> v = gload(src + size_t(Bq * n1 + t) * nk1);
                g[n1] = g[n1] + (dx ? imul(d1, v) : v);
            });
        }
        for (int s = 0; s < nH; ++s) {
            const cplx<T> *src = S.inter + size_t(S.src_job0[S.tgt_H[tg][s]] + layer) * N * nk1 + k1;
            static_for<0, A>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                h[n1] = h[n1] + gload(src + size_t(Bq * n1 + t) * nk1);
            });
        }
        if (!live) {
            static_for<0, A>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                g[n1] = czero<T>();
                h[n1] = czero<T>();
            });
        }
        rdft<T, false, A>(g);
        static_for<0, A>([&](auto k1c) {
            constexpr int kk = decltype(k1c)::value;
            Y[(t * A + kk) * C + c] = kk == 0 ? g[0] : cmul(g[kk], gtwiddle<T, false>(tw, t * kk));
        });
        if (nH > 0) {
            rdft<T, false, A>(h);
            static_for<0, A>([&](auto k1c) {
                constexpr int kk = decltype(k1c)::value;
                YH[(t * A + kk) * C + c] = kk == 0 ? h[0] : cmul(h[kk], gtwiddle<T, false>(tw, t * kk));
            });
        }
    }
    __syncthreads();
    const T scale = T(1.0 / (double(S.N1) * double(N)));
    cplx<T> *out = S.tgt_out[tg] + size_t(layer) * S.K1 * S.K2;
    for (int task = threadIdx.x; task < C * A; task += blockDim.x) {
        const int c = task & (C - 1), t = task >> lc;
        cplx<T> g[Bq], h[Bq];
        static_for<0, Bq>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            g[n2] = Y[(n2 * A + t) * C + c];
        });
        rdft<T, false, Bq>(g);
        if (nH > 0) {
            static_for<0, Bq>([&](auto n2c) {
                constexpr int n2 = decltype(n2c)::value;
                h[n2] = YH[(n2 * A + t) * C + c];
            });
            rdft<T, false, Bq>(h);
        }
        if (c < ncols) {
            static_for<0, Bq>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                const int pos = t + A * k2;
                int i2 = -1;
                if (pos <= S.k2m)
                    i2 = pos;
                else if (pos >= N - S.k2m)
                    i2 = S.K2 - (N - pos);
                if (i2 >= 0) {
                    cplx<T> v = g[k2];
                    if (nH > 0) v = v + imul(T(gload(S.D2 + i2)), h[k2]);
                    const size_t o = size_t(i2) * S.K1 + k1_0 + c;
                    if (S.xfer_centered) v = cmul(v, centre_shift<T>(S, k1_0 + c, i2, -1.0));
                    gstore(out + o, gload(out + o) + scale * v);
                }
            });
        }
    }
}

}  // namespace bld
