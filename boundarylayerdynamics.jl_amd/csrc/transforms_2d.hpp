// Fused two-dimensional transforms of whole planes, one plane per block, for
// square grids (N1 = N2 = N) with a register FFT for N.  The c2r (one (term,
// layer) plane, physical_space.jl:277-295) runs the x2 column FFTs into an
// LDS-resident intermediate Z[n2][k1] and the x1 row FFTs from there to the
// physical plane; the r2c (one (target, layer) rate plane with all its
// sources, physical_space.jl:208-246) runs the same passes in the opposite
// order.  The intermediate never goes to HBM, which removes its write and
// re-read from the two-pass path; it has to fit in LDS next to the exchange
// buffer of one round of row pairs.
#pragma once
#include "transforms_reg.hpp"

namespace bld {

// Bank spread: the row items of a pair are B consecutive lanes reading B
// consecutive entries of two Z rows; with a row stride zs ≡ B/2 (mod 16) the
// pairs of a half wave start 2·B banks apart, so their reads of 8-byte entries
// fall on disjoint banks.
__host__ __device__ constexpr int xf2d_spread(int N, int need) {
    const int bh = reg_fft_a(N) > 0 ? (N / reg_fft_a(N)) / 2 : 0;
    return need + (((bh - need) % 16) + 16) % 16;
}

// c2r layout: Z[N][zs] and the N twiddles w_N^j (one table serves both
// directions of a square grid).  Two rows of Z hold the B·(A+1) entries of a
// row pair's FFT exchange, so the x1 pass works in place in the pair's rows.
__host__ __device__ constexpr int c2r2d_zs(int N, int nk1) {
    const int A = reg_fft_a(N) > 0 ? reg_fft_a(N) : 1, Bq = N / A;
    const int ex = (Bq * (A + 1) + 1) / 2;
    return xf2d_spread(N, nk1 > ex ? nk1 : ex);
}
__host__ __device__ constexpr size_t c2r2d_lds(int N, int nk1, size_t elem) {
    return (size_t(N) * c2r2d_zs(N, nk1) + size_t(N)) * elem;
}
// threads of the c2r block: 512 where two blocks fit in one CU's LDS
__host__ __device__ constexpr int c2r2d_threads(int N, int nk1, size_t elem) {
    return c2r2d_lds(N, nk1, elem) <= 80 * 1024 ? 512 : 1024;
}
// row pairs per round of the c2r x1 pass: one first-stage item per thread
__host__ __device__ constexpr int c2r2d_pairs(int N, int NT) {
    const int A = reg_fft_a(N) > 0 ? reg_fft_a(N) : 1, Bq = N / A, NP = N / 2;
    const int rounds = (NP * Bq + NT - 1) / NT;
    return (NP + rounds - 1) / rounds;
}

// r2c layout: Z[N][zs] (the summed half spectra), the row-pair exchange
// X[P][B·(A+1)] of one round, the twiddles, the x1 derivative factors D1[k1]
__host__ __device__ constexpr int r2c2d_zs(int N, int nk1) { return xf2d_spread(N, nk1); }
__host__ __device__ constexpr int r2c2d_pairs(int N) { return N > 128 ? 32 : (N / 2 < 64 ? N / 2 : 64); }
__host__ __device__ constexpr size_t r2c2d_lds(int N, int nk1, size_t elem) {
    return (size_t(N) * r2c2d_zs(N, nk1) +
            size_t(r2c2d_pairs(N)) * (reg_fft_a(N) > 0 ? (N / reg_fft_a(N)) * (reg_fft_a(N) + 1) : 0) + size_t(N) +
            size_t(nk1)) *
           elem;
}

template <typename T, bool INV>
__device__ __forceinline__ cplx<T> ltwiddle(const cplx<T> *TW, int j) {
    cplx<T> w = TW[j];
    if (INV) w.y = -w.y;
    return w;
}

// lengths with instantiated fused transforms (larger planes do not fit in LDS)
template <typename T>
__host__ __device__ constexpr bool c2r2d_supported(int N) {
    return reg_fft_a(N) > 0 && N % 2 == 0 && N <= (sizeof(T) == 4 ? 192 : 96);
}

// Half-plane c2r (rows of one parity r of a plane of length N): Z holds the
// N/2 rows of that parity with the same row stride rule as the whole plane
__host__ __device__ constexpr int c2r2h_zs(int N, int nk1) { return c2r2d_zs(N, nk1); }
__host__ __device__ constexpr size_t c2r2h_lds(int N, int nk1, size_t elem) {
    return (size_t(N / 2) * c2r2h_zs(N, nk1) + size_t(N)) * elem;
}
// lengths with an instantiated half-plane c2r: N/2 with a register FFT, and a
// whole plane too large for two blocks per CU (N ≥ 144 in float32)
template <typename T>
__host__ __device__ constexpr bool c2r2h_supported(int N) {
    return c2r2d_supported<T>(N) && N % 4 == 0 && reg_fft_a(N / 2) > 0 && N >= (sizeof(T) == 4 ? 144 : 96);
}


// Spectral layouts with compile-time sizes: KM = k1max = k2max with every
// resolved mode inside the plane (K1 = nk1 = KM+1, K2 = 2·KM+1), the layouts
// of square grids whose physical size is N (default size N = 2·KM+2, or the
// dealiased 3·KM+3).  KM < 0: sizes read from the space descriptor.
__host__ __device__ constexpr bool xf2d_km_special(int N, int KM) {
    return (N == 192 && KM == 63) || (N == 128 && KM == 63) || (N == 144 && KM == 47) ||
           (N == 96 && (KM == 31 || KM == 47)) || (N == 72 && KM == 23) || (N == 64 && KM == 31) ||
           (N == 48 && KM == 23);
}
template <typename T>
__device__ __forceinline__ bool xf2d_layout_is(const Space<T> &S, int KM) {
    return S.k1m == KM && S.k2m == KM && S.nk1 == KM + 1 && S.K1 == KM + 1 && S.K2 == 2 * KM + 1;
}

template <typename T, int N, int NT, int KM = -1>
__device__ __forceinline__ void c2r_2d_reg(const Space<T> &S, const LaunchEntry &E, int sel, cplx<T> *smem) {
    constexpr int A = reg_fft_a(N), Bq = N / A, RS = A + 1;
    constexpr int P = c2r2d_pairs(N, NT), NP = N / 2;
    static_assert(P * Bq <= NT, "one first-stage row item per thread");
    const int job = blockIdx.x - E.block_begin;
    int term = 0;
    while (job >= S.in_job0[term + 1]) ++term;
    const int layer = job - S.in_job0[term];
    const int recipe = S.in_recipe[term];
    const int nk1 = KM >= 0 ? KM + 1 : S.nk1, zs = c2r2d_zs(N, nk1);
    const int k1m = KM >= 0 ? KM : S.k1m, k2m = KM >= 0 ? KM : S.k2m;
    const int K1 = KM >= 0 ? KM + 1 : S.K1, K2 = KM >= 0 ? 2 * KM + 1 : S.K2;
    cplx<T> *Z = smem;
    cplx<T> *TW = smem + size_t(N) * zs;
    const size_t P2 = size_t(K1) * K2;
    // the twiddle table, loaded together with the first stage's data and
    // stored to LDS behind the first round's gathers
    cplx<T> twv[(N + NT - 1) / NT];
    static_for<0, (N + NT - 1) / NT>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int e = j * NT + int(threadIdx.x);
        twv[j] = e < N ? gload(S.tw + S.p2.tw + e) : czero<T>();
    });
    const cplx<T> *u = S.state[sel][0], *vv = S.state[sel][1], *w = S.state[sel][2];

    // x2, stage 1: item (c, t), c fastest; A values at k2 positions Bq·n + t,
    // in rounds of NT items (the same count for every thread)
    {
        const int items = nk1 * Bq, rounds = (items + NT - 1) / NT;
        for (int rd = 0; rd < rounds; ++rd) {
            const int it = rd * NT + int(threadIdx.x);
            const bool have = it < items;
            const int t = (have ? it : 0) / nk1, c = (have ? it : 0) - t * nk1;
            int i2s[A];
            bool ok[A];
            static_for<0, A>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                const int pos = Bq * n1 + t;
                const bool lo = pos <= k2m, hi = pos >= N - k2m;
                ok[n1] = lo || hi;
                i2s[n1] = lo ? pos : (hi ? K2 - (N - pos) : 0);
            });
            const uint32_t base = uint32_t(layer) * uint32_t(P2) + uint32_t(c);
            cplx<T> v[A];
            // padding positions (k2 between the resolved halves) load nothing:
            // with 64 columns per wave their test is uniform across the wave
            auto gather = [&](auto value) {
                static_for<0, A>([&](auto n1c) {
                    constexpr int n1 = decltype(n1c)::value;
                    v[n1] = czero<T>();
                    if (have && ok[n1]) v[n1] = value(base + uint32_t(K1) * uint32_t(i2s[n1]), i2s[n1]);
                });
            };
            const T d1 = T(gload(S.D1 + c));
            auto d2 = [&](int i2) { return T(gload(S.D2 + i2)); };
            switch (recipe) {  // uniform per block
                case R_U: gather([&](uint32_t i, int) { return gload(u + i); }); break;
                case R_V: gather([&](uint32_t i, int) { return gload(vv + i); }); break;
                case R_W: gather([&](uint32_t i, int) { return gload(w + i); }); break;
                case R_D1U: gather([&](uint32_t i, int) { return imul(d1, gload(u + i)); }); break;
                case R_D1V: gather([&](uint32_t i, int) { return imul(d1, gload(vv + i)); }); break;
                case R_D1W: gather([&](uint32_t i, int) { return imul(d1, gload(w + i)); }); break;
                case R_D2U: gather([&](uint32_t i, int i2) { return imul(d2(i2), gload(u + i)); }); break;
                case R_D2V: gather([&](uint32_t i, int i2) { return imul(d2(i2), gload(vv + i)); }); break;
                case R_D2W: gather([&](uint32_t i, int i2) { return imul(d2(i2), gload(w + i)); }); break;
                // vorticity (derivatives.jl:566-568), the operations of term_value
                case R_VORT1: {
                    const T a = T(gload(S.alpha_i + layer + 1));
                    gather([&](uint32_t i, int i2) {
                        const cplx<T> dz = (-a) * gload(vv + i) + a * gload(vv + i + P2);
                        return imul(d2(i2), gload(w + i)) - dz;
                    });
                    break;
                }
                case R_VORT2: {
                    const T a = T(gload(S.alpha_i + layer + 1));
                    gather([&](uint32_t i, int) {
                        const cplx<T> dz = (-a) * gload(u + i) + a * gload(u + i + P2);
                        return dz - imul(d1, gload(w + i));
                    });
                    break;
                }
                case R_VORT3:
                    gather([&](uint32_t i, int i2) { return imul(d1, gload(vv + i)) - imul(d2(i2), gload(u + i)); });
                    break;
                default: gather([&](uint32_t, int i2) { return term_value(S, term, recipe, sel, layer, c, i2); });
            }
            if (rd == 0) {  // the twiddles in LDS for every round
                static_for<0, (N + NT - 1) / NT>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    const int e = j * NT + int(threadIdx.x);
                    if (e < N) TW[e] = twv[j];
                });
                __syncthreads();
            }
            if (have) {
                rdft<T, true, A>(v);
                static_for<0, A>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    Z[(t * A + k) * zs + c] = k == 0 ? v[0] : cmul(v[k], ltwiddle<T, true>(TW, t * k));
                });
            }
        }
    }
    __syncthreads();
    // x2, stage 2: item (c, t) reads rows n·A + t and writes rows t + A·k, the
    // same set, so the update is in place
    for (int it = threadIdx.x; it < nk1 * A; it += NT) {
        const int t = it / nk1, c = it - t * nk1;
        cplx<T> v[Bq];
        static_for<0, Bq>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            v[n] = Z[(n * A + t) * zs + c];
        });
        rdft<T, true, Bq>(v);
        static_for<0, Bq>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            cplx<T> val = v[k];
            if (c == 0) val.y = T(0);  // c2r uses the real part of the k1 = 0 column
            Z[(t + A * k) * zs + c] = val;
        });
    }
    __syncthreads();
    // x1: rounds of P row pairs, Hermitian completion of (row 2p) + i·(row 2p+1);
    // the pair's exchange lives in its own two rows (2·zs ≥ Bq·RS entries)
    T *out = S.in_phys[term] + size_t(layer) * N * N;
    for (int p0 = 0; p0 < NP; p0 += P) {
        const int it = threadIdx.x;
        const int pl = it / Bq, t = it - pl * Bq;
        const bool own = pl < P && p0 + pl < NP;
        const int p = min(p0 + pl, NP - 1);
        cplx<T> *R = Z + size_t(2 * p) * zs;  // rows 2p and 2p+1: the pair's region
        cplx<T> v[A];
        static_for<0, A>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            const int k = Bq * n + t;
            const bool direct = k <= k1m, mirror = !direct && N - k <= k1m;
            const bool real_only = k == 0 || (2 * k == N && direct);
            const int kk = direct ? k : (mirror ? N - k : 0);
            const cplx<T> a = R[kk], b = R[zs + kk];
            cplx<T> z = real_only ? mk(a.x, b.x) : direct ? mk(a.x - b.y, a.y + b.x) : mk(a.x + b.y, b.x - a.y);
            if (!(direct || mirror)) z = czero<T>();
            v[n] = z;
        });
        __syncthreads();  // every read of the round's rows is done before the exchange overwrites them
        if (own) {
            rdft<T, true, A>(v);
            static_for<0, A>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                R[t * RS + k] = k == 0 ? v[0] : cmul(v[k], ltwiddle<T, true>(TW, t * k));
            });
        }
        __syncthreads();
        for (int i2 = threadIdx.x; i2 < P * A; i2 += NT) {
            const int ql = i2 / A, s = i2 - ql * A;
            const int q = p0 + ql;
            if (q < NP) {
                const cplx<T> *Q = Z + size_t(2 * q) * zs;
                cplx<T> y[Bq];
                static_for<0, Bq>([&](auto nc) {
                    constexpr int n = decltype(nc)::value;
                    y[n] = Q[n * RS + s];
                });
                rdft<T, true, Bq>(y);
                T *ra = out + size_t(2 * q) * N, *rb = ra + N;
                static_for<0, Bq>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    gstore(ra + s + A * k, y[k].x);
                    gstore(rb + s + A * k, y[k].y);
                });
            }
        }
        // the next round touches other rows only: no barrier
    }
}

// Half-plane c2r: the block of job (plane, r) produces the physical rows
// 2m + r of the plane (r = 0, 1).  With n2 = 2m + r the inverse x2 transform
// of a column is x[2m+r] = Σ_{k<M} e^{2πi·mk/M}·Y_r[k], M = N/2, with
// Y_r[k] = (X[k] + (−1)^r·X[k+M])·e^{2πi·rk/N}, so the column FFTs have length
// M and the intermediate holds M rows; the x1 pass is the whole-plane one on
// those rows.  The plane's inputs are read by both blocks of the plane.
template <typename T, int N, int NT>
__device__ __forceinline__ void c2r_2h_reg(const Space<T> &S, const LaunchEntry &E, int sel, cplx<T> *smem) {
    constexpr int M = N / 2;
    constexpr int AM = reg_fft_a(M), BM = M / AM;  // column FFT
    constexpr int A = reg_fft_a(N), Bq = N / A, RS = A + 1;  // row FFT
    constexpr int NP = M / 2;  // row pairs of the half plane
    constexpr int ROUNDS = (NP * Bq + NT - 1) / NT;
    constexpr int P = (NP + ROUNDS - 1) / ROUNDS;
    static_assert(P * Bq <= NT, "one first-stage row item per thread");
    // Blocks are dealt to the 8 XCDs in turn, so blocks 8 apart share an L2:
    // the two halves of a plane are blocks 8h + x of a group of 16 (half h),
    // and the second half's gathers hit what the first half brought in.
    const int jb = blockIdx.x - E.block_begin;
    const int nb16 = (2 * E.njobs) & ~15;
    int job, r;
    if (jb < nb16) {
        job = (jb >> 4) * 8 + (jb & 7);
        r = (jb >> 3) & 1;
    } else {
        job = (nb16 >> 1) + ((jb - nb16) >> 1);
        r = (jb - nb16) & 1;
    }
    int term = 0;
    while (job >= S.in_job0[term + 1]) ++term;
    const int layer = job - S.in_job0[term];
    const int recipe = S.in_recipe[term];
    const int nk1 = S.nk1, zs = c2r2h_zs(N, nk1), k1m = S.k1m, k2m = S.k2m, K1 = S.K1, K2 = S.K2;
    cplx<T> *Z = smem;
    cplx<T> *TW = smem + size_t(M) * zs;
    const size_t P2 = size_t(K1) * K2;
    // the twiddle table, loaded together with the first round's gathers and
    // stored to LDS behind them
    constexpr int NTW = (N + NT - 1) / NT;
    cplx<T> twv[NTW];
    static_for<0, NTW>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int e = j * NT + int(threadIdx.x);
        twv[j] = e < N ? gload(S.tw + S.p2.tw + e) : czero<T>();
    });
    const cplx<T> *u = S.state[sel][0], *vv = S.state[sel][1], *w = S.state[sel][2];

    // x2, stage 1: item (c, t), c fastest; AM values Y_r[k] at k = BM·n + t,
    // in rounds of NT items (the same count for every thread)
    {
        const int items = nk1 * BM, rounds = (items + NT - 1) / NT;
        const T sg = r ? T(-1) : T(1);
        for (int rd = 0; rd < rounds; ++rd) {
            const int it = rd * NT + int(threadIdx.x);
            const bool have = it < items;
            const int t = (have ? it : 0) / nk1, c = (have ? it : 0) - t * nk1;
            const uint32_t base = uint32_t(layer) * uint32_t(P2) + uint32_t(c);
            cplx<T> v[AM];
            // resolved index of k2 position p (clamped to 0 where p is padding)
            auto idx = [&](int p, bool &ok) {
                const bool lo = p <= k2m, hi = p >= N - k2m;
                ok = lo || hi;
                return lo ? p : (hi ? K2 - (N - p) : 0);
            };
            // padding positions load nothing: with 64 columns per wave their
            // test is uniform across the wave
            auto gather = [&](auto value) {
                static_for<0, AM>([&](auto nc) {
                    constexpr int n = decltype(nc)::value;
                    bool oka, okb;
                    const int ia = idx(BM * n + t, oka), ib = idx(BM * n + t + M, okb);
                    cplx<T> a = czero<T>(), b = czero<T>();
                    if (have && oka) a = value(base + uint32_t(K1) * uint32_t(ia), ia);
                    if (have && okb) b = value(base + uint32_t(K1) * uint32_t(ib), ib);
                    v[n] = mk(a.x + sg * b.x, a.y + sg * b.y);
                });
            };
            const T d1 = T(gload(S.D1 + c));
            auto d2 = [&](int i2) { return T(gload(S.D2 + i2)); };
            switch (recipe) {  // uniform per block
                case R_U: gather([&](uint32_t i, int) { return gload(u + i); }); break;
                case R_V: gather([&](uint32_t i, int) { return gload(vv + i); }); break;
                case R_W: gather([&](uint32_t i, int) { return gload(w + i); }); break;
                case R_D1U: gather([&](uint32_t i, int) { return imul(d1, gload(u + i)); }); break;
                case R_D1V: gather([&](uint32_t i, int) { return imul(d1, gload(vv + i)); }); break;
                case R_D1W: gather([&](uint32_t i, int) { return imul(d1, gload(w + i)); }); break;
                case R_D2U: gather([&](uint32_t i, int i2) { return imul(d2(i2), gload(u + i)); }); break;
                case R_D2V: gather([&](uint32_t i, int i2) { return imul(d2(i2), gload(vv + i)); }); break;
                case R_D2W: gather([&](uint32_t i, int i2) { return imul(d2(i2), gload(w + i)); }); break;
                // vorticity (derivatives.jl:566-568), the operations of term_value
                case R_VORT1: {
                    const T al = T(gload(S.alpha_i + layer + 1));
                    gather([&](uint32_t i, int i2) {
                        const cplx<T> dz = (-al) * gload(vv + i) + al * gload(vv + i + uint32_t(P2));
                        return imul(d2(i2), gload(w + i)) - dz;
                    });
                    break;
                }
                case R_VORT2: {
                    const T al = T(gload(S.alpha_i + layer + 1));
                    gather([&](uint32_t i, int) {
                        const cplx<T> dz = (-al) * gload(u + i) + al * gload(u + i + uint32_t(P2));
                        return dz - imul(d1, gload(w + i));
                    });
                    break;
                }
                case R_VORT3:
                    gather([&](uint32_t i, int i2) { return imul(d1, gload(vv + i)) - imul(d2(i2), gload(u + i)); });
                    break;
                default: gather([&](uint32_t, int i2) { return term_value(S, term, recipe, sel, layer, c, i2); });
            }
            if (rd == 0) {  // the twiddles in LDS for every round
                static_for<0, NTW>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    const int e = j * NT + int(threadIdx.x);
                    if (e < N) TW[e] = twv[j];
                });
                __syncthreads();
            }
            if (have) {
                if (r) {  // e^{2πi·k/N}
                    static_for<1, AM>([&](auto nc) {
                        constexpr int n = decltype(nc)::value;
                        v[n] = cmul(v[n], ltwiddle<T, true>(TW, BM * n + t));
                    });
                    if (t) v[0] = cmul(v[0], ltwiddle<T, true>(TW, t));
                }
                rdft<T, true, AM>(v);
                static_for<0, AM>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    Z[(t * AM + k) * zs + c] = k == 0 ? v[0] : cmul(v[k], ltwiddle<T, true>(TW, 2 * ((t * k) % M)));
                });
            }
        }
    }
    __syncthreads();
    // x2, stage 2 in place: local rows t + AM·k
    for (int it = threadIdx.x; it < nk1 * AM; it += NT) {
        const int t = it / nk1, c = it - t * nk1;
        cplx<T> y[BM];
        static_for<0, BM>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            y[n] = Z[(n * AM + t) * zs + c];
        });
        rdft<T, true, BM>(y);
        static_for<0, BM>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            cplx<T> val = y[k];
            if (c == 0) val.y = T(0);  // c2r uses the real part of the k1 = 0 column
            Z[(t + AM * k) * zs + c] = val;
        });
    }
    __syncthreads();
    // x1 on the M local rows (physical rows 2m + r), pairs of local rows in place
    T *out = S.in_phys[term] + size_t(layer) * N * N;
    for (int p0 = 0; p0 < NP; p0 += P) {
        const int it = threadIdx.x;
        const int pl = it / Bq, t = it - pl * Bq;
        const bool own = pl < P && p0 + pl < NP;
        const int p = min(p0 + pl, NP - 1);
        cplx<T> *R = Z + size_t(2 * p) * zs;
        cplx<T> v[A];
        static_for<0, A>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            const int k = Bq * n + t;
            const bool direct = k <= k1m, mirror = !direct && N - k <= k1m;
            const bool real_only = k == 0 || (2 * k == N && direct);
            const int kk = direct ? k : (mirror ? N - k : 0);
            const cplx<T> a = R[kk], b = R[zs + kk];
            cplx<T> z = real_only ? mk(a.x, b.x) : direct ? mk(a.x - b.y, a.y + b.x) : mk(a.x + b.y, b.x - a.y);
            if (!(direct || mirror)) z = czero<T>();
            v[n] = z;
        });
        __syncthreads();
        if (own) {
            rdft<T, true, A>(v);
            static_for<0, A>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                R[t * RS + k] = k == 0 ? v[0] : cmul(v[k], ltwiddle<T, true>(TW, t * k));
            });
        }
        __syncthreads();
        for (int i2 = threadIdx.x; i2 < P * A; i2 += NT) {
            const int ql = i2 / A, s = i2 - ql * A;
            const int q = p0 + ql;
            if (q < NP) {
                const cplx<T> *Q = Z + size_t(2 * q) * zs;
                cplx<T> y[Bq];
                static_for<0, Bq>([&](auto nc) {
                    constexpr int n = decltype(nc)::value;
                    y[n] = Q[n * RS + s];
                });
                rdft<T, true, Bq>(y);
                T *ra = out + size_t(2 * (2 * q) + r) * N, *rb = out + size_t(2 * (2 * q + 1) + r) * N;
                static_for<0, Bq>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    gstore(ra + s + A * k, y[k].x);
                    gstore(rb + s + A * k, y[k].y);
                });
            }
        }
    }
}

// Fused forward transform of every source of one (target, layer) rate plane
// per block.  Per source the x1 row FFTs (two real rows per complex FFT)
// produce the half spectra of the rows, which are added into Z[n2][k1] in
// LDS, with the i·D1 prefactor where the target asks for it.  The x2 FFTs then
// run in place on Z, and the resolved modes are scaled by 1/(N1·N2) and added
// to the target.  Sources that take i·D2 after the transform form a second
// sum, transformed the same way and added afterwards.  Neither the row spectra
// nor the sums go to HBM.
template <typename T, int N, int KM = -1>
__device__ __forceinline__ void r2c_2d_reg(const Space<T> &S, const LaunchEntry &E, cplx<T> *smem) {
    constexpr int A = reg_fft_a(N), Bq = N / A, RS = A + 1, PS = Bq * RS;
    constexpr int P = r2c2d_pairs(N), NP = N / 2;
    const int job = blockIdx.x - E.block_begin;
    int tg = 0;
    while (job >= S.tgt_job0[tg + 1]) ++tg;
    const int layer = job - S.tgt_job0[tg];
    const int nk1 = KM >= 0 ? KM + 1 : S.nk1, zs = r2c2d_zs(N, nk1), k2m = KM >= 0 ? KM : S.k2m;
    const int K1 = KM >= 0 ? KM + 1 : S.K1, K2 = KM >= 0 ? 2 * KM + 1 : S.K2;
    const int nG = S.tgt_nG[tg], nH = S.tgt_nH[tg];
    cplx<T> *Z = smem;
    cplx<T> *X = smem + size_t(N) * zs;
    cplx<T> *TW = X + P * PS;
    cplx<T> *DF = TW + N;  // D1[k1] in .x
    const int tid = threadIdx.x;
    // the twiddle table for the x2 stages and the D1 factors of the row split
    // (visible after the first barrier)
    // (both stored to LDS behind the first round's row loads: square grids
    // share one twiddle table between the directions)
    constexpr int NTW = (N + kC2r2dThreads - 1) / kC2r2dThreads;
    cplx<T> twv[NTW];
    static_for<0, NTW>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int e = j * kC2r2dThreads + tid;
        twv[j] = e < N ? gload(S.tw + S.p2.tw + e) : czero<T>();
    });
    const T dfv = tid < nk1 ? T(gload(S.D1 + tid)) : T(0);
    static_assert(P * Bq <= kC2r2dThreads, "one row item per thread and round");
    bool tw_ready = false;  // block-uniform
    const T scale = T(1.0 / (double(N) * double(N)));
    cplx<T> *out = S.tgt_out[tg] + size_t(layer) * K1 * K2;

    // x1 transform of source `src` added into Z (first = overwrite)
    auto rows = [&](int src, bool dx, bool first) {
        const T *in = S.src_phys[src] + size_t(layer) * N * N;
        for (int p0 = 0; p0 < NP; p0 += P) {
            // stage 1: item (pair, t), A values of both rows at x = Bq·n + t
            {
                const bool have = tid < P * Bq;
                const int pl = (have ? tid : 0) / Bq, t = (have ? tid : 0) - pl * Bq;
                const int p = min(p0 + pl, NP - 1);
                const T *ra = in + size_t(2 * p) * N, *rb = ra + N;
                cplx<T> v[A];
                static_for<0, A>([&](auto nc) {
                    constexpr int n = decltype(nc)::value;
                    v[n] = have ? mk(gload(ra + Bq * n + t), gload(rb + Bq * n + t)) : czero<T>();
                });
                if (!tw_ready) {
                    static_for<0, NTW>([&](auto jc) {
                        constexpr int j = decltype(jc)::value;
                        const int e = j * kC2r2dThreads + tid;
                        if (e < N) TW[e] = twv[j];
                    });
                    if (tid < nk1) DF[tid] = mk(dfv, T(0));
                    __syncthreads();
                    tw_ready = true;
                }
                if (have) {
                    rdft<T, false, A>(v);
                    static_for<0, A>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        X[pl * PS + t * RS + k] = k == 0 ? v[0] : cmul(v[k], ltwiddle<T, false>(TW, t * k));
                    });
                }
            }
            __syncthreads();
            // stage 2 in place: spectrum index t + A·k of the pair at X[pl][k·RS + t]
            for (int it = tid; it < P * A; it += blockDim.x) {
                const int pl = it / A, t = it - pl * A;
                cplx<T> v[Bq];
                static_for<0, Bq>([&](auto nc) {
                    constexpr int n = decltype(nc)::value;
                    v[n] = X[pl * PS + n * RS + t];
                });
                rdft<T, false, Bq>(v);
                static_for<0, Bq>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    X[pl * PS + k * RS + t] = v[k];
                });
            }
            __syncthreads();
            // split the pair spectrum z = a + i·b into the half spectra of the rows
            for (int it = tid; it < P * nk1; it += blockDim.x) {
                const int pl = it / nk1, k = it - pl * nk1;
                const int p = p0 + pl;
                const int kn = k == 0 ? 0 : N - k;
                const cplx<T> zk = X[pl * PS + (k / A) * RS + k % A];
                const cplx<T> zn = X[pl * PS + (kn / A) * RS + kn % A];
                if (p < NP) {
                    cplx<T> ra = mk(T(0.5) * (zk.x + zn.x), T(0.5) * (zk.y - zn.y));
                    cplx<T> rb = mk(T(0.5) * (zk.y + zn.y), T(-0.5) * (zk.x - zn.x));
                    if (dx) {
                        const T d1 = DF[k].x;
                        ra = imul(d1, ra);
                        rb = imul(d1, rb);
                    }
                    cplx<T> *za = Z + size_t(2 * p) * zs + k, *zb = za + zs;
                    *za = first ? ra : *za + ra;
                    *zb = first ? rb : *zb + rb;
                }
            }
            __syncthreads();
        }
    };
    // x2 transform of Z in place; the resolved modes (times i·D2 with d2) are
    // scaled and added to the target
    auto columns = [&](bool d2) {
        const bool own1 = tid < nk1 * Bq;
        const int t1 = tid / nk1, c1 = tid - t1 * nk1;
        const int t2 = tid / nk1, c2 = tid - t2 * nk1;
        const bool own2 = tid < nk1 * A;
        // the target values this thread updates, loaded before the transform
        // so that their latency hides behind it (resolved modes only)
        cplx<T> old[Bq];
        static_for<0, Bq>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int pos = t2 + A * k;
            const bool lo = pos <= k2m, hi = pos >= N - k2m;
            const int i2 = lo ? pos : (hi ? K2 - (N - pos) : 0);
            old[k] = own2 ? gload(out + size_t(i2) * K1 + (own2 ? c2 : 0)) : czero<T>();
        });
        cplx<T> v[A];
        if (own1) {
            static_for<0, A>([&](auto nc) {
                constexpr int n = decltype(nc)::value;
                v[n] = Z[(Bq * n + t1) * zs + c1];
            });
        }
        __syncthreads();
        if (own1) {
            rdft<T, false, A>(v);
            static_for<0, A>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                Z[(t1 * A + k) * zs + c1] = k == 0 ? v[0] : cmul(v[k], ltwiddle<T, false>(TW, t1 * k));
            });
        }
        __syncthreads();
        if (own2) {
            cplx<T> w[Bq];
            static_for<0, Bq>([&](auto nc) {
                constexpr int n = decltype(nc)::value;
                w[n] = Z[(n * A + t2) * zs + c2];
            });
            rdft<T, false, Bq>(w);
            static_for<0, Bq>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const int pos = t2 + A * k;
                const bool lo = pos <= k2m, hi = pos >= N - k2m;
                if (lo || hi) {
                    const int i2 = lo ? pos : K2 - (N - pos);
                    const cplx<T> val = d2 ? imul(T(gload(S.D2 + i2)), w[k]) : w[k];
                    gstore(out + size_t(i2) * K1 + c2, old[k] + scale * val);
                }
            });
        }
        __syncthreads();
    };
    for (int s = 0; s < nG; ++s) rows(S.tgt_G[tg][s], S.tgt_Gd1[tg][s] != 0, s == 0);
    if (nG > 0) columns(false);
    for (int s = 0; s < nH; ++s) rows(S.tgt_H[tg][s], false, s == 0);
    if (nH > 0) columns(true);
}

// The whole-plane r2c for length N: the compile-time layout where the space
// has one (the c2r keeps the descriptor's sizes: compile-time sizes unroll
// its loops beyond the register budget of its 1024-thread blocks).
template <typename T, int N>
__device__ __forceinline__ void r2c_2d_any(const Space<T> &S, const LaunchEntry &E, cplx<T> *smem) {
#define BLD_KM(km)                                                        \
    if constexpr (xf2d_km_special(N, km)) {                               \
        if (xf2d_layout_is(S, km)) return r2c_2d_reg<T, N, km>(S, E, smem); \
    }
    BLD_KM(23) BLD_KM(31) BLD_KM(47) BLD_KM(63)
#undef BLD_KM
    r2c_2d_reg<T, N>(S, E, smem);
}

}  // namespace bld
