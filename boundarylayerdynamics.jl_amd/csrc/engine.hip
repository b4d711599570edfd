// Host side of the time-step engine: C ABI (include/bld.h), device memory,
// FFT plans, launch sequencing and HIP-graph capture of SSPRK33 steps.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "bld.h"
#include "kernels.hpp"

namespace bld {

static thread_local std::string g_create_error;

struct HipError {
    hipError_t code;
    std::string where;
};

#define HIPCHECK(expr)                                                     \
    do {                                                                   \
        hipError_t _e = (expr);                                            \
        if (_e != hipSuccess) throw HipError{_e, std::string(#expr)};      \
    } while (0)

struct BldError {
    int code;
    std::string msg;
};

// ---------------------------------------------------------------------------
// FFT plans and twiddles
// ---------------------------------------------------------------------------
static std::vector<int> factorize(int n) {
    std::vector<int> r;
    while (n % 4 == 0) {
        r.push_back(4);
        n /= 4;
    }
    for (int p : {2, 3, 5}) {
        while (n % p == 0) {
            r.push_back(p);
            n /= p;
        }
    }
    for (int p = 7; n > 1; p += 2) {
        while (n % p == 0) {
            r.push_back(p);
            n /= p;
        }
    }
    return r;
}

template <typename T>
struct TwiddleSet {
    std::map<int, int> offset;        // length -> offset
    std::vector<cplx<T>> host;
    int add(int n) {
        auto it = offset.find(n);
        if (it != offset.end()) return it->second;
        int off = int(host.size());
        for (int t = 0; t < n; ++t) {
            // exp(-2πi t/n) in extended precision, reduced to the first octant by symmetry
            long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
            host.push_back(mk(T(cosl(a)), T(-sinl(a))));
        }
        offset[n] = off;
        return off;
    }
};

static FftPlan make_plan(int n, int tw_offset) {
    FftPlan p{};
    p.n = n;
    std::vector<int> r = factorize(n);
    if (int(r.size()) > kMaxStages) throw BldError{BLD_EUNSUPPORTED, "transform length has too many factors"};
    p.nstages = int(r.size());
    for (size_t i = 0; i < r.size(); ++i) p.radix[i] = r[i];
    p.tw = tw_offset;
    return p;
}

// ---------------------------------------------------------------------------
// device buffer helper
// ---------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t b) {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = b;
        if (b) HIPCHECK(hipMalloc(&p, b));
    }
    template <typename U>
    U *as() const {
        return reinterpret_cast<U *>(p);
    }
};

template <typename U>
static void upload(DevBuf &d, const std::vector<U> &h) {
    d.alloc(h.size() * sizeof(U));
    if (!h.empty()) HIPCHECK(hipMemcpy(d.p, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice));
}

// ---------------------------------------------------------------------------
// engine interface
// ---------------------------------------------------------------------------
struct EngineBase {
    virtual ~EngineBase() = default;
    virtual void set_state(const double *const f[3]) = 0;
    virtual void get_state(double *const f[3]) = 0;
    virtual int rates(double *const f[3]) = 0;
    virtual void project() = 0;
    virtual int init_problem() = 0;
    virtual int step(double dt, int64_t n, bool graph) = 0;
    virtual void set_field_physical(int f, const double *v) = 0;
    virtual void get_field_physical(int f, double *v) = 0;
    virtual double time_steps(double dt, int64_t n, bool graph) = 0;
    virtual void set_profiling(bool on) = 0;
    virtual void kernel_time(int kind, double *ms, int64_t *count) = 0;
};

enum KernelKind {
    KK_LINEAR = 0, KK_C2R_X2, KK_C2R_X1, KK_ADV, KK_SGS_VISC, KK_SGS_STRESS, KK_R2C_X1, KK_R2C_X2,
    KK_STAGE, KK_CHECK, KK_COUNT
};

template <typename T>
class Engine : public EngineBase {
   public:
    explicit Engine(const bld_config &c) : cfg(c) { build(); }
    ~Engine() override {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (auto &ev : events) (void)hipEventDestroy(ev);
        if (st) (void)hipStreamDestroy(st);
    }

    // -- construction -----------------------------------------------------
    bld_config cfg;
    int K1 = 0, K2 = 0, nk = 0, n3 = 0;
    bool mixed = false;
    hipStream_t st = nullptr;
    std::vector<double> alpha_c, alpha_i, lc, li;
    DevBuf d_u[3], d_ut[3], d_du[3], d_u00[3], d_ut00[3], d_du00[3];
    DevBuf d_D1, d_D2, d_DD1, d_DD2, d_ac, d_ai, d_lc, d_li, d_gamma, d_pbuf, d_flag, d_tw, d_spaces;
    TwiddleSet<T> tw;
    std::vector<Space<T>> spaces;  // host copies
    std::vector<std::unique_ptr<DevBuf>> owned;
    int adv_idx = -1, sgs_idx = -1, xfer_idx = -1;
    DevBuf d_xfer_phys;
    LinDesc<T> lin{};
    ProjDesc<T> proj{};
    ProjDesc<double> proj00{};  // float64 (0,0) column of the mixed form
    DevBuf d_gamma0;
    // launch tables of the right-hand side
    struct Pass {
        LaunchTable lt{};
        int nblocks = 0;
        size_t smem = 0;
    };
    Pass rhs_pass[4];          // combined over adv+sgs for c2r_x2, c2r_x1, r2c_x1
    Pass r2c_x2_pass[2];       // adv, sgs separately (both accumulate into du)
    Pass xfer_pass[4];
    // graph
    hipGraphExec_t graph_exec = nullptr;
    double graph_dt = 0;
    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> events;
    std::vector<int> ev_kind;
    double kt_ms[KK_COUNT] = {};
    int64_t kt_n[KK_COUNT] = {};

    size_t nC() const { return size_t(nk) * n3; }
    size_t nI() const { return size_t(nk) * (n3 - 1); }
    size_t nf(int f) const { return f < 2 ? nC() : nI(); }

    DevBuf &own(size_t bytes) {
        owned.emplace_back(new DevBuf());
        owned.back()->alloc(bytes);
        return *owned.back();
    }

    void check_config() {
        const bld_config &c = cfg;
        if (c.abi_version != BLD_ABI_VERSION) throw BldError{BLD_EARG, "ABI version mismatch"};
        if (c.k1max < 0 || c.k2max < 0) throw BldError{BLD_EARG, "negative wavenumber range"};
        if (c.n3 < 3) throw BldError{BLD_EARG, "Due to boundary stencils, the grid needs at least 3 vertical layers"};
        if (!c.alpha_c || !c.alpha_i) throw BldError{BLD_EARG, "missing vertical scale factors"};
        for (int f = 0; f < 3; ++f)
            for (int s = 0; s < 2; ++s)
                if (c.bc[f][s].kind != BLD_BC_DIRICHLET && c.bc[f][s].kind != BLD_BC_NEUMANN)
                    throw BldError{BLD_EARG, "unknown boundary condition kind"};
        const bool w_dir = c.bc[2][0].kind == BLD_BC_DIRICHLET && c.bc[2][1].kind == BLD_BC_DIRICHLET;
        if (c.diffusion[2] && !w_dir)
            throw BldError{BLD_EUNSUPPORTED, "diffusion of vel3 supports Dirichlet boundary conditions only"};
        if (c.pressure) {
            if (!w_dir) throw BldError{BLD_EUNSUPPORTED, "The pressure solver only supports Dirichlet BCs"};
            if (c.bc[2][0].value != c.bc[2][1].value)
                throw BldError{BLD_EUNSUPPORTED, "The mean flow must be equal at both vertical boundaries."};
        }
        if (c.advection) {
            if (!w_dir) throw BldError{BLD_EUNSUPPORTED, "advection requires Dirichlet boundary conditions for vel3"};
            if (c.bc[2][0].value != 0 || c.bc[2][1].value != 0)
                throw BldError{BLD_EUNSUPPORTED, "Advection for non-zero w at boundary not implemented"};
            if (c.adv_n1 < 1 || c.adv_n2 < 1) throw BldError{BLD_EARG, "invalid advection grid"};
        }
        if (c.sgs) {
            if (!w_dir) throw BldError{BLD_EUNSUPPORTED, "the SGS model requires Dirichlet boundary conditions for vel3"};
            if (!c.sgs_lc || !c.sgs_li) throw BldError{BLD_EARG, "missing SGS length scales"};
            if (c.sgs_n1 < 1 || c.sgs_n2 < 1) throw BldError{BLD_EARG, "invalid SGS grid"};
        }
    }

    void build() {
        check_config();
        mixed = cfg.precision == BLD_FP32_MIXED;
        K1 = cfg.k1max + 1;
        K2 = 2 * cfg.k2max + 1;
        nk = K1 * K2;
        n3 = cfg.n3;
        alpha_c.assign(cfg.alpha_c, cfg.alpha_c + n3);
        alpha_i.assign(cfg.alpha_i, cfg.alpha_i + n3 + 1);
        HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        for (int f = 0; f < 3; ++f) {
            d_u[f].alloc(nf(f) * sizeof(cplx<T>));
            d_ut[f].alloc(nf(f) * sizeof(cplx<T>));
            d_du[f].alloc(nf(f) * sizeof(cplx<T>));
            HIPCHECK(hipMemset(d_u[f].p, 0, d_u[f].bytes));
            HIPCHECK(hipMemset(d_ut[f].p, 0, d_ut[f].bytes));
            HIPCHECK(hipMemset(d_du[f].p, 0, d_du[f].bytes));
            if (mixed) {
                for (DevBuf *b : {&d_u00[f], &d_ut00[f], &d_du00[f]}) {
                    b->alloc(size_t(n3) * sizeof(cplx<double>));
                    HIPCHECK(hipMemset(b->p, 0, b->bytes));
                }
            }
        }
        // derivative factors (Domains.jl:285-288, derivatives.jl:21-27)
        const double two_pi = 2 * M_PI;
        std::vector<double> D1(K1), D2(K2), DD1(K1), DD2(K2);
        for (int i = 0; i < K1; ++i) {
            D1[i] = i * (two_pi / cfg.L1);
            DD1[i] = -double(i) * i * ((two_pi / cfg.L1) * (two_pi / cfg.L1));
        }
        for (int j = 0; j < K2; ++j) {
            int k2 = j <= cfg.k2max ? j : j - K2;
            D2[j] = k2 * (two_pi / cfg.L2);
            DD2[j] = -double(k2) * k2 * ((two_pi / cfg.L2) * (two_pi / cfg.L2));
        }
        upload(d_D1, D1);
        upload(d_D2, D2);
        upload(d_DD1, DD1);
        upload(d_DD2, DD2);
        upload(d_ac, alpha_c);
        upload(d_ai, alpha_i);
        d_flag.alloc(sizeof(int));
        HIPCHECK(hipMemset(d_flag.p, 0, sizeof(int)));
        if (cfg.pressure) build_pressure();
        build_linear();
        build_proj();
        if (cfg.advection) adv_idx = add_space(SK_ADVECTION, cfg.adv_n1, cfg.adv_n2);
        if (cfg.sgs) {
            lc.assign(cfg.sgs_lc, cfg.sgs_lc + n3);
            li.assign(cfg.sgs_li, cfg.sgs_li + n3 - 1);
            upload(d_lc, lc);
            upload(d_li, li);
            sgs_idx = add_space(SK_SGS, cfg.sgs_n1, cfg.sgs_n2);
        }
        xfer_idx = add_space(SK_TRANSFER, 2 + 2 * cfg.k1max, 2 + 2 * cfg.k2max);
        upload(d_tw, tw.host);
        for (auto &s : spaces) s.tw = d_tw.as<cplx<T>>();
        d_spaces.alloc(spaces.size() * sizeof(Space<T>));
        sync_spaces();
        build_passes();
        HIPCHECK(hipDeviceSynchronize());
    }

    void sync_spaces() {
        HIPCHECK(hipMemcpy(d_spaces.p, spaces.data(), spaces.size() * sizeof(Space<T>), hipMemcpyHostToDevice));
    }

    // pressure factorisation (pressure.jl:231-296, 86-95) in float64
    void build_pressure() {
        std::vector<double> d3g3(n3, 0.0);
        for (int i = 1; i < n3; ++i) d3g3[i] -= alpha_i[i];
        for (int i = 0; i < n3 - 1; ++i) d3g3[i] -= alpha_i[i + 1];
        const double s1 = 1 / cfg.L1, s2 = 1 / cfg.L2;
        std::vector<double> gamma(size_t(n3) * nk), dv(n3);
        for (int m = 0; m < nk; ++m) {
            const int k1 = m % K1, j = m / K1;
            const int k2 = j <= cfg.k2max ? j : j - K2;
            if (k1 == 0 && k2 == 0) {
                dv = d3g3;
                dv[n3 - 1] = -3 * alpha_i[n3 - 1];
            } else {
                const double shift = 4 * M_PI * M_PI * (double(k1) * k1 * s1 * s1 + double(k2) * k2 * s2 * s2);
                for (int i = 0; i < n3; ++i) dv[i] = d3g3[i] - shift / alpha_c[i];
            }
            double g = 1 / dv[0];
            gamma[m] = g;
            for (int i = 1; i < n3; ++i) {
                const double b = alpha_i[i];
                g = 1 / (dv[i] - g * (b * b));
                gamma[size_t(i) * nk + m] = g;
            }
        }
        upload(d_gamma, gamma);
        std::vector<double> gamma0(n3);
        for (int i = 0; i < n3; ++i) gamma0[i] = gamma[size_t(i) * nk];
        upload(d_gamma0, gamma0);
        d_pbuf.alloc(size_t(n3) * nk * sizeof(cplx<double>));
    }

    void build_linear() {
        LinDesc<T> &L = lin;
        L.K1 = K1;
        L.K2 = K2;
        L.nk = nk;
        L.n3 = n3;
        L.DD1 = d_DD1.as<double>();
        L.DD2 = d_DD2.as<double>();
        L.alpha_c = d_ac.as<double>();
        L.alpha_i = d_ai.as<double>();
        for (int f = 0; f < 3; ++f) {
            L.has_diff[f] = cfg.diffusion[f];
            L.nu[f] = cfg.diffusivity[f];
            L.src[f] = cfg.source[f];
            for (int s = 0; s < 2; ++s) {
                L.bc_kind[f][s] = cfg.bc[f][s].kind;
                L.bc_val[f][s] = cfg.bc[f][s].value;
                L.bc_diff[f][s] = cfg.bc[f][s].difference;
            }
            L.state[0][f] = d_u[f].as<cplx<T>>();
            L.state[1][f] = d_ut[f].as<cplx<T>>();
            L.shadow[0][f] = mixed ? d_u00[f].as<cplx<double>>() : nullptr;
            L.shadow[1][f] = mixed ? d_ut00[f].as<cplx<double>>() : nullptr;
            L.du[f] = d_du[f].as<cplx<T>>();
            L.du00[f] = mixed ? d_du00[f].as<cplx<double>>() : nullptr;
        }
    }

    void build_proj() {
        ProjDesc<T> &D = proj;
        D.K1 = K1;
        D.K2 = K2;
        D.nk = nk;
        D.n3 = n3;
        D.has_pressure = cfg.pressure;
        D.D1 = d_D1.as<double>();
        D.D2 = d_D2.as<double>();
        D.alpha_c = d_ac.as<double>();
        D.alpha_i = d_ai.as<double>();
        D.gamma = cfg.pressure ? d_gamma.as<double>() : nullptr;
        D.wbc[0] = cfg.bc[2][0].value;
        D.wbc[1] = cfg.bc[2][1].value;
        D.pbuf = cfg.pressure ? d_pbuf.as<cplx<double>>() : nullptr;
        for (int f = 0; f < 3; ++f) {
            D.u[f] = d_u[f].as<cplx<T>>();
            D.ut[f] = d_ut[f].as<cplx<T>>();
            D.du[f] = d_du[f].as<cplx<T>>();
            D.u00[f] = mixed ? d_u00[f].as<cplx<double>>() : nullptr;
            D.ut00[f] = mixed ? d_ut00[f].as<cplx<double>>() : nullptr;
            D.du00[f] = mixed ? d_du00[f].as<cplx<double>>() : nullptr;
        }
        D.flag = d_flag.as<int>();
        D.ext_stride = 0;
        if (!mixed) return;
        ProjDesc<double> &Z = proj00;  // one column (k1 = k2 = 0) over the float64 shadows
        Z.K1 = Z.K2 = Z.nk = 1;
        Z.n3 = n3;
        Z.has_pressure = cfg.pressure;
        Z.D1 = d_D1.as<double>();  // D1[0] = D2[0] = 0
        Z.D2 = d_D2.as<double>();
        Z.alpha_c = d_ac.as<double>();
        Z.alpha_i = d_ai.as<double>();
        Z.gamma = cfg.pressure ? d_gamma0.as<double>() : nullptr;
        Z.wbc[0] = D.wbc[0];
        Z.wbc[1] = D.wbc[1];
        Z.pbuf = nullptr;
        for (int f = 0; f < 3; ++f) {
            Z.u[f] = d_u00[f].as<cplx<double>>();
            Z.ut[f] = d_ut00[f].as<cplx<double>>();
            Z.du[f] = d_du00[f].as<cplx<double>>();
            Z.u00[f] = Z.ut00[f] = Z.du00[f] = nullptr;
            Z.du_add[f] = reinterpret_cast<const cplx<float> *>(d_du[f].p);
            Z.mirror_u[f] = reinterpret_cast<cplx<float> *>(d_u[f].p);
            Z.mirror_ut[f] = reinterpret_cast<cplx<float> *>(d_ut[f].p);
        }
        Z.flag = d_flag.as<int>();
        Z.ext_stride = nk;
    }

    int add_space(int kind, int N1, int N2) {
        Space<T> S;
        std::memset(&S, 0, sizeof(S));
        S.kind = kind;
        S.N1 = N1;
        S.N2 = N2;
        S.K1 = K1;
        S.K2 = K2;
        S.k1m = std::min(N1 / 2 + 1, K1) - 1;
        S.k2m = (std::min(N2, K2) - 1) / 2;
        S.nk1 = S.k1m + 1;
        S.n3 = n3;
        S.p1 = make_plan(N1, tw.add(N1));
        S.p2 = make_plan(N2, tw.add(N2));
        S.D1 = d_D1.as<double>();
        S.D2 = d_D2.as<double>();
        S.alpha_c = d_ac.as<double>();
        S.alpha_i = d_ai.as<double>();
        for (int s = 0; s < 2; ++s)
            for (int f = 0; f < 3; ++f) S.state[s][f] = (s == 0 ? d_u[f] : d_ut[f]).template as<cplx<T>>();
        for (int f = 0; f < 3; ++f) S.tgt_out[f] = d_du[f].as<cplx<T>>();
        const size_t P = size_t(N1) * N2;
        auto set_terms = [&](std::vector<std::pair<int, int>> terms) {  // (recipe, is_I)
            S.nin = int(terms.size());
            int j = 0;
            for (int t = 0; t < S.nin; ++t) {
                S.in_recipe[t] = terms[t].first;
                S.in_layers[t] = terms[t].second ? n3 - 1 : n3;
                S.in_job0[t] = j;
                j += S.in_layers[t];
                S.in_phys[t] = own(P * S.in_layers[t] * sizeof(T)).template as<T>();
            }
            S.in_job0[S.nin] = j;
        };
        auto set_sources = [&](std::vector<int> isI) {
            S.nsrc = int(isI.size());
            int j = 0;
            for (int s = 0; s < S.nsrc; ++s) {
                S.src_layers[s] = isI[s] ? n3 - 1 : n3;
                S.src_job0[s] = j;
                j += S.src_layers[s];
                S.src_phys[s] = own(P * S.src_layers[s] * sizeof(T)).template as<T>();
            }
            S.src_job0[S.nsrc] = j;
        };
        auto set_targets = [&](int ntgt) {
            S.ntgt = ntgt;
            int j = 0;
            for (int t = 0; t < ntgt; ++t) {
                S.tgt_layers[t] = t < 2 ? n3 : n3 - 1;
                S.tgt_job0[t] = j;
                j += S.tgt_layers[t];
            }
            S.tgt_job0[ntgt] = j;
        };
        if (kind == SK_ADVECTION) {
            set_terms({{R_U, 0}, {R_V, 0}, {R_W, 1}, {R_VORT1, 1}, {R_VORT2, 1}, {R_VORT3, 0}});
            set_sources({0, 0, 1});
            set_targets(3);
            for (int t = 0; t < 3; ++t) {
                S.tgt_nG[t] = 1;
                S.tgt_G[t][0] = t;
                S.tgt_Gd1[t][0] = 0;
                S.tgt_nH[t] = 0;
            }
        } else if (kind == SK_SGS) {
            set_terms({{R_U, 0}, {R_V, 0}, {R_W, 1}, {R_D1U, 0}, {R_D2V, 0}, {R_D2U, 0}, {R_D1V, 0}, {R_D1W, 1},
                       {R_D2W, 1}});
            // R1, R2 (C), R3 (I), T11, T22, T12 (C), T13, T23 (I)
            set_sources({0, 0, 1, 0, 0, 0, 1, 1});
            set_targets(3);
            // du1 += FFT(R1 + D1·T11) + D2·FFT(T12)
            S.tgt_nG[0] = 2; S.tgt_G[0][0] = 0; S.tgt_Gd1[0][0] = 0; S.tgt_G[0][1] = 3; S.tgt_Gd1[0][1] = 1;
            S.tgt_nH[0] = 1; S.tgt_H[0][0] = 5;
            // du2 += FFT(R2 + D1·T12) + D2·FFT(T22)
            S.tgt_nG[1] = 2; S.tgt_G[1][0] = 1; S.tgt_Gd1[1][0] = 0; S.tgt_G[1][1] = 5; S.tgt_Gd1[1][1] = 1;
            S.tgt_nH[1] = 1; S.tgt_H[1][0] = 4;
            // du3 += FFT(R3 + D1·T13) + D2·FFT(T23)
            S.tgt_nG[2] = 2; S.tgt_G[2][0] = 2; S.tgt_Gd1[2][0] = 0; S.tgt_G[2][1] = 6; S.tgt_Gd1[2][1] = 1;
            S.tgt_nH[2] = 1; S.tgt_H[2][0] = 7;
            S.visc_c = own(P * n3 * sizeof(T)).template as<T>();
            S.visc_i = own(P * (n3 - 1) * sizeof(T)).template as<T>();
            S.lc = d_lc.as<double>();
            S.li = d_li.as<double>();
            for (int s = 0; s < 2; ++s) {
                S.wall_rough[s] = cfg.sgs_wall[s] == BLD_WALL_ROUGH;
                S.fstrain[s] = cfg.sgs_factor_strain[s];
                S.fstress[s] = cfg.sgs_factor_stress[s];
            }
            S.wbc[0] = cfg.bc[2][0].value;
            S.wbc[1] = cfg.bc[2][1].value;
        } else {  // transfer space: one term / one source, set per call
            S.nin = 1;
            S.in_recipe[0] = R_FIELD;
            S.in_layers[0] = n3;
            S.in_job0[0] = 0;
            S.in_job0[1] = n3;
            d_xfer_phys.alloc(P * n3 * sizeof(T));
            S.in_phys[0] = d_xfer_phys.as<T>();
            S.nsrc = 1;
            S.src_layers[0] = n3;
            S.src_job0[0] = 0;
            S.src_job0[1] = n3;
            S.src_phys[0] = d_xfer_phys.as<T>();
            S.ntgt = 1;
            S.tgt_layers[0] = n3;
            S.tgt_job0[0] = 0;
            S.tgt_job0[1] = n3;
            S.tgt_nG[0] = 1;
            S.tgt_G[0][0] = 0;
            S.tgt_Gd1[0][0] = 0;
            S.tgt_nH[0] = 0;
        }
        const size_t jobs = std::max<size_t>(S.in_job0[S.nin], S.src_job0[S.nsrc]);
        S.inter = own(jobs * size_t(N2) * S.nk1 * sizeof(cplx<T>)).template as<cplx<T>>();
        spaces.push_back(S);
        return int(spaces.size()) - 1;
    }

    // batch sizes: keep one block's LDS at <= 64 KiB so several blocks share a CU
    static int pick_batch(int n, int bufs, int avail) {
        const size_t limit = 64 * 1024;
        int b = int(limit / (size_t(bufs) * n * sizeof(cplx<T>)));
        b = std::max(1, std::min(b, 16));
        b = std::min(b, std::max(1, avail));
        int p2 = 1;  // the kernels split indices with shifts: keep B a power of two
        while (p2 * 2 <= b) p2 *= 2;
        return p2;
    }

    void add_entry(Pass &ps, int kind, int sidx) {
        const Space<T> &S = spaces[sidx];
        LaunchEntry E{};
        E.space = sidx;
        int n = 0, bufs = 2, avail = 0, njobs = 0;
        switch (kind) {
            case PASS_C2R_X2: n = S.N2; avail = S.nk1; njobs = S.in_job0[S.nin]; break;
            case PASS_C2R_X1: n = S.N1; avail = (S.N2 + 1) / 2; njobs = S.in_job0[S.nin]; break;
            case PASS_R2C_X1: n = S.N1; avail = (S.N2 + 1) / 2; njobs = S.src_job0[S.nsrc]; break;
            case PASS_R2C_X2: {
                n = S.N2;
                avail = S.nk1;
                njobs = S.tgt_job0[S.ntgt];
                bool anyH = false;
                for (int t = 0; t < S.ntgt; ++t) anyH |= S.tgt_nH[t] > 0;
                bufs = anyH ? 4 : 2;
                break;
            }
        }
        E.batch = pick_batch(n, bufs, avail);
        E.blocks_per_job = (avail + E.batch - 1) / E.batch;
        E.njobs = njobs;
        E.block_begin = ps.nblocks;
        if (ps.lt.n >= kMaxEntries) throw BldError{BLD_EARG, "too many spaces in one launch"};
        ps.lt.e[ps.lt.n++] = E;
        ps.nblocks += E.blocks_per_job * njobs;
        ps.smem = std::max(ps.smem, size_t(bufs) * E.batch * n * sizeof(cplx<T>));
    }

    void build_passes() {
        for (int k = 0; k < 4; ++k) {
            rhs_pass[k] = Pass{};
            for (int s : {adv_idx, sgs_idx})
                if (s >= 0 && k != PASS_R2C_X2) add_entry(rhs_pass[k], k, s);
            xfer_pass[k] = Pass{};
            add_entry(xfer_pass[k], k, xfer_idx);
        }
        r2c_x2_pass[0] = Pass{};
        r2c_x2_pass[1] = Pass{};
        if (adv_idx >= 0) add_entry(r2c_x2_pass[0], PASS_R2C_X2, adv_idx);
        if (sgs_idx >= 0) add_entry(r2c_x2_pass[1], PASS_R2C_X2, sgs_idx);
        for (Pass *p : {&rhs_pass[0], &rhs_pass[1], &rhs_pass[2], &rhs_pass[3], &r2c_x2_pass[0], &r2c_x2_pass[1],
                        &xfer_pass[0], &xfer_pass[1], &xfer_pass[2], &xfer_pass[3]})
            if (p->smem > 160 * 1024) throw BldError{BLD_EUNSUPPORTED, "transform length too large for LDS"};
    }

    // -- launch helpers with optional per-kernel timing ---------------------
    template <typename F>
    void timed(int kind, F &&f) {
        if (!profiling) {
            f();
            return;
        }
        hipEvent_t a, b;
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        HIPCHECK(hipEventRecord(a, st));
        f();
        HIPCHECK(hipEventRecord(b, st));
        events.push_back(a);
        events.push_back(b);
        ev_kind.push_back(kind);
    }

    void collect_profile() {
        HIPCHECK(hipStreamSynchronize(st));
        for (size_t i = 0; i < ev_kind.size(); ++i) {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, events[2 * i], events[2 * i + 1]));
            kt_ms[ev_kind[i]] += ms;
            kt_n[ev_kind[i]] += 1;
        }
        for (auto &ev : events) (void)hipEventDestroy(ev);
        events.clear();
        ev_kind.clear();
    }

    void run_pass(int kind, const Pass &p, int sel, int kk) {
        if (p.nblocks == 0) return;
        timed(kk, [&] { launch_transform<T>(kind, d_spaces.as<Space<T>>(), p.lt, p.nblocks, p.smem, sel, st); });
    }

    // compute_rates! of state `sel` (0 = u, 1 = utmp) into du
    void rhs(int sel) {
        timed(KK_LINEAR, [&] { launch_linear<T>(lin, sel, st); });
        if (adv_idx < 0 && sgs_idx < 0) return;
        run_pass(PASS_C2R_X2, rhs_pass[PASS_C2R_X2], sel, KK_C2R_X2);
        run_pass(PASS_C2R_X1, rhs_pass[PASS_C2R_X1], sel, KK_C2R_X1);
        const Space<T> *dsp = d_spaces.as<Space<T>>();
        if (adv_idx >= 0) {
            const Space<T> &S = spaces[adv_idx];
            timed(KK_ADV, [&] { launch_physics<T>(PHYS_ADV, dsp + adv_idx, size_t(S.N1) * S.N2 * n3, st); });
        }
        if (sgs_idx >= 0) {
            const Space<T> &S = spaces[sgs_idx];
            const size_t np = size_t(S.N1) * S.N2 * n3;
            timed(KK_SGS_VISC, [&] { launch_physics<T>(PHYS_SGS_VISC, dsp + sgs_idx, np, st); });
            timed(KK_SGS_STRESS, [&] { launch_physics<T>(PHYS_SGS_STRESS, dsp + sgs_idx, np, st); });
        }
        run_pass(PASS_R2C_X1, rhs_pass[PASS_R2C_X1], sel, KK_R2C_X1);
        run_pass(PASS_R2C_X2, r2c_x2_pass[0], sel, KK_R2C_X2);
        run_pass(PASS_R2C_X2, r2c_x2_pass[1], sel, KK_R2C_X2);
    }

    void stage(int s, double dt) {
        timed(KK_STAGE, [&] { launch_stage_project<T>(proj, s, dt, st, mixed ? &proj00 : nullptr); });
    }

    void check_du() {
        timed(KK_CHECK, [&] {
            for (int f = 0; f < 3; ++f) launch_check_finite<T>(d_du[f].p, nf(f), d_flag.as<int>(), st);
        });
    }

    void one_step(double dt) {
        stage(1, dt);
        rhs(1);
        stage(2, dt);
        rhs(1);
        stage(3, dt);
        rhs(0);
    }

    int read_flag() {
        int flag = 0;
        HIPCHECK(hipMemcpyAsync(&flag, d_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return flag;
    }
    void clear_flag() { HIPCHECK(hipMemsetAsync(d_flag.p, 0, sizeof(int), st)); }

    // -- state transfer -------------------------------------------------------
    void set_state(const double *const f[3]) override {
        for (int c = 0; c < 3; ++c) {
            if (!f[c]) continue;
            const size_t n = nf(c);
            std::vector<cplx<T>> h(n);
            for (size_t i = 0; i < n; ++i) h[i] = mk(T(f[c][2 * i]), T(f[c][2 * i + 1]));
            HIPCHECK(hipMemcpy(d_u[c].p, h.data(), n * sizeof(cplx<T>), hipMemcpyHostToDevice));
            if (mixed) {
                const int nl = c < 2 ? n3 : n3 - 1;
                std::vector<cplx<double>> s(n3, mk(0.0, 0.0));
                for (int l = 0; l < nl; ++l) s[l] = mk(f[c][2 * size_t(l) * nk], f[c][2 * size_t(l) * nk + 1]);
                HIPCHECK(hipMemcpy(d_u00[c].p, s.data(), n3 * sizeof(cplx<double>), hipMemcpyHostToDevice));
            }
        }
    }

    void download(const DevBuf (&src)[3], const DevBuf (&s00)[3], bool add00, double *const f[3]) {
        for (int c = 0; c < 3; ++c) {
            if (!f[c]) continue;
            const size_t n = nf(c);
            std::vector<cplx<T>> h(n);
            HIPCHECK(hipMemcpy(h.data(), src[c].p, n * sizeof(cplx<T>), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) {
                f[c][2 * i] = double(h[i].x);
                f[c][2 * i + 1] = double(h[i].y);
            }
            if (mixed) {
                const int nl = c < 2 ? n3 : n3 - 1;
                std::vector<cplx<double>> s(n3);
                HIPCHECK(hipMemcpy(s.data(), s00[c].p, n3 * sizeof(cplx<double>), hipMemcpyDeviceToHost));
                for (int l = 0; l < nl; ++l) {
                    const size_t i = size_t(l) * nk;
                    if (add00) {
                        f[c][2 * i] += s[l].x;
                        f[c][2 * i + 1] += s[l].y;
                    } else {
                        f[c][2 * i] = s[l].x;
                        f[c][2 * i + 1] = s[l].y;
                    }
                }
            }
        }
    }

    void get_state(double *const f[3]) override { download(d_u, d_u00, false, f); }

    int rates(double *const f[3]) override {
        clear_flag();
        rhs(0);
        check_du();
        const int flag = read_flag();
        download(d_du, d_du00, true, f);
        return flag ? BLD_EDIVERGED : BLD_OK;
    }

    void project() override {
        stage(0, 0.0);
        HIPCHECK(hipStreamSynchronize(st));
    }

    int init_problem() override {
        clear_flag();
        stage(0, 0.0);
        rhs(0);
        check_du();
        return read_flag() ? BLD_EDIVERGED : BLD_OK;
    }

    void capture(double dt) {
        if (graph_exec && graph_dt == dt) return;
        if (graph_exec) {
            HIPCHECK(hipGraphExecDestroy(graph_exec));
            graph_exec = nullptr;
        }
        hipGraph_t g;
        const bool prof = profiling;
        profiling = false;
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        one_step(dt);
        HIPCHECK(hipStreamEndCapture(st, &g));
        profiling = prof;
        HIPCHECK(hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0));
        HIPCHECK(hipGraphDestroy(g));
        graph_dt = dt;
    }

    int step(double dt, int64_t n, bool graph) override {
        clear_flag();
        if (graph) capture(dt);
        for (int64_t i = 0; i < n; ++i) {
            if (graph)
                HIPCHECK(hipGraphLaunch(graph_exec, st));
            else
                one_step(dt);
        }
        check_du();
        if (profiling) collect_profile();
        return read_flag() ? BLD_EDIVERGED : BLD_OK;
    }

    double time_steps(double dt, int64_t n, bool graph) override {
        if (graph) capture(dt);
        hipEvent_t a, b;
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        HIPCHECK(hipEventRecord(a, st));
        for (int64_t i = 0; i < n; ++i) {
            if (graph)
                HIPCHECK(hipGraphLaunch(graph_exec, st));
            else
                one_step(dt);
        }
        HIPCHECK(hipEventRecord(b, st));
        HIPCHECK(hipEventSynchronize(b));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (profiling) collect_profile();
        return ms;
    }

    void set_profiling(bool on) override {
        profiling = on;
        for (int k = 0; k < KK_COUNT; ++k) {
            kt_ms[k] = 0;
            kt_n[k] = 0;
        }
    }
    void kernel_time(int kind, double *ms, int64_t *count) override {
        if (kind < 0 || kind >= KK_COUNT) throw BldError{BLD_EARG, "unknown kernel kind"};
        *ms = kt_ms[kind];
        *count = kt_n[kind];
    }

    // -- physical-space transfer at the default size ----------------------
    void set_field_physical(int f, const double *v) override {
        if (f < 0 || f > 2) throw BldError{BLD_EARG, "field index must be 0, 1 or 2"};
        Space<T> &S = spaces[xfer_idx];
        const int nl = f < 2 ? n3 : n3 - 1;
        const size_t P = size_t(S.N1) * S.N2;
        std::vector<T> h(P * nl);
        for (size_t i = 0; i < h.size(); ++i) h[i] = T(v[i]);
        HIPCHECK(hipMemcpy(d_xfer_phys.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        S.src_layers[0] = nl;
        S.src_job0[1] = nl;
        S.tgt_layers[0] = nl;
        S.tgt_job0[1] = nl;
        S.tgt_out[0] = d_u[f].as<cplx<T>>();
        sync_spaces();
        rebuild_xfer();
        launch_fill<T>(d_u[f].p, nf(f), st);
        run_pass(PASS_R2C_X1, xfer_pass[PASS_R2C_X1], 0, KK_R2C_X1);
        run_pass(PASS_R2C_X2, xfer_pass[PASS_R2C_X2], 0, KK_R2C_X2);
        if (mixed) {  // float64 mean of each layer for the (0,0) shadow
            std::vector<cplx<double>> s(n3, mk(0.0, 0.0));
            std::vector<cplx<T>> h00(nl);
            for (int l = 0; l < nl; ++l) {
                double sum = 0;
                for (size_t i = 0; i < P; ++i) sum += v[size_t(l) * P + i];
                s[l] = mk(sum * (1.0 / double(P)), 0.0);
            }
            HIPCHECK(hipStreamSynchronize(st));
            HIPCHECK(hipMemcpy(d_u00[f].p, s.data(), n3 * sizeof(cplx<double>), hipMemcpyHostToDevice));
            for (int l = 0; l < nl; ++l) {
                cplx<T> z = mk(T(s[l].x), T(0));
                HIPCHECK(hipMemcpy(d_u[f].as<cplx<T>>() + size_t(l) * nk, &z, sizeof(z), hipMemcpyHostToDevice));
            }
        }
        HIPCHECK(hipStreamSynchronize(st));
    }

    void get_field_physical(int f, double *v) override {
        if (f < 0 || f > 2) throw BldError{BLD_EARG, "field index must be 0, 1 or 2"};
        Space<T> &S = spaces[xfer_idx];
        const int nl = f < 2 ? n3 : n3 - 1;
        S.in_layers[0] = nl;
        S.in_job0[1] = nl;
        S.in_field[0] = d_u[f].as<cplx<T>>();
        sync_spaces();
        rebuild_xfer();
        run_pass(PASS_C2R_X2, xfer_pass[PASS_C2R_X2], 0, KK_C2R_X2);
        run_pass(PASS_C2R_X1, xfer_pass[PASS_C2R_X1], 0, KK_C2R_X1);
        const size_t n = size_t(S.N1) * S.N2 * nl;
        std::vector<T> h(n);
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipMemcpy(h.data(), d_xfer_phys.p, n * sizeof(T), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) v[i] = double(h[i]);
    }

    void rebuild_xfer() {
        for (int k = 0; k < 4; ++k) {
            xfer_pass[k] = Pass{};
            add_entry(xfer_pass[k], k, xfer_idx);
        }
    }
};

// ---------------------------------------------------------------------------
// standalone batched LDLt (DistributedBatchLDLt + ldiv!)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_ldlt64(const double *__restrict__ gamma, const double *__restrict__ beta,
                                               cplx<double> *__restrict__ b, int nk, int nz) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nk) return;
    cplx<double> prev = b[m];
    double g = gamma[m];
    prev = mk(g * prev.x, g * prev.y);
    b[m] = prev;
    for (int i = 1; i < nz; ++i) {
        const size_t o = size_t(i) * nk + m;
        const double bt = beta[size_t(i - 1) * nk + m];
        g = gamma[o];
        const cplx<double> v = b[o];
        prev = mk(g * (v.x - bt * prev.x), g * (v.y - bt * prev.y));
        b[o] = prev;
    }
    cplx<double> next = prev;
    for (int i = nz - 2; i >= 0; --i) {
        const size_t o = size_t(i) * nk + m;
        const double gb = gamma[o] * beta[o];
        cplx<double> v = b[o];
        v = mk(v.x - gb * next.x, v.y - gb * next.y);
        b[o] = v;
        next = v;
    }
}

// float32 elimination with one float64-residual refinement step
__global__ void __launch_bounds__(64) k_ldlt32_ir(const float *__restrict__ gamma32, const double *__restrict__ dv,
                                                  const double *__restrict__ beta, const cplx<float> *__restrict__ b,
                                                  cplx<float> *__restrict__ x, cplx<float> *__restrict__ r32, int nk,
                                                  int nz) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nk) return;
    auto solve32 = [&](const cplx<float> *rhs, cplx<float> *out) {
        cplx<float> prev = rhs[m];
        float g = gamma32[m];
        prev = mk(g * prev.x, g * prev.y);
        out[m] = prev;
        for (int i = 1; i < nz; ++i) {
            const size_t o = size_t(i) * nk + m;
            const float bt = float(beta[size_t(i - 1) * nk + m]);
            g = gamma32[o];
            const cplx<float> v = rhs[o];
            prev = mk(g * (v.x - bt * prev.x), g * (v.y - bt * prev.y));
            out[o] = prev;
        }
        cplx<float> next = prev;
        for (int i = nz - 2; i >= 0; --i) {
            const size_t o = size_t(i) * nk + m;
            const float gb = gamma32[o] * float(beta[o]);
            cplx<float> v = out[o];
            v = mk(v.x - gb * next.x, v.y - gb * next.y);
            out[o] = v;
            next = v;
        }
    };
    solve32(b, x);
    // residual r = b - A x in float64, stored in float32
    cplx<double> xm = mk(0.0, 0.0), x0 = mk(double(x[m].x), double(x[m].y));
    for (int i = 0; i < nz; ++i) {
        const size_t o = size_t(i) * nk + m;
        cplx<double> xp = mk(0.0, 0.0);
        if (i + 1 < nz) xp = mk(double(x[o + nk].x), double(x[o + nk].y));
        const double d = dv[o];
        const double bl = i > 0 ? beta[size_t(i - 1) * nk + m] : 0.0;
        const double bu = i + 1 < nz ? beta[o] : 0.0;
        const cplx<float> bb = b[o];
        const double rx = double(bb.x) - (d * x0.x + bl * xm.x + bu * xp.x);
        const double ry = double(bb.y) - (d * x0.y + bl * xm.y + bu * xp.y);
        r32[o] = mk(float(rx), float(ry));
        xm = x0;
        x0 = xp;
    }
    solve32(r32, r32);
    for (int i = 0; i < nz; ++i) {
        const size_t o = size_t(i) * nk + m;
        x[o] = mk(x[o].x + r32[o].x, x[o].y + r32[o].y);
    }
}

struct Ldlt {
    int64_t nk = 0, nz = 0;
    int precision = 0;
    DevBuf gamma, gamma32, beta, dv, b64, b32, x32, r32;
    std::string err;
};

}  // namespace bld

using namespace bld;

struct bld_engine {
    std::unique_ptr<EngineBase> impl;
    std::string err;
};

struct bld_ldlt {
    Ldlt s;
};

template <typename F>
static int guarded(std::string &err, F &&f) {
    try {
        return f();
    } catch (const BldError &e) {
        err = e.msg;
        return e.code;
    } catch (const HipError &e) {
        err = std::string("HIP error: ") + hipGetErrorString(e.code) + " in " + e.where;
        return BLD_EHIP;
    } catch (const std::bad_alloc &) {
        err = "out of host memory";
        return BLD_ENOMEM;
    } catch (...) {
        err = "unknown error";
        return BLD_EARG;
    }
}

extern "C" {

int bld_create(const bld_config *cfg, bld_engine **out) {
    if (!cfg || !out) {
        g_create_error = "null argument";
        return BLD_EARG;
    }
    *out = nullptr;
    return guarded(g_create_error, [&] {
        std::unique_ptr<bld_engine> e(new bld_engine());
        if (cfg->precision == BLD_FP64)
            e->impl.reset(new Engine<double>(*cfg));
        else if (cfg->precision == BLD_FP32_MIXED)
            e->impl.reset(new Engine<float>(*cfg));
        else
            throw BldError{BLD_EARG, "unknown precision"};
        *out = e.release();
        return BLD_OK;
    });
}

void bld_destroy(bld_engine *e) { delete e; }

const char *bld_last_error(const bld_engine *e) { return e ? e->err.c_str() : g_create_error.c_str(); }

#define ENGINE_CALL(...)                                              \
    if (!e) return BLD_EARG;                                          \
    return guarded(e->err, [&]() -> int { __VA_ARGS__; });

int bld_set_state(bld_engine *e, const double *vel1, const double *vel2, const double *vel3) {
    ENGINE_CALL({
        const double *f[3] = {vel1, vel2, vel3};
        e->impl->set_state(f);
        return BLD_OK;
    })
}

int bld_get_state(bld_engine *e, double *vel1, double *vel2, double *vel3) {
    ENGINE_CALL({
        double *f[3] = {vel1, vel2, vel3};
        e->impl->get_state(f);
        return BLD_OK;
    })
}

int bld_rates(bld_engine *e, double *du1, double *du2, double *du3) {
    ENGINE_CALL({
        double *f[3] = {du1, du2, du3};
        int r = e->impl->rates(f);
        if (r == BLD_EDIVERGED) e->err = "The simulation has diverged.";
        return r;
    })
}

int bld_project(bld_engine *e) {
    ENGINE_CALL({
        e->impl->project();
        return BLD_OK;
    })
}

int bld_init_problem(bld_engine *e, double t1) {
    (void)t1;  // processes are time-independent (docs/src/time-evolution.md:21-23)
    ENGINE_CALL({
        int r = e->impl->init_problem();
        if (r == BLD_EDIVERGED) e->err = "The simulation has diverged.";
        return r;
    })
}

static void advance_time(double *t, double dt, int64_t n, double t_end) {
    if (!t) return;
    for (int64_t i = 0; i < n; ++i) {
        const double next = *t + dt;
        const double tol = std::sqrt(2.220446049250313e-16) * std::max(std::fabs(next), std::fabs(t_end));
        *t = (t_end > *t && std::fabs(next - t_end) <= tol) ? t_end : next;
    }
}

int bld_step(bld_engine *e, double dt, int64_t nsteps, double *t, double t_end) {
    ENGINE_CALL({
        if (nsteps < 0) throw BldError{BLD_EARG, "negative step count"};
        int r = e->impl->step(dt, nsteps, false);
        advance_time(t, dt, nsteps, t_end);
        if (r == BLD_EDIVERGED) e->err = "The simulation has diverged.";
        return r;
    })
}

int bld_step_graph(bld_engine *e, double dt, int64_t nsteps, double *t, double t_end) {
    ENGINE_CALL({
        if (nsteps < 0) throw BldError{BLD_EARG, "negative step count"};
        int r = e->impl->step(dt, nsteps, true);
        advance_time(t, dt, nsteps, t_end);
        if (r == BLD_EDIVERGED) e->err = "The simulation has diverged.";
        return r;
    })
}

int bld_set_field_physical(bld_engine *e, int field, const double *values) {
    ENGINE_CALL({
        if (!values) throw BldError{BLD_EARG, "null values"};
        e->impl->set_field_physical(field, values);
        return BLD_OK;
    })
}

int bld_get_field_physical(bld_engine *e, int field, double *values) {
    ENGINE_CALL({
        if (!values) throw BldError{BLD_EARG, "null values"};
        e->impl->get_field_physical(field, values);
        return BLD_OK;
    })
}

int bld_time_steps(bld_engine *e, double dt, int64_t nsteps, int use_graph, double *ms) {
    ENGINE_CALL({
        if (!ms || nsteps < 0) throw BldError{BLD_EARG, "invalid arguments"};
        *ms = e->impl->time_steps(dt, nsteps, use_graph != 0);
        return BLD_OK;
    })
}

int bld_set_profiling(bld_engine *e, int enabled) {
    ENGINE_CALL({
        e->impl->set_profiling(enabled != 0);
        return BLD_OK;
    })
}

int bld_kernel_times(bld_engine *e, int kind, double *ms_total, int64_t *launches) {
    ENGINE_CALL({
        if (!ms_total || !launches) throw BldError{BLD_EARG, "null output"};
        e->impl->kernel_time(kind, ms_total, launches);
        return BLD_OK;
    })
}

int bld_ldlt_create(int64_t nk, int64_t nz, const double *dvs, const double *evs, int precision, bld_ldlt **out) {
    if (!out) return BLD_EARG;
    *out = nullptr;
    return guarded(g_create_error, [&] {
        if (nk < 1 || nz < 1 || !dvs || (nz > 1 && !evs)) throw BldError{BLD_EARG, "invalid LDLt arguments"};
        if (precision != BLD_FP64 && precision != BLD_FP32_MIXED) throw BldError{BLD_EARG, "unknown precision"};
        std::unique_ptr<bld_ldlt> h(new bld_ldlt());
        Ldlt &s = h->s;
        s.nk = nk;
        s.nz = nz;
        s.precision = precision;
        const size_t n = size_t(nk) * nz;
        std::vector<double> g(n), beta(n, 0.0);
        for (int64_t i = 0; i + 1 < nz; ++i)
            for (int64_t m = 0; m < nk; ++m) beta[size_t(i) * nk + m] = evs[size_t(i) * nk + m];
        for (int64_t m = 0; m < nk; ++m) {  // pressure.jl:91-95
            double gg = 1 / dvs[m];
            g[m] = gg;
            for (int64_t i = 1; i < nz; ++i) {
                const double b = beta[size_t(i - 1) * nk + m];
                gg = 1 / (dvs[size_t(i) * nk + m] - gg * (b * b));
                g[size_t(i) * nk + m] = gg;
            }
        }
        upload(s.gamma, g);
        upload(s.beta, beta);
        if (precision == BLD_FP32_MIXED) {
            std::vector<float> g32(n);
            for (size_t i = 0; i < n; ++i) g32[i] = float(g[i]);
            upload(s.gamma32, g32);
            std::vector<double> dv(dvs, dvs + n);
            upload(s.dv, dv);
            s.b32.alloc(n * sizeof(cplx<float>));
            s.x32.alloc(n * sizeof(cplx<float>));
            s.r32.alloc(n * sizeof(cplx<float>));
        } else {
            s.b64.alloc(n * sizeof(cplx<double>));
        }
        *out = h.release();
        return BLD_OK;
    });
}

void bld_ldlt_destroy(bld_ldlt *s) { delete s; }

const char *bld_ldlt_last_error(const bld_ldlt *s) { return s ? s->s.err.c_str() : g_create_error.c_str(); }

static void ldlt_run(Ldlt &s, hipStream_t st) {
    const int nb = int((s.nk + 63) / 64);
    if (s.precision == BLD_FP64)
        k_ldlt64<<<nb, 64, 0, st>>>(s.gamma.as<double>(), s.beta.as<double>(), s.b64.as<cplx<double>>(), int(s.nk),
                                    int(s.nz));
    else
        k_ldlt32_ir<<<nb, 64, 0, st>>>(s.gamma32.as<float>(), s.dv.as<double>(), s.beta.as<double>(),
                                       s.b32.as<cplx<float>>(), s.x32.as<cplx<float>>(), s.r32.as<cplx<float>>(),
                                       int(s.nk), int(s.nz));
}

static void ldlt_upload(Ldlt &s, const double *b) {
    const size_t n = size_t(s.nk) * s.nz;
    if (s.precision == BLD_FP64) {
        HIPCHECK(hipMemcpy(s.b64.p, b, n * sizeof(cplx<double>), hipMemcpyHostToDevice));
    } else {
        std::vector<cplx<float>> h(n);
        for (size_t i = 0; i < n; ++i) h[i] = mk(float(b[2 * i]), float(b[2 * i + 1]));
        HIPCHECK(hipMemcpy(s.b32.p, h.data(), n * sizeof(cplx<float>), hipMemcpyHostToDevice));
    }
}

static void ldlt_download(Ldlt &s, double *x) {
    const size_t n = size_t(s.nk) * s.nz;
    if (s.precision == BLD_FP64) {
        HIPCHECK(hipMemcpy(x, s.b64.p, n * sizeof(cplx<double>), hipMemcpyDeviceToHost));
    } else {
        std::vector<cplx<float>> h(n);
        HIPCHECK(hipMemcpy(h.data(), s.x32.p, n * sizeof(cplx<float>), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            x[2 * i] = h[i].x;
            x[2 * i + 1] = h[i].y;
        }
    }
}

int bld_ldlt_solve(bld_ldlt *h, double *b) {
    if (!h || !b) return BLD_EARG;
    Ldlt &s = h->s;
    return guarded(s.err, [&] {
        ldlt_upload(s, b);
        ldlt_run(s, nullptr);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipDeviceSynchronize());
        ldlt_download(s, b);
        return BLD_OK;
    });
}

int bld_ldlt_bench(bld_ldlt *h, const double *b, int reps, double *ms_per_solve, double *x_out) {
    if (!h || !b || !ms_per_solve || reps < 1) return BLD_EARG;
    Ldlt &s = h->s;
    return guarded(s.err, [&] {
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        float total = 0;
        for (int r = 0; r < reps; ++r) {  // the FP64 solve is in place: refresh b each repetition, untimed
            ldlt_upload(s, b);
            HIPCHECK(hipEventRecord(e0, nullptr));
            ldlt_run(s, nullptr);
            HIPCHECK(hipEventRecord(e1, nullptr));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            total += ms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_per_solve = total / reps;
        if (x_out) ldlt_download(s, x_out);
        return BLD_OK;
    });
}

}  // extern "C"
