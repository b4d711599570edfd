/*
 * bld.h — C ABI of the MI355X time-step engine for BoundaryLayerDynamics.jl.
 *
 * The engine advances the reference's incompressible-flow model (Fourier in
 * x1/x2, staggered finite differences in x3) by SSPRK33 steps on one GPU.
 * It replaces, for one step, the Julia closures `rate!` / `projection!` that
 * the reference builds in src/BoundaryLayerDynamics.jl:149-165 and consumes
 * in src/ODEMethods.jl:147-166,217-234.
 *
 * Conventions
 *  - Plain C types only; every function returns 0 (BLD_OK) or a negative
 *    bld_status, and the message is available from bld_last_error().
 *  - Spectral fields are complex float64, interleaved re/im, in the Julia
 *    column-major layout of src/Grids.jl:91-93: element (i1, i2, i3) at
 *    index i1 + K1*(i2 + K2*i3) with K1 = k1max+1, K2 = 2*k2max+1, i2 ordered
 *    as k2 = [0..k2max, -k2max..-1] (src/Grids.jl:74-83).  vel1/vel2 have n3
 *    layers (C-nodes), vel3 has n3-1 layers (I-nodes).
 *  - The engine owns all device memory; it never keeps a host pointer after
 *    a call returns.  One engine per host thread; calls are synchronous.
 */
#ifndef BLD_H
#define BLD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLD_ABI_VERSION 2

typedef enum bld_status {
    BLD_OK = 0,
    BLD_EARG = -1,          /* invalid argument (sizes, pointers, config) */
    BLD_EUNSUPPORTED = -2,  /* configuration the reference rejects */
    BLD_EHIP = -3,          /* HIP runtime error */
    BLD_EDIVERGED = -4,     /* "The simulation has diverged." */
    BLD_ENOMEM = -5
} bld_status;

typedef enum bld_precision {
    BLD_FP64 = 0,           /* plain form: everything in float64 */
    BLD_FP32_MIXED = 1      /* float32 fields; float64 pressure solve, (0,0) mode and forcing */
} bld_precision;

typedef enum bld_bc_kind { BLD_BC_DIRICHLET = 0, BLD_BC_NEUMANN = 1 } bld_bc_kind;
typedef enum bld_wall_kind { BLD_WALL_FREESLIP = 0, BLD_WALL_ROUGH = 1 } bld_wall_kind;

/* Boundary condition of one velocity component at one wall
 * (ConstantValue / ConstantGradient of src/boundary_conditions.jl:36-43,53-57):
 * Dirichlet uses `value`; Neumann uses `difference` = gradient * (dx3/dζ)/n3. */
typedef struct bld_bc {
    int32_t kind;           /* bld_bc_kind */
    int32_t reserved;
    double value;
    double difference;
} bld_bc;

/* Discretised processes of one Model (src/BoundaryLayerDynamics.jl:72-97).
 * Every constant is the one the reference's init_process computes, so the GPU
 * never re-derives physics. */
typedef struct bld_config {
    int32_t abi_version;    /* BLD_ABI_VERSION */
    int32_t precision;      /* bld_precision */
    int32_t k1max, k2max, n3;
    int32_t reserved0;
    double L1, L2;          /* horizontal domain size */
    const double *alpha_c;  /* [n3]   α = n3/(dx3/dζ) at C-nodes ζ=(i-1/2)/n3 (derivatives.jl:10-19) */
    const double *alpha_i;  /* [n3+1] α at I-nodes ζ=i/n3, i = 0..n3 (walls included) */
    bld_bc bc[3][2];        /* [vel1,vel2,vel3][lower,upper] */
    /* MolecularDiffusion (molecular_diffusion.jl:31-72), per field */
    int32_t diffusion[3];
    int32_t reserved1;
    double diffusivity[3];
    /* ConstantSource (sources.jl:13-35): summed strength per field, 0 if none */
    double source[3];
    /* Pressure (pressure.jl:215-305) */
    int32_t pressure;
    /* MomentumAdvection (momentum_advection.jl:31-77): physical grid size */
    int32_t advection;
    int32_t adv_n1, adv_n2;
    /* StaticSmagorinskyModel (static_smagorinsky.jl:44-123) */
    int32_t sgs;
    int32_t sgs_n1, sgs_n2;
    int32_t sgs_wall[2];    /* bld_wall_kind, lower and upper */
    int32_t reserved2;
    const double *sgs_lc;   /* [n3]   length scale with wall damping, C-nodes */
    const double *sgs_li;   /* [n3-1] same at interior I-nodes */
    double sgs_factor_strain[2];  /* factors_strain (:92-108) */
    double sgs_factor_stress[2];  /* factors_stress */
    /* ConstantMean (sources.jl:47-77), vel1/vel2 only: projection that shifts
     * the (0,0) mode so that Σ_l weight[l]·Re q(0,0,l) = mean_value */
    int32_t mean[3];
    int32_t reserved3;
    double mean_value[3];
    const double *mean_weight;  /* [n3] 1/(α_c·L3) (sources.jl:63-66) */
} bld_config;

typedef struct bld_engine bld_engine;
typedef struct bld_ldlt bld_ldlt;

/* Model construction: allocates device state, rates, SSPRK33 buffer, FFT plans
 * and the pressure factorisation (src/BoundaryLayerDynamics.jl:79-96). */
int bld_create(const bld_config *cfg, bld_engine **out);
void bld_destroy(bld_engine *e);

/* Ensemble of n independent models (uneven grids allowed, one precision):
 * every step advances all of them, one launch per kernel class covering every
 * model, the whole SSPRK33 step captured in one HIP graph.  No reference
 * counterpart (the reference runs one Model per process). */
int bld_ensemble_create(int n, const bld_config *cfgs, bld_engine **out);
/* Number of models of an engine (1 for bld_create). */
int bld_ensemble_size(bld_engine *e);
/* Model that bld_set_state / bld_get_state / bld_rates (output) /
 * bld_set_field_physical / bld_get_field_physical act on (default 0).
 * bld_rates, bld_project, bld_init_problem and the step calls act on all. */
int bld_select_instance(bld_engine *e, int index);

/* FFT tile autotuner: times every power-of-two batch (columns / row pairs per
 * block) of every transform pass and model space on the live state, keeps the
 * fastest, and recomputes the rates of the current state (no reference
 * counterpart; the reference relies on FFTW MEASURE plans,
 * src/physical_space.jl:14). */
int bld_autotune(bld_engine *e, int reps);
/* Chosen batch and its time (µs) of transform pass (0 c2r_x2, 1 c2r_x1,
 * 2 r2c_x1, 3 r2c_x2) for the entry-th model space; pass 4: the chunk length
 * of the fused stage kernel and the time of one stage; pass 5: the transform
 * path (bit 0: whole-plane c2r, bit 1: whole-plane r2c, for the square grids
 * that fit in LDS; bit 2: c2r of the largest planes by half planes) and the
 * time of one graph-replayed SSPRK33 step (entry ignored). */
int bld_tuning(bld_engine *e, int pass, int entry, int *batch, double *us);
/* Transform path forced for testing and measurement (no reference
 * counterpart): the pass-5 bits of bld_tuning (bit 0 whole-plane c2r, bit 1
 * whole-plane r2c, bit 2 half-plane c2r blocks for planes that need
 * 1024-thread whole-plane blocks).  Model spaces a path does not fit keep the
 * two-pass transforms; the autotuned batch sizes are kept.  Results do not
 * depend on the path. */
int bld_set_transform_path(bld_engine *e, int path);

/* Learned SGS closure (no reference counterpart; it replaces the stresses of
 * the selected model's StaticSmagorinskyModel, static_smagorinsky.jl:129-249,
 * which stays configured as its wall model and per-point fallback).  params:
 * bld_mlp_nparams() floats in the bld_mlp layout; delta: Δ of the stress scale
 * Δ²|∇u|²; wall_feature: network input 10 per C-layer [n3].  params = NULL
 * restores the Smagorinsky stresses.  Features: ∇u/|∇u| at each C-node
 * (row-major ∂u_i/∂x_j), then the wall feature; outputs τ^dev/(Δ²|∇u|²) for
 * (11, 22, 33, 12, 13, 23).  Points where the output is not finite or ∇u
 * vanishes take the Smagorinsky stress; with BLD_CLOSURE_GUARD in `flags` so
 * do points where the learned stress is not dissipative (−τ:S < 0). */
#define BLD_CLOSURE_GUARD 1
int bld_sgs_closure(bld_engine *e, const float *params, double delta, const double *wall_feature, int flags);
/* Training samples of the selected model's current state: box filter of
 * `filter` grid cells in x1 and x2 on each C-layer of the default physical grid
 * (which `filter` must divide), one sample per filter box and layer.  x, t:
 * float32 [count][16] (features as above; t[0..5] = τ^dev/(Δ²|ḡ|²) with
 * τ_ij = avg(u_i u_j) − avg(u_i)·avg(u_j), t[15] = Δ²|ḡ|²).  With x or t NULL
 * only *count is set. */
int bld_closure_sample(bld_engine *e, int filter, double delta, const double *wall_feature, float *x, float *t,
                       int64_t capacity, int64_t *count);

/* Time-integration method of the step calls (ODEMethods.jl:12-166); the
 * default is SSPRK33.  AB2 takes its first step as the reference does (a
 * two-stage Runge-Kutta step that keeps f(u0)) and then requires a constant
 * dt; bld_init_problem restarts it. */
typedef enum bld_method { BLD_SSPRK33 = 0, BLD_SSPRK22 = 1, BLD_EULER = 2, BLD_AB2 = 3 } bld_method;
int bld_set_method(bld_engine *e, int method);

/* Mean profiles (MeanProfiles, logging.jl:310-525) of the selected model,
 * accumulated on the device: horizontal means sampled at the rate evaluation
 * that ends each step (and when enabled or the problem is initialised), time
 * integrated with the trapezoidal rule.  mask: one bit per bld_profile, 0
 * disables.  bld_profiles_read copies the time integral since the last reset
 * (n3 values for C-node profiles; n3+1 for I-node profiles, boundary entries
 * included) and the number of steps it covers; the mean is the integral
 * divided by the time span. */
typedef enum bld_profile {
    BLD_PROF_VEL1 = 0, BLD_PROF_VEL2, BLD_PROF_VEL3, BLD_PROF_ADV1, BLD_PROF_ADV2, BLD_PROF_ADV3,
    BLD_PROF_SGS11, BLD_PROF_SGS12, BLD_PROF_SGS13, BLD_PROF_SGS22, BLD_PROF_SGS23, BLD_PROF_SGS33
} bld_profile;
int bld_profiles_enable(bld_engine *e, uint32_t mask);
int bld_profiles_read(bld_engine *e, int profile, double *integral, int64_t *steps, int reset);

/* Message of the last failed call on `e` (or of the last failed bld_create
 * on this thread when e is NULL). */
const char *bld_last_error(const bld_engine *e);

/* State transfer in the layout above (model.state, State.jl:9). */
int bld_set_state(bld_engine *e, const double *vel1, const double *vel2, const double *vel3);
int bld_get_state(bld_engine *e, double *vel1, double *vel2, double *vel3);

/* compute_rates! (src/Processes.jl:39-62) of the current state, plus the
 * divergence check of src/BoundaryLayerDynamics.jl:155-157.  Rates are written
 * to the three host arrays (same layout as the state); pass NULL to skip. */
int bld_rates(bld_engine *e, double *du1, double *du2, double *du3);

/* apply_projections! (src/Processes.jl:67-75) on the current state. */
int bld_project(bld_engine *e);

/* ODEProblem setup (src/ODEMethods.jl:179-185): project the state and
 * evaluate the rate at t1. */
int bld_init_problem(bld_engine *e, double t1);

/* nsteps × (perform_step!(SSPRK33) + rate!) (src/ODEMethods.jl:147-166,
 * 226-231).  *t is advanced with the reference's snapping rule when t_end is
 * given (t_end <= *t disables snapping).  Every rate is checked for non-finite
 * values (src/BoundaryLayerDynamics.jl:155-157); at the first one the call
 * returns BLD_EDIVERGED with *t advanced only by the time updates the
 * reference makes before it raises (the steps completed before the step whose
 * inner rate diverged, or including the step whose trailing rate diverged).
 * The state is then unspecified (the reference keeps the last completed
 * step). */
int bld_step(bld_engine *e, double dt, int64_t nsteps, double *t, double t_end);

/* Same as bld_step but records the whole step sequence into a HIP graph the
 * first time and replays it afterwards (no host work per step). */
int bld_step_graph(bld_engine *e, double dt, int64_t nsteps, double *t, double t_end);

/* set_field!/get_field! at the default physical size (2+2k1max, 2+2k2max)
 * (src/physical_space.jl:253-312): `values` is real float64 [n1][n2][layers]
 * column-major (x1 fastest).  field: 0 vel1, 1 vel2, 2 vel3.  centered != 0:
 * the values sit at the cell centres (x_i + Δx/2), the `centered` keyword of
 * the reference (shift!/unshift!, physical_space.jl:322-336). */
int bld_set_field_physical(bld_engine *e, int field, const double *values, int centered);
int bld_get_field_physical(bld_engine *e, int field, double *values, int centered);

/* Engine timing hook: runs nsteps steps with inputs resident, returns the
 * elapsed milliseconds measured with HIP events on the engine stream.  The
 * rates are checked inside the timed steps as in bld_step; the call returns
 * BLD_EDIVERGED (with *ms set) if one was non-finite. */
int bld_time_steps(bld_engine *e, double dt, int64_t nsteps, int use_graph, double *ms);

/* Per-kernel timing: average duration (ms) of each kernel class over the last
 * bld_time_steps call when profiling is enabled (bld_set_profiling). */
int bld_set_profiling(bld_engine *e, int enabled);
int bld_kernel_times(bld_engine *e, int kind, double *ms_total, int64_t *launches);
/* Measurement hook (no counterpart in the reference): the dispatches of one
 * kernel class (same kind numbering) of a right-hand side replayed reps times
 * back to back between two HIP events on the engine stream; returns the
 * average duration of one dispatch and the dispatches per right-hand side.
 * The transform, linear-term and pointwise physics classes can be replayed;
 * the engine's rates are recomputed afterwards. */
int bld_time_kernel(bld_engine *e, int kind, int reps, double *us_per_dispatch, int64_t *dispatches);

/* Batched symmetric tridiagonal LDLt (DistributedBatchLDLt + ldiv!,
 * src/processes/pressure.jl:38-201): nk systems of size nz.  dvs [nz][nk],
 * evs [nz-1][nk] (mode index fastest).  precision: BLD_FP64 solves in float64;
 * BLD_FP32_MIXED stores the right-hand side in float32, eliminates in float64
 * and applies one step of iterative refinement with a float64 residual. */
int bld_ldlt_create(int64_t nk, int64_t nz, const double *dvs, const double *evs, int precision,
                    bld_ldlt **out);
void bld_ldlt_destroy(bld_ldlt *s);
/* In-place solve of complex float64 b [nz][nk] (interleaved), host memory. */
int bld_ldlt_solve(bld_ldlt *s, double *b);
/* Device-resident timing: uploads b, then times `reps` solves run back to back
 * between one pair of HIP events (ms per solve); x_out (optional) receives
 * the solution of b. */
int bld_ldlt_bench(bld_ldlt *s, const double *b, int reps, double *ms_per_solve, double *x_out);
const char *bld_ldlt_last_error(const bld_ldlt *s);

/* Allocation-free variant of bld_ldlt_create + bld_ldlt_solve.  No reference
 * counterpart: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops).  dvs [nz][nk], evs [nz-1][nk] and the complex
 * right-hand side b [nz][nk] (interleaved, overwritten with the solution) are
 * DEVICE pointers; the kernels run on `stream` (a hipStream_t, NULL for the
 * null stream) and the call returns without waiting for them.  The factor and
 * the float32 staging live in a fixed device array of
 * BLD_LDLT_NOALLOC_SCRATCH bytes that every GPU the library runs on holds
 * until the process ends: float64 needs 8 B per unknown (8,388,608 unknowns
 * at most), BLD_FP32_MIXED 20 B (3,355,443 at most); systems have at most
 * 4096 levels.  Larger sizes return BLD_EARG with a message.  Every call on
 * a GPU shares that GPU's array (each call overwrites all of the part it
 * uses, nothing needs clearing), so calls that overlap on one GPU race.  The
 * message of the last failed call on this thread: bld_ldlt_noalloc_last_error. */
#define BLD_LDLT_NOALLOC_SCRATCH (64 << 20)
int bld_ldlt_solve_noalloc(int64_t nk, int64_t nz, const double *dvs, const double *evs, double *b, int precision,
                           void *stream);
const char *bld_ldlt_noalloc_last_error(void);

/* Learned SGS closure network (no reference counterpart; it sits behind the
 * process API next to StaticSmagorinskyModel, static_smagorinsky.jl:39-42):
 * 10 features (padded to 16) → 64 → 64 (GELU) → 6 outputs (padded to 16),
 * bfloat16 MFMA with float32 accumulation.  Parameters are float32, row-major
 * [in][out] matrices W1[16][64], W2[64][64], W3[64][16], then b1[64], b2[64],
 * b3[16] (bld_mlp_nparams() floats).  Inputs/outputs are float32 [n][16]. */
typedef struct bld_mlp bld_mlp;
int bld_mlp_nparams(void);
int bld_mlp_create(const float *params, bld_mlp **out);
void bld_mlp_destroy(bld_mlp *h);
const char *bld_mlp_last_error(const bld_mlp *h);
int bld_mlp_get_params(bld_mlp *h, float *params);
int bld_mlp_forward(bld_mlp *h, const float *x, int64_t n, float *y);
/* Gradient of the mean squared error over the 6 outputs of (x, t). */
int bld_mlp_gradient(bld_mlp *h, const float *x, const float *t, int64_t n, float *grad, double *loss);
/* Training set (device resident) and `steps` Adam steps over consecutive
 * mini-batches of `batch` points; losses[steps] receives the batch losses. */
int bld_mlp_set_data(bld_mlp *h, const float *x, const float *t, int64_t n);
int bld_mlp_train(bld_mlp *h, int64_t steps, int64_t batch, float lr, double *losses);
/* Average milliseconds of one forward pass and one gradient pass over n points. */
int bld_mlp_time(bld_mlp *h, int64_t n, int reps, double *ms_forward, double *ms_gradient);

#ifdef __cplusplus
}
#endif

#endif /* BLD_H */
