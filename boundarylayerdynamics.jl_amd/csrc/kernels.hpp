// Descriptors and host launchers of the engine kernels.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace bld {

enum PassKind : int {
    PASS_C2R_X2 = 0, PASS_C2R_X1 = 1, PASS_R2C_X1 = 2, PASS_R2C_X2 = 3,
    PASS_C2R_2D = 4,       // fused c2r, 1024-thread blocks
    PASS_R2C_2D = 5,       // fused r2c, 1024-thread blocks
    PASS_C2R_2D_HALF = 6,  // fused c2r, 512-thread blocks (two per CU)
    PASS_C2R_2H = 7,       // half-plane c2r (even / odd rows), 512-thread blocks (two per CU)
};
constexpr int kC2r2dThreads = 1024;  // block of the fused r2c (one plane per block)
enum PhysKind : int { PHYS_ADV = 0, PHYS_SGS_VISC = 1, PHYS_SGS_STRESS = 2, PHYS_SGS_MLP = 3, PHYS_SGS_MLP_STRESS = 4 };
// C-node points per block of the learned-closure kernel (256 threads x chunks)
constexpr int kMlpChunks = 4;
// Chunk length of the float64 (0,0) column pass inside the chunked stage
// launch of chunk length CL: short chunks keep that one-column pass from
// raising the register count of the whole launch (n3 <= 64·CL levels, so
// at most 256 chunks of this length, one thread each).
__host__ __device__ constexpr int stage00_cl(int CL) { return CL <= 4 ? 1 : CL / 4; }
// consecutive levels of one mode per thread of the linear-term kernel
constexpr int kLinLevels = 8;
// consecutive C-levels of one column per thread of the advection products
constexpr int kColLevels = 16;
// consecutive C-levels of one column per thread of the SGS eddy viscosity + stress kernel
constexpr int kSgsLevels = 4;

// Spectral linear terms of one model instance.
template <typename T>
struct LinDesc {
    int K1, K2, nk, n3;
    const double *DD1, *DD2;            // -(k·2π/L)^2
    const double *alpha_c, *alpha_i;    // [n3], [n3+1]
    int has_diff[3];
    double nu[3], src[3];
    int bc_kind[3][2];
    double bc_val[3][2], bc_diff[3][2];
    const cplx<T> *state[2][3];         // input state: 0 = u, 1 = utmp
    const cplx<double> *shadow[2][3];   // float64 (0,0) shadows (mixed form) or null
    cplx<T> *du[3];
    cplx<double> *du00[3];              // float64 linear rate of the (0,0) mode (mixed form)
};

// Stage combination + projection of one model instance.
template <typename T>
struct ProjDesc {
    int K1, K2, nk, n3;
    int has_pressure;
    const double *D1, *D2;
    const double *alpha_c, *alpha_i;
    const double *gamma;                // [n3][nk] LDLt factor (pressure.jl:86-95)
    double wbc[2];                      // vel3 Dirichlet values
    cplx<double> *pbuf;                 // [n3][nk] scratch
    cplx<T> *u[3], *ut[3], *du[3];
    cplx<double> *u00[3], *ut00[3], *du00[3];
    int *flag;                          // [0] first diverged step count, [1] step count (record_diverged)
    // float64 (0,0)-mode pass of the mixed form: the float32 nonlinear rate of
    // the mode is added to du, and results are mirrored into the float32 state
    int MB, skip0;                      // chunked kernel: modes per block; skip the (0,0) column
    int ext_stride;                     // 0: no extras
    const cplx<float> *du_add[3];
    cplx<float> *mirror_u[3], *mirror_ut[3];
};

// Stage kinds of the fused stage kernels (ODEMethods.jl:100-166): the
// Runge-Kutta stages, forward Euler and the two-step Adams–Bashforth update
// (with its Runge-Kutta start), each followed by the projection.
enum StageKind : int {
    STAGE_PROJECT = 0,      // projection of u only
    STAGE_RK_1 = 1,         // utmp = u + dt·du (SSPRK33 and SSPRK22 stage 1)
    STAGE_RK3_2 = 2,        // utmp = (3u + utmp + dt·du)/4
    STAGE_RK3_3 = 3,        // u = (u + 2·utmp + 2dt·du)/3
    STAGE_EULER = 4,        // u = u + dt·du
    STAGE_RK2_2 = 5,        // u = (u + utmp + dt·du)/2
    STAGE_AB2 = 6,          // u = u + dt/2·(3du − utmp); utmp = du
    STAGE_AB2_FIRST_1 = 7,  // u = u + dt·du; utmp = du
    STAGE_AB2_FIRST_2 = 8,  // u = u + dt/2·(du − utmp)
};
__host__ __device__ constexpr bool stage_writes_ut(int s) { return s == STAGE_RK_1 || s == STAGE_RK3_2; }
__host__ __device__ constexpr bool stage_reads_ut(int s) {
    return s == STAGE_RK3_2 || s == STAGE_RK3_3 || s == STAGE_RK2_2 || s == STAGE_AB2 || s == STAGE_AB2_FIRST_2;
}
__host__ __device__ constexpr bool stage_saves_du(int s) { return s == STAGE_AB2 || s == STAGE_AB2_FIRST_1; }

// Mean profiles of one model (MeanProfiles, logging.jl:310-525)
enum ProfileKind : int {
    PROF_VEL1 = 0, PROF_VEL2, PROF_VEL3, PROF_ADV1, PROF_ADV2, PROF_ADV3,
    PROF_SGS11, PROF_SGS12, PROF_SGS13, PROF_SGS22, PROF_SGS23, PROF_SGS33, PROF_COUNT
};
template <typename T>
struct ProfDesc {
    unsigned mask;
    int n3, nk;
    const cplx<T> *u[3];
    const cplx<double> *u00[3];      // float64 (0,0) shadows of the mixed form (else null)
    const Space<T> *adv, *sgs;       // device pointers of the model's spaces (or null)
    double *prev, *acc;              // [PROF_COUNT][n3+1]: last sample, time integral
};
template <typename T>
void launch_profile_sample(const ProfDesc<T> &D, double weight, hipStream_t st);

// ConstantMean projection of one model (vel1, vel2 at C-nodes)
template <typename T>
struct CMeanDesc {
    int on[2];
    int n3, nk;
    double value[2];
    const double *w;  // [n3] weights 1/(α_c·L3)
    cplx<T> *u[2], *ut[2];
    cplx<double> *u00[2], *ut00[2];  // float64 shadows of the mixed form (else null)
};

template <typename T>
void launch_transform(int kind, const Space<T> *spaces_dev, const LaunchTable &lt, int nblocks, size_t smem, int sel,
                      hipStream_t st);
template <typename T>
hipError_t allow_large_lds();
template <typename T>
void launch_physics(int kind, const Space<T> *spaces_dev, const LaunchTable &lt, int nblocks, hipStream_t st);
// learned closure (PHYS_SGS_MLP: stresses, PHYS_SGS_MLP_STRESS: their rates)
template <typename T>
void launch_closure(int kind, const Space<T> *spaces_dev, const LaunchTable &lt, int nblocks, hipStream_t st);
// box-filtered samples (filter width F cells) of one space's physical terms
template <typename T>
void launch_closure_sample(const Space<T> *space_dev, int N1, int N2, int n3, int F, double delta2, const double *xw,
                           float *X, float *Y, hipStream_t st);
template <typename T>
void launch_constant_mean(const CMeanDesc<T> &D, int stage, hipStream_t st);
template <typename T>
void launch_linear(const LinDesc<T> *descs_dev, const LaunchTable &lt, int nblocks, int sel, int *step_count,
                   hipStream_t st);
template <typename T>
void launch_stage_chunked(int CL, const ProjDesc<T> *descs_dev, const LaunchTable &lt, int nblocks, int stage,
                          double dt, hipStream_t st, const ProjDesc<double> *descs00_dev = nullptr, int n00 = 0);
template <typename T>
void launch_stage_sequential(const ProjDesc<T> &D, int stage, double dt, hipStream_t st);
template <typename T>
void launch_check_finite(const void *a, size_t n, int *flag, hipStream_t st);
template <typename T>
void launch_fill(void *a, size_t n, hipStream_t st);

}  // namespace bld
