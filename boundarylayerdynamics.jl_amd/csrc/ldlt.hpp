// Batched symmetric tridiagonal LDLt solve (DistributedBatchLDLt + ldiv!,
// src/processes/pressure.jl:38-201) shared by the handle-based solver and the
// allocation-free entry point.
#pragma once

#include <hip/hip_runtime.h>

namespace bld {

// Tallest system of the chunked kernel (256 chunks of 16 levels).
constexpr int kLdltChunkedMaxLevels = 256 * 16;

// One chunked solve of nk systems of nz <= kLdltChunkedMaxLevels levels on
// stream st, mode index fastest in every [level][nk] array.  fp32_ir = false:
// gamma double [nz][nk], b and x complex double (x may be b).  fp32_ir = true:
// gamma float, b and x complex float, dv the float64 diagonal [nz][nk]; one
// refinement step with a float64 residual.  beta: off-diagonals, rows
// 0..nz-2 of [nz-1][nk] are read when beta_per_mode (any valid [nk] array when
// nz = 1), else one value per level [nz-1].
void ldlt_chunked_launch(bool fp32_ir, const void *gamma, const double *beta, bool beta_per_mode, const double *dv,
                         const void *b, void *x, int nk, int nz, hipStream_t st);

}  // namespace bld
