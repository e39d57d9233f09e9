// bk_diag.h — per-target diagnostics of a kriging pass (bk_diag.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/boitata_gpu.h"

struct DiagArgs {
  int64_t m;
  int stride;
  const int32_t* nidx;    // neighbour lists, stride per target
  const int32_t* ncount;  // samples used per target
  const int32_t* status;  // the pass's per-target status
  const double* tx;
  const double* ty;
  const double* tz;
  const double* sx;
  const double* sy;
  const double* sz;
  const int32_t* hole;    // may be null (no holes)
  const double* w;        // the solves' weights, [t * stride + k]
  // Outputs (any may be null).
  double* negative_weight_sum;
  int32_t* n_holes;
  double* mean_distance;
  int32_t* target_met;    // set to -1 where the target failed
};

namespace bk {
int launch_used_stats(hipStream_t s, const DiagArgs& a);
int launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t value);
}  // namespace bk
