// bk_diag.hip — per-target diagnostics of a kriging pass: the statistics of the
// samples used that _Estimator.predict(diagnostics=True) tabulates
// (crates/py/src/estimation.rs:1034-1105): neighborhood_stats over those
// samples (crates/estimation/src/neighborhood.rs:66-136; their distinct holes
// and mean distance) and the sum of the negative weights
// (crates/estimation/src/krige.rs:64-75).
//
// One wavefront per target, lane j on neighbour j (and j + 64). The mean
// distance is summed in ascending order of distance, as the reference sums
// its sorted distances: each distance's rank is counted over the others
// (ties by neighbour order, as a stable sort leaves them) and lane 0 adds them
// up in rank order.
#include "bk_diag.h"

namespace {

constexpr int kWaves = 4;
constexpr int kMaxUsed = 256;  // the most samples a search selects (BK_MAX_SEARCH), all solved on the device

__global__ void __launch_bounds__(64 * kWaves) k_used_stats(const DiagArgs a) {
  __shared__ double dist[kWaves][kMaxUsed];
  __shared__ double sorted[kWaves][kMaxUsed];
  __shared__ int32_t holes[kWaves][kMaxUsed];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t t = (int64_t)blockIdx.x * kWaves + wave;
  const bool ok = t < a.m && a.status[t] == BK_OK && a.ncount[t] <= kMaxUsed;
  const int n = ok ? a.ncount[t] : 0;
  double* d = dist[wave];
  double* srt = sorted[wave];
  int32_t* h = holes[wave];
  for (int j = lane; j < n; j += 64) {
    const int id = a.nidx[t * a.stride + j];
    // aniso.rs:114-119 euclidean(target, sample)
    const double dx = a.sx[id] - a.tx[t], dy = a.sy[id] - a.ty[t], dz = a.sz[id] - a.tz[t];
    d[j] = sqrt((dx * dx + dy * dy) + dz * dz);
    h[j] = a.hole ? a.hole[id] : -1;
  }
  __syncthreads();
  int first = 0;
  for (int j = lane; j < n; j += 64) {
    const double dj = d[j];
    int rank = 0;
    for (int q = 0; q < n; ++q) {
      const double dq = d[q];
      rank += (dq < dj || (dq == dj && q < j)) ? 1 : 0;
    }
    srt[rank] = dj;
    // A tagged hole counts at its first sample; an untagged sample counts alone.
    bool fresh = h[j] < 0;
    if (!fresh) {
      fresh = true;
      for (int q = 0; q < j; ++q) fresh = fresh && h[q] != h[j];
    }
    first += fresh ? 1 : 0;
  }
  // Distinct holes: lanes hold the counts of their neighbours.
  for (int off = 32; off > 0; off >>= 1) first += __shfl_xor(first, off, 64);
  __syncthreads();
  if (t < a.m && lane == 0) {
    if (ok) {
      double sd = -0.0;
      for (int q = 0; q < n; ++q) sd += srt[q];
      double neg = -0.0;
      const double* w = a.w + t * a.stride;
      for (int q = 0; q < n; ++q) {
        const double x = w[q];
        if (x < 0.0) neg += x;
      }
      if (a.mean_distance) a.mean_distance[t] = sd / (double)n;
      if (a.negative_weight_sum) a.negative_weight_sum[t] = n > 0 ? neg : __longlong_as_double(0x7ff8000000000000LL);
      if (a.n_holes) a.n_holes[t] = first;
    } else {
      const double nan = __longlong_as_double(0x7ff8000000000000LL);
      if (a.mean_distance) a.mean_distance[t] = nan;
      if (a.negative_weight_sum) a.negative_weight_sum[t] = nan;
      if (a.n_holes) a.n_holes[t] = 0;
      if (a.target_met) a.target_met[t] = -1;
    }
  }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

}  // namespace

namespace bk {

int launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t value) {
  if (n <= 0) return 0;
  k_fill_i32<<<(n + 255) / 256, 256, 0, s>>>(p, n, value);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_used_stats(hipStream_t s, const DiagArgs& a) {
  if (a.m <= 0) return 0;
  const int64_t blocks = (a.m + kWaves - 1) / kWaves;
  k_used_stats<<<blocks, 64 * kWaves, 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bk
