// bk_mik.hip — multiple indicator kriging kernels (crates/estimation/src/indicator.rs).
//
// One thread per target. The solves that give the weights are the ordinary
// pass kernels (bk_solve.hip) writing their weights out; these kernels turn
// weights into indicator estimates and those into conditional distributions.
// Everything that does not depend on the target (the class bounds and shapes,
// and each class's moments above each shared cutoff) is tabulated on the host,
// so a target's moments are sums of its class probabilities times table rows.
#include "bk_mik.h"

namespace {

constexpr int kThreads = 256;

int64_t blocks_for(int64_t m) { return (m + kThreads - 1) / kThreads; }

__device__ inline double nan64() { return __longlong_as_double(0x7ff8000000000000LL); }

// krige_indicators (indicator.rs:735-748): with simple kriging about f_k,
// f_k + sum_j w_j (1{z_j <= t_k} - f_k); else sum_j w_j 1{z_j <= t_k}; the
// sums run over the samples in neighbour order.
__global__ void __launch_bounds__(kThreads) k_mik_raw(const MikRawArgs a) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= a.m) return;
  int32_t st = a.v == 0 ? BK_OK : a.status[t];
  if (st == BK_OK) st = a.st_v[t];
  a.status[t] = st;
  if (st != BK_OK) return;
  const int n = a.ncount[t];
  const int32_t* id = a.nidx + t * a.stride;
  const double* w = a.w + t * a.stride;
  const int k0 = a.V == 1 ? 0 : a.v, k1 = a.V == 1 ? a.K : a.v + 1;
  for (int k = k0; k < k1; ++k) {
    const double th = a.thr[k], f = a.prop[k];
    double sum = -0.0;
    for (int j = 0; j < n; ++j) {
      const double ind = a.sv[id[j]] <= th ? 1.0 : 0.0;
      sum += a.simple ? w[j] * (ind - f) : w[j] * ind;
    }
    a.raw[(int64_t)k * a.m + t] = a.simple ? f + sum : sum;
  }
}

__global__ void __launch_bounds__(kThreads) k_mik_status(int64_t m, const int32_t* sstatus, int32_t fallback,
                                                         int32_t* status) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t < m) status[t] = sstatus[t] != BK_OK ? sstatus[t] : fallback;
}

__device__ inline double clip01(double p) { return p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p); }

// Class c's probability under the corrected cdf (indicator.rs:357-358).
__device__ inline double class_q(const MikCondArgs& a, int64_t t, int c) {
  const double lo = c == 0 ? 0.0 : a.cdf[(int64_t)(c - 1) * a.m + t];
  const double hi = c == a.K ? 1.0 : a.cdf[(int64_t)c * a.m + t];
  return hi - lo;
}

// MultipleIndicator::conditional (indicator.rs:380-437).
__global__ void __launch_bounds__(kThreads) k_mik_conditional(const MikCondArgs a) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= a.m) return;
  const int64_t m = a.m;
  const int K = a.K;
  if (a.status[t] != BK_OK) {
    const double nan = nan64();
    for (int k = 0; k < K; ++k) a.cdf[(int64_t)k * m + t] = a.raw[(int64_t)k * m + t] = nan;
    if (a.correction) a.correction[t] = nan;
    if (a.mean) a.mean[t] = nan;
    if (a.variance) a.variance[t] = nan;
    for (int c = 0; c < a.C; ++c) {
      if (a.pa) a.pa[(int64_t)c * m + t] = nan;
      if (a.ma) a.ma[(int64_t)c * m + t] = nan;
    }
    if (a.qv)
      for (int q = 0; q < a.Q; ++q) a.qv[(int64_t)q * m + t] = nan;
    if (a.violations) a.violations[t] = 0;
    if (a.n_used) a.n_used[t] = 0;
    if (a.pit) a.pit[t] = nan;
    return;
  }
  // correct_order_relations (indicator.rs:137-150): the upward running
  // maximum goes to cdf, then the downward running minimum is averaged in.
  double run = 0.0;
  for (int k = 0; k < K; ++k) {
    const double p = clip01(a.raw[(int64_t)k * m + t]);
    run = k == 0 ? p : fmax(p, run);
    a.cdf[(int64_t)k * m + t] = run;
  }
  for (int k = K - 1; k >= 0; --k) {
    const double p = clip01(a.raw[(int64_t)k * m + t]);
    run = k == K - 1 ? p : fmin(p, run);
    const int64_t o = (int64_t)k * m + t;
    a.cdf[o] = 0.5 * (a.cdf[o] + run);
  }
  double corr = -0.0;
  int viol = 0;
  for (int k = 0; k < K; ++k) {
    const double c = a.cdf[(int64_t)k * m + t], r = a.raw[(int64_t)k * m + t];
    corr += fabs(c - r);
    viol += c != r;
  }
  if (a.correction) a.correction[t] = corr;
  if (a.violations) a.violations[t] = viol;
  if (a.n_used) a.n_used[t] = a.ncount[t];
  // Moments: sum over the classes with q > 0 of q * Shape::above, one sum per
  // (cutoff, power), in class order.
  const int CC = a.C + 1;
  auto moment = [&](int ci, int j) {
    double sum = -0.0;
    for (int c = 0; c <= K; ++c) {
      const double q = class_q(a, t, c);
      if (q > 0.0) sum += q * a.table[((int64_t)c * CC + ci) * 3 + j];
    }
    return sum;
  };
  const double mean = moment(0, 1);
  if (a.mean) a.mean[t] = mean;
  if (a.variance) a.variance[t] = fmax(moment(0, 2) - mean * mean, 0.0);
  for (int ci = 0; ci < a.C; ++ci) {
    const double p = moment(ci + 1, 0);
    if (a.pa) a.pa[(int64_t)ci * m + t] = p;
    if (a.ma) a.ma[(int64_t)ci * m + t] = p > 0.0 ? moment(ci + 1, 1) / p : nan64();
  }
  if (a.pit) {
    // The held-out value as the one cutoff: F*(z) = 1 - P(Z > z).
    const double z = a.own_values[a.held[t]];
    double sum = -0.0;
    for (int c = 0; c <= K; ++c) {
      const double q = class_q(a, t, c);
      if (q > 0.0) sum += q * mik_above(a.g, a.cls[c], z, 0);
    }
    a.pit[t] = 1.0 - sum;
  }
  if (a.qv && a.Q > 0) {
    int last = 0;
    for (int c = K; c >= 0; --c)
      if (class_q(a, t, c) > 0.0) {
        last = c;
        break;
      }
    for (int qi = 0; qi < a.Q; ++qi) {
      const double u = a.quantiles[qi];
      double below = 0.0, v = a.upper;
      for (int c = 0; c <= K; ++c) {
        const double q = class_q(a, t, c);
        if (q > 0.0 && (below + q >= u || c == last)) {
          v = mik_quantile(a.g, a.cls[c], mik_clamp((u - below) / q, 0.0, 1.0));
          break;
        }
        below += q;
      }
      a.qv[(int64_t)qi * m + t] = v;
    }
  }
}

}  // namespace

namespace bk {

int launch_mik_raw(hipStream_t s, const MikRawArgs& a) {
  if (a.m <= 0) return 0;
  k_mik_raw<<<blocks_for(a.m), kThreads, 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mik_conditional(hipStream_t s, const MikCondArgs& a) {
  if (a.m <= 0) return 0;
  k_mik_conditional<<<blocks_for(a.m), kThreads, 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mik_status(hipStream_t s, int64_t m, const int32_t* sstatus, int32_t fallback, int32_t* status) {
  if (m <= 0) return 0;
  k_mik_status<<<blocks_for(m), kThreads, 0, s>>>(m, sstatus, fallback, status);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bk
