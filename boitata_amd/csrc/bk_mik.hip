// bk_mik.hip — multiple indicator kriging kernels (crates/estimation/src/indicator.rs).
//
// The solves that give the weights are the ordinary pass kernels
// (bk_solve.hip) writing their weights out; these kernels turn weights into
// indicator estimates (a group of lanes per target) and those into
// conditional distributions (a thread per target).
// Everything that does not depend on the target (the class bounds and shapes,
// and each class's moments above each shared cutoff) is tabulated on the host,
// so a target's moments are sums of its class probabilities times table rows.
#include "bk_mik.h"

namespace {

constexpr int kThreads = 256;

int64_t blocks_for(int64_t m) { return (m + kThreads - 1) / kThreads; }

__device__ inline double nan64() { return __longlong_as_double(0x7ff8000000000000LL); }

// krige_indicators (indicator.rs:735-748): with simple kriging about f_k,
// f_k + sum_j w_j (1{z_j <= t_k} - f_k); else sum_j w_j 1{z_j <= t_k}.
// G lanes per target (G >= 16, a power of two): lane r holds the weights and
// values of neighbours r, r + G, ... (a target's weights are contiguous, so
// the group's loads coalesce), and each threshold's sum is a shuffle
// reduction over the group.
constexpr int kMikPer = 4;  // neighbours per lane: G * kMikPer >= the list stride (<= 256)

template <int G>
__global__ void __launch_bounds__(kThreads) k_mik_raw(const MikRawArgs a) {
  const int64_t t = ((int64_t)blockIdx.x * kThreads + threadIdx.x) / G;
  const int r = threadIdx.x % G;
  const bool in = t < a.m;
  int32_t st = BK_SEARCH_FAILED;
  if (in) {
    st = a.v == 0 ? BK_OK : a.status[t];
    if (st == BK_OK) st = a.st_v[t];
  }
  const bool ok = in && st == BK_OK;
  const int n = ok ? a.ncount[t] : 0;
  double w[kMikPer], z[kMikPer];
#pragma unroll
  for (int e = 0; e < kMikPer; ++e) {
    const int j = r + e * G;
    w[e] = 0.0;
    z[e] = 0.0;
    if (j < n) {
      w[e] = a.w[t * a.stride + j];
      z[e] = a.sv[a.nidx[t * a.stride + j]];
    }
  }
  if (in && r == 0) a.status[t] = st;
  const int k0 = a.V == 1 ? 0 : a.v, k1 = a.V == 1 ? a.K : a.v + 1;
  for (int k = k0; k < k1; ++k) {
    const double th = a.thr[k], f = a.prop[k];
    double part = 0.0;
#pragma unroll
    for (int e = 0; e < kMikPer; ++e) {
      const double ind = z[e] <= th ? 1.0 : 0.0;
      part += a.simple ? w[e] * (ind - f) : w[e] * ind;
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, G);
    if (ok && r == 0) a.raw[(int64_t)k * a.m + t] = a.simple ? f + part : part;
  }
}

__global__ void __launch_bounds__(kThreads) k_mik_status(int64_t m, const int32_t* sstatus, int32_t fallback,
                                                         int32_t* status) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t < m) status[t] = sstatus[t] != BK_OK ? sstatus[t] : fallback;
}

__device__ inline double clip01(double p) { return p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p); }

// MultipleIndicator::conditional (indicator.rs:380-437) of target t. With
// KR > 0 (K <= KR) the corrected cdf lives in registers (every loop over the
// thresholds unrolled to KR); with KR = 0 it lives in the cdf output, read back
// per class.
template <int KR>
__device__ __forceinline__ void conditional_of(const MikCondArgs& a, int64_t t) {
  constexpr int KN = KR > 0 ? KR : 1;
  constexpr int KC = KR > 0 ? KR + 1 : 1;  // unroll counts (none for the memory form)
  const int64_t m = a.m;
  const int K = a.K;
  const int KL = KR > 0 ? KR : K;  // loop bound (guarded by K when unrolled)
  double reg[KN];
  auto put = [&](int k, double v) {
    if constexpr (KR > 0)
      reg[k] = v;
    else
      a.cdf[(int64_t)k * m + t] = v;
  };
  auto get = [&](int k) -> double {
    if constexpr (KR > 0)
      return reg[k];
    else
      return a.cdf[(int64_t)k * m + t];
  };
  // correct_order_relations (indicator.rs:137-150): the upward running
  // maximum, then the downward running minimum averaged in.
  double run = 0.0;
#pragma unroll KN
  for (int k = 0; k < KL; ++k) {
    if (k < K) {
      const double p = clip01(a.raw[(int64_t)k * m + t]);
      run = k == 0 ? p : fmax(p, run);
      put(k, run);
    }
  }
#pragma unroll KN
  for (int kk = 0; kk < KL; ++kk) {
    const int k = KL - 1 - kk;
    if (k < K) {
      const double p = clip01(a.raw[(int64_t)k * m + t]);
      run = k == K - 1 ? p : fmin(p, run);
      put(k, 0.5 * (get(k) + run));
    }
  }
  double corr = -0.0;
  int viol = 0;
#pragma unroll KN
  for (int k = 0; k < KL; ++k) {
    if (k < K) {
      const double c = get(k), r = a.raw[(int64_t)k * m + t];
      corr += fabs(c - r);
      viol += c != r;
      if constexpr (KR > 0) a.cdf[(int64_t)k * m + t] = c;
    }
  }
  if (a.correction) a.correction[t] = corr;
  if (a.violations) a.violations[t] = viol;
  if (a.n_used) a.n_used[t] = a.ncount[t];
  // Class c's probability under the corrected cdf (indicator.rs:357-358).
  auto class_q = [&](int c) -> double {
    const double lo = c == 0 ? 0.0 : get(c - 1);
    const double hi = c == K ? 1.0 : get(c);
    return hi - lo;
  };
  // Moments: sum over the classes with q > 0 of q * Shape::above, one sum per
  // (cutoff, power), in class order.
  const int CC = a.C + 1;
  auto moment = [&](int ci, int j) {
    double sum = -0.0;
#pragma unroll KC
    for (int c = 0; c <= KL; ++c) {
      if (c <= K) {
        const double q = class_q(c);
        if (q > 0.0) sum += q * a.table[((int64_t)c * CC + ci) * 3 + j];
      }
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
#pragma unroll KC
    for (int c = 0; c <= KL; ++c) {
      if (c <= K) {
        const double q = class_q(c);
        if (q > 0.0) sum += q * mik_above(a.g, a.cls[c], z, 0);
      }
    }
    a.pit[t] = 1.0 - sum;
  }
  if (a.qv && a.Q > 0) {
    int last = 0;
#pragma unroll KC
    for (int c = 0; c <= KL; ++c)
      if (c <= K && class_q(c) > 0.0) last = c;
    for (int qi = 0; qi < a.Q; ++qi) {
      const double u = a.quantiles[qi];
      double below = 0.0, v = a.upper;
      bool found = false;
#pragma unroll KC
      for (int c = 0; c <= KL; ++c) {
        if (c <= K && !found) {
          const double q = class_q(c);
          if (q > 0.0 && (below + q >= u || c == last)) {
            v = mik_quantile(a.g, a.cls[c], mik_clamp((u - below) / q, 0.0, 1.0));
            found = true;
          } else {
            below += q;
          }
        }
      }
      a.qv[(int64_t)qi * m + t] = v;
    }
  }
}

template <int KR>
__global__ void __launch_bounds__(kThreads) k_mik_conditional(const MikCondArgs a) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= a.m) return;
  const int64_t m = a.m;
  if (a.status[t] != BK_OK) {
    const double nan = nan64();
    for (int k = 0; k < a.K; ++k) a.cdf[(int64_t)k * m + t] = a.raw[(int64_t)k * m + t] = nan;
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
  conditional_of<KR>(a, t);
}

}  // namespace

namespace bk {

int launch_mik_raw(hipStream_t s, const MikRawArgs& a) {
  if (a.m <= 0) return 0;
  if (a.stride > 64 * kMikPer) return -1;
  if (a.stride <= 16)
    k_mik_raw<16><<<blocks_for(a.m * 16), kThreads, 0, s>>>(a);
  else if (a.stride <= 32)
    k_mik_raw<32><<<blocks_for(a.m * 32), kThreads, 0, s>>>(a);
  else
    k_mik_raw<64><<<blocks_for(a.m * 64), kThreads, 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mik_conditional(hipStream_t s, const MikCondArgs& a) {
  if (a.m <= 0) return 0;
  if (a.K <= 16)
    k_mik_conditional<16><<<blocks_for(a.m), kThreads, 0, s>>>(a);
  else
    k_mik_conditional<0><<<blocks_for(a.m), kThreads, 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mik_status(hipStream_t s, int64_t m, const int32_t* sstatus, int32_t fallback, int32_t* status) {
  if (m <= 0) return 0;
  k_mik_status<<<blocks_for(m), kThreads, 0, s>>>(m, sstatus, fallback, status);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bk
