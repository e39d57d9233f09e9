// bk_mik.h — multiple indicator kriging on the device: the class model of the
// conditional distribution (crates/estimation/src/indicator.rs:206-289), shared
// by the host (per-class tables that do not depend on the target) and the
// per-target kernel, and the two kernels' launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "../../include/boitata_gpu.h"

#define BK_HD __host__ __device__

// Class distribution shapes (indicator.rs:206-213).
enum { BK_SHAPE_DATA = 0, BK_SHAPE_POWER = 1, BK_SHAPE_HYPERBOLIC = 2 };

// One class of the distribution: bounds [a, b], its shape, and for a data
// shape the index range [s, e) of the sorted global values it holds.
struct MikClass {
  double a, b, w;
  int32_t kind;
  int32_t pad;
  int64_t s, e;
};

// The declustered global distribution (indicator.rs:69-81): sorted values and
// prefix sums of w, w v and w v^2 (weights summing to 1), sums[j * (n + 1) + i].
struct MikGlobal {
  const double* values;
  const double* sums;
  int64_t n;
};

// Rust's f64::powi with a run-time exponent: binary exponentiation from the
// low bit (compiler-rt __powidf2), so x^3 = x * x^2.
BK_HD inline double mik_powi(double x, int n) {
  double r = 1.0;
  for (;;) {
    if (n & 1) r = r * x;
    n /= 2;
    if (n == 0) break;
    x = x * x;
  }
  return r;
}

// indicator.rs:215-217
BK_HD inline double mik_binomial(int n, int k) {
  double acc = 1.0;
  for (int i = 0; i < k; ++i) acc = acc * (double)(n - i) / (double)(i + 1);
  return acc;
}

BK_HD inline double mik_clamp(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }

// Entries of v[lo, hi) before the first for which `below` fails (v sorted so
// that `below` holds on a prefix): slice::partition_point.
template <class F>
BK_HD inline int64_t mik_partition(const double* v, int64_t lo, int64_t hi, F below) {
  int64_t a = lo, b = hi;
  while (a < b) {
    int64_t mid = a + (b - a) / 2;
    if (below(v[mid]))
      a = mid + 1;
    else
      b = mid;
  }
  return a - lo;
}

// Shape::above (indicator.rs:220-253): E[Z^j 1{Z > c}] under the class's
// distribution.
BK_HD inline double mik_above(const MikGlobal& g, const MikClass& k, double c, int j) {
  if (k.kind == BK_SHAPE_DATA) {
    const double* s0 = g.sums;
    const double* sj = g.sums + (int64_t)j * (g.n + 1);
    int64_t p = k.s + mik_partition(g.values, k.s, k.e, [&](double v) { return v <= c; });
    return (sj[k.e] - sj[p]) / (s0[k.e] - s0[k.s]);
  }
  const double a = k.a, b = k.b, w = k.w;
  if (b <= a) return a > c ? mik_powi(a, j) : 0.0;
  if (k.kind == BK_SHAPE_POWER) {
    double x0 = mik_clamp((c - a) / (b - a), 0.0, 1.0);
    double sum = -0.0;
    for (int i = 0; i <= j; ++i) {
      double e = w + (double)i;
      sum += mik_binomial(j, i) * mik_powi(a, j - i) * mik_powi(b - a, i) * w / e * (1.0 - pow(x0, e));
    }
    return sum;
  }
  double yb = b / a, yc = mik_clamp(c, a, b) / a;
  double m = (double)j - w;
  double integral = fabs(m) < 1e-12 ? log(yb / yc) : (pow(yb, m) - pow(yc, m)) / m;
  return mik_powi(a, j) * w / (1.0 - pow(yb, -w)) * integral;
}

// Shape::quantile (indicator.rs:272-289): the value at probability u within
// the class.
BK_HD inline double mik_quantile(const MikGlobal& g, const MikClass& k, double u) {
  const double a = k.a, b = k.b;
  if (k.kind == BK_SHAPE_DATA) {
    const double* s0 = g.sums;
    double target = s0[k.s] + u * (s0[k.e] - s0[k.s]);
    int64_t i = k.s + mik_partition(s0, k.s + 1, k.e + 1, [&](double v) { return v < target; });
    return g.values[i < k.e - 1 ? i : k.e - 1];
  }
  if (b <= a) return a;
  if (k.kind == BK_SHAPE_POWER) return mik_clamp(a + (b - a) * pow(u, 1.0 / k.w), a, b);
  double r = pow(a / b, k.w);
  return mik_clamp(a * pow(1.0 - u * (1.0 - r), -1.0 / k.w), a, b);
}

// Indicator estimates of one variogram's weights (krige_indicators,
// indicator.rs:735-748): raw[k * m + t] for the thresholds k that use
// variogram v (all of them when one variogram serves every threshold).
struct MikRawArgs {
  int64_t m;
  int v, V, K;
  int simple;
  const double* thr;    // K thresholds
  const double* prop;   // K declustered proportions (the simple kriging means)
  const double* sv;     // sample values
  const int32_t* nidx;  // neighbour lists, stride per target
  const int32_t* ncount;
  int stride;
  const double* w;      // this variogram's weights, [t * stride + k]
  const int32_t* st_v;  // this variogram's per-target status
  int32_t* status;      // the target's status over the variograms so far
  double* raw;
};

// The conditional distribution per target (indicator.rs:380-437).
struct MikCondArgs {
  int64_t m;
  int K, C, Q;
  double* raw;               // K * m (NaN written where the target failed)
  const int32_t* status;
  const int32_t* ncount;
  const MikClass* cls;       // K + 1 classes
  const double* table;       // (K + 1) x (C + 1) x 3: Shape::above(class, cutoff, j); cutoff 0 is -inf
  const double* quantiles;   // Q
  MikGlobal g;
  double upper;              // global.upper
  // Cross-validation: one cutoff per target, the held-out sample's value.
  const double* own_values;  // sample values, or null
  const int64_t* held;       // held-out sample of each target
  // Outputs (any may be null but cdf).
  double* cdf;               // K * m
  double* correction;
  double* mean;
  double* variance;
  double* pa;                // C * m
  double* ma;                // C * m
  double* qv;                // Q * m
  int32_t* violations;
  int32_t* n_used;
  double* pit;
};

namespace bk {
int launch_mik_raw(hipStream_t s, const MikRawArgs& a);
int launch_mik_conditional(hipStream_t s, const MikCondArgs& a);
// Status of every target when the solves cannot run (simple kriging with a
// variogram without a sill): the search's failure, else `fallback`.
int launch_mik_status(hipStream_t s, int64_t m, const int32_t* sstatus, int32_t fallback, int32_t* status);
}  // namespace bk
