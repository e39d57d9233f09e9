// oracle.cpp — CPU restatement of boitatá's kriging hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness oracle and the timed
// CPU baseline of the engine under boitata_amd/; the engine never links, loads
// or calls it. It restates, in FP64 and in the reference's operation order
// (compiled with -ffp-contract=off so no product is fused into an FMA):
//   crates/core/src/rotation.rs:8-43           rotation_matrix, block_frame
//   crates/variogram/src/aniso.rs:75-119       Anisotropy::{new, matrix, lag}, euclidean
//   crates/variogram/src/model.rs:76-126,181-311 shape (+ the derivatives the adjoint uses)
//   crates/variogram/src/composite.rs:50-129   Variogram::{total_sill, gamma, cov, lag, ...}
//   crates/estimation/src/search.rs:214-264    Search::check
//   crates/estimation/src/search.rs:396-595    Sectors, Selector, select
//   crates/estimation/src/search.rs:647-687    neighbors_in (the scan)
//   crates/estimation/src/search.rs:692-1006   SearchTree, Space
//   crates/estimation/src/search/grid.rs:11-381 Grid (nearest-first cells)
//   crates/estimation/src/coincident.rs:11-112 Clusters
//   crates/estimation/src/kernel.rs:24-123     the stack Cholesky kernel
//   crates/estimation/src/krige.rs:83-201      krige, value, check_simple, krige_lu
//   crates/estimation/src/block.rs:18-198      Discretization, block_covariances, block_krige_points
//   crates/estimation/src/batch.rs:15-38,165-190 estimate_many, leave_one_out_at
//   crates/estimation/src/kriging_algebra.rs:320-441 Quality, prefix_quality,
//                                              calibrated_count, krige_calibrated
//   crates/estimation/src/search.rs:909-936    SearchTree::balanced
//   crates/py/src/estimation.rs:947-958        block_points
// Third-party arithmetic the reference delegates to nalgebra 0.33.3 (absent from
// the reference tree): 3x3 products are restated in nalgebra's column-axpy order,
// r_i = (m_i0 v0 + m_i1 v1) + m_i2 v2; DMatrix::lu().solve is restated as
// partial-pivoting LU with reciprocal-scaled columns and column-oriented
// triangular solves. Below about 1e-9 relative the LU outputs are parity
// unpinned (no nalgebra here to compare against).
//
// Extension (not in the reference): a structure may carry its own anisotropy
// ratios under the variogram's rotation; with every structure on the shared
// ratios each function reduces to the reference's arithmetic exactly.

#include "../include/boitata_gpu.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <queue>
#include <utility>
#include <map>
#include <vector>

namespace orc {

using P3 = std::array<double, 3>;
struct M3 {
  double m[9];  // row-major
  double operator()(int i, int j) const { return m[3 * i + j]; }
};

static const double kPi = 3.14159265358979323846;

// nalgebra Matrix3 * Vector3 (gemv, column-axpy order).
static inline P3 mv(const M3& a, const P3& v) {
  P3 r;
  for (int i = 0; i < 3; ++i) {
    double t0 = a(i, 0) * v[0];
    double t1 = a(i, 1) * v[1];
    double t2 = a(i, 2) * v[2];
    r[i] = (t0 + t1) + t2;
  }
  return r;
}

// nalgebra Matrix3 * Matrix3 (gemm as column-wise gemv).
static inline M3 mm(const M3& a, const M3& b) {
  M3 c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t0 = a(i, 0) * b(0, j);
      double t1 = a(i, 1) * b(1, j);
      double t2 = a(i, 2) * b(2, j);
      c.m[3 * i + j] = (t0 + t1) + t2;
    }
  return c;
}

static inline M3 identity() { return M3{{1, 0, 0, 0, 1, 0, 0, 0, 1}}; }

// f64::to_radians: self * (PI / 180).
static inline double to_radians(double d) { return d * (kPi / 180.0); }

// crates/core/src/rotation.rs:8-16
M3 rotation_matrix(double azimuth, double dip, double rake) {
  double a = to_radians(azimuth), d = to_radians(dip), r = to_radians(rake);
  double sa = std::sin(a), ca = std::cos(a);
  double sd = std::sin(d), cd = std::cos(d);
  double sr = std::sin(r), cr = std::cos(r);
  M3 r1{{sa, ca, 0.0, -ca, sa, 0.0, 0.0, 0.0, 1.0}};
  M3 r2{{cd, 0.0, -sd, 0.0, 1.0, 0.0, sd, 0.0, cd}};
  M3 r3{{1.0, 0.0, 0.0, 0.0, cr, sr, 0.0, -sr, cr}};
  return mm(mm(r3, r2), r1);
}

// crates/core/src/rotation.rs:40-43: rows (-R1, R0, R2).
M3 block_frame(const double rot[3]) {
  M3 r = rotation_matrix(rot[0], rot[1], rot[2]);
  M3 b;
  for (int j = 0; j < 3; ++j) {
    b.m[j] = -r(1, j);
    b.m[3 + j] = r(0, j);
    b.m[6 + j] = r(2, j);
  }
  return b;
}

// Rust `powi` lowered by LLVM's binary exponentiation (x^3 = x * x^2, ...).
static inline double powi(double x, int n) {
  double res = 0.0;
  bool have = false;
  double cur = x;
  while (n) {
    if (n & 1) {
      res = have ? res * cur : cur;
      have = true;
    }
    n >>= 1;
    if (n) cur = cur * cur;
  }
  return have ? res : 1.0;
}

// crates/variogram/src/aniso.rs:18-111 (major, semi, minor ranges).
struct Aniso {
  double azimuth = 0, dip = 0, rake = 0, major = 1, semi = 1, minor = 1;
  M3 rot = identity();
  void init() { rot = rotation_matrix(azimuth, dip, rake); }
  // aniso.rs:88-94: diag(1/major, 1/semi, 1/minor) * R
  M3 matrix() const {
    M3 d{{1.0 / major, 0.0, 0.0, 0.0, 1.0 / semi, 0.0, 0.0, 0.0, 1.0 / minor}};
    return mm(d, rot);
  }
  // aniso.rs:101-111
  double lag(const P3& p, const P3& q) const {
    P3 v{q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    P3 r = mv(rot, v);
    double s0 = r[0] / major, s1 = r[1] / semi, s2 = r[2] / minor;
    return std::sqrt((s0 * s0 + s1 * s1) + s2 * s2);
  }
};

// aniso.rs:114-119
static inline double euclidean(const P3& p, const P3& q) {
  double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
  return std::sqrt((dx * dx + dy * dy) + dz * dz);
}

// crates/variogram/src/model.rs:76-126
double shape(int model, double param, double h, double range) {
  if (!std::isfinite(h)) return 0.0;
  if (h <= 0.0) return 0.0;
  double r = std::max(range, std::numeric_limits<double>::min());
  double t = h / r;
  switch (model) {
    case BK_SPHERICAL:
      return t >= 1.0 ? 1.0 : 1.5 * t - 0.5 * powi(t, 3);
    case BK_EXPONENTIAL:
      return 1.0 - std::exp(-3.0 * t);
    case BK_GAUSSIAN:
      return 1.0 - std::exp(-3.0 * t * t);
    case BK_CUBIC:
      if (t >= 1.0) return 1.0;
      return 7.0 * powi(t, 2) - (35.0 / 4.0) * powi(t, 3) + (7.0 / 2.0) * powi(t, 5) -
             (3.0 / 4.0) * powi(t, 7);
    case BK_PENTASPHERICAL:
      if (t >= 1.0) return 1.0;
      return (15.0 / 8.0) * t - (5.0 / 4.0) * powi(t, 3) + (3.0 / 8.0) * powi(t, 5);
    case BK_CIRCULAR:
      if (t >= 1.0) return 1.0;
      return 1.0 - (2.0 / kPi) * std::acos(t) + (2.0 * t / kPi) * std::sqrt(1.0 - t * t);
    case BK_SINEHOLE: {
      double c = kPi * t;
      return 1.0 - std::sin(c) / c;
    }
    case BK_MATERN: {
      // model.rs:136-179: order snaps to 0.5, 1.5, 2.5; delta = sqrt(2 nu) * 3 * t
      double nu = param < 1.0 ? 0.5 : (param < 2.0 ? 1.5 : 2.5);
      double delta = (std::sqrt(2.0 * nu) * 3.0) * t;
      double e = std::exp(-delta);
      if (nu == 0.5) return 1.0 - e;
      if (nu == 1.5) return 1.0 - (1.0 + delta) * e;
      return 1.0 - (1.0 + delta + delta * delta / 3.0) * e;
    }
    case BK_POWER:
      return std::pow(t, param);
  }
  return 0.0;
}

// model.rs:207-311: (dg/dt, d2g/dt2) of the normalised shape; used to check the
// adjoint's closed forms.
std::pair<double, double> d_shape(int model, double param, double t) {
  switch (model) {
    case BK_SPHERICAL:
      return t >= 1.0 ? std::make_pair(0.0, 0.0) : std::make_pair(1.5 - 1.5 * t * t, -3.0 * t);
    case BK_EXPONENTIAL: {
      double e = std::exp(-3.0 * t);
      return {3.0 * e, -9.0 * e};
    }
    case BK_GAUSSIAN: {
      double e = std::exp(-3.0 * t * t);
      return {6.0 * t * e, (6.0 - 36.0 * t * t) * e};
    }
    default:
      return {std::nan(""), std::nan("")};
  }
}

struct Structure {
  int model = BK_SPHERICAL;
  double param = 0, sill = 0, range = 1;
  bool own = false;    // carries its own anisotropy ratios (extension)
  Aniso aniso;         // its own anisotropy when `own`
  bool stationary() const { return model != BK_POWER; }
};

// crates/variogram/src/composite.rs:13-129
struct Variogram {
  double nugget = 0;
  std::vector<Structure> st;
  bool has_aniso = false;
  Aniso aniso;  // the shared anisotropy (major = 1)

  double total_sill() const {
    double s = -0.0;
    for (auto& x : st) s += x.sill;
    return nugget + s;
  }
  bool is_stationary() const {
    for (auto& x : st)
      if (!x.stationary()) return false;
    return true;
  }
  bool any_own() const {
    for (auto& x : st)
      if (x.own) return true;
    return false;
  }
  // gamma with per-structure lags hs[s]; the nugget's h > 0 test uses h0.
  double gamma_lags(double h0, const double* hs) const {
    double s = -0.0;
    for (size_t k = 0; k < st.size(); ++k) s += st[k].sill * shape(st[k].model, st[k].param, hs[k], st[k].range);
    return h0 > 0.0 ? nugget + s : 0.0;
  }
  double cov_lags(double h0, const double* hs) const {
    if (h0 <= 0.0) return total_sill();
    return total_sill() - gamma_lags(h0, hs);
  }
  // composite.rs:69-85 (scalar lag, shared anisotropy)
  double gamma(double h) const {
    double hs[BK_MAX_STRUCTURES];
    for (size_t k = 0; k < st.size(); ++k) hs[k] = h;
    return gamma_lags(h, hs);
  }
  double cov(double h) const { return h <= 0.0 ? total_sill() : total_sill() - gamma(h); }
  double lag(const P3& p, const P3& q) const { return has_aniso ? aniso.lag(p, q) : euclidean(p, q); }
  // Lags of every structure between p and q; returns the nugget test lag.
  double lags(const P3& p, const P3& q, double* hs) const {
    double h = lag(p, q);
    for (size_t k = 0; k < st.size(); ++k) hs[k] = st[k].own ? st[k].aniso.lag(p, q) : h;
    return st.empty() ? h : hs[0];
  }
  // composite.rs:101-103
  double cov_points(const P3& p, const P3& q) const {
    double hs[BK_MAX_STRUCTURES];
    double h0 = lags(p, q, hs);
    return cov_lags(h0, hs);
  }
  // composite.rs:108-115
  double block_cov_points(const P3& p, const P3& q) const {
    double hs[BK_MAX_STRUCTURES];
    double h0 = lags(p, q, hs);
    if (h0 <= 0.0) return total_sill() - nugget;
    return cov_lags(h0, hs);
  }
  // composite.rs:120-129
  double block_mean_cov(const std::vector<P3>& pts) const {
    if (pts.size() == 1) return total_sill();
    double total = -0.0;
    for (auto& p : pts) {
      double row = -0.0;
      for (auto& q : pts) row += block_cov_points(p, q);
      total += row;
    }
    return total / (double)(pts.size() * pts.size());
  }
};

static Variogram make_variogram(const bk_variogram* v) {
  Variogram g;
  g.nugget = v->nugget;
  g.has_aniso = v->has_anisotropy != 0;
  if (g.has_aniso) {
    g.aniso.azimuth = v->azimuth;
    g.aniso.dip = v->dip;
    g.aniso.rake = v->rake;
    g.aniso.major = 1.0;
    g.aniso.semi = v->ratio_semi;
    g.aniso.minor = v->ratio_minor;
    g.aniso.init();
  }
  for (int k = 0; k < v->n_structures && k < BK_MAX_STRUCTURES; ++k) {
    const bk_structure& s = v->structures[k];
    Structure x;
    x.model = s.model;
    x.param = s.param;
    x.sill = s.sill;
    x.range = s.range;
    x.own = s.ratio_semi > 0.0 && s.ratio_minor > 0.0;
    if (x.own) {
      x.aniso.azimuth = v->has_anisotropy ? v->azimuth : 0.0;
      x.aniso.dip = v->has_anisotropy ? v->dip : 0.0;
      x.aniso.rake = v->has_anisotropy ? v->rake : 0.0;
      x.aniso.major = 1.0;
      x.aniso.semi = s.ratio_semi;
      x.aniso.minor = s.ratio_minor;
      x.aniso.init();
    }
    g.st.push_back(x);
  }
  return g;
}

struct Sample {
  P3 loc;
  double value;
  int32_t hole;  // -1: none
  double err;
};

enum Kind { ORDINARY = 0, SIMPLE = 1, INDICATOR = 2 };

struct Estimate {
  double value = NAN, variance = NAN, lagrange = NAN, support = NAN;
  int n_used = 0;
  std::vector<double> weights;
};

// ---------------------------------------------------------------------------
// Coincident guard, crates/estimation/src/coincident.rs:11-112
static const double EPS = 1e-6;
static inline bool close(const P3& p, const P3& q) {
  double a = p[0] - q[0], b = p[1] - q[1], c = p[2] - q[2];
  return (a * a + b * b) + c * c <= EPS * EPS;
}

struct Clusters {
  std::vector<int> of;
  std::vector<std::vector<int>> members;
  static bool find(const std::vector<Sample>& s, Clusters& out) {
    bool any = false;
    for (size_t i = 0; i < s.size() && !any; ++i)
      for (size_t j = 0; j < i; ++j)
        if (close(s[i].loc, s[j].loc)) {
          any = true;
          break;
        }
    if (!any) return false;
    out.of.clear();
    out.members.clear();
    for (size_t i = 0; i < s.size(); ++i) {
      int c = -1;
      for (size_t k = 0; k < out.members.size(); ++k)
        if (close(s[out.members[k][0]].loc, s[i].loc)) {
          c = (int)k;
          break;
        }
      if (c >= 0) {
        out.of.push_back(c);
        out.members[c].push_back((int)i);
      } else {
        out.of.push_back((int)out.members.size());
        out.members.push_back({(int)i});
      }
    }
    return true;
  }
  std::vector<Sample> merge(const std::vector<Sample>& s, const Variogram& vg) const {
    double c0 = vg.total_sill();
    std::vector<Sample> out;
    for (auto& m : members) {
      const Sample& first = s[m[0]];
      if (m.size() == 1) {
        out.push_back(first);
        continue;
      }
      double k = (double)m.size();
      double var = 0.0;
      for (int i : m) {
        var += s[i].err;
        for (int j : m) var += vg.cov_points(s[i].loc, s[j].loc);
      }
      double sum = -0.0;
      for (int i : m) sum += s[i].value;
      Sample x = first;
      x.value = sum / k;
      x.err = var / (k * k) - c0;
      out.push_back(x);
    }
    return out;
  }
};

// krige.rs:98-113
static double kind_value(int kind, double param, const std::vector<double>& w, const std::vector<Sample>& s) {
  double sum = -0.0;
  for (size_t i = 0; i < w.size(); ++i) {
    double z = s[i].value;
    double f = kind == SIMPLE ? z - param : (kind == INDICATOR ? (z <= param ? 1.0 : 0.0) : z);
    sum += w[i] * f;
  }
  if (kind == SIMPLE) return param + sum;
  if (kind == INDICATOR) return std::min(std::max(sum, 0.0), 1.0);
  return sum;
}

static void spread(const Clusters& c, Estimate& e, int kind, double param, const std::vector<Sample>& s) {
  std::vector<double> w(c.of.size());
  for (size_t i = 0; i < c.of.size(); ++i) w[i] = e.weights[c.of[i]] / (double)c.members[c.of[i]].size();
  e.weights = w;
  e.value = kind_value(kind, param, e.weights, s);
  e.n_used = (int)c.of.size();
}

// ---------------------------------------------------------------------------
// The stack Cholesky kernel, crates/estimation/src/kernel.rs:24-123.
// Returns BK_OK / BK_SINGULAR, or -1 when the system is beyond it (LU takes it).
static const int MAXK = 64;
static const double PIVOT = 1e-12;

int krige_kernel(int kind, double param, const P3& target, const std::vector<Sample>& s, const Variogram& vg,
                 Estimate& out) {
  int n = (int)s.size();
  if (n == 0 || n > MAXK || !vg.is_stationary()) return -1;
  // Frames: the shared one, and each structure's own.
  int S = (int)vg.st.size();
  M3 shared = vg.has_aniso ? vg.aniso.matrix() : identity();
  std::vector<M3> fr(S);
  for (int k = 0; k < S; ++k) fr[k] = vg.st[k].own ? vg.st[k].aniso.matrix() : shared;
  std::vector<P3> p(n);                    // shared frame
  std::vector<std::vector<P3>> ps(S, std::vector<P3>(n));
  for (int i = 0; i < n; ++i) {
    P3 d{s[i].loc[0] - target[0], s[i].loc[1] - target[1], s[i].loc[2] - target[2]};
    p[i] = vg.has_aniso ? mv(shared, d) : d;
    for (int k = 0; k < S; ++k) ps[k][i] = vg.st[k].own ? mv(fr[k], d) : p[i];
  }
  auto norm = [](const P3& a, const P3& b) {
    double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return std::sqrt((x * x + y * y) + z * z);
  };
  auto covij = [&](int i, int j) {
    double hs[BK_MAX_STRUCTURES];
    double h = norm(p[i], p[j]);
    for (int k = 0; k < S; ++k) hs[k] = vg.st[k].own ? norm(ps[k][i], ps[k][j]) : h;
    return vg.cov_lags(S ? hs[0] : h, hs);
  };
  auto cov0 = [&](int i) {
    P3 o{0.0, 0.0, 0.0};
    double hs[BK_MAX_STRUCTURES];
    double h = norm(p[i], o);
    for (int k = 0; k < S; ++k) hs[k] = vg.st[k].own ? norm(ps[k][i], o) : h;
    return vg.cov_lags(S ? hs[0] : h, hs);
  };
  double c0 = vg.total_sill();
  std::vector<double> l(n * (n + 1) / 2), c(n);
  for (int i = 0; i < n; ++i) {
    int row = i * (i + 1) / 2;
    for (int j = 0; j < i; ++j) l[row + j] = covij(i, j);
    l[row + i] = c0 + s[i].err;
    c[i] = cov0(i);
  }
  auto dot = [](const double* a, const double* b, int len) {
    double acc = -0.0;
    for (int k = 0; k < len; ++k) acc += a[k] * b[k];
    return acc;
  };
  for (int j = 0; j < n; ++j) {
    int rj = j * (j + 1) / 2;
    double d = l[rj + j] - dot(&l[rj], &l[rj], j);
    if (std::isnan(d) || d <= PIVOT * c0) return BK_SINGULAR;
    d = std::sqrt(d);
    l[rj + j] = d;
    for (int i = j + 1; i < n; ++i) {
      int ri = i * (i + 1) / 2;
      l[ri + j] = (l[ri + j] - dot(&l[ri], &l[rj], j)) / d;
    }
  }
  auto solve = [&](std::vector<double>& v) {
    for (int i = 0; i < n; ++i) {
      int ri = i * (i + 1) / 2;
      v[i] = (v[i] - dot(&l[ri], v.data(), i)) / l[ri + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double acc = v[i];
      for (int k = i + 1; k < n; ++k) acc -= l[k * (k + 1) / 2 + i] * v[k];
      v[i] = acc / l[i * (i + 1) / 2 + i];
    }
  };
  std::vector<double> a = c;
  solve(a);
  bool ordinary = kind != SIMPLE;
  std::vector<double> b(n, 0.0);
  double mu = 0.0;
  if (ordinary) {
    std::fill(b.begin(), b.end(), 1.0);
    solve(b);
    double sa = -0.0, sb = -0.0;
    for (int i = 0; i < n; ++i) sa += a[i];
    for (int i = 0; i < n; ++i) sb += b[i];
    mu = (sa - 1.0) / sb;
  }
  out.weights.resize(n);
  for (int i = 0; i < n; ++i) out.weights[i] = a[i] - mu * b[i];
  out.value = kind_value(kind, param, out.weights, s);
  double swc = -0.0;
  for (int i = 0; i < n; ++i) swc += out.weights[i] * c[i];
  out.variance = std::max(c0 - swc - mu, 0.0);
  out.n_used = n;
  out.lagrange = mu;
  out.support = c0;
  return BK_OK;
}

// ---------------------------------------------------------------------------
// Dense LU with partial pivoting: nalgebra's LU::new + solve (restated).
// a is row-major dim x dim, overwritten; b overwritten with the solution.
// Returns false when U has a zero on its diagonal (nalgebra's solve -> None).
static bool lu_solve(std::vector<double>& a, std::vector<double>& b, int dim) {
  std::vector<int> perm(dim);
  for (int i = 0; i < dim; ++i) perm[i] = i;
  for (int i = 0; i < dim; ++i) {
    int piv = i;
    double best = std::fabs(a[i * dim + i]);
    for (int r = i + 1; r < dim; ++r) {
      double v = std::fabs(a[r * dim + i]);
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    double diag = a[piv * dim + i];
    if (diag == 0.0) continue;  // nalgebra leaves the column; solve fails later
    if (piv != i) {
      for (int k = 0; k < dim; ++k) std::swap(a[i * dim + k], a[piv * dim + k]);
      std::swap(b[i], b[piv]);
    }
    double inv = 1.0 / diag;
    for (int r = i + 1; r < dim; ++r) a[r * dim + i] *= inv;
    for (int k = i + 1; k < dim; ++k) {
      double y = a[i * dim + k];
      for (int r = i + 1; r < dim; ++r) a[r * dim + k] = a[r * dim + k] + (-y) * a[r * dim + i];
    }
  }
  // L y = P b (unit diagonal), column oriented.
  for (int i = 0; i < dim; ++i) {
    double coeff = b[i];
    for (int r = i + 1; r < dim; ++r) b[r] = b[r] + (-coeff) * a[r * dim + i];
  }
  // U x = y, column oriented.
  for (int i = dim - 1; i >= 0; --i) {
    double d = a[i * dim + i];
    if (d == 0.0) return false;
    b[i] = b[i] / d;
    double coeff = b[i];
    for (int r = 0; r < i; ++r) b[r] = b[r] + (-coeff) * a[r * dim + i];
  }
  return true;
}

// krige.rs:117-124
static bool check_simple(int kind, const Variogram& vg) { return !(kind == SIMPLE && !vg.is_stationary()); }

// krige.rs:127-201
int krige_lu(int kind, double param, const P3& target, const std::vector<Sample>& s, const Variogram& vg,
             Estimate& out) {
  int n = (int)s.size();
  if (n == 0) return BK_INSUFFICIENT_DATA;
  if (!check_simple(kind, vg)) return BK_INVALID_PARAMETERS;
  std::vector<double> z(n);
  for (int i = 0; i < n; ++i) z[i] = kind == INDICATOR ? (s[i].value <= param ? 1.0 : 0.0) : s[i].value;
  double c0 = vg.total_sill();
  bool ordinary = kind != SIMPLE;
  int dim = ordinary ? n + 1 : n;
  std::vector<double> a(dim * dim, 0.0), b(dim, 0.0);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) a[i * dim + j] = vg.cov_points(s[i].loc, s[j].loc);
    a[i * dim + i] += s[i].err;
    b[i] = vg.cov_points(s[i].loc, target);
  }
  std::vector<double> rhs = b;
  if (ordinary) {
    for (int i = 0; i < n; ++i) {
      a[i * dim + n] = 1.0;
      a[n * dim + i] = 1.0;
    }
    a[n * dim + n] = 0.0;
    b[n] = 1.0;
  }
  if (!lu_solve(a, b, dim)) return BK_SINGULAR;
  out.weights.assign(b.begin(), b.begin() + n);
  double sum = -0.0;
  for (int i = 0; i < n; ++i) sum += out.weights[i] * (kind == SIMPLE ? z[i] - param : z[i]);
  if (kind == SIMPLE)
    out.value = param + sum;
  else if (kind == INDICATOR)
    out.value = std::min(std::max(sum, 0.0), 1.0);
  else
    out.value = sum;
  double swc = -0.0;
  for (int i = 0; i < n; ++i) swc += out.weights[i] * rhs[i];
  double mu = ordinary ? b[n] : 0.0;
  out.variance = std::max(c0 - swc - mu, 0.0);
  out.n_used = n;
  out.lagrange = mu;
  out.support = c0;
  return BK_OK;
}

// krige.rs:83-95
int krige(int kind, double param, const P3& target, const std::vector<Sample>& s, const Variogram& vg,
          Estimate& out) {
  Clusters c;
  if (Clusters::find(s, c)) {
    std::vector<Sample> merged = c.merge(s, vg);
    int st = krige(kind, param, target, merged, vg, out);
    if (st != BK_OK) return st;
    spread(c, out, kind, param, s);
    return BK_OK;
  }
  int st = krige_kernel(kind, param, target, s, vg, out);
  if (st >= 0) return st;
  return krige_lu(kind, param, target, s, vg, out);
}

// ---------------------------------------------------------------------------
// Block support, crates/estimation/src/block.rs
// block.rs:40-57 Discretization::points
static std::vector<P3> disc_points(const P3& center, const P3& size, int nx, int ny, int nz) {
  std::vector<P3> pts;
  double dx = size[0] / nx, dy = size[1] / ny, dz = size[2] / nz;
  double x0 = center[0] - size[0] / 2.0 + dx / 2.0;
  double y0 = center[1] - size[1] / 2.0 + dy / 2.0;
  double z0 = center[2] - size[2] / 2.0 + dz / 2.0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) pts.push_back({x0 + i * dx, y0 + j * dy, z0 + k * dz});
  return pts;
}

// block.rs:60-70 offsets + py/estimation.rs:947-958 block_points
std::vector<P3> block_points(const P3& t, const bk_block& b) {
  M3 f = block_frame(b.rotation);
  M3 ft{{f(0, 0), f(1, 0), f(2, 0), f(0, 1), f(1, 1), f(2, 1), f(0, 2), f(1, 2), f(2, 2)}};
  std::vector<P3> off = disc_points({0.0, 0.0, 0.0}, {b.size[0], b.size[1], b.size[2]}, b.nx, b.ny, b.nz);
  std::vector<P3> out;
  for (auto& o : off) {
    P3 p = mv(ft, o);
    out.push_back({t[0] + p[0], t[1] + p[1], t[2] + p[2]});
  }
  return out;
}

// block.rs:122-198
int block_krige_points(int kind, double param, const std::vector<P3>& pts, const std::vector<Sample>& s,
                       const Variogram& vg, Estimate& out) {
  int n = (int)s.size();
  if (n == 0) return BK_INSUFFICIENT_DATA;
  if (pts.empty()) return BK_INVALID_PARAMETERS;
  for (auto& p : pts)
    if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]))) return BK_INVALID_PARAMETERS;
  if (!check_simple(kind, vg)) return BK_INVALID_PARAMETERS;
  Clusters cl;
  if (Clusters::find(s, cl)) {
    int st = block_krige_points(kind, param, pts, cl.merge(s, vg), vg, out);
    if (st != BK_OK) return st;
    spread(cl, out, kind, param, s);
    return BK_OK;
  }
  bool ordinary = kind != SIMPLE;
  int dim = ordinary ? n + 1 : n;
  // block.rs:100-118
  double np = (double)pts.size();
  std::vector<double> rhs(n);
  for (int i = 0; i < n; ++i) {
    double acc = -0.0;
    for (auto& p : pts) acc += pts.size() == 1 ? vg.cov_points(s[i].loc, p) : vg.block_cov_points(s[i].loc, p);
    rhs[i] = acc / np;
  }
  double cbb = vg.block_mean_cov(pts);
  std::vector<double> b(dim, 0.0), a(dim * dim, 0.0);
  for (int i = 0; i < n; ++i) b[i] = rhs[i];
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) a[i * dim + j] = vg.cov_points(s[i].loc, s[j].loc);
    a[i * dim + i] += s[i].err;
    if (ordinary) {
      a[i * dim + n] = 1.0;
      a[n * dim + i] = 1.0;
    }
  }
  if (ordinary) b[n] = 1.0;
  std::vector<double> x = b;
  if (!lu_solve(a, x, dim)) return BK_SINGULAR;
  out.weights.assign(x.begin(), x.begin() + n);
  out.value = kind_value(kind, param, out.weights, s);
  double mu = ordinary ? x[n] : 0.0;
  double swc = -0.0;
  for (int i = 0; i < n; ++i) swc += out.weights[i] * b[i];
  out.variance = std::max(cbb - swc - mu, 0.0);
  out.n_used = n;
  out.lagrange = mu;
  out.support = cbb;
  return BK_OK;
}

// Sum of v[0 .. n) zero-padded to `lanes` (a power of two) as a butterfly:
// at each stride o = lanes/2 .. 1 every entry l becomes v[l] + v[l ^ o]; the
// sum is entry 0. This is the order of a sum across a group of lanes by xor
// shuffles.
static double butterfly_sum(std::vector<double> v, int lanes) {
  v.resize(lanes, 0.0);
  for (int o = lanes / 2; o > 0; o >>= 1) {
    std::vector<double> w(lanes);
    for (int l = 0; l < lanes; ++l) w[l] = v[l] + v[l ^ o];
    v.swap(w);
  }
  return v[0];
}

// Block kriging in the engine's FP64 order (not the reference's), kept to pin
// the engine's arithmetic where the reference's LU and this path differ by
// more than 1e-10 (ill-conditioned block systems): the right-hand side, C(B,B)
// and matrix of block_krige_points (block.rs:100-158); the Cholesky of
// kernel.rs:58-72 with its pivot test; the forward solves of kernel.rs:73-78;
// back solves that subtract the rows below from the last one up (kernel.rs:79-84
// subtracts them first to last); the Schur scalar, weights, value and variance
// of kernel.rs:86-122 with every sum a butterfly over `lanes` (a power of two,
// at least n). Returns -1 for systems with coincident samples or a failed
// pivot test: the engine solves those by LU, as block_krige_points does.
int block_krige_chol(int kind, double param, const std::vector<P3>& pts, const std::vector<Sample>& s,
                     const Variogram& vg, int lanes, Estimate& out) {
  int n = (int)s.size();
  if (n == 0) return BK_INSUFFICIENT_DATA;
  if (n > lanes || n > MAXK || !vg.is_stationary() || pts.empty()) return -1;
  Clusters cl;
  if (Clusters::find(s, cl)) return -1;
  const double c0 = vg.total_sill();
  const double np = (double)pts.size();
  std::vector<double> c(n);
  for (int i = 0; i < n; ++i) {
    double acc = -0.0;
    for (auto& p : pts) acc += pts.size() == 1 ? vg.cov_points(s[i].loc, p) : vg.block_cov_points(s[i].loc, p);
    c[i] = acc / np;
  }
  const double cbb = vg.block_mean_cov(pts);
  std::vector<double> l(n * (n + 1) / 2);
  for (int i = 0; i < n; ++i) {
    int row = i * (i + 1) / 2;
    for (int j = 0; j < i; ++j) l[row + j] = vg.cov_points(s[i].loc, s[j].loc);
    l[row + i] = c0 + s[i].err;
  }
  auto dot = [](const double* a, const double* b, int len) {
    double acc = -0.0;
    for (int k = 0; k < len; ++k) acc += a[k] * b[k];
    return acc;
  };
  for (int j = 0; j < n; ++j) {
    int rj = j * (j + 1) / 2;
    double d = l[rj + j] - dot(&l[rj], &l[rj], j);
    if (std::isnan(d) || d <= PIVOT * c0) return -1;
    d = std::sqrt(d);
    l[rj + j] = d;
    for (int i = j + 1; i < n; ++i) {
      int ri = i * (i + 1) / 2;
      l[ri + j] = (l[ri + j] - dot(&l[ri], &l[rj], j)) / d;
    }
  }
  auto solve = [&](std::vector<double>& v) {
    for (int i = 0; i < n; ++i) {
      int ri = i * (i + 1) / 2;
      v[i] = (v[i] - dot(&l[ri], v.data(), i)) / l[ri + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double acc = v[i];
      for (int k = n - 1; k > i; --k) acc -= l[k * (k + 1) / 2 + i] * v[k];
      v[i] = acc / l[i * (i + 1) / 2 + i];
    }
  };
  std::vector<double> a = c;
  solve(a);
  const bool ordinary = kind != SIMPLE;
  std::vector<double> b(n, 1.0);
  double mu = 0.0;
  if (ordinary) {
    solve(b);
    mu = (butterfly_sum(a, lanes) - 1.0) / butterfly_sum(b, lanes);
  }
  out.weights.resize(n);
  for (int i = 0; i < n; ++i) out.weights[i] = a[i] - (ordinary ? mu * b[i] : 0.0);
  std::vector<double> f(n), wc(n);
  for (int i = 0; i < n; ++i) {
    const double w = out.weights[i], z = s[i].value;
    f[i] = kind == SIMPLE ? w * (z - param) : (kind == INDICATOR ? w * (z <= param ? 1.0 : 0.0) : w * z);
    wc[i] = w * c[i];
  }
  const double sum = butterfly_sum(f, lanes);
  out.value = kind == SIMPLE ? param + sum : (kind == INDICATOR ? std::min(std::max(sum, 0.0), 1.0) : sum);
  out.variance = std::max(cbb - butterfly_sum(wc, lanes) - mu, 0.0);
  out.n_used = n;
  out.lagrange = mu;
  out.support = cbb;
  return BK_OK;
}

// ---------------------------------------------------------------------------
// Calibrated kriging, crates/estimation/src/kriging_algebra.rs:320-441.

// block.rs:100-118 block_covariances: the right-hand side (one point is point
// support) and C(B, B).
static std::vector<double> block_rhs(const std::vector<P3>& pts, const std::vector<Sample>& s, const Variogram& vg) {
  double np = (double)pts.size();
  std::vector<double> rhs(s.size());
  for (size_t i = 0; i < s.size(); ++i) {
    double acc = -0.0;
    for (auto& p : pts) acc += pts.size() == 1 ? vg.cov_points(s[i].loc, p) : vg.block_cov_points(s[i].loc, p);
    rhs[i] = acc / np;
  }
  return rhs;
}

// kriging_algebra.rs:320-336
struct Quality {
  bool some = false;
  double slope = NAN, efficiency = NAN;
  bool meets(int calibration, double target) const {
    if (!some) return false;
    return calibration == BK_CALIBRATE_SLOPE ? slope >= target : efficiency >= target;
  }
};

// kriging_algebra.rs:338-397 prefix_quality for the ordinary (one drift
// column of ones, x0 = 1) and simple (none) drifts. The factorisation and the
// triangular solves are nalgebra 0.33's, restated: Cholesky::new updates
// column j by the columns k < j in order (y <- (-L[j][k]) * L[.][k] + y) and
// divides it by the square root of its diagonal, failing unless every pivot
// is positive; solve_lower_triangular divides b[i] by L[i][i] and then
// updates the rows below by -b[i] * L[.][i].
std::vector<Quality> prefix_quality(const std::vector<Sample>& s, const Variogram& vg, const std::vector<double>& rhs,
                                    bool ordinary, double support) {
  const int n = (int)s.size();
  std::vector<Quality> q(n);
  std::vector<double> a((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[i * n + j] = vg.cov_points(s[i].loc, s[j].loc) + (i == j ? s[i].err : 0.0);
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < j; ++k) {
      double factor = -a[j * n + k];
      for (int r = j; r < n; ++r) a[r * n + j] = factor * a[r * n + k] + a[r * n + j];
    }
    double diag = a[j * n + j];
    if (!(diag > 0.0)) return q;  // every prefix None
    double denom = std::sqrt(diag);
    for (int r = j + 1; r < n; ++r) a[r * n + j] = a[r * n + j] / denom;
    a[j * n + j] = denom;
  }
  auto forward = [&](std::vector<double>& b) {
    for (int i = 0; i < n; ++i) {
      double coeff = b[i] / a[i * n + i];
      b[i] = coeff;
      for (int r = i + 1; r < n; ++r) b[r] = -coeff * a[r * n + i] + b[r];
    }
  };
  std::vector<double> y = rhs, yx(n, 1.0);
  forward(y);
  if (ordinary) forward(yx);
  double ss = 0.0, aa = 0.0, bb = 0.0;
  for (int k = 0; k < n; ++k) {
    ss += y[k] * y[k];
    double cov, mu;
    if (!ordinary) {
      cov = ss;
      mu = 0.0;
    } else {
      aa += yx[k] * yx[k];
      bb += yx[k] * y[k];
      // The 1 x 1 Cholesky of aa and its solve of (bb - x0), x0 = 1.
      if (!(aa > 0.0)) continue;
      double d = std::sqrt(aa);
      if (d <= 1e-7 * d) continue;
      double m = ((bb - 1.0) / d) / d;
      cov = ss - m * bb;
      mu = m;
    }
    q[k].some = true;
    q[k].slope = cov / (cov - mu);
    q[k].efficiency = (cov + mu) / support;
  }
  return q;
}

// kriging_algebra.rs:399-415
std::pair<int, bool> calibrated_count(int calibration, double target, int64_t min_samples,
                                      const std::vector<Quality>& q) {
  int64_t from = std::max<int64_t>(min_samples, 1);
  for (int64_t k = from - 1; k < (int64_t)q.size(); ++k)
    if (q[k].meets(calibration, target)) return {(int)(k + 1), true};
  return {(int)q.size(), false};
}

// ---------------------------------------------------------------------------
// Search, crates/estimation/src/search.rs
struct Search {
  int64_t min_samples = 4, max_samples = 16;
  double radius = INFINITY;
  bool has_aniso = false;
  Aniso aniso;
  bool octant = false;
  int64_t max_per_hole = -1, max_per_octant = -1, sectors = 0, max_per_sector = 0;
  int64_t min_sectors = -1, max_empty_sectors = -1, min_holes = -1;
  int calibration = 0;  // 0 none, BK_CALIBRATE_SLOPE, BK_CALIBRATE_EFFICIENCY (search.rs:94-99)
  double calibration_target = 0.0;
  bool balances() const { return octant || sectors > 0; }
};

static Search make_search(const bk_search* p) {
  Search s;
  s.min_samples = p->min_samples;
  s.max_samples = p->max_samples;
  s.radius = p->radius;
  s.has_aniso = p->has_anisotropy != 0;
  if (s.has_aniso) {
    s.aniso.azimuth = p->azimuth;
    s.aniso.dip = p->dip;
    s.aniso.rake = p->rake;
    s.aniso.major = 1.0;
    s.aniso.semi = p->ratio_semi;
    s.aniso.minor = p->ratio_minor;
    s.aniso.init();
  }
  s.octant = p->octant != 0;
  s.max_per_hole = p->max_per_hole;
  s.max_per_octant = p->max_per_octant;
  s.sectors = p->sectors;
  s.max_per_sector = p->max_per_sector;
  s.min_sectors = p->min_sectors;
  s.max_empty_sectors = p->max_empty_sectors;
  s.min_holes = p->min_holes;
  s.calibration = p->calibration;
  s.calibration_target = p->calibration_target;
  return s;
}

// search.rs:214-264
static bool check(const Search& s) {
  if (std::isnan(s.radius) || s.radius <= 0.0 || s.max_samples <= 0 || s.min_samples > s.max_samples ||
      s.min_samples < 0)
    return false;
  int64_t count = 0, ring = 0;
  if (s.sectors > 0) {
    if (s.octant) return false;
    if (s.sectors < 2 || s.sectors > 360 || s.max_per_sector <= 0) return false;
    count = ring = s.sectors;
  } else if (s.octant) {
    count = 8;
    ring = 4;
  }
  if (s.max_per_octant >= 0) {
    if (!s.octant) return false;
    if (s.max_per_octant == 0) return false;
  }
  bool rules = s.min_sectors >= 0 || s.max_empty_sectors >= 0;
  if (rules && !s.balances()) return false;
  if (s.min_sectors >= 0 && (s.min_sectors == 0 || s.min_sectors > count || s.min_sectors > s.max_samples))
    return false;
  if (s.max_empty_sectors >= 0 && s.max_empty_sectors >= ring) return false;
  if (s.min_holes >= 0 && (s.min_holes == 0 || s.min_holes > s.max_samples)) return false;
  // py/estimation.rs:458-472 (the Python Search constructor's checks)
  if (s.calibration < 0 || s.calibration > BK_CALIBRATE_EFFICIENCY) return false;
  if (s.calibration == BK_CALIBRATE_SLOPE && !(std::isfinite(s.calibration_target) && s.calibration_target > 0.0))
    return false;
  if (s.calibration == BK_CALIBRATE_EFFICIENCY &&
      !(std::isfinite(s.calibration_target) && s.calibration_target <= 1.0))
    return false;
  return true;
}

// search.rs:396-467
struct Sectors {
  M3 axes = identity();
  bool planar = false;
  int64_t plane = 0;  // 0: none
  int count() const { return plane > 0 ? (int)plane : (planar ? 4 : 8); }
  int of(const P3& t, const P3& s) const {
    P3 d = mv(axes, {s[0] - t[0], s[1] - t[1], s[2] - t[2]});
    if (plane > 0) {
      const double tau = 2.0 * kPi;
      double turn = std::fmod(std::atan2(-d[0], d[1]), tau);
      if (turn < 0.0) turn = turn + tau;
      double f = turn / tau * (double)plane;
      int64_t k = f >= 0.0 ? (int64_t)f : 0;  // `as usize` saturates at 0
      return (int)std::min<int64_t>(k, plane - 1);
    }
    int bx = d[0] >= 0.0, by = d[1] >= 0.0, bz = d[2] >= 0.0;
    return planar ? (bx << 1) | by : (bx << 2) | (by << 1) | bz;
  }
  int longest_empty(const std::vector<int64_t>& counts) const {
    std::vector<std::vector<int>> rings;
    if (plane > 0) {
      std::vector<int> r;
      for (int i = 0; i < plane; ++i) r.push_back(i);
      rings.push_back(r);
    } else if (planar) {
      rings.push_back({3, 1, 0, 2});
    } else {
      rings.push_back({7, 3, 1, 5});
      rings.push_back({6, 2, 0, 4});
    }
    int best = 0;
    for (auto& ring : rings) {
      bool all = true;
      for (int x : ring) all = all && counts[x] == 0;
      if (all) {
        best = std::max(best, (int)ring.size());
        continue;
      }
      int run = 0, longest = 0;
      for (int pass = 0; pass < 2; ++pass)
        for (int x : ring) {
          run = counts[x] == 0 ? run + 1 : 0;
          longest = std::max(longest, run);
        }
      best = std::max(best, longest);
    }
    return best;
  }
};

static Sectors make_sectors(const Aniso* a, bool planar, const Search& p) {
  Sectors s;
  if (a) {
    double rot[3] = {a->azimuth, a->dip, a->rake};
    s.axes = block_frame(rot);
  }
  s.planar = planar;
  s.plane = p.sectors;
  return s;
}

// search.rs:476-577
struct Selector {
  const P3* target;
  const Search* p;
  const Sectors* sec;
  std::vector<int64_t> chosen;
  std::vector<std::pair<int32_t, int64_t>> per_hole;
  int64_t untagged = 0;
  std::vector<int64_t> per_sector;
  int64_t cap;
  Selector(const P3* t, const Search* params, const Sectors* s) : target(t), p(params), sec(s) {
    per_sector.assign(p->balances() ? sec->count() : 0, 0);
    if (p->sectors > 0)
      cap = p->max_per_sector;
    else if (p->octant)
      cap = p->max_per_octant >= 0 ? p->max_per_octant : (p->max_samples + sec->count() - 1) / sec->count();
    else
      cap = std::numeric_limits<int64_t>::max();
  }
  bool full() const { return (int64_t)chosen.size() >= p->max_samples; }
  bool offer(int64_t idx, const P3& loc, int32_t hole) {
    if (full()) return true;
    int seen = -1;
    if (hole >= 0)
      for (size_t k = 0; k < per_hole.size(); ++k)
        if (per_hole[k].first == hole) {
          seen = (int)k;
          break;
        }
    if (p->max_per_hole >= 0 && seen >= 0 && per_hole[seen].second >= p->max_per_hole) return false;
    if (p->balances()) {
      int o = sec->of(*target, loc);
      if (per_sector[o] >= cap) return false;
      per_sector[o] += 1;
    }
    if (hole >= 0) {
      if (seen >= 0)
        per_hole[seen].second += 1;
      else
        per_hole.push_back({hole, 1});
    } else {
      untagged += 1;
    }
    chosen.push_back(idx);
    return full();
  }
  int finish() const {
    if ((int64_t)chosen.size() < p->min_samples) return BK_SEARCH_FAILED;
    if (p->min_sectors >= 0) {
      int64_t held = 0;
      for (auto c : per_sector) held += c > 0;
      if (held < p->min_sectors) return BK_SEARCH_FAILED;
    }
    if (p->max_empty_sectors >= 0) {
      if (sec->longest_empty(per_sector) > p->max_empty_sectors) return BK_SEARCH_FAILED;
    }
    if (p->min_holes >= 0) {
      if ((int64_t)per_hole.size() + untagged < p->min_holes) return BK_SEARCH_FAILED;
    }
    return BK_OK;
  }
};

// search.rs:704-745
struct Space {
  M3 frame = identity();
  bool has_origin = false;
  P3 origin{0, 0, 0};
  P3 project(const P3& p) const {
    P3 o = has_origin ? origin : P3{0, 0, 0};
    return mv(frame, {p[0] - o[0], p[1] - o[1], p[2] - o[2]});
  }
};

static Space make_space(const Aniso* a, const std::vector<Sample>& s) {
  Space sp;
  sp.frame = a ? a->matrix() : identity();
  for (auto& x : s)
    if (std::isfinite(x.loc[0]) && std::isfinite(x.loc[1]) && std::isfinite(x.loc[2])) {
      sp.origin = x.loc;
      sp.has_origin = true;
      break;
    }
  return sp;
}

// grid.rs:109-111
static inline double d2(const P3& p, const P3& q) {
  double a = p[0] - q[0], b = p[1] - q[1], c = p[2] - q[2];
  return (a * a + b * b) + c * c;
}

static bool planar_locs(const std::vector<Sample>& s) {
  if (s.empty()) return true;
  double z = s[0].loc[2];
  for (auto& x : s)
    if (!(x.loc[2] == z)) return false;
  return true;
}

static const Aniso* metric(const Search& p, const Variogram* vg) {
  if (p.has_aniso) return &p.aniso;
  if (vg && vg->has_aniso) return &vg->aniso;
  return nullptr;
}

// search.rs:647-687 (the scan); no domains, soft boundaries or high-grade rules.
// left_out >= 0 skips that sample as SearchTree::neighbors_around does
// (search.rs:997): the frame's origin and the planar test still see it.
int neighbors_scan(const P3& target, const std::vector<Sample>& s, const Search& p, const Variogram* vg,
                   int64_t left_out, std::vector<int64_t>& out) {
  const Aniso* a = metric(p, vg);
  Space sp = make_space(a, s);
  P3 q = sp.project(target);
  double r2 = p.radius * p.radius;
  std::vector<std::pair<double, int64_t>> cand;
  for (size_t i = 0; i < s.size(); ++i) {
    if ((int64_t)i == left_out) continue;
    double dd = d2(sp.project(s[i].loc), q);
    if (dd <= r2) cand.push_back({dd, (int64_t)i});
  }
  std::stable_sort(cand.begin(), cand.end(), [](auto& x, auto& y) { return x.first < y.first; });
  Sectors sec = make_sectors(a, p.octant && planar_locs(s), p);
  Selector sel(&target, &p, &sec);
  for (auto& c : cand)
    if (sel.offer(c.second, s[c.second].loc, s[c.second].hole)) break;
  out = sel.chosen;
  return sel.finish();
}

// crates/estimation/src/search/grid.rs
static const int64_t REACH = 12;

struct Table {
  std::vector<std::array<int64_t, 3>> offsets;
  std::vector<std::pair<uint64_t, size_t>> shells;
  Table() {
    std::vector<std::pair<uint64_t, std::array<int64_t, 3>>> keyed;
    for (int64_t x = -REACH; x <= REACH; ++x)
      for (int64_t y = -REACH; y <= REACH; ++y)
        for (int64_t z = -REACH; z <= REACH; ++z) {
          std::array<int64_t, 3> o{x, y, z};
          uint64_t m = 0;
          for (auto v : o) {
            int64_t a = std::max<int64_t>(std::llabs(v) - 1, 0);
            m += (uint64_t)(a * a);
          }
          keyed.push_back({m, o});
        }
    std::sort(keyed.begin(), keyed.end());
    for (size_t k = 0; k < keyed.size(); ++k) {
      if (!shells.empty() && shells.back().first == keyed[k].first)
        shells.back().second = k + 1;
      else
        shells.push_back({keyed[k].first, k + 1});
      offsets.push_back(keyed[k].second);
    }
  }
};

static const Table& table() {
  static Table t;
  return t;
}

// Rust `f64 as i64`: saturating, NaN -> 0.
static inline int64_t sat_i64(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 9.2233720368547758e18) return std::numeric_limits<int64_t>::max();
  if (v < -9.2233720368547758e18) return std::numeric_limits<int64_t>::min();
  return (int64_t)v;
}

struct Grid {
  P3 lo{0, 0, 0}, hi{0, 0, 0};
  double h = 1;
  int64_t dims[3] = {1, 1, 1};
  std::vector<uint32_t> start;
  std::vector<P3> ep;
  std::vector<uint32_t> ei;

  // grid.rs:83-106
  static double side(const P3& lo, const P3& hi, size_t n) {
    double ext[3];
    for (int d = 0; d < 3; ++d) ext[d] = hi[d] - lo[d];
    double longest = 0.0;
    for (double e : ext) longest = std::fmax(longest, e);
    if (longest <= 0.0) return 1.0;
    double prod = 1.0;
    int cnt = 0;
    for (double e : ext)
      if (e > longest * 1e-9) {
        prod *= e;
        cnt++;
      }
    double h = std::pow(prod / (double)std::max<size_t>(n, 1), 1.0 / (double)cnt);
    auto cells = [&](double hh) {
      double c = 1.0;
      for (double e : ext) c *= std::floor(e / hh) + 1.0;
      return c;
    };
    while (cells(h) > 4.0 * (double)n + 64.0) h *= 2.0;
    return h;
  }

  void build(const std::vector<P3>& pts) {
    // grid.rs:66-80 bounds over finite points
    P3 l{INFINITY, INFINITY, INFINITY}, u{-INFINITY, -INFINITY, -INFINITY};
    for (auto& p : pts) {
      if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]))) continue;
      for (int d = 0; d < 3; ++d) {
        l[d] = std::fmin(l[d], p[d]);
        u[d] = std::fmax(u[d], p[d]);
      }
    }
    if (!(l[0] <= u[0])) l = u = P3{0, 0, 0};
    lo = l;
    hi = u;
    h = side(lo, hi, pts.size());
    for (int d = 0; d < 3; ++d) dims[d] = sat_i64(std::floor((hi[d] - lo[d]) / h)) + 1;
    size_t cells = (size_t)(dims[0] * dims[1] * dims[2]);
    std::vector<size_t> ids(pts.size());
    start.assign(cells + 1, 0);
    for (size_t i = 0; i < pts.size(); ++i) {
      ids[i] = id(cell(pts[i]));
      start[ids[i] + 1]++;
    }
    for (size_t c = 0; c < cells; ++c) start[c + 1] += start[c];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    ep.resize(pts.size());
    ei.resize(pts.size());
    for (size_t i = 0; i < pts.size(); ++i) {
      ep[fill[ids[i]]] = pts[i];
      ei[fill[ids[i]]] = (uint32_t)i;
      fill[ids[i]]++;
    }
  }
  std::array<int64_t, 3> cell(const P3& p) const {
    std::array<int64_t, 3> c;
    for (int d = 0; d < 3; ++d) c[d] = std::min(std::max(sat_i64(std::floor((p[d] - lo[d]) / h)), (int64_t)0), dims[d] - 1);
    return c;
  }
  size_t id(const std::array<int64_t, 3>& c) const { return (size_t)(c[0] + dims[0] * (c[1] + dims[1] * c[2])); }

  using Key = std::pair<uint64_t, uint32_t>;
  using Heap = std::priority_queue<Key, std::vector<Key>, std::greater<Key>>;

  void scan(std::array<int64_t, 3> c, const P3& q, double r2, Heap& heap) const {
    for (int d = 0; d < 3; ++d)
      if (c[d] < 0 || c[d] >= dims[d]) return;
    size_t k = id(c);
    for (uint32_t e = start[k]; e < start[k + 1]; ++e) {
      double dd = d2(ep[e], q);
      if (dd <= r2) {
        uint64_t bits;
        std::memcpy(&bits, &dd, 8);
        heap.push({bits, ei[e]});
      }
    }
  }
  template <class F>
  static bool release(Heap& heap, double bound, F& visit) {
    while (!heap.empty()) {
      Key top = heap.top();
      double dd;
      std::memcpy(&dd, &top.first, 8);
      if (dd >= bound) return false;
      heap.pop();
      if (visit(dd, (int64_t)top.second)) return true;
    }
    return false;
  }
  void ring(std::array<int64_t, 3> c, int64_t r, const P3& q, double r2, Heap& heap) const {
    auto lo_ = [&](int d) { return std::max(-r, -c[d]); };
    auto hi_ = [&](int d) { return std::min(r, dims[d] - 1 - c[d]); };
    for (int64_t x = lo_(0); x <= hi_(0); ++x)
      for (int64_t y = lo_(1); y <= hi_(1); ++y) {
        if (std::llabs(x) == r || std::llabs(y) == r) {
          for (int64_t z = lo_(2); z <= hi_(2); ++z) scan({c[0] + x, c[1] + y, c[2] + z}, q, r2, heap);
        } else {
          for (int64_t z : {-r, r})
            if (z >= lo_(2) && z <= hi_(2)) scan({c[0] + x, c[1] + y, c[2] + z}, q, r2, heap);
        }
      }
  }
  // grid.rs:209-255
  template <class F>
  void nearest(const P3& q, double r2, F visit) const {
    const Table& t = table();
    auto c = cell(q);
    double h2 = h * h;
    double beyond = h2 * (double)((REACH - 1) * (REACH - 1));
    Heap heap;
    size_t begin = 0;
    for (auto& sh : t.shells) {
      double bound = std::fmin(h2 * (double)sh.first, beyond) * (1.0 - 1e-9);
      if (bound > r2) break;
      if (release(heap, bound, visit)) return;
      for (size_t k = begin; k < sh.second; ++k) {
        auto& o = t.offsets[k];
        scan({c[0] + o[0], c[1] + o[1], c[2] + o[2]}, q, r2, heap);
      }
      begin = sh.second;
    }
    int64_t far = 0;
    for (int d = 0; d < 3; ++d) far = std::max(far, std::max(c[d], dims[d] - 1 - c[d]));
    for (int64_t r = REACH + 1; r <= far; ++r) {
      double x = (double)(r - 1) * h;
      double bound = x * x * (1.0 - 1e-9);
      if (bound > r2) break;
      if (release(heap, bound, visit)) return;
      ring(c, r, q, r2, heap);
    }
    release(heap, INFINITY, visit);
  }
};

// search.rs:692-1006, without domains, soft boundaries, high grade or calibration.
struct SearchTree {
  const std::vector<Sample>* s;
  Search p;
  Space space;
  Sectors sectors;
  Grid grid;
  std::vector<P3> points;
  SearchTree(const std::vector<Sample>& samples, const Search& params, const Variogram* vg) : s(&samples), p(params) {
    const Aniso* a = metric(p, vg);
    space = make_space(a, samples);
    sectors = make_sectors(a, planar_locs(samples), p);
    points.resize(samples.size());
    for (size_t i = 0; i < samples.size(); ++i) points[i] = space.project(samples[i].loc);
    grid.build(points);
  }
  // search.rs:909-936: `chosen` round-robin over the sectors around the
  // target (each sector in selection order, a stable sort by the rank within
  // its sector), for calibrated searches that balance sectors.
  std::vector<int64_t> balanced(const P3& target, const std::vector<int64_t>& chosen) const {
    if (!p.balances() || p.calibration == 0) return chosen;
    std::vector<int64_t> seen(sectors.count(), 0);
    std::vector<std::pair<int64_t, int64_t>> ranked;
    for (int64_t i : chosen) {
      int o = sectors.of(target, (*s)[i].loc);
      seen[o] += 1;
      ranked.push_back({seen[o], i});
    }
    std::stable_sort(ranked.begin(), ranked.end(),
                     [](const std::pair<int64_t, int64_t>& a, const std::pair<int64_t, int64_t>& b) {
                       return a.first < b.first;
                     });
    std::vector<int64_t> out;
    for (auto& r : ranked) out.push_back(r.second);
    return out;
  }
  // search.rs:951-982. limit >= 0: only samples below that index are
  // candidates; sec: sectors in place of the tree's.
  int neighbors_around(const P3& target, int64_t left_out, std::vector<int64_t>& out, int64_t limit = -1,
                       const Sectors* sec = nullptr) const {
    Selector sel(&target, &p, sec ? sec : &sectors);
    if (s->empty() || p.max_samples == 0) {
      out = sel.chosen;
      return sel.finish();
    }
    P3 q = space.project(target);
    std::vector<int64_t> group;
    double at = NAN;
    bool full = false;
    auto offer = [&]() {
      bool f = false;
      for (int64_t i : group) {
        if (i == left_out || (limit >= 0 && i >= limit)) continue;
        if (sel.offer(i, (*s)[i].loc, (*s)[i].hole)) {
          f = true;
          break;
        }
      }
      group.clear();
      return f;
    };
    grid.nearest(q, p.radius * p.radius, [&](double dd, int64_t i) {
      if (dd != at && !group.empty()) {
        full = offer();
        if (full) return true;
      }
      at = dd;
      group.push_back(i);
      return false;
    });
    if (!full) offer();
    out = sel.chosen;
    return sel.finish();
  }
};

// The leading samples a calibrated search keeps (py/estimation.rs:846-897 with
// kriging_algebra.rs:421-441), for point support (pts null) or the block
// discretised by pts; all of them without calibration.
static std::pair<int, bool> calibrated_nm(const Search& p, int kind, const P3& t, const std::vector<P3>* pts,
                                          const std::vector<Sample>& sel, const Variogram& vg);

static int calibrated_n(const Search& p, int kind, const P3& t, const std::vector<P3>* pts,
                        const std::vector<Sample>& sel, const Variogram& vg) {
  return calibrated_nm(p, kind, t, pts, sel, vg).first;
}

// The count kept and whether it meets the target (py/estimation.rs:846-893).
static std::pair<int, bool> calibrated_nm(const Search& p, int kind, const P3& t, const std::vector<P3>* pts,
                                          const std::vector<Sample>& sel, const Variogram& vg) {
  if (p.calibration == 0 || sel.empty()) return {(int)sel.size(), false};
  std::vector<double> rhs;
  double support;
  if (pts) {
    rhs = block_rhs(*pts, sel, vg);
    support = vg.block_mean_cov(*pts);
  } else {
    rhs.resize(sel.size());
    for (size_t i = 0; i < sel.size(); ++i) rhs[i] = vg.cov_points(sel[i].loc, t);
    support = vg.total_sill();
  }
  auto q = prefix_quality(sel, vg, rhs, kind != SIMPLE, support);
  return calibrated_count(p.calibration, p.calibration_target, p.min_samples, q);
}

// Simple kriging has a slope of 1 by construction (kriging_algebra.rs:429-434).
static bool calibration_fits(const Search& p, int kind) {
  return !(kind == SIMPLE && p.calibration == BK_CALIBRATE_SLOPE);
}

static std::vector<Sample> make_samples(int64_t n, const double* x, const double* y, const double* z,
                                        const double* v, const double* err, const int32_t* hole) {
  std::vector<Sample> s(n);
  for (int64_t i = 0; i < n; ++i) {
    s[i].loc = {x[i], y[i], z[i]};
    s[i].value = v ? v[i] : 0.0;
    s[i].err = err ? err[i] : 0.0;
    s[i].hole = hole ? hole[i] : -1;
  }
  return s;
}

static void write_estimate(int st, const Estimate& e, int64_t t, double* value, double* variance, double* lagrange,
                           double* support, int32_t* n_used, int32_t* status) {
  bool ok = st == BK_OK;
  if (value) value[t] = ok ? e.value : NAN;
  if (variance) variance[t] = ok ? e.variance : NAN;
  if (lagrange) lagrange[t] = ok ? e.lagrange : NAN;
  if (support) support[t] = ok ? e.support : NAN;
  if (n_used) n_used[t] = ok ? e.n_used : 0;
  if (status) status[t] = st;
}


// ---------------------------------------------------------------------------
// Multiple indicator kriging, crates/estimation/src/indicator.rs.

// f64::total_cmp as an integer key.
static inline int64_t total_key(double v) {
  int64_t b;
  std::memcpy(&b, &v, sizeof(b));
  return b ^ (int64_t)((uint64_t)(b >> 63) >> 1);
}

// indicator.rs:69-81 Global, :152-212 Global::new / weight.
struct Global {
  std::vector<double> values;
  std::vector<double> sums[3];
  std::vector<std::pair<size_t, size_t>> classes;
  double lower = 0, upper = 0;
  std::vector<double> proportions;
  double weight(std::pair<size_t, size_t> r) const { return sums[0][r.second] - sums[0][r.first]; }
};

// Returns BK_OK, BK_INSUFFICIENT_DATA (no values) or BK_E_INVALID (weights).
static int global_new(const std::vector<double>& values, const double* weights, const std::vector<double>& thr,
                      const double* tails, Global& g) {
  const size_t n = values.size();
  if (n == 0) return BK_INSUFFICIENT_DATA;
  std::vector<double> w(n, 1.0);
  if (weights) w.assign(weights, weights + n);
  for (double x : w)
    if (!std::isfinite(x) || x < 0.0) return BK_E_INVALID;
  double total = -0.0;
  for (double x : w) total += x;
  if (total <= 0.0) return BK_E_INVALID;
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t i, size_t j) {
    int64_t a = total_key(values[i]), b = total_key(values[j]);
    return a != b ? a < b : i < j;
  });
  g.values.resize(n);
  for (size_t k = 0; k < n; ++k) g.values[k] = values[order[k]];
  for (int j = 0; j < 3; ++j) g.sums[j].assign(1, 0.0);
  for (size_t k = 0; k < n; ++k) {
    const double wk = w[order[k]] / total, v = g.values[k];
    for (int j = 0; j < 3; ++j) g.sums[j].push_back(g.sums[j].back() + wk * powi(v, j));
  }
  std::vector<size_t> cuts;
  for (double t : thr)
    cuts.push_back((size_t)(std::upper_bound(g.values.begin(), g.values.end(), t,
                                             [](double tt, double v) { return !(v <= tt); }) -
                            g.values.begin()));
  std::vector<size_t> bounds{0};
  bounds.insert(bounds.end(), cuts.begin(), cuts.end());
  bounds.push_back(n);
  const double first = thr.front(), last = thr.back();
  const double lo = tails ? tails[0] : g.values.front(), hi = tails ? tails[1] : g.values.back();
  g.lower = std::fmin(std::fmin(lo, g.values.front()), first);
  g.upper = std::fmax(std::fmax(hi, g.values.back()), last);
  g.proportions.clear();
  for (size_t c : cuts) g.proportions.push_back(g.sums[0][c]);
  g.classes.clear();
  for (size_t k = 0; k + 1 < bounds.size(); ++k) g.classes.push_back({bounds[k], bounds[k + 1]});
  return BK_OK;
}

// indicator.rs:137-150
static double correct_order_relations(const std::vector<double>& raw, std::vector<double>& cdf) {
  const size_t k = raw.size();
  std::vector<double> up(k), down(k);
  for (size_t i = 0; i < k; ++i) {
    double p = raw[i];
    p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);  // f64::clamp
    up[i] = down[i] = p;
  }
  for (size_t i = 1; i < k; ++i) up[i] = std::fmax(up[i], up[i - 1]);
  for (size_t i = k - 1; i-- > 0;) down[i] = std::fmin(down[i], down[i + 1]);
  cdf.resize(k);
  for (size_t i = 0; i < k; ++i) cdf[i] = 0.5 * (up[i] + down[i]);
  double correction = -0.0;
  for (size_t i = 0; i < k; ++i) correction += std::fabs(cdf[i] - raw[i]);
  return correction;
}

// indicator.rs:206-213
enum ShapeKind { SHAPE_DATA = 0, SHAPE_POWER = 1, SHAPE_HYPERBOLIC = 2 };
struct Shape {
  int kind;
  size_t s = 0, e = 0;
  double w = 1.0;
};

// indicator.rs:215-217
static double binomial(int n, int k) {
  double acc = 1.0;
  for (int i = 0; i < k; ++i) acc = acc * (double)(n - i) / (double)(i + 1);
  return acc;
}

static inline double clampd(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }

// Number of leading entries of v[lo .. hi) for which pred holds (slice::partition_point).
template <class F>
static size_t partition_point(const std::vector<double>& v, size_t lo, size_t hi, F pred) {
  size_t a = lo, b = hi;
  while (a < b) {
    size_t mid = a + (b - a) / 2;
    if (pred(v[mid]))
      a = mid + 1;
    else
      b = mid;
  }
  return a - lo;
}

// indicator.rs:220-253 Shape::above: E[Z^j 1{Z > c}] under the class distribution on [a, b].
static double shape_above(const Shape& sh, const Global& g, double a, double b, double c, int j) {
  if (sh.kind == SHAPE_DATA) {
    size_t p = sh.s + partition_point(g.values, sh.s, sh.e, [&](double v) { return v <= c; });
    return (g.sums[j][sh.e] - g.sums[j][p]) / g.weight({sh.s, sh.e});
  }
  if (b <= a) return a > c ? powi(a, j) : 0.0;
  const double w = sh.w;
  if (sh.kind == SHAPE_POWER) {
    double x0 = clampd((c - a) / (b - a), 0.0, 1.0);
    double sum = -0.0;
    for (int i = 0; i <= j; ++i) {
      double e = w + (double)i;
      sum += binomial(j, i) * powi(a, j - i) * powi(b - a, i) * w / e * (1.0 - std::pow(x0, e));
    }
    return sum;
  }
  double yb = b / a, yc = clampd(c, a, b) / a;
  double m = (double)j - w;
  double integral = std::fabs(m) < 1e-12 ? std::log(yb / yc) : (std::pow(yb, m) - std::pow(yc, m)) / m;
  return powi(a, j) * w / (1.0 - std::pow(yb, -w)) * integral;
}

// indicator.rs:272-289 Shape::quantile: the value at probability u within the class.
static double shape_quantile(const Shape& sh, const Global& g, double a, double b, double u) {
  if (sh.kind == SHAPE_DATA) {
    double target = g.sums[0][sh.s] + u * g.weight({sh.s, sh.e});
    size_t i = sh.s + partition_point(g.sums[0], sh.s + 1, sh.e + 1, [&](double c) { return c < target; });
    return g.values[std::min(i, sh.e - 1)];
  }
  if (b <= a) return a;
  if (sh.kind == SHAPE_POWER) return clampd(a + (b - a) * std::pow(u, 1.0 / sh.w), a, b);
  double r = std::pow(a / b, sh.w);
  return clampd(a * std::pow(1.0 - u * (1.0 - r), -1.0 / sh.w), a, b);
}

struct Mik {
  std::vector<double> thr;
  std::vector<Variogram> vgs;
  bool simple = false;
  int interpolation = BK_MIK_GLOBAL;
  int upper_tail = BK_TAIL_NONE;
  double omega = 1.0;
  bool has_tails = false;
  double tails[2] = {0, 0};
};

// indicator.rs:291-317 MultipleIndicator::validate.
static bool mik_valid(const bk_mik* p, const double* thr) {
  const int k = p->n_thresholds;
  if (k < 1) return false;
  for (int i = 0; i < k; ++i)
    if (!std::isfinite(thr[i])) return false;
  for (int i = 1; i < k; ++i)
    if (thr[i - 1] >= thr[i]) return false;
  if (p->n_variograms != 1 && p->n_variograms != k) return false;
  if (p->has_tails && !(std::isfinite(p->tails[0]) && std::isfinite(p->tails[1]) && p->tails[0] <= p->tails[1]))
    return false;
  if (p->upper_tail != BK_TAIL_NONE && !(std::isfinite(p->omega) && p->omega > 0.0)) return false;
  if (p->upper_tail == BK_TAIL_HYPERBOLIC && thr[k - 1] <= 0.0) return false;
  if (p->interpolation != BK_MIK_GLOBAL && p->interpolation != BK_MIK_LINEAR) return false;
  if (p->upper_tail < BK_TAIL_NONE || p->upper_tail > BK_TAIL_HYPERBOLIC) return false;
  return true;
}

static Mik make_mik(const bk_mik* p, const double* thr, const bk_variogram* vgs) {
  Mik m;
  m.thr.assign(thr, thr + p->n_thresholds);
  for (int v = 0; v < p->n_variograms; ++v) m.vgs.push_back(make_variogram(&vgs[v]));
  m.simple = p->simple != 0;
  m.interpolation = p->interpolation;
  m.upper_tail = p->upper_tail;
  m.omega = p->omega;
  m.has_tails = p->has_tails != 0;
  m.tails[0] = p->tails[0];
  m.tails[1] = p->tails[1];
  return m;
}

struct ClassRow {
  double a, b, q;
  Shape shape;
};

// indicator.rs:347-378 MultipleIndicator::classes.
static std::vector<ClassRow> mik_classes(const Mik& mk, const Global& g, const std::vector<double>& cdf) {
  const size_t k = mk.thr.size();
  std::vector<ClassRow> out;
  for (size_t c = 0; c <= k; ++c) {
    ClassRow r;
    r.a = c == 0 ? g.lower : mk.thr[c - 1];
    r.b = c == k ? g.upper : mk.thr[c];
    double lower = c == 0 ? 0.0 : cdf[c - 1];
    double upper = c == k ? 1.0 : cdf[c];
    r.q = upper - lower;
    auto range = g.classes[c];
    if (mk.upper_tail == BK_TAIL_POWER && c == k) {
      r.shape.kind = SHAPE_POWER;
      r.shape.w = mk.omega;
    } else if (mk.upper_tail == BK_TAIL_HYPERBOLIC && c == k) {
      r.shape.kind = SHAPE_HYPERBOLIC;
      r.shape.w = mk.omega;
    } else if (mk.interpolation == BK_MIK_GLOBAL && g.weight(range) > 0.0) {
      r.shape.kind = SHAPE_DATA;
      r.shape.s = range.first;
      r.shape.e = range.second;
    } else {
      r.shape.kind = SHAPE_POWER;
      r.shape.w = 1.0;
    }
    out.push_back(r);
  }
  return out;
}

// indicator.rs:83-99 Conditional.
struct Conditional {
  std::vector<double> cdf;
  double correction = 0;
  int violations = 0;
  double mean = 0, variance = 0;
  std::vector<double> probability_above, mean_above, quantile_values;
};

// indicator.rs:380-437 MultipleIndicator::conditional.
static Conditional mik_conditional(const Mik& mk, const Global& g, const std::vector<double>& raw,
                                   const std::vector<double>& cutoffs, const std::vector<double>& quantiles) {
  Conditional out;
  out.correction = correct_order_relations(raw, out.cdf);
  for (size_t i = 0; i < raw.size(); ++i) out.violations += out.cdf[i] != raw[i];
  auto classes = mik_classes(mk, g, out.cdf);
  auto moment = [&](double c, int j) {
    double sum = -0.0;
    for (auto& r : classes)
      if (r.q > 0.0) sum += r.q * shape_above(r.shape, g, r.a, r.b, c, j);
    return sum;
  };
  const double ninf = -std::numeric_limits<double>::infinity();
  out.mean = moment(ninf, 1);
  out.variance = std::fmax(moment(ninf, 2) - out.mean * out.mean, 0.0);
  for (double c : cutoffs) out.probability_above.push_back(moment(c, 0));
  for (size_t i = 0; i < cutoffs.size(); ++i) {
    double p = out.probability_above[i];
    out.mean_above.push_back(p > 0.0 ? moment(cutoffs[i], 1) / p : NAN);
  }
  for (double u : quantiles) {
    double below = 0.0;
    size_t last = 0;
    for (size_t i = classes.size(); i-- > 0;)
      if (classes[i].q > 0.0) {
        last = i;
        break;
      }
    double v = g.upper;
    for (size_t i = 0; i < classes.size(); ++i) {
      const ClassRow& r = classes[i];
      if (r.q > 0.0 && (below + r.q >= u || i == last)) {
        double within = clampd((u - below) / r.q, 0.0, 1.0);
        v = shape_quantile(r.shape, g, r.a, r.b, within);
        break;
      }
      below += r.q;
    }
    out.quantile_values.push_back(v);
  }
  return out;
}

// indicator.rs:688-749 krige_indicators, with `indicator(k, s) = s.value <= t_k`
// (:327-345 MultipleIndicator::kriged). pts: block discretisation or null.
static int krige_indicators(const Mik& mk, const P3& target, const std::vector<Sample>& s,
                            const std::vector<P3>* pts, const std::vector<double>& means, std::vector<double>& raw) {
  const int kind = mk.simple ? SIMPLE : ORDINARY;
  std::vector<std::vector<double>> weights;
  for (const Variogram& vg : mk.vgs) {
    Estimate e;
    int st = pts ? block_krige_points(kind, 0.0, *pts, s, vg, e) : krige(kind, 0.0, target, s, vg, e);
    if (st != BK_OK) return st;
    weights.push_back(e.weights);
  }
  raw.assign(means.size(), 0.0);
  for (size_t k = 0; k < means.size(); ++k) {
    const std::vector<double>& w = weights[std::min(k, weights.size() - 1)];
    const double f = means[k];
    double sum = -0.0;
    for (size_t j = 0; j < s.size(); ++j) {
      double i = s[j].value <= mk.thr[k] ? 1.0 : 0.0;
      sum += mk.simple ? w[j] * (i - f) : w[j] * i;
    }
    raw[k] = mk.simple ? f + sum : sum;
  }
  return BK_OK;
}

static void write_conditional(const Conditional* c, int st, int n, int64_t t, int64_t m, const bk_mik_out* o,
                              const std::vector<double>* raw, size_t k, size_t nc, size_t nq) {
  auto put = [&](double* a, size_t rows, const std::vector<double>* v) {
    if (!a) return;
    for (size_t r = 0; r < rows; ++r) a[r * m + t] = c ? (*v)[r] : NAN;
  };
  put(o->raw, k, raw);
  put(o->cdf, k, c ? &c->cdf : nullptr);
  put(o->probability_above, nc, c ? &c->probability_above : nullptr);
  put(o->mean_above, nc, c ? &c->mean_above : nullptr);
  put(o->quantile_values, nq, c ? &c->quantile_values : nullptr);
  if (o->correction) o->correction[t] = c ? c->correction : NAN;
  if (o->mean) o->mean[t] = c ? c->mean : NAN;
  if (o->variance) o->variance[t] = c ? c->variance : NAN;
  if (o->violations) o->violations[t] = c ? c->violations : 0;
  if (o->n_used) o->n_used[t] = c ? n : 0;
  if (o->status) o->status[t] = st;
}


// ---------------------------------------------------------------------------
// neighborhood_stats, crates/estimation/src/neighborhood.rs:18-136 (Euclidean
// distance; the anisotropic metric is not used by the diagnostics).
struct NeighborhoodStats {
  int64_t n_within = 0, n_holes = 0;
  double nearest_dist = INFINITY, mean_dist_knn = INFINITY, max_dist_knn = INFINITY;
  double value_mean = NAN, value_var = NAN, value_min = NAN, value_max = NAN, value_idw = NAN;
};

static NeighborhoodStats neighborhood_stats(const P3& target, const std::vector<Sample>& s, size_t k, double radius) {
  NeighborhoodStats st;
  if (s.empty() || k == 0) return st;
  struct Row {
    double d, v;
    int32_t hole;
  };
  std::vector<Row> d;
  for (const Sample& x : s) d.push_back({euclidean(target, x.loc), x.value, x.hole});
  std::stable_sort(d.begin(), d.end(), [](const Row& a, const Row& b) { return a.d < b.d; });
  std::vector<int32_t> seen;
  for (const Row& r : d) {
    if (!(r.d <= radius)) continue;
    st.n_within++;
    if (r.hole < 0) {
      st.n_holes++;
    } else if (std::find(seen.begin(), seen.end(), r.hole) == seen.end()) {
      seen.push_back(r.hole);
      st.n_holes++;
    }
  }
  st.nearest_dist = d[0].d;
  const size_t kk = std::min(k, d.size());
  double sd = -0.0;
  for (size_t i = 0; i < kk; ++i) sd += d[i].d;
  st.mean_dist_knn = sd / (double)kk;
  st.max_dist_knn = d[kk - 1].d;
  double sv = -0.0;
  for (size_t i = 0; i < kk; ++i) sv += d[i].v;
  st.value_mean = sv / (double)kk;
  double ss = -0.0;
  for (size_t i = 0; i < kk; ++i) ss += powi(d[i].v - st.value_mean, 2);
  st.value_var = ss / (double)kk;
  double lo = INFINITY, hi = -INFINITY;
  for (size_t i = 0; i < kk; ++i) {
    lo = std::fmin(lo, d[i].v);
    hi = std::fmax(hi, d[i].v);
  }
  st.value_min = lo;
  st.value_max = hi;
  double num = 0.0, den = 0.0;
  bool exact = false;
  double hit = 0.0;
  for (size_t i = 0; i < kk; ++i) {
    if (d[i].d < 1e-12) {
      exact = true;
      hit = d[i].v;
      break;
    }
    double w = 1.0 / (d[i].d * d[i].d);
    num += w * d[i].v;
    den += w;
  }
  st.value_idw = exact ? hit : (den > 0.0 ? num / den : st.value_mean);
  return st;
}

// krige.rs:64-75 Estimate::negative_weight_sum (NaN without weights).
static double negative_weight_sum(const std::vector<double>& w) {
  if (w.empty()) return NAN;
  double s = -0.0;
  for (double x : w)
    if (x < 0.0) s += x;
  return s;
}

}  // namespace orc

using namespace orc;

extern "C" {

typedef struct {
  int64_t n;
  const double *x, *y, *z, *value, *err;
  const int32_t* hole;
} or_samples;

static std::vector<Sample> samples_of(const or_samples* s) {
  return make_samples(s->n, s->x, s->y, s->z, s->value, s->err, s->hole);
}

void or_rotation_matrix(double az, double dip, double rake, double* out) {
  M3 r = rotation_matrix(az, dip, rake);
  std::memcpy(out, r.m, sizeof(r.m));
}

void or_block_frame(const double* rot, double* out) {
  M3 r = block_frame(rot);
  std::memcpy(out, r.m, sizeof(r.m));
}

void or_aniso_matrix(double az, double dip, double rake, double major, double semi, double minor, double* out) {
  Aniso a;
  a.azimuth = az;
  a.dip = dip;
  a.rake = rake;
  a.major = major;
  a.semi = semi;
  a.minor = minor;
  a.init();
  M3 m = a.matrix();
  std::memcpy(out, m.m, sizeof(m.m));
}

double or_shape(int model, double param, double h, double range) { return shape(model, param, h, range); }

void or_d_shape(int model, double param, double t, double* out) {
  auto d = d_shape(model, param, t);
  out[0] = d.first;
  out[1] = d.second;
}

double or_cov(const bk_variogram* v, double h) { return make_variogram(v).cov(h); }
double or_gamma(const bk_variogram* v, double h) { return make_variogram(v).gamma(h); }
double or_total_sill(const bk_variogram* v) { return make_variogram(v).total_sill(); }

double or_cov_points(const bk_variogram* v, const double* p, const double* q) {
  return make_variogram(v).cov_points({p[0], p[1], p[2]}, {q[0], q[1], q[2]});
}

double or_block_cov_points(const bk_variogram* v, const double* p, const double* q) {
  return make_variogram(v).block_cov_points({p[0], p[1], p[2]}, {q[0], q[1], q[2]});
}

int or_check_search(const bk_search* p) { return check(make_search(p)) ? 0 : -1; }

void or_block_points(const double* t, const bk_block* b, double* out) {
  auto pts = block_points({t[0], t[1], t[2]}, *b);
  for (size_t i = 0; i < pts.size(); ++i)
    for (int d = 0; d < 3; ++d) out[3 * i + d] = pts[i][d];
}

// One system: solver 0 = krige (guard, kernel, LU), 1 = kernel only (-1 when
// beyond it), 2 = LU only. weights may be NULL (else n entries).
int or_krige_system(const or_samples* sm, const bk_variogram* v, int kind, double param, const double* t,
                    int solver, double* value, double* variance, double* lagrange, double* weights,
                    int32_t* n_used) {
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
  Estimate e;
  P3 tt{t[0], t[1], t[2]};
  int st = solver == 1 ? krige_kernel(kind, param, tt, s, vg, e)
                       : (solver == 2 ? krige_lu(kind, param, tt, s, vg, e) : krige(kind, param, tt, s, vg, e));
  if (st == BK_OK) {
    *value = e.value;
    *variance = e.variance;
    *lagrange = e.lagrange;
    *n_used = e.n_used;
    if (weights)
      for (size_t i = 0; i < e.weights.size(); ++i) weights[i] = e.weights[i];
  }
  return st;
}

int or_block_krige_system(const or_samples* sm, const bk_variogram* v, int kind, double param, const double* pts,
                          int64_t npts, double* value, double* variance, double* lagrange, double* support,
                          double* weights, int32_t* n_used) {
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
  std::vector<P3> p(npts);
  for (int64_t i = 0; i < npts; ++i) p[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  Estimate e;
  int st = block_krige_points(kind, param, p, s, vg, e);
  if (st == BK_OK) {
    *value = e.value;
    *variance = e.variance;
    *lagrange = e.lagrange;
    *support = e.support;
    *n_used = e.n_used;
    if (weights)
      for (size_t i = 0; i < e.weights.size(); ++i) weights[i] = e.weights[i];
  }
  return st;
}

// Neighbour lists: use_tree 0 = the scan (search.rs:647-687), 1 = the grid tree.
int or_neighbors(const or_samples* sm, const bk_search* sp, const bk_variogram* v, int64_t m, const double* tx,
                 const double* ty, const double* tz, const int64_t* left_out, int use_tree, int32_t* idx,
                 int32_t* count, int32_t* status, int nthreads) {
  Search p = make_search(sp);
  if (!check(p)) return BK_E_INVALID;
  auto s = samples_of(sm);
  Variogram vg;
  if (v) vg = make_variogram(v);
  const Variogram* pv = v ? &vg : nullptr;
  SearchTree* tree = use_tree ? new SearchTree(s, p, pv) : nullptr;
  int64_t stride = p.max_samples;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t t = 0; t < m; ++t) {
    std::vector<int64_t> out;
    P3 q{tx[t], ty[t], tz[t]};
    int64_t lo = left_out ? left_out[t] : -1;
    int st = tree ? tree->neighbors_around(q, lo, out) : neighbors_scan(q, s, p, pv, lo, out);
    count[t] = (int32_t)out.size();
    status[t] = st;
    for (size_t k = 0; k < out.size(); ++k) idx[t * stride + k] = (int32_t)out[k];
  }
  delete tree;
  return 0;
}

// Sequential Gaussian simulation in normal scores, the node loop of
// crates/simulation/src/sgs.rs:600-750 run one node at a time (the reference's
// batches reproduce this sequential realization for any batch size): one
// search pass, no domains, trend, high-grade caps, local anisotropy or
// secondary variable. The samples are the data with their scores as values;
// vg is scaled to a total sill of 1 (unit_sill, sgs.rs:881-886). Node path[p]
// searches the data and the nodes added before it (the scan order of
// search.rs:647-687; SearchTree::add, search.rs:785-797, grows the set and its
// planar flag as the scan's planar test over the current set does); a datum at
// its location gives the node that datum's score (sgs.rs:665-674, the first
// datum at the location when the search fails: datum_index, sgs.rs:868-875);
// without neighbours the node takes the draw itself (marginal, sgs.rs:688-691);
// otherwise krige(Simple { mean: 0 }) and score = value + sqrt(max(variance, 0))
// z (sgs.rs:693-705, 466-469). Every node not on a datum takes the next draw
// from z and joins the samples with its score, no hole and no error variance.
// source[i]: the datum index, -1 kriged, -2 marginal, -3 the kriging failed
// (score NaN).
int or_sgs(const or_samples* sm, const bk_variogram* v, const bk_search* sp, int64_t m, const double* gx,
           const double* gy, const double* gz, const int64_t* path, const double* z, double* score, int32_t* source) {
  Search p = make_search(sp);
  if (!check(p)) return BK_E_INVALID;
  auto s = samples_of(sm);
  if (s.empty()) return BK_E_INVALID;
  bk_variogram vu = *v;
  double total = vu.nugget;
  for (int k = 0; k < vu.n_structures; ++k) total += vu.structures[k].sill;
  if (total > 0.0) {
    const double f = 1.0 / total;
    vu.nugget *= f;
    for (int k = 0; k < vu.n_structures; ++k) vu.structures[k].sill *= f;
  }
  Variogram vg = make_variogram(&vu);
  const int64_t n = (int64_t)s.size();
  auto same = [](const P3& a, const P3& b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; };
  int64_t draw = 0;
  for (int64_t k = 0; k < m; ++k) {
    const int64_t node = path[k];
    const P3 target{gx[node], gy[node], gz[node]};
    std::vector<int64_t> idx;
    const int st = neighbors_scan(target, s, p, &vg, -1, idx);
    int64_t datum = -1;
    if (st == BK_OK) {
      for (int64_t i : idx)
        if (same(s[i].loc, target)) {
          datum = i;
          break;
        }
    } else {
      for (int64_t i = 0; i < n; ++i)
        if (same(s[i].loc, target)) {
          datum = i;
          break;
        }
    }
    if (datum >= 0) {
      score[node] = s[datum].value;
      source[node] = (int32_t)datum;
      continue;
    }
    double sc;
    if (st != BK_OK || idx.empty()) {
      sc = z[draw];
      source[node] = -2;
    } else {
      std::vector<Sample> sel;
      for (int64_t i : idx) sel.push_back(s[i]);
      Estimate e;
      if (krige(SIMPLE, 0.0, target, sel, vg, e) == BK_OK) {
        sc = e.value + std::sqrt(std::max(e.variance, 0.0)) * z[draw];
        source[node] = -1;
      } else {
        sc = NAN;
        source[node] = -3;
      }
    }
    draw++;
    score[node] = sc;
    s.push_back(Sample{target, sc, -1, 0.0});
  }
  return 0;
}

// The realization of or_sgs computed as a plan per node (the constant and
// node weights of crates/simulation/src/shared.rs:104-143): node k of the path
// that is not on a datum searches, over one grid of the data followed by every
// such node in path order, the samples below index n + k -- those or_sgs has
// by then -- with the sectors planar while the data and those nodes share the
// first datum's elevation (search.rs:785-797); the searches and simple-kriging
// weights run in parallel (OpenMP), the draws in path order. A datum lies
// under a node where a datum has the node's location (or_sgs takes it then,
// with caps of at least 1). Used as the CPU baseline; checked against or_sgs.
int or_sgs_plan(const or_samples* sm, const bk_variogram* v, const bk_search* sp, int64_t m, const double* gx,
                const double* gy, const double* gz, const int64_t* path, const double* z, double* score,
                int32_t* source, int nthreads) {
  Search p = make_search(sp);
  if (!check(p)) return BK_E_INVALID;
  auto s = samples_of(sm);
  if (s.empty()) return BK_E_INVALID;
  bk_variogram vu = *v;
  double total = vu.nugget;
  for (int k = 0; k < vu.n_structures; ++k) total += vu.structures[k].sill;
  if (total > 0.0) {
    const double f = 1.0 / total;
    vu.nugget *= f;
    for (int k = 0; k < vu.n_structures; ++k) vu.structures[k].sill *= f;
  }
  Variogram vg = make_variogram(&vu);
  const int64_t n = (int64_t)s.size();
  auto key = [](const P3& q) {
    std::array<uint64_t, 3> k;
    for (int d = 0; d < 3; ++d) {
      const double x = q[d] + 0.0;
      std::memcpy(&k[d], &x, 8);
    }
    return k;
  };
  std::map<std::array<uint64_t, 3>, int64_t> first;
  for (int64_t i = 0; i < n; ++i) first.emplace(key(s[i].loc), i);
  std::vector<int64_t> app, datum(m, -1);
  for (int64_t k = 0; k < m; ++k) {
    const int64_t node = path[k];
    auto it = first.find(key(P3{gx[node], gy[node], gz[node]}));
    if (it != first.end())
      datum[node] = it->second;
    else
      app.push_back(node);
  }
  const int64_t m2 = (int64_t)app.size();
  for (int64_t k = 0; k < m2; ++k) s.push_back(Sample{P3{gx[app[k]], gy[app[k]], gz[app[k]]}, 0.0, -1, 0.0});
  bool planar = true;
  for (int64_t i = 1; i < n; ++i) planar = planar && s[i].loc[2] == s[0].loc[2];
  int64_t n_planar = 0;
  if (planar) {
    while (n_planar < m2 && s[n + n_planar].loc[2] == s[0].loc[2]) ++n_planar;
    n_planar = std::min(n_planar + 1, m2);
  }
  SearchTree tree(s, p, &vg);
  const Sectors flat = make_sectors(metric(p, &vg), true, p), solid = make_sectors(metric(p, &vg), false, p);
  std::vector<std::vector<int64_t>> lists(m2);
  std::vector<std::vector<double>> weights(m2);
  std::vector<double> var(m2, 0.0);
  std::vector<int> st(m2, BK_OK);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t k = 0; k < m2; ++k) {
    const P3 target = s[n + k].loc;
    std::vector<int64_t> idx;
    int r = tree.neighbors_around(target, -1, idx, n + k, k < n_planar ? &flat : &solid);
    if (r != BK_OK || idx.empty()) {
      st[k] = BK_SEARCH_FAILED;
      continue;
    }
    std::vector<Sample> sel;
    for (int64_t i : idx) sel.push_back(s[i]);
    Estimate e;
    r = krige(SIMPLE, 0.0, target, sel, vg, e);
    if (r != BK_OK) {
      st[k] = r;
      continue;
    }
    lists[k] = idx;
    weights[k] = e.weights;
    var[k] = e.variance;
  }
  for (int64_t k = 0; k < m2; ++k) {
    double sc;
    if (st[k] == BK_SEARCH_FAILED) {
      sc = z[k];
      source[app[k]] = -2;
    } else if (st[k] != BK_OK) {
      sc = NAN;
      source[app[k]] = -3;
    } else {
      double mean = 0.0;
      for (size_t j = 0; j < lists[k].size(); ++j) mean += weights[k][j] * s[lists[k][j]].value;
      sc = mean + std::sqrt(std::max(var[k], 0.0)) * z[k];
      source[app[k]] = -1;
    }
    s[n + k].value = sc;
    score[app[k]] = sc;
  }
  for (int64_t i = 0; i < m; ++i)
    if (datum[i] >= 0) {
      score[i] = s[datum[i]].value;
      source[i] = (int32_t)datum[i];
    }
  return 0;
}

// prefix_quality (kriging_algebra.rs:338-397) of the samples idx[0 .. n), in
// that order, for target t: point support, or the block `blk` centred on t.
// slope[k], efficiency[k] for the first k + 1 samples, NaN where None.
int or_prefix_quality(const or_samples* sm, const bk_variogram* v, int kind, const double* t, const bk_block* blk,
                      int n, const int64_t* idx, double* slope, double* efficiency) {
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
  std::vector<Sample> sel;
  for (int k = 0; k < n; ++k) sel.push_back(s[idx[k]]);
  P3 q{t[0], t[1], t[2]};
  std::vector<double> rhs(n);
  double support = vg.total_sill();
  if (blk) {
    auto pts = block_points(q, *blk);
    rhs = block_rhs(pts, sel, vg);
    support = vg.block_mean_cov(pts);
  } else {
    for (int k = 0; k < n; ++k) rhs[k] = vg.cov_points(sel[k].loc, q);
  }
  auto qs = prefix_quality(sel, vg, rhs, kind != SIMPLE, support);
  for (int k = 0; k < n; ++k) {
    slope[k] = qs[k].some ? qs[k].slope : NAN;
    efficiency[k] = qs[k].some ? qs[k].efficiency : NAN;
  }
  return 0;
}

// calibrated_count (kriging_algebra.rs:399-415) over qualities given as
// slope / efficiency arrays (NaN: None). Returns the count; *met whether it
// meets the target.
int or_calibrated_count(int calibration, double target, int64_t min_samples, int n, const double* slope,
                        const double* efficiency, int32_t* met) {
  std::vector<Quality> q(n);
  for (int k = 0; k < n; ++k) {
    q[k].some = !(std::isnan(slope[k]) && std::isnan(efficiency[k]));
    q[k].slope = slope[k];
    q[k].efficiency = efficiency[k];
  }
  auto r = calibrated_count(calibration, target, min_samples, q);
  if (met) *met = r.second ? 1 : 0;
  return r.first;
}

// SearchTree::balanced (search.rs:909-936) of the selection chosen[0 .. n)
// around target t.
int or_balanced(const or_samples* sm, const bk_search* sp, const bk_variogram* v, const double* t, int n,
                const int64_t* chosen, int64_t* out) {
  Search p = make_search(sp);
  if (!check(p)) return BK_E_INVALID;
  auto s = samples_of(sm);
  Variogram vg;
  if (v) vg = make_variogram(v);
  SearchTree tree(s, p, v ? &vg : nullptr);
  auto b = tree.balanced(P3{t[0], t[1], t[2]}, std::vector<int64_t>(chosen, chosen + n));
  std::copy(b.begin(), b.end(), out);
  return 0;
}

// estimate_many with krige (batch.rs:15-38).
int or_krige(const or_samples* sm, const bk_variogram* v, const bk_search* sp, int kind, double param, int64_t m,
             const double* tx, const double* ty, const double* tz, double* value, double* variance,
             double* lagrange, int32_t* n_used, int32_t* status, int nthreads) {
  Search p = make_search(sp);
  if (!check(p)) return BK_E_INVALID;
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
  if (!calibration_fits(p, kind)) return BK_E_INVALID;
  SearchTree tree(s, p, &vg);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t t = 0; t < m; ++t) {
    P3 q{tx[t], ty[t], tz[t]};
    std::vector<int64_t> chosen;
    Estimate e;
    int st = tree.neighbors_around(q, -1, chosen);
    if (st == BK_OK) {
      std::vector<Sample> sel;
      for (auto i : tree.balanced(q, chosen)) sel.push_back(s[i]);
      sel.resize(calibrated_n(p, kind, q, nullptr, sel, vg));
      st = krige(kind, param, q, sel, vg, e);
    }
    write_estimate(st, e, t, value, variance, lagrange, nullptr, n_used, status);
  }
  return 0;
}

// estimate_many with block_krige_points (block.rs:122-198, py/estimation.rs:767-784).
int or_block_krige(const or_samples* sm, const bk_variogram* v, const bk_search* sp, int kind, double param,
                   const bk_block* blk, int64_t m, const double* tx, const double* ty, const double* tz,
                   double* value, double* variance, double* lagrange, double* support, int32_t* n_used,
                   int32_t* status, int nthreads) {
  Search p = make_search(sp);
  if (!check(p)) return BK_E_INVALID;
  if (blk->nx <= 0 || blk->ny <= 0 || blk->nz <= 0) return BK_E_INVALID;
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
  if (!calibration_fits(p, kind)) return BK_E_INVALID;
  SearchTree tree(s, p, &vg);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t t = 0; t < m; ++t) {
    P3 q{tx[t], ty[t], tz[t]};
    std::vector<int64_t> chosen;
    Estimate e;
    int st = tree.neighbors_around(q, -1, chosen);
    if (st == BK_OK) {
      std::vector<Sample> sel;
      for (auto i : tree.balanced(q, chosen)) sel.push_back(s[i]);
      const std::vector<P3> pts = block_points(q, *blk);
      sel.resize(calibrated_n(p, kind, q, &pts, sel, vg));
      st = block_krige_points(kind, param, pts, sel, vg, e);
    }
    write_estimate(st, e, t, value, variance, lagrange, support, n_used, status);
  }
  return 0;
}

// leave_one_out_at (batch.rs:165-190); which == NULL means every sample.
int or_loo(const or_samples* sm, const bk_variogram* v, const bk_search* sp, int kind, double param, int64_t m,
           const int64_t* which, double* value, double* variance, int32_t* status, int nthreads) {
  Search p = make_search(sp);
  if (!check(p)) return BK_E_INVALID;
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
  if (!calibration_fits(p, kind)) return BK_E_INVALID;
  SearchTree tree(s, p, &vg);
  if (!which) m = (int64_t)s.size();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t k = 0; k < m; ++k) {
    int64_t i = which ? which[k] : k;
    const P3& q = s[i].loc;
    std::vector<int64_t> chosen;
    Estimate e;
    int st = tree.neighbors_around(q, i, chosen);
    if (st == BK_OK && chosen.empty()) st = BK_SEARCH_FAILED;
    if (st == BK_OK) {
      std::vector<Sample> sel;
      for (auto j : tree.balanced(q, chosen)) sel.push_back(s[j]);
      sel.resize(calibrated_n(p, kind, q, nullptr, sel, vg));
      st = krige(kind, param, q, sel, vg, e);
    }
    write_estimate(st, e, k, value, variance, nullptr, nullptr, nullptr, status);
  }
  return 0;
}

// Kriging with given neighbour lists (idx[t * stride + k], k < count[t]) at
// given targets: the fixed-neighbourhood systems of the leave-one-out loss.
int or_krige_lists(const or_samples* sm, const bk_variogram* v, int kind, double param, int64_t m,
                   const double* tx, const double* ty, const double* tz, const int32_t* idx, const int32_t* count,
                   int64_t stride, double* value, double* variance, double* lagrange, int32_t* status,
                   int nthreads) {
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t t = 0; t < m; ++t) {
    P3 q{tx[t], ty[t], tz[t]};
    std::vector<Sample> sel;
    for (int32_t k = 0; k < count[t]; ++k) sel.push_back(s[idx[t * stride + k]]);
    Estimate e;
    int st = krige(kind, param, q, sel, vg, e);
    write_estimate(st, e, t, value, variance, lagrange, nullptr, nullptr, status);
  }
  return 0;
}

// Block kriging with given neighbour lists on the discretisation of `blk`
// centred on each target: solver 0 = block_krige_points (block.rs:122-198, the
// reference's LU), 1 = block_krige_chol (the engine's FP64 order; lanes[t] is
// the lane group of the target's system) with the LU taking what it returns
// -1 for.
int or_block_krige_lists(const or_samples* sm, const bk_variogram* v, int kind, double param, const bk_block* blk,
                         int64_t m, const double* tx, const double* ty, const double* tz, const int32_t* idx,
                         const int32_t* count, int64_t stride, int solver, const int32_t* lanes, double* value,
                         double* variance, double* lagrange, double* support, int32_t* status, int nthreads) {
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t t = 0; t < m; ++t) {
    P3 q{tx[t], ty[t], tz[t]};
    std::vector<Sample> sel;
    for (int32_t k = 0; k < count[t]; ++k) sel.push_back(s[idx[t * stride + k]]);
    const std::vector<P3> pts = block_points(q, *blk);
    Estimate e;
    int st = solver == 1 ? block_krige_chol(kind, param, pts, sel, vg, lanes[t], e) : -1;
    if (st == -1) st = block_krige_points(kind, param, pts, sel, vg, e);
    write_estimate(st, e, t, value, variance, lagrange, support, nullptr, status);
  }
  return 0;
}

// The covariance matrix of the samples idx[0 .. n) (cov_points, error
// variances on the diagonal; block.rs:140-158), row-major into out (n * n).
void or_system_matrix(const or_samples* sm, const bk_variogram* v, int n, const int32_t* idx, double* out) {
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      out[i * n + j] = vg.cov_points(s[idx[i]].loc, s[idx[j]].loc) + (i == j ? s[idx[i]].err : 0.0);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Multiple indicator kriging exports.
extern "C" {

// MultipleIndicator::predict (indicator.rs:439-487) for one search pass.
int or_mik_predict(const or_samples* sm, const bk_mik* p, const double* thr, const bk_variogram* vgs,
                   const double* weights, const bk_search* sp, const bk_block* blk, int64_t m, const double* tx,
                   const double* ty, const double* tz, int nc, const double* cutoffs, int nq,
                   const double* quantiles, const bk_mik_out* o, int nthreads) {
  if (!mik_valid(p, thr)) return BK_E_INVALID;
  for (int i = 0; i < nc; ++i)
    if (!std::isfinite(cutoffs[i])) return BK_E_INVALID;
  for (int i = 0; i < nq; ++i)
    if (!(quantiles[i] >= 0.0 && quantiles[i] <= 1.0)) return BK_E_INVALID;
  Search sr = make_search(sp);
  if (!check(sr) || sr.calibration) return BK_E_INVALID;
  auto s = samples_of(sm);
  Mik mk = make_mik(p, thr, vgs);
  std::vector<double> values(s.size());
  for (size_t i = 0; i < s.size(); ++i) values[i] = s[i].value;
  Global g;
  int gs = global_new(values, weights, mk.thr, p->has_tails ? p->tails : nullptr, g);
  if (gs == BK_E_INVALID) return BK_E_INVALID;
  if (gs != BK_OK) return BK_E_STATE;
  const Variogram& svg = mk.vgs[mk.vgs.size() / 2];
  SearchTree tree(s, sr, &svg);
  std::vector<double> cut(cutoffs, cutoffs + nc), qs(quantiles, quantiles + nq);
  const size_t k = mk.thr.size();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t t = 0; t < m; ++t) {
    P3 q{tx[t], ty[t], tz[t]};
    std::vector<int64_t> chosen;
    int st = tree.neighbors_around(q, -1, chosen);
    std::vector<double> raw;
    std::vector<Sample> sel;
    if (st == BK_OK) {
      for (auto i : chosen) sel.push_back(s[i]);
      if (blk) {
        std::vector<P3> pts = block_points(q, *blk);
        st = krige_indicators(mk, q, sel, &pts, g.proportions, raw);
      } else {
        st = krige_indicators(mk, q, sel, nullptr, g.proportions, raw);
      }
    }
    if (st == BK_OK) {
      Conditional c = mik_conditional(mk, g, raw, cut, qs);
      write_conditional(&c, st, (int)sel.size(), t, m, o, &raw, k, nc, nq);
    } else {
      write_conditional(nullptr, st, 0, t, m, o, nullptr, k, nc, nq);
    }
  }
  return 0;
}

// MultipleIndicator::cross_validate, leave-one-out (indicator.rs:489-529) over
// the samples at which[0 .. m) (NULL: all).
int or_mik_cross_validate(const or_samples* sm, const bk_mik* p, const double* thr, const bk_variogram* vgs,
                          const double* weights, const bk_search* sp, int64_t m, const int64_t* which, double* pit,
                          const bk_mik_out* o, int nthreads) {
  if (!mik_valid(p, thr)) return BK_E_INVALID;
  Search sr = make_search(sp);
  if (!check(sr) || sr.calibration) return BK_E_INVALID;
  auto s = samples_of(sm);
  Mik mk = make_mik(p, thr, vgs);
  std::vector<double> values(s.size());
  for (size_t i = 0; i < s.size(); ++i) values[i] = s[i].value;
  Global g;
  int gs = global_new(values, weights, mk.thr, p->has_tails ? p->tails : nullptr, g);
  if (gs == BK_E_INVALID) return BK_E_INVALID;
  if (gs != BK_OK) return BK_E_STATE;
  const Variogram& svg = mk.vgs[mk.vgs.size() / 2];
  SearchTree tree(s, sr, &svg);
  if (!which) m = (int64_t)s.size();
  const size_t k = mk.thr.size();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t t = 0; t < m; ++t) {
    int64_t i = which ? which[t] : t;
    const P3& q = s[i].loc;
    std::vector<int64_t> chosen;
    int st = tree.neighbors_around(q, i, chosen);
    if (st == BK_OK && chosen.empty()) st = BK_SEARCH_FAILED;
    std::vector<double> raw;
    std::vector<Sample> sel;
    if (st == BK_OK) {
      for (auto j : chosen) sel.push_back(s[j]);
      st = krige_indicators(mk, q, sel, nullptr, g.proportions, raw);
    }
    if (st == BK_OK) {
      Conditional c = mik_conditional(mk, g, raw, {s[i].value}, {});
      if (pit) pit[t] = 1.0 - c.probability_above[0];
      write_conditional(&c, st, (int)sel.size(), t, m, o, &raw, k, 0, 0);
    } else {
      if (pit) pit[t] = NAN;
      write_conditional(nullptr, st, 0, t, m, o, nullptr, k, 0, 0);
    }
  }
  return 0;
}

// Global::new (indicator.rs:152-200) of n values: sorted values (n),
// sums (3 x (n + 1)), class bounds (2 (K + 1)), lower / upper, proportions (K).
int or_mik_global(int64_t n, const double* values, const double* weights, int k, const double* thr,
                  const double* tails, double* sorted, double* sums, int64_t* classes, double* bounds,
                  double* proportions) {
  Global g;
  int st = global_new(std::vector<double>(values, values + n), weights, std::vector<double>(thr, thr + k), tails, g);
  if (st != BK_OK) return st;
  std::copy(g.values.begin(), g.values.end(), sorted);
  for (int j = 0; j < 3; ++j) std::copy(g.sums[j].begin(), g.sums[j].end(), sums + j * (n + 1));
  for (size_t c = 0; c < g.classes.size(); ++c) {
    classes[2 * c] = (int64_t)g.classes[c].first;
    classes[2 * c + 1] = (int64_t)g.classes[c].second;
  }
  bounds[0] = g.lower;
  bounds[1] = g.upper;
  std::copy(g.proportions.begin(), g.proportions.end(), proportions);
  return 0;
}

// Shape::above (j = 0, 1, 2) or, with j < 0, Shape::quantile at u = c, of the
// class [s, e) on [a, b] of the global distribution of the n values (weights
// may be NULL). shape: 0 data, 1 power, 2 hyperbolic (exponent w).
double or_mik_shape(int64_t n, const double* values, const double* weights, int k, const double* thr,
                    const double* tails, int shape, double w, int64_t s, int64_t e, double a, double b, double c,
                    int j) {
  Global g;
  if (global_new(std::vector<double>(values, values + n), weights, std::vector<double>(thr, thr + k), tails, g) !=
      BK_OK)
    return NAN;
  Shape sh;
  sh.kind = shape;
  sh.s = (size_t)s;
  sh.e = (size_t)e;
  sh.w = w;
  return j < 0 ? shape_quantile(sh, g, a, b, c) : shape_above(sh, g, a, b, c, j);
}

// MultipleIndicator::conditional (indicator.rs:380-437) of m given kriged
// vectors raw[k * m + t] against the global distribution of the n values.
int or_mik_conditional(int64_t n, const double* values, const double* weights, const bk_mik* p, const double* thr,
                       int64_t m, const double* raw, int nc, const double* cutoffs, int nq, const double* quantiles,
                       const bk_mik_out* o) {
  if (!mik_valid(p, thr)) return BK_E_INVALID;
  Mik mk;
  mk.thr.assign(thr, thr + p->n_thresholds);
  mk.simple = p->simple != 0;
  mk.interpolation = p->interpolation;
  mk.upper_tail = p->upper_tail;
  mk.omega = p->omega;
  Global g;
  if (global_new(std::vector<double>(values, values + n), weights, mk.thr, p->has_tails ? p->tails : nullptr, g) !=
      BK_OK)
    return BK_E_INVALID;
  const size_t k = mk.thr.size();
  std::vector<double> cut(cutoffs, cutoffs + nc), qs(quantiles, quantiles + nq);
  for (int64_t t = 0; t < m; ++t) {
    std::vector<double> r(k);
    for (size_t i = 0; i < k; ++i) r[i] = raw[i * m + t];
    Conditional c = mik_conditional(mk, g, r, cut, qs);
    write_conditional(&c, BK_OK, 0, t, m, o, &r, k, nc, nq);
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Diagnostics exports.
extern "C" {

// neighborhood_stats (neighborhood.rs:66-136) of the samples idx[0 .. n) at
// target t with k nearest and radius; out: n_within, n_holes, nearest,
// mean_dist_knn, max_dist_knn, value_mean, value_var, value_min, value_max,
// value_idw.
void or_neighborhood_stats(const or_samples* sm, const double* t, int n, const int64_t* idx, int64_t k,
                           double radius, double* out) {
  auto s = samples_of(sm);
  std::vector<Sample> sel;
  for (int i = 0; i < n; ++i) sel.push_back(s[idx[i]]);
  NeighborhoodStats st = neighborhood_stats(P3{t[0], t[1], t[2]}, sel, (size_t)k, radius);
  const double v[10] = {(double)st.n_within, (double)st.n_holes, st.nearest_dist, st.mean_dist_knn,
                        st.max_dist_knn, st.value_mean, st.value_var, st.value_min, st.value_max, st.value_idw};
  std::copy(v, v + 10, out);
}

// The per-target diagnostics of _Estimator.predict(diagnostics=True)
// (py/estimation.rs:895-904,1034-1105) for one pass of point (blk NULL) or
// block kriging: the estimate, the statistics of the samples used and, for
// calibrated searches, whether they meet the target (-1 otherwise).
int or_krige_diag(const or_samples* sm, const bk_variogram* v, const bk_search* sp, int kind, double param,
                  const bk_block* blk, int64_t m, const double* tx, const double* ty, const double* tz,
                  const bk_diag_out* o, int nthreads) {
  Search p = make_search(sp);
  if (!check(p)) return BK_E_INVALID;
  auto s = samples_of(sm);
  Variogram vg = make_variogram(v);
  if (!calibration_fits(p, kind)) return BK_E_INVALID;
  SearchTree tree(s, p, &vg);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int64_t t = 0; t < m; ++t) {
    P3 q{tx[t], ty[t], tz[t]};
    std::vector<int64_t> chosen;
    Estimate e;
    std::vector<Sample> sel;
    int met = -1;
    int st = tree.neighbors_around(q, -1, chosen);
    if (st == BK_OK) {
      for (auto i : tree.balanced(q, chosen)) sel.push_back(s[i]);
      std::vector<P3> pts;
      if (blk) pts = block_points(q, *blk);
      auto nm = calibrated_nm(p, kind, q, blk ? &pts : nullptr, sel, vg);
      sel.resize(nm.first);
      if (p.calibration) met = nm.second ? 1 : 0;
      st = blk ? block_krige_points(kind, param, pts, sel, vg, e) : krige(kind, param, q, sel, vg, e);
    }
    const bool ok = st == BK_OK;
    NeighborhoodStats ns;
    if (ok) ns = neighborhood_stats(q, sel, sel.size(), INFINITY);
    if (o->value) o->value[t] = ok ? e.value : NAN;
    if (o->variance) o->variance[t] = ok ? e.variance : NAN;
    if (o->lagrange) o->lagrange[t] = ok ? e.lagrange : NAN;
    if (o->support_variance) o->support_variance[t] = ok ? e.support : NAN;
    if (o->n_used) o->n_used[t] = ok ? e.n_used : 0;
    if (o->status) o->status[t] = st;
    if (o->negative_weight_sum) o->negative_weight_sum[t] = ok ? negative_weight_sum(e.weights) : NAN;
    if (o->n_holes) o->n_holes[t] = ok ? (int32_t)ns.n_holes : 0;
    if (o->mean_distance) o->mean_distance[t] = ok ? ns.mean_dist_knn : NAN;
    if (o->target_met) o->target_met[t] = ok ? met : -1;
  }
  return 0;
}

}  // extern "C"
