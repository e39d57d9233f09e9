// bk_solve.hip — batched kriging solves (FP64 reference order; FP32 + FP64
// iterative refinement; guarded systems with coincident samples; bordered LU).
//
// Compiled with -ffp-contract=off: the FP64 kernel reproduces the reference's
// unfused arithmetic (crates/estimation/src/kernel.rs:47-113). Refinement
// code opts into contraction locally.
#include <algorithm>
#include <atomic>

#include "bk_solve.h"

// Built as two objects: BK_PART 1 holds the kriging pass (persistent solve
// kernels, guarded systems, bucketing), the FP64 gradient and the persistent
// refined gradient of systems up to 32 samples, BK_PART 2 the other refined
// gradients and the reductions; the build compiles them with different
// backend options. Without BK_PART the file holds everything.
#ifndef BK_PART
#define BK_PART 0
#endif

namespace {

constexpr double PIVOT = 1e-12;  // kernel.rs:17
constexpr double EPS2 = 1e-6 * 1e-6;  // coincident.rs:11 (EPS * EPS)

__device__ __forceinline__ int tri(int i) { return i * (i + 1) / 2; }

// Pair q of the strict lower triangle in row order: (i, j), j < i.
__device__ __forceinline__ void pair_of(int q, int& i, int& j) {
  int r = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)q)) * 0.5f);
  while (r * (r - 1) / 2 > q) --r;
  while ((r + 1) * r / 2 <= q) ++r;
  i = r;
  j = q - r * (r - 1) / 2;
}

// Sum over the NMAX lanes of one group.
template <int NMAX>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = NMAX / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, BK_WAVE);
  return v;
}

template <int NMAX>
__device__ __forceinline__ float group_sumf(float v) {
#pragma unroll
  for (int o = NMAX / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, BK_WAVE);
  return v;
}

template <int NMAX>
__device__ __forceinline__ double group_max(double v) {
#pragma unroll
  for (int o = NMAX / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, BK_WAVE));
  return v;
}

// The value of lane k of this lane's group (k a compile-time constant after
// unrolling): one v_readlane per group, no LDS round trip.
template <int NMAX>
__device__ __forceinline__ float group_bcastf(float v, int k, int g) {
  constexpr int G = BK_WAVE / NMAX;
  float out = 0.0f;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    float t = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), gg * NMAX + k));
    out = (g == gg) ? t : out;
  }
  return out;
}

template <int NMAX>
__device__ __forceinline__ double group_bcast(double v, int k, int g) {
  constexpr int G = BK_WAVE / NMAX;
  double out = 0.0;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    double t = bcast(v, gg * NMAX + k);
    out = (g == gg) ? t : out;
  }
  return out;
}

template <int NMAX>
__device__ __forceinline__ bool group_any(bool p, int g) {
  uint64_t m = __ballot(p);
  uint64_t gm = NMAX == 64 ? ~0ull : (((1ull << NMAX) - 1ull) << (g * NMAX));
  return (m & gm) != 0;
}

// Doubles of one sample record in LDS: world x, y, z, then the projected
// coordinates on every frame, padded to an odd count so that lanes reading
// consecutive records (stride 2 * SJ dwords) spread over the banks.
__host__ __device__ constexpr int rec_doubles(int nf) { return (3 + 3 * nf) | 1; }

// LDS doubles per system.
__host__ __device__ constexpr int sys_doubles(int nmax, int nf) { return nmax * (nmax + 1) / 2 + 5 * nmax + nf * 3 * nmax + 8; }

// The estimate from weights w (lane r < n holds w, z, c): krige.rs:98-113,
// kernel.rs:98-122. Writes the target's outputs from the group's lane 0.
template <int NMAX>
__device__ __forceinline__ void finish_outputs(const SolveArgs& a, int t, int r, int n, bool ok, int st_fail,
                                               double w, double z, double c, double mu, double c0v, double sup) {
  bool in = ok && r < n;
  double f = 0.0;
  if (in) {
    if (a.kind == BK_SIMPLE)
      f = w * (z - a.param);
    else if (a.kind == BK_INDICATOR)
      f = w * (z <= a.param ? 1.0 : 0.0);
    else
      f = w * z;
  }
  double sum = group_sum<NMAX>(f);
  double swc = group_sum<NMAX>(in ? w * c : 0.0);
  if (a.weights && in) a.weights[(int64_t)t * a.stride + r] = w;
  if (r == 0) {
    if (ok) {
      double val = sum;
      if (a.kind == BK_SIMPLE) val = a.param + sum;
      if (a.kind == BK_INDICATOR) val = fmin(fmax(sum, 0.0), 1.0);
      a.value[t] = val;
      a.variance[t] = fmax(c0v - swc - mu, 0.0);
      if (a.lagrange) a.lagrange[t] = mu;
      if (a.support) a.support[t] = sup;
      if (a.n_used) a.n_used[t] = n;
      a.status[t] = BK_OK;
    } else {
      a.value[t] = __longlong_as_double(0x7ff8000000000000LL);
      a.variance[t] = __longlong_as_double(0x7ff8000000000000LL);
      if (a.lagrange) a.lagrange[t] = __longlong_as_double(0x7ff8000000000000LL);
      if (a.support) a.support[t] = __longlong_as_double(0x7ff8000000000000LL);
      if (a.n_used) a.n_used[t] = 0;
      a.status[t] = st_fail;
    }
  }
}

// Loads the system's samples into LDS (world coordinates, values, error
// variances, frame coordinates relative to the target: kernel.rs:36-42) and
// reports whether two coincide (coincident.rs:13-15, 27-35).
template <int NMAX>
__device__ __forceinline__ bool load_system(const SolveArgs& a, int t, int r, int n, int g, double* W, double* V,
                                            double* E, double* PF, double& z, double& e) {
  const int nf = a.v.nf;
  double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
  if (r < n) {
    int id = a.nidx[(int64_t)t * a.stride + r];
    double x = a.sx[id], y = a.sy[id], zz = a.sz[id];
    W[r] = x;
    W[NMAX + r] = y;
    W[2 * NMAX + r] = zz;
    z = a.sv[id];
    e = a.serr ? a.serr[id] : 0.0;
    V[r] = z;
    E[r] = e;
    double d0 = x - t0, d1 = y - t1, d2 = zz - t2;
    for (int f = 0; f < nf; ++f) {
      double p0 = d0, p1 = d1, p2 = d2;
      if (!a.v.ident[f]) mv3(a.v.F[f], d0, d1, d2, p0, p1, p2);
      PF[(f * 3 + 0) * NMAX + r] = p0;
      PF[(f * 3 + 1) * NMAX + r] = p1;
      PF[(f * 3 + 2) * NMAX + r] = p2;
    }
  }
  wave_sync();
  bool close = false;
  if (r < n) {
    double x = W[r], y = W[NMAX + r], zz = W[2 * NMAX + r];
    for (int j = 0; j < r && !close; ++j) {
      double u = x - W[j], v = y - W[NMAX + j], w = zz - W[2 * NMAX + j];
      close = (u * u + v * v) + w * w <= EPS2;
    }
  }
  return group_any<NMAX>(close, g);
}

// Covariance between samples i and j in the kernel's frame (kernel.rs:43-53),
// or between world points for block support (block.rs:152-158 cov_points).
__device__ __forceinline__ double pair_cov64(const SolveArgs& a, const double* W, const double* PF, int nmax, int i,
                                             int j) {
  if (a.block)
    return cov_points64(a.v, W[i], W[nmax + i], W[2 * nmax + i], W[j], W[nmax + j], W[2 * nmax + j]);
  double h[BK_MAX_FRAMES];
  for (int f = 0; f < a.v.nf; ++f) {
    const double* P = PF + f * 3 * nmax;
    h[f] = norm3(P[i] - P[j], P[nmax + i] - P[nmax + j], P[2 * nmax + i] - P[2 * nmax + j]);
  }
  return cov64(a.v, h);
}

// Right-hand side of sample r: C(s_r, target) (kernel.rs:55), or the mean
// point-to-block covariance (block.rs:100-118).
__device__ __forceinline__ double rhs64(const SolveArgs& a, int t, const double* W, const double* PF, int nmax,
                                        int r) {
  if (a.block) {
    double x = W[r], y = W[nmax + r], z = W[2 * nmax + r];
    double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
    double acc = -0.0;
    for (int p = 0; p < a.npts; ++p) {
      double qx = t0 + a.offs[3 * p], qy = t1 + a.offs[3 * p + 1], qz = t2 + a.offs[3 * p + 2];
      acc += a.npts == 1 ? cov_points64(a.v, x, y, z, qx, qy, qz) : block_cov_points64(a.v, x, y, z, qx, qy, qz);
    }
    return acc / (double)a.npts;
  }
  double h[BK_MAX_FRAMES];
  for (int f = 0; f < a.v.nf; ++f) {
    const double* P = PF + f * 3 * nmax;
    h[f] = norm3(P[r] - 0.0, P[nmax + r] - 0.0, P[2 * nmax + r] - 0.0);
  }
  return cov64(a.v, h);
}

// The same two covariances in the fast FP64 arithmetic of the refined form.
__device__ __forceinline__ double pair_cov_fast(const SolveArgs& a, const double* W, const double* PF, int nmax,
                                                int i, int j) {
  double h[BK_MAX_FRAMES];
  if (a.block) {
    lags_world_fast(a.v, W[j] - W[i], W[nmax + j] - W[nmax + i], W[2 * nmax + j] - W[2 * nmax + i], h);
    return cov_fast64(a.v, h);
  }
  for (int f = 0; f < a.v.nf; ++f) {
    const double* P = PF + f * 3 * nmax;
    h[f] = norm3_fast(P[i] - P[j], P[nmax + i] - P[nmax + j], P[2 * nmax + i] - P[2 * nmax + j]);
  }
  return cov_fast64(a.v, h);
}

__device__ __forceinline__ double rhs_fast(const SolveArgs& a, int t, const double* W, const double* PF, int nmax,
                                           int r) {
  double h[BK_MAX_FRAMES];
  if (a.block) {
    double x = W[r], y = W[nmax + r], z = W[2 * nmax + r];
    double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
    double acc = 0.0;
    for (int p = 0; p < a.npts; ++p) {
      lags_world_fast(a.v, (t0 + a.offs[3 * p]) - x, (t1 + a.offs[3 * p + 1]) - y, (t2 + a.offs[3 * p + 2]) - z, h);
      acc += (a.npts > 1 && !(h[a.v.nug_frame] > 0.0)) ? a.v.c0 - a.v.nugget : cov_fast64(a.v, h);
    }
    return acc / (double)a.npts;
  }
  for (int f = 0; f < a.v.nf; ++f) {
    const double* P = PF + f * 3 * nmax;
    h[f] = norm3_fast(P[r], P[nmax + r], P[2 * nmax + r]);
  }
  return cov_fast64(a.v, h);
}

// Mean covariance between world point (x, y, z) and the discretisation of the
// block centred on (t0, t1, t2), nugget left out between distinct points
// (block.rs:100-118), in the fast arithmetic.
__device__ __forceinline__ double rhs_fast_world(const SolveArgs& a, double t0, double t1, double t2, double x,
                                                 double y, double z) {
  double h[BK_MAX_FRAMES];
  double acc = 0.0;
  for (int p = 0; p < a.npts; ++p) {
    lags_world_fast(a.v, (t0 + a.offs[3 * p]) - x, (t1 + a.offs[3 * p + 1]) - y, (t2 + a.offs[3 * p + 2]) - z, h);
    acc += (a.npts > 1 && !(h[a.v.nug_frame] > 0.0)) ? a.v.c0 - a.v.nugget : cov_fast64(a.v, h);
  }
  return acc / (double)a.npts;
}

// ---------------------------------------------------------------------------
// FP64, the reference's operation order.
//
// LDS per system: the packed lower triangle with the diagonal (the factor in
// place), the sample records (world coordinates and frame projections, as
// the refined kernels) and a few scalars.
__host__ __device__ constexpr int sys64_doubles(int nmax, int nf) {
  return (nmax * (nmax + 1) / 2 + 1) / 2 * 2 + nmax * rec_doubles(nf) + 8;
}

// The reference-order front of the FP64 kernels: records, coincident test,
// right-hand side c of sample r, and the matrix (each pair once, balanced over
// the lanes as in pairs_cov_fast) into L with the diagonal c0 + e. Lags are
// norm3 of frame-coordinate differences and covariances cov64, exactly the
// reference's arithmetic (kernel.rs:36-56, composite.rs:69-85); block
// support takes the world-coordinate forms (block.rs:100-158). Returns the
// guard flag (coincident samples).
template <int NMAX>
__device__ __forceinline__ bool fp64_front(const SolveArgs& a, int t, int n, int r, int g, bool active, double* L,
                                           double* PJ, int SJ, double& z, double& e, double& c) {
  const int nf = a.v.nf;
  const double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
  double wx = 0.0, wy = 0.0, wz = 0.0;
  z = 0.0;
  e = 0.0;
  double* me = PJ + r * SJ;
  if (r < n) {
    const int id = a.nidx[(int64_t)t * a.stride + r];
    wx = a.sx[id];
    wy = a.sy[id];
    wz = a.sz[id];
    z = a.sv[id];
    e = a.serr ? a.serr[id] : 0.0;
  }
  {
    const double d0 = wx - t0, d1 = wy - t1, d2 = wz - t2;
    me[0] = wx;
    me[1] = wy;
    me[2] = wz;
    for (int f = 0; f < nf; ++f) {
      double p0 = d0, p1 = d1, p2 = d2;
      if (!a.v.ident[f]) mv3(a.v.F[f], d0, d1, d2, p0, p1, p2);
      me[3 + 3 * f] = p0;
      me[4 + 3 * f] = p1;
      me[5 + 3 * f] = p2;
    }
  }
  wave_sync();
  bool close = false;
#pragma unroll
  for (int j = 0; j < NMAX; ++j) {
    if (j < r && r < n) {
      const double* pj = PJ + j * SJ;
      const double u = wx - pj[0], vv = wy - pj[1], w = wz - pj[2];
      close |= (u * u + vv * vv) + w * w <= EPS2;
    }
    if (j % 4 == 3) asm volatile("" ::: "memory");
  }
  const bool guard = group_any<NMAX>(close, g);
  const bool go = active && !guard;
  c = 0.0;
  if (go && r < n) {
    const double c0 = a.v.c0;
    L[tri(r) + r] = c0 + e;
    if (a.block) {
      double acc = -0.0;
      for (int p = 0; p < a.npts; ++p) {
        const double qx = t0 + a.offs[3 * p], qy = t1 + a.offs[3 * p + 1], qz = t2 + a.offs[3 * p + 2];
        acc += a.npts == 1 ? cov_points64(a.v, wx, wy, wz, qx, qy, qz)
                           : block_cov_points64(a.v, wx, wy, wz, qx, qy, qz);
      }
      c = acc / (double)a.npts;
    } else {
      double h[BK_MAX_FRAMES];
      for (int f = 0; f < nf; ++f) h[f] = norm3(me[3 + 3 * f] - 0.0, me[4 + 3 * f] - 0.0, me[5 + 3 * f] - 0.0);
      c = cov64(a.v, h);
    }
    const int half = (n - 1) / 2;
    const bool even = (n & 1) == 0;
    const int dmax = (even && r < n / 2) ? n / 2 : half;
    for (int d = 1; d <= dmax; ++d) {
      int j = r + d;
      j = j >= n ? j - n : j;
      const double* pj = PJ + j * SJ;
      double cij;
      if (a.block) {
        cij = cov_points64(a.v, wx, wy, wz, pj[0], pj[1], pj[2]);
      } else {
        double h[BK_MAX_FRAMES];
        for (int f = 0; f < nf; ++f)
          h[f] = norm3(me[3 + 3 * f] - pj[3 + 3 * f], me[4 + 3 * f] - pj[4 + 3 * f], me[5 + 3 * f] - pj[5 + 3 * f]);
        cij = cov64(a.v, h);
      }
      L[j < r ? tri(r) + j : tri(j) + r] = cij;
    }
  }
  wave_sync();
  return guard;
}

// The thread index, opaque to the optimiser: solve bodies compute their lane
// addresses from it, so a persistent loop does not hoist them (dozens per
// system) out of the loop and hold them in VGPRs across it.
__device__ __forceinline__ int opaque_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// Systems of a launch: the device count when the list was filled on the device.
__device__ __forceinline__ int64_t list_len(const SolveArgs& a) {
  if (!a.n_list_dev) return a.n_list;
  const int64_t n = *a.n_list_dev;
  return n < a.n_list ? n : a.n_list;
}

template <int NMAX>
__device__ __forceinline__ void krige64_blk(const SolveArgs& a, int64_t nl, int unit) {
  constexpr int G = BK_WAVE / NMAX;
  extern __shared__ double dyn[];
  const int nf = a.v.nf;
  const int SJ = rec_doubles(nf);
  const int per = sys64_doubles(NMAX, nf);
  const int tid = opaque_tid();
  const int wave = tid / BK_WAVE, lane = tid % BK_WAVE;
  const int g = lane / NMAX, r = lane % NMAX;
  double* S = dyn + (wave * G + g) * per;
  double* L = S;
  double* PJ = L + (NMAX * (NMAX + 1) / 2 + 1) / 2 * 2;
  double* misc = PJ + NMAX * SJ;
  const int64_t sys = (int64_t)unit * G + g;
  const bool active = sys < nl;
  const int t = active ? a.list[sys] : 0;
  const int n = active ? a.ncount[t] : 0;
  double z, e, c;
  const bool guard = fp64_front<NMAX>(a, t, n, r, g, active, L, PJ, SJ, z, e, c);
  if (guard && r == 0 && active) {
    a.status[t] = BK_ST_GUARD;
    a.guard[atomicAdd(a.guard_count, 1)] = t;
  }
  const bool go = active && !guard;
  const double c0 = a.v.c0;
  double reg[NMAX];
#pragma unroll
  for (int k = 0; k < NMAX; ++k) reg[k] = (go && k <= r && r < n) ? L[tri(r) + k] : 0.0;
  // Left-looking Cholesky (kernel.rs:58-72): row r of the factor stays in
  // lane r; the pivot row is read from LDS by the whole group.
  bool sing = false;
#pragma unroll
  for (int j = 0; j < NMAX; ++j) {
    const bool act = go && j < n;
    double acc = -0.0;
#pragma unroll
    for (int k = 0; k < j; ++k) acc += reg[k] * L[tri(j) + k];
    double s = reg[j] - acc;
    double d = group_bcast<NMAX>(s, j, g);
    if (act && (isnan(d) || d <= PIVOT * c0)) sing = true;
    if (act && !sing) {
      // Every lane divides (the division costs the same issued for one lane
      // or all); the pivot lane keeps the root, the lanes below the quotient.
      d = sqrt(d);
      const double x = s / d;
      if (r >= j && r < n) {
        const double v = r == j ? d : x;
        reg[j] = v;
        L[tri(r) + j] = v;
      }
    }
    wave_sync();
  }
  (void)misc;
  // Forward solves L y = c and L y = 1 (kernel.rs:73-78), column by column;
  // each new y_k goes lane to lane by readlane.
  const bool ordinary = a.kind != BK_SIMPLE;
  double ya = c, yb = 1.0, acca = -0.0, accb = -0.0;
#pragma unroll
  for (int k = 0; k < NMAX; ++k) {
    const bool act = go && !sing && k < n;
    if (act && r == k) {
      ya = (ya - acca) / reg[k];
      yb = (yb - accb) / reg[k];
    }
    double ba = group_bcast<NMAX>(ya, k, g), bb = group_bcast<NMAX>(yb, k, g);
    if (act && r > k && r < n) {
      acca += reg[k] * ba;
      accb += reg[k] * bb;
    }
  }
  // Back solves L' x = y (kernel.rs:79-84), row k of the factor from LDS.
  double xa = ya, xb = yb;
#pragma unroll
  for (int k = NMAX - 1; k >= 0; --k) {
    const bool act = go && !sing && k < n;
    if (act && r == k) {
      xa = xa / reg[k];
      xb = xb / reg[k];
    }
    double ba = group_bcast<NMAX>(xa, k, g), bb = group_bcast<NMAX>(xb, k, g);
    if (act && r < k) {
      double lk = L[tri(k) + r];
      xa -= lk * ba;
      xb -= lk * bb;
    }
  }
  // Schur scalar and weights (kernel.rs:86-97).
  const bool in = go && !sing && r < n;
  double sa = group_sum<NMAX>(in ? xa : 0.0);
  double sb = group_sum<NMAX>(in ? xb : 0.0);
  double mu = ordinary ? (sa - 1.0) / sb : 0.0;
  double w = xa - (ordinary ? mu * xb : 0.0);
  if (go && sing && a.block) {
    // block_krige_points solves by LU (block.rs:167-170), which has no pivot
    // test: a block system the Cholesky rejects goes to the LU kernel.
    if (r == 0) {
      a.status[t] = BK_ST_GUARD;
      a.guard[atomicAdd(a.guard_count, 1)] = t;
    }
  } else if (active && !guard) {
    finish_outputs<NMAX>(a, t, r, n, !sing, BK_SINGULAR, w, z, c, mu, a.block ? a.cbb : c0, a.block ? a.cbb : c0);
  }
}

template <int NMAX>
__global__ void __launch_bounds__(256) k_krige64(const SolveArgs a) {
  const int64_t nl = list_len(a);
  UnitLoop q = unit_loop((nl + BK_WAVE / NMAX - 1) / (BK_WAVE / NMAX), a.work);
  while (q.unit < q.end) {
    // The next unit is taken before this one is solved, hiding the atomic.
    const int next = unit_next(q);
    krige64_blk<NMAX>(kernarg_first<SolveArgs>(), nl, q.unit);
    wave_sync();
    q.unit = next;
  }
}

// ---------------------------------------------------------------------------
// FP32 + FP64 refinement. The system is scaled by 1/c0 and inverted in FP32 by
// Gauss-Jordan elimination without pivoting (the matrix is SPD, so pivots are
// the Cholesky pivots d_j of kernel.rs:58-72 and stay positive): lane r holds
// row r of the working matrix in registers, the pivot row goes to the group
// through LDS as float4. Every solve is then one matrix-vector product X v, so
// the initial solve and each refinement sweep are fully parallel. Residuals
// b - C x are taken in FP64 against the FP64 matrix (row r in registers for
// NMAX <= 32), the iterates are FP64, and so are the Schur scalar and the
// weights. A pivot below 1e-5 of the scaled diagonal (FP32 cannot resolve the
// reference's 1e-12 test) or a system that does not converge goes to the FP64
// kernel, which applies the reference's pivot test exactly.
template <int NMAX>
__device__ __forceinline__ bool group_all(bool p, int g) {
  uint64_t m = __ballot(p);
  uint64_t gm = NMAX == 64 ? ~0ull : (((1ull << NMAX) - 1ull) << (g * NMAX));
  return (m & gm) == gm;
}

template <int NMAX>
__device__ __forceinline__ float group_maxf(float v) {
#pragma unroll
  for (int o = NMAX / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, BK_WAVE));
  return v;
}

// Convergence of the refinement. Sweep k adds a correction d_k of relative
// size delta_k = max|d_k| / max|x_0|; the corrections still to come add up to
// about delta_k * rho, rho the contraction per sweep, ~ ||I - X C|| for the
// FP32 inverse X of C / c0. The first sweep estimates rho as the larger of
// delta_1 (the first correction is (I - X C) applied to the solution) and
// n u / d_min (u = 2^-24, d_min the smallest pivot of the elimination, which
// bounds the smallest eigenvalue of C / c0 from above), and converges when
// delta_1 * rho <= tol: a well-conditioned system (C2: d_min >= 0.1 with its
// 0.1 nugget, delta_1 ~ 1e-5) stops after one sweep with a predicted
// remainder below 1e-9, an ill-conditioned one (d_min ~ 1e-4, rho ~ 0.1)
// sweeps on. Later sweeps converge when delta_k <= tol, as a contraction
// measured from two corrections of an ill-conditioned system can understate
// the slow modes.
constexpr float kUnitRoundoff32 = 5.9604645e-8f;

// A lane's correction relative to max|x_0| (0 where x_0 = 0); delta_k is
// the group's largest over the right-hand sides.
__device__ __forceinline__ float rel_corr(float d, float m) {
  return m > 0.0f ? fabsf(d) * __builtin_amdgcn_rcpf(m) : 0.0f;
}

// The first sweep's test for a system of at most nmax samples (rho_p = nmax
// u / d_min; false for a NaN or a non-positive pivot).
__device__ __forceinline__ bool first_sweep_converged(float delta, float nmax, float dmin, float tol) {
  const float rho_p = nmax * kUnitRoundoff32 / dmin;
  return dmin > 0.0f && delta * fmaxf(delta, rho_p) <= tol;
}

// LDS layout of one system of the refined form: per sample j a record of SJ
// doubles (rec_doubles), iterates (NMAX double2), pivot rows (2 x NMAX
// floats), product vectors (NMAX float2), the FP64 matrix as a packed strict
// lower triangle, and for NMAX = 64 the FP64 rows of the wave's lanes
// ([k][lane], lane-private).
__host__ __device__ constexpr int sys32_doubles(int nmax, int nf) {
  return nmax * rec_doubles(nf) + 2 * nmax + nmax + nmax + (nmax * (nmax + 1) / 2 + 1) / 2 * 2 +
         (nmax > 32 ? nmax * nmax : 0) + 8;
}

// acc[j] += sill * shape(|P_r - P_j| / range) over the group's samples j for
// one structure whose frame starts at `off` in each record; sets bit j of
// `apart` where the lag is positive (the nugget's h > 0 test, done on the
// frame of structure 0 as composite.rs:69-85 does on the shared one).
template <int NMAX, class Shape>
__device__ __forceinline__ void row_pass(const double* PJ, int SJ, int off, const double* me, double sill, double ir,
                                         bool first, uint64_t& apart, double* acc, Shape shape) {
#pragma clang fp contract(fast)
  const double q0 = me[off], q1 = me[off + 1], q2 = me[off + 2];
#pragma unroll
  for (int j = 0; j < NMAX; ++j) {
    const double* pj = PJ + j * SJ + off;
    const double d0 = q0 - pj[0], d1 = q1 - pj[1], d2 = q2 - pj[2];
    const double h = sqrt_fast(d0 * d0 + d1 * d1 + d2 * d2);
    if (first && h > 0.0) apart |= 1ull << j;
    acc[j] += sill * shape(h * ir);
    // Keeps the compiler from hoisting every record's load to the top (which
    // would hold 3 NMAX doubles in registers).
    if (j % 4 == 3) asm volatile("" ::: "memory");
  }
}

// Row r of the covariance matrix in the fast FP64 arithmetic, without the
// diagonal: out[j] = C(s_r, s_j) for j < NMAX (values beyond n unused).
template <int NMAX>
__device__ __forceinline__ void row_cov_fast(const DevVariogram& v, const double* PJ, int SJ, int r, double* out) {
#pragma clang fp contract(fast)
  const double* me = PJ + r * SJ;
#pragma unroll
  for (int j = 0; j < NMAX; ++j) out[j] = 0.0;
  uint64_t apart = 0;
  for (int s = 0; s < v.S; ++s) {
    const int off = 3 + 3 * v.frame[s];
    const double sill = v.sill[s], ir = v.inv_range[s];
    const bool first = s == 0;
    switch (v.model[s]) {
      case BK_SPHERICAL:
        row_pass<NMAX>(PJ, SJ, off, me, sill, ir, first, apart, out,
                       [](double t) { return t >= 1.0 ? 1.0 : t * (1.5 - 0.5 * t * t); });
        break;
      case BK_EXPONENTIAL:
        row_pass<NMAX>(PJ, SJ, off, me, sill, ir, first, apart, out,
                       [](double t) { return 1.0 - exp_fast(-3.0 * t); });
        break;
      case BK_GAUSSIAN:
        row_pass<NMAX>(PJ, SJ, off, me, sill, ir, first, apart, out,
                       [](double t) { return 1.0 - exp_fast(-3.0 * t * t); });
        break;
      case BK_CUBIC:
        row_pass<NMAX>(PJ, SJ, off, me, sill, ir, first, apart, out, [](double t) {
          const double t2 = t * t;
          return t >= 1.0 ? 1.0 : t2 * (7.0 + t * (-8.75 + t2 * (3.5 - 0.75 * t2)));
        });
        break;
      case BK_PENTASPHERICAL:
        row_pass<NMAX>(PJ, SJ, off, me, sill, ir, first, apart, out, [](double t) {
          const double t2 = t * t;
          return t >= 1.0 ? 1.0 : t * (1.875 + t2 * (-1.25 + 0.375 * t2));
        });
        break;
      default:
        // Other models never reach the refined kernel (the host sends them to
        // the FP64 one); NaN would send the system there as well.
        row_pass<NMAX>(PJ, SJ, off, me, sill, ir, first, apart, out, [](double) { return __builtin_nan(""); });
    }
  }
  const double c0 = v.c0, nug = v.nugget;
#pragma unroll
  for (int j = 0; j < NMAX; ++j) out[j] = ((apart >> j) & 1) ? c0 - (nug + out[j]) : c0;
}

// Each unordered pair of the group's n samples once: lane r takes the pairs
// (r, (r + d) mod n) for d = 1 .. NMAX/2 (d = n/2 only from r < n/2 when n is
// even, d <= (n - 1)/2 otherwise), so every lane evaluates about n/2 pairs.
// The offsets d0 .. d0 + CH - 1 of one chunk: acc[d - d0] receives
// C(s_r, s_j) for valid d.
template <int CH, class Shape>
__device__ __forceinline__ void half_pass(const double* PJ, int SJ, int off, const double* me, int r, int n, int d0,
                                          double sill, double ir, bool first, uint64_t& apart, uint64_t& near,
                                          double near2, double* acc, Shape shape) {
#pragma clang fp contract(fast)
  const double q0 = me[off], q1 = me[off + 1], q2 = me[off + 2];
#pragma unroll
  for (int dd = 0; dd < CH; ++dd) {
    const int d = d0 + dd;
    // For small n, r + d can pass n twice; such d are not valid pairs and
    // read sample 0 instead of beyond the records.
    int j = r + d;
    j = j >= n ? j - n : j;
    j = j < n ? j : 0;
    const double* pj = PJ + j * SJ + off;
    const double d0v = q0 - pj[0], d1 = q1 - pj[1], d2 = q2 - pj[2];
    const double l2 = d0v * d0v + d1 * d1 + d2 * d2;
    const double h = sqrt_fast(l2);
    if (first && h > 0.0) apart |= 1ull << d;
    if (first && l2 <= near2) near |= 1ull << d;
    acc[dd] += sill * shape(h * ir);
    if (dd % 4 == 3) asm volatile("" ::: "memory");
  }
}

// The strict lower triangle of the group's covariance matrix into T (packed,
// T[i (i + 1) / 2 + j], j < i), each pair evaluated once (see half_pass), in
// chunks of at most 16 offsets so that a lane holds at most 16 partial sums.
// Returns whether one of the lane's pairs is near (within v.near2 in the
// frame of structure 0): only then can two of the samples coincide.
template <int NMAX>
__device__ __forceinline__ bool pairs_cov_fast(const DevVariogram& v, const double* PJ, int SJ, int r, int n,
                                               double* T) {
#pragma clang fp contract(fast)
  if (r >= n) return false;
  constexpr int CH = NMAX / 2 < 16 ? NMAX / 2 : 16;
  const double* me = PJ + r * SJ;
  uint64_t apart = 0, near = 0, nearv = 0;
  const double near2 = v.near2;
  const double c0 = v.c0, nug = v.nugget;
  const int half = (n - 1) / 2;
  const bool even = (n & 1) == 0;
#pragma unroll
  for (int d0 = 1; d0 <= NMAX / 2; d0 += CH) {
    double acc[CH];
#pragma unroll
    for (int dd = 0; dd < CH; ++dd) acc[dd] = 0.0;
    for (int s = 0; s < v.S; ++s) {
      const int off = 3 + 3 * v.frame[s];
      const double sill = v.sill[s], ir = v.inv_range[s];
      const bool first = s == 0;
      switch (v.model[s]) {
        case BK_SPHERICAL:
          half_pass<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, apart, near, near2, acc,
                        [](double t) { return t >= 1.0 ? 1.0 : t * (1.5 - 0.5 * t * t); });
          break;
        case BK_EXPONENTIAL:
          half_pass<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, apart, near, near2, acc,
                        [](double t) { return 1.0 - exp_fast(-3.0 * t); });
          break;
        case BK_GAUSSIAN:
          half_pass<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, apart, near, near2, acc,
                        [](double t) { return 1.0 - exp_fast(-3.0 * t * t); });
          break;
        case BK_CUBIC:
          half_pass<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, apart, near, near2, acc, [](double t) {
            const double t2 = t * t;
            return t >= 1.0 ? 1.0 : t2 * (7.0 + t * (-8.75 + t2 * (3.5 - 0.75 * t2)));
          });
          break;
        case BK_PENTASPHERICAL:
          half_pass<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, apart, near, near2, acc, [](double t) {
            const double t2 = t * t;
            return t >= 1.0 ? 1.0 : t * (1.875 + t2 * (-1.25 + 0.375 * t2));
          });
          break;
        default:
          half_pass<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, apart, near, near2, acc,
                        [](double) { return __builtin_nan(""); });
      }
    }
#pragma unroll
    for (int dd = 0; dd < CH; ++dd) {
      const int d = d0 + dd;
      const bool valid = d <= half || (even && d == n / 2 && r < n / 2);
      if (valid) {
        int j = r + d;
        j = j >= n ? j - n : j;
        const double c = ((apart >> d) & 1) ? c0 - (nug + acc[dd]) : c0;
        T[j < r ? r * (r + 1) / 2 + j : j * (j + 1) / 2 + r] = c;
        nearv |= (near >> d) & 1;
      }
    }
  }
  return nearv != 0;
}

// half_pass for the systems that a near pair sends to the FP64 kernel: a pair
// that is not near has a positive lag in every frame, so the nugget needs no
// per-pair test, the square root no guard, and the near test one flag per lane
// (offsets d >= n, possible for n < NMAX / 2 + 1, repeat pairs and are left out).
template <int CH, class Shape>
__device__ __forceinline__ void half_pass_near(const double* PJ, int SJ, int off, const double* me, int r, int n,
                                               int d0, double sill, double ir, bool first, bool& near, double near2,
                                               double* acc, Shape shape) {
#pragma clang fp contract(fast)
  const double q0 = me[off], q1 = me[off + 1], q2 = me[off + 2];
#pragma unroll
  for (int dd = 0; dd < CH; ++dd) {
    const int d = d0 + dd;
    int j = r + d;
    j = j >= n ? j - n : j;
    j = j < n ? j : 0;
    const double* pj = PJ + j * SJ + off;
    const double d0v = q0 - pj[0], d1 = q1 - pj[1], d2 = q2 - pj[2];
    const double l2 = d0v * d0v + d1 * d1 + d2 * d2;
    if (first) near |= l2 <= near2 && d < n;
    acc[dd] += sill * shape(sqrt_pos(l2) * ir);
    if (dd % 4 == 3) asm volatile("" ::: "memory");
  }
}

// pairs_cov_fast for the near-pretest kernels (see half_pass_near): the
// strict lower triangle into T, and whether one of the lane's pairs is near.
// A near system's entries are not used.
template <int NMAX>
__device__ __forceinline__ bool pairs_cov_near(const DevVariogram& v, const double* PJ, int SJ, int r, int n,
                                               double* T) {
#pragma clang fp contract(fast)
  if (r >= n) return false;
  constexpr int CH = NMAX / 2 < 16 ? NMAX / 2 : 16;
  const double* me = PJ + r * SJ;
  bool near = false;
  const double near2 = v.near2;
  const double c0 = v.c0, nug = v.nugget;
  const int half = (n - 1) / 2;
  const bool even = (n & 1) == 0;
  const ExpConstants k = exp_constants();
#pragma unroll
  for (int d0 = 1; d0 <= NMAX / 2; d0 += CH) {
    double acc[CH];
#pragma unroll
    for (int dd = 0; dd < CH; ++dd) acc[dd] = 0.0;
    for (int s = 0; s < v.S; ++s) {
      const int off = 3 + 3 * v.frame[s];
      const double sill = v.sill[s], ir = v.inv_range[s];
      const bool first = s == 0;
      switch (v.model[s]) {
        case BK_SPHERICAL:
          half_pass_near<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, near, near2, acc,
                             [](double t) { return t >= 1.0 ? 1.0 : t * (1.5 - 0.5 * t * t); });
          break;
        case BK_EXPONENTIAL:
          half_pass_near<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, near, near2, acc,
                             [&k](double t) { return 1.0 - exp_fast(-3.0 * t, k); });
          break;
        case BK_GAUSSIAN:
          half_pass_near<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, near, near2, acc,
                             [&k](double t) { return 1.0 - exp_fast(-3.0 * t * t, k); });
          break;
        case BK_CUBIC:
          half_pass_near<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, near, near2, acc, [](double t) {
            const double t2 = t * t;
            return t >= 1.0 ? 1.0 : t2 * (7.0 + t * (-8.75 + t2 * (3.5 - 0.75 * t2)));
          });
          break;
        case BK_PENTASPHERICAL:
          half_pass_near<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, near, near2, acc, [](double t) {
            const double t2 = t * t;
            return t >= 1.0 ? 1.0 : t * (1.875 + t2 * (-1.25 + 0.375 * t2));
          });
          break;
        default:
          half_pass_near<CH>(PJ, SJ, off, me, r, n, d0, sill, ir, first, near, near2, acc,
                             [](double) { return __builtin_nan(""); });
      }
    }
#pragma unroll
    for (int dd = 0; dd < CH; ++dd) {
      const int d = d0 + dd;
      const bool valid = d <= half || (even && d == n / 2 && r < n / 2);
      if (valid) {
        int j = r + d;
        j = j >= n ? j - n : j;
        T[j < r ? r * (r + 1) / 2 + j : j * (j + 1) / 2 + r] = c0 - (nug + acc[dd]);
      }
    }
  }
  return near;
}

// The covariance of sample r's record to the target (cov_fast64 on the
// record's frame coordinates, read where they lie in LDS).
__device__ __forceinline__ double rhs_record(const DevVariogram& v, const double* me) {
#pragma clang fp contract(fast)
  const int nf0 = 3 + 3 * v.nug_frame;
  if (!(norm3_fast(me[nf0], me[nf0 + 1], me[nf0 + 2]) > 0.0)) return v.c0;
  double s = 0.0;
  for (int k = 0; k < v.S; ++k) {
    const int off = 3 + 3 * v.frame[k];
    const double h = norm3_fast(me[off], me[off + 1], me[off + 2]);
    s += v.sill[k] * shape_fast64(v.model[k], v.param[k], h, v.inv_range[k]);
  }
  return v.c0 - (v.nugget + s);
}

// The pair phase of the refined kernels for the system of target t (n
// samples; lane r of group g): loads the samples as records in LDS (world
// coordinates and the projection on every frame relative to the target,
// kernel.rs:36-42), tests coincident samples (coincident.rs:13-15, 27-35; with
// report_guard the system is handed to the guarded kernel), computes the
// right-hand side c of sample r (the covariance to the target, or to the
// block) and the strict lower triangle of the FP64 matrix into T (each pair
// once, packed). Returns the guard.
// With NEAR the coincidence test is left to the FP64 kernel: the function
// returns whether a pair of the system lies within v.near2 in the frame of
// structure 0 (a superset of the coincident pairs), and the caller hands such
// a system to the FP64 kernel, which applies the exact test.
template <int NMAX, bool NEAR = false>
__device__ __forceinline__ bool refined_pairs(const SolveArgs& a, int t, int n, int r, int g, bool active,
                                              bool report_guard, double* PJ, int SJ, double* T, double& z,
                                              double& e, double& c) {
  const int nf = a.v.nf;
  const double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
  // Sample r of the system: world coordinates and its projection on every
  // frame relative to the target (kernel.rs:36-42), as a record in LDS.
  double wx = 0.0, wy = 0.0, wz = 0.0;
  z = 0.0;
  e = 0.0;
  {
#pragma clang fp contract(fast)
    double* me = PJ + r * SJ;
    if (r < n) {
      const int id = a.nidx[(int64_t)t * a.stride + r];
      wx = a.sx[id];
      wy = a.sy[id];
      wz = a.sz[id];
      z = a.sv[id];
      e = a.serr ? a.serr[id] : 0.0;
    }
    const double d0 = wx - t0, d1 = wy - t1, d2 = wz - t2;
    me[0] = wx;
    me[1] = wy;
    me[2] = wz;
    for (int f = 0; f < nf; ++f) {
      const double* F = a.v.F[f];
      const bool id = a.v.ident[f];
      me[3 + 3 * f] = id ? d0 : F[0] * d0 + F[1] * d1 + F[2] * d2;
      me[4 + 3 * f] = id ? d1 : F[3] * d0 + F[4] * d1 + F[5] * d2;
      me[5 + 3 * f] = id ? d2 : F[6] * d0 + F[7] * d1 + F[8] * d2;
    }
  }
  wave_sync();
  // Coincident samples (coincident.rs:13-15, 27-35) send the system to the
  // guarded kernel.
  auto coincident = [&]() {
    bool close = false;
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      if (j < r && r < n) {
        const double* pj = PJ + j * SJ;
        const double u = wx - pj[0], vv = wy - pj[1], w = wz - pj[2];
        close |= (u * u + vv * vv) + w * w <= EPS2;
      }
      if (j % 4 == 3) asm volatile("" ::: "memory");
    }
    return group_any<NMAX>(close, g);
  };
  // The right-hand side of sample r.
  auto rhs = [&]() {
    double cc = 0.0;
    if (a.block) {
      cc = rhs_fast_world(a, t0, t1, t2, wx, wy, wz);
    } else {
      double h[BK_MAX_FRAMES];
      const double* me = PJ + r * SJ;
      for (int f = 0; f < nf; ++f) h[f] = norm3_fast(me[3 + 3 * f], me[4 + 3 * f], me[5 + 3 * f]);
      cc = cov_fast64(a.v, h);
    }
    return cc;
  };
  bool guard;
  c = 0.0;
  if constexpr (NEAR) {
    if (active && r < n) c = a.block ? rhs_fast_world(a, t0, t1, t2, wx, wy, wz) : rhs_record(a.v, PJ + r * SJ);
    const bool near = active && pairs_cov_near<NMAX>(a.v, PJ, SJ, r, n, T);
    guard = group_any<NMAX>(near, g);
  } else {
    guard = coincident();
    const bool go = active && !guard;
    if (go && r < n) c = rhs();
    if (go) pairs_cov_fast<NMAX>(a.v, PJ, SJ, r, n, T);
  }
  if (report_guard && guard && r == 0 && active) {
    a.status[t] = BK_ST_GUARD;
    a.guard[atomicAdd(a.guard_count, 1)] = t;
  }
  wave_sync();
  return guard;
}

// The refined front for one system per NMAX lanes: refined_pairs, then row r
// of C (c0 + e on the diagonal; c0 on the diagonal and 0 elsewhere beyond n)
// in crow (NMAX <= 32) or RowL, and row r of (C / c0)^-1 in x by FP32
// Gauss-Jordan.
template <int NMAX>
__device__ __forceinline__ bool refined_front(const SolveArgs& a, int t, int n, int r, int g, int lane, bool active,
                                              bool report_guard, double* PJ, int SJ, float* P2, double* T,
                                              double* RowL, double& z, double& e, double& c, float (&x)[NMAX],
                                              double (&crow)[NMAX <= 32 ? NMAX : 1], bool& bad, float& dmin) {
  constexpr bool ROWREG = NMAX <= 32;
  const bool guard = refined_pairs<NMAX>(a, t, n, r, g, active, report_guard, PJ, SJ, T, z, e, c);
  const bool go = active && !guard;
  const bool live = go && r < n;
  const double c0 = a.v.c0;
  const double ic0 = 1.0 / c0;
  // Row r of the FP64 matrix: in registers for NMAX <= 32, else in LDS
  // ([k][lane], lane-private); the FP32 copy scaled by 1/c0 starts the
  // elimination.
#pragma unroll
  for (int k = 0; k < NMAX; ++k) {
    double v = (k == r) ? (live ? c0 + e : c0) : 0.0;
    if (live && k < n && k != r) v = k < r ? T[r * (r + 1) / 2 + k] : T[k * (k + 1) / 2 + r];
    if constexpr (ROWREG)
      crow[k] = v;
    else
      RowL[k * BK_WAVE + lane] = v;
    x[k] = (float)(v * ic0);
  }
  // Gauss-Jordan (the sweep operator): after sweeping pivot k, column k holds
  // the inverse's column. Pivots go in blocks K = {k .. k + 3}: lanes k .. k + 3
  // publish their rows, every lane inverts the 4 x 4 pivot block P by four
  // single sweeps (its pivots are those of the four single sweeps of the whole
  // matrix, so the small-pivot test is the same) and applies the block sweep,
  // x_j -= sum_q s_q row_q[j] with s = x_K P^-1, one float4 chunk of columns
  // at a time. One LDS round trip per four pivots. The four rows take P2 and
  // the right-hand-side buffer after it (4 NMAX floats), which the solves use
  // only after the elimination.
  bad = false;
  dmin = 3.4e38f;
  float* rows = P2;
#pragma unroll
  for (int k = 0; k < NMAX; k += 4) {
    const int pr = r - k;  // 0 .. 3 on the pivot rows
    const bool piv_row = pr >= 0 && pr < 4;
    if (piv_row) {
      float* dst = rows + pr * NMAX;
#pragma unroll
      for (int c4 = 0; c4 < NMAX / 4; ++c4)
        reinterpret_cast<float4*>(dst)[c4] = make_float4(x[4 * c4], x[4 * c4 + 1], x[4 * c4 + 2], x[4 * c4 + 3]);
    }
    wave_sync();
    float P[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 q = reinterpret_cast<const float4*>(rows + i * NMAX)[k / 4];
      P[i][0] = q.x;
      P[i][1] = q.y;
      P[i][2] = q.z;
      P[i][3] = q.w;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const float pv = P[p][p];
      bad |= !(pv > 1e-5f);
      dmin = fminf(dmin, pv);
      const float ip = __builtin_amdgcn_rcpf(pv);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j != p) P[p][j] *= ip;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i == p) continue;
        const float f = P[i][p];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (j != p) P[i][j] = fmaf(-f, P[p][j], P[i][j]);
        P[i][p] = -f * ip;
      }
      P[p][p] = ip;
    }
    // s = x_K P^-1; on pivot row i, s_q = [q == i] - (P^-1)_iq, which turns
    // the row (equal to row_i) into row i of P^-1 [row_0; ..; row_3]. The
    // pivot row's own row of P^-1 is blended with one-hot factors (a select
    // on the lane's row index would index the block dynamically).
    const float e0 = pr == 0 ? 1.0f : 0.0f, e1 = pr == 1 ? 1.0f : 0.0f, e2 = pr == 2 ? 1.0f : 0.0f,
                e3 = pr == 3 ? 1.0f : 0.0f;
    float own[4], s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      own[q] = fmaf(P[3][q], e3, fmaf(P[2][q], e2, fmaf(P[1][q], e1, P[0][q] * e0)));
      float acc = x[k] * P[0][q];
#pragma unroll
      for (int i = 1; i < 4; ++i) acc = fmaf(x[k + i], P[i][q], acc);
      s[q] = piv_row ? ((pr == q ? 1.0f : 0.0f) - own[q]) : acc;
    }
#pragma unroll
    for (int c4 = 0; c4 < NMAX / 4; ++c4) {
      if (c4 == k / 4) continue;
      float4 q[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) q[i] = reinterpret_cast<const float4*>(rows + i * NMAX)[c4];
      float* xc = x + 4 * c4;
      xc[0] = fmaf(-s[3], q[3].x, fmaf(-s[2], q[2].x, fmaf(-s[1], q[1].x, fmaf(-s[0], q[0].x, xc[0]))));
      xc[1] = fmaf(-s[3], q[3].y, fmaf(-s[2], q[2].y, fmaf(-s[1], q[1].y, fmaf(-s[0], q[0].y, xc[1]))));
      xc[2] = fmaf(-s[3], q[3].z, fmaf(-s[2], q[2].z, fmaf(-s[1], q[1].z, fmaf(-s[0], q[0].z, xc[2]))));
      xc[3] = fmaf(-s[3], q[3].w, fmaf(-s[2], q[2].w, fmaf(-s[1], q[1].w, fmaf(-s[0], q[0].w, xc[3]))));
      if (c4 % 2 == 1) asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) x[k + q] = piv_row ? own[q] : -s[q];
    wave_sync();
  }
  return guard;
}

template <int NMAX>
__device__ __forceinline__ void krige32_blk(const SolveArgs& a, int64_t nl, int unit) {
  constexpr int G = BK_WAVE / NMAX;
  constexpr bool ROWREG = NMAX <= 32;
  extern __shared__ double dyn[];
  const int nf = a.v.nf;
  const int SJ = rec_doubles(nf);
  const int per = sys32_doubles(NMAX, nf);
  const int tid = opaque_tid();
  const int wave = tid / BK_WAVE, lane = tid % BK_WAVE;
  const int g = lane / NMAX, r = lane % NMAX;
  double* S = dyn + (wave * G + g) * per;
  double* PJ = S;
  double* X2 = PJ + NMAX * SJ;
  float* P2 = (float*)(X2 + 2 * NMAX);
  float* B2 = P2 + 2 * NMAX;
  double* T = (double*)(B2 + 2 * NMAX);  // packed strict lower triangle
  double* RowL = T + (NMAX * (NMAX + 1) / 2 + 1) / 2 * 2;  // NMAX = 64 only: [k][lane]
  const int64_t sys = (int64_t)unit * G + g;
  const bool active = sys < nl;
  const int t = active ? a.list[sys] : 0;
  const int n = active ? a.ncount[t] : 0;
  double z, e, c;
  double crow[ROWREG ? NMAX : 1];
  float x[NMAX];
  bool bad;
  float dmin;
  const bool guard =
      refined_front<NMAX>(a, t, n, r, g, lane, active, true, PJ, SJ, P2, T, RowL, z, e, c, x, crow, bad, dmin);
  const bool go = active && !guard;
  const bool live = go && r < n;
  const double c0 = a.v.c0;
  const double ic0 = 1.0 / c0;
  // y = X v for two right-hand sides (v in FP32, one float2 per lane).
  auto apply = [&](float va, float vb, float& ya, float& yb) {
    reinterpret_cast<float2*>(B2)[r] = make_float2(va, vb);
    wave_sync();
    float sa0 = 0.0f, sb0 = 0.0f, sa1 = 0.0f, sb1 = 0.0f;
#pragma unroll
    for (int c2 = 0; c2 < NMAX / 2; ++c2) {
      float4 q = reinterpret_cast<const float4*>(B2)[c2];
      sa0 = fmaf(x[2 * c2], q.x, sa0);
      sb0 = fmaf(x[2 * c2], q.y, sb0);
      sa1 = fmaf(x[2 * c2 + 1], q.z, sa1);
      sb1 = fmaf(x[2 * c2 + 1], q.w, sb1);
      if (c2 % 4 == 3) asm volatile("" ::: "memory");
    }
    ya = sa0 + sa1;
    yb = sb0 + sb1;
    wave_sync();
  };
  const bool ordinary = a.kind != BK_SIMPLE;
  const double ba = live ? c : 0.0, bb = (live && ordinary) ? 1.0 : 0.0;
  float fa, fb;
  apply((float)(ba * ic0), (float)(bb * ic0), fa, fb);
  double xa = fa, xb = fb;
  const float tolf = (float)a.tol;
  const float ma = group_maxf<NMAX>(fabsf(fa)), mb = group_maxf<NMAX>(fabsf(fb));
  int it = 0;
  bool conv = false;
  // A flagged elimination (bad is the same on the group's lanes) goes to the
  // FP64 kernel without refinement sweeps.
  for (; it < a.max_iter && !bad; ++it) {
    reinterpret_cast<double2*>(X2)[r] = make_double2(xa, xb);
    wave_sync();
    double ra = ba, rb = bb;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) {
      double2 q = reinterpret_cast<const double2*>(X2)[k];
      double ck;
      if constexpr (ROWREG)
        ck = crow[k];
      else
        ck = RowL[k * BK_WAVE + lane];
      ra = fma(-ck, q.x, ra);
      rb = fma(-ck, q.y, rb);
      if (k % 4 == 3) asm volatile("" ::: "memory");
    }
    wave_sync();
    float da, db;
    apply((float)(ra * ic0), (float)(rb * ic0), da, db);
    xa += (double)da;
    xb += (double)db;
    bool done;
    if (it == 0) {
      const float delta = group_maxf<NMAX>(fmaxf(rel_corr(da, ma), rel_corr(db, mb)));
      done = group_all<NMAX>(!__builtin_isnan(da) && !__builtin_isnan(db), g) &&
             first_sweep_converged(delta, (float)NMAX, dmin, tolf);
    } else {
      done = group_all<NMAX>(fabsf(da) <= tolf * ma && fabsf(db) <= tolf * mb, g);
    }
    if (done) {
      conv = true;
      ++it;
      break;
    }
  }
  if (active && !guard && (bad || !conv)) {
    if (r == 0) {
      a.status[t] = BK_ST_REDO64;
      a.redo[atomicAdd(a.redo_count, 1)] = t;
    }
  }
  const bool ok = go && !bad && conv;
  // The Schur scalar 1'C^-1 1 and the weights in FP64 (kernel.rs:86-97).
  const bool in = ok && r < n;
  double sa = group_sum<NMAX>(in ? xa : 0.0);
  double sb = group_sum<NMAX>(in ? xb : 0.0);
  double mu = ordinary ? (sa - 1.0) / sb : 0.0;
  double w = xa - (ordinary ? mu * xb : 0.0);
  if (ok) {
    finish_outputs<NMAX>(a, t, r, n, true, BK_OK, w, z, c, mu, a.block ? a.cbb : c0, a.block ? a.cbb : c0);
    if (a.iters && r == 0) a.iters[t] = it;
  }
}

template <int NMAX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NMAX == 64 ? 1 : 2)))
k_krige32(const SolveArgs a) {
  const int64_t nl = list_len(a);
  UnitLoop q = unit_loop((nl + BK_WAVE / NMAX - 1) / (BK_WAVE / NMAX), a.work);
  while (q.unit < q.end) {
    // The next unit is taken before this one is solved, hiding the atomic.
    const int next = unit_next(q);
    krige32_blk<NMAX>(kernarg_first<SolveArgs>(), nl, q.unit);
    wave_sync();
    q.unit = next;
  }
}

#if BK_PART != 2
// ---------------------------------------------------------------------------
// FP32 + FP64 refinement with the elimination on the matrix cores: two
// systems of at most 32 samples per wavefront.
//
// The pair phase is refined_pairs (lane r of group g on sample r of system g).
// The matrices then live in the accumulator of v_mfma_f32_32x32x1_2b_f32,
// whose two blocks are the two systems: block b takes entry i of its
// operands from lane 32 b + i, and lane l = j + 32 h holds column j of block b
// at the rows i = 8 q + 4 h + p (q, p < 4) in accumulator register 16 b + 4 q + p.
// The FP64 matrices stay in LDS as packed triangles for the residuals (C is
// symmetric, so lane j's column is row j as well). Gauss-Jordan
// pivot k is the rank-one update X -= u v' with u_i = X_ik / d and v_j = X_kj
// off the pivot, u_k = 1 - 1/d and v_k = d + 1 (d = X_kk): it leaves 1/d on
// the diagonal, row k / d, column k / -d and the Schur complement elsewhere,
// the in-place Gauss-Jordan inverse (no pivoting: the matrix is SPD and the
// pivots are its Cholesky pivots squared). Row k of both systems is one
// accumulator register of the half-wave (k / 4) % 2; one permlane32 swap puts
// system b's row in lanes 32 b + j. Column k follows from row k: after pivots
// 0 .. k-1 of a symmetric matrix, X_ik = -X_ki for i < k and X_ik = X_ki for
// i > k. A pivot costs about a dozen VALU instructions besides the matrix
// instruction, where a lane-per-row sweep issues 32 fused multiply-adds per
// row.
//
// The pair phase (pairs_q32) keeps each sample's projections on the frames,
// relative to the target, as a record of rec_q(nf) doubles (an odd count, so
// that lanes reading consecutive records spread over the banks). Lane r takes
// the pairs (r, (r + d) mod 32), d = 1 .. 16: the same 32-sample circulant in
// every system whatever its n. The FP64 matrix of a system is Q[d][k] = C(k,
// (k + d) mod 32) with row stride QS (Q[0][k] the diagonal): a lane writes
// row d at its own column, so the stores have an immediate offset per d and
// no bank conflicts. Entry (i, j) with delta = (j - i) mod 32 lies at
// Q[delta][i] for delta < 16 and at Q[32 - delta][j] otherwise (Q[16][k] and
// Q[16][k + 16] hold the same pair). Samples beyond n, and every sample of an
// inactive system, get sentinel records 1e30 (lane + 1) in every coordinate:
// their pairs are finite and never near, and their entries are reset to the
// identity's (0 off the diagonal, c0 on it) after the pairs.
__host__ __device__ constexpr int rec_q(int nf) { return (3 * nf) | 1; }
constexpr int kQS = 35;           // row stride of Q in doubles (odd: column reads spread over the banks)
constexpr int kQSys = 17 * kQS;   // doubles of one system's Q (rows d = 0 .. 16)

__host__ __device__ constexpr int qrec_doubles(int nf) {
  // Records of both systems; later the iterates and product vectors (192
  // doubles).
  return 2 * 32 * rec_q(nf) > 192 ? 2 * 32 * rec_q(nf) : 192;
}

__host__ __device__ constexpr int sysm_doubles(int nf) { return qrec_doubles(nf) + 2 * kQSys; }

// The same for the refined gradient (k_loo_grad32m): the record area later
// holds three iterates and product inputs per row (384 doubles), then each
// sample's position and (lambda, v) (256 doubles).
__host__ __device__ constexpr int qrecl_doubles(int nf) {
  return 2 * 32 * rec_q(nf) > 384 ? 2 * 32 * rec_q(nf) : 384;
}
__host__ __device__ constexpr int sysl_doubles(int nf) { return qrecl_doubles(nf) + 2 * kQSys; }

typedef float bk_floatx32 __attribute__((ext_vector_type(32)));

// v of lane l plus v of lane l ^ 32, the same bits in both lanes.
__device__ __forceinline__ float halves_sum(float v) {
  const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}

__device__ __forceinline__ double halves_sum(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  const double x0 = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  const double x1 = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
  return x0 + x1;
}

// The value of lane l ^ 32 (h = l / 32).
__device__ __forceinline__ double other_half(double v, int h) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  const unsigned l = h ? lo[0] : lo[1], u = h ? hi[0] : hi[1];
  return __longlong_as_double((long long)(((unsigned long long)u << 32) | l));
}

// One structure's contributions to eight of the lane's pairs in pairs_circ:
// acc[dd] += f(l2), l2 the squared lag in the structure's frame (at `off` in
// the records); with `first`, the near test on that frame.
template <class Contrib>
__device__ __forceinline__ void q_pass(const double* me, const uint32_t (&pk)[4], int off, bool first,
                                       bool& near, double near2, double (&acc)[8], Contrib f) {
#pragma clang fp contract(fast)
  extern __shared__ double dyn[];
  const double q0 = me[off], q1 = me[off + 1], q2 = me[off + 2];
#pragma unroll
  for (int dd = 0; dd < 8; ++dd) {
    // The partner's record: its offset in dynamic LDS, unpacked from pk in an
    // instruction of its own per structure, so that the offsets do not stay
    // live across the structures (pk holds two per register).
    uint32_t o;
    if (dd % 2 == 0)
      asm volatile("v_and_b32 %0, 0xffff, %1" : "=v"(o) : "v"(pk[dd / 2]));
    else
      asm volatile("v_lshrrev_b32 %0, 16, %1" : "=v"(o) : "v"(pk[dd / 2]));
    const double* p = reinterpret_cast<const double*>(reinterpret_cast<const char*>(dyn) + o) + off;
    const double d0 = q0 - p[0], d1 = q1 - p[1], d2 = q2 - p[2];
    const double l2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (first) near |= l2 <= near2;
    acc[dd] += f(l2);
    if (dd % 4 == 3) asm volatile("" ::: "memory");
  }
}

// Calls f with each structure's contribution to a covariance, as a function
// of the squared lag in the structure's frame, in the fast FP64 forms with
// the constants folded: C = K + sum of the contributions, K = c0 - nugget -
// (the sills of the exponential and Gaussian structures); exponential and
// Gaussian contribute sill e^x (x clamped at -745), spherical -sill h (1.5 /
// a - 0.5 l2 / a^3) below the range and -sill beyond, cubic and
// pentaspherical -sill shape(h / a). Every contribution is 0 at l2 = 0, so K
// is the covariance of distinct coincident points (nugget left out). Models
// without a fast form contribute NaN (the host sends such variograms to the
// FP64 kernel).
template <class Each>
__device__ __forceinline__ double fast_structures(const DevVariogram& v, const ExpConstants& k, Each each) {
#pragma clang fp contract(fast)
  double K = v.c0 - v.nugget;
  for (int st = 0; st < v.S; ++st) {
    const int off = 3 * v.frame[st];
    const double sill = v.sill[st], ir = v.inv_range[st];
    const bool first = st == 0;
    switch (v.model[st]) {
      case BK_SPHERICAL: {
        const double A = 1.5 * sill * ir, B = 0.5 * sill * (ir * ir * ir), a2 = 1.0 / (ir * ir);
        each(off, first, [=](double l2) {
          const double h = sqrt_pos(l2);
          return l2 >= a2 ? -sill : -(h * (A - B * l2));
        });
        break;
      }
      case BK_EXPONENTIAL: {
        K -= sill;
        const double m = -3.0 * ir;
        each(off, first, [=, &k](double l2) { return sill * exp_core(fmax(sqrt_pos(l2) * m, -745.0), k); });
        break;
      }
      case BK_GAUSSIAN: {
        K -= sill;
        const double m = -3.0 * ir * ir;
        each(off, first, [=, &k](double l2) { return sill * exp_core(fmax(l2 * m, -745.0), k); });
        break;
      }
      case BK_CUBIC:
        each(off, first, [=](double l2) {
          const double t = sqrt_pos(l2) * ir, t2 = t * t;
          return t >= 1.0 ? -sill : -sill * (t2 * (7.0 + t * (-8.75 + t2 * (3.5 - 0.75 * t2))));
        });
        break;
      case BK_PENTASPHERICAL:
        each(off, first, [=](double l2) {
          const double t = sqrt_pos(l2) * ir, t2 = t * t;
          return t >= 1.0 ? -sill : -sill * (t * (1.875 + t2 * (-1.25 + 0.375 * t2)));
        });
        break;
      default:
        each(off, first, [](double) { return __builtin_nan(""); });
    }
  }
  return K;
}

// Pairs (r, (r + d) mod N), d = D0 + 1 .. D0 + 8, of one system of at most N
// samples (records at R, lane r on sample r): out[dd] = C of pair D0 + dd + 1
// (fast_structures). Returns whether a pair of the lane lies within near2 in
// the frame of structure 0.
template <int N, int D0>
__device__ __forceinline__ bool pairs_circ(const DevVariogram& v, const double* R, int SQ, int r, double (&out)[8]) {
#pragma clang fp contract(fast)
  extern __shared__ double dyn[];
  const double* me = R + r * SQ;
  // Byte offsets of the partners' records in dynamic LDS (below 64 KiB), two
  // per register.
  uint32_t pk[4];
  {
    const uint32_t base = (uint32_t)(reinterpret_cast<const char*>(R) - reinterpret_cast<const char*>(dyn));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t lo = base + (uint32_t)(((r + D0 + 2 * i + 1) & (N - 1)) * SQ) * 8u;
      const uint32_t hi = base + (uint32_t)(((r + D0 + 2 * i + 2) & (N - 1)) * SQ) * 8u;
      pk[i] = lo | (hi << 16);
    }
  }
#pragma unroll
  for (int dd = 0; dd < 8; ++dd) out[dd] = 0.0;
  bool near = false;
  const double near2 = v.near2;
  const ExpConstants k = exp_constants();
  const double K = fast_structures(v, k, [&](int off, bool first, auto f) {
    q_pass(me, pk, off, first, near, near2, out, f);
  });
#pragma unroll
  for (int dd = 0; dd < 8; ++dd) out[dd] += K;
  return near;
}

// Pairs d = D0 + 1 .. D0 + 8 of one system of at most 32 samples into Q rows
// D0 + 1 .. D0 + 8.
template <int D0>
__device__ __forceinline__ bool pairs_q32_part(const DevVariogram& v, const double* R, int SQ, int r, double* Qg) {
  double c[8];
  const bool near = pairs_circ<32, D0>(v, R, SQ, r, c);
#pragma unroll
  for (int dd = 0; dd < 8; ++dd) Qg[(D0 + dd + 1) * kQS + r] = c[dd];
  return near;
}

// Mean covariance between a sample (record me: F_f (w - t) on each frame,
// relative to the block centre t) and the block's discretisation points t +
// o_p (F_f o_p in a.poffs), in the fast forms of fast_structures; the nugget
// is left out for every point (npts > 1, block.rs:100-118).
__device__ __forceinline__ double rhs_block_fast(const SolveArgs& a, const double* me) {
#pragma clang fp contract(fast)
  const DevVariogram& v = a.v;
  const int nf = v.nf, P = a.npts;
  const ExpConstants k = exp_constants();
  double acc = 0.0;
  const double K = fast_structures(v, k, [&](int off, bool, auto f) {
    const double q0 = me[off], q1 = me[off + 1], q2 = me[off + 2];
    const double* o = a.poffs + off;
    for (int p = 0; p < P; ++p, o += 3 * nf) {
      const double d0 = q0 - o[0], d1 = q1 - o[1], d2 = q2 - o[2];
      acc += f(d0 * d0 + d1 * d1 + d2 * d2);
    }
  });
  return K + acc / (double)P;
}

// The record of sample r of the system of target t (n samples) for the
// circulant pair phases: the projection on each frame relative to the target
// (kernel.rs:36-42), or the sentinel 1e30 (lane + 1) in every coordinate for
// r >= n; and the sample's value z, error variance e and right-hand side c
// (the covariance to the target, or the mean over the block's points).
__device__ __forceinline__ void circ_record(const SolveArgs& a, int t, bool mine, int idx_r, int lane, double* me,
                                            double& z, double& e, double& c) {
#pragma clang fp contract(fast)
  const int nf = a.v.nf;
  const double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
  double wx = 0.0, wy = 0.0, wz = 0.0;
  z = 0.0;
  e = 0.0;
  c = 0.0;
  if (mine) {
    const int id = a.nidx[(int64_t)t * a.stride + idx_r];
    wx = a.sx[id];
    wy = a.sy[id];
    wz = a.sz[id];
    z = a.sv[id];
    e = a.serr ? a.serr[id] : 0.0;
  }
  const double d0 = wx - t0, d1 = wy - t1, d2 = wz - t2;
  const double sentinel = 1e30 * (double)(lane + 1);
  double h[BK_MAX_FRAMES];
  for (int f = 0; f < nf; ++f) {
    const double* F = a.v.F[f];
    const bool id = a.v.ident[f];
    const double p0 = id ? d0 : F[0] * d0 + F[1] * d1 + F[2] * d2;
    const double p1 = id ? d1 : F[3] * d0 + F[4] * d1 + F[5] * d2;
    const double p2 = id ? d2 : F[6] * d0 + F[7] * d1 + F[8] * d2;
    me[3 * f] = mine ? p0 : sentinel;
    me[3 * f + 1] = mine ? p1 : sentinel;
    me[3 * f + 2] = mine ? p2 : sentinel;
    h[f] = norm3_fast(p0, p1, p2);
  }
  if (mine) {
    if (!a.block)
      c = cov_fast64(a.v, h);
    else if (a.npts > 1)
      c = rhs_block_fast(a, me);
    else
      c = rhs_fast_world(a, t0, t1, t2, wx, wy, wz);
  }
}

// The pairs of one system (see pairs_q32_part), in two halves of eight
// offsets to bound the registers.
__device__ __forceinline__ bool pairs_q32(const DevVariogram& v, const double* R, int SQ, int r, double* Qg) {
  const bool n0 = pairs_q32_part<0>(v, R, SQ, r, Qg);
  const bool n1 = pairs_q32_part<8>(v, R, SQ, r, Qg);
  return n0 || n1;
}

// Offset in a system's Q of entry (i, j) (see pairs_q32).
__device__ __forceinline__ int q_entry(int i, int j) {
  const int dl = (j - i) & 31;
  return dl >= 1 && dl <= 15 ? dl * kQS + i : ((32 - dl) & 31) * kQS + j;
}

// Gauss-Jordan pivots K .. 31 of both systems in the accumulator of
// v_mfma_f32_32x32x1_2b_f32 (see krige32m_blk), one rank-one update each;
// dmin takes the smallest pivot of the lane's system (float bits as an
// integer).
template <int K>
__device__ __forceinline__ void gj_pivots32(bk_floatx32& X, int& dmin) {
  constexpr int rk = 4 * (K / 8) + (K % 4);
  const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(X[rk]), __float_as_uint(X[16 + rk]), false, false);
  const float row = __uint_as_float(((K / 4) % 2) ? p[1] : p[0]);
  // The pivot of the lane's system, lane K of its half (a swizzle broadcast
  // within 32 lanes: and-mask 0, or-mask K).
  const int pb = __builtin_amdgcn_ds_swizzle(__float_as_int(row), K << 5);
  dmin = min(dmin, pb);
  const float inv = __builtin_amdgcn_rcpf(__int_as_float(pb));
  // Operand entry i of lane 32 b + i: -u_i (from row K by the sign rule) and
  // v_i. The lane's offset from the pivot is opaque, so that these are
  // selects rather than branches on the pivots' lane masks.
  int dk = (int)(threadIdx.x & 31) - K;
  asm volatile("" : "+v"(dk));
  const float sg = dk < 0 ? inv : -inv;
  float u = row * sg;
  u = dk == 0 ? inv - 1.0f : u;
  const float v = dk == 0 ? row + 1.0f : row;
  X = __builtin_amdgcn_mfma_f32_32x32x1f32(u, v, X, 0, 0, 0);
  if constexpr (K + 1 < 32) gj_pivots32<K + 1>(X, dmin);
}

// Returns the systems this unit solved by refinement and their sweeps (x, y),
// the same in every lane.
// Byte offset in dynamic LDS of entry k of the lane's packed offsets (two
// 16-bit offsets per register), unpacked by an instruction of its own at each
// use so that the sixteen offsets stay packed across the loop over units.
__device__ __forceinline__ uint32_t unpack_off(const uint32_t (&pk)[8], int k) {
  uint32_t o;
  if (k % 2 == 0)
    asm volatile("v_and_b32 %0, 0xffff, %1" : "=v"(o) : "v"(pk[k / 2]));
  else
    asm volatile("v_lshrrev_b32 %0, 16, %1" : "=v"(o) : "v"(pk[k / 2]));
  return o;
}

__device__ __forceinline__ const double& lds_at(uint32_t byte_off) {
  extern __shared__ double dyn[];
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(dyn) + byte_off);
}

__device__ __forceinline__ int2 krige32m_blk(const SolveArgs& a, int64_t nl, int unit, const uint32_t (&ipk)[8]) {
  extern __shared__ double dyn[];
  const int nf = a.v.nf;
  const int SQ = rec_q(nf);
  const int tid = opaque_tid();
  const int wave = tid / BK_WAVE, lane = tid % BK_WAVE;
  const int g = lane >> 5, r = lane & 31;
  double* S = dyn + wave * sysm_doubles(nf);
  double* Q = S + qrec_doubles(nf);  // the two systems' matrices, kQSys doubles each
  double* Qg = Q + g * kQSys;
  const int64_t sys = (int64_t)unit * 2 + g;
  const bool active = sys < nl;
  const int t = active ? a.list[sys] : 0;
  const int n = active ? a.ncount[t] : 0;
  const bool mine = active && r < n;
  const double c0 = a.v.c0;
  const double ic0 = a.v.inv_c0;
  double z, e, c;
  circ_record(a, t, mine, r, lane, S + (g * 32 + r) * SQ, z, e, c);
  wave_sync();
  // A system with a near pair goes to the FP64 kernel (exact coincidence test).
  const bool near = group_any<32>(pairs_q32(a.v, S + g * 32 * SQ, SQ, r, Qg), g);
  const bool go = active && !near;
  const bool go0 = __builtin_amdgcn_readlane(go ? 1 : 0, 0) != 0, go1 = __builtin_amdgcn_readlane(go ? 1 : 0, 32) != 0;
  // The diagonal; entries of samples beyond n take the identity's.
  Qg[r] = mine ? c0 + e : c0;
  if (__builtin_amdgcn_readlane(n, 0) < 32 || __builtin_amdgcn_readlane(n, 32) < 32) {
#pragma unroll
    for (int dd = 1; dd <= 16; ++dd)
      if (!(r < n && ((r + dd) & 31) < n)) Qg[dd * kQS + r] = 0.0;
  }
  wave_sync();
  // Column r of both systems' matrices at the lane's rows i = 8 q + 4 g + p:
  // entry (i, r) of Q at idx (the same in every unit), also read by the
  // residuals.
  bk_floatx32 X;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double* e0 = &lds_at(unpack_off(ipk, k));
#pragma unroll
    for (int b = 0; b < 2; ++b) X[16 * b + k] = (float)(e0[b * kQSys] * ic0);
  }
  // The smallest pivot of each system as an integer over the float bits: for
  // floats >= 0 the order is the same, a negative pivot is a negative integer,
  // and a NaN pivot fails the convergence test later.
  int dmin = 0x7f7fffff;  // of the lane's system
  gj_pivots32<0>(X, dmin);
  const int dmin0 = __builtin_amdgcn_readlane(dmin, 0), dmin1 = __builtin_amdgcn_readlane(dmin, 32);
  const bool bad0 = !(dmin0 > __float_as_int(1e-5f)), bad1 = !(dmin1 > __float_as_int(1e-5f));
  // Solves. Lane (r, g) publishes its own system g's entry r of a vector; the
  // products take the lane's 16 rows of column r of both systems and the two
  // halves' partial sums: C x exactly (C is symmetric), and X' v for X v (the
  // FP32 inverse is symmetric to rounding, which the refinement absorbs).
  double* X2 = S;                    // iterates, double2 [b][j] (the records are no longer read)
  float* V2 = (float*)(S + 2 * 64);  // product inputs, float2 [b][j]
  auto apply = [&](float va, float vb, float (&ya)[2], float (&yb)[2]) {
    reinterpret_cast<float2*>(V2)[g * 32 + r] = make_float2(va, vb);
    wave_sync();
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float sa = 0.0f, sb = 0.0f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 w0 = reinterpret_cast<const float4*>(V2)[16 * b + 4 * q + 2 * g];
        const float4 w1 = reinterpret_cast<const float4*>(V2)[16 * b + 4 * q + 2 * g + 1];
        const int o = 16 * b + 4 * q;
        sa = fmaf(X[o], w0.x, sa);
        sb = fmaf(X[o], w0.y, sb);
        sa = fmaf(X[o + 1], w0.z, sa);
        sb = fmaf(X[o + 1], w0.w, sb);
        sa = fmaf(X[o + 2], w1.x, sa);
        sb = fmaf(X[o + 2], w1.y, sb);
        sa = fmaf(X[o + 3], w1.z, sa);
        sb = fmaf(X[o + 3], w1.w, sb);
      }
      ya[b] = halves_sum(sa);
      yb[b] = halves_sum(sb);
    }
    wave_sync();
  };
  const bool live = go && r < n;
  const bool ordinary = a.kind != BK_SIMPLE;
  const double ba = live ? c : 0.0, bb = (live && ordinary) ? 1.0 : 0.0;
  const double ba_o = other_half(ba, g), bb_o = other_half(bb, g);
  const double rha[2] = {g ? ba_o : ba, g ? ba : ba_o}, rhb[2] = {g ? bb_o : bb, g ? bb : bb_o};
  float fa[2], fb[2];
  apply((float)(ba * ic0), (float)(bb * ic0), fa, fb);
  double xa[2] = {fa[0], fa[1]}, xb[2] = {fb[0], fb[1]};
  const float tolf = (float)a.tol;
  float ma[2], mb[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    ma[b] = group_maxf<32>(fabsf(fa[b]));
    mb[b] = group_maxf<32>(fabsf(fb[b]));
  }
  bool conv0 = false, conv1 = false;
  int it0 = 0, it1 = 0;
  for (int it = 0; it < a.max_iter; ++it) {
    const bool more0 = go0 && !bad0 && !conv0, more1 = go1 && !bad1 && !conv1;
    if (!more0 && !more1) break;
    reinterpret_cast<double2*>(X2)[g * 32 + r] = make_double2(g ? xa[1] : xa[0], g ? xb[1] : xb[0]);
    wave_sync();
    double ra[2], rb[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      double sa = 0.0, sb = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const double2 xv = reinterpret_cast<const double2*>(X2)[32 * b + 8 * q + 4 * g + p];
          const double cv = (&lds_at(unpack_off(ipk, 4 * q + p)))[b * kQSys];
          sa = fma(cv, xv.x, sa);
          sb = fma(cv, xv.y, sb);
        }
      ra[b] = rha[b] - halves_sum(sa);
      rb[b] = rhb[b] - halves_sum(sb);
    }
    wave_sync();
    float da[2], db[2];
    apply((float)((g ? ra[1] : ra[0]) * ic0), (float)((g ? rb[1] : rb[0]) * ic0), da, db);
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      xa[b] += (double)da[b];
      xb[b] += (double)db[b];
    }
    // Every lane holds entry r of both systems' corrections (rows 0..31 twice).
    bool small0, small1;
    if (it == 0) {
      // Half g reduces system g's largest corrections.
      const float delta = group_maxf<32>(fmaxf(rel_corr(g ? da[1] : da[0], g ? ma[1] : ma[0]),
                                               rel_corr(g ? db[1] : db[0], g ? mb[1] : mb[0])));
      const bool done = first_sweep_converged(delta, 32.0f, __int_as_float(g ? dmin1 : dmin0), tolf);
      const bool nan0 = __ballot(__builtin_isnan(da[0]) || __builtin_isnan(db[0])) != 0;
      const bool nan1 = __ballot(__builtin_isnan(da[1]) || __builtin_isnan(db[1])) != 0;
      small0 = !nan0 && __builtin_amdgcn_readlane(done ? 1 : 0, 0) != 0;
      small1 = !nan1 && __builtin_amdgcn_readlane(done ? 1 : 0, 32) != 0;
    } else {
      small0 = __ballot(!(fabsf(da[0]) <= tolf * ma[0] && fabsf(db[0]) <= tolf * mb[0])) == 0;
      small1 = __ballot(!(fabsf(da[1]) <= tolf * ma[1] && fabsf(db[1]) <= tolf * mb[1])) == 0;
    }
    if (more0) {
      it0 = it + 1;
      conv0 = small0;
    }
    if (more1) {
      it1 = it + 1;
      conv1 = small1;
    }
  }
  const bool bad = g ? bad1 : bad0, conv = g ? conv1 : conv0;
  if (active && (near || bad || !conv)) {
    if (r == 0) {
      a.status[t] = BK_ST_REDO64;
      a.redo[atomicAdd(a.redo_count, 1)] = t;
    }
  }
  const bool ok = go && !bad && conv;
  // The Schur scalar 1'C^-1 1 and the weights in FP64 (kernel.rs:86-97).
  const double xag = g ? xa[1] : xa[0], xbg = g ? xb[1] : xb[0];
  const bool in = ok && r < n;
  const double sa = group_sum<32>(in ? xag : 0.0);
  const double sb = group_sum<32>(in ? xbg : 0.0);
  const double mu = ordinary ? (sa - 1.0) / sb : 0.0;
  const double w = xag - (ordinary ? mu * xbg : 0.0);
  if (ok) {
    finish_outputs<32>(a, t, r, n, true, BK_OK, w, z, c, mu, a.block ? a.cbb : c0, a.block ? a.cbb : c0);
    if (a.iters && r == 0) a.iters[t] = g ? it1 : it0;
  }
  const bool ok0 = go0 && !bad0 && conv0, ok1 = go1 && !bad1 && conv1;
  return make_int2((int)ok0 + (int)ok1, (ok0 ? it0 : 0) + (ok1 ? it1 : 0));
}

// Adds a wave's refinement totals to the pass totals.
__device__ __forceinline__ void add_refine_stats(const SolveArgs& a, int2 st) {
  if (a.refine_stats && (threadIdx.x & (BK_WAVE - 1)) == 0 && st.x > 0) {
    atomicAdd(a.refine_stats, st.x);
    atomicAdd(a.refine_stats + 1, st.y);
  }
}

// One system of 33 to 64 samples per wavefront, in the same arithmetic. The
// 64 x 64 matrix is four 32 x 32 tiles held by two accumulators of
// v_mfma_f32_32x32x1_2b_f32: A1 holds tiles (0, 0) and (1, 1) as its blocks 0
// and 1, A2 tiles (0, 1) and (1, 0). Lane l = j + 32 h thus holds columns j
// and 32 + j at the rows i = 32 I + 8 q + 4 h + p. Pivot k updates A1 with u
// and v (entry l of each from lane l) and A2 with u and v's halves swapped
// (entry l ^ 32 of v from lane l). Row k lies in the registers of row block
// k / 32 at the half-wave (k / 4) % 2, one register per column block; a
// permlane32 swap of the two gives v, another the swapped v.
__host__ __device__ constexpr int qrec64_doubles(int nf) {
  // Records (later the iterates and product vectors: 192 doubles).
  return 64 * rec_q(nf) > 192 ? 64 * rec_q(nf) : 192;
}
__host__ __device__ constexpr int sysm64_doubles(int nf) {
  // Records, then the packed triangle with the diagonal.
  return qrec64_doubles(nf) + 2080;
}

__device__ __forceinline__ int2 krige64m_blk(const SolveArgs& a, int64_t nl, int unit) {
  extern __shared__ double dyn[];
  const int nf = a.v.nf;
  const int SQ = rec_q(nf);
  const int tid = opaque_tid();
  const int wave = tid / BK_WAVE, lane = tid % BK_WAVE;
  const int j = lane & 31, h = lane >> 5;
  double* S = dyn + wave * sysm64_doubles(nf);
  double* T = S + qrec64_doubles(nf);  // packed lower triangle with the diagonal, 2080 doubles
  const bool active = unit < nl;
  const int t = active ? a.list[unit] : 0;
  const int n = active ? a.ncount[t] : 0;
  const bool mine = active && lane < n;
  double z, e, c;
  circ_record(a, t, mine, lane, lane, S + lane * SQ, z, e, c);
  wave_sync();
  // Pairs (lane, (lane + d) mod 64), d = 1 .. 32, in four parts of eight, into
  // the packed triangle; a system with a near pair goes to the FP64 kernel
  // (exact coincidence test).
  bool nearl = false;
  auto part = [&](auto d0) {
    constexpr int D0 = decltype(d0)::value;
    double cv[8];
    nearl |= pairs_circ<64, D0>(a.v, S, SQ, lane, cv);
#pragma unroll
    for (int dd = 0; dd < 8; ++dd) {
      const int jj = (lane + D0 + dd + 1) & 63;
      const int hi = max(lane, jj), lo = min(lane, jj);
      T[((hi * (hi + 1)) >> 1) + lo] = cv[dd];
    }
  };
  part(std::integral_constant<int, 0>());
  part(std::integral_constant<int, 8>());
  part(std::integral_constant<int, 16>());
  part(std::integral_constant<int, 24>());
  const bool near = __any(nearl);
  const bool go = active && !near;
  const double c0 = a.v.c0;
  const double ic0 = a.v.inv_c0;
  wave_sync();
  // The diagonal; rows beyond n take the identity's (lane r writes row r).
  {
    const int tr = lane * (lane + 1) / 2;
    if (mine) {
      T[tr + lane] = c0 + e;
    } else {
#pragma unroll
      for (int q = 0; q < 64; ++q)
        if (q <= lane) T[tr + q] = q == lane ? c0 : 0.0;
    }
  }
  wave_sync();
  // Entry (i, col) of the packed triangle for the lane's rows i = 32 I + i0 +
  // 4 h (i0 = 8 q + p) and columns col = 32 J + j: tri(i) + col for i >= col,
  // else tri(col) + i. The triangle stays in LDS for the residuals.
  auto entry = [&](int I, int J, int q, int p) {
    const int i = 32 * I + 8 * q + 4 * h + p, col = 32 * J + j;
    if (I > J) return T[i * (i + 1) / 2 + col];
    if (I < J) return T[col * (col + 1) / 2 + i];
    return T[i >= col ? i * (i + 1) / 2 + col : col * (col + 1) / 2 + i];
  };
  bk_floatx32 A1, A2;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      A1[4 * q + p] = (float)(entry(0, 0, q, p) * ic0);
      A1[16 + 4 * q + p] = (float)(entry(1, 1, q, p) * ic0);
      A2[4 * q + p] = (float)(entry(0, 1, q, p) * ic0);
      A2[16 + 4 * q + p] = (float)(entry(1, 0, q, p) * ic0);
    }
  int dmin = 0x7f7fffff;
  // Rows and columns from n on are the identity's, whose pivots (1) change
  // nothing: they are skipped, and dmin takes 1 in their place.
  const int nmax = __builtin_amdgcn_readfirstlane(n);
  if (nmax < 64) dmin = __float_as_int(1.0f);
#pragma unroll
  for (int k = 0; k < 64; ++k) {
    if (k >= nmax) continue;
    const int K = k / 32, kk = k % 32;
    const int rk = 4 * (kk / 8) + (kk % 4);
    const float R0 = K == 0 ? A1[rk] : A2[16 + rk];  // row k, columns 0..31
    const float R1 = K == 0 ? A2[rk] : A1[16 + rk];  // row k, columns 32..63
    const auto pv = __builtin_amdgcn_permlane32_swap(__float_as_uint(R0), __float_as_uint(R1), false, false);
    const auto ps = __builtin_amdgcn_permlane32_swap(__float_as_uint(R1), __float_as_uint(R0), false, false);
    const bool hi = (kk / 4) % 2;
    const float row = __uint_as_float(hi ? pv[1] : pv[0]);   // X_k,l
    const float rows = __uint_as_float(hi ? ps[1] : ps[0]);  // X_k,l^32
    const int db = __builtin_amdgcn_readlane(__float_as_int(row), k);
    dmin = min(dmin, db);
    const float inv = __builtin_amdgcn_rcpf(__int_as_float(db));
    // The lane's offset from the pivot, opaque so that the compiler compares
    // here rather than holding 64 pivots' lane masks.
    int dk = lane - k, dks = (lane ^ 32) - k;
    asm volatile("" : "+v"(dk), "+v"(dks));
    const float sg = dk < 0 ? inv : -inv;
    float u = row * sg;
    u = dk == 0 ? inv - 1.0f : u;
    const float v = dk == 0 ? row + 1.0f : row;
    const float vs = dks == 0 ? rows + 1.0f : rows;
    A1 = __builtin_amdgcn_mfma_f32_32x32x1f32(u, v, A1, 0, 0, 0);
    A2 = __builtin_amdgcn_mfma_f32_32x32x1f32(u, vs, A2, 0, 0, 0);
  }
  const bool bad = !(dmin > __float_as_int(1e-5f));
  // Solves: lane r publishes entry r of a vector; the products take the
  // lane's 32 rows of columns j and 32 + j and the two halves' partial sums.
  double* X2 = S;                   // iterates, double2 [64] (the records are no longer read)
  float* V2 = (float*)(S + 128);    // product inputs, float2 [64]
  auto apply = [&](float va, float vb, float& ya, float& yb) {
    reinterpret_cast<float2*>(V2)[lane] = make_float2(va, vb);
    wave_sync();
    float sa0 = 0.0f, sb0 = 0.0f, sa1 = 0.0f, sb1 = 0.0f;  // columns j, 32 + j
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 w0 = reinterpret_cast<const float4*>(V2)[16 * I + 4 * q + 2 * h];
        const float4 w1 = reinterpret_cast<const float4*>(V2)[16 * I + 4 * q + 2 * h + 1];
        const float wa[4] = {w0.x, w0.z, w1.x, w1.z}, wb[4] = {w0.y, w0.w, w1.y, w1.w};
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const float x0 = I == 0 ? A1[4 * q + p] : A2[16 + 4 * q + p];  // column j
          const float x1 = I == 0 ? A2[4 * q + p] : A1[16 + 4 * q + p];  // column 32 + j
          sa0 = fmaf(x0, wa[p], sa0);
          sb0 = fmaf(x0, wb[p], sb0);
          sa1 = fmaf(x1, wa[p], sa1);
          sb1 = fmaf(x1, wb[p], sb1);
        }
      }
    const float ya0 = halves_sum(sa0), yb0 = halves_sum(sb0), ya1 = halves_sum(sa1), yb1 = halves_sum(sb1);
    ya = h ? ya1 : ya0;
    yb = h ? yb1 : yb0;
    wave_sync();
  };
  const bool live = go && lane < n;
  const bool ordinary = a.kind != BK_SIMPLE;
  const double ba = live ? c : 0.0, bb = (live && ordinary) ? 1.0 : 0.0;
  float fa, fb;
  apply((float)(ba * ic0), (float)(bb * ic0), fa, fb);
  double xa = fa, xb = fb;
  const float tolf = (float)a.tol;
  const float ma = group_maxf<64>(fabsf(fa)), mb = group_maxf<64>(fabsf(fb));
  bool conv = false;
  int it = 0;
  for (; it < a.max_iter && go && !bad; ++it) {
    reinterpret_cast<double2*>(X2)[lane] = make_double2(xa, xb);
    wave_sync();
    double sa0 = 0.0, sb0 = 0.0, sa1 = 0.0, sb1 = 0.0;
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const double2 xv = reinterpret_cast<const double2*>(X2)[32 * I + 8 * q + 4 * h + p];
          const double c0v = entry(I, 0, q, p), c1v = entry(I, 1, q, p);
          sa0 = fma(c0v, xv.x, sa0);
          sb0 = fma(c0v, xv.y, sb0);
          sa1 = fma(c1v, xv.x, sa1);
          sb1 = fma(c1v, xv.y, sb1);
        }
    // Each column's partial sums over the two halves' rows: lane l's entry
    // is column l (j for h = 0, 32 + j for h = 1).
    const double ca0 = halves_sum(sa0), ca1 = halves_sum(sa1), cb0 = halves_sum(sb0), cb1 = halves_sum(sb1);
    const double ca = h ? ca1 : ca0, cb = h ? cb1 : cb0;
    wave_sync();
    float da, db;
    apply((float)((ba - ca) * ic0), (float)((bb - cb) * ic0), da, db);
    xa += (double)da;
    xb += (double)db;
    bool done;
    if (it == 0) {
      const float delta = group_maxf<64>(fmaxf(rel_corr(da, ma), rel_corr(db, mb)));
      done = __ballot(__builtin_isnan(da) || __builtin_isnan(db)) == 0 &&
             first_sweep_converged(delta, 64.0f, __int_as_float(dmin), tolf);
    } else {
      done = __ballot(!(fabsf(da) <= tolf * ma && fabsf(db) <= tolf * mb)) == 0;
    }
    if (done) {
      conv = true;
      ++it;
      break;
    }
  }
  if (active && (near || bad || !conv)) {
    if (lane == 0) {
      a.status[t] = BK_ST_REDO64;
      a.redo[atomicAdd(a.redo_count, 1)] = t;
    }
  }
  const bool ok = go && !bad && conv;
  const bool in = ok && lane < n;
  const double sa = group_sum<64>(in ? xa : 0.0);
  const double sb = group_sum<64>(in ? xb : 0.0);
  const double mu = ordinary ? (sa - 1.0) / sb : 0.0;
  const double w = xa - (ordinary ? mu * xb : 0.0);
  if (ok) {
    finish_outputs<64>(a, t, lane, n, true, BK_OK, w, z, c, mu, a.block ? a.cbb : c0, a.block ? a.cbb : c0);
    if (a.iters && lane == 0) a.iters[t] = it;
  }
  return make_int2(ok ? 1 : 0, ok ? it : 0);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_krige64m(const SolveArgs a) {
  const int64_t nl = list_len(a);
  UnitLoop q = unit_loop(nl, a.work);
  int2 st = make_int2(0, 0);
  while (q.unit < q.end) {
    const int next = unit_next(q);
    const int2 u = krige64m_blk(kernarg_first<SolveArgs>(), nl, q.unit);
    st.x += u.x;
    st.y += u.y;
    wave_sync();
    q.unit = next;
  }
  add_refine_stats(a, st);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_krige32m(const SolveArgs a) {
  const int64_t nl = list_len(a);
  UnitLoop q = unit_loop((nl + 1) / 2, a.work);
  int2 st = make_int2(0, 0);
  // Byte offsets in dynamic LDS of the lane's matrix entries (i, r) of the
  // first system, i = 8 q + 4 g + p (entry 4 q + p; the second system's lie
  // kQSys doubles on), the same in every unit.
  uint32_t ipk[8];
  {
    const int lane = threadIdx.x % BK_WAVE, g = lane >> 5, r = lane & 31;
    const uint32_t qb = (uint32_t)((threadIdx.x / BK_WAVE) * sysm_doubles(a.v.nf) + qrec_doubles(a.v.nf)) * 8u;
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
      const uint32_t lo = qb + 8u * (uint32_t)q_entry(8 * (k / 4) + 4 * g + k % 4, r);
      const uint32_t hi = qb + 8u * (uint32_t)q_entry(8 * ((k + 1) / 4) + 4 * g + (k + 1) % 4, r);
      ipk[k / 2] = lo | (hi << 16);
    }
  }
  while (q.unit < q.end) {
    // The next unit is taken before this one is solved, hiding the atomic.
    const int next = unit_next(q);
    const int2 u = krige32m_blk(kernarg_first<SolveArgs>(), nl, q.unit, ipk);
    st.x += u.x;
    st.y += u.y;
    wave_sync();
    q.unit = next;
  }
  add_refine_stats(a, st);
}
#endif

// ---------------------------------------------------------------------------
// Leave-one-out cross-validation loss and its adjoint gradient (no
// counterpart in the reference). Per held-out sample i with its fixed
// neighbours, the bordered ordinary system K w = k (w = [lambda; mu],
// k = [c; 1]) gives zhat = lambda'z and sigma2 = c0 - w'k. With the adjoint
// v = K^-1 [z; 0] (one more solve on the same factor):
//   d zhat   = v_l'(dc - dC lambda)
//   d sigma2 = dc0 - 2 lambda'dc + lambda'dC lambda
// and the loss l = (z_i - zhat)^2 + log sigma2 has
//   dl = alpha d zhat + beta d sigma2,  alpha = -2 (z_i - zhat), beta = 1 / sigma2.
// dC and dc are recomputed pair by pair from the parameters (never stored).

template <int NMAX>
__device__ __forceinline__ void group_sum_vec(double* v, int P) {
  for (int p = 0; p < P; ++p) v[p] = group_sum<NMAX>(v[p]);
}

// dg/dt of the normalised shape (crates/variogram/src/model.rs:207-294).
__device__ __forceinline__ double dshape64(int model, double t) {
  switch (model) {
    case BK_SPHERICAL:
      return t >= 1.0 ? 0.0 : 1.5 - 1.5 * t * t;
    case BK_EXPONENTIAL:
      return 3.0 * exp(-3.0 * t);
    case BK_GAUSSIAN:
      return 6.0 * t * exp(-3.0 * t * t);
    case BK_CUBIC:
      return t >= 1.0 ? 0.0 : 14.0 * t - 26.25 * t * t + 17.5 * t * t * t * t - 5.25 * t * t * t * t * t * t;
    case BK_PENTASPHERICAL:
      return t >= 1.0 ? 0.0 : 1.875 - 3.75 * t * t + 1.875 * t * t * t * t;
  }
  return 0.0;
}

__device__ __forceinline__ double shape_t64(int model, double t) {
  switch (model) {
    case BK_SPHERICAL:
      return t >= 1.0 ? 1.0 : 1.5 * t - 0.5 * t * t * t;
    case BK_EXPONENTIAL:
      return 1.0 - exp(-3.0 * t);
    case BK_GAUSSIAN:
      return 1.0 - exp(-3.0 * t * t);
    case BK_CUBIC: {
      if (t >= 1.0) return 1.0;
      double t2 = t * t;
      return 7.0 * t2 - 8.75 * t2 * t + 3.5 * t2 * t2 * t - 0.75 * t2 * t2 * t2 * t;
    }
    case BK_PENTASPHERICAL: {
      if (t >= 1.0) return 1.0;
      double t2 = t * t;
      return 1.875 * t - 1.25 * t2 * t + 0.375 * t2 * t2 * t;
    }
  }
  return 0.0;
}

// Gradient of C(p, q) with respect to the packed parameters, for world
// separation d = q - p. Layout: [0] nugget, [1 + s] sill_s, [1 + S + 3 s + a]
// range a (major, semi, minor) of structure s, [1 + 4 S + b] angle b (azimuth,
// dip, rake; per degree). Coincident points: C = c0.
__device__ __forceinline__ void cov_grad64(const GradArgs& gp, double d0, double d1, double d2, double coef,
                                           double* acc) {
  const int S = gp.S;
  if (d0 == 0.0 && d1 == 0.0 && d2 == 0.0) {
    acc[0] += coef;
    for (int s = 0; s < S; ++s) acc[1 + s] += coef;
    return;
  }
  double r0, r1, r2;
  mv3(gp.R, d0, d1, d2, r0, r1, r2);
  double q[3][3];
  for (int b = 0; b < 3; ++b) mv3(gp.dR[b], d0, d1, d2, q[b][0], q[b][1], q[b][2]);
  for (int s = 0; s < S; ++s) {
    const double* rr = gp.r[s];
    double u0 = r0 / rr[0], u1 = r1 / rr[1], u2 = r2 / rr[2];
    double t = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    double g = shape_t64(gp.model[s], t);
    double gd = dshape64(gp.model[s], t);
    acc[1 + s] += coef * (1.0 - g);
    // dC/dt = -sill g'(t); dt/dr_a = -u_a^2 / (t r_a); dt/dphi = u . (dR_phi d / r) / t
    double k = coef * gp.sill[s] * gd / t;
    acc[1 + S + 3 * s + 0] += k * u0 * u0 / rr[0];
    acc[1 + S + 3 * s + 1] += k * u1 * u1 / rr[1];
    acc[1 + S + 3 * s + 2] += k * u2 * u2 / rr[2];
    for (int b = 0; b < 3; ++b) {
      double dt = u0 * q[b][0] / rr[0] + u1 * q[b][1] / rr[1] + u2 * q[b][2] / rr[2];
      acc[1 + 4 * S + b] -= k * dt;
    }
  }
}

template <int NMAX>
__device__ __forceinline__ void loo_grad_blk(const SolveArgs& a, const GradArgs& gp, int64_t nl, int unit) {
  constexpr int G = BK_WAVE / NMAX;
  extern __shared__ double dyn[];
  const int nf = a.v.nf;
  const int per = sys_doubles(NMAX, nf);
  const int tid = opaque_tid();
  const int wave = tid / BK_WAVE, lane = tid % BK_WAVE;
  const int g = lane / NMAX, r = lane % NMAX;
  double* S = dyn + (wave * G + g) * per;
  double* L = S;
  double* W = L + NMAX * (NMAX + 1) / 2;
  double* V = W + 3 * NMAX;
  double* E = V + NMAX;
  double* PF = E + NMAX;
  double* misc = PF + nf * 3 * NMAX;
  const int64_t sys = (int64_t)unit * G + g;
  const bool active = sys < nl;
  const int t = active ? a.list[sys] : 0;
  const int n = active ? a.ncount[t] : 0;
  double z = 0.0, e = 0.0;
  bool guard = load_system<NMAX>(a, t, r, n, g, W, V, E, PF, z, e);
  const bool go = active && !guard;
  const double c0 = a.v.c0;
  const int npairs = n * (n - 1) / 2;
  if (go)
    for (int q = r; q < npairs; q += NMAX) {
      int i, j;
      pair_of(q, i, j);
      L[tri(i) + j] = pair_cov64(a, W, PF, NMAX, i, j);
    }
  double c = 0.0;
  if (go && r < n) {
    L[tri(r) + r] = c0 + e;
    c = rhs64(a, t, W, PF, NMAX, r);
  }
  wave_sync();
  double reg[NMAX];
#pragma unroll
  for (int k = 0; k < NMAX; ++k) reg[k] = (go && k <= r && r < n) ? L[tri(r) + k] : 0.0;
  bool sing = false;
#pragma unroll
  for (int j = 0; j < NMAX; ++j) {
    const bool act = go && j < n;
    double acc = -0.0;
#pragma unroll
    for (int k = 0; k < j; ++k) acc += reg[k] * L[tri(j) + k];
    double s = reg[j] - acc;
    if (act && r == j) misc[0] = s;
    wave_sync();
    double d = misc[0];
    if (act && (isnan(d) || d <= PIVOT * c0)) sing = true;
    if (act && !sing) {
      d = sqrt(d);
      if (r == j) {
        reg[j] = d;
        L[tri(j) + j] = d;
      } else if (r > j && r < n) {
        double x = s / d;
        reg[j] = x;
        L[tri(r) + j] = x;
      }
    }
    wave_sync();
  }
  // Three right-hand sides: c, 1 and the data z.
  double ya = c, yb = 1.0, yz = z, aa = -0.0, ab = -0.0, az = -0.0;
#pragma unroll
  for (int k = 0; k < NMAX; ++k) {
    const bool act = go && !sing && k < n;
    if (act && r == k) {
      ya = (ya - aa) / reg[k];
      yb = (yb - ab) / reg[k];
      yz = (yz - az) / reg[k];
      misc[1] = ya;
      misc[2] = yb;
      misc[3] = yz;
    }
    wave_sync();
    if (act && r > k && r < n) {
      aa += reg[k] * misc[1];
      ab += reg[k] * misc[2];
      az += reg[k] * misc[3];
    }
    wave_sync();
  }
  double xa = ya, xb = yb, xz = yz;
#pragma unroll
  for (int k = NMAX - 1; k >= 0; --k) {
    const bool act = go && !sing && k < n;
    if (act && r == k) {
      xa = xa / reg[k];
      xb = xb / reg[k];
      xz = xz / reg[k];
      misc[1] = xa;
      misc[2] = xb;
      misc[3] = xz;
    }
    wave_sync();
    if (act && r < k) {
      double lk = L[tri(k) + r];
      xa -= lk * misc[1];
      xb -= lk * misc[2];
      xz -= lk * misc[3];
    }
    wave_sync();
  }
  const bool in = go && !sing && r < n;
  double sa = group_sum<NMAX>(in ? xa : 0.0);
  double sb = group_sum<NMAX>(in ? xb : 0.0);
  double sz = group_sum<NMAX>(in ? xz : 0.0);
  double mu = (sa - 1.0) / sb;
  double vmu = sz / sb;
  double lam = xa - mu * xb;
  double vl = xz - vmu * xb;
  double zhat = group_sum<NMAX>(in ? lam * z : 0.0);
  double swc = group_sum<NMAX>(in ? lam * c : 0.0);
  double sigma2 = c0 - swc - mu;
  // The held-out value.
  const int64_t held = a.left_out_ids ? a.left_out_ids[t] : t;
  double zi = active ? a.sv[held] : 0.0;
  const bool ok = go && !sing && sigma2 > 0.0 && n > 0;
  double alpha = -2.0 * (zi - zhat), beta = 1.0 / sigma2;
  if (ok && r < n) {
    V[r] = lam;
    E[r] = vl;
  }
  wave_sync();
  double acc[BK_MAX_GRAD];
  for (int p = 0; p < BK_MAX_GRAD; ++p) acc[p] = 0.0;
  if (ok) {
    for (int q = r; q < npairs; q += NMAX) {
      int i, j;
      pair_of(q, i, j);
      double coef = -alpha * (E[i] * V[j] + E[j] * V[i]) + 2.0 * beta * V[i] * V[j];
      cov_grad64(gp, W[i] - W[j], W[NMAX + i] - W[NMAX + j], W[2 * NMAX + i] - W[2 * NMAX + j], coef, acc);
    }
    if (r < n) {
      double coefd = -alpha * vl * lam + beta * lam * lam;
      acc[0] += coefd;
      for (int s = 0; s < gp.S; ++s) acc[1 + s] += coefd;
      double gr = alpha * vl - 2.0 * beta * lam;
      cov_grad64(gp, a.tx[t] - W[r], a.ty[t] - W[NMAX + r], a.tz[t] - W[2 * NMAX + r], gr, acc);
    }
    if (r == 0) {
      acc[0] += beta;
      for (int s = 0; s < gp.S; ++s) acc[1 + s] += beta;
    }
  }
  group_sum_vec<NMAX>(acc, gp.P);
  if (active && r == 0) {
    // Column-major partials: column p holds entry p of every held-out sample.
    double* out = a.grad_part + t;
    const int64_t m = gp.m;
    if (ok) {
      out[0] = (zi - zhat) * (zi - zhat) + log(sigma2);
      out[m] = 1.0;
      for (int p = 0; p < gp.P; ++p) out[(2 + p) * m] = acc[p];
    } else {
      for (int p = 0; p < gp.P + 2; ++p) out[p * m] = 0.0;
    }
    a.value[t] = ok ? zhat : __longlong_as_double(0x7ff8000000000000LL);
    a.variance[t] = ok ? sigma2 : __longlong_as_double(0x7ff8000000000000LL);
    a.status[t] = ok ? BK_OK : (guard ? BK_UNSUPPORTED : BK_SINGULAR);
  }
}

template <int NMAX>
__global__ void __launch_bounds__(256) k_loo_grad(const SolveArgs a, const GradArgs gp) {
  const int64_t nl = list_len(a);
  UnitLoop q = unit_loop((nl + BK_WAVE / NMAX - 1) / (BK_WAVE / NMAX), a.work);
  while (q.unit < q.end) {
    // The next unit is taken before this one is solved, hiding the atomic.
    const int next = unit_next(q);
    loo_grad_blk<NMAX>(a, gp, nl, q.unit);
    wave_sync();
    q.unit = next;
  }
}

// ---------------------------------------------------------------------------
// Leave-one-out loss and gradient in the refined form: the forward solve of
// k_krige32 (FP32 Gauss-Jordan inverse, FP64 residuals) on three right-hand
// sides (c, 1, z), the Schur scalar, weights and the adjoint vector in FP64,
// and the adjoint accumulation over pairs in FP32 (the gradient is compared
// with finite differences at 1e-5; FP32 products carry 1e-7).

// shape(t) and its derivative in FP32 (model.rs:76-126, :181-311).
__device__ __forceinline__ void shape_d32(int model, float t, float& g, float& gd) {
#pragma clang fp contract(fast)
  switch (model) {
    case BK_SPHERICAL:
      if (t >= 1.0f) {
        g = 1.0f;
        gd = 0.0f;
      } else {
        g = t * (1.5f - 0.5f * t * t);
        gd = 1.5f - 1.5f * t * t;
      }
      return;
    case BK_EXPONENTIAL: {
      const float e = __expf(-3.0f * t);
      g = 1.0f - e;
      gd = 3.0f * e;
      return;
    }
    case BK_GAUSSIAN: {
      const float e = __expf(-3.0f * t * t);
      g = 1.0f - e;
      gd = 6.0f * t * e;
      return;
    }
    case BK_CUBIC: {
      const float t2 = t * t;
      if (t >= 1.0f) {
        g = 1.0f;
        gd = 0.0f;
      } else {
        g = t2 * (7.0f + t * (-8.75f + t2 * (3.5f - 0.75f * t2)));
        gd = t * (14.0f + t * (-26.25f + t2 * (17.5f * t - 5.25f * t2 * t)));
      }
      return;
    }
    case BK_PENTASPHERICAL: {
      const float t2 = t * t;
      if (t >= 1.0f) {
        g = 1.0f;
        gd = 0.0f;
      } else {
        g = t * (1.875f + t2 * (-1.25f + 0.375f * t2));
        gd = 1.875f + t2 * (-3.75f + 1.875f * t2);
      }
      return;
    }
  }
  g = gd = 0.0f;
}

// cov_grad64 in FP32: the gradient of C(p, q), q - p = d, weighted by coef,
// into a fixed layout with compile-time indices (registers, no scratch):
// [0] nugget, [1 + s] sill_s, [5 + 3 s + a] range a of structure s, [17 + b]
// angle b. The packed layout of the C ABI is formed at the output.
constexpr int G32_SILL = 1, G32_RANGE = 1 + BK_MAX_STRUCTURES, G32_ANGLE = 1 + 4 * BK_MAX_STRUCTURES;

__device__ __forceinline__ void cov_grad32(const GradArgs32& gp, float d0, float d1, float d2, float coef,
                                           float* acc) {
#pragma clang fp contract(fast)
  const int S = gp.S;
  if (d0 == 0.0f && d1 == 0.0f && d2 == 0.0f) {
    acc[0] += coef;
#pragma unroll
    for (int s = 0; s < BK_MAX_STRUCTURES; ++s)
      if (s < S) acc[G32_SILL + s] += coef;
    return;
  }
  const float r0 = gp.R[0] * d0 + gp.R[1] * d1 + gp.R[2] * d2;
  const float r1 = gp.R[3] * d0 + gp.R[4] * d1 + gp.R[5] * d2;
  const float r2 = gp.R[6] * d0 + gp.R[7] * d1 + gp.R[8] * d2;
  float q[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    q[b][0] = gp.dR[b][0] * d0 + gp.dR[b][1] * d1 + gp.dR[b][2] * d2;
    q[b][1] = gp.dR[b][3] * d0 + gp.dR[b][4] * d1 + gp.dR[b][5] * d2;
    q[b][2] = gp.dR[b][6] * d0 + gp.dR[b][7] * d1 + gp.dR[b][8] * d2;
  }
#pragma unroll
  for (int s = 0; s < BK_MAX_STRUCTURES; ++s) {
    if (s >= S) break;
    const float i0 = gp.inv_r[s][0], i1 = gp.inv_r[s][1], i2 = gp.inv_r[s][2];
    const float u0 = r0 * i0, u1 = r1 * i1, u2 = r2 * i2;
    const float tt = u0 * u0 + u1 * u1 + u2 * u2;
    const float rs = __builtin_amdgcn_rsqf(tt);
    const float t = tt * rs;
    float g, gd;
    shape_d32(gp.model[s], t, g, gd);
    acc[G32_SILL + s] += coef * (1.0f - g);
    const float k = coef * gp.sill[s] * gd * rs;
    acc[G32_RANGE + 3 * s + 0] += k * u0 * u0 * i0;
    acc[G32_RANGE + 3 * s + 1] += k * u1 * u1 * i1;
    acc[G32_RANGE + 3 * s + 2] += k * u2 * u2 * i2;
#pragma unroll
    for (int b = 0; b < 3; ++b)
      acc[G32_ANGLE + b] -= k * (u0 * q[b][0] * i0 + u1 * q[b][1] * i1 + u2 * q[b][2] * i2);
  }
}

// cov_grad32 with the nugget and sill terms in FP64 (coef in FP64, the
// shapes in FP32): those terms sum correlations with large weights of both
// signs where samples nearly coincide, and carry most of the cancellation.
__device__ __forceinline__ void cov_grad_mixed(const GradArgs32& gp, float d0, float d1, float d2, double coef,
                                               double* acc) {
#pragma clang fp contract(fast)
  const int S = gp.S;
  if (d0 == 0.0f && d1 == 0.0f && d2 == 0.0f) {
    acc[0] += coef;
#pragma unroll
    for (int s = 0; s < BK_MAX_STRUCTURES; ++s)
      if (s < S) acc[G32_SILL + s] += coef;
    return;
  }
  const float cf = (float)coef;
  const float r0 = gp.R[0] * d0 + gp.R[1] * d1 + gp.R[2] * d2;
  const float r1 = gp.R[3] * d0 + gp.R[4] * d1 + gp.R[5] * d2;
  const float r2 = gp.R[6] * d0 + gp.R[7] * d1 + gp.R[8] * d2;
  float q[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    q[b][0] = gp.dR[b][0] * d0 + gp.dR[b][1] * d1 + gp.dR[b][2] * d2;
    q[b][1] = gp.dR[b][3] * d0 + gp.dR[b][4] * d1 + gp.dR[b][5] * d2;
    q[b][2] = gp.dR[b][6] * d0 + gp.dR[b][7] * d1 + gp.dR[b][8] * d2;
  }
#pragma unroll
  for (int s = 0; s < BK_MAX_STRUCTURES; ++s) {
    if (s >= S) break;
    const float i0 = gp.inv_r[s][0], i1 = gp.inv_r[s][1], i2 = gp.inv_r[s][2];
    const float u0 = r0 * i0, u1 = r1 * i1, u2 = r2 * i2;
    const float tt = u0 * u0 + u1 * u1 + u2 * u2;
    const float rs = __builtin_amdgcn_rsqf(tt);
    const float t = tt * rs;
    float g, gd;
    shape_d32(gp.model[s], t, g, gd);
    acc[G32_SILL + s] += coef * (double)(1.0f - g);
    const float k = cf * gp.sill[s] * gd * rs;
    acc[G32_RANGE + 3 * s + 0] += k * u0 * u0 * i0;
    acc[G32_RANGE + 3 * s + 1] += k * u1 * u1 * i1;
    acc[G32_RANGE + 3 * s + 2] += k * u2 * u2 * i2;
#pragma unroll
    for (int b = 0; b < 3; ++b)
      acc[G32_ANGLE + b] -= k * (u0 * q[b][0] * i0 + u1 * q[b][1] * i1 + u2 * q[b][2] * i2);
  }
}

// Position in the C ABI's packed layout of fixed-layout index p (S structures).
__device__ __forceinline__ int g32_packed(int p, int S) {
  if (p < G32_RANGE) return p;  // nugget, sills
  if (p < G32_ANGLE) return 1 + S + (p - G32_RANGE);
  return 1 + 4 * S + (p - G32_ANGLE);
}

__device__ __forceinline__ bool g32_used(int p, int S) {
  if (p == 0) return true;
  if (p < G32_RANGE) return p - G32_SILL < S;
  if (p < G32_ANGLE) return (p - G32_RANGE) / 3 < S;
  return true;
}

// LDS doubles per system of the refined gradient kernel: records, iterates
// (NMAX x 4 doubles), pivot rows (2 NMAX floats), product vectors (NMAX
// float4), the packed triangle and, for NMAX = 64, the lanes' FP64 rows.
__host__ __device__ constexpr int loo32_doubles(int nmax, int nf) {
  return nmax * rec_doubles(nf) + 4 * nmax + nmax + 2 * nmax + (nmax * (nmax + 1) / 2 + 1) / 2 * 2 +
         (nmax > 32 ? nmax * nmax : 0) + 8;
}

template <int NMAX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NMAX == 64 ? 1 : 2)))
k_loo_grad32(const SolveArgs a, const GradArgs gp, const GradArgs32 g32) {
  constexpr int G = BK_WAVE / NMAX;
  constexpr bool ROWREG = NMAX <= 32;
  extern __shared__ double dyn[];
  const int nf = a.v.nf;
  const int SJ = rec_doubles(nf);
  const int per = loo32_doubles(NMAX, nf);
  // One system group per wavefront of the grid: the list length is known on
  // the host here, and this kernel's registers fit better without a
  // persistent loop.
  const int wave = threadIdx.x / BK_WAVE, lane = threadIdx.x % BK_WAVE;
  const int g = lane / NMAX, r = lane % NMAX;
  double* S = dyn + (wave * G + g) * per;
  double* PJ = S;
  double* X4 = PJ + NMAX * SJ;
  float* P2 = (float*)(X4 + 4 * NMAX);
  float* B4 = P2 + 2 * NMAX;
  double* T = (double*)(B4 + 4 * NMAX);
  double* RowL = T + (NMAX * (NMAX + 1) / 2 + 1) / 2 * 2;
  const int64_t sys = ((int64_t)blockIdx.x * (blockDim.x / BK_WAVE) + wave) * G + g;
  const bool active = sys < a.n_list;
  const int t = active ? a.list[sys] : 0;
  const int n = active ? a.ncount[t] : 0;
  double z, e, c;
  double crow[ROWREG ? NMAX : 1];
  float x[NMAX];
  bool bad;
  float dmin;
  const bool guard =
      refined_front<NMAX>(a, t, n, r, g, lane, active, false, PJ, SJ, P2, T, RowL, z, e, c, x, crow, bad, dmin);
  const bool go = active && !guard;
  const bool live = go && r < n;
  const double c0 = a.v.c0;
  const double ic0 = 1.0 / c0;
  // y = X v for three right-hand sides.
  auto apply3 = [&](float va, float vb, float vz, float& ya, float& yb, float& yz) {
    reinterpret_cast<float4*>(B4)[r] = make_float4(va, vb, vz, 0.0f);
    wave_sync();
    float sa = 0.0f, sb = 0.0f, sz = 0.0f;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) {
      const float4 q = reinterpret_cast<const float4*>(B4)[k];
      sa = fmaf(x[k], q.x, sa);
      sb = fmaf(x[k], q.y, sb);
      sz = fmaf(x[k], q.z, sz);
      if (k % 8 == 7) asm volatile("" ::: "memory");
    }
    ya = sa;
    yb = sb;
    yz = sz;
    wave_sync();
  };
  const double ba = live ? c : 0.0, bb = live ? 1.0 : 0.0, bz = live ? z : 0.0;
  float fa, fb, fz;
  apply3((float)(ba * ic0), (float)(bb * ic0), (float)(bz * ic0), fa, fb, fz);
  double xa = fa, xb = fb, xz = fz;
  const float tolf = (float)a.tol;
  const float ma = group_maxf<NMAX>(fabsf(fa)), mb = group_maxf<NMAX>(fabsf(fb)), mz = group_maxf<NMAX>(fabsf(fz));
  bool conv = false;
  for (int it = 0; it < a.max_iter && !bad; ++it) {
    double2* X = reinterpret_cast<double2*>(X4);
    X[2 * r] = make_double2(xa, xb);
    X[2 * r + 1] = make_double2(xz, 0.0);
    wave_sync();
    double ra = ba, rb = bb, rz = bz;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) {
      const double2 q0 = X[2 * k], q1 = X[2 * k + 1];
      double ck;
      if constexpr (ROWREG)
        ck = crow[k];
      else
        ck = RowL[k * BK_WAVE + lane];
      ra = fma(-ck, q0.x, ra);
      rb = fma(-ck, q0.y, rb);
      rz = fma(-ck, q1.x, rz);
      if (k % 4 == 3) asm volatile("" ::: "memory");
    }
    wave_sync();
    float da, db, dz;
    apply3((float)(ra * ic0), (float)(rb * ic0), (float)(rz * ic0), da, db, dz);
    xa += (double)da;
    xb += (double)db;
    xz += (double)dz;
    // The gradient's forward solve keeps the plain test, delta_k <= tol (its
    // registers are at the kernel's limit).
    if (group_all<NMAX>(fabsf(da) <= tolf * ma && fabsf(db) <= tolf * mb && fabsf(dz) <= tolf * mz, g)) {
      conv = true;
      break;
    }
  }
  const bool okf = go && !bad && conv;
  // A system the FP32 elimination cannot serve (a small pivot, or no
  // convergence) goes to the FP64 gradient kernel, which writes its row.
  const bool redo = go && !okf && a.redo;
  if (redo && r == 0) a.redo[atomicAdd(a.redo_count, 1)] = t;
  // Forward quantities in FP64 (kernel.rs:86-97), adjoint vector v = K^-1 [z; 0].
  const bool in = okf && r < n;
  const double sa = group_sum<NMAX>(in ? xa : 0.0);
  const double sb = group_sum<NMAX>(in ? xb : 0.0);
  const double sz = group_sum<NMAX>(in ? xz : 0.0);
  const double mu = (sa - 1.0) / sb;
  const double vmu = sz / sb;
  const double lam = xa - mu * xb;
  const double vl = xz - vmu * xb;
  const double zhat = group_sum<NMAX>(in ? lam * z : 0.0);
  const double swc = group_sum<NMAX>(in ? lam * c : 0.0);
  const double sigma2 = c0 - swc - mu;
  const int64_t held = a.left_out_ids ? a.left_out_ids[t] : t;
  const double zi = active ? a.sv[held] : 0.0;
  const bool ok = okf && sigma2 > 0.0 && n > 0;
  const double alpha = -2.0 * (zi - zhat), beta = 1.0 / sigma2;
  double2* VE = reinterpret_cast<double2*>(X4);
  wave_sync();
  if (r < NMAX) VE[r] = make_double2(ok && r < n ? lam : 0.0, ok && r < n ? vl : 0.0);
  wave_sync();
  float acc[BK_MAX_GRAD];
#pragma unroll
  for (int p = 0; p < BK_MAX_GRAD; ++p) acc[p] = 0.0f;
  if (ok && r < n) {
    const float al = (float)alpha, be = (float)beta;
    const float Vr = (float)lam, Er = (float)vl;
    const double* me = PJ + r * SJ;
    const int half = (n - 1) / 2;
    const bool even = (n & 1) == 0;
    const int dmax = (even && r < n / 2) ? n / 2 : half;
    for (int d = 1; d <= dmax; ++d) {
      int j = r + d;
      j = j >= n ? j - n : j;
      const double2 vj = VE[j];
      const float Vj = (float)vj.x, Ej = (float)vj.y;
      const float coef = -al * (Er * Vj + Ej * Vr) + 2.0f * be * Vr * Vj;
      const double* pj = PJ + j * SJ;
      cov_grad32(g32, (float)(me[0] - pj[0]), (float)(me[1] - pj[1]), (float)(me[2] - pj[2]), coef, acc);
    }
    const float coefd = (float)(-alpha * vl * lam + beta * lam * lam);
    acc[0] += coefd;
#pragma unroll
    for (int s = 0; s < BK_MAX_STRUCTURES; ++s)
      if (s < g32.S) acc[G32_SILL + s] += coefd;
    const float gr = (float)(alpha * vl - 2.0 * beta * lam);
    cov_grad32(g32, (float)(a.tx[t] - me[0]), (float)(a.ty[t] - me[1]), (float)(a.tz[t] - me[2]), gr, acc);
  }
  if (ok && r == 0) {
    acc[0] += (float)beta;
#pragma unroll
    for (int s = 0; s < BK_MAX_STRUCTURES; ++s)
      if (s < g32.S) acc[G32_SILL + s] += (float)beta;
  }
#pragma unroll
  for (int p = 0; p < BK_MAX_GRAD; ++p)
    if (g32_used(p, g32.S)) acc[p] = group_sumf<NMAX>(acc[p]);
  if (active && r == 0 && !redo) {
    double* out = a.grad_part + t;
    const int64_t m = gp.m;
    if (ok) {
      out[0] = (zi - zhat) * (zi - zhat) + log(sigma2);
      out[m] = 1.0;
#pragma unroll
      for (int p = 0; p < BK_MAX_GRAD; ++p)
        if (g32_used(p, g32.S)) out[(2 + g32_packed(p, g32.S)) * m] = (double)acc[p];
    } else {
      for (int p = 0; p < g32.P + 2; ++p) out[p * m] = 0.0;
    }
    a.value[t] = ok ? zhat : __longlong_as_double(0x7ff8000000000000LL);
    a.variance[t] = ok ? sigma2 : __longlong_as_double(0x7ff8000000000000LL);
    a.status[t] = ok ? BK_OK : (guard ? BK_UNSUPPORTED : BK_SINGULAR);
  }
}

#if BK_PART != 2
// ---------------------------------------------------------------------------
// The refined leave-one-out gradient for systems of at most 32 samples, two
// per wavefront, on the matrix cores: the front of k_krige32m (circulant
// pairs, MFMA Gauss-Jordan) with three right-hand sides, c, the ones and the
// neighbours' values z, refined in FP64 with the same stopping rule; then
// the forward quantities and the adjoint v = K^-1 [z; 0] in FP64 and the
// pair gradient of k_loo_grad32 (cov_grad32 over the pairs (r, (r + d) mod
// 32), d = 1 .. 16, the pairs of d = 16 taken once). Systems with a near pair,
// a small pivot or no convergence go to the FP64 gradient kernel through the
// redo list (it applies the exact coincidence test, so coincident systems get
// its status).
__device__ __forceinline__ void loo32m_blk(const SolveArgs& a, const GradArgs32& g32, int64_t gm, int64_t nl,
                                           int unit, const uint32_t (&ipk)[8]) {
  extern __shared__ double dyn[];
  const int nf = a.v.nf;
  const int SQ = rec_q(nf);
  const int tid = opaque_tid();
  const int wave = tid / BK_WAVE, lane = tid % BK_WAVE;
  const int g = lane >> 5, r = lane & 31;
  double* S = dyn + wave * sysl_doubles(nf);
  double* Q = S + qrecl_doubles(nf);
  double* Qg = Q + g * kQSys;
  const int64_t sys = (int64_t)unit * 2 + g;
  const bool active = sys < nl;
  const int t = active ? a.list[sys] : 0;
  const int n = active ? a.ncount[t] : 0;
  const bool mine = active && r < n;
  const double c0 = a.v.c0;
  const double ic0 = a.v.inv_c0;
  double z, e, c;
  circ_record(a, t, mine, r, lane, S + (g * 32 + r) * SQ, z, e, c);
  wave_sync();
  const bool near = group_any<32>(pairs_q32(a.v, S + g * 32 * SQ, SQ, r, Qg), g);
  const bool go = active && !near;
  const bool go0 = __builtin_amdgcn_readlane(go ? 1 : 0, 0) != 0, go1 = __builtin_amdgcn_readlane(go ? 1 : 0, 32) != 0;
  Qg[r] = mine ? c0 + e : c0;
  if (__builtin_amdgcn_readlane(n, 0) < 32 || __builtin_amdgcn_readlane(n, 32) < 32) {
#pragma unroll
    for (int dd = 1; dd <= 16; ++dd)
      if (!(r < n && ((r + dd) & 31) < n)) Qg[dd * kQS + r] = 0.0;
  }
  wave_sync();
  bk_floatx32 X;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double* e0 = &lds_at(unpack_off(ipk, k));
#pragma unroll
    for (int b = 0; b < 2; ++b) X[16 * b + k] = (float)(e0[b * kQSys] * ic0);
  }
  int dmin = 0x7f7fffff;
  gj_pivots32<0>(X, dmin);
  const int dmin0 = __builtin_amdgcn_readlane(dmin, 0), dmin1 = __builtin_amdgcn_readlane(dmin, 32);
  const bool bad0 = !(dmin0 > __float_as_int(1e-5f)), bad1 = !(dmin1 > __float_as_int(1e-5f));
  // Solves for the three right-hand sides (see krige32m_blk): iterates
  // double4 [b][j], product inputs float4 [b][j] (the records are no longer
  // read).
  double* X3 = S;
  float* V3 = (float*)(S + 4 * 64);
  auto apply = [&](float va, float vb, float vz, float (&ya)[2], float (&yb)[2], float (&yz)[2]) {
    reinterpret_cast<float4*>(V3)[g * 32 + r] = make_float4(va, vb, vz, 0.0f);
    wave_sync();
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float sa = 0.0f, sb = 0.0f, sz = 0.0f;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const float4 w = reinterpret_cast<const float4*>(V3)[32 * b + 8 * q + 4 * g + p];
          const float xv = X[16 * b + 4 * q + p];
          sa = fmaf(xv, w.x, sa);
          sb = fmaf(xv, w.y, sb);
          sz = fmaf(xv, w.z, sz);
        }
      ya[b] = halves_sum(sa);
      yb[b] = halves_sum(sb);
      yz[b] = halves_sum(sz);
    }
    wave_sync();
  };
  const bool live = go && r < n;
  const double ba = live ? c : 0.0, bb = live ? 1.0 : 0.0, bz = live ? z : 0.0;
  const double ba_o = other_half(ba, g), bb_o = other_half(bb, g), bz_o = other_half(bz, g);
  const double rha[2] = {g ? ba_o : ba, g ? ba : ba_o}, rhb[2] = {g ? bb_o : bb, g ? bb : bb_o},
               rhz[2] = {g ? bz_o : bz, g ? bz : bz_o};
  float fa[2], fb[2], fz[2];
  apply((float)(ba * ic0), (float)(bb * ic0), (float)(bz * ic0), fa, fb, fz);
  double xa[2] = {fa[0], fa[1]}, xb[2] = {fb[0], fb[1]}, xz[2] = {fz[0], fz[1]};
  const float tolf = (float)a.tol;
  float ma[2], mb[2], mz[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    ma[b] = group_maxf<32>(fabsf(fa[b]));
    mb[b] = group_maxf<32>(fabsf(fb[b]));
    mz[b] = group_maxf<32>(fabsf(fz[b]));
  }
  bool conv0 = false, conv1 = false;
  for (int it = 0; it < a.max_iter; ++it) {
    const bool more0 = go0 && !bad0 && !conv0, more1 = go1 && !bad1 && !conv1;
    if (!more0 && !more1) break;
    reinterpret_cast<double4*>(X3)[g * 32 + r] =
        make_double4(g ? xa[1] : xa[0], g ? xb[1] : xb[0], g ? xz[1] : xz[0], 0.0);
    wave_sync();
    double ra[2], rb[2], rz[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      double sa = 0.0, sb = 0.0, sz = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const double4 xv = reinterpret_cast<const double4*>(X3)[32 * b + 8 * q + 4 * g + p];
          const double cv = (&lds_at(unpack_off(ipk, 4 * q + p)))[b * kQSys];
          sa = fma(cv, xv.x, sa);
          sb = fma(cv, xv.y, sb);
          sz = fma(cv, xv.z, sz);
        }
      ra[b] = rha[b] - halves_sum(sa);
      rb[b] = rhb[b] - halves_sum(sb);
      rz[b] = rhz[b] - halves_sum(sz);
    }
    wave_sync();
    float da[2], db[2], dz[2];
    apply((float)((g ? ra[1] : ra[0]) * ic0), (float)((g ? rb[1] : rb[0]) * ic0), (float)((g ? rz[1] : rz[0]) * ic0),
          da, db, dz);
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      xa[b] += (double)da[b];
      xb[b] += (double)db[b];
      xz[b] += (double)dz[b];
    }
    bool small0, small1;
    if (it == 0) {
      const float delta = group_maxf<32>(fmaxf(fmaxf(rel_corr(g ? da[1] : da[0], g ? ma[1] : ma[0]),
                                                     rel_corr(g ? db[1] : db[0], g ? mb[1] : mb[0])),
                                               rel_corr(g ? dz[1] : dz[0], g ? mz[1] : mz[0])));
      const bool done = first_sweep_converged(delta, 32.0f, __int_as_float(g ? dmin1 : dmin0), tolf);
      const bool nan0 = __ballot(__builtin_isnan(da[0]) || __builtin_isnan(db[0]) || __builtin_isnan(dz[0])) != 0;
      const bool nan1 = __ballot(__builtin_isnan(da[1]) || __builtin_isnan(db[1]) || __builtin_isnan(dz[1])) != 0;
      small0 = !nan0 && __builtin_amdgcn_readlane(done ? 1 : 0, 0) != 0;
      small1 = !nan1 && __builtin_amdgcn_readlane(done ? 1 : 0, 32) != 0;
    } else {
      small0 = __ballot(!(fabsf(da[0]) <= tolf * ma[0] && fabsf(db[0]) <= tolf * mb[0] &&
                          fabsf(dz[0]) <= tolf * mz[0])) == 0;
      small1 = __ballot(!(fabsf(da[1]) <= tolf * ma[1] && fabsf(db[1]) <= tolf * mb[1] &&
                          fabsf(dz[1]) <= tolf * mz[1])) == 0;
    }
    if (more0) conv0 = small0;
    if (more1) conv1 = small1;
  }
  const bool bad = g ? bad1 : bad0, conv = g ? conv1 : conv0;
  const bool okf = go && !bad && conv;
  const bool redo = active && !okf && a.redo;
  if (redo && r == 0) a.redo[atomicAdd(a.redo_count, 1)] = t;
  // Forward quantities in FP64 (kernel.rs:86-97), adjoint vector v = K^-1 [z; 0].
  const double xag = g ? xa[1] : xa[0], xbg = g ? xb[1] : xb[0], xzg = g ? xz[1] : xz[0];
  const bool in = okf && r < n;
  const double sa = group_sum<32>(in ? xag : 0.0);
  const double sb = group_sum<32>(in ? xbg : 0.0);
  const double sz = group_sum<32>(in ? xzg : 0.0);
  const double mu = (sa - 1.0) / sb;
  const double vmu = sz / sb;
  const double lam = xag - mu * xbg;
  const double vl = xzg - vmu * xbg;
  const double zhat = group_sum<32>(in ? lam * z : 0.0);
  const double swc = group_sum<32>(in ? lam * c : 0.0);
  const double sigma2 = c0 - swc - mu;
  const int64_t held = a.left_out_ids ? a.left_out_ids[t] : t;
  const double zi = active ? a.sv[held] : 0.0;
  const bool ok = okf && sigma2 > 0.0 && n > 0;
  const double alpha = -2.0 * (zi - zhat), beta = 1.0 / sigma2;
  // The gradient over the pairs, from each sample's world position and its
  // (lambda, v) in LDS (the iterates are no longer read).
  double wx = 0.0, wy = 0.0, wz = 0.0;
  if (mine) {
    const int id = a.nidx[(int64_t)t * a.stride + r];
    wx = a.sx[id];
    wy = a.sy[id];
    wz = a.sz[id];
  }
  // Position and (lambda, v) of each sample: 5 doubles (an odd stride).
  double* GW = S + g * 32 * 5;
  wave_sync();
  GW[5 * r] = wx;
  GW[5 * r + 1] = wy;
  GW[5 * r + 2] = wz;
  GW[5 * r + 3] = ok && r < n ? lam : 0.0;
  GW[5 * r + 4] = ok && r < n ? vl : 0.0;
  wave_sync();
  // Per-lane sums in FP64: near-coincident samples make the terms large and
  // of both signs (FP32 increments, FP64 accumulation).
  double acc[BK_MAX_GRAD];
#pragma unroll
  for (int p = 0; p < BK_MAX_GRAD; ++p) acc[p] = 0.0;
  if (ok && r < n) {
    for (int d = 1; d <= 16; ++d) {
      const int j = (r + d) & 31;
      if (j >= n || (d == 16 && r >= 16)) continue;
      const double* pj = GW + 5 * j;
      const double coef = -alpha * (vl * pj[3] + pj[4] * lam) + 2.0 * beta * lam * pj[3];
      cov_grad_mixed(g32, (float)(wx - pj[0]), (float)(wy - pj[1]), (float)(wz - pj[2]), coef, acc);
    }
    const double coefd = -alpha * vl * lam + beta * lam * lam;
    acc[0] += coefd;
#pragma unroll
    for (int s = 0; s < BK_MAX_STRUCTURES; ++s)
      if (s < g32.S) acc[G32_SILL + s] += coefd;
    const double gr = alpha * vl - 2.0 * beta * lam;
    cov_grad_mixed(g32, (float)(a.tx[t] - wx), (float)(a.ty[t] - wy), (float)(a.tz[t] - wz), gr, acc);
  }
  if (ok && r == 0) {
    acc[0] += beta;
#pragma unroll
    for (int s = 0; s < BK_MAX_STRUCTURES; ++s)
      if (s < g32.S) acc[G32_SILL + s] += beta;
  }
  double tot[BK_MAX_GRAD];
#pragma unroll
  for (int p = 0; p < BK_MAX_GRAD; ++p) tot[p] = g32_used(p, g32.S) ? group_sum<32>(acc[p]) : 0.0;
  if (active && r == 0 && !redo) {
    double* out = a.grad_part + t;
    const int64_t m = gm;
    if (ok) {
      out[0] = (zi - zhat) * (zi - zhat) + log(sigma2);
      out[m] = 1.0;
#pragma unroll
      for (int p = 0; p < BK_MAX_GRAD; ++p)
        if (g32_used(p, g32.S)) out[(2 + g32_packed(p, g32.S)) * m] = tot[p];
    } else {
      for (int p = 0; p < g32.P + 2; ++p) out[p * m] = 0.0;
    }
    a.value[t] = ok ? zhat : __longlong_as_double(0x7ff8000000000000LL);
    a.variance[t] = ok ? sigma2 : __longlong_as_double(0x7ff8000000000000LL);
    a.status[t] = ok ? BK_OK : BK_SINGULAR;
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_loo_grad32m(const SolveArgs a, const GradArgs gp, const GradArgs32 g32) {
  const int64_t nl = list_len(a);
  UnitLoop q = unit_loop((nl + 1) / 2, a.work);
  uint32_t ipk[8];
  {
    const int lane = threadIdx.x % BK_WAVE, g = lane >> 5, r = lane & 31;
    const uint32_t qb = (uint32_t)((threadIdx.x / BK_WAVE) * sysl_doubles(a.v.nf) + qrecl_doubles(a.v.nf)) * 8u;
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
      const uint32_t lo = qb + 8u * (uint32_t)q_entry(8 * (k / 4) + 4 * g + k % 4, r);
      const uint32_t hi = qb + 8u * (uint32_t)q_entry(8 * ((k + 1) / 4) + 4 * g + (k + 1) % 4, r);
      ipk[k / 2] = lo | (hi << 16);
    }
  }
  while (q.unit < q.end) {
    const int next = unit_next(q);
    loo32m_blk(kernarg_first<SolveArgs>(), g32, gp.m, nl, q.unit, ipk);
    wave_sync();
    q.unit = next;
  }
}

#endif

// Deterministic column sums of a column-major m x cols array: block
// (b, col) sums rows [b*chunk, (b+1)*chunk) of its column (coalesced, fixed
// thread order, LDS tree), then one block per column adds the partials in
// block order. The order depends on m only.
constexpr int RED_T = 256;
constexpr int64_t RED_CHUNK = 4096;

__device__ double block_tree(double s, double* red) {
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = RED_T / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

#if BK_PART != 1
__global__ void __launch_bounds__(RED_T) k_reduce_part(int64_t m, const double* part, double* work) {
  __shared__ double red[RED_T];
  const int64_t nb = gridDim.x, b = blockIdx.x, col = blockIdx.y;
  const double* p = part + col * m;
  const int64_t lo = b * RED_CHUNK, hi = lo + RED_CHUNK < m ? lo + RED_CHUNK : m;
  double s = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += RED_T) s += p[i];
  s = block_tree(s, red);
  if (threadIdx.x == 0) work[col * nb + b] = s;
}

__global__ void __launch_bounds__(RED_T) k_reduce_final(int64_t nb, const double* work, double* out) {
  __shared__ double red[RED_T];
  const int col = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < nb; i += RED_T) s += work[col * nb + i];
  s = block_tree(s, red);
  if (threadIdx.x == 0) out[col] = s;
}
#endif

// ---------------------------------------------------------------------------
// Systems with coincident samples (coincident.rs:27-112): merged on lane 0,
// solved in FP64 by LU on the bordered system of the merged samples (the
// merged system is small and rare), weights spread back to the members.
// Also takes systems beyond the Cholesky kernels (n > 64, power models).
constexpr int LU_MAX = BK_MAX_LU + 1;
// Systems above BK_MAX_LU samples (up to the search's BK_MAX_SEARCH) take the
// same LU with its matrix in a global-memory workspace.
constexpr int LU_BIG = BK_MAX_SEARCH + 1;

// Doubles and ints of one LU system's workspace with rows of CAP entries.
__host__ __device__ constexpr int64_t lu_doubles(int cap) { return (int64_t)cap * cap + 8 * (int64_t)cap; }
__host__ __device__ constexpr int64_t lu_ints(int cap) { return 3 * (int64_t)cap + 4; }

template <int CAP>
__device__ void lu_system(const SolveArgs& a, int t, double* A, double* b, double* ws, int* cl, int lane) {
  // Runs on one wavefront. A: CAP * CAP, b: CAP, ws: 7 CAP doubles, cl: lu_ints(CAP).
  const int n0 = a.ncount[t];
  double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
  double* X = ws;                 // merged samples
  double* Y = X + CAP;
  double* Z = Y + CAP;
  double* Vv = Z + CAP;
  double* Ev = Vv + CAP;
  double* Vo = Ev + CAP;          // original values
  double* rhs = Vo + CAP;
  int* members = cl + CAP;        // cluster sizes
  int* lead = members + CAP;      // first member of each cluster
  __shared__ int sh_n, sh_fail;
  const double c0 = a.v.c0;
  if (lane == 0) {
    // Clusters::find + merge (coincident.rs:27-82), serial.
    int nc = 0;
    for (int i = 0; i < n0; ++i) {
      int id = a.nidx[(int64_t)t * a.stride + i];
      double x = a.sx[id], y = a.sy[id], z = a.sz[id];
      Vo[i] = a.sv[id];
      int c = -1;
      for (int k = 0; k < nc; ++k) {
        int l = lead[k];
        double u = a.sx[l] - x, v = a.sy[l] - y, w = a.sz[l] - z;
        if ((u * u + v * v) + w * w <= EPS2) {
          c = k;
          break;
        }
      }
      if (c < 0) {
        c = nc++;
        lead[c] = id;
        members[c] = 0;
      }
      cl[i] = c;
      members[c] += 1;
    }
    for (int k = 0; k < nc; ++k) {
      int l = lead[k];
      X[k] = a.sx[l];
      Y[k] = a.sy[l];
      Z[k] = a.sz[l];
    }
    for (int k = 0; k < nc; ++k) {
      double var = 0.0, sum = -0.0;
      int first = -1;
      for (int i = 0; i < n0; ++i)
        if (cl[i] == k) {
          if (first < 0) first = i;
          int id = a.nidx[(int64_t)t * a.stride + i];
          sum += Vo[i];
          var += a.serr ? a.serr[id] : 0.0;
          for (int j = 0; j < n0; ++j)
            if (cl[j] == k) {
              int jd = a.nidx[(int64_t)t * a.stride + j];
              var += cov_points64(a.v, a.sx[id], a.sy[id], a.sz[id], a.sx[jd], a.sy[jd], a.sz[jd]);
            }
        }
      double kk = (double)members[k];
      int fid = a.nidx[(int64_t)t * a.stride + first];
      if (members[k] == 1) {
        Vv[k] = a.sv[fid];
        Ev[k] = a.serr ? a.serr[fid] : 0.0;
      } else {
        Vv[k] = sum / kk;
        Ev[k] = var / (kk * kk) - c0;
      }
    }
    sh_n = nc;
  }
  wave_sync();
  const int n = sh_n;
  const bool ordinary = a.kind != BK_SIMPLE;
  const int dim = ordinary ? n + 1 : n;
  // Bordered system (krige.rs:151-171, block.rs:140-165) with cov_points.
  for (int q = lane; q < dim * dim; q += BK_WAVE) {
    int i = q / dim, j = q % dim;
    double v;
    if (i < n && j < n) {
      v = cov_points64(a.v, X[i], Y[i], Z[i], X[j], Y[j], Z[j]);
      if (i == j) v += Ev[i];
    } else if (i == n && j == n) {
      v = 0.0;
    } else {
      v = 1.0;
    }
    A[i * CAP + j] = v;
  }
  for (int i = lane; i < dim; i += BK_WAVE) {
    double v = 1.0;
    if (i < n) {
      if (a.block) {
        double acc = -0.0;
        for (int p = 0; p < a.npts; ++p) {
          double qx = t0 + a.offs[3 * p], qy = t1 + a.offs[3 * p + 1], qz = t2 + a.offs[3 * p + 2];
          acc += a.npts == 1 ? cov_points64(a.v, X[i], Y[i], Z[i], qx, qy, qz)
                             : block_cov_points64(a.v, X[i], Y[i], Z[i], qx, qy, qz);
        }
        v = acc / (double)a.npts;
      } else {
        v = cov_points64(a.v, X[i], Y[i], Z[i], t0, t1, t2);
      }
      rhs[i] = v;
    }
    b[i] = v;
  }
  wave_sync();
  // LU with partial pivoting, reciprocal-scaled columns (nalgebra LU::new).
  if (lane == 0) sh_fail = 0;
  wave_sync();
  for (int i = 0; i < dim; ++i) {
    if (lane == 0) {
      int piv = i;
      double best = fabs(A[i * CAP + i]);
      for (int q = i + 1; q < dim; ++q) {
        double v = fabs(A[q * CAP + i]);
        if (v > best) {
          best = v;
          piv = q;
        }
      }
      cl[2 * CAP] = piv;
    }
    wave_sync();
    int piv = cl[2 * CAP];
    double diag = A[piv * CAP + i];
    if (diag == 0.0) {
      wave_sync();
      continue;
    }
    if (piv != i) {
      for (int k = lane; k < dim; k += BK_WAVE) {
        double u = A[i * CAP + k];
        A[i * CAP + k] = A[piv * CAP + k];
        A[piv * CAP + k] = u;
      }
      if (lane == 0) {
        double u = b[i];
        b[i] = b[piv];
        b[piv] = u;
      }
    }
    wave_sync();
    double inv = 1.0 / diag;
    for (int q = i + 1 + lane; q < dim; q += BK_WAVE) A[q * CAP + i] *= inv;
    wave_sync();
    for (int q = lane; q < (dim - i - 1) * (dim - i - 1); q += BK_WAVE) {
      int rr = i + 1 + q / (dim - i - 1), k = i + 1 + q % (dim - i - 1);
      A[rr * CAP + k] = A[rr * CAP + k] + (-A[i * CAP + k]) * A[rr * CAP + i];
    }
    wave_sync();
  }
  if (lane == 0) {
    for (int i = 0; i < dim; ++i) {
      double co = b[i];
      for (int q = i + 1; q < dim; ++q) b[q] = b[q] + (-co) * A[q * CAP + i];
    }
    for (int i = dim - 1; i >= 0; --i) {
      double d = A[i * CAP + i];
      if (d == 0.0) {
        sh_fail = 1;
        break;
      }
      b[i] = b[i] / d;
      double co = b[i];
      for (int q = 0; q < i; ++q) b[q] = b[q] + (-co) * A[q * CAP + i];
    }
    double mu = ordinary ? b[n] : 0.0;
    double sup = a.block ? a.cbb : c0;
    if (sh_fail) {
      a.value[t] = __longlong_as_double(0x7ff8000000000000LL);
      a.variance[t] = __longlong_as_double(0x7ff8000000000000LL);
      if (a.lagrange) a.lagrange[t] = __longlong_as_double(0x7ff8000000000000LL);
      if (a.support) a.support[t] = __longlong_as_double(0x7ff8000000000000LL);
      if (a.n_used) a.n_used[t] = 0;
      a.status[t] = BK_SINGULAR;
    } else {
      double swc = -0.0;
      for (int i = 0; i < n; ++i) swc += b[i] * rhs[i];
      // Spread (coincident.rs:102-112): each member gets its cluster's share,
      // the value is recomputed from the members.
      double sum = -0.0;
      for (int i = 0; i < n0; ++i) {
        double wi = b[cl[i]] / (double)members[cl[i]];
        if (a.weights) a.weights[(int64_t)t * a.stride + i] = wi;
        double zz = Vo[i];
        double f = a.kind == BK_SIMPLE ? zz - a.param : (a.kind == BK_INDICATOR ? (zz <= a.param ? 1.0 : 0.0) : zz);
        sum += wi * f;
      }
      double val = sum;
      if (a.kind == BK_SIMPLE) val = a.param + sum;
      if (a.kind == BK_INDICATOR) val = fmin(fmax(sum, 0.0), 1.0);
      a.value[t] = val;
      a.variance[t] = fmax(sup - swc - mu, 0.0);
      if (a.lagrange) a.lagrange[t] = mu;
      if (a.support) a.support[t] = sup;
      if (a.n_used) a.n_used[t] = n0;
      a.status[t] = BK_OK;
    }
  }
  wave_sync();
}

// One wavefront per system of a.guard[0 .. *a.guard_count), persistent: the
// grid strides over the list.
#if BK_PART != 2
// ---------------------------------------------------------------------------
// Calibrated search, kriging_algebra.rs:338-415 (prefix_quality and
// calibrated_count), in the arithmetic of the reference, which delegates the
// factorisation and the triangular solves to nalgebra 0.33: Cholesky::new
// updates column j by the columns k < j in order (y <- (-L[j][k]) L[.][k] + y)
// and divides it by the square root of its diagonal, failing (every prefix
// None) unless each pivot is positive; solve_lower_triangular divides b[i] by
// L[i][i] and updates the rows below by -b[i] L[.][i]. The matrix and the
// right-hand side are the world-lag cov_points (plus the error variance on the
// diagonal), or the block covariances for block support. Lane r holds row r;
// the running sums over the leading parts are taken in order, as the
// reference's fold does. The kernel only trims count[t]: the solve kernels then
// krige the leading samples as krige() on samples[..n] does.
__host__ __device__ constexpr int cal_doubles(int nmax) { return 3 * nmax + (nmax * (nmax + 1) / 2 + 1) / 2 * 2; }

template <int NMAX>
__global__ void __launch_bounds__(64) k_calibrate(const SolveArgs a, const CalibArgs k) {
  constexpr int G = BK_WAVE / NMAX;
  extern __shared__ double dyn[];
  const int lane = threadIdx.x, g = lane / NMAX, r = lane % NMAX;
  const int64_t t = (int64_t)blockIdx.x * G + g;
  const bool active = t < k.m && k.sstatus[t] == BK_OK;
  const int n = active ? k.count[t] : 0;
  double* R = dyn + g * cal_doubles(NMAX);  // sample coordinates, 3 per sample
  double* L = R + 3 * NMAX;                 // packed lower triangle, L[tri(i) + j]
  double wx = 0.0, wy = 0.0, wz = 0.0, e = 0.0;
  if (r < n) {
    const int id = a.nidx[t * a.stride + r];
    wx = a.sx[id];
    wy = a.sy[id];
    wz = a.sz[id];
    e = a.serr ? a.serr[id] : 0.0;
  }
  R[3 * r] = wx;
  R[3 * r + 1] = wy;
  R[3 * r + 2] = wz;
  wave_sync();
  // Row r of the covariance matrix and its right-hand side entry.
  double c = 0.0;
  if (r < n) {
    for (int j = 0; j < r; ++j) L[tri(r) + j] = cov_points64(a.v, wx, wy, wz, R[3 * j], R[3 * j + 1], R[3 * j + 2]) + 0.0;
    L[tri(r) + r] = cov_points64(a.v, wx, wy, wz, wx, wy, wz) + e;
    const double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
    if (a.block) {
      double acc = -0.0;
      for (int p = 0; p < a.npts; ++p) {
        const double qx = t0 + a.offs[3 * p], qy = t1 + a.offs[3 * p + 1], qz = t2 + a.offs[3 * p + 2];
        acc += a.npts == 1 ? cov_points64(a.v, wx, wy, wz, qx, qy, qz)
                           : block_cov_points64(a.v, wx, wy, wz, qx, qy, qz);
      }
      c = acc / (double)a.npts;
    } else {
      c = cov_points64(a.v, wx, wy, wz, t0, t1, t2);
    }
  }
  wave_sync();
  // Cholesky, column j after the columns before it.
  bool fail = false;
  for (int j = 0; j < NMAX; ++j) {
    double v = 0.0;
    if (j < n && j <= r && r < n) {
      v = L[tri(r) + j];
      for (int q = 0; q < j; ++q) v = -L[tri(j) + q] * L[tri(r) + q] + v;
    }
    const double d = group_bcast<NMAX>(v, j, g);
    if (j < n && !(d > 0.0)) fail = true;
    if (j < n && !fail) {
      const double denom = sqrt(d);
      if (r == j) L[tri(j) + j] = denom;
      else if (r > j && r < n) L[tri(r) + j] = v / denom;
    }
    wave_sync();
  }
  // Forward solves of the right-hand side and of the column of ones.
  double y = c, yx = 1.0;
  for (int i = 0; i < NMAX; ++i) {
    double ci = 0.0, xi = 0.0;
    if (r == i && i < n && !fail) {
      const double dii = L[tri(i) + i];
      ci = y / dii;
      xi = yx / dii;
      y = ci;
      yx = xi;
    }
    ci = group_bcast<NMAX>(ci, i, g);
    xi = group_bcast<NMAX>(xi, i, g);
    if (r > i && r < n && i < n && !fail) {
      const double lri = L[tri(r) + i];
      y = -ci * lri + y;
      yx = -xi * lri + yx;
    }
  }
  // Running sums over the leading parts; lane r keeps those of the first r + 1.
  double ss = 0.0, aa = 0.0, bb = 0.0, my_s = 0.0, my_a = 0.0, my_b = 0.0;
  for (int q = 0; q < NMAX; ++q) {
    const double yq = group_bcast<NMAX>(y, q, g), xq = group_bcast<NMAX>(yx, q, g);
    if (q < n) {
      ss += yq * yq;
      aa += xq * xq;
      bb += xq * yq;
    }
    if (r == q) {
      my_s = ss;
      my_a = aa;
      my_b = bb;
    }
  }
  bool meets = false;
  if (r < n && !fail) {
    bool some = false;
    double cov = 0.0, mu = 0.0;
    if (!k.ordinary) {
      cov = my_s;
      some = true;
    } else if (my_a > 0.0) {
      // The 1 x 1 Cholesky of a and its solve of b - x0, x0 = 1.
      const double dd = sqrt(my_a);
      if (!(dd <= 1e-7 * dd)) {
        const double mm = ((my_b - 1.0) / dd) / dd;
        cov = my_s - mm * my_b;
        mu = mm;
        some = true;
      }
    }
    const double support = a.block ? a.cbb : a.v.c0;
    const double slope = cov / (cov - mu), eff = (cov + mu) / support;
    const int64_t from = k.min_samples > 1 ? k.min_samples : 1;
    meets = some && r + 1 >= from && (k.calibration == BK_CALIBRATE_SLOPE ? slope >= k.target : eff >= k.target);
  }
  const uint64_t m = __ballot(meets);
  const uint64_t gm = NMAX == 64 ? ~0ull : (((1ull << NMAX) - 1ull) << (g * NMAX));
  const uint64_t mine = m & gm;
  if (active && r == 0 && mine) {
    const int first = __ffsll((unsigned long long)mine) - 1 - g * NMAX;
    if (first + 1 < n) k.count[t] = first + 1;
  }
  if (active && r == 0 && k.met) k.met[t] = mine ? 1 : 0;
}

__global__ void __launch_bounds__(64) k_guarded(const SolveArgs a) {
  extern __shared__ double dyn[];
  double* A = dyn;
  double* b = A + LU_MAX * LU_MAX;
  double* ws = b + LU_MAX;
  int* cl = (int*)(ws + 7 * LU_MAX);
  const int64_t count = *a.guard_count;
  for (int64_t sys = blockIdx.x; sys < count; sys += gridDim.x) {
    const int t = a.guard[sys];
    if (a.ncount[t] > BK_MAX_LU) continue;  // k_guarded_big
    lu_system<LU_MAX>(a, t, A, b, ws, cl, threadIdx.x);
    __syncthreads();
  }
}

// The systems of the same list above BK_MAX_LU samples, each workgroup on its
// own slice of the global workspace lu_ws (lu_doubles(LU_BIG) doubles, then
// lu_ints(LU_BIG) ints per workgroup).
__global__ void __launch_bounds__(64) k_guarded_big(const SolveArgs a, double* lu_ws) {
  double* A = lu_ws + (int64_t)blockIdx.x * (lu_doubles(LU_BIG) + (lu_ints(LU_BIG) + 1) / 2);
  double* b = A + (int64_t)LU_BIG * LU_BIG;
  double* ws = b + LU_BIG;
  int* cl = (int*)(ws + 7 * LU_BIG);
  const int64_t count = *a.guard_count;
  for (int64_t sys = blockIdx.x; sys < count; sys += gridDim.x) {
    const int t = a.guard[sys];
    if (a.ncount[t] <= BK_MAX_LU) continue;  // k_guarded
    lu_system<LU_BIG>(a, t, A, b, ws, cl, threadIdx.x);
    __syncthreads();
  }
}

// Targets per thread of k_bucket: a workgroup takes 256 * 8 consecutive
// targets and adds to each list's global count once, so the counts see a few
// hundred atomics per million targets instead of one per wavefront.
constexpr int BUCKET_PER = 8;

__global__ void __launch_bounds__(256) k_bucket(int64_t m, const int32_t* count, const int32_t* sstatus,
                                                int32_t* status, double* value, double* variance, int32_t* lists,
                                                int32_t* counts, int64_t list_stride, int e0, int e1, int e2, int e3,
                                                int n_buckets, int stationary) {
  __shared__ int s_cnt[5], s_base[5];
  const int tid = threadIdx.x, lane = tid % BK_WAVE;
  if (tid < 5) s_cnt[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * 256 * BUCKET_PER;
  int bk[BUCKET_PER], pos[BUCKET_PER];
#pragma unroll
  for (int r = 0; r < BUCKET_PER; ++r) {
    const int64_t t = base + (int64_t)r * 256 + tid;
    const bool live = t < m;
    int st = live ? sstatus[t] : BK_SEARCH_FAILED;
    const int n = live ? count[t] : 0;
    int b = -1;
    if (live && st == BK_OK) {
      if (n == 0) {
        st = BK_INSUFFICIENT_DATA;
      } else if (!stationary || n > BK_MAX_CHOL) {
        b = n_buckets;  // the LU list
      } else {
        const int edges[4] = {e0, e1, e2, e3};
        for (int k = 0; k < n_buckets; ++k)
          if (n <= edges[k]) {
            b = k;
            break;
          }
      }
    }
    if (live && b < 0) {
      status[t] = st;
      value[t] = __longlong_as_double(0x7ff8000000000000LL);
      variance[t] = __longlong_as_double(0x7ff8000000000000LL);
    }
    // Places within the workgroup: one LDS atomic per wavefront and list.
    int p = -1;
    for (int k = 0; k <= n_buckets; ++k) {
      const uint64_t mask = __ballot(b == k);
      if (mask == 0) continue;
      const int leader = __ffsll((unsigned long long)mask) - 1;
      int w = 0;
      if (lane == leader) w = atomicAdd(&s_cnt[k], __popcll(mask));
      w = __shfl(w, leader, BK_WAVE);
      if (b == k) p = w + __popcll(mask & ((1ull << lane) - 1ull));
    }
    bk[r] = b;
    pos[r] = p;
  }
  __syncthreads();
  if (tid <= n_buckets) s_base[tid] = s_cnt[tid] > 0 ? atomicAdd(&counts[tid], s_cnt[tid]) : 0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < BUCKET_PER; ++r) {
    const int64_t t = base + (int64_t)r * 256 + tid;
    if (bk[r] >= 0) {
      lists[(int64_t)bk[r] * list_stride + s_base[bk[r]] + pos[r]] = (int32_t)t;
      status[t] = BK_OK;
    }
  }
}
#endif

}  // namespace

namespace bk {

// Dynamic LDS above the default needs the attribute; a refused request must
// not linger as the runtime's last error.
[[maybe_unused]] static void allow_lds(const void* fn, size_t bytes) {
  if (bytes > 48 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  (void)hipGetLastError();
}


// Wavefronts per workgroup: the requested count (1, 2 or 4), halved until the
// workgroup's LDS fits the 160 KiB of a CU.
[[maybe_unused]] static int fit_waves(int requested, size_t bytes_per_wave) {
  int w = requested >= 1 && requested <= 4 ? requested : 2;
  while (w > 1 && (size_t)w * bytes_per_wave > 160 * 1024) w >>= 1;
  return w;
}

// Compute units of the current device (cached per device).
[[maybe_unused]] static int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 256;
  }
  int cu = dev >= 0 && dev < 64 ? cache[dev].load(std::memory_order_relaxed) : 0;
  if (cu > 0) return cu;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) {
    (void)hipGetLastError();
    cu = 256;
  }
  if (dev >= 0 && dev < 64) cache[dev].store(cu, std::memory_order_relaxed);
  return cu;
}

// Grid of a persistent launch over `work` workgroups' worth of units: the
// workgroups the device holds at once (by the kernel's registers and LDS), at
// most `work`, as a multiple of the XCD count so that every unit range of
// unit_loop has its workgroups. The waves take their units from counters, so
// a workgroup that is not resident at first only finds its range done.
[[maybe_unused]] static unsigned persistent_grid(const void* fn, int threads, size_t sh, int64_t work) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, sh) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  const int64_t resident = (int64_t)per_cu * device_cus() / BK_XCDS * BK_XCDS;
  const int64_t want = (work + BK_XCDS - 1) / BK_XCDS * BK_XCDS;
  return (unsigned)std::max<int64_t>(std::min(resident, want), BK_XCDS);
}

// Launch shape of a solve kernel with `doubles` of LDS per system.
struct SolveShape {
  int waves;
  size_t sh;
  int64_t work;
};
template <int NMAX>
[[maybe_unused]] static SolveShape solve_shape(const SolveArgs& a, int doubles) {
  const int W = fit_waves(a.waves, (size_t)(BK_WAVE / NMAX) * doubles * sizeof(double));
  const int per_block = W * (BK_WAVE / NMAX);
  return SolveShape{W, (size_t)per_block * doubles * sizeof(double), (a.n_list + per_block - 1) / per_block};
}

#if BK_PART != 2
template <int NMAX>
static int run64(hipStream_t s, const SolveArgs& a) {
  const SolveShape q = solve_shape<NMAX>(a, sys64_doubles(NMAX, a.v.nf));
  if (q.work == 0) return 0;
  if (!a.work) return -1;  // persistent kernels take their units from this counter
  const void* fn = (const void*)k_krige64<NMAX>;
  allow_lds(fn, q.sh);
  k_krige64<NMAX><<<dim3(persistent_grid(fn, 64 * q.waves, q.sh, q.work)), dim3(64 * q.waves), q.sh, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int NMAX>
static int run32(hipStream_t s, const SolveArgs& a) {
  const SolveShape q = solve_shape<NMAX>(a, sys32_doubles(NMAX, a.v.nf));
  if (q.work == 0) return 0;
  if (!a.work) return -1;  // persistent kernels take their units from this counter
  const void* fn = (const void*)k_krige32<NMAX>;
  allow_lds(fn, q.sh);
  k_krige32<NMAX><<<dim3(persistent_grid(fn, 64 * q.waves, q.sh, q.work)), dim3(64 * q.waves), q.sh, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

static int run64m(hipStream_t s, const SolveArgs& a) {
  const SolveShape q = solve_shape<64>(a, sysm64_doubles(a.v.nf));
  if (q.work == 0) return 0;
  if (!a.work) return -1;  // persistent kernels take their units from this counter
  const void* fn = (const void*)k_krige64m;
  allow_lds(fn, q.sh);
  k_krige64m<<<dim3(persistent_grid(fn, 64 * q.waves, q.sh, q.work)), dim3(64 * q.waves), q.sh, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_loo_grad32m(hipStream_t s, const SolveArgs& a, const GradArgs& gp, const GradArgs32& g32) {
  const SolveShape q = solve_shape<32>(a, sysl_doubles(a.v.nf) / 2);
  if (q.work == 0) return 0;
  if (!a.work) return -1;
  const void* fn = (const void*)k_loo_grad32m;
  allow_lds(fn, q.sh);
  k_loo_grad32m<<<dim3(persistent_grid(fn, 64 * q.waves, q.sh, q.work)), dim3(64 * q.waves), q.sh, s>>>(a, gp, g32);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

static int run32m(hipStream_t s, const SolveArgs& a) {
  const SolveShape q = solve_shape<32>(a, sysm_doubles(a.v.nf) / 2);
  if (q.work == 0) return 0;
  if (!a.work) return -1;  // persistent kernels take their units from this counter
  const void* fn = (const void*)k_krige32m;
  allow_lds(fn, q.sh);
  k_krige32m<<<dim3(persistent_grid(fn, 64 * q.waves, q.sh, q.work)), dim3(64 * q.waves), q.sh, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int NMAX>
static int run_grad(hipStream_t s, const SolveArgs& a, const GradArgs& gp) {
  const SolveShape q = solve_shape<NMAX>(a, sys_doubles(NMAX, a.v.nf));
  if (q.work == 0) return 0;
  if (!a.work) return -1;  // persistent kernels take their units from this counter
  const void* fn = (const void*)k_loo_grad<NMAX>;
  allow_lds(fn, q.sh);
  k_loo_grad<NMAX><<<dim3(persistent_grid(fn, 64 * q.waves, q.sh, q.work)), dim3(64 * q.waves), q.sh, s>>>(a, gp);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
#endif

#if BK_PART != 1
template <int NMAX>
static int run_grad32(hipStream_t s, const SolveArgs& a, const GradArgs& gp, const GradArgs32& g32) {
  const SolveShape q = solve_shape<NMAX>(a, loo32_doubles(NMAX, a.v.nf));
  if (q.work == 0) return 0;
  const void* fn = (const void*)k_loo_grad32<NMAX>;
  allow_lds(fn, q.sh);
  k_loo_grad32<NMAX><<<dim3((unsigned)q.work), dim3(64 * q.waves), q.sh, s>>>(a, gp, g32);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_loo_grad32(hipStream_t s, const SolveArgs& a, const GradArgs& gp, const GradArgs32& g32, int nmax) {
  switch (nmax) {
    case 16:
      return run_grad32<16>(s, a, gp, g32);
    case 32:
      return run_grad32<32>(s, a, gp, g32);
    case 64:
      return run_grad32<64>(s, a, gp, g32);
  }
  return -1;
}
#endif

#if BK_PART != 2
int launch_loo_grad(hipStream_t s, const SolveArgs& a, const GradArgs& gp, int nmax) {
  switch (nmax) {
    case 16:
      return run_grad<16>(s, a, gp);
    case 32:
      return run_grad<32>(s, a, gp);
    case 64:
      return run_grad<64>(s, a, gp);
  }
  return -1;
}
#endif

#if BK_PART != 1
int64_t reduce_blocks(int64_t m) { return m <= 0 ? 1 : (m + RED_CHUNK - 1) / RED_CHUNK; }

int launch_reduce_cols(hipStream_t s, int64_t m, int cols, const double* part, double* work, double* out) {
  const int64_t nb = reduce_blocks(m);
  k_reduce_part<<<dim3((unsigned)nb, (unsigned)cols), dim3(RED_T), 0, s>>>(m, part, work);
  k_reduce_final<<<dim3((unsigned)cols), dim3(RED_T), 0, s>>>(nb, work, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
#endif

#if BK_PART != 2
int launch_calibrate(hipStream_t s, const SolveArgs& a, const CalibArgs& k, int nmax) {
  if (k.m <= 0) return 0;
  auto run = [&](auto kern, int nm) {
    const int G = BK_WAVE / nm;
    const size_t sh = (size_t)G * cal_doubles(nm) * sizeof(double);
    allow_lds((const void*)kern, sh);
    kern<<<dim3((unsigned)((k.m + G - 1) / G)), dim3(BK_WAVE), sh, s>>>(a, k);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  };
  switch (nmax) {
    case 16:
      return run(k_calibrate<16>, 16);
    case 32:
      return run(k_calibrate<32>, 32);
    case 64:
      return run(k_calibrate<64>, 64);
  }
  return -1;
}

int launch_krige64(hipStream_t s, const SolveArgs& a, int nmax) {
  switch (nmax) {
    case 16:
      return run64<16>(s, a);
    case 32:
      return run64<32>(s, a);
    case 64:
      return run64<64>(s, a);
  }
  return -1;
}

int launch_krige32(hipStream_t s, const SolveArgs& a, int nmax) {
  switch (nmax) {
    case 16:
      return run32<16>(s, a);
    case 32:
      return run32m(s, a);
    case 64:
      return run64m(s, a);
  }
  return -1;
}

// Handles a.guard[0 .. *a.guard_count); a.n_list bounds that count.
int launch_guarded(hipStream_t s, const SolveArgs& a, double* lu_ws) {
  if (a.n_list <= 0) return 0;
  size_t sh = (size_t)lu_doubles(LU_MAX) * sizeof(double) + (size_t)lu_ints(LU_MAX) * sizeof(int);
  const void* fn = (const void*)k_guarded;
  allow_lds(fn, sh);
  k_guarded<<<dim3(persistent_grid(fn, 64, sh, a.n_list)), dim3(64), sh, s>>>(a);
  if (hipGetLastError() != hipSuccess) return -2;
  if (lu_ws) {
    k_guarded_big<<<dim3(BK_LU_BIG_GROUPS), dim3(64), 0, s>>>(a, lu_ws);
    if (hipGetLastError() != hipSuccess) return -2;
  }
  return 0;
}

size_t lu_big_workspace_bytes() {
  return (size_t)BK_LU_BIG_GROUPS * (size_t)(lu_doubles(LU_BIG) + (lu_ints(LU_BIG) + 1) / 2) * sizeof(double);
}

int launch_bucket(hipStream_t s, int64_t m, const int32_t* count, const int32_t* sstatus, int32_t* status,
                  double* value, double* variance, int32_t* lists, int32_t* counts, int64_t list_stride,
                  const int* edges, int n_buckets, int stationary) {
  if (m <= 0) return 0;
  int e[4] = {0, 0, 0, 0};
  for (int k = 0; k < n_buckets && k < 4; ++k) e[k] = edges[k];
  const int64_t per_block = 256 * BUCKET_PER;
  k_bucket<<<dim3((unsigned)((m + per_block - 1) / per_block)), dim3(256), 0, s>>>(
      m, count, sstatus, status, value, variance, lists, counts, list_stride, e[0], e[1], e[2], e[3], n_buckets,
      stationary);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
#endif

}  // namespace bk
