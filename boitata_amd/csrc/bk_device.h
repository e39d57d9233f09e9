// bk_device.h — device-side covariance arithmetic shared by the search and
// solve kernels of the MI355X kriging engine (gfx950, wave64).
//
// FP64 functions follow the reference's operation order so the plain FP64
// path reproduces boitatá's rounding wherever the math library agrees:
//   crates/variogram/src/model.rs:76-126       shape
//   crates/variogram/src/composite.rs:50-115   total_sill, gamma, cov, block_cov_points
//   crates/variogram/src/aniso.rs:101-119      Anisotropy::lag, euclidean
//   crates/estimation/src/kernel.rs:36-56      the frame and the packed covariance fill
// The translation units that include this header are compiled with
// -ffp-contract=off; FP32 code opts into contraction locally.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/boitata_gpu.h"

#define BK_WAVE 64
#define BK_MAX_FRAMES (BK_MAX_STRUCTURES + 1)

// Wave-local ordering of LDS traffic between lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Persistent launches. The work units [0, units) are split into BK_XCDS
// contiguous ranges; range x is served by the workgroups with
// blockIdx % BK_XCDS == x (gfx950 hands workgroups to its 8 XCDs round robin,
// so a range stays in one XCD's L2, where neighbouring targets share
// samples). Each wavefront takes its first unit of the range by its index,
// then further ones from the range's counter (zeroed before the launch, one
// 64-byte line per counter): the time does not depend on how many workgroups
// are resident at once, a launch without units makes no atomic, and eight
// counters spread the contention.
#define BK_XCDS 8
#define BK_CTR_STRIDE 16  // int32 per counter: one 64-byte line
#define BK_CTR_INTS (BK_XCDS * BK_CTR_STRIDE)

__device__ __forceinline__ int grab_unit(int32_t* counter) {
  int v = 0;
  if ((threadIdx.x & (BK_WAVE - 1)) == 0) v = atomicAdd(counter, 1);
  return __builtin_amdgcn_readfirstlane(v);
}

struct UnitLoop {
  int unit;      // current unit (>= end: done)
  int end;       // end of this wave's range
  int base;      // first unit handed out by the counter
  int32_t* ctr;
};

__device__ __forceinline__ UnitLoop unit_loop(int64_t units, int32_t* work) {
  const int x = blockIdx.x % BK_XCDS;
  const int lo = (int)(units * x / BK_XCDS), hi = (int)(units * (x + 1) / BK_XCDS);
  const int waves = blockDim.x / BK_WAVE;
  const int blocks = ((int)gridDim.x - x + BK_XCDS - 1) / BK_XCDS;  // workgroups serving range x
  UnitLoop q;
  q.unit = lo + (int)(blockIdx.x / BK_XCDS) * waves + (int)(threadIdx.x / BK_WAVE);
  q.end = hi;
  q.base = lo + blocks * waves;
  q.ctr = work + x * BK_CTR_STRIDE;
  return q;
}

__device__ __forceinline__ int unit_next(const UnitLoop& q) { return q.base + grab_unit(q.ctr); }

// The kernel's first argument (a struct passed by value sits at offset 0 of
// the kernarg segment), through a pointer the optimiser cannot see through.
// A persistent loop that reads its arguments this way reloads them (scalar
// loads, cache hits) per work block instead of hoisting dozens of them out
// of the loop, where they outgrow the SGPR file and spill into VGPR lanes.
template <class T>
__device__ __forceinline__ const T& kernarg_first() {
  typedef const __attribute__((address_space(4))) T* KPtr;
  KPtr p = (KPtr)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return *(const T*)p;
}

__device__ __forceinline__ double bcast(double v, int lane) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ float bcastf(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int bcasti(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// Sum over the wave (butterfly; every lane gets the total).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, BK_WAVE);
  return v;
}

__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, BK_WAVE);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, BK_WAVE));
  return v;
}

// The variogram as the kernels read it. Frames: the linear maps to the
// isotropic space of each structure (kernel.rs:36-42 matrix()); a structure on
// the shared anisotropy uses the shared frame, so with shared ratios there is
// one frame and the arithmetic is the reference's.
struct DevVariogram {
  int S;              // structures
  int nf;             // frames in use
  int nug_frame;      // frame of the nugget's h > 0 test (that of structure 0)
  int stationary;
  double nugget;
  double c0;          // total_sill in the reference's order
  double inv_c0;      // 1 / c0
  // Squared lag in the frame of structure 0 below which two samples may lie
  // within 1e-6 m of each other: (1.1e-6 |F|_F)^2, since |F d| <= |F|_F |d|.
  double near2;
  int model[BK_MAX_STRUCTURES];
  int frame[BK_MAX_STRUCTURES];
  double param[BK_MAX_STRUCTURES];
  double sill[BK_MAX_STRUCTURES];
  double range[BK_MAX_STRUCTURES];   // max(range, MIN_POSITIVE)
  double inv_range[BK_MAX_STRUCTURES];
  double F[BK_MAX_FRAMES][9];        // row-major frames (matrix() of each anisotropy)
  int ident[BK_MAX_FRAMES];          // the shared frame without anisotropy: no map, euclidean lags
  // Anisotropy::lag form (aniso.rs:101-111): shared rotation, per-frame axes.
  int has_aniso;
  double rot[9];
  double lag_major[BK_MAX_FRAMES], lag_semi[BK_MAX_FRAMES], lag_minor[BK_MAX_FRAMES];
};

// r_i = (m_i0 v0 + m_i1 v1) + m_i2 v2, unfused.
__device__ __forceinline__ void mv3(const double* m, double v0, double v1, double v2, double& r0, double& r1,
                                    double& r2) {
  r0 = (m[0] * v0 + m[1] * v1) + m[2] * v2;
  r1 = (m[3] * v0 + m[4] * v1) + m[5] * v2;
  r2 = (m[6] * v0 + m[7] * v1) + m[8] * v2;
}

// Rust powi lowered by binary exponentiation.
__device__ __forceinline__ double pw3(double t) { return t * (t * t); }
__device__ __forceinline__ double pw5(double t) {
  double t2 = t * t;
  return t * (t2 * t2);
}
__device__ __forceinline__ double pw7(double t) {
  double t2 = t * t;
  return (t * t2) * (t2 * t2);
}

// model.rs:76-126, FP64, t = h / range with a true division.
__device__ __forceinline__ double shape64(int model, double param, double h, double r) {
  if (!isfinite(h) || h <= 0.0) return 0.0;
  double t = h / r;
  switch (model) {
    case BK_SPHERICAL:
      return t >= 1.0 ? 1.0 : 1.5 * t - 0.5 * pw3(t);
    case BK_EXPONENTIAL:
      return 1.0 - exp(-3.0 * t);
    case BK_GAUSSIAN:
      return 1.0 - exp(-3.0 * t * t);
    case BK_CUBIC:
      if (t >= 1.0) return 1.0;
      return 7.0 * (t * t) - (35.0 / 4.0) * pw3(t) + (7.0 / 2.0) * pw5(t) - (3.0 / 4.0) * pw7(t);
    case BK_PENTASPHERICAL:
      if (t >= 1.0) return 1.0;
      return (15.0 / 8.0) * t - (5.0 / 4.0) * pw3(t) + (3.0 / 8.0) * pw5(t);
    case BK_CIRCULAR:
      if (t >= 1.0) return 1.0;
      return 1.0 - (2.0 / 3.14159265358979323846) * acos(t) +
             (2.0 * t / 3.14159265358979323846) * sqrt(1.0 - t * t);
    case BK_SINEHOLE: {
      double c = 3.14159265358979323846 * t;
      return 1.0 - sin(c) / c;
    }
    case BK_MATERN: {
      double nu = param < 1.0 ? 0.5 : (param < 2.0 ? 1.5 : 2.5);
      double delta = (sqrt(2.0 * nu) * 3.0) * t;
      double e = exp(-delta);
      if (nu == 0.5) return 1.0 - e;
      if (nu == 1.5) return 1.0 - (1.0 + delta) * e;
      return 1.0 - (1.0 + delta + delta * delta / 3.0) * e;
    }
    case BK_POWER:
      return pow(t, param);
  }
  return 0.0;
}

// composite.rs:69-85 with a lag per frame.
__device__ __forceinline__ double cov64(const DevVariogram& v, const double* h) {
  double hn = h[v.nug_frame];
  if (hn <= 0.0) return v.c0;
  double s = -0.0;
  for (int k = 0; k < v.S; ++k) s += v.sill[k] * shape64(v.model[k], v.param[k], h[v.frame[k]], v.range[k]);
  return v.c0 - (v.nugget + s);
}

// composite.rs:108-115: nugget left out where the points coincide.
__device__ __forceinline__ double block_cov64(const DevVariogram& v, const double* h) {
  if (h[v.nug_frame] <= 0.0) return v.c0 - v.nugget;
  return cov64(v, h);
}

__device__ __forceinline__ double norm3(double a, double b, double c) { return sqrt((a * a + b * b) + c * c); }

// Anisotropy::lag per frame between world points p and q (aniso.rs:101-111),
// or euclidean (aniso.rs:114-119) without anisotropy.
__device__ __forceinline__ void lags_world(const DevVariogram& v, double px, double py, double pz, double qx,
                                           double qy, double qz, double* h) {
  double vx = qx - px, vy = qy - py, vz = qz - pz;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  bool rotated = false;
  for (int f = 0; f < v.nf; ++f) {
    if (v.ident[f]) {
      h[f] = norm3(vx, vy, vz);
      continue;
    }
    if (!rotated) {
      mv3(v.rot, vx, vy, vz, r0, r1, r2);
      rotated = true;
    }
    double s0 = r0 / v.lag_major[f], s1 = r1 / v.lag_semi[f], s2 = r2 / v.lag_minor[f];
    h[f] = norm3(s0, s1, s2);
  }
}

__device__ __forceinline__ double cov_points64(const DevVariogram& v, double px, double py, double pz, double qx,
                                               double qy, double qz) {
  double h[BK_MAX_FRAMES];
  lags_world(v, px, py, pz, qx, qy, qz, h);
  return cov64(v, h);
}

__device__ __forceinline__ double block_cov_points64(const DevVariogram& v, double px, double py, double pz,
                                                     double qx, double qy, double qz) {
  double h[BK_MAX_FRAMES];
  lags_world(v, px, py, pz, qx, qy, qz, h);
  return block_cov64(v, h);
}

// ---------------------------------------------------------------------------
// Fast FP64 covariance for the refined path: relative error about 1e-15
// (the refinement needs ~1e-10), with no IEEE division or correctly rounded
// square root. exp by Cody-Waite reduction x = k ln2 + r, |r| <= ln2 / 2, and
// exp(r) = 1 + r + r^2 q(r), q the degree-8 minimax polynomial of
// (exp(r) - 1 - r) / r^2 on |r| <= ln2 / 2 by the Remez exchange; measured
// against extended precision the relative error of exp stays below 1.4e-15.
constexpr double kExpQ[9] = {0.5000000000000002,     0.16666666666648308,    0.041666666666595006,
                             0.008333333353716447,   0.0013888888929458666,  0.00019841208757688552,
                             2.4801504975856783e-05, 2.7625101598579117e-06, 2.7626859843999097e-07};

__device__ __forceinline__ double exp_fast(double x) {
#pragma clang fp contract(fast)
  if (x < -708.0) return 0.0;
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  double k = rint(x * 1.44269504088896338700);
  double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = kExpQ[8];
#pragma unroll
  for (int i = 7; i >= 0; --i) p = p * r + kExpQ[i];
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

// exp_fast's constants as opaque scalar-register values, for loops that
// evaluate many exponentials: each Horner step is then one v_fma_f64 with a
// scalar operand (an FP64 literal rematerialised in a VGPR pair costs two
// v_mov_b32 before every step). Same arithmetic, same results as exp_fast.
struct ExpConstants {
  double c[9];  // q's coefficients (kExpQ); 1.0 is an inline constant
  double log2e, ln2_hi, ln2_lo;
};

__device__ __forceinline__ ExpConstants exp_constants() {
  ExpConstants k;
#pragma unroll
  for (int i = 0; i < 9; ++i) k.c[i] = kExpQ[i];
  k.log2e = 1.44269504088896338700;
  k.ln2_hi = 6.93147180369123816490e-01;
  k.ln2_lo = 1.90821492927058770002e-10;
#pragma unroll
  for (int i = 0; i < 9; ++i) asm volatile("" : "+s"(k.c[i]));
  asm volatile("" : "+s"(k.log2e));
  asm volatile("" : "+s"(k.ln2_hi));
  asm volatile("" : "+s"(k.ln2_lo));
  return k;
}

// exp(x) for finite x >= -745 (below that ldexp returns 0 or the smallest
// subnormal): no range test, so a caller that clamps its argument (a
// covariance's exp of minus a finite lag) has no branch.
__device__ __forceinline__ double exp_core(double x, const ExpConstants& k) {
#pragma clang fp contract(fast)
  double n = rint(x * k.log2e);
  double r = (x - n * k.ln2_hi) - n * k.ln2_lo;
  double p = k.c[8];
#pragma unroll
  for (int i = 7; i >= 0; --i) p = p * r + k.c[i];
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)n);
}

__device__ __forceinline__ double exp_fast(double x, const ExpConstants& k) {
  if (x < -708.0) return 0.0;
  return exp_core(x, k);
}

// sqrt(x) for x > 0 (x = 0 gives NaN), from the hardware reciprocal square
// root and one Newton step.
__device__ __forceinline__ double sqrt_pos(double x) {
#pragma clang fp contract(fast)
  double y = __builtin_amdgcn_rsq(x);
  double s = x * y;
  double e = fma(-s, s, x);  // x - s^2
  return s + 0.5 * e * y;
}

// sqrt(x) from the hardware reciprocal square root and one Newton step.
__device__ __forceinline__ double sqrt_fast(double x) {
#pragma clang fp contract(fast)
  if (!(x > 0.0)) return 0.0;
  double y = __builtin_amdgcn_rsq(x);
  double s = x * y;
  double e = fma(-s, s, x);  // x - s^2
  return s + 0.5 * e * y;
}

__device__ __forceinline__ double shape_fast64(int model, double param, double h, double inv_r) {
#pragma clang fp contract(fast)
  if (!(h > 0.0)) return 0.0;
  double t = h * inv_r;
  switch (model) {
    case BK_SPHERICAL:
      return t >= 1.0 ? 1.0 : t * (1.5 - 0.5 * t * t);
    case BK_EXPONENTIAL:
      return 1.0 - exp_fast(-3.0 * t);
    case BK_GAUSSIAN:
      return 1.0 - exp_fast(-3.0 * t * t);
    case BK_CUBIC: {
      if (t >= 1.0) return 1.0;
      double t2 = t * t;
      return t2 * (7.0 + t * (-8.75 + t2 * (3.5 - 0.75 * t2)));
    }
    case BK_PENTASPHERICAL: {
      if (t >= 1.0) return 1.0;
      double t2 = t * t;
      return t * (1.875 + t2 * (-1.25 + 0.375 * t2));
    }
    case BK_MATERN: {
      double nu = param < 1.0 ? 0.5 : (param < 2.0 ? 1.5 : 2.5);
      double delta = (nu == 0.5 ? 3.0 : (nu == 1.5 ? 5.196152422706632 : 6.708203932499369)) * t;
      double e = exp_fast(-delta);
      if (nu == 0.5) return 1.0 - e;
      if (nu == 1.5) return 1.0 - (1.0 + delta) * e;
      return 1.0 - (1.0 + delta + delta * delta * (1.0 / 3.0)) * e;
    }
  }
  // Circular, sine hole and power have no fast form: NaN sends the system
  // to the FP64 kernel (the host routes such variograms there directly).
  return __builtin_nan("");
}

__device__ __forceinline__ double cov_fast64(const DevVariogram& v, const double* h) {
#pragma clang fp contract(fast)
  if (!(h[v.nug_frame] > 0.0)) return v.c0;
  double s = 0.0;
  for (int k = 0; k < v.S; ++k) s += v.sill[k] * shape_fast64(v.model[k], v.param[k], h[v.frame[k]], v.inv_range[k]);
  return v.c0 - (v.nugget + s);
}

__device__ __forceinline__ double norm3_fast(double a, double b, double c) {
#pragma clang fp contract(fast)
  return sqrt_fast(a * a + b * b + c * c);
}

// Anisotropy::lag per frame (aniso.rs:101-111) in the fast arithmetic.
__device__ __forceinline__ void lags_world_fast(const DevVariogram& v, double vx, double vy, double vz, double* h) {
#pragma clang fp contract(fast)
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  bool rotated = false;
  for (int f = 0; f < v.nf; ++f) {
    if (v.ident[f]) {
      h[f] = norm3_fast(vx, vy, vz);
      continue;
    }
    if (!rotated) {
      r0 = v.rot[0] * vx + v.rot[1] * vy + v.rot[2] * vz;
      r1 = v.rot[3] * vx + v.rot[4] * vy + v.rot[5] * vz;
      r2 = v.rot[6] * vx + v.rot[7] * vy + v.rot[8] * vz;
      rotated = true;
    }
    h[f] = norm3_fast(r0 / v.lag_major[f], r1 / v.lag_semi[f], r2 / v.lag_minor[f]);
  }
}

// ---------------------------------------------------------------------------
// FP32 forms for the refined path (contraction allowed, native transcendentals).
struct DevVariogram32 {
  int S, nf, nug_frame;
  float nugget, c0;
  int model[BK_MAX_STRUCTURES];
  int frame[BK_MAX_STRUCTURES];
  float param[BK_MAX_STRUCTURES];
  float sill[BK_MAX_STRUCTURES];
  float inv_range[BK_MAX_STRUCTURES];
};

__device__ __forceinline__ float shape32(int model, float param, float h, float inv_r) {
#pragma clang fp contract(fast)
  if (!(h > 0.0f)) return 0.0f;
  float t = h * inv_r;
  switch (model) {
    case BK_SPHERICAL:
      return t >= 1.0f ? 1.0f : t * (1.5f - 0.5f * t * t);
    case BK_EXPONENTIAL:
      return 1.0f - __expf(-3.0f * t);
    case BK_GAUSSIAN:
      return 1.0f - __expf(-3.0f * t * t);
    case BK_CUBIC: {
      if (t >= 1.0f) return 1.0f;
      float t2 = t * t;
      return t2 * (7.0f + t * (-8.75f + t2 * (3.5f - 0.75f * t2)));
    }
    case BK_PENTASPHERICAL: {
      if (t >= 1.0f) return 1.0f;
      float t2 = t * t;
      return t * (1.875f + t2 * (-1.25f + 0.375f * t2));
    }
    case BK_CIRCULAR:
      if (t >= 1.0f) return 1.0f;
      return 1.0f - 0.63661977f * acosf(t) + 0.63661977f * t * sqrtf(1.0f - t * t);
    case BK_SINEHOLE: {
      float c = 3.14159265f * t;
      return 1.0f - __sinf(c) / c;
    }
    case BK_MATERN: {
      float nu = param < 1.0f ? 0.5f : (param < 2.0f ? 1.5f : 2.5f);
      float delta = sqrtf(2.0f * nu) * 3.0f * t;
      float e = __expf(-delta);
      if (nu == 0.5f) return 1.0f - e;
      if (nu == 1.5f) return 1.0f - (1.0f + delta) * e;
      return 1.0f - (1.0f + delta + delta * delta * (1.0f / 3.0f)) * e;
    }
    case BK_POWER:
      return powf(t, param);
  }
  return 0.0f;
}

__device__ __forceinline__ float cov32(const DevVariogram32& v, const float* h) {
#pragma clang fp contract(fast)
  if (!(h[v.nug_frame] > 0.0f)) return v.c0;
  float s = 0.0f;
  for (int k = 0; k < v.S; ++k) s += v.sill[k] * shape32(v.model[k], v.param[k], h[v.frame[k]], v.inv_range[k]);
  return v.c0 - (v.nugget + s);
}
