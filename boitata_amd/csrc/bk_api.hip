// bk_api.hip — the C ABI of include/boitata_gpu.h: context, device buffers and
// the per-pass pipeline (grid build, search, bucketing by neighbour count,
// batched solves, deferred systems).
//
// Host-side parameter arithmetic restates crates/core/src/rotation.rs:8-43 and
// crates/variogram/src/aniso.rs:88-94 in the reference's order so that the
// frames handed to the kernels equal the reference's to the bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/boitata_gpu.h"
#include "bk_diag.h"
#include "bk_mik.h"
#include "bk_search.h"
#include "bk_solve.h"

namespace {

struct M3 {
  double m[9];
};

M3 mm(const M3& a, const M3& b) {
  M3 c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t0 = a.m[3 * i] * b.m[j];
      double t1 = a.m[3 * i + 1] * b.m[3 + j];
      double t2 = a.m[3 * i + 2] * b.m[6 + j];
      c.m[3 * i + j] = (t0 + t1) + t2;
    }
  return c;
}

const double kPi = 3.14159265358979323846;

// rotation.rs:8-16
M3 rotation(double az, double dip, double rake) {
  double a = az * (kPi / 180.0), d = dip * (kPi / 180.0), r = rake * (kPi / 180.0);
  double sa = std::sin(a), ca = std::cos(a), sd = std::sin(d), cd = std::cos(d), sr = std::sin(r), cr = std::cos(r);
  M3 r1{{sa, ca, 0.0, -ca, sa, 0.0, 0.0, 0.0, 1.0}};
  M3 r2{{cd, 0.0, -sd, 0.0, 1.0, 0.0, sd, 0.0, cd}};
  M3 r3{{1.0, 0.0, 0.0, 0.0, cr, sr, 0.0, -sr, cr}};
  return mm(mm(r3, r2), r1);
}

// rotation.rs:40-43
M3 block_frame(double az, double dip, double rake) {
  M3 r = rotation(az, dip, rake), b;
  for (int j = 0; j < 3; ++j) {
    b.m[j] = -r.m[3 + j];
    b.m[3 + j] = r.m[j];
    b.m[6 + j] = r.m[6 + j];
  }
  return b;
}

// aniso.rs:88-94 with major = 1.
M3 aniso_matrix(double az, double dip, double rake, double semi, double minor) {
  M3 d{{1.0 / 1.0, 0.0, 0.0, 0.0, 1.0 / semi, 0.0, 0.0, 0.0, 1.0 / minor}};
  return mm(d, rotation(az, dip, rake));
}

const M3 kIdentity{{1, 0, 0, 0, 1, 0, 0, 0, 1}};

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

}  // namespace

struct bk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // Samples.
  int64_t n = 0;
  Buf x, y, z, v, e, hole;
  bool has_err = false, has_hole = false;
  bool has_origin = false;
  double origin[3] = {0, 0, 0};
  bool planar = true;
  // Grid cache, keyed by the search frame.
  bk::GridBuffers gb;
  bk::TileBuffers tiles;
  int search_mode = 0;  // uncapped searches: 0 wave per target, 1..8 tiles of that reach, -1 lane per target
  int solve_waves = 2;  // wavefronts per workgroup of the solve kernels
  // FP32+IR solves buckets of at most this many samples by FP32 elimination
  // on the matrix cores and FP64 refinement; larger buckets take the FP64 kernel.
  int refine_max = 64;
  GridDev grid{};
  bool grid_valid = false;
  double grid_F[9] = {0};
  // Grid side as a factor of the reference's rule, for searches without caps
  // (measured best for the wave search) and for capped searches (octants,
  // sectors or a hole cap: whose walk to the radius around targets with an
  // unfilled octant costs per cell visited, measured best at 1.0 - 1.4 on the
  // uneven 4M-target octant workload). grid_cf: the factor the grid was built with.
  double cell_factor = 0.65;
  double cell_factor_capped = 1.0;
  double grid_cf = 0.0;
  ShellTable tab{};
  void* tab_storage = nullptr;
  // Scratch (grow-only).
  Buf tx, ty, tz, left, nidx, ncount, sstatus, lists, counts, status, value, variance, lagrange, support, n_used,
      guard, redo, offs, which, iters, work;
  // Workspace of the LU for systems above BK_MAX_LU samples (searches with
  // max_samples > BK_MAX_LU only).
  Buf lu_big;
  // Set by bk_krige_noalloc while its scratch views stand in for the buffers
  // above: a buffer that would have to grow is an error, not an allocation.
  bool fixed_buffers = false;
  // Tuning.
  int edges[3] = {16, 32, 64};
  int max_iter = 8;
  double tol = 1e-9;
  // LOO-CV state.
  bool loo_ready = false;
  int64_t loo_m = 0;
  int loo_stride = 0;
  Buf loo_idx, loo_count, loo_status;
  // The held-out systems bucketed by neighbour count once, at prepare time:
  // lists of loo_hc[b] samples with n in (loo_edges[b-1], loo_edges[b]].
  Buf loo_lists;
  int32_t loo_hc[4] = {0, 0, 0, 0};
  int loo_edges[3] = {16, 32, 64};
  // Graph replay of point-kriging passes over device arrays: a pass repeated
  // with the same inputs and unchanged state (gen) is captured once and then
  // launched as one graph.
  bool graphs = true;
  bool capturing = false;
  uint64_t gen = 1;  // bumped whenever a buffer moves, the samples, grid or a tuning knob change
  struct Graph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;  // null: seen once, not captured yet
    uint64_t used = 0;
  };
  std::vector<Graph> graph_cache;  // a few passes (say FP32+IR and FP64 of one workload)
  uint64_t graph_clock = 0;
  int64_t graph_launches = 0;
  // Multiple indicator kriging: the global distribution of the last call on
  // the host and the device, kept while its inputs (samples, thresholds,
  // tails, weights) are the same; scratch for weights, estimates and status.
  uint64_t samples_gen = 0;
  std::vector<unsigned char> mik_key;
  std::vector<double> mik_values, mik_sums, mik_prop;
  std::vector<int64_t> mik_bounds;
  double mik_lower = 0, mik_upper = 0;
  Buf mik_gv, mik_gs, mik_par, mik_cls, mik_tab, mik_w, mik_raw, mik_st, mik_cdf, mik_out;
  // Sequential simulation (bk_sgs): a context of its own over the data and
  // the simulated nodes, and the pass's arrays.
  bk_ctx* sgs = nullptr;
  int sgs_passes = 0;  // draw passes of the last bk_sgs
  Buf sgs_x, sgs_y, sgs_z, sgs_e, sgs_h, sgs_v, sgs_s, sgs_done, sgs_limit, sgs_datum, sgs_pos, sgs_ctr, sgs_in, sgs_w,
      sgs_path, sgs_app, sgs_work, sgs_prog;
  void* sgs_tmp = nullptr;  // scan storage
  size_t sgs_tmp_bytes = 0;
  // Stage timing (bk_profile): events on the context's stream.
  bool prof = false;
  hipEvent_t ev[7] = {};
  double prof_ms[6] = {0, 0, 0, 0, 0, 0};
  int64_t prof_calls = 0;
};

namespace {
void mark(bk_ctx* c, int k) {
  if (c->prof) (void)hipEventRecord(c->ev[k], c->stream);
}

// After the stream has synchronised: add this call's stage times.
void accumulate(bk_ctx* c) {
  if (!c->prof) return;
  float t[6];
  const int from[6] = {0, 1, 2, 3, 4, 0}, to[6] = {1, 2, 3, 4, 5, 6};
  for (int k = 0; k < 6; ++k) {
    if (hipEventElapsedTime(&t[k], c->ev[from[k]], c->ev[to[k]]) != hipSuccess) return;
  }
  for (int k = 0; k < 6; ++k) c->prof_ms[k] += t[k];
  c->prof_calls += 1;
}
}  // namespace

namespace {

// Message of the last failure of a call made without a context (validation
// before bk_create, or bk_create itself), per thread.
thread_local std::string g_err;

int fail(bk_ctx* c, int code, const std::string& m) {
  (c ? c->err : g_err) = m;
  return code;
}

#define HIPOK(c, x)                                                        \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) return fail(c, BK_E_HIP, hipGetErrorString(e_)); \
  } while (0)

int ensure(bk_ctx* c, Buf& b, size_t bytes) {
  if (bytes <= b.cap) return 0;
  if (c->capturing) return fail(c, BK_E_STATE, "a buffer would move while a pass is captured");
  if (c->fixed_buffers) return fail(c, BK_E_STATE, "bk_krige_noalloc: a buffer beyond its static scratch");
  c->gen++;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
  if (bytes == 0) return 0;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) return fail(c, BK_E_NOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
  b.cap = bytes;
  return 0;
}

void release(Buf& b) {
  if (b.p) (void)hipFree(b.p);
  b = Buf();
}

bool is_device(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t at;
  hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

// An input array on the device: itself when device memory, else a copy.
template <class T>
int stage_in(bk_ctx* c, const T* src, int64_t n, Buf& b, const T** out) {
  if (!src) {
    *out = nullptr;
    return 0;
  }
  if (is_device(src)) {
    *out = src;
    return 0;
  }
  if (int r = ensure(c, b, std::max<int64_t>(n, 1) * sizeof(T))) return r;
  if (n > 0) HIPOK(c, hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  *out = (const T*)b.p;
  return 0;
}

// An output array on the device: itself when device memory, else scratch
// copied back by stage_out.
template <class T>
int out_ptr(bk_ctx* c, T* dst, int64_t n, Buf& b, T** out) {
  if (!dst) {
    *out = nullptr;
    return 0;
  }
  if (is_device(dst)) {
    *out = dst;
    return 0;
  }
  if (int r = ensure(c, b, std::max<int64_t>(n, 1) * sizeof(T))) return r;
  *out = (T*)b.p;
  return 0;
}

template <class T>
int stage_out(bk_ctx* c, T* dst, const T* dev, int64_t n) {
  if (!dst || dst == dev || n == 0) return 0;
  HIPOK(c, hipMemcpyAsync(dst, dev, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  return 0;
}

// Search::check (search.rs:214-264).
const char* check_search(const bk_search* s) {
  if (std::isnan(s->radius) || s->radius <= 0.0 || s->max_samples <= 0 || s->min_samples > s->max_samples ||
      s->min_samples < 0)
    return "need radius > 0 and 1 <= min_samples <= max_samples";
  int64_t count = 0, ring = 0;
  if (s->sectors > 0) {
    if (s->octant) return "give octant or sectors, not both";
    if (s->sectors < 2 || s->sectors > 360 || s->max_per_sector <= 0)
      return "sectors need 2 <= count <= 360 and max_per_sector >= 1";
    count = ring = s->sectors;
  } else if (s->octant) {
    count = 8;
    ring = 4;
  }
  if (s->max_per_octant >= 0) {
    if (!s->octant) return "max_per_octant needs octant";
    if (s->max_per_octant == 0) return "max_per_sector must be >= 1";
  }
  bool rules = s->min_sectors >= 0 || s->max_empty_sectors >= 0;
  if (rules && !(s->octant || s->sectors > 0)) return "min_sectors and max_empty_sectors need octant or sectors";
  if (s->min_sectors >= 0 && (s->min_sectors == 0 || s->min_sectors > count || s->min_sectors > s->max_samples))
    return "min_sectors must be at least 1 and at most the number of sectors (8 octants) and max_samples";
  if (s->max_empty_sectors >= 0 && s->max_empty_sectors >= ring)
    return "max_empty_sectors must be below the number of sectors (4 octants per half)";
  if (s->min_holes >= 0 && (s->min_holes == 0 || s->min_holes > s->max_samples))
    return "min_holes must be at least 1 and at most max_samples";
  if (s->has_anisotropy &&
      !(s->ratio_semi > 0 && std::isfinite(s->ratio_semi) && s->ratio_minor > 0 && std::isfinite(s->ratio_minor)))
    return "all ranges must be finite and > 0";
  if (s->has_anisotropy && !(std::isfinite(s->azimuth) && std::isfinite(s->dip) && std::isfinite(s->rake)))
    return "azimuth, dip and rake must be finite";
  // py/estimation.rs:458-472
  if (s->calibration < 0 || s->calibration > BK_CALIBRATE_EFFICIENCY) return "unknown calibration";
  if (s->calibration == BK_CALIBRATE_SLOPE && !(std::isfinite(s->calibration_target) && s->calibration_target > 0.0))
    return "target_slope must be finite and > 0";
  if (s->calibration == BK_CALIBRATE_EFFICIENCY &&
      !(std::isfinite(s->calibration_target) && s->calibration_target <= 1.0))
    return "target_efficiency must be finite and <= 1";
  return nullptr;
}

// Variogram::validate (composite.rs:37-48, model.rs:44-57, aniso.rs:29-47).
const char* check_variogram(const bk_variogram* v) {
  if (!(v->nugget >= 0.0 && std::isfinite(v->nugget))) return "nugget must be finite and >= 0";
  if (v->n_structures < 0 || v->n_structures > BK_MAX_STRUCTURES) return "at most 4 structures";
  for (int k = 0; k < v->n_structures; ++k) {
    const bk_structure& s = v->structures[k];
    if (s.model < BK_SPHERICAL || s.model > BK_POWER) return "unknown model";
    if (!(s.sill >= 0.0 && std::isfinite(s.sill) && s.range > 0.0 && std::isfinite(s.range)))
      return "structure sill must be finite and >= 0, range finite and > 0";
    bool own = s.ratio_semi > 0.0 || s.ratio_minor > 0.0;
    if (own && !(s.ratio_semi > 0 && std::isfinite(s.ratio_semi) && s.ratio_minor > 0 && std::isfinite(s.ratio_minor)))
      return "all ranges must be finite and > 0";
  }
  if (v->has_anisotropy) {
    if (!(v->ratio_semi > 0 && std::isfinite(v->ratio_semi) && v->ratio_minor > 0 && std::isfinite(v->ratio_minor)))
      return "all ranges must be finite and > 0";
    if (!(std::isfinite(v->azimuth) && std::isfinite(v->dip) && std::isfinite(v->rake)))
      return "azimuth, dip and rake must be finite";
  }
  return nullptr;
}

// The variogram as the kernels read it.
void dev_variogram(const bk_variogram* v, DevVariogram& d, DevVariogram32& f) {
  std::memset(&d, 0, sizeof(d));
  std::memset(&f, 0, sizeof(f));
  d.S = v->n_structures;
  d.nugget = v->nugget;
  double s = -0.0;
  for (int k = 0; k < d.S; ++k) s += v->structures[k].sill;
  d.c0 = v->nugget + s;  // composite.rs:50-52
  d.has_aniso = v->has_anisotropy;
  double az = v->has_anisotropy ? v->azimuth : 0.0, dip = v->has_anisotropy ? v->dip : 0.0,
         rk = v->has_anisotropy ? v->rake : 0.0;
  M3 rot = rotation(az, dip, rk);
  std::memcpy(d.rot, rot.m, sizeof(d.rot));
  // Frame 0: the shared anisotropy (or none); own ratios get frames of their own.
  bool shared_used = d.S == 0;
  for (int k = 0; k < d.S; ++k) {
    const bk_structure& st = v->structures[k];
    bool own = st.ratio_semi > 0.0 && st.ratio_minor > 0.0;
    if (!own) shared_used = true;
  }
  int nf = 0;
  int shared_frame = -1;
  if (shared_used) {
    shared_frame = nf++;
    M3 F = v->has_anisotropy ? aniso_matrix(az, dip, rk, v->ratio_semi, v->ratio_minor) : kIdentity;
    std::memcpy(d.F[shared_frame], F.m, sizeof(F.m));
    d.ident[shared_frame] = !v->has_anisotropy;
    d.lag_major[shared_frame] = 1.0;
    d.lag_semi[shared_frame] = v->has_anisotropy ? v->ratio_semi : 1.0;
    d.lag_minor[shared_frame] = v->has_anisotropy ? v->ratio_minor : 1.0;
  }
  d.stationary = 1;
  for (int k = 0; k < d.S; ++k) {
    const bk_structure& st = v->structures[k];
    bool own = st.ratio_semi > 0.0 && st.ratio_minor > 0.0;
    if (own) {
      int fr = nf++;
      M3 F = aniso_matrix(az, dip, rk, st.ratio_semi, st.ratio_minor);
      std::memcpy(d.F[fr], F.m, sizeof(F.m));
      d.ident[fr] = 0;
      d.lag_major[fr] = 1.0;
      d.lag_semi[fr] = st.ratio_semi;
      d.lag_minor[fr] = st.ratio_minor;
      d.frame[k] = fr;
    } else {
      d.frame[k] = shared_frame;
    }
    d.model[k] = st.model;
    d.param[k] = st.param;
    d.sill[k] = st.sill;
    d.range[k] = std::max(st.range, std::numeric_limits<double>::min());
    d.inv_range[k] = 1.0 / d.range[k];
    if (st.model == BK_POWER) d.stationary = 0;
  }
  d.nf = nf;
  d.nug_frame = d.S > 0 ? d.frame[0] : 0;
  d.inv_c0 = 1.0 / d.c0;
  {
    double fro = 0.0;
    for (int q = 0; q < 9; ++q) fro += d.F[d.nug_frame][q] * d.F[d.nug_frame][q];
    if (d.ident[d.nug_frame] || nf == 0) fro = 3.0;
    d.near2 = 1.21e-12 * fro;
  }
  f.S = d.S;
  f.nf = d.nf;
  f.nug_frame = d.nug_frame;
  f.nugget = (float)d.nugget;
  f.c0 = (float)d.c0;
  for (int k = 0; k < d.S; ++k) {
    f.model[k] = d.model[k];
    f.frame[k] = d.frame[k];
    f.param[k] = (float)d.param[k];
    f.sill[k] = (float)d.sill[k];
    f.inv_range[k] = (float)d.inv_range[k];
  }
}

// One atomic of each kind per wavefront: the first finite point of the
// wavefront, and its count of points off the first point's elevation.
__global__ void __launch_bounds__(256) k_first_finite(int64_t n, const double* x, const double* y, const double* z,
                                                      unsigned long long* out, unsigned long long* nonplanar) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < n;
  const uint64_t fin = __ballot(in && isfinite(x[i]) && isfinite(y[i]) && isfinite(z[i]));
  // search.rs:470-473: every point at the first point's elevation.
  const uint64_t off = __ballot(in && !(z[i] == z[0]));
  if ((threadIdx.x & (BK_WAVE - 1)) == 0) {
    const int64_t w0 = i;  // the wavefront's first point
    if (fin) atomicMin(out, (unsigned long long)(w0 + __builtin_ctzll(fin)));
    if (off) atomicAdd(nonplanar, (unsigned long long)__popcll(off));
  }
}

__global__ void k_copy_int(int64_t n, const int32_t* src, int32_t* dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src ? src[i] : -1;
}

__global__ void k_fill_nan(int64_t m, const int32_t* status, double* value, double* variance, double* lagrange,
                           double* support, int32_t* n_used) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  if (status[t] != BK_OK) {
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    value[t] = nan;
    variance[t] = nan;
    if (lagrange) lagrange[t] = nan;
    if (support) support[t] = nan;
    if (n_used) n_used[t] = 0;
  }
}

// The leave-one-out targets: sample locations, the sample left out.
__global__ void k_loo_targets(int64_t m, const int64_t* which, const double* x, const double* y, const double* z,
                              double* tx, double* ty, double* tz, int64_t* left) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  int64_t i = which ? which[k] : k;
  tx[k] = x[i];
  ty[k] = y[i];
  tz[k] = z[i];
  left[k] = i;
}

// batch.rs:177-182: an empty selection is no estimate.
__global__ void k_loo_nonempty(int64_t m, const int32_t* count, int32_t* sstatus) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m && sstatus[k] == BK_OK && count[k] == 0) sstatus[k] = BK_SEARCH_FAILED;
}

// A target whose candidate pool overflowed is not searched by a wider kernel
// yet: it is reported unsupported.
__global__ void k_map_overflow(int64_t m, int32_t* status) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m && status[t] == BK_ST_OVERFLOW) status[t] = BK_UNSUPPORTED;
}

unsigned blocks_of(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>((n + t - 1) / t, 1); }

// The sector of a sample from the signs of its projected offset (octants and
// planar quadrants): axes F^-1 must map each projected axis to one sector axis
// with a positive or negative scale (F = D R, axes the rows of R permuted with
// a sign). Depends on a.F, a.axes, a.balances, a.plane and a.planar.
void sector_map(SearchArgs& a) {
  a.sec_fast = 0;
  if (a.balances && !a.plane) {
    const double* f = a.F;
    const double det = f[0] * (f[4] * f[8] - f[5] * f[7]) - f[1] * (f[3] * f[8] - f[5] * f[6]) +
                       f[2] * (f[3] * f[7] - f[4] * f[6]);
    if (std::isfinite(det) && det != 0.0) {
      M3 inv{{(f[4] * f[8] - f[5] * f[7]) / det, (f[2] * f[7] - f[1] * f[8]) / det, (f[1] * f[5] - f[2] * f[4]) / det,
              (f[5] * f[6] - f[3] * f[8]) / det, (f[0] * f[8] - f[2] * f[6]) / det, (f[2] * f[3] - f[0] * f[5]) / det,
              (f[3] * f[7] - f[4] * f[6]) / det, (f[1] * f[6] - f[0] * f[7]) / det, (f[0] * f[4] - f[1] * f[3]) / det}};
      M3 axes;
      std::memcpy(axes.m, a.axes, sizeof(axes.m));
      const M3 map = mm(axes, inv);
      int perm[3], sgn[3];
      bool ok = true;
      for (int k = 0; k < 3 && ok; ++k) {
        int p = 0;
        for (int j = 1; j < 3; ++j)
          if (std::fabs(map.m[3 * k + j]) > std::fabs(map.m[3 * k + p])) p = j;
        const double big = std::fabs(map.m[3 * k + p]);
        for (int j = 0; j < 3; ++j)
          if (j != p && !(std::fabs(map.m[3 * k + j]) <= 1e-9 * big)) ok = false;
        perm[k] = p;
        sgn[k] = map.m[3 * k + p] > 0.0 ? 1 : -1;
      }
      ok = ok && perm[0] != perm[1] && perm[1] != perm[2] && perm[0] != perm[2];
      if (ok) {
        double inv_norm = 0.0;
        for (int i = 0; i < 3; ++i)
          inv_norm = std::max(inv_norm, std::fabs(inv.m[3 * i]) + std::fabs(inv.m[3 * i + 1]) +
                                            std::fabs(inv.m[3 * i + 2]));
        for (int p = 0; p < 3; ++p)
          a.sec_tol[p] = 1e-9 * (std::fabs(f[3 * p]) + std::fabs(f[3 * p + 1]) + std::fabs(f[3 * p + 2])) * inv_norm;
        uint32_t code = 0;
        for (int pat = 0; pat < 8; ++pat) {
          int bit[3];
          for (int k = 0; k < 3; ++k) {
            const int pos = (pat >> perm[k]) & 1;  // u_perm[k] >= 0
            bit[k] = sgn[k] > 0 ? pos : 1 - pos;
          }
          const int sec = a.planar ? (bit[0] << 1) | bit[1] : (bit[0] << 2) | (bit[1] << 1) | bit[2];
          code |= (uint32_t)sec << (3 * pat);
        }
        a.sec_code = code;
        a.sec_fast = 1;
      }
    }
  }
}

// Search parameters for the kernels, and the grid for the search's frame.
int prepare_search(bk_ctx* c, const bk_search* s, const bk_variogram* vg, SearchArgs& a, int64_t& stride) {
  if (const char* m = check_search(s)) return fail(c, BK_E_INVALID, m);
  std::memset(&a, 0, sizeof(a));
  // search.rs:343-348: the search's ellipsoid, else the variogram's.
  bool has = false;
  double az = 0, dip = 0, rk = 0, semi = 1, minor = 1;
  if (s->has_anisotropy) {
    has = true;
    az = s->azimuth;
    dip = s->dip;
    rk = s->rake;
    semi = s->ratio_semi;
    minor = s->ratio_minor;
  } else if (vg && vg->has_anisotropy) {
    has = true;
    az = vg->azimuth;
    dip = vg->dip;
    rk = vg->rake;
    semi = vg->ratio_semi;
    minor = vg->ratio_minor;
  }
  M3 F = has ? aniso_matrix(az, dip, rk, semi, minor) : kIdentity;
  M3 axes = has ? block_frame(az, dip, rk) : kIdentity;
  std::memcpy(a.F, F.m, sizeof(F.m));
  std::memcpy(a.axes, axes.m, sizeof(axes.m));
  for (int d = 0; d < 3; ++d) a.origin[d] = c->has_origin ? c->origin[d] : 0.0;
  a.radius2 = s->radius * s->radius;
  // A selection never holds more than the samples.
  int64_t maxs = std::min<int64_t>(s->max_samples, std::max<int64_t>(c->n, 1));
  if (maxs > BK_MAX_SEARCH)
    return fail(c, BK_E_INVALID, "max_samples above 256 (with as many samples) is beyond the device search");
  a.max_samples = (int)maxs;
  a.min_samples = (int)std::min<int64_t>(s->min_samples, std::numeric_limits<int>::max());
  a.balances = s->octant || s->sectors > 0;
  a.planar = c->planar;
  a.plane = s->sectors > 0 ? (int)s->sectors : 0;
  a.n_sectors = a.plane ? a.plane : (a.planar ? 4 : 8);
  if (s->sectors > 0)
    a.sector_cap = (int)std::min<int64_t>(s->max_per_sector, std::numeric_limits<int>::max());
  else if (s->octant)
    a.sector_cap = (int)std::min<int64_t>(
        s->max_per_octant >= 0 ? s->max_per_octant : (s->max_samples + a.n_sectors - 1) / a.n_sectors,
        std::numeric_limits<int>::max());
  else
    a.sector_cap = std::numeric_limits<int>::max();
  a.max_per_hole = s->max_per_hole >= 0 ? (int)std::min<int64_t>(s->max_per_hole, 1 << 30) : -1;
  a.octant_rank = s->octant && a.max_per_hole < 0;
  a.min_sectors = (int)s->min_sectors;
  a.max_empty_sectors = (int)s->max_empty_sectors;
  a.min_holes = (int)std::min<int64_t>(s->min_holes, 1 << 30);
  sector_map(a);
  a.sx = (const double*)c->x.p;
  a.sy = (const double*)c->y.p;
  a.sz = (const double*)c->z.p;
  a.hole = c->has_hole ? (const int32_t*)c->hole.p : nullptr;
  a.tab = c->tab;
  // Grid over the projected samples, rebuilt when the frame or the cell
  // factor of the search's kind changes.
  if (c->n > 0) {
    const double cf = (a.balances || a.max_per_hole >= 0) ? c->cell_factor_capped : c->cell_factor;
    bool same = c->grid_valid && c->grid_cf == cf && std::memcmp(c->grid_F, F.m, sizeof(F.m)) == 0;
    if (!same) {
      int r = bk::build_grid(c->stream, c->n, a.sx, a.sy, a.sz, F.m, a.origin, cf, c->gb, c->grid);
      if (r) return fail(c, BK_E_HIP, "grid build failed");
      std::memcpy(c->grid_F, F.m, sizeof(F.m));
      c->grid_cf = cf;
      c->grid_valid = true;
      c->gen++;
    }
    a.g = c->grid;
  }
  stride = a.max_samples;
  a.stride = (int)stride;
  return 0;
}

// Runs the search for m targets (device arrays) into ctx scratch.
int run_search(bk_ctx* c, SearchArgs& a, int64_t m, const double* tx, const double* ty, const double* tz,
               const int64_t* left) {
  if (int r = ensure(c, c->nidx, std::max<int64_t>(m * a.stride, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->ncount, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->sstatus, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  a.tx = tx;
  a.ty = ty;
  a.tz = tz;
  a.left_out = left;
  a.m = m;
  a.list = nullptr;
  a.idx = (int32_t*)c->nidx.p;
  a.count = (int32_t*)c->ncount.p;
  a.status = (int32_t*)c->sstatus.p;
  if (m == 0) return 0;
  const void* before[2] = {c->tiles.keys, c->tiles.ids};
  if (bk::launch_search(c->stream, a, m, c->search_mode, c->tiles)) return fail(c, BK_E_HIP, "search launch failed");
  if (before[0] != c->tiles.keys || before[1] != c->tiles.ids) {
    if (c->capturing) return fail(c, BK_E_STATE, "search buffers moved while a pass was captured");
    c->gen++;
  }
  k_map_overflow<<<blocks_of(m), 256, 0, c->stream>>>(m, a.status);
  return 0;
}

// Solves every target with status OK after the search; outputs per target.
int run_solve(bk_ctx* c, const bk_variogram* vg, const bk_search* srch, int32_t kind, double param, int32_t precision,
              int64_t m,
              const double* tx, const double* ty, const double* tz, int stride, const bk_block* blk, double* value,
              double* variance, double* lagrange, double* support, int32_t* n_used, int32_t* status,
              double* weights = nullptr, int32_t* met = nullptr) {
  SolveArgs a;
  std::memset(&a, 0, sizeof(a));
  a.weights = weights;
  dev_variogram(vg, a.v, a.v32);
  a.kind = kind;
  a.param = param;
  a.sx = (const double*)c->x.p;
  a.sy = (const double*)c->y.p;
  a.sz = (const double*)c->z.p;
  a.sv = (const double*)c->v.p;
  a.serr = c->has_err ? (const double*)c->e.p : nullptr;
  a.tx = tx;
  a.ty = ty;
  a.tz = tz;
  a.nidx = (const int32_t*)c->nidx.p;
  a.ncount = (const int32_t*)c->ncount.p;
  a.stride = stride;
  a.value = value;
  a.variance = variance;
  a.lagrange = lagrange;
  a.support = support;
  a.n_used = n_used;
  a.status = status;
  a.max_iter = c->max_iter;
  a.tol = c->tol;
  a.waves = c->solve_waves;
  if (blk) {
    // block.rs:40-70 and py/estimation.rs:947-958: offsets along the block frame.
    int npts = blk->nx * blk->ny * blk->nz;
    std::vector<double> off(3 * npts);
    M3 f = block_frame(blk->rotation[0], blk->rotation[1], blk->rotation[2]);
    double dx = blk->size[0] / blk->nx, dy = blk->size[1] / blk->ny, dz = blk->size[2] / blk->nz;
    double x0 = 0.0 - blk->size[0] / 2.0 + dx / 2.0, y0 = 0.0 - blk->size[1] / 2.0 + dy / 2.0,
           z0 = 0.0 - blk->size[2] / 2.0 + dz / 2.0;
    int k = 0;
    for (int i = 0; i < blk->nx; ++i)
      for (int j = 0; j < blk->ny; ++j)
        for (int l = 0; l < blk->nz; ++l, ++k) {
          double p0 = x0 + i * dx, p1 = y0 + j * dy, p2 = z0 + l * dz;
          // frame' * p, nalgebra order
          for (int d = 0; d < 3; ++d) {
            double t0 = f.m[d] * p0, t1 = f.m[3 + d] * p1, t2 = f.m[6 + d] * p2;
            off[3 * k + d] = (t0 + t1) + t2;
          }
        }
    // Then each offset projected on each frame, for the refined kernels'
    // right-hand sides (a record minus F_f o_p is the lag to point p).
    const int nf = a.v.nf;
    off.resize(3 * (size_t)npts * (1 + nf));
    double* po = off.data() + 3 * npts;
    for (int k2 = 0; k2 < npts; ++k2)
      for (int fr = 0; fr < nf; ++fr) {
        const double* F = a.v.F[fr];
        const double* o = off.data() + 3 * k2;
        for (int d = 0; d < 3; ++d)
          po[(k2 * nf + fr) * 3 + d] = a.v.ident[fr] ? o[d] : (F[3 * d] * o[0] + F[3 * d + 1] * o[1]) + F[3 * d + 2] * o[2];
      }
    if (int r = ensure(c, c->offs, off.size() * sizeof(double))) return r;
    HIPOK(c, hipMemcpyAsync(c->offs.p, off.data(), off.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    a.block = 1;
    a.npts = npts;
    a.offs = (const double*)c->offs.p;
    a.poffs = (const double*)c->offs.p + 3 * npts;
    a.cbb = 0.0;  // set per system: C(B,B) does not depend on the centre, computed once on the host below
    // composite.rs:120-129 on the discretisation centred at the origin.
    {
      // Evaluate on the host with the same arithmetic as the device helpers.
      auto lag = [&](double px, double py, double pz, double qx, double qy, double qz, double* h) {
        double vx = qx - px, vy = qy - py, vz = qz - pz;
        double r0 = 0, r1 = 0, r2 = 0;
        bool rot = false;
        for (int fr = 0; fr < a.v.nf; ++fr) {
          if (a.v.ident[fr]) {
            h[fr] = std::sqrt((vx * vx + vy * vy) + vz * vz);
            continue;
          }
          if (!rot) {
            r0 = (a.v.rot[0] * vx + a.v.rot[1] * vy) + a.v.rot[2] * vz;
            r1 = (a.v.rot[3] * vx + a.v.rot[4] * vy) + a.v.rot[5] * vz;
            r2 = (a.v.rot[6] * vx + a.v.rot[7] * vy) + a.v.rot[8] * vz;
            rot = true;
          }
          double s0 = r0 / a.v.lag_major[fr], s1 = r1 / a.v.lag_semi[fr], s2 = r2 / a.v.lag_minor[fr];
          h[fr] = std::sqrt((s0 * s0 + s1 * s1) + s2 * s2);
        }
      };
      auto shape = [&](int model, double prm, double h, double rr) -> double {
        if (!std::isfinite(h) || h <= 0.0) return 0.0;
        double t = h / rr;
        switch (model) {
          case BK_SPHERICAL:
            return t >= 1.0 ? 1.0 : 1.5 * t - 0.5 * (t * (t * t));
          case BK_EXPONENTIAL:
            return 1.0 - std::exp(-3.0 * t);
          case BK_GAUSSIAN:
            return 1.0 - std::exp(-3.0 * t * t);
          case BK_CUBIC: {
            if (t >= 1.0) return 1.0;
            double t2 = t * t;
            return 7.0 * t2 - (35.0 / 4.0) * (t * t2) + (7.0 / 2.0) * (t * (t2 * t2)) -
                   (3.0 / 4.0) * ((t * t2) * (t2 * t2));
          }
          case BK_PENTASPHERICAL: {
            if (t >= 1.0) return 1.0;
            double t2 = t * t;
            return (15.0 / 8.0) * t - (5.0 / 4.0) * (t * t2) + (3.0 / 8.0) * (t * (t2 * t2));
          }
          case BK_CIRCULAR:
            if (t >= 1.0) return 1.0;
            return 1.0 - (2.0 / kPi) * std::acos(t) + (2.0 * t / kPi) * std::sqrt(1.0 - t * t);
          case BK_SINEHOLE: {
            double cc = kPi * t;
            return 1.0 - std::sin(cc) / cc;
          }
          case BK_MATERN: {
            double nu = prm < 1.0 ? 0.5 : (prm < 2.0 ? 1.5 : 2.5);
            double delta = (std::sqrt(2.0 * nu) * 3.0) * t;
            double ee = std::exp(-delta);
            if (nu == 0.5) return 1.0 - ee;
            if (nu == 1.5) return 1.0 - (1.0 + delta) * ee;
            return 1.0 - (1.0 + delta + delta * delta / 3.0) * ee;
          }
          case BK_POWER:
            return std::pow(t, prm);
        }
        return 0.0;
      };
      auto bcov = [&](const double* h) {
        if (h[a.v.nug_frame] <= 0.0) return a.v.c0 - a.v.nugget;
        double ss = -0.0;
        for (int kk = 0; kk < a.v.S; ++kk)
          ss += a.v.sill[kk] * shape(a.v.model[kk], a.v.param[kk], h[a.v.frame[kk]], a.v.range[kk]);
        return a.v.c0 - (a.v.nugget + ss);
      };
      if (npts == 1) {
        a.cbb = a.v.c0;
      } else {
        double total = -0.0;
        for (int p = 0; p < npts; ++p) {
          double row = -0.0;
          for (int q = 0; q < npts; ++q) {
            double h[BK_MAX_FRAMES];
            lag(off[3 * p], off[3 * p + 1], off[3 * p + 2], off[3 * q], off[3 * q + 1], off[3 * q + 2], h);
            row += bcov(h);
          }
          total += row;
        }
        a.cbb = total / (double)((int64_t)npts * npts);
      }
    }
  }
  // A calibrated search keeps, per target, the fewest leading samples that
  // reach its target quality (the selection was put in balanced order after
  // the search).
  if (srch && srch->calibration && m > 0) {
    CalibArgs k{(const int32_t*)c->sstatus.p, (int32_t*)c->ncount.p, m, srch->calibration, srch->calibration_target,
                srch->min_samples, kind != BK_SIMPLE, met};
    if (bk::launch_calibrate(c->stream, a, k, stride <= 16 ? 16 : (stride <= 32 ? 32 : 64)))
      return fail(c, BK_E_HIP, "calibration launch failed");
  }
  // Bucket the searched targets by neighbour count. The list lengths stay on
  // the device: every solve launch is persistent and reads its count there,
  // so the pass has no host round trip and can be captured as one graph.
  const int nb = 3;
  if (int r = ensure(c, c->lists, std::max<int64_t>(m, 1) * (nb + 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->counts, 8 * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->guard, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->redo, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  int32_t* lists = (int32_t*)c->lists.p;
  int32_t* counts = (int32_t*)c->counts.p;
  // Work-unit counters of the persistent launches: a line per XCD each.
  if (int r = ensure(c, c->work, (nb + 1) * BK_CTR_INTS * sizeof(int32_t))) return r;
  int32_t* work = (int32_t*)c->work.p;
  if (bk::zero32(c->stream, counts, 8) || bk::zero32(c->stream, work, (nb + 1) * BK_CTR_INTS))
    return fail(c, BK_E_HIP, "fill launch failed");
  if (bk::launch_bucket(c->stream, m, (const int32_t*)c->ncount.p, (const int32_t*)c->sstatus.p, status, value,
                        variance, lists, counts, m, c->edges, nb, a.v.stationary))
    return fail(c, BK_E_HIP, "bucket launch failed");
  mark(c, 3);
  a.guard = (int32_t*)c->guard.p;
  a.redo = (int32_t*)c->redo.p;
  a.guard_count = counts + 4;
  a.redo_count = counts + 5;
  a.refine_stats = counts + 6;
  // The refined form covers spherical, exponential, Gaussian, cubic and
  // pentaspherical structures (fast FP64 covariances); a variogram with any
  // other model takes the FP64 kernel in either precision.
  bool refined = precision == BK_FP32_IR;
  for (int k = 0; k < vg->n_structures; ++k) {
    const int md = vg->structures[k].model;
    if (md != BK_SPHERICAL && md != BK_EXPONENTIAL && md != BK_GAUSSIAN && md != BK_CUBIC && md != BK_PENTASPHERICAL)
      refined = false;
  }
  // Bucket b holds n in (edges[b-1], edges[b]]; one that no neighbour count
  // can reach is not launched.
  for (int b = 0; b < nb; ++b) {
    const int lo = b ? c->edges[b - 1] : 0;
    if (c->edges[b] <= lo || lo >= stride) continue;
    a.list = lists + (int64_t)b * m;
    a.n_list = m;
    a.n_list_dev = counts + b;
    a.work = work + b * BK_CTR_INTS;
    const bool ref_b = refined && c->edges[b] <= c->refine_max;
    int r = ref_b ? bk::launch_krige32(c->stream, a, c->edges[b]) : bk::launch_krige64(c->stream, a, c->edges[b]);
    if (r) return fail(c, BK_E_HIP, "solve launch failed");
  }
  mark(c, 4);
  // Systems the FP32 form handed back: FP64 over the redo list.
  if (refined) {
    SolveArgs b = a;
    b.list = a.redo;
    b.n_list = m;
    b.n_list_dev = a.redo_count;
    b.work = work + nb * BK_CTR_INTS;
    if (bk::launch_krige64(c->stream, b, 64)) return fail(c, BK_E_HIP, "solve launch failed");
  }
  // The LU list (n > 64, power models), then the guarded systems (merged
  // coincident samples, block systems the Cholesky rejects) of every solve
  // kernel. Systems above BK_MAX_LU samples take the LU in global memory.
  double* lu_ws = nullptr;
  if (stride > BK_MAX_LU) {
    if (int r = ensure(c, c->lu_big, bk::lu_big_workspace_bytes())) return r;
    lu_ws = (double*)c->lu_big.p;
  }
  if (!a.v.stationary || stride > BK_MAX_CHOL) {
    SolveArgs b = a;
    b.guard = lists + (int64_t)nb * m;
    b.guard_count = counts + nb;
    b.n_list = m;
    if (bk::launch_guarded(c->stream, b, lu_ws)) return fail(c, BK_E_HIP, "guarded launch failed");
  }
  {
    SolveArgs b = a;
    b.n_list = m;
    if (bk::launch_guarded(c->stream, b, lu_ws)) return fail(c, BK_E_HIP, "guarded launch failed");
  }
  mark(c, 5);
  return 0;
}

int finalize(bk_ctx* c, int64_t m, int32_t* status, double* value, double* variance, double* lagrange,
             double* support, int32_t* n_used) {
  if (m == 0) return 0;
  k_fill_nan<<<blocks_of(m), 256, 0, c->stream>>>(m, status, value, variance, lagrange, support, n_used);
  HIPOK(c, hipGetLastError());
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of bk_krige_noalloc: one static device array per GPU (each device
// holds its own copy until the process ends), carved into the buffers of one
// pass of at most BK_STATIC_MAX_TARGETS targets, BK_STATIC_MAX_SAMPLES
// samples and BK_STATIC_MAX_NEIGHBOURS neighbours per target. Host code never
// uses the array's name as an address: bk_create takes each device's address
// once with hipGetSymbolAddress.
namespace {
constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
constexpr int64_t kStM = BK_STATIC_MAX_TARGETS, kStN = BK_STATIC_MAX_SAMPLES, kStK = BK_STATIC_MAX_NEIGHBOURS;
constexpr int64_t kStCells = 4 * kStN + 64;  // the grid's side rule at cell factor 1 (grid.rs:83-106)
constexpr size_t kStScan = 1 << 20;          // hipCUB scan storage over kStCells + 1 counts
struct StaticPlan {
  size_t nidx, ncount, sstatus, status, lists, counts, guard, redo, work, ids, entries, cell_of, px, bounds, start,
      fill, scan, total;
};
constexpr StaticPlan static_plan() {
  StaticPlan p{};
  size_t o = 0;
  auto take = [&o](size_t bytes) {
    const size_t at = o;
    o += al256(bytes);
    return at;
  };
  p.nidx = take(kStM * kStK * 4);
  p.ncount = take(kStM * 4);
  p.sstatus = take(kStM * 4);
  p.status = take(kStM * 4);
  p.lists = take(kStM * 4 * 4);
  p.counts = take(8 * 4);
  p.guard = take(kStM * 4);
  p.redo = take(kStM * 4);
  p.work = take(4 * BK_CTR_INTS * 4);
  p.ids = take((3 * kStM + 1) * 4);
  p.entries = take(kStN * 32);
  p.cell_of = take(kStN * 4);
  p.px = take(3 * kStN * 8);
  p.bounds = take((1024 * 6 + 16) * 8);
  p.start = take((kStCells + 1) * 4);
  p.fill = take((kStCells + 1) * 4);
  p.scan = take(kStScan);
  p.total = o;
  return p;
}
constexpr StaticPlan kStPlan = static_plan();
static_assert(kStPlan.total <= 48u * 1024 * 1024, "static scratch above 48 MiB");

__device__ __align__(256) unsigned char scratch[kStPlan.total];

std::mutex g_scratch_mu;         // guards g_scratch
void* g_scratch[64] = {};        // scratch's address on each device
std::mutex g_scratch_calls[64];  // one bk_krige_noalloc at a time per device (they share its scratch)

// Takes the device's address of scratch once (the device current).
void scratch_setup(int device) {
  if (device < 0 || device >= 64) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p)
    g_scratch[device] = p;
  else
    (void)hipGetLastError();
}
}  // namespace

extern "C" {

const char* bk_version(void) { return "boitata-mi355x 0.1.0 (gfx950)"; }

int bk_create(bk_ctx** out, int device) {
  if (!out) return fail(nullptr, BK_E_INVALID, "no output pointer");
  *out = nullptr;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess)
    return fail(nullptr, BK_E_HIP, std::string("device ") + std::to_string(device) + ": " + hipGetErrorString(e));
  bk_ctx* c = new bk_ctx();
  c->device = device;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(nullptr, BK_E_HIP, std::string("stream: ") + hipGetErrorString(e));
  }
  if (bk::upload_shell_table(c->tab, &c->tab_storage)) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(nullptr, BK_E_HIP, "shell table upload failed");
  }
  scratch_setup(device);
  *out = c;
  return 0;
}

void bk_destroy(bk_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (Buf* b : {&c->x, &c->y, &c->z, &c->v, &c->e, &c->hole, &c->tx, &c->ty, &c->tz, &c->left, &c->nidx, &c->ncount,
                 &c->sstatus, &c->lists, &c->counts, &c->status, &c->value, &c->variance, &c->lagrange, &c->support,
                 &c->n_used, &c->guard, &c->redo, &c->offs, &c->which, &c->iters, &c->work, &c->loo_idx, &c->loo_count,
                 &c->loo_status, &c->loo_lists, &c->mik_gv, &c->mik_gs, &c->mik_par, &c->mik_cls, &c->mik_tab,
                 &c->mik_w, &c->mik_raw, &c->mik_st, &c->mik_cdf, &c->mik_out, &c->lu_big, &c->sgs_x, &c->sgs_y,
                 &c->sgs_z, &c->sgs_e, &c->sgs_h, &c->sgs_v, &c->sgs_s, &c->sgs_done, &c->sgs_limit, &c->sgs_datum,
                 &c->sgs_pos, &c->sgs_ctr, &c->sgs_in, &c->sgs_w, &c->sgs_path, &c->sgs_app, &c->sgs_work, &c->sgs_prog})
    release(*b);
  if (c->sgs) bk_destroy(c->sgs);
  if (c->sgs_tmp) (void)hipFree(c->sgs_tmp);
  bk::free_grid(c->gb);
  bk::free_tiles(c->tiles);
  for (auto& g : c->graph_cache)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (c->tab_storage) (void)hipFree(c->tab_storage);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* bk_last_error(const bk_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int bk_check_search(bk_ctx* c, const bk_search* s) {
  if (!s) return fail(c, BK_E_INVALID, "no search");
  if (const char* m = check_search(s)) return fail(c, BK_E_INVALID, m);
  return 0;
}

int bk_check_variogram(bk_ctx* c, const bk_variogram* v) {
  if (!v) return fail(c, BK_E_INVALID, "no variogram");
  if (const char* m = check_variogram(v)) return fail(c, BK_E_INVALID, m);
  return 0;
}

int bk_set_samples(bk_ctx* c, int64_t n, const double* x, const double* y, const double* z, const double* value,
                   const double* error_var, const int32_t* hole) {
  if (!c) return BK_E_INVALID;
  if (n < 0 || n > (int64_t)2147483647) return fail(c, BK_E_INVALID, "need 0 <= n < 2^31 samples");
  if (n > 0 && (!x || !y || !z || !value)) return fail(c, BK_E_INVALID, "coordinates and values are required");
  HIPOK(c, hipSetDevice(c->device));
  size_t bytes = std::max<int64_t>(n, 1) * sizeof(double);
  for (Buf* b : {&c->x, &c->y, &c->z, &c->v, &c->e})
    if (int r = ensure(c, *b, bytes)) return r;
  if (int r = ensure(c, c->hole, std::max<int64_t>(n, 1) * sizeof(int32_t))) return r;
  auto copy = [&](Buf& b, const void* src, size_t sz) -> int {
    if (n == 0) return 0;
    HIPOK(c, hipMemcpyAsync(b.p, src, n * sz, is_device(src) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            c->stream));
    return 0;
  };
  if (int r = copy(c->x, x, 8)) return r;
  if (int r = copy(c->y, y, 8)) return r;
  if (int r = copy(c->z, z, 8)) return r;
  if (int r = copy(c->v, value, 8)) return r;
  c->has_err = error_var != nullptr;
  if (c->has_err)
    if (int r = copy(c->e, error_var, 8)) return r;
  c->has_hole = hole != nullptr;
  if (c->has_hole)
    if (int r = copy(c->hole, hole, 4)) return r;
  c->n = n;
  c->gen++;
  c->samples_gen++;
  c->grid_valid = false;
  c->loo_ready = false;
  // Origin: the first sample with finite coordinates; planar: one elevation.
  c->has_origin = false;
  c->planar = true;
  if (n > 0) {
    if (int r = ensure(c, c->counts, 8 * sizeof(int32_t))) return r;
    unsigned long long init[2] = {~0ull, 0ull};
    unsigned long long* d = (unsigned long long*)c->counts.p;
    HIPOK(c, hipMemcpyAsync(d, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    k_first_finite<<<blocks_of(n), 256, 0, c->stream>>>(n, (const double*)c->x.p, (const double*)c->y.p,
                                                        (const double*)c->z.p, d, d + 1);
    unsigned long long res[2];
    HIPOK(c, hipMemcpyAsync(res, d, sizeof(res), hipMemcpyDeviceToHost, c->stream));
    HIPOK(c, hipStreamSynchronize(c->stream));
    c->planar = res[1] == 0;
    if (res[0] != ~0ull) {
      c->has_origin = true;
      double o[3];
      HIPOK(c, hipMemcpy(&o[0], (const double*)c->x.p + res[0], 8, hipMemcpyDeviceToHost));
      HIPOK(c, hipMemcpy(&o[1], (const double*)c->y.p + res[0], 8, hipMemcpyDeviceToHost));
      HIPOK(c, hipMemcpy(&o[2], (const double*)c->z.p + res[0], 8, hipMemcpyDeviceToHost));
      std::memcpy(c->origin, o, sizeof(o));
    }
  }
  HIPOK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int bk_neighbors(bk_ctx* c, const bk_search* s, const bk_variogram* vg, int64_t m, const double* tx, const double* ty,
                 const double* tz, const int64_t* left_out, int32_t* idx, int32_t* count, int32_t* status) {
  if (!c || !s) return BK_E_INVALID;
  if (m < 0) return fail(c, BK_E_INVALID, "m < 0");
  HIPOK(c, hipSetDevice(c->device));
  SearchArgs a;
  int64_t stride;
  if (int r = prepare_search(c, s, vg, a, stride)) return r;
  const double *dx, *dy, *dz;
  const int64_t* dl;
  if (int r = stage_in(c, tx, m, c->tx, &dx)) return r;
  if (int r = stage_in(c, ty, m, c->ty, &dy)) return r;
  if (int r = stage_in(c, tz, m, c->tz, &dz)) return r;
  if (int r = stage_in(c, left_out, m, c->left, &dl)) return r;
  if (int r = run_search(c, a, m, dx, dy, dz, dl)) return r;
  // The caller's layout has stride max_samples (not clamped to the sample count).
  int64_t user_stride = s->max_samples;
  if (user_stride == stride) {
    if (is_device(idx))
      HIPOK(c, hipMemcpyAsync(idx, c->nidx.p, m * stride * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
    else
      HIPOK(c, hipMemcpyAsync(idx, c->nidx.p, m * stride * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  } else {
    HIPOK(c, hipMemcpy2DAsync(idx, user_stride * sizeof(int32_t), c->nidx.p, stride * sizeof(int32_t),
                              stride * sizeof(int32_t), m, is_device(idx) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                              c->stream));
  }
  HIPOK(c, hipMemcpyAsync(count, c->ncount.p, m * sizeof(int32_t),
                          is_device(count) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
  HIPOK(c, hipMemcpyAsync(status, c->sstatus.p, m * sizeof(int32_t),
                          is_device(status) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
  HIPOK(c, hipStreamSynchronize(c->stream));
  // A pool overflow is not yet searched by a wider kernel.
  return 0;
}

// The pass on the context's stream, without the closing synchronisation:
// search, bucketing, solves and output copies. It neither allocates nor waits
// on the host when the buffers already fit and the grid is current, which is
// what lets it be captured into a graph.
static int krige_enqueue(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param,
                         const bk_block* blk, int32_t precision, int64_t m, const double* tx, const double* ty,
                         const double* tz, const int64_t* which_loo, bool loo, double* value, double* variance,
                         double* lagrange, double* support, int32_t* n_used, int32_t* status) {
  mark(c, 0);
  SearchArgs a;
  int64_t stride;
  if (int r = prepare_search(c, s, vg, a, stride)) return r;
  const double *dx, *dy, *dz;
  const int64_t* dl = nullptr;
  if (loo) {
    const int64_t* dw = nullptr;
    if (which_loo) {
      if (int r = stage_in(c, which_loo, m, c->which, &dw)) return r;
    } else {
      m = c->n;
    }
    for (Buf* b : {&c->tx, &c->ty, &c->tz})
      if (int r = ensure(c, *b, std::max<int64_t>(m, 1) * sizeof(double))) return r;
    if (int r = ensure(c, c->left, std::max<int64_t>(m, 1) * sizeof(int64_t))) return r;
    if (m > 0)
      k_loo_targets<<<blocks_of(m), 256, 0, c->stream>>>(m, dw, (const double*)c->x.p, (const double*)c->y.p,
                                                         (const double*)c->z.p, (double*)c->tx.p, (double*)c->ty.p,
                                                         (double*)c->tz.p, (int64_t*)c->left.p);
    dx = (const double*)c->tx.p;
    dy = (const double*)c->ty.p;
    dz = (const double*)c->tz.p;
    dl = (const int64_t*)c->left.p;
  } else {
    if (int r = stage_in(c, tx, m, c->tx, &dx)) return r;
    if (int r = stage_in(c, ty, m, c->ty, &dy)) return r;
    if (int r = stage_in(c, tz, m, c->tz, &dz)) return r;
  }
  double *ov, *ovar, *olag = nullptr, *osup = nullptr;
  int32_t *onu = nullptr, *ost;
  if (int r = out_ptr(c, value, m, c->value, &ov)) return r;
  if (int r = out_ptr(c, variance, m, c->variance, &ovar)) return r;
  if (int r = out_ptr(c, lagrange, m, c->lagrange, &olag)) return r;
  if (int r = out_ptr(c, support, m, c->support, &osup)) return r;
  if (int r = out_ptr(c, n_used, m, c->n_used, &onu)) return r;
  if (status) {
    if (int r = out_ptr(c, status, m, c->status, &ost)) return r;
  } else {
    if (int r = ensure(c, c->status, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
    ost = (int32_t*)c->status.p;
  }
  mark(c, 1);
  if (int r = run_search(c, a, m, dx, dy, dz, dl)) return r;
  if (loo && m > 0) k_loo_nonempty<<<blocks_of(m), 256, 0, c->stream>>>(m, (const int32_t*)c->ncount.p,
                                                                       (int32_t*)c->sstatus.p);
  // search.rs:909-936: a calibrated search that balances sectors takes them in turn.
  if (s->calibration && (s->octant || s->sectors > 0) && m > 0)
    if (bk::launch_balance(c->stream, a, m)) return fail(c, BK_E_HIP, "balance launch failed");
  mark(c, 2);
  if (int r = run_solve(c, vg, s, kind, param, precision, m, dx, dy, dz, (int)stride, blk, ov, ovar, olag, osup, onu,
                        ost))
    return r;
  if (int r = finalize(c, m, ost, ov, ovar, olag, osup, onu)) return r;
  if (int r = stage_out(c, value, ov, m)) return r;
  if (int r = stage_out(c, variance, ovar, m)) return r;
  if (int r = stage_out(c, lagrange, olag, m)) return r;
  if (int r = stage_out(c, support, osup, m)) return r;
  if (int r = stage_out(c, n_used, onu, m)) return r;
  if (status)
    if (int r = stage_out(c, status, ost, m)) return r;
  mark(c, 6);
  return 0;
}

// Inputs and state that fix a captured pass: the variogram, search and call
// arguments, the array addresses and the context's generation.
static std::vector<unsigned char> pass_key(const bk_ctx* c, const bk_variogram* vg, const bk_search* s,
                                           int32_t kind, double param, int32_t precision, int64_t m,
                                           const void* const* ptrs, int n_ptrs) {
  std::vector<unsigned char> k;
  auto put = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    k.insert(k.end(), b, b + n);
  };
  put(vg, sizeof(*vg));
  put(s, sizeof(*s));
  put(&kind, sizeof(kind));
  put(&param, sizeof(param));
  put(&precision, sizeof(precision));
  put(&m, sizeof(m));
  put(ptrs, n_ptrs * sizeof(void*));
  put(&c->gen, sizeof(c->gen));
  return k;
}

static int krige_common(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param,
                        const bk_block* blk, int32_t precision, int64_t m, const double* tx, const double* ty,
                        const double* tz, const int64_t* which_loo, bool loo, double* value, double* variance,
                        double* lagrange, double* support, int32_t* n_used, int32_t* status) {
  if (!c || !vg || !s) return BK_E_INVALID;
  if (m < 0) return fail(c, BK_E_INVALID, "m < 0");
  if (kind < BK_ORDINARY || kind > BK_INDICATOR) return fail(c, BK_E_INVALID, "unknown kind");
  if (precision != BK_FP64 && precision != BK_FP32_IR) return fail(c, BK_E_INVALID, "unknown precision");
  if (const char* msg = check_variogram(vg)) return fail(c, BK_E_INVALID, msg);
  bool stationary = true;
  for (int k = 0; k < vg->n_structures; ++k) stationary = stationary && vg->structures[k].model != BK_POWER;
  if (kind == BK_SIMPLE && !stationary)
    return fail(c, BK_E_INVALID,
                "simple kriging needs a variogram with a sill: every structure must be stationary (not power); use "
                "ordinary kriging");
  if (s->calibration == BK_CALIBRATE_SLOPE && kind == BK_SIMPLE)
    return fail(c, BK_E_INVALID, "simple kriging has a slope of regression of 1; calibrate it with target_efficiency");
  if (s->calibration && s->max_samples > BK_MAX_CHOL)
    return fail(c, BK_E_INVALID, "calibrated searches select at most 64 samples on the device");
  if (blk) {
    if (blk->nx <= 0 || blk->ny <= 0 || blk->nz <= 0)
      return fail(c, BK_E_INVALID, "block discretization must be at least 1 per axis");
    for (int d = 0; d < 3; ++d)
      if (!(std::isfinite(blk->size[d]) && blk->size[d] >= 0.0))
        return fail(c, BK_E_INVALID, "block size must be finite and >= 0");
  }
  HIPOK(c, hipSetDevice(c->device));
  // A point-kriging pass over device arrays, repeated with the same inputs,
  // runs as a graph: the first repeat is captured, later ones launch it.
  bool graph_ok = c->graphs && !c->prof && !loo && !blk && m > 0 && c->n > 0 && c->search_mode <= 0 &&
                        is_device(tx) && is_device(ty) && is_device(tz) && is_device(value) &&
                        is_device(variance) && (!lagrange || is_device(lagrange)) &&
                        (!n_used || is_device(n_used)) && (!status || is_device(status));
  const void* ptrs[8] = {tx, ty, tz, value, variance, lagrange, n_used, status};
  bk_ctx::Graph* entry = nullptr;
  if (graph_ok) {
    std::vector<unsigned char> key = pass_key(c, vg, s, kind, param, precision, m, ptrs, 8);
    for (auto& g : c->graph_cache)
      if (g.key == key) entry = &g;
    if (entry && entry->exec) {
      entry->used = ++c->graph_clock;
      HIPOK(c, hipGraphLaunch(entry->exec, c->stream));
      HIPOK(c, hipStreamSynchronize(c->stream));
      c->graph_launches++;
      return 0;
    }
    if (entry) {
      // Second pass with these inputs: capture it.
      hipGraph_t g = nullptr;
      c->capturing = true;
      hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
      int r = e == hipSuccess ? krige_enqueue(c, vg, s, kind, param, blk, precision, m, tx, ty, tz, which_loo, loo,
                                              value, variance, lagrange, support, n_used, status)
                              : BK_E_HIP;
      if (e == hipSuccess) e = hipStreamEndCapture(c->stream, &g);
      c->capturing = false;
      hipGraphExec_t x = nullptr;
      if (r == 0 && e == hipSuccess && g) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      if (r == 0 && e == hipSuccess && x) {
        entry->exec = x;
        entry->used = ++c->graph_clock;
        HIPOK(c, hipGraphLaunch(x, c->stream));
        HIPOK(c, hipStreamSynchronize(c->stream));
        c->graph_launches++;
        return 0;
      }
      // Not capturable here (a runtime that refuses a call made by the pass):
      // run passes as usual from now on.
      if (x) (void)hipGraphExecDestroy(x);
      (void)hipGetLastError();
      c->err.clear();
      (void)bk_set_graphs(c, 0);
      graph_ok = false;
    }
  }
  if (int r = krige_enqueue(c, vg, s, kind, param, blk, precision, m, tx, ty, tz, which_loo, loo, value, variance,
                            lagrange, support, n_used, status))
    return r;
  HIPOK(c, hipStreamSynchronize(c->stream));
  accumulate(c);
  if (graph_ok) {
    // Remember the pass (with the state it left) for a capture next time,
    // in place of the least recently used of at most four.
    const size_t kMax = 4;
    bk_ctx::Graph* slot = nullptr;
    if (c->graph_cache.size() < kMax) {
      c->graph_cache.emplace_back();
      slot = &c->graph_cache.back();
    } else {
      slot = &c->graph_cache[0];
      for (auto& g : c->graph_cache)
        if (g.used < slot->used) slot = &g;
      if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
    }
    slot->key = pass_key(c, vg, s, kind, param, precision, m, ptrs, 8);
    slot->exec = nullptr;
    slot->used = ++c->graph_clock;
  }
  return 0;
}

int bk_krige(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param, int32_t precision,
             int64_t m, const double* tx, const double* ty, const double* tz, double* value, double* variance,
             double* lagrange, int32_t* n_used, int32_t* status) {
  return krige_common(c, vg, s, kind, param, nullptr, precision, m, tx, ty, tz, nullptr, false, value, variance,
                      lagrange, nullptr, n_used, status);
}

int bk_krige_noalloc(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param,
                     int32_t precision, int64_t m, const double* tx, const double* ty, const double* tz, double* value,
                     double* variance, double* lagrange, int32_t* n_used, int32_t* status) {
  if (!c || !vg || !s) return BK_E_INVALID;
  if (m < 0 || m > BK_STATIC_MAX_TARGETS)
    return fail(c, BK_E_INVALID, "bk_krige_noalloc: m must be in [0, BK_STATIC_MAX_TARGETS = 65536]");
  if (c->n > BK_STATIC_MAX_SAMPLES)
    return fail(c, BK_E_INVALID, "bk_krige_noalloc: at most BK_STATIC_MAX_SAMPLES = 262144 samples");
  if (std::min<int64_t>(s->max_samples, std::max<int64_t>(c->n, 1)) > BK_STATIC_MAX_NEIGHBOURS)
    return fail(c, BK_E_INVALID, "bk_krige_noalloc: max_samples above BK_STATIC_MAX_NEIGHBOURS = 64");
  if (s->calibration) return fail(c, BK_E_INVALID, "bk_krige_noalloc: calibrated searches are not served");
  HIPOK(c, hipSetDevice(c->device));
  if (!is_device(tx) || !is_device(ty) || !is_device(tz) || !is_device(value) || !is_device(variance) ||
      (lagrange && !is_device(lagrange)) || (n_used && !is_device(n_used)) || (status && !is_device(status)))
    return fail(c, BK_E_INVALID, "bk_krige_noalloc: targets and outputs must be device arrays");
  if (c->device < 0 || c->device >= 64 || !g_scratch[c->device])
    return fail(c, BK_E_STATE, "bk_krige_noalloc: the static scratch of this device is not available");
  if (bk::grid_scan_bytes(kStCells) > kStScan)
    return fail(c, BK_E_STATE, "bk_krige_noalloc: the grid scan needs more storage than its scratch holds");
  std::lock_guard<std::mutex> lk(g_scratch_calls[c->device]);
  unsigned char* base = (unsigned char*)g_scratch[c->device];
  // The context's buffers and grid give way to views of the scratch for
  // this call, and come back afterwards.
  Buf* bufs[9] = {&c->nidx, &c->ncount, &c->sstatus, &c->status, &c->lists, &c->counts, &c->guard, &c->redo, &c->work};
  const size_t offs[9] = {kStPlan.nidx,  kStPlan.ncount, kStPlan.sstatus, kStPlan.status, kStPlan.lists,
                          kStPlan.counts, kStPlan.guard, kStPlan.redo,   kStPlan.work};
  const size_t caps[9] = {(size_t)kStM * kStK * 4, (size_t)kStM * 4, (size_t)kStM * 4, (size_t)kStM * 4,
                          (size_t)kStM * 16,       32,               (size_t)kStM * 4, (size_t)kStM * 4,
                          4 * BK_CTR_INTS * 4};
  Buf saved[9];
  for (int k = 0; k < 9; ++k) {
    saved[k] = *bufs[k];
    bufs[k]->p = base + offs[k];
    bufs[k]->cap = caps[k];
  }
  bk::GridBuffers gb_saved = c->gb;
  GridDev grid_saved = c->grid;
  const bool grid_valid_saved = c->grid_valid;
  double grid_F_saved[9];
  std::memcpy(grid_F_saved, c->grid_F, sizeof(grid_F_saved));
  const double grid_cf_saved = c->grid_cf;
  bk::TileBuffers tiles_saved = c->tiles;
  const bool graphs_saved = c->graphs;
  bk::GridBuffers& g = c->gb;
  g = bk::GridBuffers();
  g.entries = (double4*)(base + kStPlan.entries);
  g.cell_of = (uint32_t*)(base + kStPlan.cell_of);
  g.px = (double*)(base + kStPlan.px);
  g.bounds = (double*)(base + kStPlan.bounds);
  g.start = (uint32_t*)(base + kStPlan.start);
  g.fill = (uint32_t*)(base + kStPlan.fill);
  g.scan_tmp = base + kStPlan.scan;
  g.cap_n = kStN;
  g.cap_cells = kStCells;
  g.scan_tmp_bytes = kStScan;
  g.fixed = true;
  c->tiles = bk::TileBuffers();
  c->tiles.ids = (int32_t*)(base + kStPlan.ids);
  c->tiles.cap_ids = 3 * kStM + 1;
  c->tiles.fixed = true;
  c->grid_valid = false;
  c->graphs = false;
  c->fixed_buffers = true;
  int r = krige_common(c, vg, s, kind, param, nullptr, precision, m, tx, ty, tz, nullptr, false, value, variance,
                       lagrange, nullptr, n_used, status);
  if (r == 0 && hipStreamSynchronize(c->stream) != hipSuccess) r = fail(c, BK_E_HIP, "bk_krige_noalloc: sync failed");
  c->fixed_buffers = false;
  for (int k = 0; k < 9; ++k) *bufs[k] = saved[k];
  c->gb = gb_saved;
  c->grid = grid_saved;
  c->grid_valid = grid_valid_saved;
  std::memcpy(c->grid_F, grid_F_saved, sizeof(grid_F_saved));
  c->grid_cf = grid_cf_saved;
  c->tiles = tiles_saved;
  c->graphs = graphs_saved;
  c->gen++;
  return r;
}

int bk_block_krige(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param,
                   const bk_block* block, int32_t precision, int64_t m, const double* tx, const double* ty,
                   const double* tz, double* value, double* variance, double* lagrange, double* support,
                   int32_t* n_used, int32_t* status) {
  if (!block) return fail(c, BK_E_INVALID, "no block");
  return krige_common(c, vg, s, kind, param, block, precision, m, tx, ty, tz, nullptr, false, value, variance,
                      lagrange, support, n_used, status);
}

int bk_loo(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param, int32_t precision,
           int64_t m, const int64_t* which, double* value, double* variance, int32_t* status) {
  return krige_common(c, vg, s, kind, param, nullptr, precision, which ? m : (c ? c->n : 0), nullptr, nullptr, nullptr,
                      which, true, value, variance, nullptr, nullptr, nullptr, status);
}

int bk_block_loo(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param,
                 const bk_block* block, int32_t precision, int64_t m, const int64_t* which, double* value,
                 double* variance, int32_t* status) {
  if (!block) return c ? fail(c, BK_E_INVALID, "no block") : BK_E_INVALID;
  return krige_common(c, vg, s, kind, param, block, precision, which ? m : (c ? c->n : 0), nullptr, nullptr, nullptr,
                      which, true, value, variance, nullptr, nullptr, nullptr, status);
}

int bk_loo_prepare(bk_ctx* c, const bk_search* s, const bk_variogram* vg) {
  if (!c || !s) return BK_E_INVALID;
  if (s->calibration)
    return fail(c, BK_E_INVALID, "the leave-one-out gradient needs fixed neighbour sets: a search without calibration");
  HIPOK(c, hipSetDevice(c->device));
  SearchArgs a;
  int64_t stride;
  if (int r = prepare_search(c, s, vg, a, stride)) return r;
  const int64_t m = c->n;
  for (Buf* b : {&c->tx, &c->ty, &c->tz})
    if (int r = ensure(c, *b, std::max<int64_t>(m, 1) * sizeof(double))) return r;
  if (int r = ensure(c, c->left, std::max<int64_t>(m, 1) * sizeof(int64_t))) return r;
  if (m > 0)
    k_loo_targets<<<blocks_of(m), 256, 0, c->stream>>>(m, nullptr, (const double*)c->x.p, (const double*)c->y.p,
                                                       (const double*)c->z.p, (double*)c->tx.p, (double*)c->ty.p,
                                                       (double*)c->tz.p, (int64_t*)c->left.p);
  if (int r = run_search(c, a, m, (const double*)c->tx.p, (const double*)c->ty.p, (const double*)c->tz.p,
                         (const int64_t*)c->left.p))
    return r;
  if (m > 0) k_loo_nonempty<<<blocks_of(m), 256, 0, c->stream>>>(m, (const int32_t*)c->ncount.p, (int32_t*)c->sstatus.p);
  if (int r = ensure(c, c->loo_idx, std::max<int64_t>(m * stride, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->loo_count, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->loo_status, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  if (m > 0) {
    HIPOK(c, hipMemcpyAsync(c->loo_idx.p, c->nidx.p, m * stride * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
    HIPOK(c, hipMemcpyAsync(c->loo_count.p, c->ncount.p, m * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
    HIPOK(c, hipMemcpyAsync(c->loo_status.p, c->sstatus.p, m * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
  }
  // The sets do not move with the parameters, so neither do the buckets.
  const int nb = 3;
  if (int r = ensure(c, c->loo_lists, std::max<int64_t>(m, 1) * (nb + 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->counts, 8 * sizeof(int32_t))) return r;
  for (Buf* b : {&c->value, &c->variance})
    if (int r = ensure(c, *b, std::max<int64_t>(m, 1) * sizeof(double))) return r;
  if (int r = ensure(c, c->status, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  if (bk::zero32(c->stream, (int32_t*)c->counts.p, 8)) return fail(c, BK_E_HIP, "fill launch failed");
  DevVariogram dv;
  DevVariogram32 dv32;
  bool stationary = true;
  if (vg) {
    dev_variogram(vg, dv, dv32);
    stationary = dv.stationary != 0;
  }
  if (bk::launch_bucket(c->stream, m, (const int32_t*)c->loo_count.p, (const int32_t*)c->loo_status.p,
                        (int32_t*)c->status.p, (double*)c->value.p, (double*)c->variance.p,
                        (int32_t*)c->loo_lists.p, (int32_t*)c->counts.p, m, c->edges, nb, stationary))
    return fail(c, BK_E_HIP, "bucket launch failed");
  int32_t hc[8];
  HIPOK(c, hipMemcpyAsync(hc, c->counts.p, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPOK(c, hipStreamSynchronize(c->stream));
  for (int b = 0; b <= nb; ++b) c->loo_hc[b] = hc[b];
  for (int b = 0; b < nb; ++b) c->loo_edges[b] = c->edges[b];
  c->loo_ready = true;
  c->loo_m = m;
  c->loo_stride = (int)stride;
  return 0;
}

int bk_loo_loss_grad(bk_ctx* c, const bk_variogram* vg, int32_t precision, double* loss, double* grad,
                     int64_t* n_used) {
  if (!c || !vg) return BK_E_INVALID;
  if (!c->loo_ready) return fail(c, BK_E_STATE, "call bk_loo_prepare first");
  if (const char* msg = check_variogram(vg)) return fail(c, BK_E_INVALID, msg);
  for (int k = 0; k < vg->n_structures; ++k) {
    int md = vg->structures[k].model;
    if (md != BK_SPHERICAL && md != BK_EXPONENTIAL && md != BK_GAUSSIAN && md != BK_CUBIC && md != BK_PENTASPHERICAL)
      return fail(c, BK_E_INVALID, "the gradient covers spherical, exponential, gaussian, cubic and pentaspherical");
  }
  if (precision != BK_FP64 && precision != BK_FP32_IR) return fail(c, BK_E_INVALID, "unknown precision");
  HIPOK(c, hipSetDevice(c->device));
  const int64_t m = c->loo_m;
  SolveArgs a;
  std::memset(&a, 0, sizeof(a));
  dev_variogram(vg, a.v, a.v32);
  GradArgs gp;
  std::memset(&gp, 0, sizeof(gp));
  gp.S = vg->n_structures;
  gp.P = 4 + 4 * gp.S;
  gp.m = m;
  double az = vg->has_anisotropy ? vg->azimuth : 0.0, dip = vg->has_anisotropy ? vg->dip : 0.0,
         rk = vg->has_anisotropy ? vg->rake : 0.0;
  for (int k = 0; k < gp.S; ++k) {
    const bk_structure& st = vg->structures[k];
    bool own = st.ratio_semi > 0.0 && st.ratio_minor > 0.0;
    double rs = own ? st.ratio_semi : (vg->has_anisotropy ? vg->ratio_semi : 1.0);
    double rm = own ? st.ratio_minor : (vg->has_anisotropy ? vg->ratio_minor : 1.0);
    gp.model[k] = st.model;
    gp.sill[k] = st.sill;
    gp.r[k][0] = st.range;
    gp.r[k][1] = st.range * rs;
    gp.r[k][2] = st.range * rm;
  }
  {
    // R = r3 r2 r1 (rotation.rs:8-16) and its derivative per degree of each angle.
    const double d2r = kPi / 180.0;
    double a1 = az * d2r, a2 = dip * d2r, a3 = rk * d2r;
    double sa = std::sin(a1), ca = std::cos(a1), sd = std::sin(a2), cd = std::cos(a2), sr = std::sin(a3),
           cr = std::cos(a3);
    M3 r1{{sa, ca, 0.0, -ca, sa, 0.0, 0.0, 0.0, 1.0}}, dr1{{ca, -sa, 0.0, sa, ca, 0.0, 0.0, 0.0, 0.0}};
    M3 r2{{cd, 0.0, -sd, 0.0, 1.0, 0.0, sd, 0.0, cd}}, dr2{{-sd, 0.0, -cd, 0.0, 0.0, 0.0, cd, 0.0, -sd}};
    M3 r3{{1.0, 0.0, 0.0, 0.0, cr, sr, 0.0, -sr, cr}}, dr3{{0.0, 0.0, 0.0, 0.0, -sr, cr, 0.0, -cr, -sr}};
    M3 R = mm(mm(r3, r2), r1);
    M3 D[3] = {mm(mm(r3, r2), dr1), mm(mm(r3, dr2), r1), mm(mm(dr3, r2), r1)};
    std::memcpy(gp.R, R.m, sizeof(R.m));
    for (int b = 0; b < 3; ++b)
      for (int q = 0; q < 9; ++q) gp.dR[b][q] = D[b].m[q] * d2r;
  }
  // The FP32 copy for the refined kernel.
  GradArgs32 g32;
  std::memset(&g32, 0, sizeof(g32));
  g32.S = gp.S;
  g32.P = gp.P;
  for (int k = 0; k < gp.S; ++k) {
    g32.model[k] = gp.model[k];
    g32.sill[k] = (float)gp.sill[k];
    for (int q = 0; q < 3; ++q) g32.inv_r[k][q] = (float)(1.0 / gp.r[k][q]);
  }
  for (int q = 0; q < 9; ++q) g32.R[q] = (float)gp.R[q];
  for (int b = 0; b < 3; ++b)
    for (int q = 0; q < 9; ++q) g32.dR[b][q] = (float)gp.dR[b][q];
  a.max_iter = c->max_iter;
  a.tol = c->tol;
  a.waves = c->solve_waves;
  a.kind = BK_ORDINARY;
  a.sx = (const double*)c->x.p;
  a.sy = (const double*)c->y.p;
  a.sz = (const double*)c->z.p;
  a.sv = (const double*)c->v.p;
  a.serr = c->has_err ? (const double*)c->e.p : nullptr;
  a.tx = a.sx;
  a.ty = a.sy;
  a.tz = a.sz;
  a.nidx = (const int32_t*)c->loo_idx.p;
  a.ncount = (const int32_t*)c->loo_count.p;
  a.stride = c->loo_stride;
  const int cols = gp.P + 2;
  for (Buf* b : {&c->value, &c->variance})
    if (int r = ensure(c, *b, std::max<int64_t>(m, 1) * sizeof(double))) return r;
  if (int r = ensure(c, c->status, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->support, std::max<int64_t>(m * cols, 32) * sizeof(double))) return r;
  a.value = (double*)c->value.p;
  a.variance = (double*)c->variance.p;
  a.status = (int32_t*)c->status.p;
  a.grad_part = (double*)c->support.p;
  const int nb = 3;
  if (int r = ensure(c, c->counts, 8 * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->guard, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->redo, std::max<int64_t>(m, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->work, (nb + 1) * BK_CTR_INTS * sizeof(int32_t))) return r;
  HIPOK(c, hipMemsetAsync(c->counts.p, 0, 8 * sizeof(int32_t), c->stream));
  HIPOK(c, hipMemsetAsync(c->work.p, 0, (nb + 1) * BK_CTR_INTS * sizeof(int32_t), c->stream));
  // Rows of samples outside every bucket stay zero (they are out of the loss).
  if (m > 0) HIPOK(c, hipMemsetAsync(c->support.p, 0, m * cols * sizeof(double), c->stream));
  int32_t* counts = (int32_t*)c->counts.p;
  a.guard = (int32_t*)c->guard.p;
  a.guard_count = counts + 4;
  a.redo = (int32_t*)c->redo.p;
  a.redo_count = counts + 5;
  // The buckets of bk_loo_prepare, with the edges they were made with.
  for (int b = 0; b < nb; ++b) {
    if (c->loo_hc[b] == 0) continue;
    a.list = (const int32_t*)c->loo_lists.p + (int64_t)b * m;
    a.n_list = c->loo_hc[b];
    a.work = (int32_t*)c->work.p + b * BK_CTR_INTS;
    const bool ref_b = precision == BK_FP32_IR && c->loo_edges[b] <= c->refine_max;
    const int r = !ref_b                   ? bk::launch_loo_grad(c->stream, a, gp, c->loo_edges[b])
                  : c->loo_edges[b] == 32 ? bk::launch_loo_grad32m(c->stream, a, gp, g32)
                                          : bk::launch_loo_grad32(c->stream, a, gp, g32, c->loo_edges[b]);
    if (r) return fail(c, BK_E_HIP, "gradient launch failed");
  }
  // Systems the refined kernel handed back: the FP64 gradient over the redo
  // list (its length stays on the device).
  if (precision == BK_FP32_IR) {
    SolveArgs b = a;
    b.list = a.redo;
    b.n_list = m;
    b.n_list_dev = a.redo_count;
    b.work = (int32_t*)c->work.p + nb * BK_CTR_INTS;
    if (bk::launch_loo_grad(c->stream, b, gp, 64)) return fail(c, BK_E_HIP, "gradient launch failed");
  }
  // Systems beyond the Cholesky kernels (n > 64) stay out of the loss.
  const int64_t rb = bk::reduce_blocks(m);
  if (int r = ensure(c, c->offs, (64 + cols * rb) * sizeof(double))) return r;
  if (bk::launch_reduce_cols(c->stream, m, cols, a.grad_part, (double*)c->offs.p + 64, (double*)c->offs.p))
    return fail(c, BK_E_HIP, "reduction launch failed");
  double sums[BK_MAX_GRAD + 2];
  HIPOK(c, hipMemcpyAsync(sums, c->offs.p, cols * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPOK(c, hipStreamSynchronize(c->stream));
  double used = sums[1];
  if (n_used) *n_used = (int64_t)used;
  if (used <= 0) return fail(c, BK_E_STATE, "no held-out sample was estimated");
  if (loss) *loss = sums[0] / used;
  if (grad)
    for (int p = 0; p < gp.P; ++p) grad[p] = sums[2 + p] / used;
  return 0;
}

// Tuning knobs (not part of the reference interface): bucket edges, FP32+IR
// sweeps and tolerance, grid cell factor.
}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Multiple indicator kriging (crates/estimation/src/indicator.rs): one search
// with the median threshold's variogram, then per variogram the ordinary
// solves writing their weights, the indicator estimates from those weights,
// and the conditional distributions.

// MultipleIndicator::validate (indicator.rs:291-317).
const char* check_mik(const bk_mik* p, const double* thr) {
  const int k = p->n_thresholds;
  if (k < 1) return "need at least one finite threshold";
  for (int i = 0; i < k; ++i)
    if (!std::isfinite(thr[i])) return "need at least one finite threshold";
  for (int i = 1; i < k; ++i)
    if (thr[i - 1] >= thr[i]) return "thresholds must be strictly increasing";
  if (p->n_variograms != 1 && p->n_variograms != k) return "need one variogram, or one per threshold";
  if (p->has_tails && !(std::isfinite(p->tails[0]) && std::isfinite(p->tails[1]) && p->tails[0] <= p->tails[1]))
    return "tails must be finite with lower <= upper";
  if (p->interpolation != BK_MIK_GLOBAL && p->interpolation != BK_MIK_LINEAR) return "unknown interpolation";
  if (p->upper_tail < BK_TAIL_NONE || p->upper_tail > BK_TAIL_HYPERBOLIC) return "unknown upper tail";
  if (p->upper_tail != BK_TAIL_NONE && !(std::isfinite(p->omega) && p->omega > 0.0))
    return "the upper tail exponent must be > 0";
  if (p->upper_tail == BK_TAIL_HYPERBOLIC && thr[k - 1] <= 0.0)
    return "a hyperbolic upper tail needs a last threshold > 0";
  return nullptr;
}

// A host copy of a parameter array given in host or device memory.
template <class T>
int host_array(bk_ctx* c, const T* src, int64_t n, std::vector<T>& out) {
  out.resize(std::max<int64_t>(n, 0));
  if (n <= 0 || !src) return 0;
  if (is_device(src)) {
    HIPOK(c, hipMemcpy(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
  } else {
    std::memcpy(out.data(), src, n * sizeof(T));
  }
  return 0;
}

// f64::total_cmp as an integer key.
inline int64_t total_order_key(double v) {
  int64_t b;
  std::memcpy(&b, &v, sizeof(b));
  return b ^ (int64_t)((uint64_t)(b >> 63) >> 1);
}

// Global::new (indicator.rs:152-200) of the fitted samples with declustering
// weights w (empty: all 1), kept in the context while its inputs repeat.
int mik_global(bk_ctx* c, const bk_mik* p, const std::vector<double>& thr, const std::vector<double>& w) {
  std::vector<unsigned char> key;
  auto put = [&](const void* q, size_t n) {
    const unsigned char* b = (const unsigned char*)q;
    key.insert(key.end(), b, b + n);
  };
  put(&c->samples_gen, sizeof(c->samples_gen));
  put(thr.data(), thr.size() * sizeof(double));
  put(&p->has_tails, sizeof(p->has_tails));
  put(p->tails, sizeof(p->tails));
  const int64_t nw = (int64_t)w.size();
  put(&nw, sizeof(nw));
  put(w.data(), w.size() * sizeof(double));
  if (key == c->mik_key) return 0;
  c->mik_key.clear();
  const int64_t n = c->n;
  std::vector<double> values(n);
  HIPOK(c, hipMemcpy(values.data(), c->v.p, n * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> wt = w.empty() ? std::vector<double>(n, 1.0) : w;
  for (double x : wt)
    if (!std::isfinite(x) || x < 0.0) return fail(c, BK_E_INVALID, "weights must be finite and >= 0");
  double total = -0.0;
  for (double x : wt) total += x;
  if (total <= 0.0) return fail(c, BK_E_INVALID, "weights must not all be 0");
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int64_t i, int64_t j) {
    const int64_t a = total_order_key(values[i]), b = total_order_key(values[j]);
    return a != b ? a < b : i < j;
  });
  std::vector<double>& v = c->mik_values;
  std::vector<double>& sums = c->mik_sums;
  v.resize(n);
  sums.assign(3 * (n + 1), 0.0);
  for (int64_t k = 0; k < n; ++k) v[k] = values[order[k]];
  for (int64_t k = 0; k < n; ++k) {
    const double wk = wt[order[k]] / total;
    for (int j = 0; j < 3; ++j) sums[j * (n + 1) + k + 1] = sums[j * (n + 1) + k] + wk * mik_powi(v[k], j);
  }
  const int K = (int)thr.size();
  c->mik_bounds.assign(1, 0);
  c->mik_prop.resize(K);
  for (int k = 0; k < K; ++k) {
    const int64_t cut = mik_partition(v.data(), 0, n, [&](double x) { return x <= thr[k]; });
    c->mik_bounds.push_back(cut);
    c->mik_prop[k] = sums[cut];
  }
  c->mik_bounds.push_back(n);
  const double lo = p->has_tails ? p->tails[0] : v.front(), hi = p->has_tails ? p->tails[1] : v.back();
  c->mik_lower = std::fmin(std::fmin(lo, v.front()), thr.front());
  c->mik_upper = std::fmax(std::fmax(hi, v.back()), thr.back());
  if (int r = ensure(c, c->mik_gv, std::max<int64_t>(n, 1) * sizeof(double))) return r;
  if (int r = ensure(c, c->mik_gs, sums.size() * sizeof(double))) return r;
  HIPOK(c, hipMemcpyAsync(c->mik_gv.p, v.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPOK(c, hipMemcpyAsync(c->mik_gs.p, sums.data(), sums.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPOK(c, hipStreamSynchronize(c->stream));
  c->mik_key = key;
  return 0;
}

int mik_common(bk_ctx* c, const bk_mik* p, const double* thresholds, const bk_variogram* vgs,
                      const double* weights, const bk_search* s, const bk_block* blk, int32_t precision, int64_t m,
                      const double* tx, const double* ty, const double* tz, bool loo, const int64_t* which,
                      int32_t nc, const double* cutoffs, int32_t nq, const double* quantiles, double* pit,
                      const bk_mik_out* out) {
  if (!c) return BK_E_INVALID;
  if (!p || !thresholds || !vgs || !s || !out) return fail(c, BK_E_INVALID, "null argument");
  if (m < 0) return fail(c, BK_E_INVALID, "m < 0");
  if (nc < 0 || nq < 0 || (nc > 0 && !cutoffs) || (nq > 0 && !quantiles))
    return fail(c, BK_E_INVALID, "cutoffs or quantiles missing");
  if (precision != BK_FP64 && precision != BK_FP32_IR) return fail(c, BK_E_INVALID, "unknown precision");
  HIPOK(c, hipSetDevice(c->device));
  std::vector<double> thr, cut, qs, w;
  if (int r = host_array(c, thresholds, (int64_t)p->n_thresholds, thr)) return r;
  if (const char* msg = check_mik(p, thr.data())) return fail(c, BK_E_INVALID, msg);
  if (int r = host_array(c, cutoffs, (int64_t)nc, cut)) return r;
  if (int r = host_array(c, quantiles, (int64_t)nq, qs)) return r;
  for (double x : cut)
    if (!std::isfinite(x)) return fail(c, BK_E_INVALID, "cutoffs must be finite");
  for (double x : qs)
    if (!(x >= 0.0 && x <= 1.0)) return fail(c, BK_E_INVALID, "quantiles must be in [0, 1]");
  if (const char* msg = check_search(s)) return fail(c, BK_E_INVALID, msg);
  if (s->calibration)
    return fail(c, BK_E_INVALID, "multiple indicator kriging takes a search without target_slope or target_efficiency");
  const int K = p->n_thresholds, V = p->n_variograms;
  bool stationary = true;
  for (int v = 0; v < V; ++v) {
    if (const char* msg = check_variogram(&vgs[v])) return fail(c, BK_E_INVALID, msg);
    for (int k = 0; k < vgs[v].n_structures; ++k) stationary = stationary && vgs[v].structures[k].model != BK_POWER;
  }
  if (blk) {
    if (blk->nx <= 0 || blk->ny <= 0 || blk->nz <= 0)
      return fail(c, BK_E_INVALID, "block discretization must be at least 1 per axis");
    for (int d = 0; d < 3; ++d)
      if (!(std::isfinite(blk->size[d]) && blk->size[d] >= 0.0))
        return fail(c, BK_E_INVALID, "block size must be finite and >= 0");
  }
  if (c->n == 0) return fail(c, BK_E_STATE, "no samples");
  if (weights)
    if (int r = host_array(c, weights, c->n, w)) return r;
  if (int r = mik_global(c, p, thr, w)) return r;
  // The classes (indicator.rs:347-378) and their moments above each shared
  // cutoff (column 0: -inf), which no target changes.
  const int CC = nc + 1;
  std::vector<MikClass> cls(K + 1);
  const MikGlobal hg{c->mik_values.data(), c->mik_sums.data(), c->n};
  for (int k = 0; k <= K; ++k) {
    MikClass& q = cls[k];
    std::memset(&q, 0, sizeof(q));
    q.a = k == 0 ? c->mik_lower : thr[k - 1];
    q.b = k == K ? c->mik_upper : thr[k];
    q.s = c->mik_bounds[k];
    q.e = c->mik_bounds[k + 1];
    q.w = 1.0;
    const double weight = c->mik_sums[q.e] - c->mik_sums[q.s];
    if (k == K && p->upper_tail == BK_TAIL_POWER) {
      q.kind = BK_SHAPE_POWER;
      q.w = p->omega;
    } else if (k == K && p->upper_tail == BK_TAIL_HYPERBOLIC) {
      q.kind = BK_SHAPE_HYPERBOLIC;
      q.w = p->omega;
    } else if (p->interpolation == BK_MIK_GLOBAL && weight > 0.0) {
      q.kind = BK_SHAPE_DATA;
    } else {
      q.kind = BK_SHAPE_POWER;
    }
  }
  std::vector<double> tab((size_t)(K + 1) * CC * 3);
  for (int k = 0; k <= K; ++k)
    for (int ci = 0; ci < CC; ++ci) {
      const double at = ci == 0 ? -std::numeric_limits<double>::infinity() : cut[ci - 1];
      for (int j = 0; j < 3; ++j) tab[((size_t)k * CC + ci) * 3 + j] = mik_above(hg, cls[k], at, j);
    }
  std::vector<double> par(2 * K + nq);
  std::copy(thr.begin(), thr.end(), par.begin());
  std::copy(c->mik_prop.begin(), c->mik_prop.end(), par.begin() + K);
  std::copy(qs.begin(), qs.end(), par.begin() + 2 * K);
  if (int r = ensure(c, c->mik_par, par.size() * sizeof(double))) return r;
  if (int r = ensure(c, c->mik_cls, cls.size() * sizeof(MikClass))) return r;
  if (int r = ensure(c, c->mik_tab, tab.size() * sizeof(double))) return r;
  HIPOK(c, hipMemcpyAsync(c->mik_par.p, par.data(), par.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPOK(c, hipMemcpyAsync(c->mik_cls.p, cls.data(), cls.size() * sizeof(MikClass), hipMemcpyHostToDevice, c->stream));
  HIPOK(c, hipMemcpyAsync(c->mik_tab.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const double* dthr = (const double*)c->mik_par.p;
  // Targets and the search of the median threshold's variogram (indicator.rs:319-321).
  SearchArgs a;
  int64_t stride;
  if (int r = prepare_search(c, s, &vgs[V / 2], a, stride)) return r;
  const double *dx, *dy, *dz;
  const int64_t* dl = nullptr;
  if (loo) {
    const int64_t* dw = nullptr;
    if (which) {
      if (int r = stage_in(c, which, m, c->which, &dw)) return r;
    } else {
      m = c->n;
    }
    for (Buf* b : {&c->tx, &c->ty, &c->tz})
      if (int r = ensure(c, *b, std::max<int64_t>(m, 1) * sizeof(double))) return r;
    if (int r = ensure(c, c->left, std::max<int64_t>(m, 1) * sizeof(int64_t))) return r;
    if (m > 0)
      k_loo_targets<<<blocks_of(m), 256, 0, c->stream>>>(m, dw, (const double*)c->x.p, (const double*)c->y.p,
                                                         (const double*)c->z.p, (double*)c->tx.p, (double*)c->ty.p,
                                                         (double*)c->tz.p, (int64_t*)c->left.p);
    dx = (const double*)c->tx.p;
    dy = (const double*)c->ty.p;
    dz = (const double*)c->tz.p;
    dl = (const int64_t*)c->left.p;
  } else {
    if (m > 0 && (!tx || !ty || !tz)) return fail(c, BK_E_INVALID, "targets missing");
    if (int r = stage_in(c, tx, m, c->tx, &dx)) return r;
    if (int r = stage_in(c, ty, m, c->ty, &dy)) return r;
    if (int r = stage_in(c, tz, m, c->tz, &dz)) return r;
  }
  if (int r = run_search(c, a, m, dx, dy, dz, dl)) return r;
  if (loo && m > 0)
    k_loo_nonempty<<<blocks_of(m), 256, 0, c->stream>>>(m, (const int32_t*)c->ncount.p, (int32_t*)c->sstatus.p);
  // Outputs: device arrays in place; host arrays through one scratch block.
  const int64_t M = std::max<int64_t>(m, 1);
  const int64_t dbl = (int64_t)K * M * 2 + M * 4 + (int64_t)nc * M * 2 + (int64_t)nq * M, i32 = 3 * M;
  if (int r = ensure(c, c->mik_out, dbl * sizeof(double) + i32 * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->mik_w, M * stride * sizeof(double))) return r;
  if (int r = ensure(c, c->mik_st, 2 * M * sizeof(int32_t))) return r;
  for (Buf* b : {&c->value, &c->variance})
    if (int r = ensure(c, *b, M * sizeof(double))) return r;
  double* dpos = (double*)c->mik_out.p;
  int32_t* ipos = (int32_t*)(dpos + dbl);
  auto place_d = [&](double* user, int64_t n, bool need) -> double* {
    double* own = dpos;
    dpos += n;
    if (user && is_device(user)) return user;
    return user || need ? own : nullptr;
  };
  auto place_i = [&](int32_t* user, bool need) -> int32_t* {
    int32_t* own = ipos;
    ipos += M;
    if (user && is_device(user)) return user;
    return user || need ? own : nullptr;
  };
  double* d_raw = place_d(out->raw, (int64_t)K * M, true);
  double* d_cdf = place_d(out->cdf, (int64_t)K * M, true);
  double* d_corr = place_d(out->correction, M, false);
  double* d_mean = place_d(out->mean, M, false);
  double* d_var = place_d(out->variance, M, false);
  double* d_pit = place_d(pit, M, false);
  double* d_pa = place_d(out->probability_above, (int64_t)nc * M, false);
  double* d_ma = place_d(out->mean_above, (int64_t)nc * M, false);
  double* d_qv = place_d(out->quantile_values, (int64_t)nq * M, false);
  int32_t* d_viol = place_i(out->violations, false);
  int32_t* d_nu = place_i(out->n_used, false);
  int32_t* d_st = place_i(out->status, true);
  int32_t* st_v = (int32_t*)c->mik_st.p;
  const int kind = p->simple ? BK_SIMPLE : BK_ORDINARY;
  if (m > 0) {
    if (p->simple && !stationary) {
      // krige.rs:117-124: simple kriging needs a sill; every searched target fails.
      if (bk::launch_mik_status(c->stream, m, (const int32_t*)c->sstatus.p, BK_INVALID_PARAMETERS, d_st))
        return fail(c, BK_E_HIP, "status launch failed");
    } else {
      for (int v = 0; v < V; ++v) {
        if (int r = run_solve(c, &vgs[v], nullptr, kind, 0.0, precision, m, dx, dy, dz, (int)stride, blk,
                              (double*)c->value.p, (double*)c->variance.p, nullptr, nullptr, nullptr, st_v,
                              (double*)c->mik_w.p))
          return r;
        MikRawArgs ra;
        ra.m = m;
        ra.v = v;
        ra.V = V;
        ra.K = K;
        ra.simple = p->simple ? 1 : 0;
        ra.thr = dthr;
        ra.prop = dthr + K;
        ra.sv = (const double*)c->v.p;
        ra.nidx = (const int32_t*)c->nidx.p;
        ra.ncount = (const int32_t*)c->ncount.p;
        ra.stride = (int)stride;
        ra.w = (const double*)c->mik_w.p;
        ra.st_v = st_v;
        ra.status = d_st;
        ra.raw = d_raw;
        if (bk::launch_mik_raw(c->stream, ra)) return fail(c, BK_E_HIP, "indicator launch failed");
      }
    }
    MikCondArgs ca;
    std::memset(&ca, 0, sizeof(ca));
    ca.m = m;
    ca.K = K;
    ca.C = nc;
    ca.Q = nq;
    ca.raw = d_raw;
    ca.status = d_st;
    ca.ncount = (const int32_t*)c->ncount.p;
    ca.cls = (const MikClass*)c->mik_cls.p;
    ca.table = (const double*)c->mik_tab.p;
    ca.quantiles = dthr + 2 * K;
    ca.g = MikGlobal{(const double*)c->mik_gv.p, (const double*)c->mik_gs.p, c->n};
    ca.upper = c->mik_upper;
    ca.own_values = loo ? (const double*)c->v.p : nullptr;
    ca.held = dl;
    ca.cdf = d_cdf;
    ca.correction = d_corr;
    ca.mean = d_mean;
    ca.variance = d_var;
    ca.pa = d_pa;
    ca.ma = d_ma;
    ca.qv = d_qv;
    ca.violations = d_viol;
    ca.n_used = d_nu;
    ca.pit = loo ? d_pit : nullptr;
    if (bk::launch_mik_conditional(c->stream, ca)) return fail(c, BK_E_HIP, "conditional launch failed");
  }
  if (int r = stage_out(c, out->raw, d_raw, (int64_t)K * m)) return r;
  if (int r = stage_out(c, out->cdf, d_cdf, (int64_t)K * m)) return r;
  if (int r = stage_out(c, out->correction, d_corr, m)) return r;
  if (int r = stage_out(c, out->mean, d_mean, m)) return r;
  if (int r = stage_out(c, out->variance, d_var, m)) return r;
  if (int r = stage_out(c, pit, d_pit, m)) return r;
  if (int r = stage_out(c, out->probability_above, d_pa, (int64_t)nc * m)) return r;
  if (int r = stage_out(c, out->mean_above, d_ma, (int64_t)nc * m)) return r;
  if (int r = stage_out(c, out->quantile_values, d_qv, (int64_t)nq * m)) return r;
  if (int r = stage_out(c, out->violations, d_viol, m)) return r;
  if (int r = stage_out(c, out->n_used, d_nu, m)) return r;
  if (int r = stage_out(c, out->status, d_st, m)) return r;
  HIPOK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// Diagnostics of a kriging pass (py/estimation.rs:895-904,1034-1105): the
// ordinary pass with the solves writing their weights and the calibration its
// outcome, then the statistics of the samples used.
int diag_common(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param,
                const bk_block* blk, int32_t precision, int64_t m, const double* tx, const double* ty,
                const double* tz, const bk_diag_out* out) {
  if (!c) return BK_E_INVALID;
  if (!vg || !s || !out) return fail(c, BK_E_INVALID, "null argument");
  if (m < 0) return fail(c, BK_E_INVALID, "m < 0");
  if (m > 0 && (!tx || !ty || !tz)) return fail(c, BK_E_INVALID, "targets missing");
  if (kind < BK_ORDINARY || kind > BK_INDICATOR) return fail(c, BK_E_INVALID, "unknown kind");
  if (precision != BK_FP64 && precision != BK_FP32_IR) return fail(c, BK_E_INVALID, "unknown precision");
  if (const char* msg = check_variogram(vg)) return fail(c, BK_E_INVALID, msg);
  bool stationary = true;
  for (int k = 0; k < vg->n_structures; ++k) stationary = stationary && vg->structures[k].model != BK_POWER;
  if (kind == BK_SIMPLE && !stationary)
    return fail(c, BK_E_INVALID,
                "simple kriging needs a variogram with a sill: every structure must be stationary (not power); use "
                "ordinary kriging");
  if (s->calibration == BK_CALIBRATE_SLOPE && kind == BK_SIMPLE)
    return fail(c, BK_E_INVALID, "simple kriging has a slope of regression of 1; calibrate it with target_efficiency");
  if (s->calibration && s->max_samples > BK_MAX_CHOL)
    return fail(c, BK_E_INVALID, "calibrated searches select at most 64 samples on the device");
  if (blk) {
    if (blk->nx <= 0 || blk->ny <= 0 || blk->nz <= 0)
      return fail(c, BK_E_INVALID, "block discretization must be at least 1 per axis");
    for (int d = 0; d < 3; ++d)
      if (!(std::isfinite(blk->size[d]) && blk->size[d] >= 0.0))
        return fail(c, BK_E_INVALID, "block size must be finite and >= 0");
  }
  HIPOK(c, hipSetDevice(c->device));
  SearchArgs a;
  int64_t stride;
  if (int r = prepare_search(c, s, vg, a, stride)) return r;
  const double *dx, *dy, *dz;
  if (int r = stage_in(c, tx, m, c->tx, &dx)) return r;
  if (int r = stage_in(c, ty, m, c->ty, &dy)) return r;
  if (int r = stage_in(c, tz, m, c->tz, &dz)) return r;
  const int64_t M = std::max<int64_t>(m, 1);
  double *ov, *ovar, *olag, *osup, *onws, *omd;
  int32_t *onu, *ost, *oh, *omet;
  if (int r = out_ptr(c, out->value, m, c->value, &ov)) return r;
  if (int r = out_ptr(c, out->variance, m, c->variance, &ovar)) return r;
  if (int r = out_ptr(c, out->lagrange, m, c->lagrange, &olag)) return r;
  if (int r = out_ptr(c, out->support_variance, m, c->support, &osup)) return r;
  if (int r = out_ptr(c, out->n_used, m, c->n_used, &onu)) return r;
  // The value, variance and status always run through device memory; the rest
  // of the host outputs through one scratch block.
  if (!ov) {
    if (int r = ensure(c, c->value, M * sizeof(double))) return r;
    ov = (double*)c->value.p;
  }
  if (!ovar) {
    if (int r = ensure(c, c->variance, M * sizeof(double))) return r;
    ovar = (double*)c->variance.p;
  }
  if (int r = ensure(c, c->mik_out, 2 * M * sizeof(double) + 3 * M * sizeof(int32_t))) return r;
  double* dscr = (double*)c->mik_out.p;
  int32_t* iscr = (int32_t*)(dscr + 2 * M);
  auto place_d = [&](double* user, double* own) { return user && is_device(user) ? user : (user ? own : nullptr); };
  auto place_i = [&](int32_t* user, int32_t* own, bool need) {
    return user && is_device(user) ? user : (user || need ? own : nullptr);
  };
  onws = place_d(out->negative_weight_sum, dscr);
  omd = place_d(out->mean_distance, dscr + M);
  ost = place_i(out->status, iscr, true);
  oh = place_i(out->n_holes, iscr + M, false);
  omet = place_i(out->target_met, iscr + 2 * M, false);
  if (int r = ensure(c, c->mik_w, M * stride * sizeof(double))) return r;
  if (int r = run_search(c, a, m, dx, dy, dz, nullptr)) return r;
  if (s->calibration && (s->octant || s->sectors > 0) && m > 0)
    if (bk::launch_balance(c->stream, a, m)) return fail(c, BK_E_HIP, "balance launch failed");
  // -1 where the search is not calibrated or the target is not solved: the
  // calibration writes 0 / 1 for the targets it sees, the statistics kernel
  // -1 for those that failed.
  if (omet && m > 0)
    if (bk::launch_fill_i32(c->stream, omet, m, -1)) return fail(c, BK_E_HIP, "fill launch failed");
  if (int r = run_solve(c, vg, s, kind, param, precision, m, dx, dy, dz, (int)stride, blk, ov, ovar, olag, osup, onu,
                        ost, (double*)c->mik_w.p, s->calibration ? omet : nullptr))
    return r;
  if (int r = finalize(c, m, ost, ov, ovar, olag, osup, onu)) return r;
  DiagArgs da;
  std::memset(&da, 0, sizeof(da));
  da.m = m;
  da.stride = (int)stride;
  da.nidx = (const int32_t*)c->nidx.p;
  da.ncount = (const int32_t*)c->ncount.p;
  da.status = ost;
  da.tx = dx;
  da.ty = dy;
  da.tz = dz;
  da.sx = (const double*)c->x.p;
  da.sy = (const double*)c->y.p;
  da.sz = (const double*)c->z.p;
  da.hole = c->has_hole ? (const int32_t*)c->hole.p : nullptr;
  da.w = (const double*)c->mik_w.p;
  da.negative_weight_sum = onws;
  da.n_holes = oh;
  da.mean_distance = omd;
  da.target_met = omet;
  if (bk::launch_used_stats(c->stream, da)) return fail(c, BK_E_HIP, "diagnostics launch failed");
  if (int r = stage_out(c, out->value, ov, m)) return r;
  if (int r = stage_out(c, out->variance, ovar, m)) return r;
  if (int r = stage_out(c, out->lagrange, olag, m)) return r;
  if (int r = stage_out(c, out->support_variance, osup, m)) return r;
  if (int r = stage_out(c, out->n_used, onu, m)) return r;
  if (int r = stage_out(c, out->status, ost, m)) return r;
  if (int r = stage_out(c, out->negative_weight_sum, onws, m)) return r;
  if (int r = stage_out(c, out->mean_distance, omd, m)) return r;
  if (int r = stage_out(c, out->n_holes, oh, m)) return r;
  if (int r = stage_out(c, out->target_met, omet, m)) return r;
  HIPOK(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace

extern "C" {

int bk_krige_diagnostics(bk_ctx* c, const bk_variogram* vg, const bk_search* search, int32_t kind, double param,
                         const bk_block* block, int32_t precision, int64_t m, const double* tx, const double* ty,
                         const double* tz, const bk_diag_out* out) {
  return diag_common(c, vg, search, kind, param, block, precision, m, tx, ty, tz, out);
}

int bk_krige_weights(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param,
                     const bk_block* blk, int32_t precision, int64_t m, const double* tx, const double* ty,
                     const double* tz, int32_t* idx, double* weights, int32_t* count, int32_t* status) {
  if (!c) return BK_E_INVALID;
  if (!vg || !s || !idx || !weights || !count) return fail(c, BK_E_INVALID, "null argument");
  if (m < 0) return fail(c, BK_E_INVALID, "m < 0");
  if (m > 0 && (!tx || !ty || !tz)) return fail(c, BK_E_INVALID, "targets missing");
  if (kind < BK_ORDINARY || kind > BK_INDICATOR) return fail(c, BK_E_INVALID, "unknown kind");
  if (precision != BK_FP64 && precision != BK_FP32_IR) return fail(c, BK_E_INVALID, "unknown precision");
  if (const char* msg = check_variogram(vg)) return fail(c, BK_E_INVALID, msg);
  for (int k = 0; k < vg->n_structures; ++k)
    if (kind == BK_SIMPLE && vg->structures[k].model == BK_POWER)
      return fail(c, BK_E_INVALID, "simple kriging needs a variogram with a sill: every structure must be stationary");
  if (s->calibration == BK_CALIBRATE_SLOPE && kind == BK_SIMPLE)
    return fail(c, BK_E_INVALID, "simple kriging has a slope of regression of 1; calibrate it with target_efficiency");
  if (s->calibration && s->max_samples > BK_MAX_CHOL)
    return fail(c, BK_E_INVALID, "calibrated searches select at most 64 samples on the device");
  if (blk) {
    if (blk->nx <= 0 || blk->ny <= 0 || blk->nz <= 0)
      return fail(c, BK_E_INVALID, "block discretization must be at least 1 per axis");
    for (int d = 0; d < 3; ++d)
      if (!(std::isfinite(blk->size[d]) && blk->size[d] >= 0.0))
        return fail(c, BK_E_INVALID, "block size must be finite and >= 0");
  }
  HIPOK(c, hipSetDevice(c->device));
  SearchArgs a;
  int64_t stride;
  if (int r = prepare_search(c, s, vg, a, stride)) return r;
  const double *dx, *dy, *dz;
  if (int r = stage_in(c, tx, m, c->tx, &dx)) return r;
  if (int r = stage_in(c, ty, m, c->ty, &dy)) return r;
  if (int r = stage_in(c, tz, m, c->tz, &dz)) return r;
  const int64_t M = std::max<int64_t>(m, 1);
  for (Buf* b : {&c->value, &c->variance})
    if (int r = ensure(c, *b, M * sizeof(double))) return r;
  double *ow, *ov = (double*)c->value.p, *ovar = (double*)c->variance.p;
  int32_t *oc, *ost;
  if (int r = out_ptr(c, weights, m * stride, c->mik_w, &ow)) return r;
  if (int r = out_ptr(c, count, m, c->n_used, &oc)) return r;
  if (status) {
    if (int r = out_ptr(c, status, m, c->status, &ost)) return r;
  } else {
    if (int r = ensure(c, c->status, M * sizeof(int32_t))) return r;
    ost = (int32_t*)c->status.p;
  }
  if (int r = run_search(c, a, m, dx, dy, dz, nullptr)) return r;
  if (s->calibration && (s->octant || s->sectors > 0) && m > 0)
    if (bk::launch_balance(c->stream, a, m)) return fail(c, BK_E_HIP, "balance launch failed");
  if (int r = run_solve(c, vg, s, kind, param, precision, m, dx, dy, dz, (int)stride, blk, ov, ovar, nullptr, nullptr,
                        oc, ost, ow))
    return r;
  if (int r = finalize(c, m, ost, ov, ovar, nullptr, nullptr, oc)) return r;
  // The samples of each target in the order the estimator received them.
  if (m > 0)
    HIPOK(c, hipMemcpyAsync(idx, c->nidx.p, m * stride * sizeof(int32_t), hipMemcpyDefault, c->stream));
  if (int r = stage_out(c, weights, ow, m * stride)) return r;
  if (int r = stage_out(c, count, oc, m)) return r;
  if (status)
    if (int r = stage_out(c, status, ost, m)) return r;
  HIPOK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int bk_mik_predict(bk_ctx* c, const bk_mik* mik, const double* thresholds, const bk_variogram* variograms,
                   const double* weights, const bk_search* search, const bk_block* block, int32_t precision,
                   int64_t m, const double* tx, const double* ty, const double* tz, int32_t n_cutoffs,
                   const double* cutoffs, int32_t n_quantiles, const double* quantiles, const bk_mik_out* out) {
  return mik_common(c, mik, thresholds, variograms, weights, search, block, precision, m, tx, ty, tz, false, nullptr,
                    n_cutoffs, cutoffs, n_quantiles, quantiles, nullptr, out);
}

int bk_mik_cross_validate(bk_ctx* c, const bk_mik* mik, const double* thresholds, const bk_variogram* variograms,
                          const double* weights, const bk_search* search, int32_t precision, int64_t m,
                          const int64_t* which, double* pit, const bk_mik_out* out) {
  if (!out) return c ? fail(c, BK_E_INVALID, "null argument") : BK_E_INVALID;
  bk_mik_out o = *out;
  o.probability_above = o.mean_above = o.quantile_values = nullptr;
  return mik_common(c, mik, thresholds, variograms, weights, search, nullptr, precision, which ? m : (c ? c->n : 0),
                    nullptr, nullptr, nullptr, true, which, 0, nullptr, 0, nullptr, pit, &o);
}

int bk_set_tuning(bk_ctx* c, const int32_t* edges3, int32_t max_iter, double tol, double cell_factor) {
  if (!c) return BK_E_INVALID;
  if (edges3) {
    for (int k = 0; k < 3; ++k)
      if (edges3[k] != 16 && edges3[k] != 32 && edges3[k] != 64)
        return fail(c, BK_E_INVALID, "bucket edges are 16, 32 or 64");
    if (!(edges3[0] <= edges3[1] && edges3[1] <= edges3[2] && edges3[2] == 64))
      return fail(c, BK_E_INVALID, "bucket edges must rise to 64");
    for (int k = 0; k < 3; ++k) c->edges[k] = edges3[k];
  }
  if (max_iter > 0) c->max_iter = max_iter;
  if (tol > 0) c->tol = tol;
  c->gen++;
  if (cell_factor > 0) {
    c->cell_factor = c->cell_factor_capped = cell_factor;
    c->grid_valid = false;
  }
  return 0;
}

int bk_set_search_tuning(bk_ctx* c, int32_t mode) {
  if (!c) return BK_E_INVALID;
  if (mode < -1 || mode > 8) return fail(c, BK_E_INVALID, "search mode is -1, 0 or a tile reach of 1 to 8 cells");
  c->search_mode = mode;
  c->gen++;
  return 0;
}

int bk_set_refine_limit(bk_ctx* c, int32_t max_n) {
  if (!c) return BK_E_INVALID;
  if (max_n <= 0) return 0;
  if (max_n != 16 && max_n != 32 && max_n != 64) return fail(c, BK_E_INVALID, "refine limit must be 16, 32 or 64");
  c->refine_max = max_n;
  c->gen++;
  return 0;
}

int bk_set_solve_tuning(bk_ctx* c, int32_t waves) {
  if (!c) return BK_E_INVALID;
  if (waves != 1 && waves != 2 && waves != 4) return fail(c, BK_E_INVALID, "waves per workgroup is 1, 2 or 4");
  c->solve_waves = waves;
  c->gen++;
  return 0;
}

int bk_set_graphs(bk_ctx* c, int32_t enable) {
  if (!c) return BK_E_INVALID;
  c->graphs = enable != 0;
  (void)hipSetDevice(c->device);
  for (auto& g : c->graph_cache)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  c->graph_cache.clear();
  return 0;
}

int bk_get_stat(bk_ctx* c, int32_t which, int64_t* value) {
  if (!c || !value) return BK_E_INVALID;
  switch (which) {
    case BK_STAT_GRAPH_LAUNCHES:
      *value = c->graph_launches;
      return 0;
    case BK_STAT_SGS_PASSES:
      *value = c->sgs_passes;
      return 0;
    case BK_STAT_REFINED:
    case BK_STAT_REFINE_SWEEPS:
    case BK_STAT_REFINE_REDO: {
      // Pass totals kept on the device with the bucket counts.
      *value = 0;
      if (!c->counts.p) return 0;
      HIPOK(c, hipSetDevice(c->device));
      int32_t h[8] = {0};
      HIPOK(c, hipMemcpyAsync(h, c->counts.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
      HIPOK(c, hipStreamSynchronize(c->stream));
      *value = which == BK_STAT_REFINED ? h[6] : (which == BK_STAT_REFINE_SWEEPS ? h[7] : h[5]);
      return 0;
    }
    case BK_STAT_SEARCH_SPILL:
    case BK_STAT_SEARCH_RETRY: {
      // The count stays on the device during the pass; read it now.
      *value = 0;
      const int32_t* src = which == BK_STAT_SEARCH_SPILL ? c->tiles.spill_count : c->tiles.retry_count;
      if (!src) return 0;
      HIPOK(c, hipSetDevice(c->device));
      int32_t h = 0;
      HIPOK(c, hipMemcpyAsync(&h, src, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      HIPOK(c, hipStreamSynchronize(c->stream));
      *value = h;
      return 0;
    }
  }
  return fail(c, BK_E_INVALID, "unknown statistic");
}

int bk_profile(bk_ctx* c, int32_t enable) {
  if (!c) return BK_E_INVALID;
  HIPOK(c, hipSetDevice(c->device));
  if (enable && !c->ev[0])
    for (auto& e : c->ev) HIPOK(c, hipEventCreate(&e));
  c->prof = enable != 0;
  for (double& v : c->prof_ms) v = 0.0;
  c->prof_calls = 0;
  return 0;
}

int bk_profile_read(bk_ctx* c, double* ms6, int64_t* calls) {
  if (!c) return BK_E_INVALID;
  for (int k = 0; k < 6; ++k) {
    if (ms6) ms6[k] = c->prof_ms[k];
    c->prof_ms[k] = 0.0;
  }
  if (calls) *calls = c->prof_calls;
  c->prof_calls = 0;
  return 0;
}


int bk_autotune(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t kind, double param,
                int32_t precision, int64_t m, const double* tx, const double* ty, const double* tz, double* report) {
  if (!c || !vg || !s) return BK_E_INVALID;
  if (m <= 0 || !tx || !ty || !tz) return fail(c, BK_E_INVALID, "autotune needs targets");
  HIPOK(c, hipSetDevice(c->device));
  struct Scratch {
    void* p[3] = {nullptr, nullptr, nullptr};
    ~Scratch() {
      for (void* q : p)
        if (q) (void)hipFree(q);
    }
  } out;
  HIPOK(c, hipMalloc(&out.p[0], m * sizeof(double)));
  HIPOK(c, hipMalloc(&out.p[1], m * sizeof(double)));
  HIPOK(c, hipMalloc(&out.p[2], m * sizeof(int32_t)));
  const bool was_prof = c->prof;
  c->gen++;  // the knobs below change without the setters
  if (int r = bk_profile(c, 1)) return r;
  (void)bk_profile_read(c, nullptr, nullptr);
  // One untimed pass (grid build, buffer growth), then two timed ones.
  auto trial = [&](double& search_ms, double& solve_ms) -> int {
    for (int k = 0; k < 3; ++k) {
      if (k == 1) (void)bk_profile_read(c, nullptr, nullptr);
      if (int r = krige_common(c, vg, s, kind, param, nullptr, precision, m, tx, ty, tz, nullptr, false,
                               (double*)out.p[0], (double*)out.p[1], nullptr, nullptr, nullptr, (int32_t*)out.p[2]))
        return r;
    }
    double ms[6];
    int64_t calls = 0;
    (void)bk_profile_read(c, ms, &calls);
    if (calls <= 0) return fail(c, BK_E_STATE, "no timed calls");
    search_ms = (ms[0] + ms[1]) / (double)calls;
    solve_ms = (ms[2] + ms[3] + ms[4]) / (double)calls;
    return 0;
  };
  double def_search = 0.0, def_solve = 0.0;
  if (int r = trial(def_search, def_solve)) return r;
  // The search: the grid cell size of this search's kind (capped or not)
  // and, without caps, the search's mode.
  const bool capped = s->octant || s->sectors > 0 || s->max_per_hole >= 0;
  double& knob = capped ? c->cell_factor_capped : c->cell_factor;
  const double cfs[] = {0.45, 0.55, 0.65, 0.8, 1.0, 1.4};
  const int modes[] = {0, -1, 2};
  const int n_modes = capped ? 1 : 3;
  double best_cf = knob, best_search = def_search;
  int best_mode = c->search_mode;
  for (double cf : cfs)
    for (int mi = 0; mi < n_modes; ++mi) {
      knob = cf;
      c->grid_valid = false;
      if (!capped) c->search_mode = modes[mi];
      double sm = 0.0, so = 0.0;
      if (int r = trial(sm, so)) return r;
      if (sm < best_search) {
        best_search = sm;
        best_cf = cf;
        best_mode = c->search_mode;
      }
    }
  knob = best_cf;
  c->grid_valid = false;
  c->search_mode = best_mode;
  // The solves: workgroup size, bucket edges and, for FP32+IR, the largest
  // bucket the refined kernel takes (the rest go to the FP64 kernel).
  const int ws[] = {1, 2, 4};
  const int edge_sets[4][3] = {{16, 32, 64}, {32, 32, 64}, {16, 64, 64}, {64, 64, 64}};
  const int limits[2] = {32, 64};
  const int n_limits = precision == BK_FP32_IR ? 2 : 1;
  int best_w = c->solve_waves, best_e[3] = {c->edges[0], c->edges[1], c->edges[2]}, best_l = c->refine_max;
  double best_solve = def_solve;
  for (int w : ws)
    for (const auto& es : edge_sets)
      for (int li = 0; li < n_limits; ++li) {
        c->solve_waves = w;
        for (int k = 0; k < 3; ++k) c->edges[k] = es[k];
        if (n_limits > 1) c->refine_max = limits[li];
        double sm = 0.0, so = 0.0;
        if (int r = trial(sm, so)) return r;
        // The refine limit moves off its default only for a clear gain (a
        // workload without large systems times the same either way).
        if (c->refine_max != 32) so *= 1.05;
        if (so < best_solve) {
          best_solve = so;
          best_w = w;
          best_l = c->refine_max;
          for (int k = 0; k < 3; ++k) best_e[k] = es[k];
        }
      }
  c->solve_waves = best_w;
  c->refine_max = best_l;
  for (int k = 0; k < 3; ++k) c->edges[k] = best_e[k];
  c->gen++;
  if (!was_prof) (void)bk_profile(c, 0);
  if (report) {
    const double rep[BK_AUTOTUNE_REPORT] = {best_cf,     (double)best_mode, (double)best_w, (double)best_e[0],
                                            (double)best_e[1], (double)best_e[2], def_search,  def_solve,
                                            best_search, best_solve,      (double)best_l};
    std::memcpy(report, rep, sizeof(rep));
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Sequential Gaussian simulation in normal scores (crates/simulation/src/
// sgs.rs:600-750 with one search pass and no domains, trend, high-grade caps,
// local anisotropy or secondary variable; the plan of each node, constant
// and node weights, as crates/simulation/src/shared.rs:104-143 keeps it).
// Every node's position is known before any is drawn, so node k of the
// visiting order that is not on a datum searches the samples below index
// n + k of "the data, then the nodes in that order": exactly the data and the
// nodes the sequential loop has added by then. All nodes are searched and
// their simple-kriging (mean 0) weights solved at once; only the draws are
// sequential. Node k's score, sum_j w_j s_j + sqrt(max(var, 0)) z_k over its
// neighbours' scores (z_k for a node without neighbours), is evaluated in
// passes: a node is drawn once every simulated neighbour was drawn in an
// earlier pass (70-80 passes for 10^4 to 10^6 nodes of a random path).
namespace {

// The nodes in visiting order after the data: coordinates, score 0 (drawn
// later), no error variance and no hole; first_off[0] becomes the first
// rank whose elevation differs from the first sample's (m2 if none).
__global__ void k_sgs_gather(int64_t m2, int64_t n, const int64_t* app, const double* gx, const double* gy,
                             const double* gz, double* x, double* y, double* z, double* v, double* e, int32_t* hole,
                             int32_t* first_off) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= m2) return;
  const int64_t node = app[k];
  x[n + k] = gx[node];
  y[n + k] = gy[node];
  z[n + k] = gz[node];
  v[n + k] = 0.0;
  e[n + k] = 0.0;
  if (hole) hole[n + k] = -1;
  if (!(gz[node] == z[0])) atomicMin(first_off, (int32_t)k);
}

__global__ void k_sgs_set(int32_t* p, int32_t v) { p[0] = v; }

__global__ void k_sgs_limits(int64_t m2, int64_t n, int32_t* limit) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < m2) limit[k] = (int32_t)(n + k);
}

// One pass of the draws over the nodes of ranks [0, m2) in visiting order.
// s: the scores of the data (first n) and of the nodes drawn so far. state[k]:
// the pass that drew node k (> 0); while it waits, -2 - b with b the rank of
// the neighbour it waits on, -1 before its first pass. prog[k]: that
// neighbour's place in its list (the earlier ones were drawn). A node still
// blocked costs one read of its own state and one of its blocker's. A
// neighbour drawn in this pass may not be visible yet: the node waits for the
// next pass.
__device__ __forceinline__ bool sgs_drawn_before(int32_t state, int32_t pass) { return state > 0 && state < pass; }

// Node k in pass `pass`: drawn, or still waiting (true).
__device__ __forceinline__ bool sgs_step(int64_t k, int64_t n, int stride, const int32_t* nidx, const int32_t* count,
                                         const int32_t* status, const double* w, const double* variance,
                                         const double* z, double* s, int32_t* state, int32_t* prog, int32_t pass) {
  const int32_t sk = state[k];
  if (sk > 0) return false;
  if (sk <= -2 && !sgs_drawn_before(state[-2 - sk], pass)) return true;
  const int st = status[k];
  const int cnt = st == BK_OK ? count[k] : 0;
  const int32_t* id = nidx + k * (int64_t)stride;
  int j = prog[k];
  int64_t q = -1;
  for (; j < cnt; ++j) {
    q = (int64_t)id[j] - n;
    if (q >= 0 && !sgs_drawn_before(state[q], pass)) break;
  }
  if (j < cnt) {
    prog[k] = j;
    state[k] = -2 - (int32_t)q;
    return true;
  }
  double v;
  if (cnt > 0) {
    const double* wk = w + k * (int64_t)stride;
    double mean = 0.0;
    for (int i = 0; i < cnt; ++i) mean += wk[i] * s[id[i]];
    v = mean + sqrt(fmax(variance[k], 0.0)) * z[k];
  } else if (st == BK_OK || st == BK_SEARCH_FAILED || st == BK_INSUFFICIENT_DATA) {
    v = z[k];  // no neighbours: the marginal (sgs.rs:688-691, 725-728)
  } else {
    v = __longlong_as_double(0x7ff8000000000000LL);
  }
  s[n + k] = v;
  state[k] = pass;
  return false;
}

// One pass over every node. With `clear`, thread 0 zeroes *left; with `left`,
// the pass adds the nodes still waiting after it (one atomic per block).
__global__ void __launch_bounds__(256) k_sgs_pass(int64_t m2, int64_t n, int stride, const int32_t* nidx,
                                                  const int32_t* count, const int32_t* status, const double* w,
                                                  const double* variance, const double* z, double* s, int32_t* state,
                                                  int32_t* prog, int32_t pass, int32_t* clear, int32_t* left) {
  __shared__ int32_t part[4];
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (clear && k == 0) *clear = 0;
  const bool waits = k < m2 && sgs_step(k, n, stride, nidx, count, status, w, variance, z, s, state, prog, pass);
  if (!left) return;
  const int c = __popcll(__ballot(waits));
  if ((threadIdx.x & (BK_WAVE - 1)) == 0) part[threadIdx.x / BK_WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(left, t);
  }
}

// Node k at an earlier node's location: that node is its first neighbour
// (distance 0). The grid's nodes must be distinct.
__global__ void k_sgs_dup(int64_t m2, int64_t n, int stride, const int32_t* nidx, const int32_t* count,
                          const int32_t* status, const double* x, const double* y, const double* z, int32_t* dup) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= m2 || status[k] != BK_OK || count[k] == 0) return;
  const int32_t j = nidx[k * (int64_t)stride];
  if (j >= n && x[j] == x[n + k] && y[j] == y[n + k] && z[j] == z[n + k]) dup[0] = 1;
}

// Scores and sources in grid order: a node on a datum takes the datum's score
// (source = its index); others their draw (source -1 kriged, -2 marginal,
// -3 the kriging failed: score NaN).
__global__ void k_sgs_out(int64_t m, int64_t n, const int64_t* datum, const int32_t* pos, const double* s,
                          const int32_t* status, const int32_t* count, double* score, int32_t* source) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t d = datum[i];
  if (d >= 0) {
    score[i] = s[d];
    source[i] = (int32_t)d;
    return;
  }
  const int32_t k = pos[i];
  const int st = status[k];
  score[i] = s[n + k];
  source[i] = st == BK_OK && count[k] > 0 ? -1
              : (st == BK_OK || st == BK_SEARCH_FAILED || st == BK_INSUFFICIENT_DATA ? -2 : -3);
}

}  // namespace

int bk_sgs(bk_ctx* c, const bk_variogram* vg, const bk_search* s, int32_t precision, int64_t m, const double* gx,
           const double* gy, const double* gz, const int64_t* path, const double* z, double* score, int32_t* source) {
  if (!c) return BK_E_INVALID;
  if (!vg || !s) return fail(c, BK_E_INVALID, "null argument");
  if (m < 0) return fail(c, BK_E_INVALID, "m < 0");
  if (m > 0 && (!gx || !gy || !gz || !path || !z || !score || !source)) return fail(c, BK_E_INVALID, "null argument");
  if (precision != BK_FP64 && precision != BK_FP32_IR) return fail(c, BK_E_INVALID, "unknown precision");
  if (const char* msg = check_variogram(vg)) return fail(c, BK_E_INVALID, msg);
  if (const char* msg = check_search(s)) return fail(c, BK_E_INVALID, msg);
  for (int k = 0; k < vg->n_structures; ++k)
    if (vg->structures[k].model == BK_POWER)
      return fail(c, BK_E_INVALID, "simple kriging of scores needs a variogram with a sill (no power structure)");
  if (s->calibration) return fail(c, BK_E_INVALID, "the simulation takes searches without calibration");
  if (s->max_per_hole == 0 || (s->octant && s->max_per_octant == 0))
    return fail(c, BK_E_INVALID, "caps of 0 are not supported by the simulation");
  if (c->n <= 0) return fail(c, BK_E_INVALID, "no conditioning data");
  HIPOK(c, hipSetDevice(c->device));
  HIPOK(c, hipStreamSynchronize(c->stream));
  // sgs.rs:881-886: the scores' variogram with a total sill of 1.
  bk_variogram vu = *vg;
  {
    double total = vu.nugget;
    for (int k = 0; k < vu.n_structures; ++k) total += vu.structures[k].sill;
    if (total > 0.0) {
      const double f = 1.0 / total;
      vu.nugget *= f;
      for (int k = 0; k < vu.n_structures; ++k) vu.structures[k].sill *= f;
    }
  }
  const int64_t n = c->n;
  // Inputs on the device.
  const double *dgx = nullptr, *dgy = nullptr, *dgz = nullptr, *dzin = nullptr;
  const int64_t* dpath = nullptr;
  if (int r = stage_in(c, gx, m, c->tx, &dgx)) return r;
  if (int r = stage_in(c, gy, m, c->ty, &dgy)) return r;
  if (int r = stage_in(c, gz, m, c->tz, &dgz)) return r;
  if (int r = stage_in(c, path, m, c->sgs_path, &dpath)) return r;
  if (int r = stage_in(c, z, m, c->sgs_in, &dzin)) return r;
  const int64_t M = std::max<int64_t>(m, 1);
  if (int r = ensure(c, c->sgs_datum, M * sizeof(int64_t))) return r;
  if (int r = ensure(c, c->sgs_pos, M * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->sgs_app, M * sizeof(int64_t))) return r;
  if (int r = ensure(c, c->sgs_work, (3 * M + 2) * sizeof(int32_t))) return r;
  // The datum under each node (the data's grid in the search frame), and the
  // order in which the other nodes join the samples: their ranks along the path.
  int64_t m2 = 0;
  if (m > 0) {
    SearchArgs a0;
    int64_t stride0;
    if (int r = prepare_search(c, s, &vu, a0, stride0)) return r;
    a0.sx = (const double*)c->x.p;
    a0.sy = (const double*)c->y.p;
    a0.sz = (const double*)c->z.p;
    int32_t out[2] = {0, 0};
    if (bk::sgs_order(c->stream, a0, m, dgx, dgy, dgz, dpath, (int64_t*)c->sgs_datum.p, (int32_t*)c->sgs_pos.p,
                      (int64_t*)c->sgs_app.p, (int32_t*)c->sgs_work.p, &c->sgs_tmp, &c->sgs_tmp_bytes, out))
      return fail(c, BK_E_HIP, "simulation order failed");
    if (out[0]) return fail(c, BK_E_INVALID, "path must be a permutation of the nodes");
    m2 = out[1];
  }
  const int64_t N = n + m2;
  if (N > 2147483647LL) return fail(c, BK_E_INVALID, "data and nodes beyond 2^31 samples");
  // The samples of the search: the data, then the nodes in visiting order.
  for (Buf* b : {&c->sgs_x, &c->sgs_y, &c->sgs_z, &c->sgs_e, &c->sgs_v, &c->sgs_s})
    if (int r = ensure(c, *b, std::max<int64_t>(N, 1) * sizeof(double))) return r;
  if (int r = ensure(c, c->sgs_h, std::max<int64_t>(N, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->sgs_ctr, 4 * sizeof(int32_t))) return r;
  for (auto pr : {std::make_pair(&c->sgs_x, &c->x), std::make_pair(&c->sgs_y, &c->y), std::make_pair(&c->sgs_z, &c->z),
                  std::make_pair(&c->sgs_v, &c->v), std::make_pair(&c->sgs_e, &c->e)})
    if (pr.first != &c->sgs_e || c->has_err)
      HIPOK(c, hipMemcpyAsync(pr.first->p, pr.second->p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (c->has_hole)
    HIPOK(c, hipMemcpyAsync(c->sgs_h.p, c->hole.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
  // search.rs:785-797: the sectors stay planar while every sample added lies
  // at the first sample's elevation. Nodes [0, n_planar) search planar.
  int64_t n_planar = 0;
  if (m2 > 0) {
    int32_t* first_off = (int32_t*)c->sgs_ctr.p + 3;
    k_sgs_set<<<1, 1, 0, c->stream>>>(first_off, (int32_t)m2);
    k_sgs_gather<<<blocks_of(m2), 256, 0, c->stream>>>(m2, n, (const int64_t*)c->sgs_app.p, dgx, dgy, dgz,
                                                        (double*)c->sgs_x.p, (double*)c->sgs_y.p, (double*)c->sgs_z.p,
                                                        (double*)c->sgs_v.p, (double*)c->sgs_e.p,
                                                        c->has_hole ? (int32_t*)c->sgs_h.p : nullptr, first_off);
    HIPOK(c, hipGetLastError());
    int32_t first = 0;
    HIPOK(c, hipMemcpyAsync(&first, first_off, sizeof(first), hipMemcpyDeviceToHost, c->stream));
    HIPOK(c, hipStreamSynchronize(c->stream));
    if (c->planar) n_planar = std::min<int64_t>((int64_t)first + 1, m2);
  }
  HIPOK(c, hipStreamSynchronize(c->stream));
  if (!c->sgs)
    if (int r = bk_create(&c->sgs, c->device)) return fail(c, r, std::string("simulation context: ") + g_err);
  bk_ctx* k = c->sgs;
  k->cell_factor = c->cell_factor;
  k->cell_factor_capped = c->cell_factor_capped;
  k->search_mode = c->search_mode;
  k->solve_waves = c->solve_waves;
  k->refine_max = c->refine_max;
  k->max_iter = c->max_iter;
  k->tol = c->tol;
  for (int b = 0; b < 3; ++b) k->edges[b] = c->edges[b];
  if (int r = bk_set_samples(k, N, (const double*)c->sgs_x.p, (const double*)c->sgs_y.p, (const double*)c->sgs_z.p,
                             (const double*)c->sgs_v.p, c->has_err ? (const double*)c->sgs_e.p : nullptr,
                             c->has_hole ? (const int32_t*)c->sgs_h.p : nullptr))
    return fail(c, r, std::string("simulation samples: ") + k->err);
  int64_t stride = 1;
  const double* tx = (const double*)k->x.p + n;
  const double* ty = (const double*)k->y.p + n;
  const double* tz = (const double*)k->z.p + n;
  if (m2 > 0) {
    // Each node's neighbours among the samples before it.
    SearchArgs a;
    if (int r = prepare_search(k, s, &vu, a, stride)) return fail(c, r, k->err);
    if (int r = ensure(k, k->nidx, m2 * stride * sizeof(int32_t))) return fail(c, r, k->err);
    if (int r = ensure(k, k->ncount, m2 * sizeof(int32_t))) return fail(c, r, k->err);
    if (int r = ensure(k, k->sstatus, m2 * sizeof(int32_t))) return fail(c, r, k->err);
    if (int r = ensure(c, c->sgs_limit, m2 * sizeof(int32_t))) return r;
    int32_t* limit = (int32_t*)c->sgs_limit.p;
    k_sgs_limits<<<blocks_of(m2), 256, 0, k->stream>>>(m2, n, limit);
    const int64_t cut[3] = {0, n_planar, m2};
    for (int part = 0; part < 2; ++part) {
      const int64_t r0 = cut[part], r1 = cut[part + 1];
      if (r1 <= r0) continue;
      SearchArgs b = a;
      b.planar = part == 0;
      b.n_sectors = b.plane ? b.plane : (b.planar ? 4 : 8);
      if (s->octant)
        b.sector_cap = (int)std::min<int64_t>(
            s->max_per_octant >= 0 ? s->max_per_octant : (s->max_samples + b.n_sectors - 1) / b.n_sectors,
            std::numeric_limits<int>::max());
      sector_map(b);
      b.tx = tx + r0;
      b.ty = ty + r0;
      b.tz = tz + r0;
      b.m = r1 - r0;
      b.list = nullptr;
      b.left_out = nullptr;
      b.limit = limit + r0;
      b.idx = (int32_t*)k->nidx.p + r0 * stride;
      b.count = (int32_t*)k->ncount.p + r0;
      b.status = (int32_t*)k->sstatus.p + r0;
      if (bk::launch_search(k->stream, b, r1 - r0, k->search_mode, k->tiles)) return fail(c, BK_E_HIP, "search launch failed");
    }
    k_map_overflow<<<blocks_of(m2), 256, 0, k->stream>>>(m2, (int32_t*)k->sstatus.p);
    // Simple kriging (mean 0) weights and variances from positions alone.
    for (Buf* b : {&k->value, &k->variance})
      if (int r = ensure(k, *b, m2 * sizeof(double))) return fail(c, r, k->err);
    if (int r = ensure(k, k->n_used, m2 * sizeof(int32_t))) return fail(c, r, k->err);
    if (int r = ensure(k, k->status, m2 * sizeof(int32_t))) return fail(c, r, k->err);
    if (int r = ensure(c, c->sgs_w, m2 * stride * sizeof(double))) return r;
    if (int r = run_solve(k, &vu, s, BK_SIMPLE, 0.0, precision, m2, tx, ty, tz, (int)stride, nullptr,
                          (double*)k->value.p, (double*)k->variance.p, nullptr, nullptr, (int32_t*)k->n_used.p,
                          (int32_t*)k->status.p, (double*)c->sgs_w.p))
      return fail(c, r, k->err);
  }
  // The draws, in passes.
  if (int r = ensure(c, c->sgs_done, std::max<int64_t>(m2, 1) * sizeof(int32_t))) return r;
  if (int r = ensure(c, c->sgs_prog, std::max<int64_t>(m2, 1) * sizeof(int32_t))) return r;
  double* sc = (double*)c->sgs_s.p;
  HIPOK(c, hipMemcpyAsync(sc, c->v.p, n * sizeof(double), hipMemcpyDeviceToDevice, k->stream));
  int32_t* dup = (int32_t*)c->sgs_ctr.p + 2;
  if (m2 > 0) {
    int32_t* ctr = (int32_t*)c->sgs_ctr.p;  // [0]: nodes still waiting after the last pass; [2]: duplicate
    if (bk::zero32(k->stream, (int32_t*)c->sgs_prog.p, m2) || bk::zero32(k->stream, ctr, 4))
      return fail(c, BK_E_HIP, "fill launch failed");
    HIPOK(c, hipMemsetAsync(c->sgs_done.p, 0xff, m2 * sizeof(int32_t), k->stream));  // state -1
    k_sgs_dup<<<blocks_of(m2), 256, 0, k->stream>>>(m2, n, (int)stride, (const int32_t*)k->nidx.p,
                                                     (const int32_t*)k->ncount.p, (const int32_t*)k->status.p,
                                                     (const double*)k->x.p, (const double*)k->y.p,
                                                     (const double*)k->z.p, dup);
    int32_t pass = 0, left = 1;
    while (left > 0) {
      // Each pass draws at least the first node still waiting (its simulated
      // neighbours come before it), so m2 passes always end.
      if (pass > m2) return fail(c, BK_E_HIP, "simulation passes did not finish");
      // The first pass of eight clears the count, the last one counts.
      for (int q = 0; q < 8; ++q) {
        ++pass;
        k_sgs_pass<<<blocks_of(m2), 256, 0, k->stream>>>(
            m2, n, (int)stride, (const int32_t*)k->nidx.p, (const int32_t*)k->ncount.p, (const int32_t*)k->status.p,
            (const double*)c->sgs_w.p, (const double*)k->variance.p, dzin, sc, (int32_t*)c->sgs_done.p,
            (int32_t*)c->sgs_prog.p, pass, q == 0 ? ctr : nullptr, q == 7 ? ctr : nullptr);
      }
      HIPOK(c, hipGetLastError());
      HIPOK(c, hipMemcpyAsync(&left, ctr, sizeof(left), hipMemcpyDeviceToHost, k->stream));
      HIPOK(c, hipStreamSynchronize(k->stream));
    }
    c->sgs_passes = pass;
    int32_t d = 0;
    HIPOK(c, hipMemcpyAsync(&d, dup, sizeof(d), hipMemcpyDeviceToHost, k->stream));
    HIPOK(c, hipStreamSynchronize(k->stream));
    if (d) return fail(c, BK_E_INVALID, "grid nodes must be distinct");
  }
  if (m > 0) {
    double* osc;
    int32_t* osrc;
    if (int r = out_ptr(c, score, m, c->value, &osc)) return r;
    if (int r = out_ptr(c, source, m, c->n_used, &osrc)) return r;
    const int32_t* kst = m2 > 0 ? (const int32_t*)k->status.p : nullptr;
    const int32_t* kcnt = m2 > 0 ? (const int32_t*)k->n_used.p : nullptr;
    k_sgs_out<<<blocks_of(m), 256, 0, k->stream>>>(m, n, (const int64_t*)c->sgs_datum.p, (const int32_t*)c->sgs_pos.p,
                                                   sc, kst, kcnt, osc, osrc);
    HIPOK(c, hipGetLastError());
    HIPOK(c, hipStreamSynchronize(k->stream));
    if (int r = stage_out(c, score, osc, m)) return r;
    if (int r = stage_out(c, source, osrc, m)) return r;
    HIPOK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}
