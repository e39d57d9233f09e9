// bk_solve.h — batched kriging systems on the device.
//
// One system of n <= NMAX samples per group of NMAX lanes (64 / NMAX systems
// per wavefront); lane r owns row r of the covariance matrix. The FP64 kernel
// restates crates/estimation/src/kernel.rs:24-123 (frame, packed covariance,
// left-looking Cholesky with the 1e-12 * sill pivot test, two triangular solve
// pairs, the ordinary-kriging constraint by the Schur scalar 1'C^-1 1); the
// block form restates crates/estimation/src/block.rs:100-198 with the bordered
// system solved through the same factor.
#pragma once

#include <stdint.h>

#include "bk_device.h"

// Internal per-target status: the system has coincident samples and goes to
// the guarded kernel; or the FP32 factor failed / refinement did not converge
// and the system goes to the FP64 kernel.
#define BK_ST_GUARD 101
#define BK_ST_REDO64 102
#define BK_ST_LU 103

// Workgroups of the LU kernel for systems above BK_MAX_LU samples.
#define BK_LU_BIG_GROUPS 32

struct SolveArgs {
  DevVariogram v;
  DevVariogram32 v32;
  int kind;
  double param;
  int block;            // block support: RHS averaged over the discretisation
  int npts;
  const double* offs;   // npts * 3 offsets from the block centre (world axes)
  const double* poffs;  // the offsets projected on each frame: F_f o_p at (p * nf + f) * 3 (refined kernels)
  double cbb;           // C(B, B) of the discretisation (block.rs:100-118)
  const double* sx;
  const double* sy;
  const double* sz;
  const double* sv;
  const double* serr;   // may be null
  const double* tx;
  const double* ty;
  const double* tz;
  const int32_t* nidx;  // neighbour lists, stride per target
  const int32_t* ncount;
  int stride;
  const int32_t* list;  // systems of this launch: target ids
  int64_t n_list;       // their number, or a bound on it when n_list_dev is set
  const int32_t* n_list_dev;  // the number on the device (written by an earlier kernel), or null
  int32_t* work;        // BK_CTR_INTS of work-unit counters (zeroed before the launch); a unit is 64 / NMAX systems
  double* value;
  double* variance;
  double* lagrange;     // may be null
  double* support;      // may be null
  double* weights;      // may be null: kriging weights per target, in neighbour order at [t * stride + k]
  int32_t* n_used;      // may be null
  int32_t* status;
  int32_t* guard;       // targets handed to the guarded (merge + LU) kernel
  int32_t* guard_count;
  int32_t* redo;        // FP32 systems handed to the FP64 kernel
  int32_t* redo_count;
  int waves;            // wavefronts per workgroup of the solve kernels (1, 2 or 4)
  int max_iter;         // FP32+IR refinement sweeps
  double tol;           // FP32+IR: stop when the correction is below tol * |x|
  // Per-system diagnostics of the refined form (may be null): sweeps taken.
  int32_t* iters;
  // Pass totals of the refined form (may be null): [0] systems it solved,
  // [1] their refinement sweeps (zeroed before the pass).
  int32_t* refine_stats;
  // Leave-one-out gradient: the sample held out by each system, and per
  // system (loss, used, gradient[P]).
  const int64_t* left_out_ids;
  double* grad_part;
};

#define BK_MAX_GRAD (4 + 4 * BK_MAX_STRUCTURES)

// The parameters of the leave-one-out gradient (layout in boitata_gpu.h).
struct GradArgs {
  int S;
  int P;     // 4 + 4 S
  int64_t m; // held-out samples: the column length of grad_part
  int model[BK_MAX_STRUCTURES];
  double sill[BK_MAX_STRUCTURES];
  double r[BK_MAX_STRUCTURES][3];  // major, semi, minor ranges (m)
  double R[9];                     // rotation_matrix(azimuth, dip, rake)
  double dR[3][9];                 // its derivatives per degree of each angle
};

// The same parameters in FP32 for the refined gradient kernel: inverse
// ranges instead of ranges.
struct GradArgs32 {
  int S;
  int P;
  int model[BK_MAX_STRUCTURES];
  float sill[BK_MAX_STRUCTURES];
  float inv_r[BK_MAX_STRUCTURES][3];
  float R[9];
  float dR[3][9];
};

// Calibrated searches (kriging_algebra.rs:338-415): per target, the quality of
// kriging from each leading part of its selection, and the count trimmed to
// the fewest leading samples, from min_samples, that reach the target.
struct CalibArgs {
  const int32_t* sstatus;  // search status per target (only BK_OK targets are calibrated)
  int32_t* count;          // the selections' counts, trimmed in place
  int64_t m;
  int calibration;         // BK_CALIBRATE_SLOPE or BK_CALIBRATE_EFFICIENCY
  double target;
  int64_t min_samples;
  int ordinary;            // the unbiasedness column of ones (not simple kriging)
  int32_t* met;            // may be null: 1 where a leading part reaches the target, else 0
};

namespace bk {
// One FP64 system per NMAX lanes (NMAX >= the selections' size, <= 64).
int launch_calibrate(hipStream_t s, const SolveArgs& a, const CalibArgs& k, int nmax);
int launch_krige64(hipStream_t s, const SolveArgs& a, int nmax);
int launch_krige32(hipStream_t s, const SolveArgs& a, int nmax);
// Solves the systems of a.guard[0 .. *a.guard_count) by bordered LU (merged
// coincident samples, n > 64, power models): up to BK_MAX_LU samples with the
// matrix in LDS, larger ones (up to BK_MAX_SEARCH) in lu_ws, which must hold
// lu_big_workspace_bytes() when a list may hold such systems (else NULL).
int launch_guarded(hipStream_t s, const SolveArgs& a, double* lu_ws);
size_t lu_big_workspace_bytes();
int launch_loo_grad(hipStream_t s, const SolveArgs& a, const GradArgs& gp, int nmax);
// The refined form: FP32 Gauss-Jordan forward with FP64 refinement, FP32
// adjoint accumulation.
int launch_loo_grad32(hipStream_t s, const SolveArgs& a, const GradArgs& gp, const GradArgs32& g32, int nmax);
// The refined form for systems of at most 32 samples with the elimination on
// the matrix cores (two systems per wavefront, persistent over a.work).
int launch_loo_grad32m(hipStream_t s, const SolveArgs& a, const GradArgs& gp, const GradArgs32& g32);
// Column sums of part (cols columns of m values each, column-major) into
// out[0..cols), in an order fixed by m alone; work holds
// cols * reduce_blocks(m) doubles.
int64_t reduce_blocks(int64_t m);
int launch_reduce_cols(hipStream_t s, int64_t m, int cols, const double* part, double* work, double* out);
int launch_lu(hipStream_t s, const SolveArgs& a);
// Buckets targets by neighbour count after the search (status OK): lists[b]
// holds targets with n in (edge[b-1], edge[b]]; others get their status.
int launch_bucket(hipStream_t s, int64_t m, const int32_t* count, const int32_t* sstatus, int32_t* status,
                  double* value, double* variance, int32_t* lists, int32_t* counts, int64_t list_stride,
                  const int* edges, int n_buckets, int stationary);
}  // namespace bk
