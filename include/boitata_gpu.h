/*
 * boitata_gpu.h — C ABI of the MI355X kriging engine.
 *
 * The engine replaces the per-target estimator closure and the search tree of
 * boitatá's estimation hot path:
 *   - crates/estimation/src/batch.rs:15-38      estimate_many (search + krige per target)
 *   - crates/estimation/src/batch.rs:165-190    leave_one_out_at
 *   - crates/estimation/src/krige.rs:83-95      krige (OK / SK / IK, point support)
 *   - crates/estimation/src/block.rs:122-198    block_krige_points
 *   - crates/estimation/src/search.rs:951-982   SearchTree::neighbors_around
 * Parameters mirror crates/variogram/src/composite.rs:13-18 (Variogram),
 * crates/variogram/src/aniso.rs:18-26 (Angles), crates/estimation/src/search.rs:31-81
 * (Search) and the Python constructors crates/py/src/variogram.rs:330-362 and
 * crates/py/src/estimation.rs:415-531.
 *
 * Conventions
 *   - Every array argument may be host memory or device memory of the context's
 *     device (detected per pointer). Results are written in place; the library
 *     keeps nothing but the state given to bk_set_samples.
 *   - Functions return 0 on success or a negative BK_E* code; the message is in
 *     bk_last_error(ctx). Per-target failures are not errors: they are reported in
 *     the status array (the reference's EstimError per target, crates/estimation/src/error.rs:4-16)
 *     and the target's value and variance are NaN.
 *   - A context owns one HIP stream on one device and is not thread-safe; calls are
 *     synchronous. Separate contexts may be used from separate threads.
 */
#ifndef BOITATA_GPU_H
#define BOITATA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-target status, crates/estimation/src/error.rs:4-16. */
enum {
  BK_OK = 0,
  BK_SEARCH_FAILED = 1,      /* too few samples, or a sector / hole rule broken */
  BK_SINGULAR = 2,           /* pivot <= 1e-12 * sill (kernel.rs:58-65) or LU breakdown */
  BK_INSUFFICIENT_DATA = 3,  /* no samples */
  BK_INVALID_PARAMETERS = 4,
  BK_UNSUPPORTED = 5         /* a capped search with more than 65,536 candidates pending at once, or with
                                bk_krige_noalloc more than 512 (the device pools) */
};

/* Return codes of the API calls. */
enum {
  BK_E_INVALID = -1,   /* invalid parameters (Search::check, Variogram::validate) */
  BK_E_HIP = -2,       /* a HIP runtime error */
  BK_E_STATE = -3,     /* call order: no samples set, no factor saved */
  BK_E_NOMEM = -4
};

/* Kriging kind, crates/estimation/src/krige.rs:20-28. */
enum { BK_ORDINARY = 0, BK_SIMPLE = 1, BK_INDICATOR = 2 };

/* Arithmetic of the solve.
 * BK_FP64: FP64 throughout. Point systems follow kernel.rs:24-123 (covariances
 *   in its operation order, left-looking Cholesky with its pivot test; the back
 *   substitution subtracts from the last row up). Block systems take the same
 *   Cholesky and Schur steps where block.rs:167-170 uses a bordered LU, and go
 *   to that LU when the Cholesky pivot test fails; merged coincident samples,
 *   systems above 64 samples and power models take the bordered LU.
 * BK_FP32_IR: covariances in FP64 (a fast exp and square root), an FP32
 *   Gauss-Jordan inverse of C / c0 on the matrix cores (MFMA), FP64 residual
 *   refinement, and the Schur scalar and weights in FP64. The refinement stops
 *   when the correction still to come is predicted below the tolerance
 *   (relative to the first iterate; 1e-9 by default): after the first sweep
 *   from its correction delta and the contraction estimated as
 *   max(delta, n 2^-24 / smallest pivot), after later sweeps from their own
 *   correction. Size buckets above the refine limit
 *   (bk_set_refine_limit), variograms with a circular, sine-hole, Matern or
 *   power structure, and systems whose FP32 pivot falls below 1e-5 of the
 *   scaled sill or that do not converge take the FP64 path. */
enum {
  BK_FP64 = 0,
  BK_FP32_IR = 1
};

/* Variogram model, crates/variogram/src/model.rs:11-30. */
enum {
  BK_SPHERICAL = 0,
  BK_EXPONENTIAL = 1,
  BK_GAUSSIAN = 2,
  BK_CUBIC = 3,
  BK_PENTASPHERICAL = 4,
  BK_CIRCULAR = 5,
  BK_SINEHOLE = 6,
  BK_MATERN = 7,   /* param = order (snaps to 0.5, 1.5, 2.5) */
  BK_POWER = 8     /* param = exponent; not stationary */
};

#define BK_MAX_STRUCTURES 4
#define BK_MAX_CHOL 64     /* systems solved by Cholesky on the device (kernel.rs:11 MAX) */
#define BK_MAX_LU 79       /* larger systems and power models: bordered LU with the matrix in LDS */
#define BK_MAX_SEARCH 256  /* most samples one search selects (bk_search.max_samples); systems above
                              BK_MAX_LU take the same LU with the matrix in device memory */

/* One nested structure, crates/variogram/src/model.rs:34-41. `range` runs along
 * the major axis. ratio_semi / ratio_minor > 0 give this structure its own
 * anisotropy ratios under the variogram's rotation (an extension: the reference
 * shares one anisotropy, composite.rs:17); values <= 0 take the variogram's. */
typedef struct {
  int32_t model;
  int32_t reserved;
  double param;
  double sill;
  double range;
  double ratio_semi;
  double ratio_minor;
} bk_structure;

/* Nugget plus nested structures with an optional anisotropy of major = 1
 * (crates/py/src/variogram.rs:164-177: rotation (azimuth, dip, rake) in degrees,
 * ratios (semi, minor) of the major range). */
typedef struct {
  double nugget;
  int32_t n_structures;
  int32_t has_anisotropy;
  bk_structure structures[BK_MAX_STRUCTURES];
  double azimuth, dip, rake;
  double ratio_semi, ratio_minor;
} bk_variogram;

/* Search neighbourhood, crates/estimation/src/search.rs:31-81. Optional counts
 * are -1 when absent. Without an anisotropy of its own the search measures
 * distance in the variogram's (search.rs:343-348). */
typedef struct {
  int64_t min_samples;
  int64_t max_samples;
  double radius;              /* in the ellipsoid metric: metres along its major axis */
  int32_t has_anisotropy;
  int32_t octant;
  double azimuth, dip, rake;
  double ratio_semi, ratio_minor;
  int64_t max_per_hole;       /* -1: none */
  int64_t max_per_octant;     /* -1: max_samples / count, rounded up */
  int64_t sectors;            /* 0: none; else the number of plane sectors */
  int64_t max_per_sector;
  int64_t min_sectors;        /* -1: none */
  int64_t max_empty_sectors;  /* -1: none */
  int64_t min_holes;          /* -1: none */
  /* Calibrated search (search.rs:59-61,94-99; py/estimation.rs:458-472): each
   * target is kriged from the fewest leading samples, from min_samples, whose
   * kriging reaches the target slope of regression or efficiency
   * (kriging_algebra.rs:344-441), or from all of them. A search that balances
   * octants or sectors then takes the sectors in turn (search.rs:909-936). */
  int32_t calibration;        /* 0: none, BK_CALIBRATE_SLOPE or BK_CALIBRATE_EFFICIENCY */
  int32_t reserved;
  double calibration_target;  /* slope: finite and > 0; efficiency: finite and <= 1 */
} bk_search;

#define BK_CALIBRATE_SLOPE 1
#define BK_CALIBRATE_EFFICIENCY 2

/* Block support, crates/estimation/src/block.rs:18-71 and
 * crates/py/src/estimation.rs:947-958: a block of `size` along the axes of
 * block_frame(rotation), discretised nx * ny * nz. */
typedef struct {
  double size[3];
  int32_t nx, ny, nz;
  int32_t reserved;
  double rotation[3];
} bk_block;

typedef struct bk_ctx bk_ctx;

/* Context on HIP device `device`. On failure *out is NULL and the message is
 * in bk_last_error(NULL). */
int bk_create(bk_ctx** out, int device);
void bk_destroy(bk_ctx* ctx);
const char* bk_last_error(const bk_ctx* ctx);
/* Library version string; never fails. */
const char* bk_version(void);

/* Validation, as crates/estimation/src/search.rs:214-264 (Search::check) and
 * crates/variogram/src/composite.rs:37-48 (Variogram::validate). ctx may be
 * NULL (no device needed); the message is then read with bk_last_error(NULL),
 * which holds the last context-free failure of the calling thread. */
int bk_check_search(bk_ctx* ctx, const bk_search* search);
int bk_check_variogram(bk_ctx* ctx, const bk_variogram* vg);

/* The fitted samples (crates/estimation/src/lib.rs:70-82, crates/py/src/estimation.rs:1244-1310).
 * error_var may be NULL (all 0); hole may be NULL (no holes), -1 marks a sample
 * without a hole. Copied to the device. */
int bk_set_samples(bk_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                   const double* value, const double* error_var, const int32_t* hole);

/* Neighbour selection of SearchTree::neighbors_around (search.rs:951-982) for m
 * targets: idx[t * max_samples + k], k < count[t], in selection order.
 * left_out (may be NULL) holds per target a sample index taken out of the
 * search, or -1. status[t] is BK_OK or BK_SEARCH_FAILED. */
int bk_neighbors(bk_ctx* ctx, const bk_search* search, const bk_variogram* vg, int64_t m,
                 const double* tx, const double* ty, const double* tz, const int64_t* left_out,
                 int32_t* idx, int32_t* count, int32_t* status);

/* estimate_many with the krige estimator (batch.rs:15-38, krige.rs:83-95).
 * param is the SK mean or the IK threshold. lagrange, n_used, status may be NULL. */
int bk_krige(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t kind,
             double param, int32_t precision, int64_t m, const double* tx, const double* ty,
             const double* tz, double* value, double* variance, double* lagrange,
             int32_t* n_used, int32_t* status);

/* bk_krige with no device allocation (no counterpart in the reference), for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc. The pass's scratch (neighbour lists,
 * buckets, counters, the cell grid of the search) lives in one static device
 * array per GPU of about 44 MiB, which each GPU the library runs on holds until
 * the process ends; the call allocates nothing and calls no allocating
 * routine. Limits: m <= BK_STATIC_MAX_TARGETS, at most BK_STATIC_MAX_SAMPLES
 * samples set on the context, max_samples <= BK_STATIC_MAX_NEIGHBOURS, no
 * calibrated search; targets and outputs must be device arrays (lagrange,
 * n_used and status may be NULL). Beyond a limit the call returns BK_E_INVALID
 * and computes nothing. Results equal bk_krige's. Calls on one GPU share its
 * array: the library serialises them within a process; calls that overlap on
 * one GPU from several processes race. */
#define BK_STATIC_MAX_TARGETS 65536
#define BK_STATIC_MAX_SAMPLES 262144
#define BK_STATIC_MAX_NEIGHBOURS 64
int bk_krige_noalloc(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t kind, double param,
                     int32_t precision, int64_t m, const double* tx, const double* ty, const double* tz,
                     double* value, double* variance, double* lagrange, int32_t* n_used, int32_t* status);

/* estimate_many with block_krige_points (block.rs:122-198) on the discretisation
 * of `block` centred on each target (py/estimation.rs:947-958). support_variance
 * (C(B,B)) may be NULL. */
int bk_block_krige(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t kind,
                   double param, const bk_block* block, int32_t precision, int64_t m,
                   const double* tx, const double* ty, const double* tz, double* value,
                   double* variance, double* lagrange, double* support_variance,
                   int32_t* n_used, int32_t* status);

/* leave_one_out_at (batch.rs:165-190) over the samples at `which` (m of them;
 * NULL = all samples, m ignored). */
int bk_loo(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t kind,
           double param, int32_t precision, int64_t m, const int64_t* which, double* value,
           double* variance, int32_t* status);

/* leave_one_out_at (batch.rs:165-190) with block_krige_points as the
 * estimator (crates/py/src/estimation.rs:1826-1834, BlockKriging.cross_validate):
 * each sample at `which` (m of them; NULL = all) is kriged on the block of
 * `block` centred on its location, from its neighbours without itself. */
int bk_block_loo(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t kind, double param,
                 const bk_block* block, int32_t precision, int64_t m, const int64_t* which, double* value,
                 double* variance, int32_t* status);

/* weights_many (batch.rs:76-95) with the krige or block_krige_points estimator
 * (block NULL: point support): per target, the samples it was kriged from in
 * the order the estimator received them, idx[t * S + k] for k < count[t], and
 * their kriging weights weights[t * S + k] (merged coincident samples share
 * their cluster's weight, coincident.rs:102-112), where S = min(max_samples,
 * the number of samples), at least 1. count[t] is 0 and status[t] (may be
 * NULL) not BK_OK where the reference gives None. */
int bk_krige_weights(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t kind, double param,
                     const bk_block* block, int32_t precision, int64_t m, const double* tx, const double* ty,
                     const double* tz, int32_t* idx, double* weights, int32_t* count, int32_t* status);

/* Sequential Gaussian simulation in normal scores: the node loop of
 * crates/simulation/src/sgs.rs:600-750 (sequential, with its search tree grown
 * by SearchTree::add, search.rs:785-797) for one search pass, without domains,
 * trend, high-grade caps, local anisotropy or a secondary variable. The
 * context's samples are the conditioning data with their normal scores as
 * values; vg is the scores' variogram, scaled to a total sill of 1 as sgs.rs
 * does. The m grid nodes (distinct) are visited in the order `path` (a
 * permutation of 0 .. m-1, the reference's shuffled path); z holds m standard
 * normal draws, the k-th node of the path that is not on a datum taking z[k]
 * (the reference draws from its generator in that order). Per node: score[i],
 * and source[i] >= 0 the datum it lies on (score = that datum's score), -1
 * simple-kriged from its neighbours, -2 drawn from the marginal (no
 * neighbours), -3 its kriging failed (score NaN; the reference returns an
 * error). Searches with calibration or caps of 0 are rejected. */
int bk_sgs(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t precision, int64_t m,
           const double* gx, const double* gy, const double* gz, const int64_t* path, const double* z, double* score,
           int32_t* source);

/* Leave-one-out cross-validation loss and its gradient (no counterpart in the
 * reference). The neighbour sets are those of bk_loo with `search`, fixed before
 * the call by bk_loo_prepare, so they do not move with the parameters.
 *   loss = mean_i (z_i - zhat_-i)^2 + mean_i log sigma2_-i
 * over the held-out samples whose system solved. grad packs the derivatives in
 * this order: nugget, sill_s (s < S), then per structure its three ranges in
 * metres (major = range, semi = range * ratio_semi, minor = range * ratio_minor),
 * then azimuth, dip, rake in degrees: 4 + 4 S values (12 for two structures). */
int bk_loo_prepare(bk_ctx* ctx, const bk_search* search, const bk_variogram* vg);
int bk_loo_loss_grad(bk_ctx* ctx, const bk_variogram* vg, int32_t precision, double* loss,
                     double* grad, int64_t* n_used);

/* Per-target diagnostics of a kriging pass: the estimate's fields and the
 * statistics of the samples used, as _Estimator.predict(diagnostics=True)
 * tabulates them (crates/py/src/estimation.rs:1034-1105, with
 * neighborhood_stats over the samples used, crates/estimation/src/neighborhood.rs:66-136).
 * Any pointer may be NULL. */
typedef struct {
  double* value;
  double* variance;
  double* lagrange;
  double* support_variance;      /* C(v, v): the sill for points, C(B, B) for blocks */
  int32_t* n_used;
  int32_t* status;
  double* negative_weight_sum;   /* sum of the negative kriging weights (krige.rs:64-75) */
  int32_t* n_holes;              /* distinct holes among the samples used; untagged samples count one each */
  double* mean_distance;         /* mean Euclidean distance from the target to the samples used */
  int32_t* target_met;           /* calibrated searches: 1 when the samples kept reach the target, else 0; -1 when
                                    the search is not calibrated or the target failed */
} bk_diag_out;

/* estimate_many with the krige or block_krige_points estimator (block NULL:
 * point support) and the diagnostics of each target (py/estimation.rs:895-904,
 * 1034-1105). Values are NaN where status != BK_OK. */
int bk_krige_diagnostics(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t kind, double param,
                         const bk_block* block, int32_t precision, int64_t m, const double* tx, const double* ty,
                         const double* tz, const bk_diag_out* out);

/* Multiple indicator kriging, crates/estimation/src/indicator.rs:53-67
 * (MultipleIndicator), :35-50 (Interpolation, UpperTail) and
 * crates/py/src/indicator.rs:116-192 (the Python constructor). */
enum { BK_MIK_GLOBAL = 0, BK_MIK_LINEAR = 1 };
enum { BK_TAIL_NONE = 0, BK_TAIL_POWER = 1, BK_TAIL_HYPERBOLIC = 2 };

typedef struct {
  int32_t n_thresholds;   /* K >= 1, strictly increasing */
  int32_t n_variograms;   /* 1 (one for every threshold) or K (one per threshold) */
  int32_t simple;         /* simple kriging about the declustered proportions; else ordinary */
  int32_t interpolation;  /* within a class: BK_MIK_GLOBAL (the declustered data) or BK_MIK_LINEAR */
  int32_t upper_tail;     /* model of the class above the last threshold: BK_TAIL_* */
  int32_t has_tails;      /* tails[] bound the distribution; else the data minimum and maximum */
  double omega;           /* exponent of the upper tail model */
  double tails[2];        /* lower, upper: widened to cover the data and thresholds */
} bk_mik;

/* Outputs of multiple indicator kriging; any pointer may be NULL. Arrays with
 * several entries per target are entry-major: entry k of target t at [k * m + t]. */
typedef struct {
  double* raw;                /* K * m: kriged P(Z <= t_k), before the correction */
  double* cdf;                /* K * m: after the order-relation correction */
  double* correction;         /* m: sum over k of |cdf - raw| */
  double* mean;               /* m: E-type estimate */
  double* variance;           /* m */
  double* probability_above;  /* n_cutoffs * m */
  double* mean_above;         /* n_cutoffs * m; NaN where nothing lies above the cutoff */
  double* quantile_values;    /* n_quantiles * m */
  int32_t* violations;        /* m: thresholds whose kriged probability the correction changed */
  int32_t* n_used;            /* m: samples found by the search */
  int32_t* status;            /* m: BK_OK, or the first failure among the target's search and solves */
} bk_mik_out;

/* MultipleIndicator::predict for one search pass (indicator.rs:439-487):
 * per target, the search of variograms[n_variograms / 2] (:319-321); each
 * variogram's kriging weights on those samples (krige_indicators, :688-749,
 * with block_krige_points on the discretisation of `block` when it is not
 * NULL); the indicator estimates; then the conditional distribution
 * (:380-437) against the declustered global distribution of the fitted
 * samples (Global::new, :152-200; weights may be NULL: all 1). Values are NaN
 * where status != BK_OK. The search must not be calibrated. */
int bk_mik_predict(bk_ctx* ctx, const bk_mik* mik, const double* thresholds, const bk_variogram* variograms,
                   const double* weights, const bk_search* search, const bk_block* block, int32_t precision,
                   int64_t m, const double* tx, const double* ty, const double* tz, int32_t n_cutoffs,
                   const double* cutoffs, int32_t n_quantiles, const double* quantiles, const bk_mik_out* out);

/* MultipleIndicator::cross_validate, leave-one-out form (indicator.rs:489-529)
 * for one pass: each sample at `which` (m of them; NULL = all, m ignored) is
 * re-estimated without itself; pit[t] = F*(z) at its own value (NaN where
 * unestimated). probability_above, mean_above and quantile_values of `out`
 * are not written. */
int bk_mik_cross_validate(bk_ctx* ctx, const bk_mik* mik, const double* thresholds,
                          const bk_variogram* variograms, const double* weights, const bk_search* search,
                          int32_t precision, int64_t m, const int64_t* which, double* pit, const bk_mik_out* out);

/* Tuning of the device path (no counterpart in the reference): bucket edges of
 * the neighbour counts (three of 16, 32, 64, rising to 64; NULL keeps them),
 * FP32+IR refinement sweeps and relative tolerance, and the grid cell size as
 * a factor of the reference's side rule (search/grid.rs:83-106), set for both
 * kinds of search (the defaults are 0.65 without caps and 1.0 for searches
 * with octants, sectors or a hole cap; bk_autotune tunes the kind of the
 * search it is given). Values <= 0 keep the current setting. Results do not
 * depend on the cell factor. */
int bk_set_tuning(bk_ctx* ctx, const int32_t* edges3, int32_t max_iter, double tol, double cell_factor);

/* Search tuning (no counterpart in the reference) for searches without caps
 * (no octants, sectors or hole cap, max_samples <= 64). mode 0 (default):
 * one wavefront per target walking the grid's shells 64 cells at a time with
 * the selection held across the lanes; 1..8: tiles of 64 targets sorted along
 * a Morton curve of their grid cells, each staging the samples of the cells
 * within `mode` cells of its cells in LDS; -1: one lane per target walking
 * cell by cell. Targets the first two cannot settle are searched by the
 * last. Results do not depend on the mode. */
int bk_set_search_tuning(bk_ctx* ctx, int32_t mode);

/* Workgroup size of the solve kernels (no counterpart in the reference): 1, 2
 * (default) or 4 wavefronts. Results do not depend on it. */
int bk_set_solve_tuning(bk_ctx* ctx, int32_t waves);

/* Largest size bucket (16, 32 or 64 samples; default 64) that BK_FP32_IR
 * solves by FP32 elimination with FP64 refinement (no counterpart in the
 * reference). Systems in larger buckets take the FP64 kernel. Values <= 0
 * keep the setting. Applies to kriging passes and to the leave-one-out
 * gradient. */
int bk_set_refine_limit(bk_ctx* ctx, int32_t max_n);

/* Autotuner (no counterpart in the reference): times krige passes over the m
 * targets given (a sample of a workload) with HIP events and keeps, in the
 * context, the fastest grid cell factor and uncapped search mode (by the
 * search stage) and the fastest solve workgroup size and bucket edges (by
 * the bucket, solve and deferred stages). The cell factor, search mode and
 * workgroup size do not change results. The bucket edges and the refine
 * limit choose the kernel that solves a size class: in BK_FP64 its results
 * are the same bit for bit; in BK_FP32_IR the refined kernels agree with each
 * other and with FP64 to the refinement's tolerance, not bit for bit.
 * report (may be NULL) receives BK_AUTOTUNE_REPORT doubles: cell factor,
 * search mode, waves, edges[3], then search and solve milliseconds per pass
 * before and after tuning, then the refine limit (bk_set_refine_limit; for
 * FP32+IR it is tuned too). */
#define BK_AUTOTUNE_REPORT 11
int bk_autotune(bk_ctx* ctx, const bk_variogram* vg, const bk_search* search, int32_t kind, double param,
                int32_t precision, int64_t m, const double* tx, const double* ty, const double* tz,
                double* report);

/* Graph replay of bk_krige (no counterpart in the reference; on by default).
 * A bk_krige call whose targets and outputs are device arrays, made again with
 * the same arguments while the context's samples, buffers and tuning are
 * unchanged, is captured into a HIP graph (search, bucketing, persistent
 * solves, deferred systems: no host round trip inside the pass) and later
 * calls launch that graph. Results do not depend on it; stage timing
 * (bk_profile) and the tiled search keep calls on the ordinary path. */
int bk_set_graphs(bk_ctx* ctx, int32_t enable);

/* Counters (no counterpart in the reference): BK_STAT_SEARCH_SPILL, targets
 * the last wave, tiled or sector search handed back to the walk without a
 * shell table; BK_STAT_SEARCH_RETRY, targets of the last capped search whose
 * candidate pool outgrew its 512 LDS entries and which were searched again
 * with a pool of 65,536 entries in device memory (beyond that, BK_UNSUPPORTED);
 * BK_STAT_GRAPH_LAUNCHES, bk_krige calls run as a graph so far;
 * BK_STAT_SGS_PASSES, the draw passes of the last bk_sgs. Of the last
 * kriging pass: BK_STAT_REFINED, systems solved by FP32 elimination and FP64
 * refinement; BK_STAT_REFINE_SWEEPS, their refinement sweeps in total;
 * BK_STAT_REFINE_REDO, systems the refined form handed to the FP64 kernel
 * (a small pivot, no convergence, or two samples within a few micrometres). */
#define BK_STAT_SEARCH_SPILL 0
#define BK_STAT_GRAPH_LAUNCHES 1
#define BK_STAT_REFINED 2
#define BK_STAT_REFINE_SWEEPS 3
#define BK_STAT_REFINE_REDO 4
#define BK_STAT_SEARCH_RETRY 5
#define BK_STAT_SGS_PASSES 6
int bk_get_stat(bk_ctx* ctx, int32_t which, int64_t* value);

/* Stage timing with HIP events recorded on the context's stream (no
 * counterpart in the reference). bk_profile(ctx, 1) starts accumulating;
 * bk_profile_read returns the summed milliseconds of each stage since the last
 * read and clears them: ms[0] grid build, ms[1] search kernel, ms[2] bucketing,
 * ms[3] solve kernels of the size buckets, ms[4] deferred systems (FP64 redo,
 * merged and LU systems), ms[5] the whole call; calls = calls timed. */
int bk_profile(bk_ctx* ctx, int32_t enable);
int bk_profile_read(bk_ctx* ctx, double* ms6, int64_t* calls);

#ifdef __cplusplus
}
#endif

#endif /* BOITATA_GPU_H */
