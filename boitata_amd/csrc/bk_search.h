// bk_search.h — neighbour search of the MI355X kriging engine.
//
// Replaces SearchTree (crates/estimation/src/search.rs:692-1006) and its cell
// grid (crates/estimation/src/search/grid.rs). The result is defined by the
// reference alone: candidates within the radius in increasing
// (IEEE bits of the FP64 squared distance in the search frame, sample index),
// offered greedily to the selector of search.rs:476-577. The device grid is an
// acceleration structure only; its cell size is free.
#pragma once

#include <stdint.h>

#include "bk_device.h"

#define BK_POOL 512       // candidates held per target between releases (LDS)
#define BK_POOL_BIG 65536 // the same in global memory, for the targets that overflow BK_POOL
#define BK_BIG_BLOCKS 32  // blocks (pools) of the global-memory pass
#define BK_MAX_KNN 64     // selections without caps up to this size run one target per lane
// Cells of the shell table on each side of the target's cell; a walk goes on
// by rings beyond it. The reference's table reaches 12 (grid.rs:11); the
// selections do not depend on the reach, and 24 keeps the walks of sparse
// neighbourhoods (simulation nodes early on the path) in the wave searches
// instead of the lane-per-target ring walk.
#define BK_REACH 24

// Internal per-target status: the candidate pool overflowed (BK_POOL, then
// BK_POOL_BIG entries pending between two releases).
#define BK_ST_OVERFLOW 100

struct GridDev {
  double lo[3];
  double h;
  int dims[3];
  const uint32_t* start;      // cells + 1
  const double4* entries;     // projected x, y, z; w holds the sample index bits
};

struct ShellTable {
  const char4* offsets;       // sorted by m = sum max(0, |o| - 1)^2
  const uint32_t* end;        // end of each shell in offsets
  const uint32_t* m;          // m of each shell
  int n;
};

struct SearchArgs {
  GridDev g;
  ShellTable tab;
  double F[9];                // search frame (Anisotropy::matrix() or identity)
  double origin[3];           // first finite sample (search.rs:714-734)
  double radius2;
  int max_samples, min_samples;
  int balances;               // octant or plane sectors
  int octant_rank;            // octants without a hole cap: prune by per-octant rank
  int n_sectors;              // 8, 4 (planar octants) or the plane sector count
  int plane;                  // plane sectors (atan2 in the major/semi plane)
  int planar;
  int sector_cap;
  int max_per_hole;           // -1 none
  int min_sectors, max_empty_sectors, min_holes;  // -1 none
  double axes[9];             // block_frame of the metric's angles (search.rs:404-417)
  // Octant or quadrant of a sample from its projected offset u = E - q when
  // axes F^-1 is a signed permutation times a positive diagonal (always, for
  // F = D R): sec_code holds the sector of each sign pattern of u (3 bits per
  // pattern (u0 >= 0) | (u1 >= 0) << 1 | (u2 >= 0) << 2); a component with
  // |u_k| <= sec_tol[k] (|E| + |q|) (1-norms) takes the world computation.
  int sec_fast;
  uint32_t sec_code;
  double sec_tol[3];
  const double* sx;
  const double* sy;
  const double* sz;
  const int32_t* hole;        // may be null
  const double* tx;
  const double* ty;
  const double* tz;
  const int64_t* left_out;    // may be null
  const int32_t* limit;       // per target: only samples below this index are candidates (may be null)
  const int32_t* list;        // targets to search (null: 0..m)
  int64_t m;                  // their number, or a bound on it when m_dev is set
  const int32_t* m_dev;       // the number on the device (a spill count), or null
  int32_t* ovf;               // targets whose LDS pool overflows are added here (may be null)
  int32_t* ovf_count;
  int32_t* idx;               // m * stride
  int32_t* count;
  int32_t* status;
  int stride;
};

// Launchers (bk_search.hip).
namespace bk {
// Projects samples, bounds them and counting-sorts them into cells of
// side cell_factor * (the reference's side rule), on the device.
struct GridBuffers {
  double4* entries = nullptr;
  uint32_t* start = nullptr;
  uint32_t* cell_of = nullptr;
  uint32_t* fill = nullptr;
  double* px = nullptr;       // projected samples (SoA, reused by the scratch of the build)
  double* bounds = nullptr;   // partial min / max
  int64_t cap_n = 0, cap_cells = 0, cap_blocks = 0;
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  // Buffers of fixed capacity owned by the caller: the build allocates
  // nothing, widens its cells until they fit cap_cells (the result does not
  // depend on the cell size) and fails where n or the scan's storage does not fit.
  bool fixed = false;
};
int build_grid(hipStream_t s, int64_t n, const double* x, const double* y, const double* z, const double F[9],
               const double origin[3], double cell_factor, GridBuffers& gb, GridDev& out);
// Bytes of hipCUB scan storage the grid build needs for up to `cells` cells.
size_t grid_scan_bytes(int64_t cells);
void free_grid(GridBuffers& gb);
int upload_shell_table(ShellTable& tab, void** storage);
// Scratch of the tiled search: Morton keys and target ids (sorted copies),
// the spill list and its count, the radix sort's temporary storage.
struct TileBuffers {
  uint64_t* keys = nullptr;
  int32_t* ids = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int64_t cap_keys = 0, cap_ids = 0;
  const int32_t* spill_count = nullptr;  // device count of the targets the last wave or tile search left to k_knn
  bool fixed = false;                    // caller-owned ids of fixed capacity: the searches allocate nothing
  void* big = nullptr;                   // BK_BIG_BLOCKS global-memory selector pools (allocated on first use)
  const int32_t* retry_count = nullptr;  // device count of the targets the last capped search searched again
};
void free_tiles(TileBuffers& tb);
// SearchTree::balanced order of each target's selection (search.rs:909-936),
// in place; selections of at most 64 samples.
int launch_balance(hipStream_t s, const SearchArgs& a, int64_t m);
// p[0 .. n) = 0 by a kernel (graph-safe).
int zero32(hipStream_t s, int32_t* p, int64_t n);
// Sequential simulation bookkeeping over m nodes visited in `path`: datum[i],
// the first datum at node i's exact location (-1 none; a's grid holds the
// data); for the nodes not on a datum, pos[node] their rank along the path
// and app[rank] the node. work: 3 m + 2 ints. out[0]: 1 when the path is not
// a permutation of 0 .. m-1; out[1]: the nodes not on a datum.
int sgs_order(hipStream_t s, const SearchArgs& a, int64_t m, const double* gx, const double* gy, const double* gz,
              const int64_t* path, int64_t* datum, int32_t* pos, int64_t* app, int32_t* work, void** tmp,
              size_t* tmp_bytes, int32_t out[2]);
// Searches without caps by `mode`: 0 one wavefront per target (k_knn_wave),
// 1..8 sorted tiles staging the cells within `mode` cells (k_knn_tile), -1
// one lane per target (k_knn). Capped searches always run k_search.
int launch_search(hipStream_t s, const SearchArgs& a, int64_t n_targets, int mode, TileBuffers& tb);
bool knn_path(const SearchArgs& a);
}  // namespace bk
