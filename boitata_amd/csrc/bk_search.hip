// bk_search.hip — device grid and nearest-first neighbour search.
//
// The selection restates crates/estimation/src/search.rs:476-577 (Selector)
// and :951-1005 (SearchTree::neighbors_around / offer) over candidates ordered
// by (FP64 squared distance bits, sample index), the order the reference's grid
// releases them in (crates/estimation/src/search/grid.rs:13-15,115-127,209-255).
// Distances are computed exactly as the reference computes them: the frame
// product in nalgebra's order and d2 = ((a*a + b*b) + c*c), unfused (this file
// is compiled with -ffp-contract=off), so the neighbour sets are bit-identical.
//
// One wavefront per target. Cells are visited shell by shell (offsets within
// 12 cells sorted by the squared lower bound of their distance, then rings);
// before a shell is scanned every pooled candidate below its bound is released
// in order to the selector. Candidates travel in batches of at most 64, one per
// lane: sorted across the wave, merged into a sorted LDS pool, and pruned where
// the reference could never select them (beyond max_samples, a full octant or
// hole, or ranked past what is left of an octant's cap).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "bk_search.h"

namespace {

__device__ __forceinline__ uint64_t dbits(double d) { return (uint64_t)__double_as_longlong(d); }

__device__ __forceinline__ bool key_less(uint64_t ka, uint32_t ia, uint64_t kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = __shfl_xor((unsigned)(v & 0xffffffffu), m, BK_WAVE);
  uint32_t hi = __shfl_xor((unsigned)(v >> 32), m, BK_WAVE);
  return ((uint64_t)hi << 32) | lo;
}

// Ascending bitonic sort of one candidate per lane, by (key, idx).
__device__ __forceinline__ void wave_sort(uint64_t& k, uint32_t& i, int& sec, int& hl, int lane) {
#pragma unroll
  for (int size = 2; size <= BK_WAVE; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      uint64_t ok = shfl_xor_u64(k, stride);
      uint32_t oi = __shfl_xor(i, stride, BK_WAVE);
      int os = __shfl_xor(sec, stride, BK_WAVE);
      int oh = __shfl_xor(hl, stride, BK_WAVE);
      bool ascending = (lane & size) == 0;
      bool lower = (lane & stride) == 0;
      bool other_less = key_less(ok, oi, k, i);
      bool take = (lower == ascending) ? other_less : key_less(k, i, ok, oi);
      if (take) {
        k = ok;
        i = oi;
        sec = os;
        hl = oh;
      }
    }
  }
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int o = 1; o < BK_WAVE; o <<= 1) {
    int u = __shfl_up(v, o, BK_WAVE);
    if (lane >= o) v += u;
  }
  return v;
}

__device__ __forceinline__ int clamp_cell(double v, int dim) {
  // Rust `as i64` saturates (NaN -> 0) before the clamp (grid.rs:183-188).
  double f = floor(v);
  f = fmax(f, -1.0);
  f = fmin(f, (double)dim);
  int c = (int)f;
  return c < 0 ? 0 : (c > dim - 1 ? dim - 1 : c);
}

// The selector's state for one target: a pool of POOL candidates held
// between releases, sorted by (key, idx), and the selection so far. In LDS
// with BK_POOL entries; in global memory with BK_POOL_BIG for the targets
// whose pool overflows the first.
template <int POOL>
struct SearchPool {
  static constexpr int kPool = POOL;
  uint64_t pk[POOL];
  uint32_t pi[POOL];
  int16_t ps[POOL];  // sector (< 360)
  int ph[POOL];
  uint64_t bk[BK_WAVE];
  uint32_t bi[BK_WAVE];
  int cs[BK_WAVE];
  int ex[BK_WAVE];
  int chosen[BK_MAX_SEARCH];
  int hole_id[BK_MAX_SEARCH];
  int hole_cnt[BK_MAX_SEARCH];
  int16_t per_sector[360];  // selected per sector (<= max_samples)
  uint64_t thr_sec[8];
  int scalar[4];
};
using SearchLds = SearchPool<BK_POOL>;
using SearchBig = SearchPool<BK_POOL_BIG>;

struct State {
  int P;          // pooled candidates
  int n;          // chosen
  int n_holes;    // distinct holes in hole_id (with a hole cap)
  int untagged;
  int overflow;
  uint64_t thr;   // with a pool full of sure picks: keys above are useless
  uint32_t thr_i;
};

// Number of pool entries with (key, idx) < (k, i).
template <class Pool>
__device__ __forceinline__ int pool_lower(const Pool& L, int P, uint64_t k, uint32_t i) {
  int lo = 0, hi = P;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (key_less(L.pk[mid], L.pi[mid], k, i))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

template <class Pool>
__device__ __forceinline__ int batch_lower(const Pool& L, int nb, uint64_t k, uint32_t i) {
  int lo = 0, hi = nb;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (key_less(L.bk[mid], L.bi[mid], k, i))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Keeps the pool entries whose flag is set, in order. Chunks of one entry per
// lane from the bottom: an entry only moves down, so a chunk's destinations
// never reach entries not yet read. keep(in, j, sector) is called once per
// chunk by the whole wave (it may use wave-wide operations).
template <class Pool, class Keep>
__device__ __forceinline__ int pool_compact(Pool& L, int P, int lane, Keep keep) {
  int out = 0;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int base = 0; base < P; base += BK_WAVE) {
    int j = base + lane;
    bool in = j < P;
    uint64_t k = 0;
    uint32_t ii = 0;
    int s = -1, h = -1;
    if (in) {
      k = L.pk[j];
      ii = L.pi[j];
      s = L.ps[j];
      h = L.ph[j];
    }
    bool kp = keep(in, j, s);
    uint64_t mask = __ballot(kp);
    int pos = out + __popcll(mask & lt);
    wave_sync();
    if (kp) {
      L.pk[pos] = k;
      L.pi[pos] = ii;
      L.ps[pos] = (int16_t)s;
      L.ph[pos] = h;
    }
    wave_sync();
    out += __popcll(mask);
  }
  return out;
}

// Sectors::of (search.rs:428-441): the sector of sample i around the target at
// world (t0, t1, t2).
__device__ __forceinline__ int sector_at(const SearchArgs& a, double t0, double t1, double t2, uint32_t i) {
  double d0, d1, d2v;
  mv3(a.axes, a.sx[i] - t0, a.sy[i] - t1, a.sz[i] - t2, d0, d1, d2v);
  if (a.plane) {
    // search.rs:430-434
    const double tau = 6.283185307179586;
    double turn = fmod(atan2(-d0, d1), tau);
    if (turn < 0.0) turn = turn + tau;
    double f = turn / tau * (double)a.plane;
    int k = f >= 0.0 ? (f < 2147483647.0 ? (int)f : 2147483647) : 0;
    return k < a.plane - 1 ? k : a.plane - 1;
  }
  int bx = d0 >= 0.0, by = d1 >= 0.0, bz = d2v >= 0.0;
  return a.planar ? (bx << 1) | by : (bx << 2) | (by << 1) | bz;
}

// Batches up to this size are ranked by reading the others' keys lane by lane
// rather than sorted across the wave.
constexpr int kSmallBatch = 16;

template <class Pool>
struct Ctx {
  const SearchArgs* a;
  Pool* L;
  State st;
  int lane;
  double q0, q1, q2;
  double t0, t1, t2;
  int64_t left_out;
  uint32_t limit;  // samples from this index on are not candidates
  int capped;
  __device__ bool full() const { return st.n >= a->max_samples; }

  __device__ int sector_of(uint32_t i) const { return sector_at(*a, t0, t1, t2, i); }

  __device__ int hole_count(int h) const {
    for (int k = 0; k < st.n_holes; ++k)
      if (L->hole_id[k] == h) return L->hole_cnt[k];
    return 0;
  }

  // After a merge: drop what can never be selected, and set the thresholds.
  __device__ void prune() {
    int rem = a->max_samples - st.n;
    if (!capped) {
      if (st.P > rem) st.P = rem;
      if (st.P == rem && rem > 0) {
        st.thr = L->pk[st.P - 1];
        st.thr_i = L->pi[st.P - 1];
      }
      return;
    }
    if (a->octant_rank) {
      // Per-octant rank among pooled candidates: rank >= what is left of the
      // octant's cap is never selected (no hole cap, so an earlier candidate
      // of the octant is selected unless the selection is full first).
      int nsec = a->n_sectors;
      int running[8];
      for (int o = 0; o < 8; ++o) running[o] = 0;
      int sure = 0;
      const uint64_t lt = (1ull << lane) - 1ull;
      Pool* Lc = L;
      const int cap = a->sector_cap;
      // Thresholds: an octant holding all it may still take makes larger keys
      // of that octant useless; its threshold is the key of its last kept
      // entry, whose rank is what is left of its cap minus one (the kept
      // entries of an octant are its ranks 0, 1, ... in pool order).
      if (lane < 8) L->thr_sec[lane] = ~0ull;
      wave_sync();
      st.P = pool_compact(*L, st.P, lane, [&](bool in, int j, int s) {
        int rank = 0;
        for (int o = 0; o < nsec; ++o) {
          uint64_t mask = __ballot(in && s == o);
          if (in && s == o) rank = running[o] + __popcll(mask & lt);
          running[o] += __popcll(mask);
        }
        bool k = in && rank < cap - Lc->per_sector[s];
        // Sure picks in order; beyond what is left of max_samples none is taken.
        uint64_t km = __ballot(k);
        int before = sure + __popcll(km & lt);
        sure += __popcll(km);
        bool kept = k && before < rem;
        if (kept && rank == cap - Lc->per_sector[s] - 1) Lc->thr_sec[s] = Lc->pk[j];
        return kept;
      });
      wave_sync();
      // A pool of `rem` sure picks makes larger keys useless.
      if (st.P == rem && rem > 0) {
        st.thr = L->pk[st.P - 1];
        st.thr_i = L->pi[st.P - 1];
      }
    }
  }

  // Merges this lane's candidate (have) into the pool.
  __device__ void merge(bool have, uint64_t key, uint32_t idx, int sec, int hl) {
    uint64_t mask = __ballot(have);
    if (mask == 0) return;
    int nb = __popcll(mask);
    // Final slot of each batch entry, against the pool before it moves.
    int bpos = -1;
    if (nb <= kSmallBatch) {
      // Few candidates: each counts the batch entries below it (the others'
      // keys read lane by lane) instead of a 64-lane sort; the ranks give the
      // same sorted batch, entries staying in their lanes.
      int rank = 0;
      uint64_t m = mask;
      for (int q = 0; q < nb; ++q) {
        const int src = __builtin_ctzll(m);
        m &= m - 1;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, src);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(key >> 32), src);
        const uint32_t iq = (uint32_t)__builtin_amdgcn_readlane((int)idx, src);
        rank += key_less(((uint64_t)hi << 32) | lo, iq, key, idx) ? 1 : 0;
      }
      if (st.P + nb > Pool::kPool) {
        st.overflow = 1;
        return;
      }
      if (have) {
        L->bk[rank] = key;
        L->bi[rank] = idx;
      }
      wave_sync();
      if (have) bpos = rank + pool_lower(*L, st.P, key, idx);
    } else {
      if (!have) {
        key = ~0ull;
        idx = 0xffffffffu;
        sec = 0;
        hl = -1;
      }
      wave_sort(key, idx, sec, hl, lane);
      if (st.P + nb > Pool::kPool) {
        st.overflow = 1;
        return;
      }
      L->bk[lane] = key;
      L->bi[lane] = idx;
      wave_sync();
      bpos = lane < nb ? lane + pool_lower(*L, st.P, key, idx) : -1;
    }
    // Pool entries move up by the batch entries below them; chunks from the
    // top, so a destination never holds an entry not yet read.
    for (int base = ((st.P - 1) / BK_WAVE) * BK_WAVE; base >= 0 && st.P > 0; base -= BK_WAVE) {
      int j = base + lane;
      int pos = -1;
      uint64_t k = 0;
      uint32_t ii = 0;
      int s = 0, h = -1;
      if (j < st.P) {
        k = L->pk[j];
        ii = L->pi[j];
        s = L->ps[j];
        h = L->ph[j];
        pos = j + batch_lower(*L, nb, k, ii);
      }
      wave_sync();
      if (pos >= 0) {
        L->pk[pos] = k;
        L->pi[pos] = ii;
        L->ps[pos] = (int16_t)s;
        L->ph[pos] = h;
      }
      wave_sync();
    }
    if (bpos >= 0) {
      L->pk[bpos] = key;
      L->pi[bpos] = idx;
      L->ps[bpos] = (int16_t)sec;
      L->ph[bpos] = hl;
    }
    wave_sync();
    st.P += nb;
    prune();
  }

  // Scans the cell `id` of this lane (if valid); candidates go to the pool.
  __device__ void scan_cells(bool valid, int id) {
    int cs = 0, cnt = 0;
    if (valid) {
      cs = (int)a->g.start[id];
      cnt = (int)a->g.start[id + 1] - cs;
    }
    int inc = wave_incl_scan(cnt, lane);
    int total = bcasti(inc, BK_WAVE - 1);
    if (total == 0) return;
    L->cs[lane] = cs;
    L->ex[lane] = inc - cnt;
    wave_sync();
    for (int base = 0; base < total; base += BK_WAVE) {
      int f = base + lane;
      bool have = false;
      uint64_t key = 0;
      uint32_t idx = 0;
      int sec = 0, hl = -1;
      if (f < total) {
        int lo = 0, hi = BK_WAVE - 1;
        while (lo < hi) {  // largest k with ex[k] <= f
          int mid = (lo + hi + 1) >> 1;
          if (L->ex[mid] <= f)
            lo = mid;
          else
            hi = mid - 1;
        }
        int e = L->cs[lo] + (f - L->ex[lo]);
        double4 E = a->g.entries[e];
        double u0 = E.x - q0, u1 = E.y - q1, u2 = E.z - q2;
        double d2 = (u0 * u0 + u1 * u1) + u2 * u2;
        idx = (uint32_t)__double_as_longlong(E.w);
        if (d2 <= a->radius2 && (int64_t)idx != left_out && idx < limit) {
          key = dbits(d2);
          have = key_less(key, idx, st.thr, st.thr_i);
          if (have && a->balances) {
            sec = sector_of(idx);
            if (L->per_sector[sec] >= a->sector_cap) have = false;
            if (have && a->octant_rank && key > L->thr_sec[sec]) have = false;
          }
          if (have && a->hole) {
            hl = a->hole[idx];
            if (a->max_per_hole >= 0 && hl >= 0 && hole_count(hl) >= a->max_per_hole) have = false;
          }
        }
      }
      merge(have, key, idx, sec, hl);
      if (st.overflow) return;
    }
    wave_sync();
  }

  // Offers every pooled candidate with d2 < bound, in order.
  __device__ void release(double bound) {
    if (st.P == 0) return;
    uint64_t b = dbits(bound);
    int r = pool_lower(*L, st.P, b, 0);
    if (r == 0) return;
    if (!capped) {
      int take = min(r, a->max_samples - st.n);
      for (int k = lane; k < take; k += BK_WAVE) L->chosen[st.n + k] = (int)L->pi[k];
      st.n += take;
    } else if (a->octant_rank) {
      // Octants without a hole cap: after prune every pooled entry is a sure
      // pick (its rank in its octant is below what is left of the octant's
      // cap), so the released entries are taken in order up to max_samples,
      // as the one-by-one Selector::offer would take them.
      int take = min(r, a->max_samples - st.n);
      const int nsec = a->n_sectors;
      for (int base = 0; base < take; base += BK_WAVE) {
        int e = base + lane;
        bool in = e < take;
        int s = in ? L->ps[e] : -1;
        if (in) L->chosen[st.n + e] = (int)L->pi[e];
        for (int o = 0; o < nsec; ++o) {
          int add = __popcll(__ballot(in && s == o));
          if (lane == 0 && add) L->per_sector[o] = (int16_t)(L->per_sector[o] + add);
        }
      }
      wave_sync();
      st.n += take;
    } else {
      if (lane == 0) {
        int n = st.n, nh = st.n_holes, un = st.untagged;
        for (int e = 0; e < r && n < a->max_samples; ++e) {
          int hl = L->ph[e];
          int slot = -1;
          if (hl >= 0)
            for (int k = 0; k < nh; ++k)
              if (L->hole_id[k] == hl) {
                slot = k;
                break;
              }
          if (a->max_per_hole >= 0 && slot >= 0 && L->hole_cnt[slot] >= a->max_per_hole) continue;
          if (a->balances) {
            int s = L->ps[e];
            if (L->per_sector[s] >= a->sector_cap) continue;
            L->per_sector[s] = (int16_t)(L->per_sector[s] + 1);
          }
          if (hl >= 0) {
            if (slot >= 0)
              L->hole_cnt[slot] += 1;
            else {
              L->hole_id[nh] = hl;
              L->hole_cnt[nh] = 1;
              nh++;
            }
          } else {
            un++;
          }
          L->chosen[n++] = (int)L->pi[e];
        }
        L->scalar[0] = n;
        L->scalar[1] = nh;
        L->scalar[2] = un;
      }
      wave_sync();
      st.n = L->scalar[0];
      st.n_holes = L->scalar[1];
      st.untagged = L->scalar[2];
    }
    st.P = pool_compact(*L, st.P, lane, [&](bool in, int j, int) { return in && j >= r; });
    st.thr = ~0ull;
    st.thr_i = 0xffffffffu;
    if (!full()) prune();
  }
};

// Selector::finish (search.rs:547-576): the status of a selection of n
// samples chosen[0 .. n) holding per_sector[s] samples of sector s.
__device__ int selection_status(const SearchArgs& a, int n, const int16_t* per_sector, const int* chosen) {
  if (n < a.min_samples) return BK_SEARCH_FAILED;
  if (a.min_sectors >= 0) {
    int held = 0;
    for (int s = 0; s < a.n_sectors; ++s) held += per_sector[s] > 0;
    if (held < a.min_sectors) return BK_SEARCH_FAILED;
  }
  if (a.max_empty_sectors >= 0) {
    // search.rs:445-467: longest run of empty sectors around each ring.
    int best = 0;
    const int rings = a.plane ? 1 : (a.planar ? 1 : 2);
    for (int ri = 0; ri < rings; ++ri) {
      const int len = a.plane ? a.plane : 4;
      auto at = [&](int k) {
        if (a.plane) return k;
        if (a.planar) {
          const int r0[4] = {3, 1, 0, 2};
          return r0[k];
        }
        const int r1[4] = {7, 3, 1, 5};
        const int r2[4] = {6, 2, 0, 4};
        return ri == 0 ? r1[k] : r2[k];
      };
      bool all = true;
      for (int k = 0; k < len; ++k) all = all && per_sector[at(k)] == 0;
      if (all) {
        best = max(best, len);
        continue;
      }
      int run = 0, longest = 0;
      for (int pass = 0; pass < 2; ++pass)
        for (int k = 0; k < len; ++k) {
          run = per_sector[at(k)] == 0 ? run + 1 : 0;
          longest = max(longest, run);
        }
      best = max(best, longest);
    }
    if (best > a.max_empty_sectors) return BK_SEARCH_FAILED;
  }
  if (a.min_holes >= 0) {
    int distinct = 0;
    for (int k = 0; k < n; ++k) {
      const int h = a.hole ? a.hole[chosen[k]] : -1;
      if (h < 0) {
        distinct++;
        continue;
      }
      bool seen = false;
      for (int j = 0; j < k && !seen; ++j) seen = a.hole[chosen[j]] == h;
      distinct += !seen;
    }
    if (distinct < a.min_holes) return BK_SEARCH_FAILED;
  }
  return BK_OK;
}

// The selection of target t by one wavefront (L: its selector state). A
// target whose pool overflows is added to a.ovf when that list is given.
template <class Pool>
__device__ void search_target(const SearchArgs& a, Pool& L, int64_t t) {
  const int lane = threadIdx.x;
  Ctx<Pool> c;
  c.a = &a;
  c.L = &L;
  c.lane = lane;
  c.st = State{0, 0, 0, 0, 0, ~0ull, 0xffffffffu};
  c.capped = a.balances || a.max_per_hole >= 0;
  c.left_out = a.left_out ? a.left_out[t] : -1;
  c.limit = a.limit ? (uint32_t)a.limit[t] : 0xffffffffu;
  c.t0 = a.tx[t];
  c.t1 = a.ty[t];
  c.t2 = a.tz[t];
  for (int k = lane; k < 360; k += BK_WAVE) L.per_sector[k] = 0;
  if (lane < 8) L.thr_sec[lane] = ~0ull;
  wave_sync();
  // search.rs:738-740: the target in the search frame.
  mv3(a.F, c.t0 - a.origin[0], c.t1 - a.origin[1], c.t2 - a.origin[2], c.q0, c.q1, c.q2);
  const int n_samples_nonempty = a.g.start != nullptr;
  if (n_samples_nonempty && a.max_samples > 0) {
    const GridDev& g = a.g;
    int c0 = clamp_cell((c.q0 - g.lo[0]) / g.h, g.dims[0]);
    int c1 = clamp_cell((c.q1 - g.lo[1]) / g.h, g.dims[1]);
    int c2 = clamp_cell((c.q2 - g.lo[2]) / g.h, g.dims[2]);
    double h2 = g.h * g.h;
    double beyond = h2 * (double)((BK_REACH - 1) * (BK_REACH - 1));
    const double slack = 1.0 - 1e-9;
    bool done = false;
    uint32_t begin = 0;
    for (int s = 0; s < a.tab.n && !done; ++s) {
      double bound = fmin(h2 * (double)a.tab.m[s], beyond) * slack;
      if (bound > a.radius2) break;
      c.release(bound);
      if (c.full()) {
        done = true;
        break;
      }
      uint32_t end = a.tab.end[s];
      for (uint32_t chunk = begin; chunk < end; chunk += BK_WAVE) {
        uint32_t o = chunk + lane;
        bool valid = o < end;
        int id = 0;
        if (valid) {
          char4 off = a.tab.offsets[o];
          int x = c0 + off.x, y = c1 + off.y, z = c2 + off.z;
          valid = x >= 0 && x < g.dims[0] && y >= 0 && y < g.dims[1] && z >= 0 && z < g.dims[2];
          id = x + g.dims[0] * (y + g.dims[1] * z);
        }
        c.scan_cells(valid, id);
        if (c.st.overflow) break;
      }
      if (c.st.overflow) done = true;
      begin = end;
    }
    if (!done) {
      int far = max(max(max(c0, g.dims[0] - 1 - c0), max(c1, g.dims[1] - 1 - c1)), max(c2, g.dims[2] - 1 - c2));
      for (int r = BK_REACH + 1; r <= far && !done; ++r) {
        double x = (double)(r - 1) * g.h;
        double bound = x * x * slack;
        if (bound > a.radius2) break;
        c.release(bound);
        if (c.full()) {
          done = true;
          break;
        }
        int xlo = max(-r, -c0), xhi = min(r, g.dims[0] - 1 - c0);
        int ylo = max(-r, -c1), yhi = min(r, g.dims[1] - 1 - c1);
        int zlo = max(-r, -c2), zhi = min(r, g.dims[2] - 1 - c2);
        int nx = xhi - xlo + 1, ny = yhi - ylo + 1, nz = zhi - zlo + 1;
        int pairs = nx * ny;
        for (int pb = 0; pb < pairs && !c.st.overflow; pb += BK_WAVE) {
          int p = pb + lane;
          bool valid = p < pairs;
          int x0 = xlo + (valid ? p % nx : 0), y0 = ylo + (valid ? p / nx : 0);
          bool side = valid && (abs(x0) == r || abs(y0) == r);
          int zc = !valid ? 0 : (side ? nz : (int)(zlo <= -r) + (int)(r <= zhi));
          int zmax = zc;
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) zmax = max(zmax, __shfl_xor(zmax, o, BK_WAVE));
          for (int k = 0; k < zmax && !c.st.overflow; ++k) {
            bool v = k < zc;
            int z = side ? zlo + k : ((zlo <= -r && k == 0) ? -r : r);
            int id = (x0 + c0) + g.dims[0] * ((y0 + c1) + g.dims[1] * (z + c2));
            c.scan_cells(v, v ? id : 0);
          }
        }
        if (c.st.overflow) done = true;
      }
    }
    if (!done) c.release(INFINITY);
  }
  wave_sync();
  // Selector::finish (search.rs:547-576).
  int n = c.st.n;
  int* out = a.idx + (int64_t)t * a.stride;
  for (int k = lane; k < n; k += BK_WAVE) out[k] = L.chosen[k];
  if (lane == 0) {
    const int status = c.st.overflow ? BK_ST_OVERFLOW : selection_status(a, n, L.per_sector, L.chosen);
    a.count[t] = n;
    a.status[t] = status;
    if (c.st.overflow && a.ovf) a.ovf[atomicAdd(a.ovf_count, 1)] = (int32_t)t;
  }
}

// One block per target, or with a.m_dev (a list whose length is on the
// device) blocks striding over the list.
__global__ void __launch_bounds__(BK_WAVE) k_search(const SearchArgs a) {
  __shared__ SearchLds L;
  const int64_t end = a.m_dev ? (int64_t)*a.m_dev : (int64_t)blockIdx.x + 1;
  for (int64_t b = blockIdx.x; b < end; b += gridDim.x) {
    search_target(a, L, a.list ? a.list[b] : b);
    wave_sync();
  }
}

// The targets whose LDS pool overflowed (a.list, length *a.m_dev), again with
// a pool of BK_POOL_BIG entries in global memory: block b uses ws[b].
__global__ void __launch_bounds__(BK_WAVE) k_search_big(const SearchArgs a, SearchBig* ws) {
  SearchBig& L = ws[blockIdx.x];
  const int64_t end = (int64_t)*a.m_dev;
  for (int64_t b = blockIdx.x; b < end; b += gridDim.x) {
    search_target(a, L, (int64_t)a.list[b]);
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// Searches without caps (no octants, sectors or hole cap): the selection is
// the K = max_samples smallest (d2 bits, index) keys within the radius, in
// that order (search.rs:519-543 reduces to "take until full"). One target per
// lane: the lane walks the same shells with a max-heap of K keys in LDS and
// stops once a shell's lower bound exceeds the heap's largest distance, since
// no point of that shell or a later one can enter (grid.rs:209-255 bounds).

struct KnnHeap {
  uint64_t* k;  // [K][64], interleaved by lane
  uint32_t* i;
  int lane;
  __device__ uint64_t key(int e) const { return k[e * BK_WAVE + lane]; }
  __device__ uint32_t idx(int e) const { return i[e * BK_WAVE + lane]; }
  __device__ void put(int e, uint64_t kk, uint32_t ii) {
    k[e * BK_WAVE + lane] = kk;
    i[e * BK_WAVE + lane] = ii;
  }
  // Max-heap on (key, idx).
  __device__ void push(int& size, uint64_t kk, uint32_t ii) {
    int e = size++;
    while (e > 0) {
      int p = (e - 1) >> 1;
      uint64_t pk = key(p);
      uint32_t pi = idx(p);
      if (!key_less(pk, pi, kk, ii)) break;
      put(e, pk, pi);
      e = p;
    }
    put(e, kk, ii);
  }
  __device__ void replace_top(int size, uint64_t kk, uint32_t ii) {
    int e = 0;
    for (;;) {
      int l = 2 * e + 1;
      if (l >= size) break;
      int r = l + 1;
      uint64_t ck = key(l);
      uint32_t ci = idx(l);
      if (r < size) {
        uint64_t rk = key(r);
        uint32_t ri = idx(r);
        if (key_less(ck, ci, rk, ri)) {
          ck = rk;
          ci = ri;
          l = r;
        }
      }
      if (!key_less(kk, ii, ck, ci)) break;
      put(e, ck, ci);
      e = l;
    }
    put(e, kk, ii);
  }
};

// Target b of the launch (lane-wise), the launch holding mm targets.
__device__ __forceinline__ void knn_lane(const SearchArgs& a, int64_t b, int64_t mm) {
  extern __shared__ uint64_t knn_dyn[];
  const int lane = threadIdx.x;
  const int K = a.max_samples;
  KnnHeap H{knn_dyn, (uint32_t*)(knn_dyn + (size_t)K * BK_WAVE), lane};
  const bool live = b < mm;
  const int64_t t = live ? (a.list ? a.list[b] : b) : 0;
  const int64_t left_out = (live && a.left_out) ? a.left_out[t] : -1;
  const uint32_t limit = (live && a.limit) ? (uint32_t)a.limit[t] : 0xffffffffu;
  double q0 = 0, q1 = 0, q2 = 0;
  if (live) mv3(a.F, a.tx[t] - a.origin[0], a.ty[t] - a.origin[1], a.tz[t] - a.origin[2], q0, q1, q2);
  int size = 0;
  uint64_t top = 0;  // largest key while the heap is full
  uint32_t top_i = 0;
  const GridDev& g = a.g;
  bool go = live && g.start != nullptr && K > 0;
  int c0 = 0, c1 = 0, c2 = 0;
  if (go) {
    c0 = clamp_cell((q0 - g.lo[0]) / g.h, g.dims[0]);
    c1 = clamp_cell((q1 - g.lo[1]) / g.h, g.dims[1]);
    c2 = clamp_cell((q2 - g.lo[2]) / g.h, g.dims[2]);
  }
  const double h2 = g.h * g.h;
  const double beyond = h2 * (double)((BK_REACH - 1) * (BK_REACH - 1));
  const double slack = 1.0 - 1e-9;
  auto visit = [&](int x, int y, int z) {
    if (x < 0 || x >= g.dims[0] || y < 0 || y >= g.dims[1] || z < 0 || z >= g.dims[2]) return;
    int id = x + g.dims[0] * (y + g.dims[1] * z);
    uint32_t e0 = g.start[id], e1 = g.start[id + 1];
    for (uint32_t e = e0; e < e1; ++e) {
      double4 E = g.entries[e];
      double u0 = E.x - q0, u1 = E.y - q1, u2 = E.z - q2;
      double d2 = (u0 * u0 + u1 * u1) + u2 * u2;
      uint32_t ii = (uint32_t)__double_as_longlong(E.w);
      if (!(d2 <= a.radius2) || (int64_t)ii == left_out || ii >= limit) continue;
      uint64_t kk = dbits(d2);
      if (size < K) {
        H.push(size, kk, ii);
        if (size == K) {
          top = H.key(0);
          top_i = H.idx(0);
        }
      } else if (key_less(kk, ii, top, top_i)) {
        H.replace_top(size, kk, ii);
        top = H.key(0);
        top_i = H.idx(0);
      }
    }
  };
  uint32_t begin = 0;
  bool open = go;
  for (int s = 0; s < a.tab.n; ++s) {
    uint32_t end = a.tab.end[s];
    double bound = fmin(h2 * (double)a.tab.m[s], beyond) * slack;
    if (open && (bound > a.radius2 || (size == K && bound > __longlong_as_double((long long)top)))) open = false;
    if (!__any(open)) break;
    if (open)
      for (uint32_t o = begin; o < end; ++o) {
        char4 off = a.tab.offsets[o];
        visit(c0 + off.x, c1 + off.y, c2 + off.z);
      }
    begin = end;
  }
  if (__any(open)) {
    int far = 0;
    if (open)
      far = max(max(max(c0, g.dims[0] - 1 - c0), max(c1, g.dims[1] - 1 - c1)), max(c2, g.dims[2] - 1 - c2));
    for (int r = BK_REACH + 1;; ++r) {
      double x = (double)(r - 1) * g.h;
      double bound = x * x * slack;
      if (open && (r > far || bound > a.radius2 || (size == K && bound > __longlong_as_double((long long)top))))
        open = false;
      if (!__any(open)) break;
      if (!open) continue;
      int xlo = max(-r, -c0), xhi = min(r, g.dims[0] - 1 - c0);
      int ylo = max(-r, -c1), yhi = min(r, g.dims[1] - 1 - c1);
      int zlo = max(-r, -c2), zhi = min(r, g.dims[2] - 1 - c2);
      for (int dx = xlo; dx <= xhi; ++dx)
        for (int dy = ylo; dy <= yhi; ++dy) {
          if (abs(dx) == r || abs(dy) == r) {
            for (int dz = zlo; dz <= zhi; ++dz) visit(c0 + dx, c1 + dy, c2 + dz);
          } else {
            if (zlo <= -r) visit(c0 + dx, c1 + dy, c2 - r);
            if (r <= zhi) visit(c0 + dx, c1 + dy, c2 + r);
          }
        }
    }
  }
  if (!live) return;
  // Heap sort: the selection in increasing (d2 bits, index).
  const int n = size;
  for (int end = n - 1; end > 0; --end) {
    uint64_t mk = H.key(0);
    uint32_t mi = H.idx(0);
    uint64_t lk = H.key(end);
    uint32_t li = H.idx(end);
    H.put(end, mk, mi);
    H.replace_top(end, lk, li);
  }
  int* out = a.idx + t * a.stride;
  for (int k = 0; k < n; ++k) out[k] = (int)H.idx(k);
  // Selector::finish (search.rs:547-576) for the rules that apply here.
  int status = n < a.min_samples ? BK_SEARCH_FAILED : BK_OK;
  if (status == BK_OK && a.min_holes >= 0) {
    int distinct = 0;
    for (int k = 0; k < n; ++k) {
      int h = a.hole ? a.hole[out[k]] : -1;
      if (h < 0) {
        distinct++;
        continue;
      }
      bool seen = false;
      for (int j = 0; j < k && !seen; ++j) seen = a.hole[out[j]] == h;
      distinct += !seen;
    }
    if (distinct < a.min_holes) status = BK_SEARCH_FAILED;
  }
  a.count[t] = n;
  a.status[t] = status;
}

// One lane per target; persistent over the targets when their number is read
// on the device (a spill count).
__global__ void __launch_bounds__(BK_WAVE) k_knn(const SearchArgs a) {
  int64_t mm = a.m;
  if (a.m_dev) mm = min<int64_t>(*a.m_dev, a.m);
  for (int64_t base = (int64_t)blockIdx.x * BK_WAVE; base < mm; base += (int64_t)gridDim.x * BK_WAVE) {
    knn_lane(a, base + threadIdx.x, mm);
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// The same selections with one wavefront per target. The K best keys so far
// are held sorted across the lanes (lane l has rank l, +inf padding), so the
// K-th key is a wave-uniform threshold. The wave takes the shell table's cell
// offsets 64 at a time (lane l one cell), offers every entry of those cells,
// keeps the survivors (within the radius, below the threshold) in an LDS
// buffer, and merges them into the list with bitonic networks across lanes.
// It stops at the first offset whose shell's lower bound exceeds the
// threshold (or the radius), exactly where k_knn stops, since the offsets are
// sorted by that bound; visiting cells of a later shell first would only offer
// keys the threshold rejects. A target still open after the table (its walk
// would continue on rings) is handed to k_knn through the spill list.

__device__ __forceinline__ void cx_keys(uint64_t& k, uint32_t& i, int lane, int stride, bool up) {
  const uint64_t ok = shfl_xor_u64(k, stride);
  const uint32_t oi = __shfl_xor(i, stride, BK_WAVE);
  // The lower lane keeps the smaller key when sorting up. Keys are distinct
  // except padding entries, which are identical, so "the other is not
  // smaller" means "the other is larger" or the same entry.
  const bool want_small = ((lane & stride) == 0) == up;
  const bool take = key_less(ok, oi, k, i) == want_small;
  k = take ? ok : k;
  i = take ? oi : i;
}

// Inclusive prefix sum over groups of G lanes (lane is the index within the group).
template <int G>
__device__ __forceinline__ uint32_t group_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, G);
    v += lane >= o ? u : 0u;
  }
  return v;
}

// Ascending bitonic sort of one key per lane over groups of G lanes (lane is
// the index within the group).
template <int G>
__device__ __forceinline__ void group_sort_keys(uint64_t& k, uint32_t& i, int lane) {
#pragma unroll
  for (int size = 2; size <= G; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) cx_keys(k, i, lane, stride, (lane & size) == 0);
  }
}

// A (sorted ascending across the group) becomes the G smallest of A and B (B
// sorted ascending), sorted.
template <int G>
__device__ __forceinline__ void group_merge_keys(uint64_t& ak, uint32_t& ai, uint64_t bk, uint32_t bi, int lane) {
  // Reverse B, take the pairwise minimum: a bitonic sequence of the G smallest.
  const int src = G - 1 - lane;
  const uint32_t lo = __shfl((unsigned)(bk & 0xffffffffu), src, G);
  const uint32_t hi = __shfl((unsigned)(bk >> 32), src, G);
  const uint64_t rk = ((uint64_t)hi << 32) | lo;
  const uint32_t ri = __shfl(bi, src, G);
  if (key_less(rk, ri, ak, ai)) {
    ak = rk;
    ai = ri;
  }
#pragma unroll
  for (int stride = G / 2; stride > 0; stride >>= 1) cx_keys(ak, ai, lane, stride, true);
}

constexpr uint64_t KEY_INF = ~0ull;

// k_knn_wave ranks batches up to this size entry by entry (about 24 VALU per
// entry) and sorts larger ones into the list by networks (about 130 VALU).
constexpr int kSmallMerge = 6;

// The bits of a wave ballot that belong to this lane's group of G lanes.
template <int G>
__device__ __forceinline__ uint64_t group_ballot(bool p, int h) {
  const uint64_t m = __ballot(p);
  return G == BK_WAVE ? m : (m >> (h * G)) & ((1ull << G) - 1ull);
}

// One target per group of G lanes (G = 64, or 32 when max_samples <= 32: two
// targets per wavefront, whose networks, ballots and broadcasts then serve
// both at once). Once K keys are held, a cell whose box lies beyond the K-th
// key is skipped whole: every entry in it would be rejected one by one.
// The walk waits on shuffles and gathers in turn, so it runs at seven
// wavefronts per SIMD (at most 72 registers).
template <int G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
k_knn_wave(const SearchArgs a, int32_t* spill, int32_t* spill_count) {
  constexpr int PER = BK_WAVE / G;  // targets per wavefront
  __shared__ uint64_t buf_k[4][2 * BK_WAVE];
  __shared__ uint32_t buf_i[4][2 * BK_WAVE];
  const int wave = threadIdx.x / BK_WAVE, wl = threadIdx.x % BK_WAVE;
  const int h = wl / G, lane = wl % G;
  const int64_t b = ((int64_t)blockIdx.x * 4 + wave) * PER + h;
  if (((int64_t)blockIdx.x * 4 + wave) * PER >= a.m) return;  // whole wave
  const bool live = b < a.m;
  const int64_t t = live ? (a.list ? a.list[b] : b) : 0;
  const int64_t left_out = (live && a.left_out) ? a.left_out[t] : -1;
  const uint32_t limit = (live && a.limit) ? (uint32_t)a.limit[t] : 0xffffffffu;
  const int K = a.max_samples;
  const GridDev& g = a.g;
  double q0, q1, q2;
  mv3(a.F, a.tx[t] - a.origin[0], a.ty[t] - a.origin[1], a.tz[t] - a.origin[2], q0, q1, q2);
  uint64_t ak = KEY_INF;
  uint32_t ai = 0xffffffffu;
  uint64_t* bk = buf_k[wave] + h * 2 * G;
  uint32_t* bi = buf_i[wave] + h * 2 * G;
  bool open = live && g.start != nullptr && K > 0;
  if (__any(open)) {
    // The target in cells, by one reciprocal: the cell only starts the walk,
    // and a rounding of it is within the slack of the bounds below.
    const double ihc = 1.0 / g.h;
    const double s0 = (q0 - g.lo[0]) * ihc, s1 = (q1 - g.lo[1]) * ihc, s2 = (q2 - g.lo[2]) * ihc;
    const int c0 = clamp_cell(s0, g.dims[0]);
    const int c1 = clamp_cell(s1, g.dims[1]);
    const int c2 = clamp_cell(s2, g.dims[2]);
    const double h2 = g.h * g.h;
    const double beyond = h2 * (double)((BK_REACH - 1) * (BK_REACH - 1));
    const double slack = 1.0 - 1e-9;
    const uint32_t n_off = a.tab.end[a.tab.n - 1];
    // The target's place in its cell, in cells (outside [0, 1) for a target
    // beyond the grid). A cell dx > 0 cells away along an axis has its near
    // face dx - fr cells away, one dx < 0 away |dx| - 1 + fr; up and dn are
    // fr and 1 - fr plus 1e-5 cells for the rounding of the cell assignment
    // and of these FP32 values, so the box distance is never overestimated.
    const double ih2 = ihc * ihc;
    const float fr0 = (float)(s0 - (double)c0);
    const float fr1 = (float)(s1 - (double)c1);
    const float fr2 = (float)(s2 - (double)c2);
    const float up0 = fr0 + 1e-5f, up1 = fr1 + 1e-5f, up2 = fr2 + 1e-5f;
    const float dn0 = (1.0f - fr0) + 1e-5f, dn1 = (1.0f - fr1) + 1e-5f, dn2 = (1.0f - fr2) + 1e-5f;
    int nbuf = 0;  // the same in the group's lanes
    auto merge = [&]() {
      if (nbuf <= kSmallMerge) {
        // Few entries: every list entry and every batch entry counts the
        // entries below it (the batch's keys read from the group's lanes)
        // and is written at its rank in the merged list, in the buffer's
        // upper half; the G smallest of the list and the batch, sorted, as
        // the networks below give them.
        const uint64_t mk = lane < nbuf ? bk[lane] : KEY_INF;
        const uint32_t mi = lane < nbuf ? bi[lane] : 0xffffffffu;
        int below = 0, posb = 0;
        for (int q = 0; q < nbuf; ++q) {
          const uint32_t lo = __shfl((unsigned)(uint32_t)mk, q, G);
          const uint32_t hi = __shfl((unsigned)(uint32_t)(mk >> 32), q, G);
          const uint32_t iq = __shfl(mi, q, G);
          const uint64_t kq = ((uint64_t)hi << 32) | lo;
          below += key_less(kq, iq, ak, ai) ? 1 : 0;
          const int na = __popcll(group_ballot<G>(key_less(ak, ai, kq, iq), h));
          if (lane == q) posb += na;
          if (lane < nbuf && key_less(kq, iq, mk, mi)) posb += 1;
        }
        const int posa = lane + below;
        if (posa < G) {
          bk[G + posa] = ak;
          bi[G + posa] = ai;
        }
        if (lane < nbuf && posb < G) {
          bk[G + posb] = mk;
          bi[G + posb] = mi;
        }
        wave_sync();
        ak = bk[G + lane];
        ai = bi[G + lane];
        wave_sync();
        nbuf = 0;
        return;
      }
      uint64_t k = lane < nbuf ? bk[lane] : KEY_INF;
      uint32_t i = lane < nbuf ? bi[lane] : 0xffffffffu;
      // At most G / 2 entries in every group: sorting groups of G / 2 lanes
      // leaves the padding above them, so the G lanes are sorted too.
      if (__any(nbuf > G / 2))
        group_sort_keys<G>(k, i, lane);
      else
        group_sort_keys<G / 2>(k, i, lane & (G / 2 - 1));
      if (group_ballot<G>(ak != KEY_INF, h) == 0) {
        // Nothing held yet (the first merge): the sorted batch is the list.
        ak = k;
        ai = i;
      } else {
        group_merge_keys<G>(ak, ai, k, i, lane);
      }
      // Keep what did not fit in this merge.
      const int rest = nbuf > G ? nbuf - G : 0;
      uint64_t kk = 0;
      uint32_t ii = 0;
      if (lane < rest) {
        kk = bk[G + lane];
        ii = bi[G + lane];
      }
      wave_sync();
      if (lane < rest) {
        bk[lane] = kk;
        bi[lane] = ii;
      }
      wave_sync();
      nbuf = rest;
    };
    // The threshold: the K-th key held (rank K - 1), +inf until K are held.
    uint64_t tk;
    uint32_t ti;
    auto threshold = [&]() {
      const uint32_t tlo = __shfl((unsigned)(ak & 0xffffffffu), K - 1, G);
      const uint32_t thi = __shfl((unsigned)(ak >> 32), K - 1, G);
      tk = ((uint64_t)thi << 32) | tlo;
      ti = __shfl(ai, K - 1, G);
    };
    for (uint32_t o0 = 0; o0 < n_off; o0 += G) {
      threshold();
      const double tcap = tk != KEY_INF ? fmin(a.radius2, __longlong_as_double((long long)tk)) : a.radius2;
      // Lower bound of every offset from o0 on (sorted by m).
      const char4 f0 = a.tab.offsets[o0];
      const int m0 = max(abs(f0.x) - 1, 0) * max(abs(f0.x) - 1, 0) + max(abs(f0.y) - 1, 0) * max(abs(f0.y) - 1, 0) +
                     max(abs(f0.z) - 1, 0) * max(abs(f0.z) - 1, 0);
      const double bound = fmin(h2 * (double)m0, beyond) * slack;
      if (open && bound > tcap) open = false;
      if (!__any(open)) break;
      // The threshold in squared cells, rounded up.
      const float tcc = (float)(tcap * ih2) * (1.0f + 1e-6f);
      // Lane l: cell o0 + l, whose entries are [e, e + cnt) of the CSR,
      // unless its box lies beyond the threshold.
      uint32_t e = 0, cnt = 0;
      const uint32_t o = o0 + lane;
      if (open && o < n_off) {
        const char4 f = a.tab.offsets[o];
        const int x = c0 + f.x, y = c1 + f.y, z = c2 + f.z;
        const float gx = fmaxf((float)abs(f.x) - (f.x > 0 ? up0 : dn0), 0.0f);
        const float gy = fmaxf((float)abs(f.y) - (f.y > 0 ? up1 : dn1), 0.0f);
        const float gz = fmaxf((float)abs(f.z) - (f.z > 0 ? up2 : dn2), 0.0f);
        const bool far = (gx * gx + gy * gy) + gz * gz > tcc;
        if (x >= 0 && x < g.dims[0] && y >= 0 && y < g.dims[1] && z >= 0 && z < g.dims[2] && !far) {
          const int id = x + g.dims[0] * (y + g.dims[1] * z);
          e = g.start[id];
          cnt = g.start[id + 1] - e;
        }
      }
      // The batch's entries, cell after cell, G per round: entry f of the
      // concatenation lies in the cell of the first lane whose inclusive
      // count exceeds f (a binary search over the group's lanes).
      const uint32_t incl = group_incl_scan<G>(cnt, lane);
      const uint32_t excl = incl - cnt;
      const uint32_t total = __shfl(incl, G - 1, G);
      for (uint32_t f = lane; __any(f - lane < total); f += G) {
        int pos = 0;
#pragma unroll
        for (int step = G / 2; step > 0; step >>= 1) {
          const uint32_t probe = __shfl(incl, pos + step - 1, G);
          pos += probe <= f ? step : 0;
        }
        const uint32_t entry = __shfl(e, pos, G) + (f - __shfl(excl, pos, G));
        bool keep = false;
        uint64_t kk = 0;
        uint32_t ii = 0;
        if (f < total) {
          const double4 E = g.entries[entry];
          const double u0 = E.x - q0, u1 = E.y - q1, u2 = E.z - q2;
          const double d2 = (u0 * u0 + u1 * u1) + u2 * u2;
          ii = (uint32_t)__double_as_longlong(E.w);
          kk = dbits(d2);
          keep = d2 <= a.radius2 && (int64_t)ii != left_out && ii < limit && key_less(kk, ii, tk, ti);
        }
        const uint64_t mask = group_ballot<G>(keep, h);
        if (__any(mask != 0)) {
          const int slot = nbuf + __popcll(mask & ((1ull << lane) - 1ull));
          if (keep) {
            bk[slot] = kk;
            bi[slot] = ii;
          }
          nbuf += __popcll(mask);
          wave_sync();
          if (__any(nbuf >= G)) {
            if (nbuf >= G) merge();
            // Entries of the rest of the batch must beat the K-th key held
            // now: the first batch, whose threshold starts at +inf, then
            // keeps only what can still enter the list.
            threshold();
          }
        }
      }
      if (__any(nbuf > 0)) {
        if (nbuf > 0) merge();
      }
    }
  }
  if (!live) return;
  if (open) {
    if (lane == 0) spill[atomicAdd(spill_count, 1)] = (int32_t)t;
    return;
  }
  // The selection: ranks below K with a finite key, in order.
  const bool have = lane < K && ak != KEY_INF;
  const int n = __popcll(group_ballot<G>(have, h));
  int* out = a.idx + t * a.stride;
  if (have) out[lane] = (int)ai;
  bi[lane] = ai;
  wave_sync();
  if (lane == 0) {
    int status = n < a.min_samples ? BK_SEARCH_FAILED : BK_OK;
    if (status == BK_OK && a.min_holes >= 0) {
      int distinct = 0;
      for (int k = 0; k < n; ++k) {
        const int hk = a.hole ? a.hole[bi[k]] : -1;
        if (hk < 0) {
          distinct++;
          continue;
        }
        bool seen = false;
        for (int j = 0; j < k && !seen; ++j) seen = a.hole[bi[j]] == hk;
        distinct += !seen;
      }
      if (distinct < a.min_holes) status = BK_SEARCH_FAILED;
    }
    a.count[t] = n;
    a.status[t] = status;
  }
}

// ---------------------------------------------------------------------------
// Octant (or planar quadrant) searches without a hole cap, cap <= S per
// sector. Selector::offer (search.rs:519-543) takes a candidate, in (d2,
// index) order, iff its sector holds fewer than cap and the selection is not
// full; so the selection is the first max_samples entries, in that order, of
// the union of every sector's cap smallest keys within the radius (before
// the selection is full each sector takes exactly its first cap candidates,
// and what the walk offers after that lies beyond all of them). One target
// per wavefront: lanes S o .. S o + S - 1 hold sector o's smallest keys,
// sorted; candidates below their sector's cap-th key gather in per-sector
// LDS buffers, merged into the lists by bitonic networks over S lanes, all
// sectors at once. The walk follows the shell table as k_knn_wave does and
// stops once every sector holds cap keys and the shell's lower bound passes
// the largest cap-th key (or the radius); a target still open after the
// table goes to k_search.

// The sector of projected offset u (entry E, target q) where the signs
// decide it; -1 where a component lies within rounding of 0.
__device__ __forceinline__ int sector_fast(const SearchArgs& a, double u0, double u1, double u2, double scale) {
  const bool amb = fabs(u0) <= a.sec_tol[0] * scale || fabs(u1) <= a.sec_tol[1] * scale ||
                   fabs(u2) <= a.sec_tol[2] * scale;
  const int pat = (u0 >= 0.0 ? 1 : 0) | (u1 >= 0.0 ? 2 : 0) | (u2 >= 0.0 ? 4 : 0);
  return amb ? -1 : (int)((a.sec_code >> (3 * pat)) & 7u);
}

// Cells within this many cell sizes of a coordinate plane through the target
// count as straddling it (far above the rounding of the cell assignment, of
// the FP32 cell coordinates and of the sign tests).
constexpr float kSideMargin = 1e-3f;

// Four wavefronts per SIMD (at most 128 registers): the walk waits on
// shuffles, LDS and gathers in turn.
template <int S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_sectors(const SearchArgs a, int32_t* spill, int32_t* spill_count) {
  constexpr int NG = BK_WAVE / S;    // sector groups
  constexpr int CAPB = BK_WAVE + S;  // buffer entries per sector: < S held + one round
  __shared__ uint64_t buf_k[4][NG][CAPB];
  __shared__ uint32_t buf_i[4][NG][CAPB];
  __shared__ uint32_t buf_n[4][NG];
  __shared__ uint64_t thr_k[4][NG];
  __shared__ uint32_t thr_i[4][NG];
  __shared__ int16_t held[4][8];
  __shared__ int chosen[4][BK_WAVE];
  const int wave = threadIdx.x / BK_WAVE, lane = threadIdx.x % BK_WAVE;
  const int64_t b = (int64_t)blockIdx.x * 4 + wave;
  if (b >= a.m) return;  // whole wave
  const int64_t t = a.list ? a.list[b] : b;
  const int64_t left_out = a.left_out ? a.left_out[t] : -1;
  const uint32_t limit = a.limit ? (uint32_t)a.limit[t] : 0xffffffffu;
  const int o = lane / S, j = lane % S;
  const int cap = a.sector_cap;
  const int K = a.max_samples;
  const GridDev& g = a.g;
  const double t0 = a.tx[t], t1 = a.ty[t], t2 = a.tz[t];
  double q0, q1, q2;
  mv3(a.F, t0 - a.origin[0], t1 - a.origin[1], t2 - a.origin[2], q0, q1, q2);
  const double qabs = fabs(q0) + fabs(q1) + fabs(q2);
  uint64_t lk = KEY_INF;  // this lane's list entry
  uint32_t li = 0xffffffffu;
  if (lane < NG) {
    buf_n[wave][lane] = 0;
    thr_k[wave][lane] = KEY_INF;
    thr_i[wave][lane] = 0xffffffffu;
  }
  wave_sync();
  // Every buffer into the lists, S entries of each sector per step from the
  // top of its buffer; then the thresholds (the cap-th key of each sector).
  auto flush = [&]() {
    for (;;) {
      const uint32_t nb = buf_n[wave][o];
      if (!__any(nb > 0)) break;
      const uint32_t take = nb < (uint32_t)S ? nb : (uint32_t)S;
      uint64_t k = KEY_INF;
      uint32_t i = 0xffffffffu;
      if ((uint32_t)j < take) {
        k = buf_k[wave][o][nb - take + j];
        i = buf_i[wave][o][nb - take + j];
      }
      wave_sync();
      if (j == 0) buf_n[wave][o] = nb - take;
      group_sort_keys<S>(k, i, j);
      group_merge_keys<S>(lk, li, k, i, j);
      wave_sync();
    }
    if (j == cap - 1) {
      thr_k[wave][o] = lk;
      thr_i[wave][o] = li;
    }
    wave_sync();
  };
  bool open = g.start != nullptr && K > 0;
  if (open) {
    const int c0 = clamp_cell((q0 - g.lo[0]) / g.h, g.dims[0]);
    const int c1 = clamp_cell((q1 - g.lo[1]) / g.h, g.dims[1]);
    const int c2 = clamp_cell((q2 - g.lo[2]) / g.h, g.dims[2]);
    const double h2 = g.h * g.h;
    const double beyond = h2 * (double)((BK_REACH - 1) * (BK_REACH - 1));
    const double slack = 1.0 - 1e-9;
    const uint32_t n_off = a.tab.end[a.tab.n - 1];
    // The target's place in cells (outside [0, dims) beyond the grid), for
    // the cells' boxes relative to it.
    const double qc0 = (q0 - g.lo[0]) / g.h, qc1 = (q1 - g.lo[1]) / g.h, qc2 = (q2 - g.lo[2]) / g.h;
    const double ih2 = 1.0 / h2;
    for (uint32_t o0 = 0; o0 < n_off; o0 += BK_WAVE) {
      // The largest cap-th key over the sectors, +inf while one holds fewer.
      uint64_t tmax = lane < a.n_sectors ? thr_k[wave][lane] : 0ull;
#pragma unroll
      for (int x = 1; x < 8; x <<= 1) {
        const uint64_t other = shfl_xor_u64(tmax, x);
        tmax = other > tmax ? other : tmax;
      }
      tmax = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(tmax >> 32), 0) << 32) |
             (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)tmax, 0);
      const char4 f0 = a.tab.offsets[o0];
      const int m0 = max(abs(f0.x) - 1, 0) * max(abs(f0.x) - 1, 0) + max(abs(f0.y) - 1, 0) * max(abs(f0.y) - 1, 0) +
                     max(abs(f0.z) - 1, 0) * max(abs(f0.z) - 1, 0);
      const double bound = fmin(h2 * (double)m0, beyond) * slack;
      if (bound > a.radius2 || (tmax != KEY_INF && bound > __longlong_as_double((long long)tmax))) {
        open = false;
        break;
      }
      // Lane l: cell o0 + l, whose entries are [e, e + cnt) of the CSR,
      // unless every sector its entries can fall in already holds cap keys
      // nearer than the cell's box: where the sectors are sign patterns of
      // the projected offset (sec_fast), a cell at least kSideMargin cells
      // from a coordinate plane through the target lies on one side of it,
      // and a cell straddling a plane can hold both signs.
      uint32_t e = 0, cnt = 0;
      const uint32_t oc = o0 + lane;
      if (oc < n_off) {
        const char4 f = a.tab.offsets[oc];
        const int x = c0 + f.x, y = c1 + f.y, z = c2 + f.z;
        if (x >= 0 && x < g.dims[0] && y >= 0 && y < g.dims[1] && z >= 0 && z < g.dims[2]) {
          // Cells on one side of every plane (most cells once the near
          // ones are done) lie in one sector.
          bool skip = false;
          if (a.sec_fast) {
            const float l0 = (float)((double)x - qc0), l1 = (float)((double)y - qc1), l2 = (float)((double)z - qc2);
            const bool p0 = l0 > kSideMargin, p1 = l1 > kSideMargin, p2 = l2 > kSideMargin;
            const bool n0 = l0 + 1.0f < -kSideMargin, n1 = l1 + 1.0f < -kSideMargin, n2 = l2 + 1.0f < -kSideMargin;
            if ((p0 || n0) && (p1 || n1) && (p2 || n2)) {
              // The gap to the box along each axis, less the margin: a
              // lower bound of every entry's squared distance, in cells.
              const float g0 = fmaxf((p0 ? l0 : -(l0 + 1.0f)) - kSideMargin, 0.0f);
              const float g1 = fmaxf((p1 ? l1 : -(l1 + 1.0f)) - kSideMargin, 0.0f);
              const float g2 = fmaxf((p2 ? l2 : -(l2 + 1.0f)) - kSideMargin, 0.0f);
              const int pat = (p0 ? 1 : 0) | (p1 ? 2 : 0) | (p2 ? 4 : 0);
              const int sec = (int)((a.sec_code >> (3 * pat)) & 7u);
              const uint64_t th = thr_k[wave][sec];
              // The sector's cap-th distance in squared cells, rounded up.
              skip = th != KEY_INF &&
                     ((g0 * g0 + g1 * g1) + g2 * g2) * (1.0f - 1e-5f) >
                         (float)(__longlong_as_double((long long)th) * ih2) * (1.0f + 1e-5f);
            }
          }
          if (!skip) {
            const int id = x + g.dims[0] * (y + g.dims[1] * z);
            e = g.start[id];
            cnt = g.start[id + 1] - e;
          }
        }
      }
      const uint32_t incl = group_incl_scan<BK_WAVE>(cnt, lane);
      const uint32_t excl = incl - cnt;
      const uint32_t total = __shfl(incl, BK_WAVE - 1, BK_WAVE);
      for (uint32_t base = 0; base < total; base += BK_WAVE) {
        const uint32_t f = base + lane;
        int pos = 0;
#pragma unroll
        for (int step = BK_WAVE / 2; step > 0; step >>= 1) {
          const uint32_t probe = __shfl(incl, pos + step - 1, BK_WAVE);
          pos += probe <= f ? step : 0;
        }
        const uint32_t entry = __shfl(e, pos, BK_WAVE) + (f - __shfl(excl, pos, BK_WAVE));
        if (f < total) {
          const double4 E = g.entries[entry];
          const double u0 = E.x - q0, u1 = E.y - q1, u2 = E.z - q2;
          const double d2 = (u0 * u0 + u1 * u1) + u2 * u2;
          const uint32_t ii = (uint32_t)__double_as_longlong(E.w);
          const uint64_t kk = dbits(d2);
          bool keep = d2 <= a.radius2 && (int64_t)ii != left_out && ii < limit && (tmax == KEY_INF || kk <= tmax);
          if (keep) {
            int sec = a.sec_fast ? sector_fast(a, u0, u1, u2, fabs(E.x) + fabs(E.y) + fabs(E.z) + qabs) : -1;
            if (sec < 0) sec = sector_at(a, t0, t1, t2, ii);
            keep = key_less(kk, ii, thr_k[wave][sec], thr_i[wave][sec]);
            if (keep) {
              const uint32_t slot = atomicAdd(&buf_n[wave][sec], 1u);
              buf_k[wave][sec][slot] = kk;
              buf_i[wave][sec][slot] = ii;
            }
          }
        }
        wave_sync();
        const bool full = lane < NG && buf_n[wave][lane] >= (uint32_t)S;
        if (__any(full)) flush();
      }
    }
    flush();
  }
  if (open) {
    if (lane == 0) spill[atomicAdd(spill_count, 1)] = (int32_t)t;
    return;
  }
  // The selection: the sectors' first cap keys, sorted, the first K of them.
  uint64_t k = j < cap ? lk : KEY_INF;
  uint32_t i = j < cap ? li : 0xffffffffu;
  uint32_t sec = (uint32_t)o;
  // Ascending sort of (k, i) across the wave with the sector alongside.
#pragma unroll
  for (int size = 2; size <= BK_WAVE; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const uint64_t ok = shfl_xor_u64(k, stride);
      const uint32_t oi = __shfl_xor(i, stride, BK_WAVE);
      const uint32_t os = __shfl_xor(sec, stride, BK_WAVE);
      const bool want_small = ((lane & stride) == 0) == ((lane & size) == 0);
      const bool take = key_less(ok, oi, k, i) == want_small;
      k = take ? ok : k;
      i = take ? oi : i;
      sec = take ? os : sec;
    }
  }
  const bool have = lane < K && k != KEY_INF;
  const int n = __popcll(__ballot(have));
  if (lane < 8) held[wave][lane] = 0;
  wave_sync();
  for (int s = 0; s < a.n_sectors; ++s) {
    const int cs = __popcll(__ballot(have && sec == (uint32_t)s));
    if (lane == 0) held[wave][s] = (int16_t)cs;
  }
  int* out = a.idx + t * a.stride;
  if (have) {
    out[lane] = (int)i;
    chosen[wave][lane] = (int)i;
  }
  wave_sync();
  if (lane == 0) {
    a.count[t] = n;
    a.status[t] = selection_status(a, n, held[wave], chosen[wave]);
  }
}

// ---------------------------------------------------------------------------
// The same selections on tiles of 64 targets that lie close in the grid
// (targets sorted by the Morton code of their cell). The wave copies the
// entries of every cell within `reach` cells of the tile's cell box into LDS
// in chunks (cells along x are contiguous in the CSR, so each (y, z) row of
// the box is one coalesced range), and each lane offers the staged
// candidates to its own heap, as k_knn does cell by cell. Exactness: a sample
// outside the box lies beyond the nearest face of the box that is not a face
// of the grid, at distance at least `safe` from the target; a lane whose K-th
// key lies below safe^2 (or, with fewer than K found, whose radius does) has
// its exact selection. Any other lane goes to the spill list, which k_knn
// searches afterwards.

__device__ __forceinline__ uint64_t morton_spread(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

__global__ void k_cell_keys(int64_t m, const SearchArgs a, uint64_t* keys, int32_t* ids) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t t = a.list ? a.list[i] : i;
  const GridDev& g = a.g;
  double q0, q1, q2;
  mv3(a.F, a.tx[t] - a.origin[0], a.ty[t] - a.origin[1], a.tz[t] - a.origin[2], q0, q1, q2);
  const uint32_t c0 = (uint32_t)clamp_cell((q0 - g.lo[0]) / g.h, g.dims[0]);
  const uint32_t c1 = (uint32_t)clamp_cell((q1 - g.lo[1]) / g.h, g.dims[1]);
  const uint32_t c2 = (uint32_t)clamp_cell((q2 - g.lo[2]) / g.h, g.dims[2]);
  keys[i] = morton_spread(c0) | (morton_spread(c1) << 1) | (morton_spread(c2) << 2);
  ids[i] = (int32_t)t;
}

constexpr int TILE_CHUNK = 256;
constexpr int TILE_SPAN = 3;

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, BK_WAVE));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, BK_WAVE));
  return v;
}

__global__ void __launch_bounds__(BK_WAVE) k_knn_tile(const SearchArgs a, int reach, int32_t* spill,
                                                      int32_t* spill_count) {
  extern __shared__ uint64_t knn_dyn[];
  const int lane = threadIdx.x;
  const int K = a.max_samples;
  KnnHeap H{knn_dyn, (uint32_t*)(knn_dyn + (size_t)K * BK_WAVE), lane};
  double4* stage = (double4*)(knn_dyn + (size_t)K * BK_WAVE + ((size_t)K * BK_WAVE + 1) / 2);
  const int64_t b = (int64_t)blockIdx.x * BK_WAVE + lane;
  const bool live = b < a.m;
  const int64_t t = live ? a.list[b] : 0;
  const int64_t left_out = (live && a.left_out) ? a.left_out[t] : -1;
  const uint32_t limit = (live && a.limit) ? (uint32_t)a.limit[t] : 0xffffffffu;
  const GridDev& g = a.g;
  double q0 = 0, q1 = 0, q2 = 0;
  if (live) mv3(a.F, a.tx[t] - a.origin[0], a.ty[t] - a.origin[1], a.tz[t] - a.origin[2], q0, q1, q2);
  int c[3] = {0, 0, 0};
  if (live) {
    c[0] = clamp_cell((q0 - g.lo[0]) / g.h, g.dims[0]);
    c[1] = clamp_cell((q1 - g.lo[1]) / g.h, g.dims[1]);
    c[2] = clamp_cell((q2 - g.lo[2]) / g.h, g.dims[2]);
  }
  // A Morton run can jump across the grid: lanes farther than TILE_SPAN cells
  // from the cell of the tile's first target take the walk instead, so the
  // box stays small. The staged box: the remaining cells widened by `reach`,
  // inside the grid.
  const int first = __builtin_amdgcn_readfirstlane(lane);
  int ref[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) ref[d] = __shfl(c[d], first, BK_WAVE);
  const bool near = live && abs(c[0] - ref[0]) <= TILE_SPAN && abs(c[1] - ref[1]) <= TILE_SPAN &&
                    abs(c[2] - ref[2]) <= TILE_SPAN;
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = max(wave_min_i(near ? c[d] : INT_MAX) - reach, 0);
    hi[d] = min(wave_max_i(near ? c[d] : INT_MIN) + reach, g.dims[d] - 1);
  }
  // Distance from the target to the nearest face of the box inside the grid,
  // less a margin for the rounding of the cell assignment.
  const double qq[3] = {q0, q1, q2};
  double safe = INFINITY;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (lo[d] > 0) safe = fmin(safe, qq[d] - (g.lo[d] + (double)lo[d] * g.h));
    if (hi[d] < g.dims[d] - 1) safe = fmin(safe, (g.lo[d] + (double)(hi[d] + 1) * g.h) - qq[d]);
  }
  safe -= 1e-6 * g.h;
  const double safe2 = safe > 0.0 ? safe * safe : 0.0;
  int size = 0;
  uint64_t top = 0;
  uint32_t top_i = 0;
  const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
  const int rows = ny * nz;
  int row = 0;
  uint32_t e = 0, e_end = 0;
  auto row_range = [&](int rr) {
    const int y = lo[1] + rr % ny, z = lo[2] + rr / ny;
    const int id = lo[0] + g.dims[0] * (y + g.dims[1] * z);
    e = g.start[id];
    e_end = g.start[id + nx];
  };
  if (rows > 0) row_range(0);
  while (row < rows) {
    // Fill one chunk.
    int fill = 0;
    while (fill < TILE_CHUNK && row < rows) {
      const int take = (int)min((uint32_t)(TILE_CHUNK - fill), e_end - e);
      for (int k = lane; k < take; k += BK_WAVE) stage[fill + k] = g.entries[e + k];
      fill += take;
      e += take;
      if (e == e_end && ++row < rows) row_range(row);
    }
    __syncthreads();
    if (near) {
      for (int k = 0; k < fill; ++k) {
        const double4 E = stage[k];
        const double u0 = E.x - q0, u1 = E.y - q1, u2 = E.z - q2;
        const double d2 = (u0 * u0 + u1 * u1) + u2 * u2;
        const uint32_t ii = (uint32_t)__double_as_longlong(E.w);
        if (!(d2 <= a.radius2) || (int64_t)ii == left_out || ii >= limit) continue;
        const uint64_t kk = dbits(d2);
        if (size < K) {
          H.push(size, kk, ii);
          if (size == K) {
            top = H.key(0);
            top_i = H.idx(0);
          }
        } else if (key_less(kk, ii, top, top_i)) {
          H.replace_top(size, kk, ii);
          top = H.key(0);
          top_i = H.idx(0);
        }
      }
    }
    __syncthreads();
  }
  if (!live) return;
  const bool exact = near && (size == K ? __longlong_as_double((long long)top) < safe2 : a.radius2 < safe2);
  if (!exact) {
    spill[atomicAdd(spill_count, 1)] = (int32_t)t;
    return;
  }
  const int n = size;
  for (int end = n - 1; end > 0; --end) {
    uint64_t mk = H.key(0);
    uint32_t mi = H.idx(0);
    uint64_t lk = H.key(end);
    uint32_t li = H.idx(end);
    H.put(end, mk, mi);
    H.replace_top(end, lk, li);
  }
  int* out = a.idx + t * a.stride;
  for (int k = 0; k < n; ++k) out[k] = (int)H.idx(k);
  int status = n < a.min_samples ? BK_SEARCH_FAILED : BK_OK;
  if (status == BK_OK && a.min_holes >= 0) {
    int distinct = 0;
    for (int k = 0; k < n; ++k) {
      int h = a.hole ? a.hole[out[k]] : -1;
      if (h < 0) {
        distinct++;
        continue;
      }
      bool seen = false;
      for (int j = 0; j < k && !seen; ++j) seen = a.hole[out[j]] == h;
      distinct += !seen;
    }
    if (distinct < a.min_holes) status = BK_SEARCH_FAILED;
  }
  a.count[t] = n;
  a.status[t] = status;
}

// ---------------------------------------------------------------------------
// Grid build.
// SearchTree::balanced (search.rs:909-936) for calibrated searches that balance
// octants or sectors: each target's selection taken round-robin over its
// sectors, i.e. stably sorted by the rank of each sample within its sector,
// so every leading part stays balanced. One wavefront per target, one lane
// per selected sample (at most 64).
__global__ void __launch_bounds__(256) k_balance(const SearchArgs a) {
  const int wave = threadIdx.x / BK_WAVE, lane = threadIdx.x % BK_WAVE;
  const int64_t t = (int64_t)blockIdx.x * 4 + wave;
  if (t >= a.m || a.status[t] != BK_OK) return;  // whole wave
  const int n = a.count[t];
  int32_t* idx = a.idx + t * a.stride;
  const bool live = lane < n;
  const int i = live ? idx[lane] : 0;
  const int sec = live ? sector_at(a, a.tx[t], a.ty[t], a.tz[t], (uint32_t)i) : -1;
  int rank = 1;  // 1 + earlier samples of the same sector
  for (int j = 0; j < BK_WAVE; ++j) {
    const int sj = __shfl(sec, j, BK_WAVE);
    rank += (j < lane && sj == sec) ? 1 : 0;
  }
  int pos = 0;  // place in the stable order by rank
  for (int j = 0; j < BK_WAVE; ++j) {
    const int rj = __shfl(rank, j, BK_WAVE);
    pos += (j < n && (rj < rank || (rj == rank && j < lane))) ? 1 : 0;
  }
  if (live) idx[pos] = i;
}

// Zero fill as a kernel rather than hipMemsetAsync: an enqueued pass may be
// captured into a graph and replayed, and its nodes are kept to kernels and
// copies.
__global__ void k_zero32(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_project(int64_t n, const double* x, const double* y, const double* z, const double* F,
                          double o0, double o1, double o2, double* px, double* py, double* pz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r0, r1, r2;
  mv3(F, x[i] - o0, y[i] - o1, z[i] - o2, r0, r1, r2);
  px[i] = r0;
  py[i] = r1;
  pz[i] = r2;
}

// Partial min / max of the finite projected points, one pair per block.
__global__ void __launch_bounds__(256) k_bounds(int64_t n, const double* px, const double* py, const double* pz,
                                                double* part) {
  __shared__ double red[6][256];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double p[3] = {px[i], py[i], pz[i]};
    if (isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]))
      for (int d = 0; d < 3; ++d) {
        lo[d] = fmin(lo[d], p[d]);
        hi[d] = fmax(hi[d], p[d]);
      }
  }
  for (int d = 0; d < 3; ++d) {
    red[d][threadIdx.x] = lo[d];
    red[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + s]);
        red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_cell_count(int64_t n, const double* px, const double* py, const double* pz, double lo0,
                             double lo1, double lo2, double h, int d0, int d1, int d2, uint32_t* cell_of,
                             uint32_t* count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c0 = clamp_cell((px[i] - lo0) / h, d0);
  int c1 = clamp_cell((py[i] - lo1) / h, d1);
  int c2 = clamp_cell((pz[i] - lo2) / h, d2);
  uint32_t id = (uint32_t)(c0 + d0 * (c1 + d1 * c2));
  cell_of[i] = id;
  atomicAdd(&count[id], 1u);
}

__global__ void k_scatter(int64_t n, const double* px, const double* py, const double* pz, const uint32_t* cell_of,
                          const uint32_t* start, uint32_t* fill, double4* entries) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = cell_of[i];
  uint32_t pos = start[c] + atomicAdd(&fill[c], 1u);
  entries[pos] = make_double4(px[i], py[i], pz[i], __longlong_as_double((long long)i));
}

}  // namespace

namespace bk {

#define HIPCHK(x)                      \
  do {                                 \
    if ((x) != hipSuccess) return -2;  \
  } while (0)

// grid.rs:83-106: side for about one point per cell over the spanned axes,
// doubled until there are at most 4 n + 64 cells.
static double reference_side(const double lo[3], const double hi[3], int64_t n) {
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
  double h = std::pow(prod / (double)std::max<int64_t>(n, 1), 1.0 / (double)cnt);
  auto cells = [&](double hh) {
    double c = 1.0;
    for (double e : ext) c *= std::floor(e / hh) + 1.0;
    return c;
  };
  while (cells(h) > 4.0 * (double)n + 64.0) h *= 2.0;
  return h;
}

void free_grid(GridBuffers& gb) {
  (void)hipFree(gb.entries);
  (void)hipFree(gb.start);
  (void)hipFree(gb.cell_of);
  (void)hipFree(gb.fill);
  (void)hipFree(gb.px);
  (void)hipFree(gb.bounds);
  (void)hipFree(gb.scan_tmp);
  gb = GridBuffers();
}

int build_grid(hipStream_t s, int64_t n, const double* x, const double* y, const double* z, const double F[9],
               const double origin[3], double cell_factor, GridBuffers& gb, GridDev& out) {
  const int T = 256;
  int64_t blocks = (n + T - 1) / T;
  int64_t red_blocks = std::min<int64_t>(std::max<int64_t>(blocks, 1), 1024);
  if (gb.fixed && (n > gb.cap_n || !gb.bounds)) return -3;
  if (n > gb.cap_n) {
    (void)hipFree(gb.entries);
    (void)hipFree(gb.cell_of);
    (void)hipFree(gb.px);
    HIPCHK(hipMalloc(&gb.entries, n * sizeof(double4)));
    HIPCHK(hipMalloc(&gb.cell_of, n * sizeof(uint32_t)));
    HIPCHK(hipMalloc(&gb.px, 3 * n * sizeof(double)));
    gb.cap_n = n;
  }
  // Partial bounds of up to 1024 blocks, then the frame.
  if (!gb.bounds) HIPCHK(hipMalloc(&gb.bounds, (1024 * 6 + 16) * sizeof(double)));
  double* dF = gb.bounds + 1024 * 6;
  double* px = gb.px;
  double* py = gb.px + n;
  double* pz = gb.px + 2 * n;
  HIPCHK(hipMemcpyAsync(dF, F, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  k_project<<<dim3((unsigned)blocks), dim3(T), 0, s>>>(n, x, y, z, dF, origin[0], origin[1], origin[2], px, py, pz);
  k_bounds<<<dim3((unsigned)red_blocks), dim3(256), 0, s>>>(n, px, py, pz, gb.bounds);
  double part[1024 * 6];
  HIPCHK(hipMemcpyAsync(part, gb.bounds, red_blocks * 6 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t b = 0; b < red_blocks; ++b)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::fmin(lo[d], part[b * 6 + d]);
      hi[d] = std::fmax(hi[d], part[b * 6 + 3 + d]);
    }
  if (!(lo[0] <= hi[0])) {
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = 0.0;
  }
  double h = reference_side(lo, hi, n) * cell_factor;
  int dims[3];
  int64_t cells = 1;
  for (;;) {
    cells = 1;
    for (int d = 0; d < 3; ++d) {
      double k = std::floor((hi[d] - lo[d]) / h);
      dims[d] = (int)std::min(k, 1e9) + 1;
      cells *= dims[d];
    }
    if (!gb.fixed || cells <= gb.cap_cells) break;
    h *= 2.0;
  }
  if (cells > gb.cap_cells) {
    (void)hipFree(gb.start);
    (void)hipFree(gb.fill);
    HIPCHK(hipMalloc(&gb.start, (cells + 1) * sizeof(uint32_t)));
    HIPCHK(hipMalloc(&gb.fill, (cells + 1) * sizeof(uint32_t)));
    gb.cap_cells = cells;
  }
  HIPCHK(hipMemsetAsync(gb.fill, 0, (cells + 1) * sizeof(uint32_t), s));
  k_cell_count<<<dim3((unsigned)blocks), dim3(T), 0, s>>>(n, px, py, pz, lo[0], lo[1], lo[2], h, dims[0], dims[1],
                                                         dims[2], gb.cell_of, gb.fill);
  size_t need = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, gb.fill, gb.start, (int)(cells + 1), s));
  if (need > gb.scan_tmp_bytes) {
    if (gb.fixed) return -3;
    (void)hipFree(gb.scan_tmp);
    HIPCHK(hipMalloc(&gb.scan_tmp, need));
    gb.scan_tmp_bytes = need;
  }
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(gb.scan_tmp, need, gb.fill, gb.start, (int)(cells + 1), s));
  HIPCHK(hipMemsetAsync(gb.fill, 0, (cells + 1) * sizeof(uint32_t), s));
  k_scatter<<<dim3((unsigned)blocks), dim3(T), 0, s>>>(n, px, py, pz, gb.cell_of, gb.start, gb.fill, gb.entries);
  HIPCHK(hipGetLastError());
  for (int d = 0; d < 3; ++d) {
    out.lo[d] = lo[d];
    out.dims[d] = dims[d];
  }
  out.h = h;
  out.start = gb.start;
  out.entries = gb.entries;
  return 0;
}

size_t grid_scan_bytes(int64_t cells) {
  size_t need = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, need, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)(cells + 1),
                                       (hipStream_t)0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return need;
}

// grid.rs:39-64: offsets within BK_REACH sorted by m = sum max(0, |o| - 1)^2.
int upload_shell_table(ShellTable& tab, void** storage) {
  struct K {
    uint32_t m;
    int x, y, z;
  };
  std::vector<K> keyed;
  for (int x = -BK_REACH; x <= BK_REACH; ++x)
    for (int y = -BK_REACH; y <= BK_REACH; ++y)
      for (int z = -BK_REACH; z <= BK_REACH; ++z) {
        auto f = [](int v) {
          int a = std::max(std::abs(v) - 1, 0);
          return (uint32_t)(a * a);
        };
        keyed.push_back({f(x) + f(y) + f(z), x, y, z});
      }
  std::sort(keyed.begin(), keyed.end(), [](const K& a, const K& b) {
    if (a.m != b.m) return a.m < b.m;
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    return a.z < b.z;
  });
  std::vector<char4> off;
  std::vector<uint32_t> ends, ms;
  for (size_t k = 0; k < keyed.size(); ++k) {
    off.push_back(make_char4((signed char)keyed[k].x, (signed char)keyed[k].y, (signed char)keyed[k].z, 0));
    if (!ms.empty() && ms.back() == keyed[k].m)
      ends.back() = (uint32_t)(k + 1);
    else {
      ms.push_back(keyed[k].m);
      ends.push_back((uint32_t)(k + 1));
    }
  }
  size_t bytes = off.size() * sizeof(char4) + 2 * ends.size() * sizeof(uint32_t);
  char* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, bytes));
  HIPCHK(hipMemcpy(d, off.data(), off.size() * sizeof(char4), hipMemcpyHostToDevice));
  uint32_t* de = (uint32_t*)(d + off.size() * sizeof(char4));
  uint32_t* dm = de + ends.size();
  HIPCHK(hipMemcpy(de, ends.data(), ends.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dm, ms.data(), ms.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  tab.offsets = (const char4*)d;
  tab.end = de;
  tab.m = dm;
  tab.n = (int)ms.size();
  *storage = d;
  return 0;
}

bool knn_path(const SearchArgs& a) { return !a.balances && a.max_per_hole < 0 && a.max_samples <= BK_MAX_KNN; }

template <class T>
static int grow(T*& p, int64_t& cap, int64_t n) {
  if (n <= cap) return 0;
  (void)hipFree(p);
  p = nullptr;
  cap = 0;
  HIPCHK(hipMalloc((void**)&p, (size_t)n * sizeof(T)));
  cap = n;
  return 0;
}

// The targets a wave or tile search handed back (spill[0 .. *count), at most
// m): k_knn over them, persistent, reading the count on the device so that
// no host round trip sits between the two kernels.
static int launch_spill(hipStream_t s, const SearchArgs& a, int64_t m, const int32_t* spill, const int32_t* count) {
  SearchArgs d = a;
  d.list = spill;
  d.m = m;
  d.m_dev = count;
  const size_t heap = (size_t)a.max_samples * BK_WAVE * (sizeof(uint64_t) + sizeof(uint32_t));
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_knn, BK_WAVE, heap) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess) {
    (void)hipGetLastError();
    cus = 256;
  }
  const int64_t grid = std::min<int64_t>((m + BK_WAVE - 1) / BK_WAVE, (int64_t)per_cu * cus);
  k_knn<<<dim3((unsigned)std::max<int64_t>(grid, 1)), dim3(BK_WAVE), heap, s>>>(d);
  HIPCHK(hipGetLastError());
  return 0;
}

// Uncapped selections on Morton-sorted tiles with LDS-staged candidates, the
// targets the tile cannot settle by k_knn.
static int launch_tiles(hipStream_t s, const SearchArgs& a, int64_t m, int reach, TileBuffers& tb) {
  if (m > 2147483647LL) return -1;
  if (grow(tb.keys, tb.cap_keys, 2 * m)) return -1;
  if (grow(tb.ids, tb.cap_ids, 3 * m + 1)) return -1;
  uint64_t* kin = tb.keys;
  uint64_t* kout = tb.keys + m;
  int32_t* iin = tb.ids;
  int32_t* iout = tb.ids + m;
  int32_t* spill = tb.ids + 2 * m;
  int32_t* count = tb.ids + 3 * m;
  k_cell_keys<<<dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s>>>(m, a, kin, iin);
  size_t need = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, iin, iout, (int)m, 0, 64, s));
  if (need > tb.tmp_bytes) {
    (void)hipFree(tb.tmp);
    tb.tmp = nullptr;
    tb.tmp_bytes = 0;
    HIPCHK(hipMalloc(&tb.tmp, need));
    tb.tmp_bytes = need;
  }
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(tb.tmp, need, kin, kout, iin, iout, (int)m, 0, 64, s));
  k_zero32<<<1, 64, 0, s>>>(count, 1);
  SearchArgs c = a;
  c.list = iout;
  c.m = m;
  const size_t heap = (size_t)a.max_samples * BK_WAVE * (sizeof(uint64_t) + sizeof(uint32_t));
  const size_t sh = heap + (size_t)TILE_CHUNK * sizeof(double4);
  k_knn_tile<<<dim3((unsigned)((m + BK_WAVE - 1) / BK_WAVE)), dim3(BK_WAVE), sh, s>>>(c, reach, spill, count);
  HIPCHK(hipGetLastError());
  tb.spill_count = count;
  return launch_spill(s, a, m, spill, count);
}

int launch_balance(hipStream_t s, const SearchArgs& a, int64_t m) {
  if (m <= 0) return 0;
  SearchArgs c = a;
  c.m = m;
  k_balance<<<dim3((unsigned)((m + 3) / 4)), dim3(256), 0, s>>>(c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// The first datum (lowest index) at each node's exact location: the node
// projects to the same cell as that datum (the same projection), whose
// entries are compared in world coordinates.
__global__ void k_sgs_datum(const SearchArgs a, int64_t m, const double* gx, const double* gy, const double* gz,
                            int64_t* datum) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double x = gx[i], y = gy[i], z = gz[i];
  int64_t best = -1;
  const GridDev& g = a.g;
  if (g.start) {
    double q0, q1, q2;
    mv3(a.F, x - a.origin[0], y - a.origin[1], z - a.origin[2], q0, q1, q2);
    const int c0 = clamp_cell((q0 - g.lo[0]) / g.h, g.dims[0]);
    const int c1 = clamp_cell((q1 - g.lo[1]) / g.h, g.dims[1]);
    const int c2 = clamp_cell((q2 - g.lo[2]) / g.h, g.dims[2]);
    const int id = c0 + g.dims[0] * (c1 + g.dims[1] * c2);
    for (uint32_t e = g.start[id]; e < g.start[id + 1]; ++e) {
      const uint32_t k = (uint32_t)__double_as_longlong(g.entries[e].w);
      if (a.sx[k] == x && a.sy[k] == y && a.sz[k] == z && (best < 0 || (int64_t)k < best)) best = k;
    }
  }
  datum[i] = best;
}

// flag[k]: path[k] is a node not on a datum; seen[i]: times node i is on
// the path (bad[0] set when the path is not a permutation).
__global__ void k_sgs_flags(int64_t m, const int64_t* path, const int64_t* datum, int32_t* flag, int32_t* seen,
                            int32_t* bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int64_t node = path[k];
  if (node < 0 || node >= m) {
    bad[0] = 1;
    flag[k] = 0;
    return;
  }
  flag[k] = datum[node] < 0 ? 1 : 0;
  atomicAdd(&seen[node], 1);
}

__global__ void k_sgs_place(int64_t m, const int64_t* path, const int32_t* flag, const int32_t* scan,
                            const int32_t* seen, int32_t* pos, int64_t* app, int32_t* bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  if (seen[k] != 1) bad[0] = 1;
  if (k == m - 1) bad[1] = scan[k] + flag[k];
  const int64_t node = path[k];
  if (flag[k] && node >= 0 && node < m) {
    pos[node] = scan[k];
    app[scan[k]] = node;
  }
}

int sgs_order(hipStream_t s, const SearchArgs& a, int64_t m, const double* gx, const double* gy, const double* gz,
              const int64_t* path, int64_t* datum, int32_t* pos, int64_t* app, int32_t* work, void** tmp,
              size_t* tmp_bytes, int32_t out[2]) {
  if (m <= 0 || m > 2147483647LL) return -1;
  int32_t* flag = work;
  int32_t* scan = work + m;
  int32_t* seen = work + 2 * m;
  int32_t* bad = work + 3 * m;
  const unsigned blocks = (unsigned)((m + 255) / 256);
  if (zero32(s, seen, m) || zero32(s, bad, 2)) return -1;
  k_sgs_datum<<<dim3(blocks), dim3(256), 0, s>>>(a, m, gx, gy, gz, datum);
  k_sgs_flags<<<dim3(blocks), dim3(256), 0, s>>>(m, path, datum, flag, seen, bad);
  size_t need = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, flag, scan, (int)m, s));
  if (need > *tmp_bytes) {
    (void)hipFree(*tmp);
    *tmp = nullptr;
    *tmp_bytes = 0;
    HIPCHK(hipMalloc(tmp, need));
    *tmp_bytes = need;
  }
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(*tmp, need, flag, scan, (int)m, s));
  k_sgs_place<<<dim3(blocks), dim3(256), 0, s>>>(m, path, flag, scan, seen, pos, app, bad);
  HIPCHK(hipMemcpyAsync(out, bad, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int zero32(hipStream_t s, int32_t* p, int64_t n) {
  if (n <= 0) return 0;
  k_zero32<<<dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0, s>>>(p, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

void free_tiles(TileBuffers& tb) {
  (void)hipFree(tb.keys);
  (void)hipFree(tb.ids);
  (void)hipFree(tb.tmp);
  (void)hipFree(tb.big);
  tb = TileBuffers();
}

// One wavefront per target, the targets left open after the shell table by
// k_knn.
static int launch_waves(hipStream_t s, const SearchArgs& a, int64_t m, TileBuffers& tb) {
  if (m > 2147483647LL) return -1;
  if (tb.fixed && 3 * m + 1 > tb.cap_ids) return -1;
  if (grow(tb.ids, tb.cap_ids, 3 * m + 1)) return -1;
  int32_t* spill = tb.ids + 2 * m;
  int32_t* count = tb.ids + 3 * m;
  k_zero32<<<1, 64, 0, s>>>(count, 1);
  SearchArgs c = a;
  c.m = m;
  if (a.max_samples <= 32)
    k_knn_wave<32><<<dim3((unsigned)((m + 7) / 8)), dim3(256), 0, s>>>(c, spill, count);
  else
    k_knn_wave<64><<<dim3((unsigned)((m + 3) / 4)), dim3(256), 0, s>>>(c, spill, count);
  HIPCHK(hipGetLastError());
  tb.spill_count = count;
  return launch_spill(s, a, m, spill, count);
}

// Lanes per sector of k_sectors for this search, or 0 when it does not apply
// (plane sectors, a hole cap, or more than 64 list slots).
static int sector_lanes(const SearchArgs& a) {
  if (!a.balances || a.plane || a.max_per_hole >= 0) return 0;
  if (a.n_sectors == 8 && a.sector_cap <= 8) return 8;
  if (a.n_sectors == 4 && a.sector_cap <= 16) return 16;
  return 0;
}

// The targets of a capped search whose LDS pool overflowed (ovf, length on
// the device), again with global-memory pools (allocated on first use).
static int launch_big(hipStream_t s, const SearchArgs& a, int64_t m, int32_t* ovf, const int32_t* ovf_count,
                      TileBuffers& tb) {
  if (!tb.big) HIPCHK(hipMalloc(&tb.big, (size_t)BK_BIG_BLOCKS * sizeof(SearchBig)));
  SearchArgs d = a;
  d.list = ovf;
  d.m = m;
  d.m_dev = ovf_count;
  d.ovf = nullptr;
  d.ovf_count = nullptr;
  k_search_big<<<dim3(BK_BIG_BLOCKS), dim3(BK_WAVE), 0, s>>>(d, (SearchBig*)tb.big);
  HIPCHK(hipGetLastError());
  tb.retry_count = ovf_count;
  return 0;
}

// Capped searches by k_sectors, the targets it leaves open by k_search, and
// those whose pool overflows there by k_search_big (not with fixed buffers).
static int launch_sectors(hipStream_t s, const SearchArgs& a, int64_t m, int lanes, TileBuffers& tb) {
  if (m > 2147483647LL) return -1;
  const bool big = !tb.fixed;
  if (grow(tb.ids, tb.cap_ids, 3 * m + (big ? 2 : 1))) return -1;
  int32_t* spill = tb.ids + 2 * m;
  int32_t* count = tb.ids + 3 * m;
  int32_t* ovf = big ? tb.ids : nullptr;
  int32_t* ovf_count = big ? tb.ids + 3 * m + 1 : nullptr;
  k_zero32<<<1, 64, 0, s>>>(count, 1);
  if (big) k_zero32<<<1, 64, 0, s>>>(ovf_count, 1);
  SearchArgs c = a;
  c.m = m;
  const dim3 grid((unsigned)((m + 3) / 4));
  if (lanes == 8)
    k_sectors<8><<<grid, dim3(256), 0, s>>>(c, spill, count);
  else
    k_sectors<16><<<grid, dim3(256), 0, s>>>(c, spill, count);
  HIPCHK(hipGetLastError());
  tb.spill_count = count;
  SearchArgs d = a;
  d.list = spill;
  d.m = m;
  d.m_dev = count;
  d.ovf = ovf;
  d.ovf_count = ovf_count;
  k_search<<<dim3((unsigned)std::min<int64_t>(m, 4096)), dim3(BK_WAVE), 0, s>>>(d);
  HIPCHK(hipGetLastError());
  return big ? launch_big(s, a, m, ovf, ovf_count, tb) : 0;
}

int launch_search(hipStream_t s, const SearchArgs& a, int64_t n_targets, int mode, TileBuffers& tb) {
  if (n_targets <= 0) return 0;
  tb.spill_count = nullptr;
  tb.retry_count = nullptr;
  if (const int lanes = sector_lanes(a)) {
    if (!(tb.fixed && 3 * n_targets + 1 > tb.cap_ids)) return launch_sectors(s, a, n_targets, lanes, tb);
  }
  if (knn_path(a)) {
    if (mode > 0 && a.list == nullptr && !tb.fixed) return launch_tiles(s, a, n_targets, mode, tb);
    if (mode > 0 && tb.fixed) mode = 0;
    if (mode == 0) return launch_waves(s, a, n_targets, tb);
    // Selections without caps: one target per lane.
    SearchArgs c = a;
    c.m = n_targets;
    size_t sh = (size_t)a.max_samples * BK_WAVE * (sizeof(uint64_t) + sizeof(uint32_t));
    k_knn<<<dim3((unsigned)((n_targets + BK_WAVE - 1) / BK_WAVE)), dim3(BK_WAVE), sh, s>>>(c);
    HIPCHK(hipGetLastError());
    return 0;
  }
  // Capped searches: a block per target; targets whose pool overflows are
  // listed for k_search_big (one launch of at most 2^31 - 1 targets).
  const bool big = !tb.fixed && n_targets <= 2147483647LL;
  int32_t* ovf = nullptr;
  int32_t* ovf_count = nullptr;
  if (big) {
    if (grow(tb.ids, tb.cap_ids, n_targets + 1)) return -1;
    ovf = tb.ids;
    ovf_count = tb.ids + n_targets;
    k_zero32<<<1, 64, 0, s>>>(ovf_count, 1);
  }
  for (int64_t b = 0; b < n_targets; b += 2147483647LL) {
    int64_t nb = std::min<int64_t>(n_targets - b, 2147483647LL);
    SearchArgs c = a;
    c.ovf = ovf;
    c.ovf_count = ovf_count;
    if (b) {
      if (c.list)
        c.list += b;
      else
        return -1;  // more than 2^31 targets need a list
    }
    k_search<<<dim3((unsigned)nb), dim3(BK_WAVE), 0, s>>>(c);
  }
  HIPCHK(hipGetLastError());
  return big ? launch_big(s, a, n_targets, ovf, ovf_count, tb) : 0;
}

}  // namespace bk
