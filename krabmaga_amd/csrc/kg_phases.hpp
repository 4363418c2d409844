// kg_phases.hpp -- the bodies of the binning and flock kernels, written as
// __host__ __device__ functions of (block, thread) so that a host driver can
// execute them block by block, phase by phase, with the exact index
// arithmetic and f32 operations of the device build.
#pragma once

#include "kg_types.hpp"

namespace kg {

KG_HD uint32_t atom_add(uint32_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(p, v);
#else
  const uint32_t o = *p;
  *p = o + v;
  return o;
#endif
}

KG_HD void atom_or(uint32_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicOr(p, v);
#else
  *p |= v;
#endif
}

// Spreads consecutive work items over the 8 XCDs so that each XCD (and its
// L2) sees a contiguous run of tiles: hardware dispatches block b to XCD
// b % 8.
KG_HD uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t q = nb / 8u;
  if (b >= q * 8u) return b;
  return (b % 8u) * q + b / 8u;
}

// Last r with prefix[r] <= item (prefix has nreps + 1 entries).
KG_HD int find_rep(const uint32_t* prefix, int nreps, uint32_t item) {
  int lo = 0, hi = nreps;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---- set_object_location path: histogram and scatter of appended agents ----

KG_HD void count_append_one(const RepDev* reps, int nreps, const uint32_t* app_prefix, Buf A,
                            uint32_t i, uint32_t* count, uint32_t* err) {
  const int r = find_rep(app_prefix, nreps, i);
  const Geom& g = reps[r].g;
  int32_t cx, cy;
  discretize(g, A.x[i], A.y[i], &cx, &cy);
  if (!cell_valid(g, cx, cy)) {
    atom_or(err, 1u);
    return;
  }
  atom_add(&count[reps[r].bin_base + bin_of(g, cx, cy)], 1u);
}

KG_HD void scatter_append_one(const RepDev* reps, int nreps, const uint32_t* app_prefix, Buf A,
                              uint32_t i, uint32_t* cursor, Buf R2) {
  const int r = find_rep(app_prefix, nreps, i);
  const Geom& g = reps[r].g;
  const float x = A.x[i], y = A.y[i];
  int32_t cx, cy;
  discretize(g, x, y, &cx, &cy);
  if (!cell_valid(g, cx, cy)) return;
  const uint32_t pos = atom_add(&cursor[reps[r].bin_base + bin_of(g, cx, cy)], 1u);
  R2.x[pos] = x;
  R2.y[pos] = y;
  R2.hx[pos] = A.hx[i];
  R2.hy[pos] = A.hy[i];
  R2.id[pos] = A.id[i];
}

// ---- rebinning by gather ----------------------------------------------------
// After a flock step the birds of a tile's new bins are the tile's old birds
// that stayed in it plus birds that left neighbouring tiles: the flock
// kernel lists every bird that leaves its tile in that tile's outbox.  A
// work-group per tile collects its candidates, ranks each inside its new
// bin by (cell row, id) -- the order the bin sort gives -- and writes it to
// its final slot, so the scatter and the bin sort need not run.  A tile that
// cannot (more than kGatherCap candidates or new birds, an outbox past
// kOutboxCap, a bin past kGatherBinMax birds, or collected counts that
// differ from the histogram's, as when a bird jumps past the neighbouring
// tiles) sets the gate word, and then the scatter and the bin sort redo the
// whole rebinning of that step.

constexpr uint32_t kGatherThreads = 256;
constexpr uint32_t kGatherCap = 2048;     // candidates and new birds of a tile, at most
constexpr uint32_t kGatherRegs = kGatherCap / kGatherThreads;  // candidates per thread
constexpr uint32_t kGatherBinMax = 512;   // a fuller bin sends the step to the bin sort
constexpr uint32_t kGatherMaxBins = 256;  // bins per tile the gather handles
constexpr uint32_t kOutboxCap = 256;      // birds listed per tile and step
constexpr uint32_t kGatherSources = 16;   // neighbouring tiles read, at most

struct Outbox {
  uint32_t* n = nullptr;     // per tile (work item): birds that left it (> kOutboxCap: overflow)
  uint32_t* slot = nullptr;  // per tile: kOutboxCap read-buffer indices
};

// The gate of a gathered step: the flag a tile that cannot rebin sets (the
// scatter and the bin sort run only if it is set), the count of steps they
// redid, the flag the gather clears for the next step (or none: the scan
// clears this step's), and the arrival counter of the fallback's grid wait
// (the gather zeroes it; or none).
struct Gate {
  uint32_t* flag = nullptr;
  uint32_t* redone = nullptr;
  uint32_t* clear = nullptr;
  uint32_t* bar = nullptr;
};

// The scan done inside the gather, for a layout whose flock step left the
// sums of the scan tiles (whole tiles each; every replica's bins start on a
// scan tile): this step's bin counts and sums (or, with prefixed, the sums'
// exclusive prefixes), the next step's (cleared here), the scatter cursors
// to fill (or none: the fallback counts this step's counts down instead),
// the new offsets, the schedule-step counter to advance, the bin count and
// the last tile.  count == null: the offsets come from the scan
// kernels.
struct GatherScan {
  uint32_t* count = nullptr;
  const uint32_t* tile_sum = nullptr;
  uint32_t* count_next = nullptr;
  uint32_t* tile_sum_next = nullptr;
  uint32_t* cursor = nullptr;
  uint64_t* gstep = nullptr;
  uint32_t nbins = 0;
  uint32_t last_item = 0;
  int prefixed = 0;
};

// Entry i (0..3) of the tile columns (or rows) of n whose birds can enter
// column c in one step, or -1: c - 1, c, c + 1 cyclically and, across the
// seam of a torus whose last tile column may be narrower than a step, the
// partner at distance two (n - 2 for 0, 0 for n - 2); each value once.
KG_HD int32_t gather_axis_at(int32_t c, int32_t n, int i) {
  const int32_t a0 = (c + n - 1) % n, a1 = c, a2 = (c + 1) % n;
  const int32_t p = n >= 3 ? (c == 0 ? n - 2 : (c == n - 2 ? 0 : -1)) : -1;
  if (i == 0) return a0;
  if (i == 1) return a1 == a0 ? -1 : a1;
  if (i == 2) return (a2 == a0 || a2 == a1) ? -1 : a2;
  return (p < 0 || p == a0 || p == a1 || p == a2) ? -1 : p;
}

// Source q (< kGatherSources) of local tile `tile`: the local tile index, or
// -1 (not a neighbour, or the tile itself).
KG_HD int32_t gather_source(const Geom& g, uint32_t tile, uint32_t q) {
  const int32_t tcx = (int32_t)(tile % (uint32_t)g.ntx), tcy = (int32_t)(tile / (uint32_t)g.ntx);
  const int32_t sx = gather_axis_at(tcx, g.ntx, (int)(q & 3u));
  const int32_t sy = gather_axis_at(tcy, g.nty, (int)(q >> 2));
  if (sx < 0 || sy < 0 || (sx == tcx && sy == tcy)) return -1;
  return sy * g.ntx + sx;
}

// Candidate k of a tile: its old birds [lo, lo + nown), then the outbox
// entries of its sources (work items sitem, counts scnt, kGatherSources).
KG_HD uint32_t gather_src(const Outbox& ob, const uint32_t* sitem, const uint32_t* scnt,
                          uint32_t lo, uint32_t nown, uint32_t k) {
  if (k < nown) return lo + k;
  uint32_t m = k - nown, q = 0;
  while (m >= scnt[q]) m -= scnt[q++];
  return ob.slot[(size_t)sitem[q] * kOutboxCap + m];
}

// The tile-local bin of a bird at (x, y) if it lies in local tile `tile`,
// else kGatherMaxBins; *cy: its cell row.
KG_HD uint32_t gather_bin(const Geom& g, uint32_t tile, float x, float y, int32_t* cy) {
  int32_t cx;
  discretize(g, x, y, &cx, cy);
  if (!cell_valid(g, cx, *cy)) return kGatherMaxBins;
  const uint32_t nb = bin_of(g, cx, *cy);
  if ((nb >> (g.tx_log2 + g.tr_log2)) != tile) return kGatherMaxBins;
  return nb & ((1u << (g.tx_log2 + g.tr_log2)) - 1u);
}

// Rank of entry e among keys[lo, hi) of its bin by (cell row, id), ties by
// entry.
KG_HD uint32_t gather_rank(const uint64_t* key, uint32_t lo, uint32_t hi, uint32_t e,
                           uint64_t k) {
  uint32_t rank = 0;
  for (uint32_t j = lo; j < hi; ++j) {
    const uint64_t kj = key[j];
    rank += (kj < k || (kj == k && j < e)) ? 1u : 0u;
  }
  return rank;
}

// ---- rebinning of the flock kernel's output (aligned with R) ---------------

struct TileRange {
  int r;
  uint32_t lo, hi;  // read-buffer range of the tile
};

KG_HD TileRange tile_range(const RepDev* reps, int nreps, const uint32_t* tile_prefix,
                           uint32_t item, const uint32_t* offset) {
  TileRange tr;
  tr.r = find_rep(tile_prefix, nreps, item);
  const uint32_t tile = item - tile_prefix[tr.r];
  const Geom& g = reps[tr.r].g;
  const uint32_t b0 = reps[tr.r].bin_base + (tile << (g.tx_log2 + g.tr_log2));
  const uint32_t b1 = b0 + (1u << (g.tx_log2 + g.tr_log2));
  tr.lo = offset[b0];
  tr.hi = offset[b1];
  return tr;
}

KG_HD void scatter_aligned_one(const RepDev& rp, Buf R, Buf W, uint32_t i, uint32_t* cursor,
                               Buf R2) {
  const Geom& g = rp.g;
  const float x = W.x[i], y = W.y[i];
  int32_t cx, cy;
  discretize(g, x, y, &cx, &cy);
  if (!cell_valid(g, cx, cy)) return;  // reported by the flock kernel
  const uint32_t pos = atom_add(&cursor[rp.bin_base + bin_of(g, cx, cy)], 1u);
  R2.x[pos] = x;
  R2.y[pos] = y;
  R2.hx[pos] = W.hx[i];
  R2.hy[pos] = W.hy[i];
  R2.id[pos] = R.id[i];
}

// ---- bin order -------------------------------------------------------------
//
// A bin holds hb rows of one column; the scatters leave its entries in
// arbitrary order.  Every rebinning ends by sorting each bin by (cell y, id),
// ties by position, so the entries of one bag are a contiguous run in
// ascending id -- the canonical bag order -- and a query walks the run.

KG_HD uint64_t bin_key(const Geom& g, float y, uint32_t id) {
  return ((uint64_t)(uint32_t)disc_axis(y, g.d, g.inv_d) << 32) | id;
}

// Bins of at most kSortSmall entries are sorted by one lane: keys in
// registers (fixed indices only, so nothing lands in scratch), a bin already
// in order left untouched, otherwise every entry loaded before any is
// stored at its rank (ties by position).
constexpr uint32_t kSortSmall = 8;

KG_HD void sort_small_bin(const Geom& g, Buf R, uint32_t s, uint32_t m) {
  uint64_t key[kSortSmall];
#pragma unroll
  for (uint32_t u = 0; u < kSortSmall; ++u)
    key[u] = u < m ? bin_key(g, R.y[s + u], R.id[s + u]) : ~0ull;
  bool sorted = true;
#pragma unroll
  for (uint32_t u = 1; u < kSortSmall; ++u)
    if (u < m && key[u] < key[u - 1]) sorted = false;
  if (sorted) return;
  float x[kSortSmall], y[kSortSmall], hx[kSortSmall], hy[kSortSmall];
  uint32_t id[kSortSmall], rank[kSortSmall];
#pragma unroll
  for (uint32_t u = 0; u < kSortSmall; ++u) {
    const uint32_t q = u < m ? s + u : s;
    x[u] = R.x[q];
    y[u] = R.y[q];
    hx[u] = R.hx[q];
    hy[u] = R.hy[q];
    id[u] = R.id[q];
    rank[u] = 0;
#pragma unroll
    for (uint32_t v = 0; v < kSortSmall; ++v)
      rank[u] += (v < m) && (key[v] < key[u] || (key[v] == key[u] && v < u));
  }
#pragma unroll
  for (uint32_t u = 0; u < kSortSmall; ++u) {
    if (u < m) {
      const uint32_t q = s + rank[u];
      R.x[q] = x[u];
      R.y[q] = y[u];
      R.hx[q] = hx[u];
      R.hy[q] = hy[u];
      R.id[q] = id[u];
    }
  }
}

// ---- flock: fused neighbour query + Bird::step, one block per tile ---------
//
// LDS tile path (toroidal fields whose window reaches at most one tile
// further): the tile's own bins plus the halo cells the own agents' windows
// reach are staged into LDS -- bins in row order, so each staged column is in
// (cell y, id) order, the reference's bag-scan order -- and every own agent
// scans its window from LDS.  A tile whose staged set exceeds the LDS
// capacity, up to kFlockChunkFactor times it, is stepped in chunks of
// consecutive own columns, each staged with its own halo columns
// (flock_chunk_end).  Larger tiles, tiles with a single own column past the
// capacity, and fields outside the LDS path's reach take the global path
// (query_global per agent).

struct FlockTile {
  int r;
  uint32_t tile;
  int32_t tcy;
  int32_t c0, c1, y0, y1;          // own cells [c0,c1) x [y0,y1)
  int32_t sc_lo, ry_lo, ry_hi;     // staged columns start at sc_lo; staged rows
  int32_t ncols;                   // staged columns
  uint32_t tb0, tb1;               // the tile's bins
  uint64_t step;
};

KG_HD FlockTile flock_tile(const RepDev* reps, int nreps, const uint32_t* tile_prefix,
                           uint32_t item, uint64_t gstep) {
  FlockTile ft;
  ft.r = find_rep(tile_prefix, nreps, item);
  const RepDev& rp = reps[ft.r];
  const Geom& g = rp.g;
  const int32_t dd = rp.disc_dist;
  ft.tile = item - tile_prefix[ft.r];
  const int32_t tcx = (int32_t)(ft.tile % (uint32_t)g.ntx);
  ft.tcy = (int32_t)(ft.tile / (uint32_t)g.ntx);
  ft.c0 = tcx << g.tx_log2;
  ft.c1 = imin(ft.c0 + g.tx, g.dw);
  ft.y0 = ft.tcy * g.ty;
  ft.y1 = imin(ft.y0 + g.ty, g.dh);
  ft.tb0 = rp.bin_base + (ft.tile << (g.tx_log2 + g.tr_log2));
  ft.tb1 = ft.tb0 + (1u << (g.tx_log2 + g.tr_log2));
  // staged cells: own cells plus every cell an own agent's clamped window
  // reaches (field_2d.rs:411-416)
  ft.sc_lo = imax(0, ft.c0 - dd);
  const int32_t sc_hi = imax(ft.c1 - 1, imin(g.max_x - 1, ft.c1 - 1 + dd));
  ft.ry_lo = imax(0, ft.y0 - dd);
  ft.ry_hi = imax(ft.y1 - 1, imin(g.max_y - 1, ft.y1 - 1 + dd));
  ft.ncols = sc_hi - ft.sc_lo + 1;
  ft.step = rp.step_base + gstep;
  return ft;
}

// Staged column t (< ncols): up to three read-buffer runs -- the band of the
// tile row below, the tile's own row band, the band of the tile row above.
// Returns the column's total length.
KG_HD uint32_t flock_segments(const RepDev& rp, const FlockTile& ft, int t,
                              const uint32_t* offset, uint32_t* seg_start, uint32_t* seg_len) {
  const Geom& g = rp.g;
  const int32_t col = ft.sc_lo + t;
  const int32_t ttx = col >> g.tx_log2;
  const int32_t lc = col & (g.tx - 1);
  uint32_t len = 0;
  for (int k = 0; k < 3; ++k) {
    uint32_t st = 0, ln = 0;
    const int32_t tty = ft.tcy - 1 + k;
    if (tty >= 0 && tty < g.nty) {
      const int32_t by0 = tty * g.ty;
      const int32_t lo_row = imax(ft.ry_lo, by0);
      const int32_t hi_row = imin(ft.ry_hi, by0 + g.ty - 1);
      if (lo_row <= hi_row) {
        const uint32_t tt = (uint32_t)tty * (uint32_t)g.ntx + (uint32_t)ttx;
        const uint32_t cb = rp.bin_base + col_bin0(g, tt, lc);
        const uint32_t b_lo = cb + (uint32_t)((lo_row - by0) >> g.hb_log2);
        const uint32_t b_hi = cb + (uint32_t)((hi_row - by0) >> g.hb_log2) + 1u;
        st = offset[b_lo];
        ln = offset[b_hi] - st;
      }
    }
    seg_start[3 * t + k] = st;
    seg_len[3 * t + k] = ln;
    len += ln;
  }
  return len;
}

struct FlockSmem {
  float *x, *y, *hx, *hy;
  uint32_t *id, *gi;
  int32_t* cy;
  uint16_t *col, *own;
};

KG_HD size_t flock_smem_bytes(int cap) {
  return (size_t)cap * (7 * sizeof(uint32_t) + 2 * sizeof(uint16_t));
}

KG_HD FlockSmem flock_carve(uint8_t* p, int cap) {
  FlockSmem s;
  s.x = (float*)p;
  s.y = s.x + cap;
  s.hx = s.y + cap;
  s.hy = s.hx + cap;
  s.id = (uint32_t*)(s.hy + cap);
  s.gi = s.id + cap;
  s.cy = (int32_t*)(s.gi + cap);
  s.col = (uint16_t*)(s.cy + cap);
  s.own = s.col + cap;
  return s;
}

// A tile staged in chunks holds at most this many capacities' worth of
// staged items; a larger one takes the global path, where its birds spread
// over the whole GPU instead of running chunk after chunk in one work-group.
constexpr uint32_t kFlockChunkFactor = 2;

// Staged columns [*sa, *sb) (relative to sc_lo) that own staged columns
// [oa, ob) need: the own columns plus the columns their clamped windows reach.
KG_HD void flock_chunk_cols(const RepDev& rp, const FlockTile& ft, int32_t oa, int32_t ob,
                            int32_t* sa, int32_t* sb) {
  const int32_t dd = rp.disc_dist;
  const int32_t a = ft.sc_lo + oa, b = ft.sc_lo + ob;
  *sa = imax(0, a - dd) - ft.sc_lo;
  *sb = imax(b - 1, imin(rp.g.max_x - 1, b - 1 + dd)) - ft.sc_lo + 1;
}

// Staged items that own columns [oa, ob) need.
KG_HD uint32_t flock_chunk_size(const RepDev& rp, const FlockTile& ft, const uint32_t* col_base,
                                int32_t oa, int32_t ob) {
  int32_t sa, sb;
  flock_chunk_cols(rp, ft, oa, ob, &sa, &sb);
  return col_base[sb] - col_base[sa];
}

// End of the chunk of own columns that starts at oa: the most columns (up to
// ob_max) whose staged items fit cap; oa + 1 at least (the caller has checked
// that every single column fits).
KG_HD int32_t flock_chunk_end(const RepDev& rp, const FlockTile& ft, const uint32_t* col_base,
                              int32_t oa, int32_t ob_max, uint32_t cap) {
  int32_t ob = oa + 1;
  while (ob < ob_max && flock_chunk_size(rp, ft, col_base, oa, ob + 1) <= cap) ++ob;
  return ob;
}

// Stages item k (< total staged) into LDS slot dst.
KG_HD void flock_stage(const RepDev& rp, const FlockTile& ft, uint32_t k, uint32_t dst, Buf R,
                       const uint32_t* seg_start, const uint32_t* seg_len,
                       const uint32_t* col_base, FlockSmem S) {
  int lo = 0, hi = ft.ncols;  // column c with col_base[c] <= k < col_base[c+1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (col_base[mid] <= k) lo = mid;
    else hi = mid;
  }
  uint32_t off = k - col_base[lo];
  int seg = 0;
  while (off >= seg_len[3 * lo + seg]) {
    off -= seg_len[3 * lo + seg];
    ++seg;
  }
  const uint32_t gi = seg_start[3 * lo + seg] + off;
  const float y = R.y[gi];
  S.x[dst] = R.x[gi];
  S.y[dst] = y;
  S.hx[dst] = R.hx[gi];
  S.hy[dst] = R.hy[gi];
  S.id[dst] = R.id[gi];
  S.gi[dst] = gi;
  S.cy[dst] = disc_axis(y, rp.g.d, rp.g.inv_d);
  S.col[dst] = (uint16_t)lo;
}

// Appends LDS slot k to the own list if it holds one of the tile's agents in
// own staged columns [oa, ob).
KG_HD void flock_own(const FlockTile& ft, uint32_t k, int32_t oa, int32_t ob, FlockSmem S,
                     uint32_t* nown) {
  const int32_t ky = S.cy[k];
  const int32_t c = S.col[k];
  if (c >= oa && c < ob && ky >= ft.y0 && ky < ft.y1) {
    const uint32_t o = atom_add(nown, 1u);
    S.own[o] = (uint16_t)k;
  }
}

// Where a flock kernel records a bird's new place besides the bin counts:
// - the outboxes (or none) listing the birds that leave their tile, with
//   the work item of the replica's first tile (a bird's old tile is
//   item_base + its local tile) and the count to take slots from (a
//   work-group's LDS counter, or null for the outbox's own count);
// - the sums of the scan tiles (or none; each scan tile holds whole tiles)
//   and a histogram of the work-group's own tile's bins in LDS (or none):
//   with one, a bird that stays in its tile is counted there (the work-group
//   adds it to the bins and the sum at its end), any other bird in its bin
//   and its scan tile's sum at once.
struct FlockSinks {
  Outbox ob;
  uint32_t item_base;
  uint32_t* lds_n;
  uint32_t* lds_hist = nullptr;
  uint32_t* tile_sum = nullptr;
};

// Writes the agent's new state into the write buffer and counts its new bin;
// an agent that leaves its tile (old_tile, local to the replica) is listed
// in the tile's outbox, if there are outboxes.
KG_HD void flock_emit(const Geom& g, uint32_t bin_base, Buf W, uint32_t gi, float nx, float ny,
                      float nhx, float nhy, uint32_t* count_next, uint32_t* err,
                      uint32_t old_tile, const FlockSinks& lv) {
  W.x[gi] = nx;
  W.y[gi] = ny;
  W.hx[gi] = nhx;
  W.hy[gi] = nhy;
  int32_t cx, cy;
  discretize(g, nx, ny, &cx, &cy);
  if (!cell_valid(g, cx, cy)) {
    atom_or(err, 1u);
    return;
  }
  const uint32_t nb = bin_of(g, cx, cy);
  const bool stays = (nb >> (g.tx_log2 + g.tr_log2)) == old_tile;
  if (lv.lds_hist && stays) {
    atom_add(&lv.lds_hist[nb & ((1u << (g.tx_log2 + g.tr_log2)) - 1u)], 1u);
  } else {
    atom_add(&count_next[bin_base + nb], 1u);
    if (lv.tile_sum) atom_add(&lv.tile_sum[(bin_base + nb) / (uint32_t)kScanTile], 1u);
  }
  if (lv.ob.n && !stays) {
    const uint32_t item = lv.item_base + old_tile;
    const uint32_t o = atom_add(lv.lds_n ? lv.lds_n : &lv.ob.n[item], 1u);
    // past the capacity the count alone tells the gather to give up
    if (o < kOutboxCap) lv.ob.slot[(size_t)item * kOutboxCap + o] = gi;
  }
}

// Bird::step for own agent k of the staged tile, neighbours from LDS; LDS
// slot j holds staged item j + base.
KG_HD void flock_agent_lds(const RepDev& rp, const FlockTile& ft, uint32_t k,
                           const uint32_t* col_base, uint32_t base, FlockSmem S, Buf W,
                           uint32_t* count_next, uint32_t* err, const FlockSinks& lv) {
  const Geom& g = rp.g;
  const float T = rp.sq_thresh;
  const int32_t dd = rp.disc_dist;
  const float qx = S.x[k], qy = S.y[k];
  const uint32_t qid = S.id[k];
  const int32_t cx = ft.sc_lo + S.col[k];
  const int32_t cy = S.cy[k];
  // clamped window (field_2d.rs:401-416)
  const int32_t min_i = imax(0, cx - dd), max_i = imin(cx + dd, g.max_x - 1);
  const int32_t min_j = imax(0, cy - dd), max_j = imin(cy + dd, g.max_y - 1);
  // windows of up to 3 x 3 cells: every cell's check_circle once per agent
  const bool packed = max_i - min_i < 3 && max_j - min_j < 3;
  uint32_t wchk = 0;
  if (packed) {
    const int32_t bxs[3] = {min_i, imin(min_i + 1, max_i), imin(min_i + 2, max_i)};
    const int32_t bys[3] = {min_j, imin(min_j + 1, max_j), imin(min_j + 2, max_j)};
    wchk = window_checks(g, bxs, bys, max_i - min_i + 1, max_j - min_j + 1, qx, qy, T);
  }
  FlockAcc acc;
  for (int32_t i = min_i; i <= max_i; ++i) {
    const int c = i - ft.sc_lo;
    uint32_t lo = col_base[c] - base, hi = col_base[c + 1] - base;
    while (lo < hi) {  // first slot with cell y >= min_j
      const uint32_t mid = (lo + hi) >> 1;
      if (S.cy[mid] < min_j) lo = mid + 1;
      else hi = mid;
    }
    const uint32_t end = col_base[c + 1] - base;
    int32_t cur = -1;
    int chk = 0;
    for (uint32_t e = lo; e < end; ++e) {
      const int32_t ey_c = S.cy[e];
      if (ey_c > max_j) break;
      if (packed) {
        chk = window_check_at(wchk, i - min_i, ey_c - min_j);
      } else if (ey_c != cur) {
        cur = ey_c;
        chk = check_circle_sq(g, i, ey_c, qx, qy, T);
      }
      if (chk == -1) continue;
      // field_2d.rs:432 (toroidal distance) doubles as bird.rs:64-65
      const float dx = toroidal_distance(qx, S.x[e], g.w);
      const float dy = toroidal_distance(qy, S.y[e], g.h);
      if (chk == 1 || dx * dx + dy * dy <= T) acc.add(S.id[e] == qid, dx, dy, S.hx[e], S.hy[e]);
    }
  }
  float nx, ny, nhx, nhy;
  flock_finish(acc, rp.p, qid, ft.step, qx, qy, S.hx[k], S.hy[k], g.w, &nx, &ny, &nhx, &nhy);
  flock_emit(g, rp.bin_base, W, S.gi[k], nx, ny, nhx, nhy, count_next, err, ft.tile, lv);
}

// Bird::step for read-buffer agent gi, neighbours straight from HBM.
KG_HD void flock_agent_global(const RepDev& rp, uint64_t step, uint32_t gi, Buf R,
                              const uint32_t* offset, Buf W, uint32_t* count_next, uint32_t* err,
                              const FlockSinks& lv) {
  const Geom& g = rp.g;
  const float qx = R.x[gi], qy = R.y[gi];
  const uint32_t qid = R.id[gi];
  FlockAcc acc;
  query_global(g, R, offset + rp.bin_base, qx, qy, rp.p.radius, rp.sq_thresh, false,
               [&](uint32_t e) {
                 const float dx = toroidal_distance(qx, R.x[e], g.w);
                 const float dy = toroidal_distance(qy, R.y[e], g.h);
                 acc.add(R.id[e] == qid, dx, dy, R.hx[e], R.hy[e]);
               });
  float nx, ny, nhx, nhy;
  flock_finish(acc, rp.p, qid, step, qx, qy, R.hx[gi], R.hy[gi], g.w, &nx, &ny, &nhx, &nhy);
  // the old tile from the old position (fewer registers held over the walk)
  int32_t ocx, ocy;
  discretize(g, qx, qy, &ocx, &ocy);
  flock_emit(g, rp.bin_base, W, gi, nx, ny, nhx, nhy, count_next, err, tile_of(g, ocx, ocy), lv);
}

}  // namespace kg
