// kg_rt.hpp -- tile-resident agent step: the whole Schedule::step of the
// flockers model (neighbour query, Bird::step, rebinning, buffer swap) in one
// kernel launch per tile class.
//
// Layout.  Every tile of a replica owns a region of the stayer arrays
// (capacity scap, base sbase) holding its agents sorted by (bin, cell y, id)
// -- i.e. each column of the tile is in the reference's bag-scan order -- and
// a table boff of per-bin starts (tx * tr + 1 entries).  An agent whose new
// cell lies in another tile is appended to that tile's inbox (capacity icap,
// base ibase, atomic counter icnt); inbox entries that do not fit go to the
// engine-wide spill list (capacity = all agents, so nothing is ever lost).
// Stayer arrays, boff, counts and inbox payloads are double-buffered by the
// parity of the step counter; inbox/spill counters are triple-buffered by
// step % 3 so that the counters a step writes were cleared by the step
// before (each tile clears its own counter two steps ahead).
//
// Per step and tile (one work-group): stage the tile's own stayers plus the
// halo rows/columns its agents' clamped windows reach (already sorted) into
// LDS, gather the inbox/spill entries ("extras") of the 3 x 3 tile block that
// fall in the staged rectangle and sort them per column, index every staged
// column by cell row, run every own agent's query as a merge of the two
// sorted sequences (so neighbours are visited in bag-scan order,
// field_2d.rs:418-435), compute Bird::step (bird.rs:38-154) into an LDS
// result slot, then write the new states: stayers sorted into the other
// parity's region, movers into their destination tile's inbox.
//
// Tile classes.  Each tile is planned into the smallest of up to four
// classes whose LDS capacities hold it (capacities grow geometrically).  The
// class-0 launch covers every tile and returns at once for the others; each
// further class runs its tile list with larger capacities.  The class
// launches of a step touch disjoint tiles and run side by side; a tile of a
// class below the last that outgrew its capacities since the plan defers
// itself before writing anything, and one more launch, after the class
// launches, runs the deferred tiles with the last class's capacities.  A
// tile whose working set still exceeds its launch's capacities runs the same
// step from global memory and sends all its agents through the inboxes.
// The global path, spilling and frequent deferral make the host re-plan.
#pragma once

#include "kg_phases.hpp"

namespace kg {

constexpr int kRtThreads = 256;
// Tile classes of one plan: class 0 covers every tile in one launch, classes
// 1.. run their tile lists.
constexpr int kRtMaxClasses = 4;
// Launch modes of k_rt_step: a class below the last (its tiles may defer),
// the last class (no deferral), the deferred tiles of the step.
constexpr int kRtModeClass = 0;
constexpr int kRtModeFinal = 1;
constexpr int kRtModeDeferred = 2;
constexpr int kRtMaxBins = 1024;  // tx * tr of a tile on the LDS path

constexpr uint32_t kRtFlagSpill = 1u;    // an inbox overflowed into the spill list
constexpr uint32_t kRtFlagGlobal = 2u;   // a tile took the global path
constexpr uint32_t kRtFlagRange = 4u;    // a location left the bag grid
constexpr uint32_t kRtFlagStayCap = 8u;  // a tile's stayers exceeded its capacity
constexpr uint32_t kRtFlagDefer = 16u;   // a normal tile outgrew its class and was deferred

// Tiles that outgrow their class's capacities defer to the same step's
// launch for deferred tiles through a per-parity list ([0] count, then ids);
// that launch has kRtOvfMax work-groups.
constexpr uint32_t kRtOvfMax = 256;
// Deferrals per step past which the host re-plans a field: any (a step
// with a deferred tile pays one more work-group's latency after its class
// launches; a field's re-plan costs about one step's worth every few dozen).
constexpr uint32_t kRtReplanDeferred = 1;
// The same for an ensemble, whose re-plan rebuilds every run's layout (many
// milliseconds) while a step's deferred-tile launch costs tens of
// microseconds.
constexpr uint32_t kRtEnsReplanDeferred = kRtOvfMax / 2;

// Per-tile plan entry, two 16-byte loads.
struct alignas(16) RtMeta {
  uint32_t sbase;  // first stayer slot
  uint32_t scap;   // stayer capacity
  uint32_t ibase;  // first inbox slot
  uint32_t icap;   // inbox capacity
  uint32_t cls;    // 0 normal, 1 dense
  uint32_t pad[3];
};

struct RtDev {
  const RtMeta* meta;  // per global tile
  uint32_t* cnt[2];
  uint32_t* boff[2];
  uint32_t* icnt[3];
  uint32_t* scnt[3];      // one counter each
  Buf stay[2];
  Buf inbox[2];
  Buf spill[2];
  uint32_t* spill_rep[2];  // replica of each spill entry
  uint32_t spill_cap;
  uint32_t* flags;  // [0] kRtFlag* bits, [1] most tiles deferred in one step
  uint64_t* gstep;  // schedule step of launch k = step_base + *gstep + k
  uint32_t* ovf[2];  // deferred normal tiles, by step parity (kRtOvfMax + 1 each)
};

// LDS capacities of one launch.
struct RtCaps {
  int runs;    // staged stayers
  int extras;  // staged inbox/spill entries
  int own;     // own agents
  int nb;      // max tx * tr over the replicas (<= kRtMaxBins)
  int nbc;     // bin-table entries per staged column (staged bins + 1)
  int cols;    // staged columns (<= kMaxStageCols)
  int threads; // work-group size of the launch: 256, 512 or 1024
};

struct RtTile {
  int r;
  uint32_t gt;       // global tile id
  uint32_t tbase;    // global id of the replica's tile 0
  int32_t tcx, tcy;
  int32_t c0, c1, y0, y1;           // own cells [c0,c1) x [y0,y1)
  int32_t sc_lo, sc_hi, ry_lo, ry_hi;
  int32_t ncols;
  uint64_t step;
  int pin, ph_in, ph_out, ph_reset;
};

KG_HD RtTile rt_tile(const RepDev* reps, int nreps, const uint32_t* tile_prefix, uint32_t item,
                     uint64_t gstep) {
  RtTile rt;
  rt.r = find_rep(tile_prefix, nreps, item);
  const RepDev& rp = reps[rt.r];
  const Geom& g = rp.g;
  const int32_t dd = rp.disc_dist;
  rt.gt = item;
  rt.tbase = tile_prefix[rt.r];
  const uint32_t tile = item - rt.tbase;
  rt.tcx = (int32_t)(tile % (uint32_t)g.ntx);
  rt.tcy = (int32_t)(tile / (uint32_t)g.ntx);
  rt.c0 = rt.tcx << g.tx_log2;
  rt.c1 = imin(rt.c0 + g.tx, g.dw);
  rt.y0 = rt.tcy * g.ty;
  rt.y1 = imin(rt.y0 + g.ty, g.dh);
  rt.sc_lo = imax(0, rt.c0 - dd);
  rt.sc_hi = imax(rt.c1 - 1, imin(g.max_x - 1, rt.c1 - 1 + dd));
  rt.ry_lo = imax(0, rt.y0 - dd);
  rt.ry_hi = imax(rt.y1 - 1, imin(g.max_y - 1, rt.y1 - 1 + dd));
  rt.ncols = rt.sc_hi - rt.sc_lo + 1;
  rt.step = rp.step_base + gstep;
  rt.pin = (int)(gstep & 1u);
  rt.ph_in = (int)(gstep % 3u);
  rt.ph_out = (int)((gstep + 1u) % 3u);
  rt.ph_reset = (int)((gstep + 2u) % 3u);
  return rt;
}

// Global id of the tile at tile-grid (tx_, ty_) of the replica, or -1.
KG_HD int64_t rt_tile_at(const Geom& g, const RtTile& rt, int32_t ttx, int32_t tty) {
  if (ttx < 0 || ttx >= g.ntx || tty < 0 || tty >= g.nty) return -1;
  return (int64_t)rt.tbase + (int64_t)tty * g.ntx + ttx;
}

// Start of tile tt's bin-offset table (tx * tr + 1 entries) in D.boff[p]:
// the tiles of a replica have equal bin counts and consecutive tables.
KG_HD uint32_t rt_boff_at(const RepDev& rp, const RtTile& rt, uint32_t tt) {
  return rp.boff_base + (tt - rt.tbase) * (uint32_t)(rp.g.tx * rp.g.tr + 1);
}

// Rows of a staged column: up to three segments, the bins of tile rows
// tcy-1, tcy, tcy+1 that cover [ry_lo, ry_hi].  The same for every column.
struct RtSegs {
  int32_t blo[3];   // first staged bin (within its tile column)
  int32_t nbin[3];  // staged bins of the segment (0 if none)
  int32_t pre[3];   // bin-table index of the segment's first bin
  int32_t nbc;      // bin-table entries per column (staged bins + 1)
};

KG_HD RtSegs rt_segs(const Geom& g, const RtTile& rt) {
  RtSegs sg;
  int32_t pre = 0;
  for (int k = 0; k < 3; ++k) {
    const int32_t tty = rt.tcy - 1 + k;
    const int32_t by0 = tty * g.ty;
    const int32_t lo_row = imax(rt.ry_lo, by0);
    const int32_t hi_row = imin(rt.ry_hi, by0 + g.ty - 1);
    const bool on = tty >= 0 && tty < g.nty && lo_row <= hi_row;
    sg.blo[k] = on ? (lo_row - by0) >> g.hb_log2 : 0;
    sg.nbin[k] = on ? ((hi_row - by0) >> g.hb_log2) + 1 - sg.blo[k] : 0;
    sg.pre[k] = pre;
    pre += sg.nbin[k];
  }
  sg.nbc = pre + 1;
  return sg;
}

// Element k (0..2) of a three-entry table without indexing memory (a
// per-lane index into a local array would put the array in scratch).
KG_HD int32_t rt_pick3(const int32_t* a, int32_t k) {
  return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]);
}

// Bin-table index of staged row j (ry_lo <= j <= ry_hi).
KG_HD int32_t rt_seg_bin(const Geom& g, const RtTile& rt, const RtSegs& sg, int32_t j) {
  const int32_t k = (j >> (g.tr_log2 + g.hb_log2)) - (rt.tcy - 1);
  const int32_t b = (j & (g.ty - 1)) >> g.hb_log2;
  return rt_pick3(sg.pre, k) + b - rt_pick3(sg.blo, k);
}

// Whether the tile's staged rectangle fits the launch's tables.
KG_HD bool rt_fits_tables(const RtTile& rt, const RtSegs& sg, const RtCaps& caps) {
  return rt.ncols <= caps.cols && rt.ncols <= kMaxStageCols && sg.nbc <= caps.nbc;
}

// Staged column t: up to three sorted stayer runs (tile rows tcy-1, tcy,
// tcy+1) covering rows [ry_lo, ry_hi].  Returns the column's length.  The
// table loads of all three rows are issued before any is used.
KG_HD uint32_t rt_segments(const RepDev& rp, const RtTile& rt, const RtDev& D, int t,
                           uint32_t* seg_start, uint32_t* seg_len) {
  const Geom& g = rp.g;
  const int32_t col = rt.sc_lo + t;
  const int32_t ttx = col >> g.tx_log2;
  const int32_t lc = col & (g.tx - 1);
  bool on[3];
  uint32_t tt[3], blo[3], bhi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int32_t tty = rt.tcy - 1 + k;
    const int64_t t64 = rt_tile_at(g, rt, ttx, tty);
    const int32_t by0 = tty * g.ty;
    const int32_t lo_row = imax(rt.ry_lo, by0);
    const int32_t hi_row = imin(rt.ry_hi, by0 + g.ty - 1);
    on[k] = t64 >= 0 && lo_row <= hi_row;
    tt[k] = on[k] ? (uint32_t)t64 : rt.gt;
    blo[k] = on[k] ? ((uint32_t)lc << g.tr_log2) + (uint32_t)((lo_row - by0) >> g.hb_log2) : 0u;
    bhi[k] = on[k] ? ((uint32_t)lc << g.tr_log2) + (uint32_t)((hi_row - by0) >> g.hb_log2) + 1u
                   : 0u;
  }
  const uint32_t* bo = D.boff[rt.pin];
  uint32_t sb[3], lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t base = rt_boff_at(rp, rt, tt[k]);
    sb[k] = D.meta[tt[k]].sbase;
    lo[k] = bo[base + blo[k]];
    hi[k] = bo[base + bhi[k]];
  }
  uint32_t len = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t ln = on[k] ? hi[k] - lo[k] : 0u;
    seg_start[3 * t + k] = on[k] ? sb[k] + lo[k] : 0u;
    seg_len[3 * t + k] = ln;
    len += ln;
  }
  return len;
}

// Sources of extras: the inboxes of the 3 x 3 tile block (k = 0..8, row
// major from (tcx-1, tcy-1)) and the spill list (k = 9).
KG_HD uint32_t rt_extra_count(const RepDev& rp, const RtTile& rt, const RtDev& D, int k,
                              uint32_t* base) {
  if (k == 9) {
    uint32_t n = *D.scnt[rt.ph_in];
    if (n > D.spill_cap) n = D.spill_cap;
    *base = 0;
    return n;
  }
  const int64_t tt = rt_tile_at(rp.g, rt, rt.tcx - 1 + k % 3, rt.tcy - 1 + k / 3);
  if (tt < 0) {
    *base = 0;
    return 0;
  }
  uint32_t n = D.icnt[rt.ph_in][tt];
  const uint32_t icap = D.meta[tt].icap, ibase = D.meta[tt].ibase;
  if (n > icap) n = icap;
  *base = ibase;
  return n;
}

// Whether a cell belongs to the tile's staged working set (own cells, or a
// cell some own agent's clamped window reaches).
KG_HD bool rt_keep(const RtTile& rt, int32_t cx, int32_t cy) {
  if (cx >= rt.c0 && cx < rt.c1 && cy >= rt.y0 && cy < rt.y1) return true;
  return cx >= rt.sc_lo && cx <= rt.sc_hi && cy >= rt.ry_lo && cy <= rt.ry_hi;
}

KG_HD bool rt_own(const RtTile& rt, int32_t cx, int32_t cy) {
  return cx >= rt.c0 && cx < rt.c1 && cy >= rt.y0 && cy < rt.y1;
}

// LDS of the tile path: results and the output lists live through the
// whole step; the staging arrays are dead once every query has run.
struct RtSmem {
  // results, one per own agent
  float *rx, *ry, *rhx, *rhy;
  uint32_t* rid;
  // output: stayers in classification order, bin histogram and starts
  int32_t* ocy;
  uint32_t* oid;
  uint32_t* hist;   // nb + 1
  uint32_t* bbase;  // nb + 1
  uint32_t* bcur;   // nb
  uint16_t* obin;
  uint16_t* osrc;   // result slot of each stayer
  uint16_t* operm;
  // staging
  float *sx, *sy, *shx, *shy;
  uint32_t* sid;
  int32_t* scy;
  float *ex, *ey, *ehx, *ehy;
  uint32_t* eid;
  int32_t* ecy;
  uint32_t* own;     // own agents: (column << 16) | staged index, or | 0x8000 | extra slot
  uint16_t* ecol;
  uint16_t* eperm;   // sorted position -> extra slot
  uint16_t* bintab;  // per staged column: start of each staged bin, relative to the column
  uint16_t* ecbase;  // per staged column + 1: first sorted extra of the column
};

KG_HD size_t rt_smem_res(const RtCaps& c) {
  return (size_t)c.own * 34 + (size_t)(3 * c.nb + 2) * 4;
}
KG_HD size_t rt_smem_bytes(const RtCaps& c) {
  size_t n = rt_smem_res(c);
  n = (n + 15) & ~(size_t)15;
  n += (size_t)c.runs * 24 + (size_t)c.extras * 28 + (size_t)c.own * 4 +
       (size_t)c.cols * (size_t)c.nbc * 2 + (size_t)(c.cols + 2) * 2;
  return (n + 15) & ~(size_t)15;
}

KG_HD RtSmem rt_carve(uint8_t* p, const RtCaps& c) {
  RtSmem s;
  float* f = (float*)p;
  s.rx = f; f += c.own;
  s.ry = f; f += c.own;
  s.rhx = f; f += c.own;
  s.rhy = f; f += c.own;
  s.rid = (uint32_t*)f; f += c.own;
  s.ocy = (int32_t*)f; f += c.own;
  s.oid = (uint32_t*)f; f += c.own;
  s.hist = (uint32_t*)f; f += c.nb + 1;
  s.bbase = (uint32_t*)f; f += c.nb + 1;
  s.bcur = (uint32_t*)f; f += c.nb;
  uint16_t* h = (uint16_t*)f;
  s.obin = h; h += c.own;
  s.osrc = h; h += c.own;
  s.operm = h; h += c.own;
  f = (float*)(p + ((rt_smem_res(c) + 15) & ~(size_t)15));
  s.sx = f; f += c.runs;
  s.sy = f; f += c.runs;
  s.shx = f; f += c.runs;
  s.shy = f; f += c.runs;
  s.sid = (uint32_t*)f; f += c.runs;
  s.scy = (int32_t*)f; f += c.runs;
  s.ex = f; f += c.extras;
  s.ey = f; f += c.extras;
  s.ehx = f; f += c.extras;
  s.ehy = f; f += c.extras;
  s.eid = (uint32_t*)f; f += c.extras;
  s.ecy = (int32_t*)f; f += c.extras;
  s.own = (uint32_t*)f; f += c.own;
  h = (uint16_t*)f;
  s.ecol = h; h += c.extras;
  s.eperm = h; h += c.extras;
  s.bintab = h; h += (size_t)c.cols * (size_t)c.nbc;
  s.ecbase = h;
  return s;
}

// Global slot, staged column and segment (0 above, 1 own tile, 2 below) of
// run item k.
KG_HD uint32_t rt_stage_src(const RtTile& rt, uint32_t k, const uint32_t* seg_start,
                            const uint32_t* seg_len, const uint32_t* col_base, int* col,
                            int* seg_out) {
  int lo = 0, hi = rt.ncols;
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
  *col = lo;
  *seg_out = seg;
  return seg_start[3 * lo + seg] + off;
}

// Writes staged item k; own-tile items of own columns join the own list.
KG_HD void rt_stage_put(const RepDev& rp, const RtTile& rt, uint32_t k, int col, int seg, float x,
                        float y, float hx, float hy, uint32_t id, int cap_own, uint32_t* nown,
                        RtSmem S) {
  S.sx[k] = x;
  S.sy[k] = y;
  S.shx[k] = hx;
  S.shy[k] = hy;
  S.sid[k] = id;
  S.scy[k] = disc_axis(y, rp.g.d, rp.g.inv_d);
  const int32_t cx = rt.sc_lo + col;
  if (seg == 1 && cx >= rt.c0 && cx < rt.c1) {
    const uint32_t o = atom_add(nown, 1u);
    if (o < (uint32_t)cap_own) S.own[o] = ((uint32_t)col << 16) | k;
  }
}

// Stages run item k into LDS.
KG_HD void rt_stage(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t k,
                    const uint32_t* seg_start, const uint32_t* seg_len, const uint32_t* col_base,
                    int cap_own, uint32_t* nown, RtSmem S) {
  int col, seg;
  const uint32_t gi = rt_stage_src(rt, k, seg_start, seg_len, col_base, &col, &seg);
  const Buf& B = D.stay[rt.pin];
  rt_stage_put(rp, rt, k, col, seg, B.x[gi], B.y[gi], B.hx[gi], B.hy[gi], B.id[gi], cap_own, nown,
               S);
}

// Source k (0..9) and global slot of extra number j (flattened over the 10
// sources).
KG_HD uint32_t rt_extra_src(uint32_t j, const uint32_t* src_base, const uint32_t* src_pre, int* k) {
  int q = 0;
  while (j >= src_pre[q + 1]) ++q;
  *k = q;
  return src_base[q] + (j - src_pre[q]);
}

// Appends an extra to the LDS extras if it is in the staged rectangle (and
// to the own list if in an own cell).  Returns false if the extras capacity
// is exceeded.
KG_HD bool rt_extra_put(const RepDev& rp, const RtTile& rt, float x, float y, float hx, float hy,
                        uint32_t id, int cap_e, uint32_t* ne, int cap_own, uint32_t* nown,
                        RtSmem S) {
  int32_t cx, cy;
  discretize(rp.g, x, y, &cx, &cy);
  if (!rt_keep(rt, cx, cy)) return true;
  const uint32_t slot = atom_add(ne, 1u);
  if (slot >= (uint32_t)cap_e) return false;
  S.ex[slot] = x;
  S.ey[slot] = y;
  S.ehx[slot] = hx;
  S.ehy[slot] = hy;
  S.eid[slot] = id;
  S.ecy[slot] = cy;
  S.ecol[slot] = (uint16_t)(cx - rt.sc_lo);
  if (rt_own(rt, cx, cy)) {
    const uint32_t o = atom_add(nown, 1u);
    if (o < (uint32_t)cap_own) S.own[o] = ((uint32_t)(cx - rt.sc_lo) << 16) | 0x8000u | slot;
  }
  return true;
}

// Reads extra number j and, if it is in the staged rectangle, appends it to
// the LDS extras.  Returns false if the extras capacity is exceeded.
KG_HD bool rt_gather_extra(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t j,
                           const uint32_t* src_base, const uint32_t* src_pre, int cap_e,
                           uint32_t* ne, int cap_own, uint32_t* nown, RtSmem S) {
  int k;
  const uint32_t gi = rt_extra_src(j, src_base, src_pre, &k);
  const Buf& B = (k == 9) ? D.spill[rt.pin] : D.inbox[rt.pin];
  if (k == 9 && D.spill_rep[rt.pin][gi] != (uint32_t)rt.r) return true;
  return rt_extra_put(rp, rt, B.x[gi], B.y[gi], B.hx[gi], B.hy[gi], B.id[gi], cap_e, ne, cap_own,
                      nown, S);
}

// Run items k0 and k1 and extra j (each skipped when past its end) with all
// global loads issued before any LDS write.  Returns false if the extras
// capacity is exceeded.
KG_HD bool rt_fetch(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t k0, uint32_t k1,
                    uint32_t total, uint32_t j, uint32_t next, const uint32_t* seg_start,
                    const uint32_t* seg_len, const uint32_t* col_base, const uint32_t* src_base,
                    const uint32_t* src_pre, const RtCaps& caps, uint32_t* ne, uint32_t* nown,
                    RtSmem S) {
  const bool h0 = k0 < total, h1 = k1 < total, he = j < next;
  int c0 = 0, c1 = 0, s0 = 0, s1 = 0, ek = 0;
  uint32_t g0 = 0, g1 = 0, ge = 0;
  if (h0) g0 = rt_stage_src(rt, k0, seg_start, seg_len, col_base, &c0, &s0);
  if (h1) g1 = rt_stage_src(rt, k1, seg_start, seg_len, col_base, &c1, &s1);
  if (he) ge = rt_extra_src(j, src_base, src_pre, &ek);
  const Buf& B = D.stay[rt.pin];
  const Buf& E = (ek == 9) ? D.spill[rt.pin] : D.inbox[rt.pin];
  float x0 = 0, y0 = 0, hx0 = 0, hy0 = 0, x1 = 0, y1 = 0, hx1 = 0, hy1 = 0;
  float xe = 0, ye = 0, hxe = 0, hye = 0;
  uint32_t id0 = 0, id1 = 0, ide = 0, repe = (uint32_t)rt.r;
  if (h0) {
    x0 = B.x[g0]; y0 = B.y[g0]; hx0 = B.hx[g0]; hy0 = B.hy[g0]; id0 = B.id[g0];
  }
  if (h1) {
    x1 = B.x[g1]; y1 = B.y[g1]; hx1 = B.hx[g1]; hy1 = B.hy[g1]; id1 = B.id[g1];
  }
  if (he) {
    xe = E.x[ge]; ye = E.y[ge]; hxe = E.hx[ge]; hye = E.hy[ge]; ide = E.id[ge];
    if (ek == 9) repe = D.spill_rep[rt.pin][ge];
  }
  if (h0) rt_stage_put(rp, rt, k0, c0, s0, x0, y0, hx0, hy0, id0, caps.own, nown, S);
  if (h1) rt_stage_put(rp, rt, k1, c1, s1, x1, y1, hx1, hy1, id1, caps.own, nown, S);
  if (he && repe == (uint32_t)rt.r)
    return rt_extra_put(rp, rt, xe, ye, hxe, hye, ide, caps.extras, ne, caps.own, nown, S);
  return true;
}

// Bin table entry p (column p / nbc, entry p % nbc): the start of staged
// bin x of the column relative to the column's first staged item; entry
// nbc - 1 is the column's length.  From the bin offsets of the segment's tile
// and the lengths of the segments before it.
KG_HD void rt_bintab(const RepDev& rp, const RtTile& rt, const RtDev& D, const RtSegs& sg,
                     uint32_t p, const uint32_t* seg_len, RtSmem S) {
  const Geom& g = rp.g;
  const int c = (int)(p / (uint32_t)sg.nbc);
  const int32_t x = (int32_t)(p % (uint32_t)sg.nbc);
  const uint32_t* len = seg_len + 3 * c;
  if (x == sg.nbc - 1) {
    S.bintab[p] = (uint16_t)(len[0] + len[1] + len[2]);
    return;
  }
  // x < the staged bins, so k <= 2
  const int k = x < sg.pre[0] + sg.nbin[0] ? 0 : (x < sg.pre[1] + sg.nbin[1] ? 1 : 2);
  const uint32_t v0 = k > 0 ? len[0] : 0u, v1 = k > 1 ? len[1] : 0u;
  uint32_t v = v0 + v1;
  const int32_t b = x - rt_pick3(sg.pre, k);
  if (b > 0) {
    const int32_t col = rt.sc_lo + c;
    const int32_t ttx = col >> g.tx_log2;
    const int32_t lc = col & (g.tx - 1);
    const uint32_t tt = rt.tbase + (uint32_t)((rt.tcy - 1 + k) * g.ntx + ttx);
    const uint32_t* bo = D.boff[rt.pin] + rt_boff_at(rp, rt, tt);
    const uint32_t b0 = ((uint32_t)lc << g.tr_log2) + (uint32_t)rt_pick3(sg.blo, k);
    v += bo[b0 + (uint32_t)b] - bo[b0];
  }
  S.bintab[p] = (uint16_t)v;
}

// Sorts extra j among all extras by (column, cell y, id, slot).
KG_HD void rt_rank_extra(uint32_t j, uint32_t ne, RtSmem S) {
  const uint32_t c = S.ecol[j];
  const int32_t cy = S.ecy[j];
  const uint32_t id = S.eid[j];
  uint32_t rank = 0;
  for (uint32_t q = 0; q < ne; ++q) {
    const uint32_t qc = S.ecol[q];
    const int32_t qy = S.ecy[q];
    const uint32_t qid = S.eid[q];
    rank += (qc < c) || (qc == c && (qy < cy || (qy == cy && (qid < id || (qid == id && q < j)))));
  }
  S.eperm[rank] = (uint16_t)j;
}

// Start of column c among the sorted extras: the extras of lower columns.
KG_HD void rt_ecbase(int c, uint32_t ne, RtSmem S) {
  uint32_t n = 0;
  for (uint32_t q = 0; q < ne; ++q) n += S.ecol[q] < (uint32_t)c;
  S.ecbase[c] = (uint16_t)n;
}

// One neighbour candidate of query (qx, qy) in bag-scan order: the
// check_circle of its cell (from the window's packed checks when `packed`,
// else computed when the cell row changes), then the distance test and the
// accumulation.
KG_HD void rt_visit(const Geom& g, int32_t i, float T, bool packed, uint32_t wchk, int32_t min_i,
                    int32_t min_j, float qx, float qy, uint32_t qid, float ex_, float ey_,
                    float ehx_, float ehy_, uint32_t eid_, int32_t ecy_, int32_t* cur, int* chk,
                    FlockAcc* acc) {
  if (packed) {
    *chk = window_check_at(wchk, i - min_i, ecy_ - min_j);
  } else if (ecy_ != *cur) {
    *cur = ecy_;
    *chk = check_circle_sq(g, i, ecy_, qx, qy, T);
  }
  if (*chk == -1) return;
  const float dx = toroidal_distance(qx, ex_, g.w);
  const float dy = toroidal_distance(qy, ey_, g.h);
  if (*chk == 1 || dx * dx + dy * dy <= T) acc->add(eid_ == qid, dx, dy, ehx_, ehy_);
}

// Bird::step for own agent number o: neighbours come from the staged runs
// (the row table gives the window rows' range of each column) merged with
// the column's sorted extras; the new state goes to result slot o.  For
// windows of up to 3 x 3 cells every cell's check_circle is computed once per
// agent (window_checks) instead of at each cell row the walk enters.
KG_HD void rt_agent_lds(const RepDev& rp, const RtTile& rt, const RtSegs& sg, uint32_t o,
                        const uint32_t* col_base, RtSmem S) {
  const Geom& g = rp.g;
  const float T = rp.sq_thresh;
  const int32_t dd = rp.disc_dist;
  const uint32_t ref = S.own[o];
  const uint32_t j0 = ref & 0x7FFFu;
  float qx, qy, qhx, qhy;
  uint32_t qid;
  int32_t cy;
  const int32_t cx = rt.sc_lo + (int32_t)(ref >> 16);
  if (ref & 0x8000u) {
    qx = S.ex[j0]; qy = S.ey[j0]; qhx = S.ehx[j0]; qhy = S.ehy[j0]; qid = S.eid[j0];
    cy = S.ecy[j0];
  } else {
    qx = S.sx[j0]; qy = S.sy[j0]; qhx = S.shx[j0]; qhy = S.shy[j0]; qid = S.sid[j0];
    cy = S.scy[j0];
  }
  const int32_t min_i = imax(0, cx - dd), max_i = imin(cx + dd, g.max_x - 1);
  const int32_t min_j = imax(0, cy - dd), max_j = imin(cy + dd, g.max_y - 1);
  const int32_t x_lo = rt_seg_bin(g, rt, sg, min_j), x_hi = rt_seg_bin(g, rt, sg, max_j) + 1;
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
    const int c = i - rt.sc_lo;
    const uint16_t* tab = S.bintab + (size_t)c * (size_t)sg.nbc;
    const uint32_t base = col_base[c];
    uint32_t a = base + tab[x_lo], a_end = base + tab[x_hi];
    // skip the rows of min_j's bin below the window; rows ascend along the
    // column, so the items below min_j among the next eight are a prefix
    // (eight loads per round trip instead of one)
    while (a < a_end) {
      uint32_t n = 0;
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) {
        const uint32_t q = a + u < a_end ? a + u : a;
        n += (a + u < a_end && S.scy[q] < min_j) ? 1u : 0u;
      }
      a += n;
      if (n < 8u) break;
    }
    uint32_t b = S.ecbase[c], b_end = S.ecbase[c + 1];
    while (b < b_end && S.ecy[S.eperm[b]] < min_j) ++b;
    int32_t cur = -1;
    int chk = 0;
    if (!(b < b_end && S.ecy[S.eperm[b]] <= max_j)) {
      // no extras in this column's window rows (the common case): the
      // stayers alone, two items per round trip with all their fields
      for (; a < a_end; a += 2) {
        const uint32_t a1 = a + 1 < a_end ? a + 1 : a;
        const int32_t ecy0 = S.scy[a], ecy1 = S.scy[a1];
        const float ex0 = S.sx[a], ey0 = S.sy[a], ehx0 = S.shx[a], ehy0 = S.shy[a];
        const float ex1 = S.sx[a1], ey1 = S.sy[a1], ehx1 = S.shx[a1], ehy1 = S.shy[a1];
        const uint32_t eid0 = S.sid[a], eid1 = S.sid[a1];
        if (ecy0 > max_j) break;
        rt_visit(g, i, T, packed, wchk, min_i, min_j, qx, qy, qid, ex0, ey0, ehx0, ehy0, eid0,
                 ecy0, &cur, &chk, &acc);
        if (a + 1 >= a_end || ecy1 > max_j) break;
        rt_visit(g, i, T, packed, wchk, min_i, min_j, qx, qy, qid, ex1, ey1, ehx1, ehy1, eid1,
                 ecy1, &cur, &chk, &acc);
      }
      continue;
    }
    for (;;) {
      const bool ha = a < a_end && S.scy[a] <= max_j;
      uint32_t e = 0;
      bool hb = false;
      if (b < b_end) {
        e = S.eperm[b];
        hb = S.ecy[e] <= max_j;
      }
      if (!ha && !hb) break;
      bool take_a = ha;
      if (ha && hb)
        take_a = S.scy[a] < S.ecy[e] || (S.scy[a] == S.ecy[e] && S.sid[a] <= S.eid[e]);
      const uint32_t src = take_a ? a : e;
      const float* px = take_a ? S.sx : S.ex;
      const float* py = take_a ? S.sy : S.ey;
      const float* phx = take_a ? S.shx : S.ehx;
      const float* phy = take_a ? S.shy : S.ehy;
      const uint32_t* pid = take_a ? S.sid : S.eid;
      const int32_t* pcy = take_a ? S.scy : S.ecy;
      const float ex_ = px[src], ey_ = py[src], ehx_ = phx[src], ehy_ = phy[src];
      const uint32_t eid_ = pid[src];
      const int32_t ecy_ = pcy[src];
      a += take_a ? 1u : 0u;
      b += take_a ? 0u : 1u;
      rt_visit(g, i, T, packed, wchk, min_i, min_j, qx, qy, qid, ex_, ey_, ehx_, ehy_, eid_, ecy_,
               &cur, &chk, &acc);
    }
  }
  float nx, ny, nhx, nhy;
  flock_finish(acc, rp.p, qid, rt.step, qx, qy, qhx, qhy, g.w, &nx, &ny, &nhx, &nhy);
  S.rx[o] = nx;
  S.ry[o] = ny;
  S.rhx[o] = nhx;
  S.rhy[o] = nhy;
  S.rid[o] = qid;
}

// Sends one agent to its destination tile's inbox (or the spill list).
KG_HD void rt_push(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t dest, float x,
                   float y, float hx, float hy, uint32_t id) {
  const int pout = rt.pin ^ 1;
  const uint32_t icap = D.meta[dest].icap, ibase = D.meta[dest].ibase;  // loaded with the atomic
  const uint32_t slot = atom_add(&D.icnt[rt.ph_out][dest], 1u);
  if (slot < icap) {
    const Buf& B = D.inbox[pout];
    const uint32_t gi = ibase + slot;
    B.x[gi] = x; B.y[gi] = y; B.hx[gi] = hx; B.hy[gi] = hy; B.id[gi] = id;
    return;
  }
  const uint32_t s = atom_add(D.scnt[rt.ph_out], 1u);
  atom_or(D.flags, kRtFlagSpill);
  if (s < D.spill_cap) {
    const Buf& B = D.spill[pout];
    B.x[s] = x; B.y[s] = y; B.hx[s] = hx; B.hy[s] = hy; B.id[s] = id;
    D.spill_rep[pout][s] = (uint32_t)rt.r;
  }
  (void)rp;
}

// Classifies result o: a stayer joins the LDS output list and the bin
// histogram, a mover is pushed at once.  Out-of-grid agents are dropped and
// flagged.
KG_HD void rt_classify(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t o,
                       uint32_t* nstay, RtSmem S) {
  const Geom& g = rp.g;
  const float x = S.rx[o], y = S.ry[o];
  int32_t cx, cy;
  discretize(g, x, y, &cx, &cy);
  if (!cell_valid(g, cx, cy)) {
    atom_or(D.flags, kRtFlagRange);
    return;
  }
  const uint32_t dest = rt.tbase + tile_of(g, cx, cy);
  if (dest != rt.gt) {
    rt_push(rp, rt, D, dest, x, y, S.rhx[o], S.rhy[o], S.rid[o]);
    return;
  }
  const uint32_t slot = atom_add(nstay, 1u);
  const uint32_t bin = ((uint32_t)(cx & (g.tx - 1)) << g.tr_log2) +
                       (uint32_t)((cy & (g.ty - 1)) >> g.hb_log2);
  S.ocy[slot] = cy;
  S.oid[slot] = S.rid[o];
  S.osrc[slot] = (uint16_t)o;
  S.obin[slot] = (uint16_t)bin;
  atom_add(&S.hist[bin], 1u);
}

// Places stayer s inside its bin by (cell y, id, slot) and writes it out.
// Stayers sorted past the tile's capacity go to its own inbox instead (the
// tile keeps the first scap of its sorted stayers; boff is clipped).
KG_HD void rt_write_stayer(const RepDev& rp, const RtTile& rt, const RtDev& D, RtMeta me,
                           uint32_t s, RtSmem S) {
  const uint32_t b = S.obin[s];
  const uint32_t lo = S.bbase[b], hi = S.bbase[b + 1];
  const int32_t cy = S.ocy[s];
  const uint32_t id = S.oid[s];
  uint32_t rank = 0;
  for (uint32_t q = lo; q < hi; ++q) {
    const uint32_t o = S.operm[q];
    const int32_t qy = S.ocy[o];
    const uint32_t qid = S.oid[o];
    rank += (qy < cy) || (qy == cy && (qid < id || (qid == id && o < s)));
  }
  const uint32_t pos = lo + rank;
  const uint32_t r = S.osrc[s];
  if (pos >= me.scap) {
    atom_or(D.flags, kRtFlagStayCap);
    rt_push(rp, rt, D, rt.gt, S.rx[r], S.ry[r], S.rhx[r], S.rhy[r], id);
    return;
  }
  const Buf& B = D.stay[rt.pin ^ 1];
  const uint32_t gi = me.sbase + pos;
  B.x[gi] = S.rx[r]; B.y[gi] = S.ry[r]; B.hx[gi] = S.rhx[r]; B.hy[gi] = S.rhy[r]; B.id[gi] = id;
}

// ---- global path (tiles beyond the LDS capacities) --------------------------

// Visits, in bag-scan order, the neighbours of (qx, qy) among the stayers
// and extras of the read parity; visit(x, y, hx, hy, id).
template <class Visit>
KG_HD void rt_query_global(const RepDev& rp, const RtTile& rt, const RtDev& D, float qx, float qy,
                           Visit&& visit) {
  const Geom& g = rp.g;
  const float T = rp.sq_thresh;
  const int32_t dd = rp.disc_dist;
  int32_t cx, cy;
  discretize(g, qx, qy, &cx, &cy);
  const int32_t min_i = imax(0, cx - dd), max_i = imin(cx + dd, g.max_x - 1);
  const int32_t min_j = imax(0, cy - dd), max_j = imin(cy + dd, g.max_y - 1);
  uint32_t src_base[10], src_n[10];
  for (int k = 0; k < 10; ++k) src_n[k] = rt_extra_count(rp, rt, D, k, &src_base[k]);
  const Buf& SB = D.stay[rt.pin];
  for (int32_t i = min_i; i <= max_i; ++i) {
    const int32_t ttx = i >> g.tx_log2;
    const int32_t lc = i & (g.tx - 1);
    for (int32_t j = min_j; j <= max_j; ++j) {
      const int check = check_circle_sq(g, i, j, qx, qy, T);
      if (check == -1) continue;
      // stayers of cell (i, j): part of one bin, sorted by (cell y, id)
      const int32_t tty = j / g.ty;
      const int64_t tt = rt_tile_at(g, rt, ttx, tty);
      uint32_t a = 0, a_end = 0;
      if (tt >= 0) {
        const uint32_t* bo = D.boff[rt.pin] + rt_boff_at(rp, rt, (uint32_t)tt);
        const uint32_t b = ((uint32_t)lc << g.tr_log2) + (uint32_t)((j - tty * g.ty) >> g.hb_log2);
        const uint32_t sb = D.meta[tt].sbase;
        a = sb + bo[b];
        a_end = sb + bo[b + 1];
        while (a < a_end && disc_axis(SB.y[a], g.d, g.inv_d) < j) ++a;
      }
      // extras of cell (i, j) in ascending (id, source, index) by selection
      uint64_t last = 0;  // (id + 1) << 32 | position of the last taken extra
      bool have_last = false;
      for (;;) {
        const bool ha = a < a_end && disc_axis(SB.y[a], g.d, g.inv_d) == j;
        uint64_t best = ~0ull;
        float bx = 0, by = 0, bhx = 0, bhy = 0;
        uint32_t bid = 0;
        uint32_t pos = 0;
        for (int k = 0; k < 10; ++k) {
          const Buf& B = (k == 9) ? D.spill[rt.pin] : D.inbox[rt.pin];
          for (uint32_t q = 0; q < src_n[k]; ++q, ++pos) {
            const uint32_t gi = src_base[k] + q;
            if (k == 9 && D.spill_rep[rt.pin][gi] != (uint32_t)rt.r) continue;
            const uint32_t id = B.id[gi];
            const uint64_t key = ((uint64_t)id << 32) | pos;
            if (have_last && key <= last) continue;
            if (key >= best) continue;
            int32_t ex, ey;
            discretize(g, B.x[gi], B.y[gi], &ex, &ey);
            if (ex != i || ey != j) continue;
            best = key;
            bx = B.x[gi]; by = B.y[gi]; bhx = B.hx[gi]; bhy = B.hy[gi]; bid = id;
          }
        }
        const bool hb = best != ~0ull;
        if (!ha && !hb) break;
        bool take_a = ha && (!hb || SB.id[a] <= bid);
        float ex_, ey_, ehx_, ehy_;
        uint32_t eid_;
        if (take_a) {
          ex_ = SB.x[a]; ey_ = SB.y[a]; ehx_ = SB.hx[a]; ehy_ = SB.hy[a]; eid_ = SB.id[a];
          ++a;
        } else {
          ex_ = bx; ey_ = by; ehx_ = bhx; ehy_ = bhy; eid_ = bid;
          last = best;
          have_last = true;
        }
        const float dx = toroidal_distance(qx, ex_, g.w);
        const float dy = toroidal_distance(qy, ey_, g.h);
        if (check == 1 || dx * dx + dy * dy <= T) visit(dx, dy, ehx_, ehy_, eid_);
      }
    }
  }
}

// Own agent number o of the tile on the global path: stayers first, then
// the own tile's inbox, then spill entries in the tile.  Returns false past
// the end.
KG_HD bool rt_own_global(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t o, float* x,
                         float* y, float* hx, float* hy, uint32_t* id, bool* skip) {
  *skip = false;
  const uint32_t ns = D.cnt[rt.pin][rt.gt];
  if (o < ns) {
    const Buf& B = D.stay[rt.pin];
    const uint32_t gi = D.meta[rt.gt].sbase + o;
    *x = B.x[gi]; *y = B.y[gi]; *hx = B.hx[gi]; *hy = B.hy[gi]; *id = B.id[gi];
    return true;
  }
  o -= ns;
  uint32_t ib;
  const uint32_t ni = rt_extra_count(rp, rt, D, 4, &ib);
  if (o < ni) {
    const Buf& B = D.inbox[rt.pin];
    const uint32_t gi = ib + o;
    *x = B.x[gi]; *y = B.y[gi]; *hx = B.hx[gi]; *hy = B.hy[gi]; *id = B.id[gi];
    return true;
  }
  o -= ni;
  uint32_t sb;
  const uint32_t nsp = rt_extra_count(rp, rt, D, 9, &sb);
  if (o < nsp) {
    const Buf& B = D.spill[rt.pin];
    *x = B.x[o]; *y = B.y[o]; *hx = B.hx[o]; *hy = B.hy[o]; *id = B.id[o];
    int32_t cx, cy;
    discretize(rp.g, *x, *y, &cx, &cy);
    *skip = D.spill_rep[rt.pin][o] != (uint32_t)rt.r || !rt_own(rt, cx, cy);
    return true;
  }
  return false;
}

KG_HD void rt_agent_global(const RepDev& rp, const RtTile& rt, const RtDev& D, float qx, float qy,
                           float qhx, float qhy, uint32_t qid, float* nx, float* ny, float* nhx,
                           float* nhy) {
  FlockAcc acc;
  rt_query_global(rp, rt, D, qx, qy, [&](float dx, float dy, float ehx, float ehy, uint32_t eid) {
    acc.add(eid == qid, dx, dy, ehx, ehy);
  });
  flock_finish(acc, rp.p, qid, rt.step, qx, qy, qhx, qhy, rp.g.w, nx, ny, nhx, nhy);
}

// Global path of a whole tile, thread t of nthreads: every own agent is
// stepped from global memory and pushed to its destination's inbox (its own
// tile's inbox for stayers).
KG_HD void rt_tile_global(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t t,
                          uint32_t nthreads) {
  for (uint32_t o = t;; o += nthreads) {
    float x, y, hx, hy;
    uint32_t id;
    bool skip;
    if (!rt_own_global(rp, rt, D, o, &x, &y, &hx, &hy, &id, &skip)) break;
    if (skip) continue;
    float ax, ay, ahx, ahy;
    rt_agent_global(rp, rt, D, x, y, hx, hy, id, &ax, &ay, &ahx, &ahy);
    int32_t cx, cy;
    discretize(rp.g, ax, ay, &cx, &cy);
    if (!cell_valid(rp.g, cx, cy)) {
      atom_or(D.flags, kRtFlagRange);
      continue;
    }
    rt_push(rp, rt, D, rt.tbase + tile_of(rp.g, cx, cy), ax, ay, ahx, ahy, id);
  }
}

}  // namespace kg
