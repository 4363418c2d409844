// kg_rt.hpp -- tile-resident agent step: the whole Schedule::step of the
// flockers model (neighbour query, Bird::step, rebinning, buffer swap) in one
// kernel launch.
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
// fall in the staged rectangle and sort them per column, run every own
// agent's query as a merge of the two sorted sequences (so neighbours are
// visited in bag-scan order, field_2d.rs:418-435), compute Bird::step
// (bird.rs:38-154), and write the new states: stayers sorted into the other
// parity's region, movers into their destination tile's inbox.  Tiles whose
// working set does not fit the LDS capacities run the same step from global
// memory and send all their agents through the inboxes.
#pragma once

#include "kg_phases.hpp"

namespace kg {

constexpr int kRtThreads = 256;
constexpr int kRtOwnPerThread = 4;
constexpr int kRtMaxOwn = kRtThreads * kRtOwnPerThread;
constexpr int kRtMaxBins = 1024;  // tx * tr of a tile on the LDS path

constexpr uint32_t kRtFlagSpill = 1u;    // an inbox overflowed into the spill list
constexpr uint32_t kRtFlagGlobal = 2u;   // a tile took the global path
constexpr uint32_t kRtFlagRange = 4u;    // a location left the bag grid
constexpr uint32_t kRtFlagStayCap = 8u;  // a tile's stayers exceeded its capacity

struct RtDev {
  const uint32_t* sbase;  // per global tile
  const uint32_t* scap;
  const uint32_t* ibase;
  const uint32_t* icap;
  const uint32_t* tboff;  // start of the tile's boff table
  uint32_t* cnt[2];
  uint32_t* boff[2];
  uint32_t* icnt[3];
  uint32_t* scnt[3];      // one counter each
  Buf stay[2];
  Buf inbox[2];
  Buf spill[2];
  uint32_t* spill_rep[2];  // replica of each spill entry
  uint32_t spill_cap;
  uint32_t* flags;
  uint32_t* done;
  uint64_t* gstep;
};

// LDS capacities of one launch.
struct RtCaps {
  int runs;   // staged stayers
  int extras; // staged inbox/spill entries
  int own;    // own agents (<= kRtMaxOwn)
  int nb;     // max tx * tr over the replicas (<= kRtMaxBins)
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

// Staged column t: up to three sorted stayer runs (tile rows tcy-1, tcy,
// tcy+1) covering rows [ry_lo, ry_hi].  Returns the column's length.
KG_HD uint32_t rt_segments(const RepDev& rp, const RtTile& rt, const RtDev& D, int t,
                           uint32_t* seg_start, uint32_t* seg_len) {
  const Geom& g = rp.g;
  const int32_t col = rt.sc_lo + t;
  const int32_t ttx = col >> g.tx_log2;
  const int32_t lc = col & (g.tx - 1);
  uint32_t len = 0;
  for (int k = 0; k < 3; ++k) {
    uint32_t st = 0, ln = 0;
    const int32_t tty = rt.tcy - 1 + k;
    const int64_t tt = rt_tile_at(g, rt, ttx, tty);
    if (tt >= 0) {
      const int32_t by0 = tty * g.ty;
      const int32_t lo_row = imax(rt.ry_lo, by0);
      const int32_t hi_row = imin(rt.ry_hi, by0 + g.ty - 1);
      if (lo_row <= hi_row) {
        const uint32_t* bo = D.boff[rt.pin] + D.tboff[tt];
        const uint32_t b_lo = ((uint32_t)lc << g.tr_log2) + (uint32_t)((lo_row - by0) >> g.hb_log2);
        const uint32_t b_hi =
            ((uint32_t)lc << g.tr_log2) + (uint32_t)((hi_row - by0) >> g.hb_log2) + 1u;
        st = D.sbase[tt] + bo[b_lo];
        ln = bo[b_hi] - bo[b_lo];
      }
    }
    seg_start[3 * t + k] = st;
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
  if (n > D.icap[tt]) n = D.icap[tt];
  *base = D.ibase[tt];
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

// LDS of the tile path.  Phase A (query) and phase B (output) overlay.
struct RtSmem {
  // phase A
  float *sx, *sy, *shx, *shy;
  uint32_t* sid;
  int32_t* scy;
  uint16_t* scol;
  float *ex, *ey, *ehx, *ehy;
  uint32_t* eid;
  int32_t* ecy;
  uint16_t* ecol;
  uint16_t* eperm;
  uint16_t* own;  // own agents: staged index, or 0x8000 | extra index
  // phase B
  float *ox, *oy, *ohx, *ohy;
  uint32_t* oid;
  int32_t* ocy;
  uint16_t* obin;
  uint16_t* operm;
  uint32_t* hist;   // nb + 1
  uint32_t* bbase;  // nb + 1
  uint32_t* bcur;   // nb
};

KG_HD size_t rt_smem_a(const RtCaps& c) {
  return (size_t)c.runs * 26 + (size_t)c.extras * 30 + (size_t)c.own * 2;
}
KG_HD size_t rt_smem_b(const RtCaps& c) {
  return (size_t)c.own * 30 + (size_t)(3 * c.nb + 2) * 4;
}
KG_HD size_t rt_smem_bytes(const RtCaps& c) {
  const size_t a = rt_smem_a(c), b = rt_smem_b(c);
  return ((a > b ? a : b) + 15) & ~(size_t)15;
}

KG_HD RtSmem rt_carve(uint8_t* p, const RtCaps& c) {
  RtSmem s;
  // phase A: 4-byte arrays first, then 2-byte arrays
  float* f = (float*)p;
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
  uint16_t* h = (uint16_t*)f;
  s.scol = h; h += c.runs;
  s.ecol = h; h += c.extras;
  s.eperm = h; h += c.extras;
  s.own = h;
  // phase B
  f = (float*)p;
  s.ox = f; f += c.own;
  s.oy = f; f += c.own;
  s.ohx = f; f += c.own;
  s.ohy = f; f += c.own;
  s.oid = (uint32_t*)f; f += c.own;
  s.ocy = (int32_t*)f; f += c.own;
  s.hist = (uint32_t*)f; f += c.nb + 1;
  s.bbase = (uint32_t*)f; f += c.nb + 1;
  s.bcur = (uint32_t*)f; f += c.nb;
  h = (uint16_t*)f;
  s.obin = h; h += c.own;
  s.operm = h;
  return s;
}

// Stages run item k into LDS.
KG_HD void rt_stage(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t k,
                    const uint32_t* seg_start, const uint32_t* seg_len, const uint32_t* col_base,
                    RtSmem S) {
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
  const Buf& B = D.stay[rt.pin];
  const uint32_t gi = seg_start[3 * lo + seg] + off;
  const float y = B.y[gi];
  S.sx[k] = B.x[gi];
  S.sy[k] = y;
  S.shx[k] = B.hx[gi];
  S.shy[k] = B.hy[gi];
  S.sid[k] = B.id[gi];
  S.scy[k] = f2i(floorf(y / rp.g.d));
  S.scol[k] = (uint16_t)lo;
}

// Reads extra number j (flattened over the 10 sources) and, if it is in the
// staged rectangle, appends it to the LDS extras.  Returns false if the
// extras capacity is exceeded.
KG_HD bool rt_gather_extra(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t j,
                           const uint32_t* src_base, const uint32_t* src_pre, int cap_e,
                           uint32_t* ne, RtSmem S) {
  int k = 0;
  while (j >= src_pre[k + 1]) ++k;
  const uint32_t gi = src_base[k] + (j - src_pre[k]);
  const Buf& B = (k == 9) ? D.spill[rt.pin] : D.inbox[rt.pin];
  if (k == 9 && D.spill_rep[rt.pin][gi] != (uint32_t)rt.r) return true;
  const float x = B.x[gi], y = B.y[gi];
  int32_t cx, cy;
  discretize(rp.g, x, y, &cx, &cy);
  if (!rt_keep(rt, cx, cy)) return true;
  const uint32_t slot = atom_add(ne, 1u);
  if (slot >= (uint32_t)cap_e) return false;
  S.ex[slot] = x;
  S.ey[slot] = y;
  S.ehx[slot] = B.hx[gi];
  S.ehy[slot] = B.hy[gi];
  S.eid[slot] = B.id[gi];
  S.ecy[slot] = cy;
  S.ecol[slot] = (uint16_t)(cx - rt.sc_lo);
  return true;
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

// Start of column c in the sorted extras.
KG_HD uint32_t rt_extra_col_base(int c, uint32_t ne, RtSmem S) {
  uint32_t n = 0;
  for (uint32_t q = 0; q < ne; ++q) n += S.ecol[q] < (uint32_t)c;
  return n;
}

// Bird::step for one own agent with neighbours merged from the staged runs
// and the sorted extras.
KG_HD void rt_agent_lds(const RepDev& rp, const RtTile& rt, uint16_t ref, const uint32_t* col_base,
                        const uint32_t* ecol_base, RtSmem S, float* nx, float* ny, float* nhx,
                        float* nhy, uint32_t* nid) {
  const Geom& g = rp.g;
  const float T = rp.sq_thresh;
  const int32_t dd = rp.disc_dist;
  float qx, qy, qhx, qhy;
  uint32_t qid;
  int32_t cx, cy;
  if (ref & 0x8000u) {
    const uint32_t j = ref & 0x7FFFu;
    qx = S.ex[j]; qy = S.ey[j]; qhx = S.ehx[j]; qhy = S.ehy[j]; qid = S.eid[j];
    cx = rt.sc_lo + S.ecol[j];
    cy = S.ecy[j];
  } else {
    qx = S.sx[ref]; qy = S.sy[ref]; qhx = S.shx[ref]; qhy = S.shy[ref]; qid = S.sid[ref];
    cx = rt.sc_lo + S.scol[ref];
    cy = S.scy[ref];
  }
  const int32_t min_i = imax(0, cx - dd), max_i = imin(cx + dd, g.max_x - 1);
  const int32_t min_j = imax(0, cy - dd), max_j = imin(cy + dd, g.max_y - 1);
  FlockAcc acc;
  for (int32_t i = min_i; i <= max_i; ++i) {
    const int c = i - rt.sc_lo;
    // runs: first slot with cell y >= min_j
    uint32_t a = col_base[c], a_end = col_base[c + 1];
    {
      uint32_t lo = a, hi = a_end;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (S.scy[mid] < min_j) lo = mid + 1;
        else hi = mid;
      }
      a = lo;
    }
    uint32_t b = ecol_base[c], b_end = ecol_base[c + 1];
    while (b < b_end && S.ecy[S.eperm[b]] < min_j) ++b;
    int32_t cur = -1;
    int chk = 0;
    for (;;) {
      const bool ha = a < a_end && S.scy[a] <= max_j;
      const bool hb = b < b_end && S.ecy[S.eperm[b]] <= max_j;
      if (!ha && !hb) break;
      bool take_a = ha;
      if (ha && hb) {
        const uint32_t e = S.eperm[b];
        take_a = S.scy[a] < S.ecy[e] || (S.scy[a] == S.ecy[e] && S.sid[a] <= S.eid[e]);
      }
      float ex_, ey_, ehx_, ehy_;
      uint32_t eid_;
      int32_t ecy_;
      if (take_a) {
        ex_ = S.sx[a]; ey_ = S.sy[a]; ehx_ = S.shx[a]; ehy_ = S.shy[a]; eid_ = S.sid[a];
        ecy_ = S.scy[a];
        ++a;
      } else {
        const uint32_t e = S.eperm[b];
        ex_ = S.ex[e]; ey_ = S.ey[e]; ehx_ = S.ehx[e]; ehy_ = S.ehy[e]; eid_ = S.eid[e];
        ecy_ = S.ecy[e];
        ++b;
      }
      if (ecy_ != cur) {
        cur = ecy_;
        chk = check_circle_sq(g, i, ecy_, qx, qy, T);
      }
      if (chk == -1) continue;
      const float dx = toroidal_distance(qx, ex_, g.w);
      const float dy = toroidal_distance(qy, ey_, g.h);
      if (chk == 1 || dx * dx + dy * dy <= T) acc.add(eid_ == qid, dx, dy, ehx_, ehy_);
    }
  }
  flock_finish(acc, rp.p, qid, rt.step, qx, qy, qhx, qhy, g.w, nx, ny, nhx, nhy);
  *nid = qid;
}

// Sends one agent to its destination tile's inbox (or the spill list).
KG_HD void rt_push(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t dest, float x,
                   float y, float hx, float hy, uint32_t id) {
  const int pout = rt.pin ^ 1;
  const uint32_t slot = atom_add(&D.icnt[rt.ph_out][dest], 1u);
  if (slot < D.icap[dest]) {
    const Buf& B = D.inbox[pout];
    const uint32_t gi = D.ibase[dest] + slot;
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

// Classifies a new state: a stayer goes to the LDS output list (returns its
// slot), a mover is pushed at once (returns -1).  Out-of-grid agents are
// dropped and flagged.
KG_HD int64_t rt_classify(const RepDev& rp, const RtTile& rt, const RtDev& D, bool all_movers,
                          float x, float y, float hx, float hy, uint32_t id, uint32_t* nstay,
                          RtSmem S) {
  const Geom& g = rp.g;
  int32_t cx, cy;
  discretize(g, x, y, &cx, &cy);
  if (!cell_valid(g, cx, cy)) {
    atom_or(D.flags, kRtFlagRange);
    return -1;
  }
  const uint32_t dest = rt.tbase + tile_of(g, cx, cy);
  if (dest != rt.gt || all_movers) {
    rt_push(rp, rt, D, dest, x, y, hx, hy, id);
    return -1;
  }
  const uint32_t slot = atom_add(nstay, 1u);
  S.ox[slot] = x; S.oy[slot] = y; S.ohx[slot] = hx; S.ohy[slot] = hy; S.oid[slot] = id;
  S.ocy[slot] = cy;
  S.obin[slot] = (uint16_t)(((uint32_t)(cx & (g.tx - 1)) << g.tr_log2) +
                            (uint32_t)((cy & (g.ty - 1)) >> g.hb_log2));
  return slot;
}

// Places stayer s inside its bin by (cell y, id, slot) and writes it out.
// Stayers sorted past the tile's capacity go to its own inbox instead (the
// tile keeps the first scap of its sorted stayers; boff is clipped).
KG_HD void rt_write_stayer(const RepDev& rp, const RtTile& rt, const RtDev& D, uint32_t s,
                           RtSmem S) {
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
  if (pos >= D.scap[rt.gt]) {
    atom_or(D.flags, kRtFlagStayCap);
    rt_push(rp, rt, D, rt.gt, S.ox[s], S.oy[s], S.ohx[s], S.ohy[s], id);
    return;
  }
  const Buf& B = D.stay[rt.pin ^ 1];
  const uint32_t gi = D.sbase[rt.gt] + pos;
  B.x[gi] = S.ox[s]; B.y[gi] = S.oy[s]; B.hx[gi] = S.ohx[s]; B.hy[gi] = S.ohy[s]; B.id[gi] = id;
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
        const uint32_t* bo = D.boff[rt.pin] + D.tboff[tt];
        const uint32_t b = ((uint32_t)lc << g.tr_log2) + (uint32_t)((j - tty * g.ty) >> g.hb_log2);
        a = D.sbase[tt] + bo[b];
        a_end = D.sbase[tt] + bo[b + 1];
        while (a < a_end && f2i(floorf(SB.y[a] / g.d)) < j) ++a;
      }
      // extras of cell (i, j) in ascending (id, source, index) by selection
      uint64_t last = 0;  // (id + 1) << 32 | position of the last taken extra
      bool have_last = false;
      for (;;) {
        const bool ha = a < a_end && f2i(floorf(SB.y[a] / g.d)) == j;
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
    const uint32_t gi = D.sbase[rt.gt] + o;
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

}  // namespace kg
