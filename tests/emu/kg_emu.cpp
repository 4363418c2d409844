// kg_emu.cpp -- host-side driver of the device kernel phases (tests only).
//
// Executes the bodies of the count / scan / scatter / flock kernels
// (krabmaga_amd/csrc/kg_phases.hpp) on the CPU, block by block and phase by
// phase, over host arrays.  Built with AddressSanitizer so that every index
// the kernels compute is checked against its buffer, and compared with the
// oracle, before the same code runs on the GPU.  The order in which agents
// are scattered into their bins can be shuffled, to show that results do not
// depend on the in-bin order the device's atomics produce.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "kg_geom.hpp"
#include "kg_phases.hpp"
#include "kg_rt.hpp"

using namespace kg;

namespace {

struct HostBuf {
  std::vector<float> x, y, hx, hy;
  std::vector<uint32_t> id;
  void resize(size_t n) {
    x.assign(n, 0.f);
    y.assign(n, 0.f);
    hx.assign(n, 0.f);
    hy.assign(n, 0.f);
    id.assign(n, 0u);
  }
  Buf view() { return Buf{x.data(), y.data(), hx.data(), hy.data(), id.data()}; }
};

}  // namespace

struct EmuRt {
  bool active = false;
  RtPlan plan;
  int ncls = 2;
  RtCaps caps[kRtMaxClasses]{};
  std::vector<uint32_t> dense;  // class lists 1.., grouped by class
  uint32_t off[kRtMaxClasses + 1] = {};
  HostBuf stay[2], inbox[2], spill[2];
  std::vector<uint32_t> spill_rep[2];
  std::vector<uint32_t> cnt[2], boff[2], icnt[3];
  uint32_t scnt[3] = {0, 0, 0};
  uint32_t flags = 0;
  uint64_t gstep = 0;
  std::vector<uint32_t> ovf[2];
  RtDev dev() {
    RtDev D;
    D.meta = plan.meta.data();
    for (int p = 0; p < 2; ++p) {
      D.cnt[p] = cnt[p].data();
      D.boff[p] = boff[p].data();
      D.stay[p] = stay[p].view();
      D.inbox[p] = inbox[p].view();
      D.spill[p] = spill[p].view();
      D.spill_rep[p] = spill_rep[p].data();
    }
    for (int k = 0; k < 3; ++k) {
      D.icnt[k] = icnt[k].data();
      D.scnt[k] = &scnt[k];
    }
    D.spill_cap = (uint32_t)spill[0].x.size();
    D.flags = &flags;
    D.gstep = &gstep;
    for (int p = 0; p < 2; ++p) {
      ovf[p].resize(kRtOvfMax + 1, 0u);
      D.ovf[p] = ovf[p].data();
    }
    return D;
  }
};

struct emu_field {
  EmuRt rt;
  Geom g{};
  RepDev rep{};
  uint32_t cap = 0;
  int lds_cap = 1024;
  HostBuf R, R2, W, A;
  std::vector<uint32_t> count, offset, offset2, cursor;
  uint32_t err = 0;
  uint32_t nR = 0, nA = 0;
  bool w_aligned = false;
  bool force_global = false;
  uint64_t shuffle_seed = 0;  // 0 = scatter in index order
  uint64_t lds_tiles = 0, global_tiles = 0, deferred_tiles = 0, chunked_tiles = 0;
  // gather rebinning, as the device field runs it by default
  bool gather = true;
  bool outbox_filled = false;
  std::vector<uint32_t> outbox_n, outbox_slot;
  uint64_t gather_steps = 0, gather_fallbacks = 0;
  bool gather_ok() const { return gather && (uint32_t)(g.tx * g.tr) <= kGatherMaxBins; }
  // flock counts into the scan's tile sums (the LDS histogram of a tile's
  // stayers), as the device field's runs take them
  bool summed_step = false;
  std::vector<uint32_t> tsum;
  bool summed_ok() const {
    const uint32_t nbt = (uint32_t)(g.tx * g.tr);
    return gather_ok() && (g.nbins + kScanTile - 1) / kScanTile <= kScanInlineTiles &&
           (uint32_t)kScanTile % nbt == 0;
  }
};

extern "C" {

emu_field* emu_reset_geom(emu_field* f, float w, float h, float d, int tor, uint32_t cap,
                          uint32_t bin_rows, int lds_cap, int keep_geom, int);

emu_field* emu_create(float w, float h, float d, int tor, uint32_t cap, uint32_t bin_rows,
                      int lds_cap) {
  return emu_reset_geom(new emu_field(), w, h, d, tor, cap, bin_rows, lds_cap, 0, 0);
}

// Recreates the field with the tile-resident geometry (tx, ty, hb; 0 = auto).
emu_field* emu_create_rt(float w, float h, float d, int tor, uint32_t cap, uint32_t tx,
                         uint32_t ty, uint32_t hb) {
  emu_field* f = new emu_field();
  std::string why;
  if (!make_geom_rt(w, h, d, tor, cap, tx, ty, hb, &f->g, &why, true)) {
    delete f;
    return nullptr;
  }
  return emu_reset_geom(f, w, h, d, tor, cap, 0, 0, 1, 0);
}

emu_field* emu_reset_geom(emu_field* f, float w, float h, float d, int tor, uint32_t cap,
                          uint32_t bin_rows, int lds_cap, int keep_geom, int) {
  std::string why;
  if (!keep_geom && !make_geom(w, h, d, tor, cap, bin_rows, &f->g, &why)) {
    delete f;
    return nullptr;
  }
  f->cap = cap;
  f->lds_cap = lds_cap > 0 ? lds_cap : pick_lds_cap(f->g, cap);
  f->rep.g = f->g;
  kg_flock_params dflt{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0.7f, 10.0f, 0u, 0ull};
  set_rep_params(&f->rep, dflt);
  f->R.resize(cap);
  f->R2.resize(cap);
  f->W.resize(cap);
  f->A.resize(cap);
  f->count.assign(f->g.nbins, 0);
  f->offset.assign(f->g.nbins + 1, 0);
  f->offset2.assign(f->g.nbins + 1, 0);
  f->cursor.assign(f->g.nbins, 0);
  f->outbox_n.assign(f->g.ntiles, 0xFFFFFFFFu);  // written by every flock step
  f->tsum.assign((f->g.nbins + kScanTile - 1) / kScanTile, 0u);
  f->outbox_slot.assign((size_t)f->g.ntiles * kOutboxCap, 0xFFFFFFFFu);
  return f;
}

void emu_destroy(emu_field* f) { delete f; }

// ---- tile-resident path ---------------------------------------------------

static std::vector<uint32_t> order(uint32_t lo, uint32_t hi, uint64_t seed);

// Output phase of a tile, as rt_output in kg_kernels.hip: the stayers are
// classified and counted per bin already.
static void emu_rt_output(const RepDev& rp, const RtTile& rt, const RtDev& D, const RtMeta& me,
                          uint32_t bo_own, int nb_tile, uint32_t nstay, RtSmem S) {
  const int pout = rt.pin ^ 1;
  uint32_t run = 0;
  for (int b = 0; b < nb_tile; ++b) {
    S.bbase[b] = run;
    S.bcur[b] = run;
    run += S.hist[b];
  }
  S.bbase[nb_tile] = nstay;
  for (uint32_t s2 = 0; s2 < nstay; ++s2) S.operm[S.bcur[S.obin[s2]]++] = (uint16_t)s2;
  const uint32_t scap = me.scap;
  for (uint32_t s2 = 0; s2 < nstay; ++s2) rt_write_stayer(rp, rt, D, me, s2, S);
  for (int b = 0; b <= nb_tile; ++b)
    D.boff[pout][bo_own + b] = S.bbase[b] < scap ? S.bbase[b] : scap;
  D.cnt[pout][rt.gt] = nstay < scap ? nstay : scap;
}

// One work-group of k_rt_step (load = 0) or k_rt_load (load = 1) for the
// tile `item` in class `cls`, with the threads of each phase run in turn.
static void emu_rt_block(emu_field* f, uint32_t item, int cls, int load, uint64_t load_g,
                         std::vector<uint8_t>& smem, bool deferred = false) {
  const bool last = cls == f->rt.ncls - 1;
  EmuRt& E = f->rt;
  RtDev D = E.dev();
  const RtCaps caps = E.caps[cls];
  const uint32_t prefix[2] = {0u, f->g.ntiles};
  const uint64_t g = load ? load_g : E.gstep;
  const RtTile rt = rt_tile(&f->rep, 1, prefix, item, g);
  const RtMeta me = D.meta[rt.gt];
  if (!deferred && (int)me.cls != cls) return;
  const RepDev& rp = f->rep;
  smem.assign(rt_smem_bytes(caps), 0x5A);
  const RtSmem S = rt_carve(smem.data(), caps);
  const int nb_tile = rp.g.tx * rp.g.tr;
  const int pout = rt.pin ^ 1;
  const uint32_t bo_own = rt_boff_at(rp, rt, rt.gt);
  Buf Ra = f->R.view();
  if (load) {
    const uint32_t tile = rt.gt - rt.tbase;
    const uint32_t b0 = rp.bin_base + (tile << (rp.g.tx_log2 + rp.g.tr_log2));
    const uint32_t lo = f->offset[b0], hi = f->offset[b0 + (uint32_t)nb_tile];
    const uint32_t nown = hi - lo;
    if (nown > (uint32_t)caps.own || nb_tile > caps.nb) {
      for (uint32_t i = lo; i < hi; ++i) {
        int32_t cx, cy;
        discretize(rp.g, Ra.x[i], Ra.y[i], &cx, &cy);
        rt_push(rp, rt, D, rt.tbase + tile_of(rp.g, cx, cy), Ra.x[i], Ra.y[i], Ra.hx[i],
                Ra.hy[i], Ra.id[i]);
      }
      for (int b = 0; b <= nb_tile; ++b) D.boff[pout][bo_own + b] = 0u;
      D.cnt[pout][rt.gt] = 0u;
      ++f->global_tiles;
      return;
    }
    for (int b = 0; b <= nb_tile; ++b) S.hist[b] = 0u;
    uint32_t nstay = 0;
    for (uint32_t o = 0; o < nown; ++o) {
      S.rx[o] = Ra.x[lo + o];
      S.ry[o] = Ra.y[lo + o];
      S.rhx[o] = Ra.hx[lo + o];
      S.rhy[o] = Ra.hy[lo + o];
      S.rid[o] = Ra.id[lo + o];
      rt_classify(rp, rt, D, o, &nstay, S);
    }
    emu_rt_output(rp, rt, D, me, bo_own, nb_tile, nstay, S);
    return;
  }
  uint32_t seg_start[3 * kMaxStageCols], seg_len[3 * kMaxStageCols];
  uint32_t col_base[kMaxStageCols + 1];
  uint32_t src_base[10], src_pre[11];
  uint32_t ne = 0, nown = 0;
  bool fail = false;
  const RtSegs sg = rt_segs(rp.g, rt);
  const bool lds_ok = rp.lds_ok != 0 && rt_fits_tables(rt, sg, caps) && nb_tile <= caps.nb;
  if (lds_ok)
    for (int b = 0; b <= nb_tile; ++b) S.hist[b] = 0u;
  uint32_t total = 0;
  for (int t = 0; t < rt.ncols && lds_ok; ++t) {
    col_base[t] = total;
    total += rt_segments(rp, rt, D, t, seg_start, seg_len);
  }
  col_base[rt.ncols] = total;
  src_pre[0] = 0;
  for (int k = 0; k < 10; ++k) src_pre[k + 1] = src_pre[k] + rt_extra_count(rp, rt, D, k, &src_base[k]);
  bool global_path = !lds_ok || total > (uint32_t)caps.runs;
  if (!global_path) {
    const uint32_t next = src_pre[10];
    const uint32_t kT = (uint32_t)caps.threads;
    for (uint32_t t = 0; t < kT; ++t)
      if (!rt_fetch(rp, rt, D, t, t + kT, total, t, next, seg_start, seg_len, col_base,
                    src_base, src_pre, caps, &ne, &nown, S))
        fail = true;
    for (uint32_t k = 2 * kT; k < total; ++k)
      rt_stage(rp, rt, D, k, seg_start, seg_len, col_base, caps.own, &nown, S);
    for (uint32_t j = kT; j < next; ++j)
      if (!rt_gather_extra(rp, rt, D, j, src_base, src_pre, caps.extras, &ne, caps.own, &nown, S))
        fail = true;
    for (uint32_t q = 0; q < (uint32_t)(rt.ncols * sg.nbc); ++q)
      rt_bintab(rp, rt, D, sg, q, seg_len, S);
    global_path = fail || nown > (uint32_t)caps.own;
  }
  if (global_path && !last && lds_ok) {
    const int par = (int)(g & 1u);
    const uint32_t slot = D.ovf[par][0]++;
    *D.flags |= kRtFlagDefer;
    if (slot < kRtOvfMax) {
      D.ovf[par][1 + slot] = rt.gt;
      ++f->deferred_tiles;
      return;
    }
  }
  if (!global_path) {
    ++f->lds_tiles;
    for (uint32_t j = 0; j < ne; ++j) rt_rank_extra(j, ne, S);
    for (int c = 0; c <= rt.ncols; ++c) rt_ecbase(c, ne, S);
    uint32_t nstay = 0;
    for (uint32_t o = 0; o < nown; ++o) {
      rt_agent_lds(rp, rt, sg, o, col_base, S);
      rt_classify(rp, rt, D, o, &nstay, S);
    }
    // the staging arrays are dead from here on
    const size_t res = (rt_smem_res(caps) + 15) & ~(size_t)15;
    std::fill(smem.begin() + (ptrdiff_t)res, smem.end(), 0x3C);
    emu_rt_output(rp, rt, D, me, bo_own, nb_tile, nstay, S);
  } else {
    ++f->global_tiles;
    *D.flags |= kRtFlagGlobal;
    rt_tile_global(rp, rt, D, 0, 1);
    for (int b = 0; b <= nb_tile; ++b) D.boff[pout][bo_own + b] = 0u;
    D.cnt[pout][rt.gt] = 0u;
  }
  D.icnt[rt.ph_reset][rt.gt] = 0u;
  if (rt.gt == 0) *D.scnt[rt.ph_reset] = 0u;
}

// Enters the tile-resident layout from the current read buffer.  caps_* > 0
// override the planned capacities of both classes; `mode` bit 0 makes the
// per-tile capacities far smaller than the counts (stay overflow, inbox
// overflow, spill), bit 1 gives the normal class a tiny LDS budget so that
// most tiles run in the dense class.
int emu_rt_enter(emu_field* f, int caps_runs, int caps_own, int caps_extras, int mode) {
  if (f->w_aligned || f->nA) return -6;
  EmuRt& E = f->rt;
  const uint32_t nt = f->g.ntiles;
  const uint32_t nbt = (uint32_t)(f->g.tx * f->g.tr);
  std::vector<uint32_t> counts(nt), nb(nt, nbt);
  for (uint32_t t = 0; t < nt; ++t) counts[t] = f->offset[(t + 1) * nbt] - f->offset[t * nbt];
  std::vector<double> expect;
  uint64_t total = 0;
  for (uint32_t c : counts) total += c;
  rt_expected_counts(f->g, total, &expect);
  rt_plan(counts, nb, expect, &E.plan);
  if (mode & 1) {
    // capacities far below the counts: exercises stay-capacity overflow,
    // inbox overflow and the spill list
    uint64_t s2 = 0, ib = 0;
    for (uint32_t t = 0; t < nt; ++t) {
      E.plan.scap[t] = std::max<uint32_t>(4u, counts[t] / 2u);
      E.plan.icap[t] = 4u;
      E.plan.sbase[t] = (uint32_t)s2;
      E.plan.ibase[t] = (uint32_t)ib;
      E.plan.meta[t] = RtMeta{E.plan.sbase[t], E.plan.scap[t], E.plan.ibase[t], E.plan.icap[t], 0u,
                              {0u, 0u, 0u}};
      s2 += E.plan.scap[t];
      ib += E.plan.icap[t];
    }
    E.plan.stay_total = s2;
    E.plan.inbox_total = ib;
  }
  RtClassPlan cls;
  rt_classes(&E.plan, counts, {RtRepInfo{f->g, f->rep.disc_dist, 0u}}, (mode & 2) ? 2048 : 0,
             157696, &cls);
  E.ncls = cls.ncls;
  for (int k = 0; k < kRtMaxClasses; ++k) E.caps[k] = cls.caps[k];
  for (int k = 0; k <= kRtMaxClasses; ++k) E.off[k] = cls.off[k];
  E.dense = cls.dense;
  for (int c = 0; c < ((mode & 4) ? 1 : E.ncls); ++c) {
    if (caps_runs > 0) E.caps[c].runs = caps_runs;
    if (caps_own > 0) E.caps[c].own = caps_own;
    if (caps_extras > 0) E.caps[c].extras = caps_extras;
  }
  for (int p = 0; p < 2; ++p) {
    E.stay[p].resize(E.plan.stay_total);
    E.inbox[p].resize(E.plan.inbox_total);
    E.spill[p].resize(f->cap);
    E.spill_rep[p].assign(f->cap, 0u);
    E.cnt[p].assign(nt, 0u);
    E.boff[p].assign(E.plan.boff_total, 0u);
  }
  for (int k = 0; k < 3; ++k) {
    E.icnt[k].assign(nt, 0u);
    E.scnt[k] = 0;
  }
  E.flags = 0;
  E.gstep = 1;
  for (int p = 0; p < 2; ++p) E.ovf[p].assign(kRtOvfMax + 1, 0u);
  std::vector<uint8_t> smem;
  for (int c = 0; c < E.ncls; ++c)
    for (uint32_t item : order(0, nt, f->shuffle_seed))
      emu_rt_block(f, item, c, 1, E.gstep - 1, smem);
  E.active = true;
  return 0;
}

int emu_rt_steps(emu_field* f, const kg_flock_params* p, uint64_t step0, uint32_t nsteps) {
  EmuRt& E = f->rt;
  if (!E.active) return -6;
  set_rep_params(&f->rep, *p);
  f->rep.step_base = step0 - E.gstep;
  std::vector<uint8_t> smem;
  for (uint32_t s2 = 0; s2 < nsteps; ++s2) {
    // the dense-class launch runs after the normal one; tiles are
    // independent within a step, so any order is valid
    const uint64_t sh = f->shuffle_seed ? f->shuffle_seed + s2 : 0;
    for (uint32_t item : order(0, f->g.ntiles, sh)) emu_rt_block(f, item, 0, 0, 0, smem);
    for (int c = 1; c < E.ncls; ++c)
      for (uint32_t k : order(E.off[c], E.off[c + 1], sh)) emu_rt_block(f, E.dense[k], c, 0, 0, smem);
    const int par = (int)(E.gstep & 1u);
    const uint32_t nd = std::min<uint32_t>(E.ovf[par][0], kRtOvfMax);
    for (uint32_t i = 0; i < nd; ++i)
      emu_rt_block(f, E.ovf[par][1 + i], E.ncls - 1, 0, 0, smem, true);
    E.ovf[par ^ 1][0] = 0;
    E.gstep += 1;
  }
  return (E.flags & kRtFlagRange) ? -3 : 0;
}

uint32_t emu_rt_flags(emu_field* f) { return f->rt.flags; }

// Work-groups run on the LDS path, on the global path, and deferred.
void emu_rt_counts(emu_field* f, uint64_t* lds, uint64_t* glob, uint64_t* deferred,
                   uint32_t* ndense, int* ncls) {
  *lds = f->lds_tiles;
  *glob = f->global_tiles;
  *deferred = f->deferred_tiles;
  *ndense = (uint32_t)f->rt.dense.size();
  *ncls = f->rt.ncls;
}

// All agents of the tile-resident layout (any order).
uint32_t emu_rt_read(emu_field* f, uint32_t* ids, float* x, float* y, float* hx, float* hy) {
  EmuRt& E = f->rt;
  RtDev D = E.dev();
  const uint32_t prefix[2] = {0u, f->g.ntiles};
  uint32_t n = 0;
  auto put = [&](const Buf& B, uint32_t gi) {
    ids[n] = B.id[gi];
    x[n] = B.x[gi];
    y[n] = B.y[gi];
    hx[n] = B.hx[gi];
    hy[n] = B.hy[gi];
    ++n;
  };
  for (uint32_t item = 0; item < f->g.ntiles; ++item) {
    const RtTile rt = rt_tile(&f->rep, 1, prefix, item, E.gstep);
    for (uint32_t o = 0; o < D.cnt[rt.pin][rt.gt]; ++o) put(D.stay[rt.pin], D.meta[rt.gt].sbase + o);
    uint32_t ib;
    const uint32_t ni = rt_extra_count(f->rep, rt, D, 4, &ib);
    for (uint32_t o = 0; o < ni; ++o) put(D.inbox[rt.pin], ib + o);
    if (item == 0) {
      uint32_t sb;
      const uint32_t nsp = rt_extra_count(f->rep, rt, D, 9, &sb);
      for (uint32_t o = 0; o < nsp; ++o) put(D.spill[rt.pin], o);
    }
  }
  return n;
}

// geom: max_x, max_y, dw, dh, tx, tr, hb, ty, ntx, nty, ntiles, nbins, lds_cap
void emu_geom(emu_field* f, int64_t* out) {
  const Geom& g = f->g;
  const int64_t v[] = {g.max_x, g.max_y, g.dw, g.dh, g.tx, g.tr, g.hb, g.ty,
                       g.ntx,   g.nty,   g.ntiles, g.nbins, f->lds_cap};
  std::memcpy(out, v, sizeof(v));
}

void emu_options(emu_field* f, int force_global, uint64_t shuffle_seed) {
  f->force_global = force_global != 0;
  f->shuffle_seed = shuffle_seed;
}

// gather rebinning on or off (on by default, as on the device)
void emu_set_gather(emu_field* f, int on) { f->gather = on != 0; }

void emu_gather_stats(emu_field* f, uint64_t* steps, uint64_t* fallbacks) {
  *steps = f->gather_steps;
  *fallbacks = f->gather_fallbacks;
}

void emu_stats(emu_field* f, uint64_t* lds_tiles, uint64_t* global_tiles,
               uint64_t* chunked_tiles) {
  *lds_tiles = f->lds_tiles;
  *global_tiles = f->global_tiles;
  *chunked_tiles = f->chunked_tiles;
}

int emu_set(emu_field* f, const uint32_t* ids, const float* x, const float* y, const float* hx,
            const float* hy, uint32_t n) {
  if (f->w_aligned || f->nA + n > f->cap) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    f->A.x[f->nA + i] = x[i];
    f->A.y[f->nA + i] = y[i];
    f->A.hx[f->nA + i] = hx[i];
    f->A.hy[f->nA + i] = hy[i];
    f->A.id[f->nA + i] = ids[i];
  }
  f->nA += n;
  return 0;
}

static std::vector<uint32_t> order(uint32_t lo, uint32_t hi, uint64_t seed) {
  std::vector<uint32_t> o(hi - lo);
  std::iota(o.begin(), o.end(), lo);
  if (seed) {
    std::mt19937_64 rng(seed + lo * 7919ull);
    std::shuffle(o.begin(), o.end(), rng);
  }
  return o;
}

// k_rebin_gather for work item `item`, phase by phase (candidates taken in
// a shuffled order, as the device's atomics may take them); true if the
// tile could not rebin (the gate word).
static bool emu_gather_tile(emu_field* f, uint32_t item, Buf R, Buf W, Buf R2) {
  const Geom& g = f->rep.g;
  const uint32_t tile = item;  // one replica
  const uint32_t nbt = 1u << (g.tx_log2 + g.tr_log2);
  const uint32_t tb0 = tile * nbt;
  const uint32_t lo = f->offset[tb0], nown = f->offset[tb0 + nbt] - lo;
  const uint32_t nlo = f->offset2[tb0];
  std::vector<uint32_t> bstart(kGatherMaxBins + 1), fill(kGatherMaxBins, 0u);
  std::vector<uint64_t> key(kGatherCap);
  uint32_t sitem[kGatherSources], scnt[kGatherSources];
  for (uint32_t b = 0; b <= nbt; ++b) bstart[b] = f->offset2[tb0 + b] - nlo;
  for (uint32_t q = 0; q < kGatherSources; ++q) {
    const int32_t sl = gather_source(g, tile, q);
    sitem[q] = sl >= 0 ? (uint32_t)sl : 0u;
    scnt[q] = sl >= 0 ? f->outbox_n[(uint32_t)sl] : 0u;
  }
  uint32_t fail = 0, ncand = nown;
  for (uint32_t q = 0; q < kGatherSources; ++q) {
    ncand += scnt[q];
    if (scnt[q] > kOutboxCap) fail = 1u;
  }
  if (ncand > kGatherCap || bstart[nbt] > kGatherCap) fail = 1u;
  for (uint32_t b = 0; b < nbt; ++b)
    if (bstart[b + 1] - bstart[b] > kGatherBinMax) fail = 1u;
  if (fail) return true;
  // each candidate stays with its thread from load to store
  struct Held {
    uint32_t b = kGatherMaxBins, slot = 0;
    float x = 0, y = 0, hx = 0, hy = 0;
    uint32_t id = 0;
  };
  const Outbox ob{f->outbox_n.data(), f->outbox_slot.data()};
  std::vector<Held> held(ncand);
  for (uint32_t k : order(0, ncand, f->shuffle_seed)) {
    const uint32_t src = gather_src(ob, sitem, scnt, lo, nown, k);
    int32_t row;
    const uint32_t b = gather_bin(g, tile, W.x[src], W.y[src], &row);
    if (b >= nbt) continue;
    const uint32_t slot = fill[b]++;
    if (slot >= bstart[b + 1] - bstart[b]) {
      fail = 1u;
      continue;
    }
    Held& h = held[k];
    h.b = b;
    h.slot = slot;
    h.x = W.x[src];
    h.y = W.y[src];
    h.hx = W.hx[src];
    h.hy = W.hy[src];
    h.id = R.id[src];
    key[bstart[b] + slot] = ((uint64_t)(uint32_t)row << 32) | h.id;
  }
  for (uint32_t b = 0; b < nbt; ++b)
    if (fill[b] != bstart[b + 1] - bstart[b]) fail = 1u;
  if (fail) return true;
  for (const Held& h : held) {
    if (h.b >= nbt) continue;
    const uint32_t e = bstart[h.b] + h.slot;
    const uint32_t pos =
        nlo + bstart[h.b] + gather_rank(key.data(), bstart[h.b], bstart[h.b + 1], e, key[e]);
    R2.x[pos] = h.x;
    R2.y[pos] = h.y;
    R2.hx[pos] = h.hx;
    R2.hy[pos] = h.hy;
    R2.id[pos] = h.id;
  }
  return false;
}

int emu_lazy_update(emu_field* f) {
  const uint32_t prefix[2] = {0u, f->g.ntiles};
  const uint32_t app_prefix[2] = {0u, f->nA};
  const bool gather = f->w_aligned && f->outbox_filled && f->nA == 0;
  f->outbox_filled = false;
  if (f->summed_step) {
    // the tile sums the flock step left equal the scan's own, and each
    // tile's first new offset as the gather finds it (the earlier scan
    // tiles' sums plus the earlier bins of its own scan tile) equals the
    // exclusive scan's
    const uint32_t nbt = 1u << (f->g.tx_log2 + f->g.tr_log2);
    std::vector<uint32_t> ex(f->g.nbins + 1, 0u);
    for (uint32_t b = 0; b < f->g.nbins; ++b) ex[b + 1] = ex[b] + f->count[b];
    for (uint32_t q = 0; q < f->tsum.size(); ++q) {
      uint32_t sum = 0;
      for (uint32_t b = q * kScanTile; b < std::min<uint32_t>(f->g.nbins, (q + 1) * kScanTile); ++b)
        sum += f->count[b];
      if (sum != f->tsum[q]) return -8;
    }
    for (uint32_t tile = 0; tile < f->g.ntiles; ++tile) {
      const uint32_t tb0 = tile * nbt, q = tb0 / kScanTile;
      uint32_t nlo = 0;
      for (uint32_t i = 0; i < q; ++i) nlo += f->tsum[i];
      for (uint32_t b = q * kScanTile; b < tb0; ++b) nlo += f->count[b];
      if (nlo != ex[tb0]) return -9;
    }
    std::fill(f->tsum.begin(), f->tsum.end(), 0u);
    f->summed_step = false;
  }
  Buf A = f->A.view();
  for (uint32_t i = 0; i < f->nA; ++i)
    count_append_one(&f->rep, 1, app_prefix, A, i, f->count.data(), &f->err);
  // exclusive scan (the k_scan contract: offsets, cursors, counts zeroed)
  uint32_t run = 0;
  for (uint32_t b = 0; b < f->g.nbins; ++b) {
    f->offset2[b] = run;
    f->cursor[b] = run;
    run += f->count[b];
    f->count[b] = 0;
  }
  f->offset2[f->g.nbins] = run;
  Buf R = f->R.view(), W = f->W.view(), R2 = f->R2.view();
  bool redo = !gather;
  if (gather) {
    ++f->gather_steps;
    for (uint32_t item : order(0, f->g.ntiles, f->shuffle_seed))
      if (emu_gather_tile(f, item, R, W, R2)) redo = true;
    if (redo) ++f->gather_fallbacks;
  }
  if (f->w_aligned && redo) {
    std::vector<uint32_t> items = order(0, f->g.ntiles, f->shuffle_seed);
    for (uint32_t item : items) {
      const TileRange tr = tile_range(&f->rep, 1, prefix, item, f->offset.data());
      for (uint32_t i : order(tr.lo, tr.hi, f->shuffle_seed))
        scatter_aligned_one(f->rep, R, W, i, f->cursor.data(), R2);
    }
  }
  for (uint32_t i : order(0, f->nA, f->shuffle_seed))
    scatter_append_one(&f->rep, 1, app_prefix, A, i, f->cursor.data(), R2);
  // k_sort_bins: each bin into (cell y, id) order, ties by position
  std::vector<uint32_t> idx;
  std::vector<uint64_t> key;
  for (uint32_t b = 0; b < f->g.nbins && redo; ++b) {
    const uint32_t s = f->offset2[b], m = f->offset2[b + 1] - s;
    if (m <= 1) continue;
    if (m <= kSortSmall) {  // one lane on the device
      sort_small_bin(f->g, R2, s, m);
      continue;
    }
    idx.resize(m);
    key.resize(m);
    for (uint32_t j = 0; j < m; ++j) {
      idx[j] = j;
      key[j] = bin_key(f->g, R2.y[s + j], R2.id[s + j]);
    }
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t c) { return key[a] < key[c]; });
    Buf T = W;  // the write buffer is consumed: scratch, as on the device
    for (uint32_t j = 0; j < m; ++j) {
      T.id[s + j] = R2.id[s + idx[j]];
      T.x[s + j] = R2.x[s + idx[j]];
      T.y[s + j] = R2.y[s + idx[j]];
      T.hx[s + j] = R2.hx[s + idx[j]];
      T.hy[s + j] = R2.hy[s + idx[j]];
    }
    for (uint32_t j = 0; j < m; ++j) {
      R2.id[s + j] = T.id[s + j];
      R2.x[s + j] = T.x[s + j];
      R2.y[s + j] = T.y[s + j];
      R2.hx[s + j] = T.hx[s + j];
      R2.hy[s + j] = T.hy[s + j];
    }
  }
  std::swap(f->R, f->R2);
  std::swap(f->offset, f->offset2);
  f->nR = run;
  f->nA = 0;
  f->w_aligned = false;
  return f->err ? -3 : 0;
}

int emu_flock_step(emu_field* f, const kg_flock_params* p, uint64_t step) {
  if (f->nA || f->w_aligned) return -6;
  set_rep_params(&f->rep, *p);
  f->rep.step_base = step;
  const uint32_t prefix[2] = {0u, f->g.ntiles};
  Buf R = f->R.view(), W = f->W.view();
  std::vector<uint32_t> seg_start(3 * kMaxStageCols), seg_len(3 * kMaxStageCols),
      col_base(kMaxStageCols + 1);
  std::vector<uint8_t> smem(flock_smem_bytes(f->lds_cap));
  std::vector<uint8_t> gflag(f->g.ntiles, 0);
  const bool ob_on = f->gather_ok();
  f->outbox_filled = ob_on;
  const bool summed = f->summed_ok();
  f->summed_step = summed;
  uint32_t* const tsum = summed ? f->tsum.data() : nullptr;
  std::vector<uint32_t> hist(kGatherMaxBins);
  for (uint32_t item = 0; item < f->g.ntiles; ++item) {
    const FlockTile ft = flock_tile(&f->rep, 1, prefix, item, 0);
    const RepDev& rp = f->rep;
    bool use_lds = !f->force_global && rp.lds_ok != 0 && ft.ncols <= kMaxStageCols;
    uint32_t L = 0;
    if (use_lds) {
      uint32_t run = 0;
      for (int t = 0; t < ft.ncols; ++t) {
        const uint32_t len =
            flock_segments(rp, ft, t, f->offset.data(), seg_start.data(), seg_len.data());
        col_base[t] = run;
        run += len;
      }
      col_base[ft.ncols] = run;
      L = run;
    }
    const int32_t own_lo = ft.c0 - ft.sc_lo, own_hi = ft.c1 - ft.sc_lo;
    const uint32_t cap = (uint32_t)f->lds_cap;
    if (use_lds && L > cap) {
      use_lds = L <= kFlockChunkFactor * cap;
      for (int32_t c = own_lo; use_lds && c < own_hi; ++c)
        if (flock_chunk_size(rp, ft, col_base.data(), c, c + 1) > cap) use_lds = false;
    }
    if (use_lds) {
      ++f->lds_tiles;
      if (L > cap) ++f->chunked_tiles;
      const FlockSmem S = flock_carve(smem.data(), f->lds_cap);
      uint32_t outn = 0;  // the tile's leavers (an LDS counter on the device)
      std::fill(hist.begin(), hist.end(), 0u);
      const FlockSinks lv{ob_on ? Outbox{f->outbox_n.data(), f->outbox_slot.data()} : Outbox{}, 0u,
                          &outn, summed ? hist.data() : nullptr, tsum};
      for (int32_t oa = own_lo; oa < own_hi;) {
        const int32_t ob =
            L <= cap ? own_hi : flock_chunk_end(rp, ft, col_base.data(), oa, own_hi, cap);
        int32_t sa, sb;
        flock_chunk_cols(rp, ft, oa, ob, &sa, &sb);
        const uint32_t base = col_base[sa], n = col_base[sb] - base;
        if (n > cap) return -7;  // a chunk past the capacity: a planning error
        std::fill(smem.begin(), smem.end(), 0xA5);
        uint32_t nown = 0;
        for (uint32_t k = 0; k < n; ++k) {
          flock_stage(rp, ft, base + k, k, R, seg_start.data(), seg_len.data(), col_base.data(), S);
          flock_own(ft, k, oa, ob, S, &nown);
        }
        for (uint32_t o = 0; o < nown; ++o)
          flock_agent_lds(rp, ft, S.own[o], col_base.data(), base, S, W, f->count.data(),
                          &f->err, lv);
        oa = ob;
      }
      if (ob_on) f->outbox_n[item] = outn;
      if (summed)
        for (uint32_t b = 0; b < ft.tb1 - ft.tb0; ++b) {
          f->count[ft.tb0 + b] += hist[b];
          tsum[ft.tb0 / kScanTile] += hist[b];
        }
    } else {
      ++f->global_tiles;
      gflag[item] = 1;
      if (ob_on) f->outbox_n[item] = 0u;
    }
  }
  // k_flock_global: the flagged tiles' birds, one per thread, in any order
  for (uint32_t gi : order(0, f->nR, f->shuffle_seed)) {
    int32_t cx, cy;
    discretize(f->rep.g, R.x[gi], R.y[gi], &cx, &cy);
    if (!cell_valid(f->rep.g, cx, cy) || !gflag[tile_of(f->rep.g, cx, cy)]) continue;
    const uint32_t tile = tile_of(f->rep.g, cx, cy);
    flock_agent_global(
        f->rep, f->rep.step_base, gi, R, f->offset.data(), W, f->count.data(), &f->err,
        FlockSinks{ob_on ? Outbox{f->outbox_n.data(), f->outbox_slot.data()} : Outbox{}, 0u,
                   nullptr, nullptr, tsum});
  }
  f->w_aligned = true;
  return f->err ? -3 : 0;
}

// Copies the read buffer (storage order).
uint32_t emu_read(emu_field* f, uint32_t* ids, float* x, float* y, float* hx, float* hy) {
  for (uint32_t i = 0; i < f->nR; ++i) {
    ids[i] = f->R.id[i];
    x[i] = f->R.x[i];
    y[i] = f->R.y[i];
    hx[i] = f->R.hx[i];
    hy[i] = f->R.hy[i];
  }
  return f->nR;
}

// get_neighbors_within_distance / relax on the read buffer (query_global).
uint32_t emu_query(emu_field* f, float x, float y, float dist, int relax, uint32_t* out,
                   uint32_t cap) {
  uint32_t n = 0;
  Buf R = f->R.view();
  query_global(f->g, R, f->offset.data(), x, y, dist, sq_threshold(dist), relax != 0,
               [&](uint32_t e) {
                 if (n < cap) out[n] = R.id[e];
                 ++n;
               });
  return n;
}

float emu_sq_threshold(float d) { return sq_threshold(d); }

}  // extern "C"
