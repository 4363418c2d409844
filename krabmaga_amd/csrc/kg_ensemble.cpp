// kg_ensemble.cpp -- many flockers runs in one tile-resident layout.
//
// explore_parallel! (src/explore/model_exploration.rs:354-432) gives every
// run its own State and simulates the runs side by side on CPU cores.  Here
// the runs are replicas of one device layout: each run keeps its own
// geometry (Field2D::new), flock parameters, Philox key and schedule step in
// a replica descriptor, its tiles are a contiguous range of the global tile
// list, and one launch per schedule step (per tile class) advances all of
// them.  Nothing in a run's arithmetic depends on the other runs, so its
// results are the bits it would produce alone.
//
// Membership changes (admit, retire) and capacity overflows rebuild the
// layout: the surviving runs are unloaded into their slots of an append
// buffer, new runs are initialised into theirs, everything is binned
// (count / scan / scatter) and loaded into a fresh tile-resident plan.
//
// The layout has two forms with identical results, as a single field has:
// the tile-resident plan above (one k_rt_step launch per tile class and
// step), and the bin-grouped read buffer itself, stepped by the
// three-kernel path over every replica at once (one k_flock launch over all
// tiles, one scan over all bins, one scatter and one bin sort).  By default
// ensembles whose runs average 32,768 agents or more take the bin-grouped
// form; KG_RUN_BINNED and KG_RUN_TILE_RESIDENT choose either.
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "kg_common.hpp"
#include "kg_rt_state.hpp"

using namespace kg;

namespace {

struct EnsRun {
  uint32_t id = 0;
  kg_run_desc desc{};
  Geom g{};
  uint64_t steps_done = 0;  // schedule steps completed
  bool fresh = true;        // initial state not generated yet
  bool dead = false;        // retired; dropped at the next build
  uint32_t app_base = 0;    // slot in the append buffer of the current layout
  uint32_t tile0 = 0, ntiles = 0;
};

constexpr uint32_t kRtCheckEvery = 64;

size_t device_lds_limit() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      v > 8192)
    return (size_t)v - 6144;  // static LDS of the tile-resident kernel
  return 65536 - 6144;
}

bool run_geom(const kg_run_desc& d, const kg_tuning& tu, Geom* g, RepDev* rep, std::string* why) {
  if (d.toroidal != 1) {
    *why = "ensemble runs must be toroidal";
    return false;
  }
  if (d.agents == 0 || d.agents > (1u << 30)) {
    *why = "agents must be in [1, 2^30]";
    return false;
  }
  if (!params_valid(d.params, why)) return false;
  if (!make_geom_rt(d.width, d.height, d.discretization, 1, d.agents, tu.tile_cols, tu.tile_rows,
                    tu.bin_rows, g, why))
    return false;
  RepDev r{};
  r.g = *g;
  set_rep_params(&r, d.params);
  if (!r.lds_ok) {
    *why = "the neighbour window is wider than a tile of this geometry";
    return false;
  }
  if (rep) *rep = r;
  return true;
}

}  // namespace

struct kg_ensemble {
  hipStream_t stream = nullptr;
  kg_tuning tuning{};
  size_t lds_limit = 0;
  std::vector<EnsRun> runs;  // layout order
  uint32_t next_id = 0;
  bool dirty = true;  // membership changed since the layout was built
  bool built = false;

  // device memory, grown on demand
  DevBuf A, A2, R;  // unload target, build input, binned read buffer
  size_t cap_a = 0, cap_a2 = 0, cap_r = 0;
  uint32_t *count = nullptr, *offset = nullptr, *cursor = nullptr;
  ScanState ss{};
  size_t cap_bins = 0, cap_scan = 0;
  RepDev* d_reps = nullptr;
  uint32_t* d_prefix = nullptr;  // tile prefix (nreps + 1), then append prefix (nreps + 1)
  uint32_t* rep_cursor = nullptr;
  size_t cap_reps = 0;
  uint32_t* tcount = nullptr;
  size_t cap_tiles = 0;
  uint32_t* err = nullptr;

  // bin-grouped form: the second read buffer, the write buffer, the next
  // bin offsets, the per-tile global-path flags and the step counter
  bool binned = false;  // the layout's form
  int want = -1;        // requested form: -1 by run size, 0 tile-resident, 1 bin-grouped
  DevBuf R2, W;
  size_t cap_r2 = 0, cap_w = 0;
  uint32_t* offset2 = nullptr;  // with the bin arrays (cap_bins + 1)
  uint32_t* gflag = nullptr;
  size_t cap_gflag = 0;
  uint64_t* gstep = nullptr;
  int lds_cap = 1024;
  int parity = 0;  // bin-grouped steps taken since the build, mod 2
  // two-step graphs by parity + 2 * gather
  std::vector<std::pair<int, hipGraphExec_t>> bgraphs;
  // gather rebinning (every run's tiles of at most kGatherMaxBins bins): the
  // lists of leavers, the gate words (a flag per step parity, steps redone,
  // the fallback's grid-wait counter) and the redone
  // count at the last check; a layout whose steps fail it drops it
  bool gather = false;
  Outbox outbox{};
  size_t cap_outbox = 0;
  uint32_t* gate = nullptr;
  uint32_t redone_seen = 0;
  // gathered steps whose counts go straight into the scan tiles' sums (every
  // run's bins start on a scan tile): the flock step counts into
  // counts(parity) and tsum[parity], one block turns the sums into
  // prefixes and the gather does the rest of the scan, clearing the other
  // parity's counts, sums and gate flag for the next step
  bool summed = false;
  uint32_t* count2 = nullptr;      // with the bin arrays (cap_bins)
  uint32_t* tsum[2] = {nullptr, nullptr};  // with the bin arrays (one per scan tile)
  uint32_t* counts(int p) const { return p ? count2 : count; }

  // current layout
  std::vector<RepDev> h_reps;
  std::vector<uint32_t> h_prefix;
  uint32_t nreps = 0, ntiles = 0, nbins = 0, nagents = 0;
  uint64_t dead_tiles = 0;  // tiles of retired runs still in the layout
  RtState rt;
  bool replan = false;
  uint64_t builds = 0, flags_seen = 0, steps_run = 0;

  WorkList wl() const {
    WorkList w;
    w.reps = d_reps;
    w.nreps = (int)nreps;
    w.tile_prefix = d_prefix;
    w.ntiles = ntiles;
    w.app_prefix = d_prefix + nreps + 1;
    w.napp = nagents;
    return w;
  }

  size_t bytes_held() const {
    return (cap_a + cap_a2 + cap_r + cap_r2 + cap_w) * 20 + cap_bins * 20 +
           (cap_bins / kScanTile + 2) * 12 + cap_gflag * 4 +
           cap_outbox * (kOutboxCap + 1) * 4 +
           rt.stay_alloc * 40 + rt.inbox_alloc * 40 + rt.spill_alloc * 44 +
           rt.boff_alloc * 8 + cap_tiles * 48;
  }

  void drop_bgraphs() {
    for (auto& g : bgraphs) (void)hipGraphExecDestroy(g.second);
    bgraphs.clear();
  }
};

namespace {

// Every run's bins start on a scan tile, so the scan tiles' sums that the
// flock step leaves are each one run's.
uint64_t scan_tile_align(uint64_t bins) {
  return (bins + kScanTile - 1) / kScanTile * kScanTile;
}

template <class T>
hipError_t regrow(T** p, size_t* cap, size_t want) {
  if (want <= *cap) return hipSuccess;
  (void)hipFree(*p);
  *p = nullptr;
  const size_t n = std::max(want, *cap + *cap / 2);
  hipError_t e = dmalloc(p, n);
  if (e == hipSuccess) *cap = n;
  else *cap = 0;
  return e;
}

// Grows one agent buffer (its contents are not kept).
int reserve_buf(kg_ensemble* e, DevBuf* b, size_t* cap, size_t want) {
  if (want <= *cap) return KG_OK;
  KG_HIP(hipStreamSynchronize(e->stream));
  const size_t n = std::max(want, *cap + *cap / 2);
  b->release();
  *cap = 0;
  const hipError_t err = b->alloc(n);
  if (err != hipSuccess)
    return set_error(KG_E_NOMEM, "ensemble agent buffer: %s", hipGetErrorString(err));
  *cap = n;
  return KG_OK;
}

// Grows the bin arrays, replica tables and per-tile counts (contents not
// kept).
int reserve(kg_ensemble* e, size_t bins, size_t reps, size_t tiles) {
  hipError_t err = hipSuccess;
  if (bins > e->cap_bins) {
    KG_HIP(hipStreamSynchronize(e->stream));
    const size_t n = std::max(bins, e->cap_bins + e->cap_bins / 2);
    for (uint32_t* p : {e->count, e->count2, e->offset, e->offset2, e->cursor, e->ss.partial,
                        e->tsum[0], e->tsum[1]})
      (void)hipFree(p);
    e->count = e->count2 = e->offset = e->offset2 = e->cursor = e->ss.partial = nullptr;
    e->tsum[0] = e->tsum[1] = nullptr;
    e->cap_bins = 0;
    const size_t nscan = (n + kScanTile - 1) / kScanTile + 1;
    // offset and offset2 have the same size: the bin-grouped steps swap them
    if ((err = dmalloc(&e->count, n)) != hipSuccess || (err = dmalloc(&e->offset, n + 1)) != hipSuccess ||
        (err = dmalloc(&e->offset2, n + 1)) != hipSuccess ||
        (err = dmalloc(&e->cursor, n)) != hipSuccess ||
        (err = dmalloc(&e->ss.partial, nscan)) != hipSuccess ||
        (err = dmalloc(&e->count2, n)) != hipSuccess ||
        (err = dmalloc(&e->tsum[0], nscan)) != hipSuccess ||
        (err = dmalloc(&e->tsum[1], nscan)) != hipSuccess)
      return set_error(KG_E_NOMEM, "ensemble bin arrays: %s", hipGetErrorString(err));
    // the scan leaves the counts at zero, and a summed run leaves both
    // arrays there; a fresh array starts there
    KG_HIP(hipMemsetAsync(e->count, 0, n * sizeof(uint32_t), e->stream));
    KG_HIP(hipMemsetAsync(e->count2, 0, n * sizeof(uint32_t), e->stream));
    e->cap_bins = n;
  }
  if (reps > e->cap_reps) {
    KG_HIP(hipStreamSynchronize(e->stream));
    const size_t n = std::max(reps, e->cap_reps + e->cap_reps / 2);
    (void)hipFree(e->d_reps);
    (void)hipFree(e->d_prefix);
    (void)hipFree(e->rep_cursor);
    e->d_reps = nullptr;
    e->d_prefix = e->rep_cursor = nullptr;
    e->cap_reps = 0;
    if ((err = dmalloc(&e->d_reps, n)) != hipSuccess ||
        (err = dmalloc(&e->d_prefix, 2 * (n + 1))) != hipSuccess ||
        (err = dmalloc(&e->rep_cursor, n)) != hipSuccess)
      return set_error(KG_E_NOMEM, "ensemble replica tables: %s", hipGetErrorString(err));
    e->cap_reps = n;
  }
  if ((err = regrow(&e->tcount, &e->cap_tiles, tiles)) != hipSuccess)
    return set_error(KG_E_NOMEM, "ensemble tile counts: %s", hipGetErrorString(err));
  return KG_OK;
}

// Unloads tiles [tile0, tile0 + ntiles) of the current layout into `dst`
// at their runs' append slots.
int unload(kg_ensemble* e, Buf dst, uint32_t tile0, uint32_t ntiles) {
  KG_HIP(launch_fill_u32(e->rep_cursor, 0u, e->nreps, e->stream));
  KG_HIP(launch_rt_unload(e->wl(), e->rt.dev(), dst, e->rep_cursor, tile0, ntiles, e->stream));
  return KG_OK;
}

hipError_t copy_buf(Buf dst, uint32_t dbase, Buf src, uint32_t sbase, uint32_t n, hipStream_t s) {
  const size_t f = sizeof(float) * n;
  hipError_t e;
  if ((e = hipMemcpyAsync(dst.x + dbase, src.x + sbase, f, hipMemcpyDeviceToDevice, s))) return e;
  if ((e = hipMemcpyAsync(dst.y + dbase, src.y + sbase, f, hipMemcpyDeviceToDevice, s))) return e;
  if ((e = hipMemcpyAsync(dst.hx + dbase, src.hx + sbase, f, hipMemcpyDeviceToDevice, s))) return e;
  if ((e = hipMemcpyAsync(dst.hy + dbase, src.hy + sbase, f, hipMemcpyDeviceToDevice, s))) return e;
  return hipMemcpyAsync(dst.id + dbase, src.id + sbase, sizeof(uint32_t) * n,
                        hipMemcpyDeviceToDevice, s);
}

// Halves the tile height of the runs holding the fullest tiles (at least
// half the fullest tile's count), for runs whose flocks have outgrown every
// LDS class; false if none of them can shrink.
constexpr int kShrinkAttempts = 8;

bool shrink_crowded(kg_ensemble* e, const std::vector<uint32_t>& counts) {
  uint32_t top = 0;
  for (uint32_t c : counts) top = std::max(top, c);
  bool any = false;
  for (EnsRun& r : e->runs) {
    uint32_t m = 0;
    for (uint32_t t = r.tile0; t < r.tile0 + r.ntiles; ++t) m = std::max(m, counts[t]);
    if (2ull * m < top || r.g.ty <= 1) continue;
    const uint32_t ty = (uint32_t)r.g.ty / 2;
    const uint32_t hb = std::min<uint32_t>((uint32_t)r.g.hb, ty);
    Geom g;
    std::string why;
    if (!make_geom_rt(r.desc.width, r.desc.height, r.desc.discretization, 1, r.desc.agents,
                      (uint32_t)r.g.tx, ty, hb, &g, &why))
      continue;
    RepDev probe{};
    probe.g = g;
    set_rep_params(&probe, r.desc.params);
    if (!probe.lds_ok) continue;
    r.g = g;
    any = true;
  }
  return any;
}

// The form of the next layout: the requested one, or by the runs' mean size
// (the single-field crossover of the two paths lies between 16k and 64k).
constexpr uint64_t kBinnedMeanAgents = 32768;

bool want_binned(const kg_ensemble* e) {
  if (e->want >= 0) return e->want == 1;
  uint64_t agents = 0, live = 0;
  for (const EnsRun& r : e->runs)
    if (!r.dead) {
      agents += r.desc.agents;
      ++live;
    }
  return live && agents >= kBinnedMeanAgents * live;
}

// Buffers of the bin-grouped form for `agents` and `tiles` (the second
// offset table comes with the bin arrays).
int reserve_binned(kg_ensemble* e, size_t agents, size_t tiles) {
  if (int rc = reserve_buf(e, &e->R2, &e->cap_r2, agents)) return rc;
  if (int rc = reserve_buf(e, &e->W, &e->cap_w, agents)) return rc;
  hipError_t err;
  if ((err = regrow(&e->gflag, &e->cap_gflag, tiles + 2)) != hipSuccess)
    return set_error(KG_E_NOMEM, "ensemble bin-grouped buffers: %s", hipGetErrorString(err));
  if (!e->gstep && (err = dmalloc(&e->gstep, 1)) != hipSuccess)
    return set_error(KG_E_NOMEM, "ensemble step counter: %s", hipGetErrorString(err));
  if (tiles > e->cap_outbox) {
    (void)hipFree(e->outbox.n);
    (void)hipFree(e->outbox.slot);
    e->outbox = Outbox{};
    e->cap_outbox = 0;
    const size_t n = std::max(tiles, e->cap_outbox + e->cap_outbox / 2);
    if ((err = dmalloc(&e->outbox.n, n)) != hipSuccess ||
        (err = dmalloc(&e->outbox.slot, n * kOutboxCap)) != hipSuccess)
      return set_error(KG_E_NOMEM, "ensemble outboxes: %s", hipGetErrorString(err));
    e->cap_outbox = n;
  }
  if (!e->gate && (err = dmalloc(&e->gate, 4)) != hipSuccess)
    return set_error(KG_E_NOMEM, "ensemble gate: %s", hipGetErrorString(err));
  return KG_OK;
}

// Builds the layout for the current run list: every run's agents go through
// the append buffer and the bins, then into a fresh tile-resident plan or,
// for the bin-grouped form, through the bin sort.
int rebuild(kg_ensemble* e) {
  KG_HIP(hipStreamSynchronize(e->stream));
  const uint64_t G = e->rt.G;
  // retired runs leave; the survivors keep their slots in the old layout
  // for the unload below
  e->runs.erase(std::remove_if(e->runs.begin(), e->runs.end(),
                               [](const EnsRun& r) { return r.dead; }),
                e->runs.end());
  // new slots; survivors of a built layout are unloaded first
  std::vector<uint32_t> old_base(e->runs.size());
  uint64_t agents = 0, bins = 0, tiles = 0;
  for (size_t i = 0; i < e->runs.size(); ++i) {
    EnsRun& r = e->runs[i];
    old_base[i] = r.app_base;
    agents += r.desc.agents;
    bins += r.g.nbins;
    tiles += r.g.ntiles;
  }
  if (agents >= (1ull << 31) || bins >= (1ull << 31) || tiles >= (1ull << 31))
    return set_error(KG_E_ADMIT, "ensemble too large for 32-bit indices");
  // the old layout (replica tables, A) stays intact until its survivors
  // are unloaded into A and copied into the build buffer A2
  if (int rc = reserve_buf(e, &e->A2, &e->cap_a2, std::max<uint64_t>(agents, 1))) return rc;
  const bool have_old = e->built && e->nreps > 0;
  const bool old_binned = have_old && e->binned;
  if (!old_binned)
    if (int rc = reserve_buf(e, &e->R, &e->cap_r, std::max<uint64_t>(agents, 1))) return rc;
  // survivors of a bin-grouped layout are already together in the read
  // buffer, at their append slots
  if (have_old && !e->binned) {
    if (int rc = unload(e, e->A.b, 0, e->ntiles)) return rc;
  }
  const Buf old_src = e->binned ? e->R.b : e->A.b;
  const bool binned = want_binned(e);
  uint32_t base = 0;
  for (size_t i = 0; i < e->runs.size(); ++i) {
    EnsRun& r = e->runs[i];
    if (r.fresh) {
      KG_HIP(launch_init_uniform(e->A2.b, base, r.desc.agents, r.desc.width, r.desc.height,
                                 r.desc.init_seed, r.desc.heading_mode, e->stream));
    } else {
      KG_HIP(copy_buf(e->A2.b, base, old_src, old_base[i], r.desc.agents, e->stream));
    }
    r.app_base = base;
    base += r.desc.agents;
  }
  // the survivors have left the old read buffer: it may grow now
  if (old_binned)
    if (int rc = reserve_buf(e, &e->R, &e->cap_r, std::max<uint64_t>(agents, 1))) return rc;
  e->drop_bgraphs();
  RtPlan plan;
  RtClassPlan cls;
  WorkList w{};
  for (int attempt = 0;; ++attempt) {
    uint64_t lbins = 0, ltiles = 0;
    for (const EnsRun& r : e->runs) {
      lbins = scan_tile_align(lbins) + r.g.nbins;
      ltiles += r.g.ntiles;
    }
    if (lbins >= (1ull << 31) || ltiles >= (1ull << 31))
      return set_error(KG_E_ADMIT, "ensemble too large for 32-bit indices");
    // the old layout is no longer read: grow the tables and the unload buffer
    if (int rc = reserve(e, std::max<uint64_t>(lbins, 1), std::max<size_t>(e->runs.size(), 1),
                         std::max<uint64_t>(ltiles, 1)))
      return rc;
    if (int rc = reserve_buf(e, &e->A, &e->cap_a, std::max<uint64_t>(agents, 1))) return rc;
    // replica table
    const uint32_t nreps = (uint32_t)e->runs.size();
    e->h_reps.assign(nreps, RepDev{});
    e->h_prefix.assign(2 * (nreps + 1), 0u);
    std::vector<RtRepInfo> infos(nreps);
    uint32_t bin_base = 0, tile0 = 0, boff_base = 0;
    for (uint32_t i = 0; i < nreps; ++i) {
      EnsRun& r = e->runs[i];
      RepDev& rp = e->h_reps[i];
      rp.g = r.g;
      set_rep_params(&rp, r.desc.params);
      rp.bin_base = bin_base;
      rp.app_base = r.app_base;
      rp.app_n = r.desc.agents;
      rp.boff_base = boff_base;
      // a run's next schedule step is steps_done: the first tile-resident
      // launch after this rebuild runs step G, the bin-grouped step
      // counter restarts at 0
      rp.step_base = binned ? r.steps_done : r.steps_done - G;
      r.tile0 = tile0;
      r.ntiles = r.g.ntiles;
      infos[i] = RtRepInfo{r.g, rp.disc_dist, tile0};
      e->h_prefix[i] = tile0;
      e->h_prefix[nreps + 1 + i] = r.app_base;
      bin_base = (uint32_t)scan_tile_align(bin_base + r.g.nbins);
      boff_base += r.g.ntiles * (uint32_t)(r.g.tx * r.g.tr + 1);
      tile0 += r.g.ntiles;
    }
    e->h_prefix[nreps] = tile0;
    e->h_prefix[2 * nreps + 1] = base;
    e->nreps = nreps;
    e->ntiles = tile0;
    e->nbins = nreps ? e->h_reps[nreps - 1].bin_base + e->runs[nreps - 1].g.nbins : 0;
    e->nagents = base;
    e->dead_tiles = 0;
    e->built = true;
    e->dirty = false;
    e->replan = false;
    if (nreps == 0) return KG_OK;
    KG_HIP(hipMemcpyAsync(e->d_reps, e->h_reps.data(), nreps * sizeof(RepDev),
                          hipMemcpyHostToDevice, e->stream));
    KG_HIP(hipMemcpyAsync(e->d_prefix, e->h_prefix.data(), e->h_prefix.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, e->stream));
    // Field::lazy_update (field_2d.rs:905-921) of every run's agents at once
    w = e->wl();
    KG_HIP(launch_count_append(w, e->A2.b, e->count, e->err, e->stream));
    KG_HIP(launch_scan(e->count, e->offset, e->cursor, e->nbins, e->ss, nullptr, e->stream));
    KG_HIP(launch_scatter_append(w, e->A2.b, e->cursor, e->R.b, e->stream));
    // plan from the per-tile counts
    KG_HIP(launch_tile_counts(w, e->offset, e->tcount, e->stream));
    std::vector<uint32_t> counts(e->ntiles), nb(e->ntiles);
    KG_HIP(hipMemcpyAsync(counts.data(), e->tcount, e->ntiles * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, e->stream));
    uint32_t errflag = 0;
    KG_HIP(hipMemcpyAsync(&errflag, e->err, sizeof(errflag), hipMemcpyDeviceToHost, e->stream));
    KG_HIP(hipStreamSynchronize(e->stream));
    if (errflag)
      return set_error(KG_E_RANGE, "an agent discretized outside its run's bag grid");
    if (binned) break;
    std::vector<double> expect(e->ntiles);
    for (uint32_t i = 0; i < nreps; ++i) {
      const EnsRun& r = e->runs[i];
      std::vector<double> ex;
      rt_expected_counts(r.g, r.desc.agents, &ex);
      std::copy(ex.begin(), ex.end(), expect.begin() + r.tile0);
      std::fill(nb.begin() + r.tile0, nb.begin() + r.tile0 + r.ntiles,
                (uint32_t)(r.g.tx * r.g.tr));
    }
    plan = RtPlan{};
    rt_plan(counts, nb, expect, &plan);
    cls = RtClassPlan{};
    const size_t lds_normal = e->tuning.lds_normal ? e->tuning.lds_normal : (size_t)0;
    rt_classes(&plan, counts, infos, lds_normal, e->lds_limit, &cls);
    if (cls.fits) break;
    // tiles too crowded for every LDS class: shorter tiles for the runs that
    // hold the fullest ones, and bin again
    if (attempt >= kShrinkAttempts || !shrink_crowded(e, counts))
      return set_error(KG_E_ADMIT, "the runs' tiles do not fit the LDS budget of the tile-resident step");
  }
  e->binned = binned;
  if (binned) {
    // every bin into (cell y, id) order, the consumed build buffer as scratch
    if (int rc = reserve_binned(e, std::max<uint64_t>(agents, 1), e->ntiles)) return rc;
    KG_HIP(launch_sort_bins(w, e->R.b, e->offset, e->A2.b, e->stream));
    KG_HIP(hipMemsetAsync(e->gstep, 0, sizeof(uint64_t), e->stream));
    int cap = 256;
    for (const EnsRun& r : e->runs) cap = std::max(cap, pick_lds_cap(r.g, r.desc.agents));
    e->lds_cap = clamp_lds_cap(cap);
    e->parity = 0;
    e->gather = rebin_gather_enabled();
    for (const RepDev& rp : e->h_reps)
      if ((uint32_t)(rp.g.tx * rp.g.tr) > kGatherMaxBins) e->gather = false;
    KG_HIP(hipMemsetAsync(e->gate, 0, 4 * sizeof(uint32_t), e->stream));
    e->redone_seen = 0;
  } else {
    if (int rc = e->rt.setup(plan, cls, e->nagents, e->stream)) return rc;
    if (int rc = e->rt.reset(G, e->stream)) return rc;
    const RtDev D = e->rt.dev();
    for (int c = 0; c < e->rt.ncls; ++c)
      KG_HIP(launch_rt_load(w, D, e->rt.caps[c], c, e->rt.dense + e->rt.off[c],
                            e->rt.off[c + 1] - e->rt.off[c], e->R.b, e->offset, G - 1,
                            e->stream));
  }
  for (EnsRun& r : e->runs) r.fresh = false;
  ++e->builds;
  return KG_OK;
}

// One bin-grouped schedule step of every replica: k_flock over all tiles,
// the scan over all bins (it advances the step counter), the gather (the
// scatter and the bin sort only if it cannot), then the buffers swap.
// k_flock_global's work-groups in an ensemble step: a layout of many runs
// seldom has a global-path tile, and a grid of one work-group per run of
// birds costs about 20 us at 23M birds just to launch and leave.
constexpr uint32_t kEnsGlobalGrid = 2048;

hipError_t binned_step(kg_ensemble* e, const WorkList& w) {
  hipError_t err;
  const int p = e->parity;
  const bool summed = e->summed;
  uint32_t* gate = e->gather ? e->gate + (summed ? p : 0) : nullptr;
  uint32_t* count = summed ? e->counts(p) : e->count;
  if ((err = launch_flock(w, e->R.b, e->offset, e->W.b, count, e->gstep, e->err, e->lds_cap,
                          e->gflag, e->nagents, e->stream,
                          e->gather ? e->outbox : Outbox{}, false,
                          summed ? e->tsum[p] : nullptr, kEnsGlobalGrid)) != hipSuccess)
    return err;
  const uint32_t nscan = (e->nbins + kScanTile - 1) / kScanTile;
  if (summed) err = launch_scan_prefix(e->tsum[p], nscan, e->stream);
  else err = launch_scan(e->count, e->offset2, e->cursor, e->nbins, e->ss, e->gstep, e->stream, gate);
  if (err != hipSuccess) return err;
  const Gate gt{gate, gate ? e->gate + 2 : nullptr, summed ? e->gate + (p ^ 1) : nullptr,
                gate ? e->gate + 3 : nullptr};
  if (gate) {
    GatherScan sc;
    if (summed) {
      sc.count = count;
      sc.tile_sum = e->tsum[p];
      sc.count_next = e->counts(p ^ 1);
      sc.tile_sum_next = e->tsum[p ^ 1];
      sc.gstep = e->gstep;
      sc.nbins = e->nbins;
      sc.last_item = e->ntiles - 1;
      sc.prefixed = 1;
    }
    if ((err = launch_rebin_gather(w, e->R.b, e->W.b, e->offset, e->offset2, e->outbox, e->R2.b,
                                   gt, sc, e->stream)) != hipSuccess)
      return err;
    // what the gather could not rebin, in one launch (the write buffer,
    // consumed by the scatter, is the sort's scratch)
    if ((err = launch_rebin_fallback(w, e->R.b, e->W.b, e->offset,
                                     summed ? e->counts(p) : e->cursor, e->R2.b, e->offset2,
                                     e->W.b, gt, e->err, summed, e->stream)) != hipSuccess)
      return err;
  } else {
    if ((err = launch_scatter_aligned(w, e->R.b, e->W.b, e->offset, e->cursor, e->R2.b,
                                      e->stream)) != hipSuccess)
      return err;
    // the write buffer has been consumed: it is the sort's scratch
    if ((err = launch_sort_bins(w, e->R2.b, e->offset2, e->W.b, e->stream)) != hipSuccess)
      return err;
  }
  std::swap(e->R, e->R2);
  std::swap(e->cap_r, e->cap_r2);
  std::swap(e->offset, e->offset2);
  e->parity ^= 1;
  return hipSuccess;
}

// `n` bin-grouped steps (graph replays of two steps when allowed: the
// buffers are back in place after two).
int binned_steps(kg_ensemble* e, uint32_t n, bool use_graph) {
  const WorkList w = e->wl();
  uint32_t left = n;
  // every gathered step is summed: the sums and both gate flags start at
  // zero, and so do both count arrays outside a run
  e->summed = e->gather;
  if (e->summed) {
    const uint32_t nscan = (e->nbins + kScanTile - 1) / kScanTile;
    for (int k = 0; k < 2; ++k)
      KG_HIP(hipMemsetAsync(e->tsum[k], 0, nscan * sizeof(uint32_t), e->stream));
    KG_HIP(hipMemsetAsync(e->gate, 0, 2 * sizeof(uint32_t), e->stream));
  }
  if (use_graph && left >= 2) {
    hipGraphExec_t exec = nullptr;
    const int key = e->parity + (e->gather ? 2 : 0);
    for (auto& g : e->bgraphs)
      if (g.first == key) exec = g.second;
    if (!exec) {
      hipGraph_t graph = nullptr;
      KG_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      hipError_t err = binned_step(e, w);
      if (err == hipSuccess) err = binned_step(e, w);
      const hipError_t e2 = hipStreamEndCapture(e->stream, &graph);
      if (err != hipSuccess || e2 != hipSuccess)
        return set_error(KG_E_HIP, "graph capture: %s",
                         hipGetErrorString(err != hipSuccess ? err : e2));
      err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (err != hipSuccess)
        return set_error(KG_E_HIP, "graph instantiate: %s", hipGetErrorString(err));
      e->bgraphs.push_back({key, exec});
    }
    while (left >= 2) {
      KG_HIP(hipGraphLaunch(exec, e->stream));
      left -= 2;
    }
  }
  for (; left > 0; --left) KG_HIP(binned_step(e, w));
  if (e->summed) {
    // the last step's counts were read, not cleared
    KG_HIP(launch_fill_u32(e->counts(e->parity ^ 1), 0u, e->nbins, e->stream));
    e->summed = false;
  }
  e->steps_run += n;
  for (EnsRun& r : e->runs) r.steps_done += n;
  return KG_OK;
}

hipError_t launch_step(kg_ensemble* e, const WorkList& w, const RtDev& D, uint32_t koff) {
  return rt_launch_classes(w, D, e->rt, koff, e->stream);
}

// `n` schedule steps of every run (graph replays of 8 steps when allowed).
int steps(kg_ensemble* e, uint32_t n, bool use_graph) {
  const RtDev D = e->rt.dev();
  const WorkList w = e->wl();
  uint32_t left = n;
  constexpr int kChunk = 8;
  if (use_graph && left >= kChunk) {
    hipGraphExec_t exec = nullptr;
    for (auto& gch : e->rt.graphs)
      if (gch.first == kChunk) exec = gch.second;
    if (!exec) {
      hipGraph_t graph = nullptr;
      KG_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      hipError_t err = hipSuccess;
      for (int k = 0; k < kChunk && err == hipSuccess; ++k) err = launch_step(e, w, D, (uint32_t)k);
      if (err == hipSuccess) err = launch_add_u64(D.gstep, kChunk, e->stream);
      const hipError_t e2 = hipStreamEndCapture(e->stream, &graph);
      if (err != hipSuccess || e2 != hipSuccess)
        return set_error(KG_E_HIP, "graph capture: %s",
                         hipGetErrorString(err != hipSuccess ? err : e2));
      err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (err != hipSuccess)
        return set_error(KG_E_HIP, "graph instantiate: %s", hipGetErrorString(err));
      e->rt.graphs.push_back({kChunk, exec});
    }
    while (left >= (uint32_t)kChunk) {
      KG_HIP(hipGraphLaunch(exec, e->stream));
      left -= kChunk;
    }
  }
  if (left > 0) {
    for (uint32_t k = 0; k < left; ++k) KG_HIP(launch_step(e, w, D, k));
    KG_HIP(launch_add_u64(D.gstep, left, e->stream));
  }
  e->rt.G += n;
  e->steps_run += n;
  for (EnsRun& r : e->runs) r.steps_done += n;
  return KG_OK;
}

int check_flags(kg_ensemble* e) {
  KG_HIP(hipStreamSynchronize(e->stream));
  if (e->binned) {
    uint32_t errflag = 0;
    KG_HIP(hipMemcpy(&errflag, e->err, sizeof(errflag), hipMemcpyDeviceToHost));
    if (errflag & 2u)
      return set_error(KG_E_HIP, "the rebinning fallback's grid-wide wait timed out");
    if (errflag) return set_error(KG_E_RANGE, "an agent left its run's bag grid");
    if (e->gather) {
      // steps the gather could not rebin since the last check: a layout
      // that keeps failing it pays for it and for the scatter and the sort
      uint32_t redone = 0;
      KG_HIP(hipMemcpy(&redone, e->gate + 2, sizeof(redone), hipMemcpyDeviceToHost));
      if (redone != e->redone_seen) e->gather = false;
      e->redone_seen = redone;
    }
    return KG_OK;
  }
  uint32_t fl2[2] = {0, 0};
  KG_HIP(hipMemcpy(fl2, e->rt.flags, sizeof(fl2), hipMemcpyDeviceToHost));
  const uint32_t fl = fl2[0];
  e->flags_seen |= fl;
  if (fl & kRtFlagRange)
    return set_error(KG_E_RANGE, "an agent left its run's bag grid");
  if ((fl & (kRtFlagSpill | kRtFlagGlobal)) || fl2[1] >= kRtEnsReplanDeferred) e->replan = true;
  return KG_OK;
}

EnsRun* find_run(kg_ensemble* e, uint32_t id, size_t* idx = nullptr) {
  for (size_t i = 0; i < e->runs.size(); ++i)
    if (e->runs[i].id == id && !e->runs[i].dead) {
      if (idx) *idx = i;
      return &e->runs[i];
    }
  return nullptr;
}

}  // namespace

extern "C" {

int kg_run_footprint(const kg_run_desc* run, const kg_tuning* tuning, uint64_t* bytes) {
  clear_error();
  if (!run || !bytes) return set_error(KG_E_INVALID, "NULL argument");
  const kg_tuning tu = tuning ? *tuning : kg_tuning{};
  Geom g;
  std::string why;
  if (!run_geom(*run, tu, &g, nullptr, &why)) return set_error(KG_E_INVALID, "%s", why.c_str());
  const uint64_t n = run->agents;
  // agent buffers (append x2, binned) and stayer regions (two parities,
  // about 2.5x the count) and inboxes, all with the growth headroom of half
  // again that reserve_buf and the layout give them; spill, bins (count,
  // offset, cursor) and the per-tile tables
  *bytes = n * 20 * 3 * 3 / 2 + n * 40 * 3 * 3 / 2 + n * 44 + ((uint64_t)g.nbins + kScanTile) * 20 +
           (uint64_t)g.ntiles * ((uint64_t)(g.tx * g.tr + 1) * 8 + 64 + (kOutboxCap + 1) * 4) *
               3 / 2;
  return KG_OK;
}

int kg_ensemble_create(const kg_tuning* tuning, kg_ensemble** out) {
  clear_error();
  if (!out) return set_error(KG_E_INVALID, "out is NULL");
  *out = nullptr;
  kg_ensemble* e = new kg_ensemble();
  if (tuning) e->tuning = *tuning;
  e->lds_limit = device_lds_limit();
  hipError_t err;
  if ((err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess ||
      (err = dmalloc(&e->err, 1)) != hipSuccess ||
      (err = hipMemsetAsync(e->err, 0, sizeof(uint32_t), e->stream)) != hipSuccess ||
      (err = hipStreamSynchronize(e->stream)) != hipSuccess) {
    kg_ensemble_destroy(e);
    return set_error(KG_E_NOMEM, "ensemble setup: %s", hipGetErrorString(err));
  }
  e->rt.G = 1;
  *out = e;
  return KG_OK;
}

void kg_ensemble_destroy(kg_ensemble* e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  e->drop_bgraphs();
  e->rt.release();
  e->A.release();
  e->A2.release();
  e->R.release();
  e->R2.release();
  e->W.release();
  for (void* p : {(void*)e->count, (void*)e->offset, (void*)e->offset2, (void*)e->cursor,
                  (void*)e->ss.partial, (void*)e->d_reps, (void*)e->d_prefix,
                  (void*)e->rep_cursor, (void*)e->tcount, (void*)e->err, (void*)e->gflag,
                  (void*)e->gstep, (void*)e->outbox.n, (void*)e->outbox.slot, (void*)e->gate})
    (void)hipFree(p);
  for (void* p : {(void*)e->count2, (void*)e->tsum[0], (void*)e->tsum[1]}) (void)hipFree(p);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int kg_ensemble_admit(kg_ensemble* e, const kg_run_desc* runs, uint32_t n, uint32_t* ids) {
  clear_error();
  if (!e || (n && !runs)) return set_error(KG_E_INVALID, "NULL argument");
  std::vector<EnsRun> add(n);
  for (uint32_t i = 0; i < n; ++i) {
    std::string why;
    if (!run_geom(runs[i], e->tuning, &add[i].g, nullptr, &why))
      return set_error(KG_E_INVALID, "run %u: %s", i, why.c_str());
    add[i].desc = runs[i];
  }
  for (uint32_t i = 0; i < n; ++i) {
    add[i].id = e->next_id++;
    if (ids) ids[i] = add[i].id;
    e->runs.push_back(add[i]);
  }
  if (n) e->dirty = true;
  return KG_OK;
}

int kg_ensemble_retire(kg_ensemble* e, uint32_t id) {
  clear_error();
  if (!e) return set_error(KG_E_INVALID, "ensemble is NULL");
  size_t idx = 0;
  EnsRun* r = find_run(e, id, &idx);
  if (!r) return set_error(KG_E_INVALID, "no run %u", id);
  if (e->built && !e->dirty && idx < e->nreps) {
    // the run stays in the layout as a dead replica (its tiles exit at
    // once) until the next build, so retiring costs no rebuild; once dead
    // tiles are a third of the layout, the next run compacts it
    r->dead = true;
    e->h_reps[idx].dead = 1u;
    KG_HIP(launch_set_rep(e->d_reps + idx, e->h_reps[idx], e->stream));
    e->dead_tiles += r->ntiles;
    if (3ull * e->dead_tiles > e->ntiles) e->dirty = true;
    return KG_OK;
  }
  e->runs.erase(e->runs.begin() + (ptrdiff_t)idx);
  e->dirty = true;
  return KG_OK;
}

int kg_ensemble_run(kg_ensemble* e, uint32_t nsteps, int flags) {
  clear_error();
  if (!e) return set_error(KG_E_INVALID, "ensemble is NULL");
  if ((flags & KG_RUN_BINNED) && (flags & KG_RUN_TILE_RESIDENT))
    return set_error(KG_E_INVALID, "KG_RUN_BINNED and KG_RUN_TILE_RESIDENT exclude each other");
  if (nsteps == 0) return KG_OK;
  e->want = (flags & KG_RUN_BINNED) ? 1 : ((flags & KG_RUN_TILE_RESIDENT) ? 0 : -1);
  if (e->built && !e->dirty && e->nreps > 0 && want_binned(e) != e->binned) e->dirty = true;
  if (e->dirty || !e->built)
    if (int rc = rebuild(e)) return rc;
  if (e->nreps == 0) return KG_OK;
  const bool graph = (flags & KG_RUN_GRAPH) != 0;
  while (nsteps > 0) {
    const uint32_t chunk = std::min(nsteps, kRtCheckEvery);
    if (int rc = e->binned ? binned_steps(e, chunk, graph) : steps(e, chunk, graph)) return rc;
    nsteps -= chunk;
    if (int rc = check_flags(e)) return rc;
    if (e->replan)
      if (int rc = rebuild(e)) return rc;
  }
  return KG_OK;
}

int kg_ensemble_steps(kg_ensemble* e, uint32_t id, uint64_t* steps_out) {
  clear_error();
  if (!e || !steps_out) return set_error(KG_E_INVALID, "NULL argument");
  const EnsRun* r = find_run(e, id);
  if (!r) return set_error(KG_E_INVALID, "no run %u", id);
  *steps_out = r->steps_done;
  return KG_OK;
}

int kg_ensemble_read(kg_ensemble* e, uint32_t id, uint32_t* ids, float* x, float* y, float* hx,
                     float* hy, size_t cap, size_t* n) {
  clear_error();
  if (!e || !n) return set_error(KG_E_INVALID, "NULL argument");
  {
    const EnsRun* r0 = find_run(e, id);
    if (!r0 || r0->dead) return set_error(KG_E_INVALID, "no run %u", id);
    *n = r0->desc.agents;
    if (cap < r0->desc.agents)
      return set_error(KG_E_CAPACITY, "buffer holds %zu of %u agents", cap, r0->desc.agents);
  }
  // a rebuild drops retired runs from the run list: look the run up after it
  if (e->dirty || !e->built)
    if (int rc = rebuild(e)) return rc;
  size_t idx = 0;
  EnsRun* r = find_run(e, id, &idx);
  if (!r) return set_error(KG_E_INVALID, "no run %u", id);
  // the bin-grouped read buffer holds every run at its append slots; the
  // tile-resident layout is unloaded there first
  if (!e->binned)
    if (int rc = unload(e, e->A.b, r->tile0, r->ntiles)) return rc;
  const Buf src = e->binned ? e->R.b : e->A.b;
  const size_t m = r->desc.agents;
  const uint32_t b = r->app_base;
  if (ids) KG_HIP(hipMemcpyAsync(ids, src.id + b, m * 4, hipMemcpyDeviceToHost, e->stream));
  if (x) KG_HIP(hipMemcpyAsync(x, src.x + b, m * 4, hipMemcpyDeviceToHost, e->stream));
  if (y) KG_HIP(hipMemcpyAsync(y, src.y + b, m * 4, hipMemcpyDeviceToHost, e->stream));
  if (hx) KG_HIP(hipMemcpyAsync(hx, src.hx + b, m * 4, hipMemcpyDeviceToHost, e->stream));
  if (hy) KG_HIP(hipMemcpyAsync(hy, src.hy + b, m * 4, hipMemcpyDeviceToHost, e->stream));
  uint32_t got = (uint32_t)m;
  if (!e->binned)
    KG_HIP(hipMemcpyAsync(&got, e->rep_cursor + idx, sizeof(got), hipMemcpyDeviceToHost,
                          e->stream));
  KG_HIP(hipStreamSynchronize(e->stream));
  if (got != r->desc.agents)
    return set_error(KG_E_STATE, "run %u holds %u agents, expected %u", id, got, r->desc.agents);
  return KG_OK;
}

int kg_ensemble_sync(kg_ensemble* e) {
  clear_error();
  if (!e) return set_error(KG_E_INVALID, "ensemble is NULL");
  KG_HIP(hipStreamSynchronize(e->stream));
  return KG_OK;
}

void* kg_ensemble_stream(kg_ensemble* e) { return e ? (void*)e->stream : nullptr; }

int kg_ensemble_stats(kg_ensemble* e, uint64_t out[8]) {
  clear_error();
  if (!e || !out) return set_error(KG_E_INVALID, "NULL argument");
  uint64_t agents = 0, live = 0;
  for (const EnsRun& r : e->runs)
    if (!r.dead) {
      agents += r.desc.agents;
      ++live;
    }
  const uint64_t v[8] = {live,      agents,         e->ntiles,     e->rt.ndense,
                         e->builds,      e->flags_seen, e->steps_run, e->bytes_held()};
  std::memcpy(out, v, sizeof(v));
  return KG_OK;
}

}  // extern "C"
