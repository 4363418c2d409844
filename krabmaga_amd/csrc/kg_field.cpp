// kg_field.cpp -- C ABI of one device-resident Field2D<Bird> (krabgpu.h).
//
// A kg_field is a replica table of one: the same kernels drive the batched
// ensemble.  The read buffer lives in one of two layouts:
//   bin-grouped (API layout)   R, R2 read buffer and rebinning target,
//                              W     write buffer produced by the flock
//                                    kernel, aligned with R,
//                              A     write buffer filled by
//                                    set_object_location / init;
//   tile-resident (kg_rt.hpp)  used by kg_schedule_run: one launch per
//                              schedule step.
// API calls that read or write the buffers move the field back to the
// bin-grouped layout first.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "kg_rt_state.hpp"

using namespace kg;

namespace {

struct GraphKey {
  uint8_t params[sizeof(kg_flock_params)];
  int parity;
  int gather;         // the steps rebin by gather
  int inline_global;  // global-path tiles stepped inside k_flock
  bool operator<(const GraphKey& o) const {
    if (parity != o.parity) return parity < o.parity;
    if (gather != o.gather) return gather < o.gather;
    if (inline_global != o.inline_global) return inline_global < o.inline_global;
    return std::memcmp(params, o.params, sizeof(params)) < 0;
  }
};

}  // namespace

struct kg_field {
  kg_field_desc desc{};
  Geom g{};
  uint32_t cap = 0;
  int lds_cap = 1024;
  hipStream_t stream = nullptr;

  DevBuf R, R2, W, A;
  int parity = 0;  // flips with every rebinning (R <-> R2)
  uint32_t *count = nullptr, *offset = nullptr, *offset2 = nullptr, *cursor = nullptr;
  ScanState ss{};
  uint32_t* err = nullptr;
  uint64_t* gstep = nullptr;
  uint64_t* zero64 = nullptr;
  uint32_t* prefix = nullptr;  // [0..1] tile prefix, [2..3] append prefix
  RepDev* d_rep = nullptr;
  uint32_t* scratch = nullptr;  // small query outputs
  uint32_t* tcount = nullptr;   // per-tile agent counts (tile-resident planning)
  // per tile: k_flock left its birds to the global path; then the step
  // stamp and the count of global-path tiles so far
  uint32_t* gflag = nullptr;
  // gather rebinning (tiles of at most kGatherMaxBins bins): the flock
  // kernel's lists of leavers, the gate words (a flag per step parity, the
  // steps redone, the fallback's grid-wait counter) and whether the last
  // flock step filled the lists
  bool gather_ok = false;
  Outbox outbox{};
  uint32_t* gate = nullptr;
  bool outbox_filled = false;
  // steps run with the gather; the counts the host last saw (pinned, copied
  // after every run): [0] steps the gather could not rebin, [1] global-path
  // tiles; the first gather step of a field is checked on its own
  uint64_t gather_steps = 0, gather_steps_copied = 0;
  uint32_t* seen = nullptr;
  bool gather_probed = false;
  // no tile has taken the global path since the probe: k_flock steps such a
  // tile itself and k_flock_global is not launched
  bool global_inline = false;
  // the steps of the current run add their counts to the scan's tile sums
  // and the gather does the scan: bin counts and sums by step parity, the
  // gather clearing the other step's (count2 is the second count array)
  bool summed = false;
  uint32_t* tsum[2] = {nullptr, nullptr};
  uint32_t* count2 = nullptr;
  uint32_t* counts(int p) const { return p ? count2 : count; }
  RepDev h_rep{};

  uint32_t nR = 0, nW = 0, nA = 0;
  bool w_aligned = false;
  bool poisoned = false;
  uint32_t poison_bits = 0;  // the device error flag's bits that poisoned it

  // Field2D bookkeeping (field_2d.rs:277, :290-293)
  size_t nagents = 0;
  bool dens_check = false;
  size_t density_estimation = 0;

  std::map<GraphKey, hipGraphExec_t> graphs;

  // tile-resident layout
  RtState rt;
  bool rt_mode = false;     // the read buffer lives in the tile-resident layout
  bool rt_replan = false;   // its capacities were exceeded: re-plan before the next run
  bool rt_unsuitable = false;  // the last plan did not fit the LDS budget
  uint64_t rt_plans = 0, rt_flags_seen = 0, rt_steps_run = 0;
  size_t lds_limit = 65536;

  // HIP-event timing of individual kernel launches (kg_field_profile)
  bool prof_on = false;
  bool capturing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  double prof_ms[KG_KERNEL_COUNT] = {};
  uint64_t prof_n[KG_KERNEL_COUNT] = {};

  WorkList worklist() const {
    WorkList wl;
    wl.reps = d_rep;
    wl.nreps = 1;
    wl.tile_prefix = prefix;
    wl.ntiles = g.ntiles;
    wl.app_prefix = prefix + 2;
    wl.napp = nA;
    return wl;
  }
};

namespace {

hipEvent_t take_event(kg_field* f) {
  if (!f->ev_pool.empty()) {
    hipEvent_t e = f->ev_pool.back();
    f->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one kernel launch with events on the field's stream when
// profiling is on (never inside a graph capture).
struct ProfScope {
  kg_field* f;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(kg_field* f_, int kind_) : f(f_), kind(kind_) {
    if (!f->prof_on || f->capturing) return;
    a = take_event(f);
    b = take_event(f);
    if (a) (void)hipEventRecord(a, f->stream);
  }
  ~ProfScope() {
    if (!a || !b) return;
    (void)hipEventRecord(b, f->stream);
    f->ev_pending.push_back({kind, {a, b}});
  }
};

int prof_collect(kg_field* f) {
  if (f->ev_pending.empty()) return KG_OK;
  KG_HIP(hipStreamSynchronize(f->stream));
  for (auto& pe : f->ev_pending) {
    float ms = 0.f;
    KG_HIP(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
    f->prof_ms[pe.first] += ms;
    f->prof_n[pe.first] += 1;
    f->ev_pool.push_back(pe.second.first);
    f->ev_pool.push_back(pe.second.second);
  }
  f->ev_pending.clear();
  return KG_OK;
}

int check_poison(kg_field* f) {
  if (f->poisoned && (f->poison_bits & 2u))
    return set_error(KG_E_HIP,
                     "the rebinning fallback's grid-wide wait timed out in an earlier step; "
                     "the field is unusable");
  if (f->poisoned)
    return set_error(KG_E_RANGE,
                     "a location fell outside the bag grid in an earlier step; the field is "
                     "unusable");
  return KG_OK;
}

int ensure_atomic(kg_field* f);

// Reads the device error flag (synchronises the stream).
int sync_only(kg_field* f) {
  KG_HIP(hipStreamSynchronize(f->stream));
  uint32_t e = 0;
  KG_HIP(hipMemcpy(&e, f->err, sizeof(e), hipMemcpyDeviceToHost));
  if (e) {
    f->poisoned = true;
    f->poison_bits |= e;
  }
  return check_poison(f);
}

// Brings the read buffer into the bin-grouped layout, synchronises and
// reads the device error flag.
int sync_and_check(kg_field* f) {
  if (int rc = ensure_atomic(f)) return rc;
  KG_HIP(hipStreamSynchronize(f->stream));
  uint32_t e = 0;
  KG_HIP(hipMemcpy(&e, f->err, sizeof(e), hipMemcpyDeviceToHost));
  if (e) {
    f->poisoned = true;
    f->poison_bits |= e;
  }
  return check_poison(f);
}

// Stream-ordered update of the device replica descriptor (a kernel argument
// copy, so no host buffer has to outlive the call).
int upload_rep(kg_field* f) {
  KG_HIP(launch_set_rep(f->d_rep, f->h_rep, f->stream));
  return KG_OK;
}

// Field::lazy_update (field_2d.rs:905-921): bins the write buffer into a
// fresh read buffer.  `step_gstep` advances the schedule-step counter.
int lazy_update_impl(kg_field* f, uint64_t* step_gstep) {
  WorkList wl = f->worklist();
  // a flock step's output with its lists of leavers and nothing appended:
  // the gather rebins it, the scatter and the bin sort run only if it
  // sets the gate
  const bool gather = f->w_aligned && f->outbox_filled && f->nA == 0;
  const int p = f->parity;
  uint32_t* gate = gather ? f->gate + p : nullptr;
  f->outbox_filled = false;
  const bool summed = gather && f->summed;
  // one replica: the append work list needs no device prefix
  if (f->nA) {
    ProfScope ps(f, KG_KERNEL_COUNT_APPEND);
    KG_HIP(launch_count_append(wl, f->A.b, f->count, f->err, f->stream));
  }
  if (!summed) {
    ProfScope ps(f, KG_KERNEL_SCAN);
    KG_HIP(launch_scan(f->count, f->offset2, f->cursor, f->g.nbins, f->ss, step_gstep, f->stream,
                       gate));
  }
  const Gate gt{gate, gather ? f->gate + 2 : nullptr, summed ? f->gate + (p ^ 1) : nullptr,
                gather ? f->gate + 3 : nullptr};
  if (f->w_aligned) {
    ProfScope ps(f, KG_KERNEL_SCATTER);
    if (gather) {
      GatherScan sc;
      if (summed) {
        sc.count = f->counts(p);
        sc.tile_sum = f->tsum[p];
        sc.count_next = f->counts(p ^ 1);
        sc.tile_sum_next = f->tsum[p ^ 1];
        sc.gstep = step_gstep;
        sc.nbins = f->g.nbins;
        sc.last_item = f->g.ntiles - 1;
      }
      KG_HIP(launch_rebin_gather(wl, f->R.b, f->W.b, f->offset, f->offset2, f->outbox,
                                 f->R2.b, gt, sc, f->stream));
      // what the gather could not rebin: scatter and bin sort in one launch
      // (no appended agents in a gathered step); the write buffer has been
      // consumed by then and is the sort's scratch
      KG_HIP(launch_rebin_fallback(wl, f->R.b, f->W.b, f->offset,
                                   summed ? f->counts(p) : f->cursor, f->R2.b, f->offset2,
                                   f->W.b, gt, f->err, summed, f->stream));
    } else {
      KG_HIP(launch_scatter_aligned(wl, f->R.b, f->W.b, f->offset, f->cursor, f->R2.b,
                                    f->stream));
    }
  }
  if (!gather) {
    if (f->nA) {
      ProfScope ps(f, KG_KERNEL_SCATTER_APPEND);
      KG_HIP(launch_scatter_append(wl, f->A.b, f->cursor, f->R2.b, f->stream));
    }
    // the write buffer has been consumed: it is the sort's scratch
    ProfScope ps(f, KG_KERNEL_SORT_BINS);
    KG_HIP(launch_sort_bins(wl, f->R2.b, f->offset2, f->W.b, f->stream));
  }
  std::swap(f->R, f->R2);
  std::swap(f->offset, f->offset2);
  f->parity ^= 1;
  f->nR = (f->w_aligned ? f->nW : 0) + f->nA;
  f->nW = 0;
  f->nA = 0;
  f->w_aligned = false;
  if (!f->dens_check) {
    f->density_estimation = f->nagents / ((size_t)f->g.dw * (size_t)f->g.dh);
    f->dens_check = true;
  }
  return KG_OK;
}

int flock_impl(kg_field* f, const uint64_t* gstep) {
  ProfScope ps(f, KG_KERNEL_FLOCK);
  const bool summed = f->gather_ok && f->summed;
  KG_HIP(launch_flock(f->worklist(), f->R.b, f->offset, f->W.b,
                      summed ? f->counts(f->parity) : f->count, gstep, f->err,
                      f->lds_cap, f->gflag, (uint32_t)f->nR, f->stream,
                      f->gather_ok ? f->outbox : Outbox{}, f->global_inline,
                      summed ? f->tsum[f->parity] : nullptr));
  f->outbox_filled = f->gather_ok;
  f->nW = f->nR;
  f->w_aligned = true;
  return KG_OK;
}

int set_params(kg_field* f, const kg_flock_params* p, uint64_t step_base) {
  std::string why;
  if (!p) return set_error(KG_E_INVALID, "params is NULL");
  if (!params_valid(*p, &why)) return set_error(KG_E_INVALID, "%s", why.c_str());
  set_rep_params(&f->h_rep, *p);
  f->h_rep.step_base = step_base;
  return upload_rep(f);
}

// ---- tile-resident layout -------------------------------------------------

// Deferred tiles in one step that ask for a new plan; KG_RT_REPLAN_DEFERRED
// overrides it for tuning.
uint32_t rt_replan_deferred() {
  static const uint32_t v = [] {
    const char* e = std::getenv("KG_RT_REPLAN_DEFERRED");
    const long n = e ? std::atol(e) : 0;
    return n >= 1 && n <= (long)kRtOvfMax ? (uint32_t)n : kRtReplanDeferred;
  }();
  return v;
}

// Reads the tile-resident flags (synchronises): a location out of the grid
// poisons the field; exceeded capacities ask for a new plan.
int rt_check_flags(kg_field* f) {
  if (!f->rt_mode) return KG_OK;
  KG_HIP(hipStreamSynchronize(f->stream));
  uint32_t fl2[2] = {0, 0};
  KG_HIP(hipMemcpy(fl2, f->rt.flags, sizeof(fl2), hipMemcpyDeviceToHost));
  const uint32_t fl = fl2[0];
  f->rt_flags_seen |= fl;
  if (fl & kRtFlagRange) f->poisoned = true;
  // stay overflow is absorbed by the inboxes and a few deferred tiles by
  // the dense launch; many deferrals mean the plan no longer fits the
  // density (left alone they grow into global-path tiles), and spilling or
  // the global path costs every tile: those ask for a new plan
  if ((fl & (kRtFlagSpill | kRtFlagGlobal)) || fl2[1] >= rt_replan_deferred()) f->rt_replan = true;
  return check_poison(f);
}

// LDS budget of a class-0 work-group: 0 sizes it to keep a CU full for the
// class's work-group size (rt_lds_budget); KG_RT_LDS_NORMAL (bytes) fixes
// it for tuning.
size_t rt_lds_normal() {
  static const size_t v = [] {
    const char* e = std::getenv("KG_RT_LDS_NORMAL");
    const long n = e ? std::atol(e) : 0;
    return n >= 4096 && n <= 160 * 1024 ? (size_t)n : (size_t)0;
  }();
  return v;
}

// Moves the bin-grouped read buffer into the tile-resident layout, with
// capacities planned from the current per-tile counts.
int rt_enter(kg_field* f) {
  const uint32_t nt = f->g.ntiles;
  const uint32_t nbt = (uint32_t)(f->g.tx * f->g.tr);
  if (!f->tcount) KG_HIP(dmalloc(&f->tcount, std::max<uint32_t>(nt, 1u)));
  KG_HIP(launch_tile_counts(f->worklist(), f->offset, f->tcount, f->stream));
  std::vector<uint32_t> counts(nt), nb(nt, nbt);
  KG_HIP(hipMemcpyAsync(counts.data(), f->tcount, nt * sizeof(uint32_t), hipMemcpyDeviceToHost,
                        f->stream));
  KG_HIP(hipStreamSynchronize(f->stream));
  RtPlan plan;
  std::vector<double> expect;
  uint64_t total = 0;
  for (uint32_t c : counts) total += c;
  rt_expected_counts(f->g, total, &expect);
  rt_plan(counts, nb, expect, &plan);
  RtClassPlan cls;
  rt_classes(&plan, counts, {RtRepInfo{f->g, f->h_rep.disc_dist, 0u}}, rt_lds_normal(),
             f->lds_limit, &cls);
  ++f->rt_plans;
  if (!cls.fits) {
    // tiles too crowded for the LDS working set: stay bin-grouped
    f->rt_unsuitable = true;
    return KG_OK;
  }
  if (int rc = f->rt.setup(plan, cls, f->cap, f->stream)) return rc;
  if (int rc = f->rt.reset(1, f->stream)) return rc;
  {
    ProfScope ps(f, KG_KERNEL_RT_LOAD);
    const RtDev D = f->rt.dev();
    for (int c = 0; c < f->rt.ncls; ++c)
      KG_HIP(launch_rt_load(f->worklist(), D, f->rt.caps[c], c, f->rt.dense + f->rt.off[c],
                            f->rt.off[c + 1] - f->rt.off[c], f->R.b, f->offset, f->rt.G - 1,
                            f->stream));
  }
  f->rt_mode = true;
  f->rt_replan = false;
  return KG_OK;
}

// Moves the tile-resident read buffer back to the bin-grouped layout.
int ensure_atomic(kg_field* f) {
  if (!f->rt_mode) return KG_OK;
  if (int rc = rt_check_flags(f)) return rc;
  KG_HIP(launch_fill_u32(f->scratch, 0u, 1, f->stream));
  KG_HIP(launch_rt_unload(f->worklist(), f->rt.dev(), f->A.b, f->scratch, 0, f->g.ntiles,
                          f->stream));
  uint32_t n = 0;
  KG_HIP(hipMemcpyAsync(&n, f->scratch, sizeof(n), hipMemcpyDeviceToHost, f->stream));
  KG_HIP(hipStreamSynchronize(f->stream));
  f->rt_mode = false;
  f->nA = n;
  f->nW = 0;
  f->w_aligned = false;
  return lazy_update_impl(f, nullptr);
}

// One tile-resident schedule step: one launch per tile class.
hipError_t rt_launch_step(kg_field* f, const WorkList& wl, const RtDev& D, uint32_t koff) {
  return rt_launch_classes(wl, D, f->rt, koff, f->stream);
}

// `nsteps` tile-resident schedule steps (graph replays of 8 steps when
// use_graph is set and profiling is off).  Step k of a batch runs schedule
// step step_base + *gstep + k; one 1-thread kernel then advances *gstep by
// the batch size, so no work-group waits on another.
int rt_steps(kg_field* f, uint32_t nsteps, int use_graph) {
  const RtDev D = f->rt.dev();
  const WorkList wl = f->worklist();
  uint32_t left = nsteps;
  constexpr int kChunk = 8;
  if (use_graph && !f->prof_on && left >= kChunk) {
    hipGraphExec_t exec = nullptr;
    for (auto& gch : f->rt.graphs)
      if (gch.first == kChunk) exec = gch.second;
    if (!exec) {
      hipGraph_t graph = nullptr;
      KG_HIP(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipSuccess;
      for (int k = 0; k < kChunk && e == hipSuccess; ++k) e = rt_launch_step(f, wl, D, (uint32_t)k);
      if (e == hipSuccess) e = launch_add_u64(D.gstep, kChunk, f->stream);
      const hipError_t e2 = hipStreamEndCapture(f->stream, &graph);
      if (e != hipSuccess || e2 != hipSuccess)
        return set_error(KG_E_HIP, "graph capture: %s",
                         hipGetErrorString(e != hipSuccess ? e : e2));
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return set_error(KG_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
      f->rt.graphs.push_back({kChunk, exec});
    }
    while (left >= (uint32_t)kChunk) {
      KG_HIP(hipGraphLaunch(exec, f->stream));
      left -= kChunk;
    }
  }
  if (left > 0) {
    for (uint32_t k = 0; k < left; ++k) {
      ProfScope ps(f, KG_KERNEL_RT_STEP);
      KG_HIP(rt_launch_step(f, wl, D, k));
    }
    KG_HIP(launch_add_u64(D.gstep, left, f->stream));
  }
  f->rt.G += nsteps;
  f->rt_steps_run += nsteps;
  return KG_OK;
}

}  // namespace

extern "C" {

int kg_field_create(const kg_field_desc* desc, kg_field** out) {
  clear_error();
  if (!desc || !out) return set_error(KG_E_INVALID, "desc and out must not be NULL");
  *out = nullptr;
  if (desc->capacity == 0 || desc->capacity > (1u << 31))
    return set_error(KG_E_INVALID, "capacity must be in [1, 2^31]");
  if (desc->bin_rows & (desc->bin_rows - 1))
    return set_error(KG_E_INVALID, "bin_rows must be 0 or a power of two");
  kg_field* f = new kg_field();
  f->desc = *desc;
  f->cap = desc->capacity;
  std::string why;
  if ((desc->tile_cols & (desc->tile_cols - 1)) || (desc->tile_rows & (desc->tile_rows - 1))) {
    delete f;
    return set_error(KG_E_INVALID, "tile_cols and tile_rows must be 0 or powers of two");
  }
  if (!make_geom_rt(desc->width, desc->height, desc->discretization, desc->toroidal,
                    desc->capacity, desc->tile_cols, desc->tile_rows, desc->bin_rows, &f->g,
                    &why, true)) {
    delete f;
    return set_error(KG_E_INVALID, "%s", why.c_str());
  }
  f->lds_cap = clamp_lds_cap(pick_lds_cap(f->g, desc->capacity));
  {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 8192)
      f->lds_limit = (size_t)v - 6144;  // static LDS of the tile-resident kernel
  }
  f->h_rep.g = f->g;
  f->h_rep.bin_base = 0;
  kg_flock_params dflt{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0.7f, 10.0f, 0u, 0ull};
  set_rep_params(&f->h_rep, dflt);

  auto fail = [&](hipError_t e) {
    kg_field_destroy(f);
    return set_error(KG_E_NOMEM, "device allocation failed: %s", hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
  if ((e = f->R.alloc(f->cap)) != hipSuccess) return fail(e);
  if ((e = f->R2.alloc(f->cap)) != hipSuccess) return fail(e);
  if ((e = f->W.alloc(f->cap)) != hipSuccess) return fail(e);
  if ((e = f->A.alloc(f->cap)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->gflag, f->g.ntiles + 2u)) != hipSuccess) return fail(e);
  const uint32_t nb = f->g.nbins;
  const uint32_t nscan = (nb + kScanTile - 1) / kScanTile;
  if ((e = dmalloc(&f->count, nb)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->offset, nb + 1)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->offset2, nb + 1)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->cursor, nb)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->ss.partial, nscan + 1)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->err, 1)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->gstep, 1)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->zero64, 1)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->prefix, 4)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->d_rep, 1)) != hipSuccess) return fail(e);
  if ((e = dmalloc(&f->scratch, 4)) != hipSuccess) return fail(e);
  f->gather_ok = rebin_gather_enabled() && (uint32_t)(f->g.tx * f->g.tr) <= kGatherMaxBins;
  if (f->gather_ok) {
    if ((e = dmalloc(&f->outbox.n, f->g.ntiles)) != hipSuccess) return fail(e);
    if ((e = dmalloc(&f->outbox.slot, (size_t)f->g.ntiles * kOutboxCap)) != hipSuccess)
      return fail(e);
    if ((e = dmalloc(&f->gate, 4)) != hipSuccess) return fail(e);
    if ((e = hipMemset(f->gate, 0, 4 * sizeof(uint32_t))) != hipSuccess) return fail(e);
    if ((e = dmalloc(&f->count2, nb)) != hipSuccess) return fail(e);
    if ((e = hipMemset(f->count2, 0, nb * sizeof(uint32_t))) != hipSuccess) return fail(e);
    if ((e = hipHostMalloc((void**)&f->seen, 2 * sizeof(uint32_t))) != hipSuccess)
      return fail(e);
    f->seen[0] = f->seen[1] = 0;
    const uint32_t nscan = (f->g.nbins + kScanTile - 1) / kScanTile;
    for (int k = 0; k < 2; ++k)
      if ((e = dmalloc(&f->tsum[k], nscan)) != hipSuccess) return fail(e);
  }
  // one-time clears (never inside a captured graph)
  const std::pair<void*, size_t> clears[] = {
      {f->count, sizeof(uint32_t) * nb},          {f->offset, sizeof(uint32_t) * (nb + 1)},
      {f->offset2, sizeof(uint32_t) * (nb + 1)},  {f->ss.partial, sizeof(uint32_t) * (nscan + 1)},
      {f->err, sizeof(uint32_t)},                 {f->gstep, sizeof(uint64_t)},
      {f->gflag, sizeof(uint32_t) * (f->g.ntiles + 2u)},
      {f->zero64, sizeof(uint64_t)}};
  for (const auto& c : clears)
    if ((e = hipMemset(c.first, 0, c.second)) != hipSuccess) return fail(e);
  const uint32_t pre[4] = {0u, f->g.ntiles, 0u, 0u};
  if ((e = hipMemcpy(f->prefix, pre, sizeof(pre), hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(f->d_rep, &f->h_rep, sizeof(RepDev), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e);
  *out = f;
  return KG_OK;
}

void kg_field_destroy(kg_field* f) {
  if (!f) return;
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  for (auto& kv : f->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& pe : f->ev_pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (hipEvent_t e : f->ev_pool) (void)hipEventDestroy(e);
  f->rt.release();
  f->R.release();
  f->R2.release();
  f->W.release();
  f->A.release();
  (void)hipFree(f->count);
  (void)hipFree(f->offset);
  (void)hipFree(f->offset2);
  (void)hipFree(f->cursor);
  (void)hipFree(f->ss.partial);
  (void)hipFree(f->err);
  (void)hipFree(f->gstep);
  (void)hipFree(f->zero64);
  (void)hipFree(f->prefix);
  (void)hipFree(f->d_rep);
  (void)hipFree(f->scratch);
  (void)hipFree(f->tcount);
  (void)hipFree(f->gflag);
  (void)hipFree(f->outbox.n);
  (void)hipFree(f->outbox.slot);
  (void)hipFree(f->gate);
  (void)hipFree(f->tsum[0]);
  (void)hipFree(f->tsum[1]);
  (void)hipFree(f->count2);
  if (f->seen) (void)hipHostFree(f->seen);
  if (f->stream) (void)hipStreamDestroy(f->stream);
  delete f;
}

int kg_set_object_locations(kg_field* f, const uint32_t* ids, const float* x, const float* y,
                            const float* hx, const float* hy, size_t n) {
  clear_error();
  if (!f) return set_error(KG_E_INVALID, "field is NULL");
  if (int rc = check_poison(f)) return rc;
  if (int rc = ensure_atomic(f)) return rc;
  if (n == 0) return KG_OK;
  if (!ids || !x || !y) return set_error(KG_E_INVALID, "ids, x and y must not be NULL");
  if (f->w_aligned)
    return set_error(KG_E_STATE,
                     "the write buffer holds the output of a flock step; call kg_lazy_update "
                     "first");
  if ((uint64_t)f->nA + n > f->cap)
    return set_error(KG_E_CAPACITY, "write buffer capacity %u exceeded", f->cap);
  // range check on the host: the reference panics on an out-of-range bag
  for (size_t i = 0; i < n; ++i) {
    int32_t cx, cy;
    discretize(f->g, x[i], y[i], &cx, &cy);
    if (!(cx >= 0 && cx < f->g.dw && cy >= 0 && cy < f->g.dh))
      return set_error(KG_E_RANGE, "location (%g, %g) is outside the bag grid", (double)x[i],
                       (double)y[i]);
  }
  const size_t bytes = n * sizeof(float);
  const uint32_t o = f->nA;
  KG_HIP(hipMemcpyAsync(f->A.b.x + o, x, bytes, hipMemcpyHostToDevice, f->stream));
  KG_HIP(hipMemcpyAsync(f->A.b.y + o, y, bytes, hipMemcpyHostToDevice, f->stream));
  KG_HIP(hipMemcpyAsync(f->A.b.id + o, ids, n * sizeof(uint32_t), hipMemcpyHostToDevice, f->stream));
  if (hx) KG_HIP(hipMemcpyAsync(f->A.b.hx + o, hx, bytes, hipMemcpyHostToDevice, f->stream));
  else KG_HIP(hipMemsetAsync(f->A.b.hx + o, 0, bytes, f->stream));
  if (hy) KG_HIP(hipMemcpyAsync(f->A.b.hy + o, hy, bytes, hipMemcpyHostToDevice, f->stream));
  else KG_HIP(hipMemsetAsync(f->A.b.hy + o, 0, bytes, f->stream));
  KG_HIP(hipStreamSynchronize(f->stream));  // host buffers are only borrowed
  f->nA += (uint32_t)n;
  if (!f->dens_check) f->nagents += n;
  return KG_OK;
}

int kg_init_uniform(kg_field* f, uint32_t n, uint64_t seed, int heading_mode) {
  clear_error();
  if (!f) return set_error(KG_E_INVALID, "field is NULL");
  if (int rc = check_poison(f)) return rc;
  if (int rc = ensure_atomic(f)) return rc;
  if (heading_mode != 0 && heading_mode != 1)
    return set_error(KG_E_INVALID, "heading_mode must be 0 or 1");
  if (f->w_aligned)
    return set_error(KG_E_STATE, "the write buffer holds the output of a flock step");
  if ((uint64_t)f->nA + n > f->cap)
    return set_error(KG_E_CAPACITY, "write buffer capacity %u exceeded", f->cap);
  KG_HIP(launch_init_uniform(f->A.b, f->nA, n, f->g.w, f->g.h, seed, heading_mode, f->stream));
  f->nA += n;
  if (!f->dens_check) f->nagents += n;
  return KG_OK;
}

int kg_remove_object_location(kg_field* f, uint32_t id, float x, float y) {
  clear_error();
  if (!f) return set_error(KG_E_INVALID, "field is NULL");
  if (int rc = check_poison(f)) return rc;
  if (int rc = ensure_atomic(f)) return rc;
  if (f->w_aligned) {
    KG_HIP(launch_materialize(f->R.b, f->W.b, f->nW, f->A.b, f->stream));
    f->nA = f->nW;
    f->nW = 0;
    f->w_aligned = false;
  }
  if (f->nA == 0) return KG_OK;
  KG_HIP(launch_remove(f->d_rep, f->A.b, f->nA, id, x, y, f->scratch, f->stream));
  uint32_t after = 0;
  KG_HIP(hipMemcpyAsync(&after, f->scratch, sizeof(after), hipMemcpyDeviceToHost, f->stream));
  KG_HIP(hipStreamSynchronize(f->stream));
  if (!f->dens_check) f->nagents -= (f->nA - after);
  f->nA = after;
  return KG_OK;
}

int kg_lazy_update(kg_field* f) {
  clear_error();
  if (!f) return set_error(KG_E_INVALID, "field is NULL");
  if (int rc = check_poison(f)) return rc;
  if (int rc = ensure_atomic(f)) return rc;
  return lazy_update_impl(f, nullptr);
}

static int query_common(kg_field* f, float x, float y, float dist, int relax, uint32_t* ids,
                        size_t cap, size_t* n_out) {
  if (!f || !n_out) return set_error(KG_E_INVALID, "field and n_out must not be NULL");
  if (int rc = sync_and_check(f)) return rc;
  uint32_t* d_out = nullptr;
  const uint32_t dcap = (uint32_t)std::min<size_t>(cap, 0xFFFFFFFFu);
  KG_HIP(dmalloc(&d_out, dcap));
  hipError_t e = launch_query(f->d_rep, f->R.b, f->offset, x, y, dist, sq_threshold(dist), relax,
                              d_out, dcap, f->scratch, f->stream);
  uint32_t n = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&n, f->scratch, 4, hipMemcpyDeviceToHost, f->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
  if (e == hipSuccess && ids && cap)
    e = hipMemcpy(ids, d_out, std::min<size_t>(n, cap) * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(d_out);
  if (e != hipSuccess) return set_error(KG_E_HIP, "query failed: %s", hipGetErrorString(e));
  *n_out = n;
  if (n > cap) return set_error(KG_E_CAPACITY, "%u neighbours do not fit in %zu", n, cap);
  return KG_OK;
}

int kg_get_neighbors_within_distance(kg_field* f, float x, float y, float dist, uint32_t* ids,
                                     size_t cap, size_t* n_out) {
  clear_error();
  return query_common(f, x, y, dist, 0, ids, cap, n_out);
}

int kg_get_neighbors_within_relax_distance(kg_field* f, float x, float y, float dist,
                                           uint32_t* ids, size_t cap, size_t* n_out) {
  clear_error();
  return query_common(f, x, y, dist, 1, ids, cap, n_out);
}

int kg_get_objects(kg_field* f, float x, float y, int unbuffered, uint32_t* ids, size_t cap,
                   size_t* n_out) {
  clear_error();
  if (!f || !n_out) return set_error(KG_E_INVALID, "field and n_out must not be NULL");
  if (int rc = sync_and_check(f)) return rc;
  uint32_t* d_out = nullptr;
  const uint32_t dcap = (uint32_t)std::min<size_t>(cap, 0xFFFFFFFFu);
  KG_HIP(dmalloc(&d_out, dcap));
  hipError_t e;
  if (!unbuffered) {
    e = launch_objects_read(f->d_rep, f->R.b, f->offset, x, y, d_out, dcap, f->scratch, f->stream);
  } else if (f->w_aligned) {
    Buf wb = f->W.b;
    wb.id = f->R.b.id;
    e = launch_objects_write(f->d_rep, wb, f->nW, 1, x, y, d_out, dcap, f->scratch, f->stream);
  } else {
    e = launch_objects_write(f->d_rep, f->A.b, f->nA, 0, x, y, d_out, dcap, f->scratch, f->stream);
  }
  uint32_t n = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&n, f->scratch, 4, hipMemcpyDeviceToHost, f->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
  if (e == hipSuccess && ids && cap)
    e = hipMemcpy(ids, d_out, std::min<size_t>(n, cap) * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(d_out);
  if (e != hipSuccess) return set_error(KG_E_HIP, "get_objects failed: %s", hipGetErrorString(e));
  *n_out = n;
  if (n > cap) return set_error(KG_E_CAPACITY, "%u objects do not fit in %zu", n, cap);
  return KG_OK;
}

int kg_num_objects_at_location(kg_field* f, float x, float y, size_t* n) {
  clear_error();
  const int rc = kg_get_objects(f, x, y, 0, nullptr, 0, n);
  if (rc == KG_E_CAPACITY) {
    clear_error();
    return KG_OK;
  }
  return rc;
}

int kg_num_empty_bags(kg_field* f, size_t* n) {
  clear_error();
  if (!f || !n) return set_error(KG_E_INVALID, "field and n must not be NULL");
  if (int rc = sync_and_check(f)) return rc;
  KG_HIP(hipMemsetAsync(f->scratch, 0, 4, f->stream));
  KG_HIP(launch_count_cells(f->d_rep, f->R.b, f->offset, f->nR, f->scratch, f->stream));
  uint32_t distinct = 0;
  KG_HIP(hipMemcpyAsync(&distinct, f->scratch, 4, hipMemcpyDeviceToHost, f->stream));
  KG_HIP(hipStreamSynchronize(f->stream));
  *n = (size_t)f->g.dw * (size_t)f->g.dh - distinct;
  return KG_OK;
}

// Empty bags of the read buffer on the device: *count of them, and -- when
// xy is not NULL -- their not_discretize positions in bag index order in
// device arrays *xy (x) and *xy + count (y), which the caller frees.
static int empty_bags_dev(kg_field* f, uint32_t* count, float** xy) {
  const uint64_t ncells = (uint64_t)f->g.dw * (uint64_t)f->g.dh;
  const uint64_t nw64 = (ncells + 31) / 32;
  if (nw64 >= 0xFFFFFFFFull) return set_error(KG_E_INVALID, "bag grid too large");
  const uint32_t nwords = (uint32_t)nw64;
  const uint32_t nscan = (nwords + kScanTile - 1) / kScanTile;
  uint32_t *occ = nullptr, *cnt = nullptr, *off = nullptr, *cur = nullptr, *part = nullptr;
  float* out = nullptr;
  uint32_t nR = 0;
  hipError_t e = hipMemcpy(&nR, f->offset + f->g.nbins, 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = dmalloc(&occ, nwords);
  if (e == hipSuccess) e = dmalloc(&cnt, nwords);
  if (e == hipSuccess) e = dmalloc(&off, nwords + 1);
  if (e == hipSuccess) e = dmalloc(&cur, nwords);
  if (e == hipSuccess) e = dmalloc(&part, nscan + 1);
  if (e == hipSuccess) e = launch_fill_u32(occ, 0u, nwords, f->stream);
  if (e == hipSuccess) e = launch_mark_bags(f->d_rep, f->R.b, nR, occ, f->stream);
  if (e == hipSuccess) e = launch_count_empty(occ, nwords, ncells, cnt, f->stream);
  ScanState ss;
  ss.partial = part;
  if (e == hipSuccess) e = launch_scan(cnt, off, cur, nwords, ss, nullptr, f->stream);
  uint32_t total = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&total, off + nwords, 4, hipMemcpyDeviceToHost, f->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
  if (e == hipSuccess && xy) {
    e = dmalloc(&out, 2 * (size_t)total);
    if (e == hipSuccess)
      e = launch_emit_empty(f->d_rep, occ, nwords, ncells, off, out, out + total, f->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
  }
  (void)hipFree(occ);
  (void)hipFree(cnt);
  (void)hipFree(off);
  (void)hipFree(cur);
  (void)hipFree(part);
  if (e != hipSuccess) {
    (void)hipFree(out);
    return set_error(KG_E_HIP, "empty bags: %s", hipGetErrorString(e));
  }
  *count = total;
  if (xy) *xy = out;
  return KG_OK;
}

int kg_get_empty_bags(kg_field* f, float* x, float* y, size_t cap, size_t* n_out) {
  clear_error();
  if (!f || !n_out) return set_error(KG_E_INVALID, "field and n_out must not be NULL");
  if (int rc = sync_and_check(f)) return rc;
  uint32_t total = 0;
  float* xy = nullptr;
  if (int rc = empty_bags_dev(f, &total, &xy)) return rc;
  const size_t m = std::min<size_t>(total, cap);
  hipError_t e = hipSuccess;
  if (m && x) e = hipMemcpy(x, xy, m * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && m && y) e = hipMemcpy(y, xy + total, m * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(xy);
  if (e != hipSuccess) return set_error(KG_E_HIP, "get_empty_bags: %s", hipGetErrorString(e));
  *n_out = total;
  if (total > cap) return set_error(KG_E_CAPACITY, "%u empty bags do not fit in %zu", total, cap);
  return KG_OK;
}

int kg_get_random_empty_bag(kg_field* f, uint64_t seed, float* x, float* y, int* found) {
  clear_error();
  if (!f || !x || !y || !found) return set_error(KG_E_INVALID, "arguments must not be NULL");
  if (int rc = sync_and_check(f)) return rc;
  uint32_t total = 0;
  float* xy = nullptr;
  if (int rc = empty_bags_dev(f, &total, &xy)) return rc;
  *found = total != 0;
  if (!total) {
    (void)hipFree(xy);
    return KG_OK;
  }
  // index = floor(u * total / 2^64) for the 64-bit word u of Philox(ctr =
  // {0, 0, 0, "BAGS"}, key = seed)
  const u32x4 o = philox4x32_10(0u, 0u, 0u, 0x42414753u, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t u = ((uint64_t)o.v[1] << 32) | o.v[0];
  const uint64_t idx = (uint64_t)(((unsigned __int128)u * total) >> 64);
  hipError_t e = hipMemcpy(x, xy + idx, sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(y, xy + total + idx, sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(xy);
  if (e != hipSuccess) return set_error(KG_E_HIP, "get_random_empty_bag: %s", hipGetErrorString(e));
  return KG_OK;
}

int kg_iter_objects(kg_field* f, int unbuffered, kg_object_fn fn, void* user) {
  clear_error();
  if (!f || !fn) return set_error(KG_E_INVALID, "field and fn must not be NULL");
  size_t n = 0;
  const int buffer = unbuffered ? 1 : 0;
  int rc = kg_read_agents(f, buffer, nullptr, nullptr, nullptr, nullptr, nullptr, 0, &n);
  if (rc != KG_OK && rc != KG_E_CAPACITY) return rc;
  std::vector<uint32_t> id(n);
  std::vector<float> x(n), y(n), hx(n), hy(n);
  if ((rc = kg_read_agents(f, buffer, id.data(), x.data(), y.data(), hx.data(), hy.data(), n, &n)))
    return rc;
  // bag index order (cx * dh + cy); inside a bag the read buffer and the
  // flock step's output are in id order, set_object_location entries in
  // insertion order
  const bool by_id = buffer == 0 || f->w_aligned;
  std::vector<uint64_t> key(n);
  for (size_t i = 0; i < n; ++i) {
    int32_t cx, cy;
    discretize(f->g, x[i], y[i], &cx, &cy);
    key[i] = (uint64_t)cx * (uint64_t)f->g.dh + (uint64_t)cy;
  }
  std::vector<size_t> ord(n);
  for (size_t i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
    if (key[a] != key[b]) return key[a] < key[b];
    return by_id && id[a] < id[b];
  });
  for (size_t k : ord) {
    const float lx = (float)(int32_t)(key[k] / (uint64_t)f->g.dh) * f->g.d;
    const float ly = (float)(int32_t)(key[k] % (uint64_t)f->g.dh) * f->g.d;
    fn(user, lx, ly, id[k], x[k], y[k], hx[k], hy[k]);
  }
  return KG_OK;
}

int kg_nagents(kg_field* f, size_t* n) {
  clear_error();
  if (!f || !n) return set_error(KG_E_INVALID, "field and n must not be NULL");
  *n = f->nagents;
  return KG_OK;
}

int kg_count(kg_field* f, int buffer, size_t* n) {
  clear_error();
  if (!f || !n) return set_error(KG_E_INVALID, "field and n must not be NULL");
  if (buffer == 0) {
    if (int rc = sync_and_check(f)) return rc;
    uint32_t total = 0;
    KG_HIP(hipMemcpy(&total, f->offset + f->g.nbins, 4, hipMemcpyDeviceToHost));
    *n = total;
  } else {
    *n = f->w_aligned ? f->nW : f->nA;
  }
  return KG_OK;
}

int kg_neighbor_sets(kg_field* f, float dist, uint32_t* row_ptr, uint32_t* ids, size_t cap,
                     size_t* total) {
  clear_error();
  if (!f || !row_ptr || !total) return set_error(KG_E_INVALID, "field, row_ptr, total required");
  if (int rc = sync_and_check(f)) return rc;
  uint32_t nR = 0;
  KG_HIP(hipMemcpy(&nR, f->offset + f->g.nbins, 4, hipMemcpyDeviceToHost));
  const float T = sq_threshold(dist);
  uint32_t* d_cnt = nullptr;
  KG_HIP(dmalloc(&d_cnt, nR + 1));
  std::vector<uint32_t> cnt(nR);
  hipError_t e = launch_neighbor_count(f->d_rep, f->R.b, f->offset, nR, dist, T, d_cnt, f->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
  if (e == hipSuccess && nR)
    e = hipMemcpy(cnt.data(), d_cnt, nR * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    (void)hipFree(d_cnt);
    return set_error(KG_E_HIP, "neighbour count failed: %s", hipGetErrorString(e));
  }
  uint64_t run = 0;
  for (uint32_t i = 0; i < nR; ++i) {
    row_ptr[i] = (uint32_t)run;
    run += cnt[i];
  }
  row_ptr[nR] = (uint32_t)run;
  *total = run;
  if (run > cap || run > 0xFFFFFFFFull) {
    (void)hipFree(d_cnt);
    return set_error(KG_E_CAPACITY, "%llu neighbour ids do not fit in %zu",
                     (unsigned long long)run, cap);
  }
  uint32_t* d_ids = nullptr;
  e = hipMemcpy(d_cnt, row_ptr, (nR + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = dmalloc(&d_ids, run);
  if (e == hipSuccess)
    e = launch_neighbor_fill(f->d_rep, f->R.b, f->offset, nR, dist, T, d_cnt, d_ids, f->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
  if (e == hipSuccess && run)
    e = hipMemcpy(ids, d_ids, run * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(d_cnt);
  (void)hipFree(d_ids);
  if (e != hipSuccess) return set_error(KG_E_HIP, "neighbour fill failed: %s", hipGetErrorString(e));
  return KG_OK;
}

int kg_flock_step(kg_field* f, const kg_flock_params* p, uint64_t step) {
  clear_error();
  if (!f) return set_error(KG_E_INVALID, "field is NULL");
  if (int rc = check_poison(f)) return rc;
  if (int rc = ensure_atomic(f)) return rc;
  if (f->nW || f->nA || f->w_aligned)
    return set_error(KG_E_STATE, "the write buffer must be empty before a flock step");
  if (int rc = set_params(f, p, step)) return rc;
  return flock_impl(f, f->zero64);
}

static int run_direct_step(kg_field* f) {
  if (int rc = flock_impl(f, f->gstep)) return rc;
  return lazy_update_impl(f, f->gstep);
}

int kg_schedule_run(kg_field* f, const kg_flock_params* p, uint64_t first_step, uint32_t nsteps,
                    int use_graph) {
  clear_error();
  if (!f) return set_error(KG_E_INVALID, "field is NULL");
  if (int rc = check_poison(f)) return rc;
  if (nsteps == 0) return KG_OK;
  if (f->w_aligned) return set_error(KG_E_STATE, "a flock step output is pending");
  // schedule.rs:349-351: step 0 first calls state.update(0)
  if (first_step == 0) {
    if (int rc = ensure_atomic(f)) return rc;
    if (int rc = lazy_update_impl(f, nullptr)) return rc;
  }
  if (f->nA) return set_error(KG_E_STATE, "the write buffer must be empty");
  if (!p) return set_error(KG_E_INVALID, "params is NULL");
  if ((use_graph & KG_RUN_BINNED) && (use_graph & KG_RUN_TILE_RESIDENT))
    return set_error(KG_E_INVALID, "KG_RUN_BINNED and KG_RUN_TILE_RESIDENT exclude each other");
  {
    RepDev probe = f->h_rep;
    set_rep_params(&probe, *p);
    // by default small fields take the one-launch tile-resident step (fewer
    // launches per step) and large ones the three-kernel path (more of its
    // work-groups fit a CU): measured at constant density on MI355X, the two
    // cross between 16k and 64k birds
    constexpr uint32_t kRtDefaultMaxAgents = 32768;
    const bool want_rt = (use_graph & KG_RUN_TILE_RESIDENT) ||
                         (!(use_graph & KG_RUN_BINNED) && f->nR < kRtDefaultMaxAgents);
    if (want_rt && probe.lds_ok) {
      // tile-resident path: the class launches of each step; the capacity
      // flags are read every kRtCheckEvery steps (32 measured ~1.5 % faster
      // than 64 at 1M boids: deferrals are caught sooner) and the layout
      // re-planned when a tile overflowed
      constexpr uint32_t kRtCheckEvery = 32;
      if (f->rt_mode)
        if (int rc = rt_check_flags(f)) return rc;
      if (f->rt_mode && (f->rt_replan || probe.disc_dist != f->h_rep.disc_dist))
        if (int rc = ensure_atomic(f)) return rc;
      if (int rc = set_params(f, p, 0)) return rc;
      if (!f->rt_mode) {
        f->rt_unsuitable = false;
        if (int rc = rt_enter(f)) return rc;
      }
      while (f->rt_mode && nsteps > 0) {
        const uint32_t chunk = std::min(nsteps, kRtCheckEvery);
        // schedule step of the next launch = step_base + *gstep
        if (int rc = set_params(f, p, first_step - f->rt.G)) return rc;
        if (int rc = rt_steps(f, chunk, use_graph & KG_RUN_GRAPH)) return rc;
        first_step += chunk;
        nsteps -= chunk;
        if (nsteps == 0) break;
        if (int rc = rt_check_flags(f)) return rc;
        if (f->rt_replan) {
          if (int rc = ensure_atomic(f)) return rc;
          if (int rc = set_params(f, p, 0)) return rc;
          if (int rc = rt_enter(f)) return rc;
        }
      }
      if (nsteps == 0) return KG_OK;
      // the plan did not fit: the remaining steps run bin-grouped
    }
  }
  if (int rc = ensure_atomic(f)) return rc;
  if (int rc = set_params(f, p, first_step)) return rc;
  // a field whose steps fail the gather (tiles too full, long jumps) pays
  // for it and for the scatter and the sort: it rebins by scatter and sort
  // from then on.  The first gather step is checked at once (one
  // synchronisation per field), later ones from the count copied after each
  // run, which may lag a run behind: it is only a hint.
  // Likewise a field whose tiles took the global path (counted by k_flock)
  // goes back to launching k_flock_global.
  const volatile uint32_t* seen = f->seen;
  if (f->gather_ok && f->gather_steps_copied >= 16 &&
      16ull * seen[0] >= f->gather_steps_copied)
    f->gather_ok = false;
  if (f->global_inline && seen[1] != 0) f->global_inline = false;
  KG_HIP(hipMemsetAsync(f->gstep, 0, sizeof(uint64_t), f->stream));
  uint32_t left = nsteps;
  if (f->gather_ok && !f->gather_probed) {
    if (int rc = run_direct_step(f)) return rc;
    --left;
    uint32_t counts[2] = {0, 0};
    KG_HIP(hipMemcpyAsync(&counts[0], f->gate + 2, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          f->stream));
    KG_HIP(hipMemcpyAsync(&counts[1], f->gflag + f->g.ntiles + 1, sizeof(uint32_t),
                          hipMemcpyDeviceToHost, f->stream));
    KG_HIP(hipStreamSynchronize(f->stream));
    f->gather_probed = true;
    if (counts[0]) f->gather_ok = false;
    f->global_inline = counts[1] == 0;
  }
  // steps whose counts go straight into the scan's tile sums: gathered
  // steps on a field without global-path tiles (their atomics would crowd
  // one sum), whose tiles fit the LDS histogram and whole scan tiles
  const uint32_t nbt = (uint32_t)(f->g.tx * f->g.tr);
  const uint32_t nscan = (f->g.nbins + kScanTile - 1) / kScanTile;
  f->summed = f->gather_ok && f->global_inline && nscan <= kScanInlineTiles &&
              nbt <= kGatherMaxBins && (uint32_t)kScanTile % nbt == 0;
  if (f->summed) {
    for (int k = 0; k < 2; ++k)
      KG_HIP(hipMemsetAsync(f->tsum[k], 0, nscan * sizeof(uint32_t), f->stream));
    KG_HIP(hipMemsetAsync(f->gate, 0, 2 * sizeof(uint32_t), f->stream));
  }
  if ((use_graph & KG_RUN_GRAPH) && left >= 2 && !f->prof_on) {
    GraphKey key;
    std::memcpy(key.params, p, sizeof(kg_flock_params));
    key.parity = f->parity;
    key.gather = f->gather_ok ? 1 : 0;
    key.inline_global = f->global_inline ? 1 : 0;
    auto it = f->graphs.find(key);
    hipGraphExec_t exec = nullptr;
    if (it == f->graphs.end()) {
      // capture two steps: the read buffer returns to the same array pair
      hipGraph_t graph = nullptr;
      KG_HIP(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
      f->capturing = true;
      int rc = run_direct_step(f);
      if (rc == KG_OK) rc = run_direct_step(f);
      f->capturing = false;
      hipError_t e = hipStreamEndCapture(f->stream, &graph);
      if (rc) return rc;
      if (e != hipSuccess) return set_error(KG_E_HIP, "graph capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return set_error(KG_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
      f->graphs[key] = exec;
    } else {
      exec = it->second;
    }
    while (left >= 2) {
      KG_HIP(hipGraphLaunch(exec, f->stream));
      left -= 2;
    }
  }
  while (left > 0) {
    if (int rc = run_direct_step(f)) return rc;
    --left;
  }
  if (f->gather_ok) {
    f->gather_steps += nsteps;
    KG_HIP(hipMemcpyAsync(f->seen, f->gate + 2, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          f->stream));
    f->gather_steps_copied = f->gather_steps;
  }
  if (f->global_inline)
    KG_HIP(hipMemcpyAsync(f->seen + 1, f->gflag + f->g.ntiles + 1, sizeof(uint32_t),
                          hipMemcpyDeviceToHost, f->stream));
  if (f->summed) {
    // the last step's counts were read, not cleared: bins are counted from
    // zero in the main array outside a run
    KG_HIP(launch_fill_u32(f->counts(f->parity ^ 1), 0u, f->g.nbins, f->stream));
    f->summed = false;
  }
  return KG_OK;
}

int kg_read_agents(kg_field* f, int buffer, uint32_t* ids, float* x, float* y, float* hx,
                   float* hy, size_t cap, size_t* n) {
  clear_error();
  if (!f || !n) return set_error(KG_E_INVALID, "field and n must not be NULL");
  if (int rc = sync_and_check(f)) return rc;
  Buf src;
  uint32_t cnt;
  if (buffer == 0) {
    KG_HIP(hipMemcpy(&cnt, f->offset + f->g.nbins, 4, hipMemcpyDeviceToHost));
    src = f->R.b;
  } else if (f->w_aligned) {
    cnt = f->nW;
    src = f->W.b;
    src.id = f->R.b.id;
  } else {
    cnt = f->nA;
    src = f->A.b;
  }
  *n = cnt;
  if (cnt > cap) return set_error(KG_E_CAPACITY, "%u agents do not fit in %zu", cnt, cap);
  const size_t fb = (size_t)cnt * sizeof(float);
  if (ids) KG_HIP(hipMemcpy(ids, src.id, (size_t)cnt * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (x) KG_HIP(hipMemcpy(x, src.x, fb, hipMemcpyDeviceToHost));
  if (y) KG_HIP(hipMemcpy(y, src.y, fb, hipMemcpyDeviceToHost));
  if (hx) KG_HIP(hipMemcpy(hx, src.hx, fb, hipMemcpyDeviceToHost));
  if (hy) KG_HIP(hipMemcpy(hy, src.hy, fb, hipMemcpyDeviceToHost));
  return KG_OK;
}

int kg_field_sync(kg_field* f) {
  clear_error();
  if (!f) return set_error(KG_E_INVALID, "field is NULL");
  if (int rc = sync_only(f)) return rc;
  return rt_check_flags(f);
}

void* kg_field_stream(kg_field* f) { return f ? (void*)f->stream : nullptr; }

int kg_field_profile(kg_field* f, int enable) {
  clear_error();
  if (!f) return set_error(KG_E_INVALID, "field is NULL");
  if (int rc = prof_collect(f)) return rc;
  f->prof_on = enable != 0;
  for (int k = 0; k < KG_KERNEL_COUNT; ++k) {
    f->prof_ms[k] = 0.0;
    f->prof_n[k] = 0;
  }
  return KG_OK;
}

int kg_field_profile_read(kg_field* f, int kernel, double* total_ms, uint64_t* launches) {
  clear_error();
  if (!f || !total_ms || !launches) return set_error(KG_E_INVALID, "NULL argument");
  if (kernel < 0 || kernel >= KG_KERNEL_COUNT) return set_error(KG_E_INVALID, "unknown kernel");
  if (int rc = prof_collect(f)) return rc;
  *total_ms = f->prof_ms[kernel];
  *launches = f->prof_n[kernel];
  return KG_OK;
}

int kg_field_stats(kg_field* f, uint64_t out[16]) {
  clear_error();
  if (!f || !out) return set_error(KG_E_INVALID, "NULL argument");
  if (int rc = sync_only(f)) return rc;
  uint32_t fl = 0;
  if (f->rt.allocated) KG_HIP(hipMemcpy(&fl, f->rt.flags, sizeof(fl), hipMemcpyDeviceToHost));
  f->rt_flags_seen |= fl;
  const RtCaps& c = f->rt.caps[0];
  const uint64_t v[16] = {f->rt_mode,
                          f->rt_unsuitable,
                          f->rt_plans,
                          f->rt_flags_seen,
                          f->g.ntiles,
                          (uint64_t)f->g.tx,
                          (uint64_t)f->g.ty,
                          (uint64_t)f->g.hb,
                          (uint64_t)c.runs,
                          (uint64_t)c.own,
                          (uint64_t)c.extras,
                          f->rt.allocated ? rt_lds_bytes(c) : 0,
                          f->lds_limit,
                          f->rt.plan.max_scap,
                          f->rt.ndense,
                          f->rt_steps_run};
  std::memcpy(out, v, sizeof(v));
  return KG_OK;
}

}  // extern "C"
