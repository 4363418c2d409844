// kg_rt_state.cpp -- allocation and reset of the tile-resident layout.
#include "kg_rt_state.hpp"

namespace kg {

RtDev RtState::dev() const {
  RtDev D;
  D.meta = meta;
  for (int p = 0; p < 2; ++p) {
    D.cnt[p] = cnt[p];
    D.boff[p] = boff[p];
    D.stay[p] = stay[p].b;
    D.inbox[p] = inbox[p].b;
    D.spill[p] = spill[p].b;
    D.spill_rep[p] = spill_rep[p];
  }
  for (int k = 0; k < 3; ++k) {
    D.icnt[k] = icnt[k];
    D.scnt[k] = scnt + k;
  }
  D.spill_cap = (uint32_t)spill_alloc;
  D.flags = flags;
  D.gstep = gstep;
  D.ovf[0] = ovf;
  D.ovf[1] = ovf + (kRtOvfMax + 1);
  return D;
}

void RtState::drop_graphs() {
  for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
  graphs.clear();
}

void RtState::release() {
  drop_graphs();
  for (int k = 0; k < kRtMaxClasses; ++k) {
    if (aux[k]) (void)hipStreamDestroy(aux[k]);
    if (ev_done[k]) (void)hipEventDestroy(ev_done[k]);
  }
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  (void)hipFree(meta);
  (void)hipFree(dense);
  for (uint32_t* p : {cnt[0], cnt[1], boff[0], boff[1], icnt[0], icnt[1], icnt[2], scnt,
                      spill_rep[0], spill_rep[1], flags, ovf})
    (void)hipFree(p);
  (void)hipFree(gstep);
  for (int p = 0; p < 2; ++p) {
    stay[p].release();
    inbox[p].release();
    spill[p].release();
  }
  *this = RtState();
}

int RtState::setup(const RtPlan& pl, const RtClassPlan& cls, uint64_t spill_cap, hipStream_t s) {
  const uint64_t nt = pl.sbase.size();
  drop_graphs();
  if (!allocated || pl.stay_total > stay_alloc || pl.inbox_total > inbox_alloc ||
      pl.boff_total > boff_alloc || spill_cap > spill_alloc || nt > tiles_alloc) {
    // grow every array together, by half again what this plan needs (at
    // least what is held): ensembles re-plan as runs join, and each
    // reallocation frees and allocates every array
    auto grown = [](uint64_t need, uint64_t held) {
      return std::max<uint64_t>(need + need / 2 + 64, held);
    };
    const uint64_t st = grown(pl.stay_total, stay_alloc), ib = grown(pl.inbox_total, inbox_alloc);
    const uint64_t bo = grown(pl.boff_total, boff_alloc), sp = grown(spill_cap, spill_alloc);
    const uint64_t ti = grown(nt, tiles_alloc);
    release();
    stay_alloc = st;
    inbox_alloc = ib;
    boff_alloc = bo;
    spill_alloc = sp;
    tiles_alloc = ti;
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) {
      if (e == hipSuccess) e = r;
    };
    ok(dmalloc(&meta, tiles_alloc));
    ok(dmalloc(&dense, tiles_alloc));
    for (int p = 0; p < 2; ++p) {
      ok(dmalloc(&cnt[p], tiles_alloc));
      ok(dmalloc(&boff[p], boff_alloc));
      ok(stay[p].alloc(stay_alloc));
      ok(inbox[p].alloc(inbox_alloc));
      ok(spill[p].alloc(spill_alloc));
      ok(dmalloc(&spill_rep[p], spill_alloc));
    }
    for (int k = 0; k < 3; ++k) ok(dmalloc(&icnt[k], tiles_alloc));
    ok(dmalloc(&scnt, 3));
    ok(dmalloc(&flags, 2));
    ok(dmalloc(&ovf, 2 * (kRtOvfMax + 1)));
    ok(dmalloc(&gstep, 1));
    if (e != hipSuccess) {
      release();
      return set_error(KG_E_NOMEM, "tile-resident layout allocation failed: %s",
                       hipGetErrorString(e));
    }
    allocated = true;
  }
  if (!ev_fork) {
    KG_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    for (int k = 0; k < kRtMaxClasses; ++k) {
      KG_HIP(hipStreamCreateWithFlags(&aux[k], hipStreamNonBlocking));
      KG_HIP(hipEventCreateWithFlags(&ev_done[k], hipEventDisableTiming));
    }
  }
  plan = pl;
  ntiles = (uint32_t)nt;
  KG_HIP(hipMemcpyAsync(meta, plan.meta.data(), nt * sizeof(RtMeta), hipMemcpyHostToDevice, s));
  dense_host = cls.dense;
  ndense = (uint32_t)dense_host.size();
  ncls = cls.ncls;
  for (int k = 0; k < kRtMaxClasses; ++k) caps[k] = cls.caps[k];
  for (int k = 0; k <= kRtMaxClasses; ++k) off[k] = cls.off[k];
  if (ndense)
    KG_HIP(hipMemcpyAsync(dense, dense_host.data(), ndense * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
  // the plan vectors are members and stay unchanged until the copies ran
  KG_HIP(hipStreamSynchronize(s));
  return KG_OK;
}

int RtState::reset(uint64_t g, hipStream_t s) {
  for (int k = 0; k < 3; ++k) KG_HIP(launch_fill_u32(icnt[k], 0u, ntiles, s));
  KG_HIP(launch_fill_u32(scnt, 0u, 3, s));
  KG_HIP(launch_fill_u32(flags, 0u, 2, s));
  KG_HIP(launch_fill_u32(ovf, 0u, 2 * (kRtOvfMax + 1), s));
  KG_HIP(launch_set_u64(gstep, g, s));
  G = g;
  return KG_OK;
}

hipError_t rt_launch_classes(const WorkList& wl, const RtDev& D, const RtState& rt, uint32_t koff,
                             hipStream_t s) {
  hipError_t e;
  const int last = rt.ncls - 1;
  if (last == 0) return launch_rt_step(wl, D, rt.caps[0], koff, 0, rt.dense, 0, kRtModeFinal, s);
  // fork: classes 1.. on the auxiliary streams, class 0 on s
  if ((e = hipEventRecord(rt.ev_fork, s)) != hipSuccess) return e;
  for (int k = 1; k <= last; ++k) {
    if ((e = hipStreamWaitEvent(rt.aux[k], rt.ev_fork, 0)) != hipSuccess) return e;
    if ((e = launch_rt_step(wl, D, rt.caps[k], koff, k, rt.dense + rt.off[k],
                            rt.off[k + 1] - rt.off[k], k == last ? kRtModeFinal : kRtModeClass,
                            rt.aux[k])) != hipSuccess)
      return e;
    if ((e = hipEventRecord(rt.ev_done[k], rt.aux[k])) != hipSuccess) return e;
  }
  if ((e = launch_rt_step(wl, D, rt.caps[0], koff, 0, rt.dense, 0, kRtModeClass, s)) != hipSuccess)
    return e;
  // join the classes that may defer, run their deferred tiles with the last
  // class's capacities, then join the last class
  for (int k = 1; k < last; ++k)
    if ((e = hipStreamWaitEvent(s, rt.ev_done[k], 0)) != hipSuccess) return e;
  if ((e = launch_rt_step(wl, D, rt.caps[last], koff, last, rt.dense, 0, kRtModeDeferred, s)) !=
      hipSuccess)
    return e;
  return hipStreamWaitEvent(s, rt.ev_done[last], 0);
}

}  // namespace kg
