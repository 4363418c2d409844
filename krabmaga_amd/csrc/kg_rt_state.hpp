// kg_rt_state.hpp -- device memory of the tile-resident layout (kg_rt.hpp)
// and the host mirror of its step counter, shared by the plain field and
// the ensemble engine.
#pragma once

#include "kg_common.hpp"

namespace kg {

struct DevBuf {
  Buf b{};
  hipError_t alloc(size_t cap) {
    hipError_t e;
    if ((e = dmalloc(&b.x, cap)) != hipSuccess) return e;
    if ((e = dmalloc(&b.y, cap)) != hipSuccess) return e;
    if ((e = dmalloc(&b.hx, cap)) != hipSuccess) return e;
    if ((e = dmalloc(&b.hy, cap)) != hipSuccess) return e;
    return dmalloc(&b.id, cap);
  }
  void release() {
    (void)hipFree(b.x);
    (void)hipFree(b.y);
    (void)hipFree(b.hx);
    (void)hipFree(b.hy);
    (void)hipFree(b.id);
    b = Buf{};
  }
};

struct RtState {
  bool allocated = false;
  RtPlan plan;
  int ncls = 2;
  RtCaps caps[kRtMaxClasses]{};
  uint32_t* dense = nullptr;  // device tile lists of classes 1.., grouped by class
  uint32_t off[kRtMaxClasses + 1] = {};
  uint32_t ndense = 0;
  std::vector<uint32_t> dense_host;  // kept until the upload ran
  uint32_t ntiles = 0;
  uint64_t stay_alloc = 0, inbox_alloc = 0, boff_alloc = 0, spill_alloc = 0, tiles_alloc = 0;
  RtMeta* meta = nullptr;  // per tile
  uint32_t *cnt[2] = {nullptr, nullptr}, *boff[2] = {nullptr, nullptr};
  uint32_t* icnt[3] = {nullptr, nullptr, nullptr};
  uint32_t* scnt = nullptr;  // 3 counters
  DevBuf stay[2], inbox[2], spill[2];
  uint32_t* spill_rep[2] = {nullptr, nullptr};
  uint32_t* flags = nullptr;
  uint64_t* gstep = nullptr;
  uint32_t* ovf = nullptr;  // 2 x (kRtOvfMax + 1)
  uint64_t G = 0;  // value of *gstep once queued work has run
  std::vector<std::pair<int, hipGraphExec_t>> graphs;  // (steps, exec)
  // the class launches of a step run side by side on these streams
  hipStream_t aux[kRtMaxClasses] = {};
  hipEvent_t ev_fork = nullptr, ev_done[kRtMaxClasses] = {};

  RtDev dev() const;
  // (Re)allocates for `plan` (growing only) and uploads the plan tables and
  // the class split.
  int setup(const RtPlan& plan, const RtClassPlan& cls, uint64_t spill_cap, hipStream_t s);
  // Zeroes the counters and flags and sets the step counter to G.
  int reset(uint64_t G, hipStream_t s);
  void drop_graphs();
  void release();
};

hipError_t launch_set_u64(uint64_t* p, uint64_t v, hipStream_t s);

// One tile-resident schedule step, launch koff of a batch: the class
// launches of the plan in `rt` side by side (they touch disjoint tiles), then
// one launch for the tiles they deferred.
hipError_t rt_launch_classes(const WorkList& wl, const RtDev& D, const RtState& rt, uint32_t koff,
                             hipStream_t s);
hipError_t launch_add_u64(uint64_t* p, uint64_t v, hipStream_t s);

}  // namespace kg
