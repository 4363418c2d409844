// kg_launch.hpp -- host-side launchers of the agent-step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "kg_rt.hpp"

namespace kg {

// Bytes of dynamic LDS the flock kernel needs for `cap` staged agents.
size_t flock_lds_bytes(int cap);

hipError_t launch_set_rep(RepDev* dst, const RepDev& v, hipStream_t s);
hipError_t launch_init_uniform(Buf A, uint32_t base, uint32_t n, float w, float h, uint64_t seed,
                               int heading_mode, hipStream_t s);
hipError_t launch_count_append(const WorkList& wl, Buf A, uint32_t* count, uint32_t* err,
                               hipStream_t s);
// One work-group: sums[0..n) becomes its exclusive prefix.
hipError_t launch_scan_prefix(uint32_t* sums, uint32_t n, hipStream_t s);
// gate (or null): this step's gate flag, cleared for the gather that follows
hipError_t launch_scan(uint32_t* count, uint32_t* offset, uint32_t* cursor, uint32_t nbins,
                       const ScanState& ss, uint64_t* gstep, hipStream_t s,
                       uint32_t* gate = nullptr);

// Moves every agent of W (aligned with R) to its new bin's cursor in R2.
hipError_t launch_scatter_aligned(const WorkList& wl, Buf R, Buf W, const uint32_t* offset_cur,
                                  uint32_t* cursor, Buf R2, hipStream_t s);
// The gather rebinning (one work-group per tile; tiles of at most
// kGatherMaxBins bins) of the flock step that filled the outboxes: R2 in
// bin order with every bin sorted, or gate.flag set for
// launch_rebin_fallback to redo it.  sc.count: the scan is done here too (offset_new is
// written); else offset_new holds the scan's offsets.
hipError_t launch_rebin_gather(const WorkList& wl, Buf R, Buf W, const uint32_t* offset_cur,
                               uint32_t* offset_new, const Outbox& ob, Buf R2, const Gate& gate,
                               const GatherScan& sc, hipStream_t s);
hipError_t launch_scatter_append(const WorkList& wl, Buf A, uint32_t* cursor, Buf R2,
                                 hipStream_t s);
// Sorts every bin of R (offsets `offset`) by (cell y, id); T is scratch of
// R's capacity whose contents are lost.
hipError_t launch_sort_bins(const WorkList& wl, Buf R, const uint32_t* offset, Buf T,
                            hipStream_t s);
// The scatter and the bin sort of a gathered step in one launch, run only if
// gate.flag is set (gate.redone counts such steps, gate.bar is the grid
// wait's counter, zeroed by the gather): the scatter as
// launch_scatter_aligned (count_down: cursor holds the step's bin counts,
// each counted down to zero from offset_new), then the sort of R2's bins by
// offset_new with T as scratch.  A grid wait that times out sets bit 2 of
// err.
hipError_t launch_rebin_fallback(const WorkList& wl, Buf R, Buf W, const uint32_t* offset_cur,
                                 uint32_t* cursor, Buf R2, const uint32_t* offset_new, Buf T,
                                 const Gate& gate, uint32_t* err, bool count_down,
                                 hipStream_t s);
// gflag: one word per tile plus two (the last counts global-path tiles and
// starts at zero), nagents: R's count; ob: the outboxes for the gather
// rebinning that follows, or none; inline_global: each global-path tile is
// stepped by its own work-group, and k_flock_global is not launched;
// tile_sum: the scan's tile sums to add the counts to (zero on entry; tiles
// of at most kGatherMaxBins bins, whole tiles per scan tile), or none;
// global_grid: k_flock_global's work-groups, each looping over runs of
// birds, or 0 for one per run.
hipError_t launch_flock(const WorkList& wl, Buf R, const uint32_t* offset, Buf W,
                        uint32_t* count_next, const uint64_t* gstep, uint32_t* err, int lds_cap,
                        uint32_t* gflag, uint32_t nagents, hipStream_t s,
                        const Outbox& ob = Outbox{}, bool inline_global = false,
                        uint32_t* tile_sum = nullptr, uint32_t global_grid = 0);
hipError_t launch_materialize(Buf R, Buf W, uint32_t n, Buf A, hipStream_t s);

// Tile-resident step (one launch per schedule step); with load != 0 the
// launch instead moves the bin-grouped read buffer (Ra, offset_a) into the
// tile-resident layout that step `load_g + 1` reads.
size_t rt_lds_bytes(const RtCaps& c);
hipError_t launch_rt_step(const WorkList& wl, const RtDev& D, const RtCaps& caps, uint32_t koff,
                          int cls, const uint32_t* list, uint32_t nlist, int mode, hipStream_t s);
hipError_t launch_tile_counts(const WorkList& wl, const uint32_t* offset, uint32_t* count,
                              hipStream_t s);
hipError_t launch_rt_load(const WorkList& wl, const RtDev& D, const RtCaps& caps, int cls,
                          const uint32_t* list, uint32_t nlist, Buf Ra, const uint32_t* offset_a,
                          uint64_t load_g, hipStream_t s);
hipError_t launch_rt_unload(const WorkList& wl, const RtDev& D, Buf A, uint32_t* rep_cursor,
                            uint32_t tile0, uint32_t ntiles, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s);

// Single-replica queries (API and parity use; not on the hot path).
hipError_t launch_query(const RepDev* rep, Buf R, const uint32_t* offset, float x, float y,
                        float dist, float T, int relax, uint32_t* out, uint32_t cap,
                        uint32_t* nout, hipStream_t s);
hipError_t launch_neighbor_count(const RepDev* rep, Buf R, const uint32_t* offset, uint32_t n,
                                 float dist, float T, uint32_t* counts, hipStream_t s);
hipError_t launch_neighbor_fill(const RepDev* rep, Buf R, const uint32_t* offset, uint32_t n,
                                float dist, float T, const uint32_t* row_ptr, uint32_t* ids,
                                hipStream_t s);
hipError_t launch_objects_read(const RepDev* rep, Buf R, const uint32_t* offset, float x, float y,
                               uint32_t* out, uint32_t cap, uint32_t* nout, hipStream_t s);
hipError_t launch_objects_write(const RepDev* rep, Buf B, uint32_t n, int by_id, float x, float y,
                                uint32_t* out, uint32_t cap, uint32_t* nout, hipStream_t s);
hipError_t launch_count_cells(const RepDev* rep, Buf R, const uint32_t* offset, uint32_t n,
                              uint32_t* distinct, hipStream_t s);
// Empty bags of the read buffer (bitmap `occ` of ceil(cells / 32) words,
// zeroed by the caller; `cnt`, `off` one word and one word + 1 per bitmap
// word; x, y receive off[nwords] positions).
hipError_t launch_mark_bags(const RepDev* rep, Buf R, uint32_t n, uint32_t* occ, hipStream_t s);
hipError_t launch_count_empty(const uint32_t* occ, uint32_t nwords, uint64_t ncells,
                              uint32_t* cnt, hipStream_t s);
hipError_t launch_emit_empty(const RepDev* rep, const uint32_t* occ, uint32_t nwords,
                             uint64_t ncells, const uint32_t* off, float* x, float* y,
                             hipStream_t s);
hipError_t launch_remove(const RepDev* rep, Buf A, uint32_t n, uint32_t id, float x, float y,
                         uint32_t* n_after, hipStream_t s);

}  // namespace kg
