// kg_geom.hpp -- host-side geometry and parameter helpers (no HIP calls).
#pragma once

#include <string>
#include <vector>

#include "../../include/krabgpu.h"
#include "kg_rt.hpp"

namespace kg {

// Largest float s with sqrtf(s) <= dist (NaN for a NaN radius).  Comparing
// the squared distance the reference forms against it is exactly the
// reference's `distance(..) <= dist` test.
float sq_threshold(float dist);

// Geometry of a field (field_2d.rs:304-322, :403-404) with a device binning
// chosen for `expected_agents` (0 = unknown) or forced by `bin_rows`.
// Returns false if the arguments cannot describe a field.
bool make_geom(float w, float h, float d, int tor, uint64_t expected_agents, uint32_t bin_rows,
               Geom* g, std::string* why);

// Fills the per-step parts of a replica descriptor from flock params.
void set_rep_params(RepDev* r, const kg_flock_params& p);

// LDS capacity (staged agents per tile) for a geometry and expected count.
int pick_lds_cap(const Geom& g, uint64_t expected_agents);

bool params_valid(const kg_flock_params& p, std::string* why);

// Tile-resident layout plan: per-tile stayer and inbox capacities and bases
// from per-tile agent counts and the counts expected from the tiles' areas;
// nb[t] is tx * tr of tile t's replica.
struct RtPlan {
  std::vector<uint32_t> sbase, scap, ibase, icap, tboff;
  std::vector<RtMeta> meta;  // (sbase, scap, ibase, icap) per tile, as uploaded
  uint64_t stay_total = 0, inbox_total = 0, boff_total = 0;
  uint32_t max_scap = 0, max_icap = 0, max_nb = 0;
};
void rt_plan(const std::vector<uint32_t>& counts, const std::vector<uint32_t>& nb,
             const std::vector<double>& expect, RtPlan* plan);

// Agents each tile of `g` would hold if `agents` were spread uniformly.
void rt_expected_counts(const Geom& g, uint64_t agents, std::vector<double>* expect);

// One replica of a tile-resident plan: its geometry, window reach and first
// global tile.
struct RtRepInfo {
  Geom g;
  int32_t dd;
  uint32_t tile0;
};

// Per-tile LDS needs of a plan for the current per-tile counts: own agents
// (the count with 30 % + 32 of headroom) and staged stayers (that plus the
// neighbours' halo shares, with 10 % headroom).
void rt_needs(const RtPlan& plan, const std::vector<uint32_t>& counts,
              const std::vector<RtRepInfo>& reps, std::vector<uint32_t>* own,
              std::vector<uint32_t>* staged);

// Capacities of one launch for the given needs, covering every replica's
// geometry.
RtCaps rt_caps_for(const std::vector<RtRepInfo>& reps, uint32_t own, uint32_t staged);

// Splits the tiles into classes (sets plan->meta[].cls): class 0's
// capacities fit `lds_normal` bytes of LDS per work-group (0: rt_lds_budget
// of its work-group size), further classes
// grow geometrically to the fullest tile, and the last one also has room for
// tiles that outgrow class 0 before the next plan.  fits = every class fits
// `lds_limit`.
struct RtClassPlan {
  int ncls = 2;
  RtCaps caps[kRtMaxClasses];
  std::vector<uint32_t> dense;        // tiles of classes 1.., grouped by class
  uint32_t off[kRtMaxClasses + 1] = {};  // class k >= 1: dense[off[k] .. off[k + 1])
  bool fits = false;
};
// Dynamic LDS of a class-0 work-group of `threads` (256 or 512): what keeps
// a CU's waves busy, as measured.
size_t rt_lds_budget(int threads);
void rt_classes(RtPlan* plan, const std::vector<uint32_t>& counts,
                const std::vector<RtRepInfo>& reps, size_t lds_normal, size_t lds_limit,
                RtClassPlan* out);

// Geometry for the tile-resident layout: tiles of about 448 agents at the
// expected density unless tx, ty, hb are given (powers of two).  The tile
// height is the power of two at or below the wanted height, or the nearest
// one with rows_nearest (single fields: taller tiles share the per-tile
// work over more birds; ensembles keep the shorter tiles, which stay inside
// the LDS classes as their runs' flocks densify).
bool make_geom_rt(float w, float h, float d, int tor, uint64_t expected_agents, uint32_t tx,
                  uint32_t ty, uint32_t hb, Geom* g, std::string* why, bool rows_nearest = false);

}  // namespace kg
