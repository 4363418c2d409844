// kg_geom.hpp -- host-side geometry and parameter helpers (no HIP calls).
#pragma once

#include <string>

#include "../../include/krabgpu.h"
#include "kg_types.hpp"

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

}  // namespace kg
