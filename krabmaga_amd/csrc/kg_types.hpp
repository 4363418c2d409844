// kg_types.hpp -- replica descriptors and work lists shared by the kernels,
// the host runtime and the host-side driver of the kernel phases.
//
// The kernels operate on a table of replicas (RepDev): a plain Field2D is a
// table of one.  All replicas share the global agent arrays and the global
// bin arrays; replica r owns bins [bin_base, bin_base + g.nbins) and its read
// buffer range follows from the offsets of those bins.
#pragma once

#include "kg_device.hpp"

namespace kg {

struct RepDev {
  Geom g;
  FlockParams p;
  float sq_thresh;     // largest s with sqrtf(s) <= p.radius
  int32_t disc_dist;   // floor(p.radius / g.d)
  int32_t lds_ok;      // the LDS tile path applies (toroidal, window fits one tile of halo)
  uint32_t bin_base;   // first global bin of this replica
  uint32_t app_base;   // first append-buffer entry of this replica
  uint32_t app_n;      // append-buffer entries of this replica
  uint32_t boff_base;  // tile-resident layout: start of tile 0's bin-offset table
  uint32_t dead;       // retired replica: its tiles do no work (ensembles)
  uint64_t step_base;  // schedule step of this replica = step_base + *gstep
};

// Work lists are prefix arrays over replicas: item i belongs to the replica
// r with prefix[r] <= i < prefix[r + 1].
struct WorkList {
  const RepDev* reps;     // device table
  int nreps;
  const uint32_t* tile_prefix;  // device, nreps + 1 (tiles of each replica)
  uint32_t ntiles;              // tile_prefix[nreps]
  const uint32_t* app_prefix;   // device, nreps + 1 (append entries)
  uint32_t napp;
};

struct ScanState {
  uint32_t* partial;  // one sum per scan tile, plus the total
};

constexpr int kScanThreads = 256;
constexpr int kScanItems = 2;
constexpr int kScanTile = kScanThreads * kScanItems;
// Scan tiles the two-launch scan sums inline (past them it adds a prefix pass).
constexpr uint32_t kScanInlineTiles = 2048;
constexpr int kFlockThreads = 256;
constexpr int kMaxStageCols = 136;  // tx <= 128 plus a halo of up to 4 columns

}  // namespace kg
