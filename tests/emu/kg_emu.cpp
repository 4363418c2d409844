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

struct emu_field {
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
  uint64_t lds_tiles = 0, global_tiles = 0;
};

extern "C" {

emu_field* emu_create(float w, float h, float d, int tor, uint32_t cap, uint32_t bin_rows,
                      int lds_cap) {
  emu_field* f = new emu_field();
  std::string why;
  if (!make_geom(w, h, d, tor, cap, bin_rows, &f->g, &why)) {
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
  return f;
}

void emu_destroy(emu_field* f) { delete f; }

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

void emu_stats(emu_field* f, uint64_t* lds_tiles, uint64_t* global_tiles) {
  *lds_tiles = f->lds_tiles;
  *global_tiles = f->global_tiles;
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

int emu_lazy_update(emu_field* f) {
  const uint32_t prefix[2] = {0u, f->g.ntiles};
  const uint32_t app_prefix[2] = {0u, f->nA};
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
  if (f->w_aligned) {
    std::vector<uint32_t> items = order(0, f->g.ntiles, f->shuffle_seed);
    for (uint32_t item : items) {
      const TileRange tr = tile_range(&f->rep, 1, prefix, item, f->offset.data());
      for (uint32_t i : order(tr.lo, tr.hi, f->shuffle_seed))
        scatter_aligned_one(f->rep, R, W, i, f->cursor.data(), R2);
    }
  }
  for (uint32_t i : order(0, f->nA, f->shuffle_seed))
    scatter_append_one(&f->rep, 1, app_prefix, A, i, f->cursor.data(), R2);
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
      use_lds = L <= (uint32_t)f->lds_cap;
    }
    if (use_lds) {
      ++f->lds_tiles;
      std::fill(smem.begin(), smem.end(), 0xA5);
      const FlockSmem S = flock_carve(smem.data(), f->lds_cap);
      for (uint32_t k = 0; k < L; ++k)
        flock_stage(rp, ft, k, R, seg_start.data(), seg_len.data(), col_base.data(), S);
      uint32_t nown = 0;
      for (uint32_t k = 0; k < L; ++k) flock_rank(ft, k, col_base.data(), S, &nown);
      for (uint32_t o = 0; o < nown; ++o)
        flock_agent_lds(rp, ft, S.own[o], col_base.data(), S, W, f->count.data(), &f->err);
    } else {
      ++f->global_tiles;
      const uint32_t lo = f->offset[ft.tb0], hi = f->offset[ft.tb1];
      for (uint32_t gi = lo; gi < hi; ++gi)
        flock_agent_global(rp, ft.step, gi, R, f->offset.data(), W, f->count.data(), &f->err);
    }
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
