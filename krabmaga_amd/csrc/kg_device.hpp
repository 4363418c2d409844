// kg_device.hpp -- device building blocks of the agent-step kernels.
//
// Data layout of one field in HBM (struct-of-arrays, f32/u32):
//   read buffer  R: x, y, hx, hy, id    grouped by bin, unordered inside a bin
//   write buffer W: x, y, hx, hy        the flock kernel's output, aligned
//                                       with R (ids stay in R.id)
//   append buffer A: x, y, hx, hy, id   set_object_location / init, in
//                                       insertion order
//   offset[b] .. offset[b+1] is bin b's range in R (bin numbering in Geom).
//
// Inside a bin the order is arbitrary (atomic scatter); every consumer that
// needs the reference's bag order sorts by (cell, id) itself.
#pragma once

#include "kg_math.hpp"

namespace kg {

constexpr int kWave = 64;

struct Buf {
  float* x;
  float* y;
  float* hx;
  float* hy;
  uint32_t* id;
};

struct FlockParams {
  float cohesion, avoidance, consistency, randomness, momentum, jump, radius;
  uint32_t reserved;
  uint64_t seed;
};

KG_HD uint32_t tile_of(const Geom& g, int32_t cx, int32_t cy) {
  return (uint32_t)(cy >> (g.tr_log2 + g.hb_log2)) * (uint32_t)g.ntx + (uint32_t)(cx >> g.tx_log2);
}

// First bin of column-in-tile lc of tile t.
KG_HD uint32_t col_bin0(const Geom& g, uint32_t t, int32_t lc) {
  return ((t << g.tx_log2) + (uint32_t)lc) << g.tr_log2;
}

KG_HD uint32_t bin_of(const Geom& g, int32_t cx, int32_t cy) {
  const uint32_t t = tile_of(g, cx, cy);
  const int32_t lc = cx & (g.tx - 1);
  const int32_t lr = (cy & (g.ty - 1)) >> g.hb_log2;
  return col_bin0(g, t, lc) + (uint32_t)lr;
}

KG_HD bool cell_valid(const Geom& g, int32_t cx, int32_t cy) {
  return cx >= 0 && cx < g.dw && cy >= 0 && cy < g.dh;
}

// field_2d.rs:934-972 with every `distance(..) <= dis` replaced by the exact
// equivalent on squared distances (s <= T and s > T; NaN fails both, as it
// fails both comparisons of the original).
KG_HD int check_circle_sq(const Geom& g, int32_t bx, int32_t by, float lx,
                                               float ly, float T) {
  const float nw_x = (float)bx * g.d;
  const float nw_y = (float)by * g.d;
  const float ne_y = fminf(nw_y + g.d, g.h);
  const float sw_x = fminf(nw_x + g.d, g.w);
  const bool tor = g.tor != 0;
  const float s_nw = dist2(nw_x, nw_y, lx, ly, g.w, g.h, tor);
  const float s_ne = dist2(nw_x, ne_y, lx, ly, g.w, g.h, tor);
  const float s_sw = dist2(sw_x, nw_y, lx, ly, g.w, g.h, tor);
  const float s_se = dist2(sw_x, ne_y, lx, ly, g.w, g.h, tor);
  if (s_nw <= T && s_ne <= T && s_sw <= T && s_se <= T) return 1;
  if (s_nw > T && s_ne > T && s_sw > T && s_se > T) return -1;
  return 0;
}

// check_circle_sq of every cell of a window of at most 3 x 3 cells, packed
// two bits per cell at bit 2 * (3 * kc + kr) for cell (bxs[kc], bys[kr]):
// 1 = all four corners within T, 2 = all four beyond T, 0 = otherwise.
// A corner's squared distance is fl(sx + sy) with sx the squared x term of
// its column edge and sy the squared y term of its row edge; fl(+) is
// monotone, so the largest of the four sums is fl(max sx + max sy) and the
// smallest fl(min sx + min sy), and the two comparisons below decide exactly
// what the four comparisons of check_circle_sq decide (a NaN query makes
// both terms of a pair NaN, so both tests fail, as all eight do there).
// ncol, nrow <= 3 are the window's column and row counts; columns and rows
// are given as cell indices (already wrapped).
KG_HD uint32_t window_checks(const Geom& g, const int32_t* bxs, const int32_t* bys, int ncol,
                             int nrow, float lx, float ly, float T) {
  const bool tor = g.tor != 0;
  float xmx[3], xmn[3], ymx[3], ymn[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float nw_x = (float)bxs[k] * g.d;
    const float sw_x = fminf(nw_x + g.d, g.w);
    const float nw_y = (float)bys[k] * g.d;
    const float ne_y = fminf(nw_y + g.d, g.h);
    const float ax = tor ? toroidal_distance(nw_x, lx, g.w) : nw_x - lx;
    const float bx = tor ? toroidal_distance(sw_x, lx, g.w) : sw_x - lx;
    const float ay = tor ? toroidal_distance(nw_y, ly, g.h) : nw_y - ly;
    const float by = tor ? toroidal_distance(ne_y, ly, g.h) : ne_y - ly;
    const float ax2 = ax * ax, bx2 = bx * bx, ay2 = ay * ay, by2 = by * by;
    xmx[k] = fmaxf(ax2, bx2);
    xmn[k] = fminf(ax2, bx2);
    ymx[k] = fmaxf(ay2, by2);
    ymn[k] = fminf(ay2, by2);
  }
  uint32_t bits = 0;
#pragma unroll
  for (int kc = 0; kc < 3; ++kc) {
#pragma unroll
    for (int kr = 0; kr < 3; ++kr) {
      const float hi = xmx[kc] + ymx[kr];
      const float lo = xmn[kc] + ymn[kr];
      const uint32_t c = hi <= T ? 1u : (lo > T ? 2u : 0u);
      bits |= (kc < ncol && kr < nrow ? c : 0u) << (2 * (3 * kc + kr));
    }
  }
  return bits;
}

// The check_circle_sq value of window cell (kc, kr) from window_checks' bits.
KG_HD int window_check_at(uint32_t bits, int kc, int kr) {
  const uint32_t c = (bits >> (2 * (3 * kc + kr))) & 3u;
  return c == 1u ? 1 : (c == 2u ? -1 : 0);
}

// Accumulators of Bird::step's neighbour loop (bird.rs:54-81).
struct FlockAcc {
  float x_avoid = 0.0f, y_avoid = 0.0f;
  float x_cohe = 0.0f, y_cohe = 0.0f;
  float x_cons = 0.0f, y_cons = 0.0f;
  int32_t count = 0;
  uint32_t nfound = 0;  // vec.len() including self

  // One element of the neighbour vector.  dx, dy are
  // toroidal_distance(self, elem) per axis (bird.rs:64-65).
  KG_HD void add(bool is_self, float dx, float dy, float ehx, float ehy) {
    nfound += 1;
    if (is_self) return;
    count += 1;
    const float square = dx * dx + dy * dy;
    x_avoid += dx / (square * square + 1.0f);
    y_avoid += dy / (square * square + 1.0f);
    x_cohe += dx;
    y_cohe += dy;
    x_cons += ehx;
    y_cons += ehy;
  }
};

// bird.rs:83-153: combine the terms, normalise to jump, wrap.  Writes the
// new position and heading.
KG_HD void flock_finish(const FlockAcc& a, const FlockParams& p, uint32_t id,
                                             uint64_t step, float px, float py, float hx, float hy,
                                             float width, float* nx, float* ny, float* nhx,
                                             float* nhy) {
  float avoid_x = 0.0f, avoid_y = 0.0f, cohe_x = 0.0f, cohe_y = 0.0f;
  float rand_x = 0.0f, rand_y = 0.0f, cons_x = 0.0f, cons_y = 0.0f;
  if (a.nfound != 0) {
    float x_avoid = a.x_avoid, y_avoid = a.y_avoid, x_cohe = a.x_cohe, y_cohe = a.y_cohe;
    float x_cons = a.x_cons, y_cons = a.y_cons;
    if (a.count > 0) {
      const float c = (float)a.count;
      x_avoid /= c;
      y_avoid /= c;
      x_cohe /= c;
      y_cohe /= c;
      x_cons /= c;
      y_cons /= c;
      cons_x = x_cons / c;
      cons_y = y_cons / c;
    } else {
      cons_x = x_cons;
      cons_y = y_cons;
    }
    avoid_x = 400.0f * x_avoid;
    avoid_y = 400.0f * y_avoid;
    cohe_x = -x_cohe / 10.0f;
    cohe_y = -y_cohe / 10.0f;
    float r1, r2;
    step_uniforms(p.seed, id, step, &r1, &r2);
    const float x_rand = r1 * 2.0f - 1.0f;
    const float y_rand = r2 * 2.0f - 1.0f;
    const float sq = sqrtf(x_rand * x_rand + y_rand * y_rand);
    rand_x = 0.05f * x_rand / sq;
    rand_y = 0.05f * y_rand / sq;
  }
  float dx = p.cohesion * cohe_x + p.avoidance * avoid_x + p.consistency * cons_x +
             p.randomness * rand_x + p.momentum * hx;
  float dy = p.cohesion * cohe_y + p.avoidance * avoid_y + p.consistency * cons_y +
             p.randomness * rand_y + p.momentum * hy;
  const float dis = sqrtf(dx * dx + dy * dy);
  if (dis > 0.0f) {
    dx = dx / dis * p.jump;
    dy = dy / dis * p.jump;
  }
  *nhx = dx;
  *nhy = dy;
  *nx = toroidal_transform(px + dx, width);
  *ny = toroidal_transform(py + dy, width);  // bird.rs:147 wraps y by width
}

// Window of field_2d.rs:401-416.
struct Window {
  int32_t min_i, max_i, min_j, max_j;
};

KG_HD Window query_window(const Geom& g, float x, float y, int32_t dd) {
  int32_t cx, cy;
  discretize(g, x, y, &cx, &cy);
  Window wdw;
  wdw.min_i = cx - dd;
  wdw.max_i = cx + dd;
  wdw.min_j = cy - dd;
  wdw.max_j = cy + dd;
  if (g.tor) {
    wdw.min_i = imax(0, wdw.min_i);
    wdw.max_i = imin(wdw.max_i, g.max_x - 1);
    wdw.min_j = imax(0, wdw.min_j);
    wdw.max_j = imin(wdw.max_j, g.max_y - 1);
  }
  return wdw;
}

// Visits, in the reference's bag-scan order, every agent of the read buffer
// that get_neighbors_within_distance (field_2d.rs:386-440, relax = false) or
// get_neighbors_within_relax_distance (:472-516, relax = true) returns for
// (x, y, dist).  Bags are read from the binned read buffer; inside a bag
// agents are visited by ascending id (equal ids in unspecified order).
// visit(e) receives the read-buffer index.  Works for any radius and both
// toroidal modes.
//
// A bin holds rows of one column sorted by (cell y, id), so the bag of
// (bx, by) is the run of cell y == by.  Its start is found by a linear scan
// from the bin's start when the bin is small, by bisection otherwise; the
// window's next row in the same bin starts where this row's run ended, so a
// column costs one lookup per bin it touches.
template <class Visit>
KG_HD void query_global(const Geom& g, const Buf& R, const uint32_t* __restrict__ offset,
                             float x, float y, float dist, float T, bool relax, Visit&& visit) {
  if (dist <= 0.0f) return;
  const int32_t dd = f2i(floorf(dist / g.d));
  const Window wdw = query_window(g, x, y, dd);
  const bool tor = g.tor != 0;
  // windows of up to 3 x 3 cells: every cell's check_circle at once
  const bool packed = !relax && wdw.max_i - wdw.min_i < 3 && wdw.max_j - wdw.min_j < 3;
  uint32_t wchk = 0;
  if (packed) {
    int32_t bxs[3], bys[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      bxs[k] = t_transform(imin(wdw.min_i + k, wdw.max_i), g.max_x);
      bys[k] = t_transform(imin(wdw.min_j + k, wdw.max_j), g.max_y);
    }
    wchk = window_checks(g, bxs, bys, wdw.max_i - wdw.min_i + 1, wdw.max_j - wdw.min_j + 1, x, y,
                         T);
  }
  for (int64_t i = wdw.min_i; i < (int64_t)wdw.max_i + 1; ++i) {
    const int32_t bx = t_transform((int32_t)i, g.max_x);
    uint32_t run_b = 0xFFFFFFFFu;  // bin and row whose run ended at run_end
    int32_t run_by = 0;
    uint32_t run_end = 0;
    for (int64_t j = wdw.min_j; j < (int64_t)wdw.max_j + 1; ++j) {
      const int32_t by = t_transform((int32_t)j, g.max_y);
      const int check = relax ? 1
                        : packed ? window_check_at(wchk, (int)(i - wdw.min_i), (int)(j - wdw.min_j))
                                 : check_circle_sq(g, bx, by, x, y, T);
      if (check == -1) continue;
      const uint32_t b = bin_of(g, bx, by);
      const uint32_t e = offset[b + 1];
      uint32_t q;
      if (b == run_b && by == run_by + 1) {
        q = run_end;
      } else {
        uint32_t lo = offset[b], hi = e;
        if (hi - lo <= 8u) {
          while (lo < hi && disc_axis(R.y[lo], g.d, g.inv_d) < by) ++lo;
        } else {
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (disc_axis(R.y[mid], g.d, g.inv_d) < by) lo = mid + 1;
            else hi = mid;
          }
        }
        q = lo;
      }
      for (; q < e; ++q) {
        const float ex = R.x[q], ey = R.y[q];
        if (disc_axis(ey, g.d, g.inv_d) != by) break;
        if (check == 1 || dist2(x, y, ex, ey, g.w, g.h, tor) <= T) visit(q);
      }
      run_b = b;
      run_by = by;
      run_end = q;
    }
  }
}

#if defined(__HIPCC__)
// ---- block-wide helpers (blockDim.x == kThreads, multiple of 64) ----------

template <int kThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_warp,
                                                         uint32_t* total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += t;
  }
  if (lane == kWave - 1) s_warp[wid] = incl;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / kWave; ++w) {
    const uint32_t t = s_warp[w];
    if (w < wid) wbase += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return wbase + incl - v;
}
#endif

}  // namespace kg
