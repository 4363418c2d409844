// kg_geom.cpp -- field geometry, device binning choice, replica parameters
// and the exact squared-distance threshold (host code without HIP calls).
#include "kg_geom.hpp"

#include <algorithm>
#include <cmath>

namespace kg {

float sq_threshold(float dist) {
  if (dist != dist) return dist;
  if (!(dist > 0.0f)) return -1.0f;
  if (std::isinf(dist)) return INFINITY;
  float t = dist * dist;
  while (t > 0.0f && std::sqrt(t) > dist) t = std::nextafter(t, 0.0f);
  for (;;) {
    const float up = std::nextafter(t, INFINITY);
    if (std::sqrt(up) <= dist) t = up;
    else break;
  }
  return t;
}

static int ilog2_floor(uint64_t v) {
  int k = 0;
  while ((v >> (k + 1)) != 0) ++k;
  return k;
}

bool make_geom(float w, float h, float d, int tor, uint64_t expected_agents, uint32_t bin_rows,
               Geom* g, std::string* why) {
  if (!(w > 0.0f) || !(h > 0.0f) || !(d > 0.0f) || std::isinf(w) || std::isinf(h) ||
      std::isinf(d)) {
    *why = "width, height and discretization must be finite and positive";
    return false;
  }
  // field_2d.rs:317-318 and :403-404, f32 arithmetic with Rust `as i32`
  const float cw = std::ceil(w / d), ch = std::ceil(h / d);
  if (!(cw < 1.0e7f) || !(ch < 1.0e7f)) {
    *why = "bag grid too large (ceil(w/d) and ceil(h/d) must stay below 1e7)";
    return false;
  }
  g->w = w;
  g->h = h;
  g->d = d;
  g->inv_d = 1.0f / d;
  g->tor = tor ? 1 : 0;
  g->max_x = (int32_t)cw;
  g->max_y = (int32_t)ch;
  g->dw = g->max_x + 1;
  g->dh = g->max_y + 1;
  const double cells = (double)g->dw * (double)g->dh;
  if (cells * 2.0 > 4.0e9) {
    *why = "bag grid too large for 32-bit bin indices";
    return false;
  }
  // Binning heuristic: about 2 agents per bin, about 512 own agents per tile,
  // 8 columns per tile; every size a power of two.
  const double dens = expected_agents ? (double)expected_agents / cells : 0.1;
  int hb_log2;
  if (bin_rows) {
    hb_log2 = ilog2_floor(bin_rows);
  } else {
    const double want = std::clamp(2.0 / std::max(dens, 1e-6), 1.0, 64.0);
    hb_log2 = ilog2_floor((uint64_t)want);
  }
  int tx_log2 = dens > 4.0 ? 2 : 3;
  const double tile_cells = 512.0 / std::max(dens, 1e-6);
  double ty_want = tile_cells / (double)(1 << tx_log2);
  ty_want = std::clamp(ty_want, (double)(1 << hb_log2), 1024.0);
  int ty_log2 = ilog2_floor((uint64_t)ty_want);
  if (ty_log2 < hb_log2) ty_log2 = hb_log2;
  // do not make tiles much taller than the grid
  while (ty_log2 > hb_log2 && (1 << (ty_log2 - 1)) >= g->dh) --ty_log2;
  while (tx_log2 > 0 && (1 << (tx_log2 - 1)) >= g->dw) --tx_log2;
  g->hb_log2 = hb_log2;
  g->tx_log2 = tx_log2;
  g->tr_log2 = ty_log2 - hb_log2;
  g->hb = 1 << g->hb_log2;
  g->tx = 1 << g->tx_log2;
  g->tr = 1 << g->tr_log2;
  g->ty = g->tr * g->hb;
  g->ntx = (g->dw + g->tx - 1) / g->tx;
  g->nty = (g->dh + g->ty - 1) / g->ty;
  const uint64_t ntiles = (uint64_t)g->ntx * (uint64_t)g->nty;
  const uint64_t nbins = ntiles * (uint64_t)g->tx * (uint64_t)g->tr;
  if (nbins >= (1ull << 31)) {
    *why = "too many device bins";
    return false;
  }
  g->ntiles = (uint32_t)ntiles;
  g->nbins = (uint32_t)nbins;
  return true;
}

int pick_lds_cap(const Geom& g, uint64_t expected_agents) {
  const double cells = (double)g.dw * (double)g.dh;
  const double dens = expected_agents ? (double)expected_agents / cells : 0.1;
  const double staged = dens * (double)(g.tx + 2) * (double)(g.ty + 2 * g.hb);
  // room for the tile-to-tile spread of a uniform field; a tile past it
  // takes the global path (clamp_lds_cap rounds up to what the LDS of a CU
  // holds at the same number of work-groups)
  const double need = std::ceil((staged * 1.3 + 64.0) / 64.0) * 64.0;
  return (int)std::min(4096.0, std::max(256.0, need));
}

bool params_valid(const kg_flock_params& p, std::string* why) {
  if (p.reserved != 0) {
    *why = "kg_flock_params.reserved must be 0";
    return false;
  }
  return true;
}

void set_rep_params(RepDev* r, const kg_flock_params& p) {
  r->p.cohesion = p.cohesion;
  r->p.avoidance = p.avoidance;
  r->p.consistency = p.consistency;
  r->p.randomness = p.randomness;
  r->p.momentum = p.momentum;
  r->p.jump = p.jump;
  r->p.radius = p.radius;
  r->p.reserved = 0;
  r->p.seed = p.seed;
  r->sq_thresh = sq_threshold(p.radius);
  // field_2d.rs:401: (dist/self.discretization).floor() as i32
  const float q = std::floor(p.radius / r->g.d);
  int32_t dd;
  if (q != q) dd = 0;
  else if (q >= 2147483648.0f) dd = INT32_MAX;
  else if (q <= -2147483648.0f) dd = INT32_MIN;
  else dd = (int32_t)q;
  r->disc_dist = dd;
  r->lds_ok = (r->g.tor != 0 && p.radius > 0.0f && dd >= 0 && dd <= r->g.tx && dd <= r->g.ty &&
               r->g.tx + 2 * dd <= kMaxStageCols)
                  ? 1
                  : 0;
}

}  // namespace kg

namespace kg {

void rt_plan(const std::vector<uint32_t>& counts, const std::vector<uint32_t>& nb,
             const std::vector<double>& expect, RtPlan* plan) {
  const size_t nt = counts.size();
  plan->sbase.resize(nt);
  plan->scap.resize(nt);
  plan->ibase.resize(nt);
  plan->icap.resize(nt);
  plan->tboff.resize(nt);
  plan->meta.resize(nt);
  uint64_t s = 0, ib = 0, bo = 0;
  plan->max_scap = plan->max_icap = plan->max_nb = 0;
  for (size_t t = 0; t < nt; ++t) {
    const uint32_t c = counts[t];
    // Flocks build large-scale density waves: after a few hundred steps the
    // fullest tile holds 2.5-4.5 times the mean of uniform tiles.  A tile
    // gets half its count again, and at least 2.5 times its share of the
    // agents by area; what still overflows goes to its inbox.
    const uint32_t by_area = (uint32_t)std::min(4.0e9, std::ceil(2.5 * expect[t]));
    uint32_t sc = std::max<uint32_t>(c + std::max<uint32_t>(32u, c / 2u), by_area + 32u);
    sc = (sc + 15u) & ~15u;
    uint32_t ic = std::max<uint32_t>(32u, sc / 4u);
    ic = (ic + 7u) & ~7u;
    plan->sbase[t] = (uint32_t)s;
    plan->scap[t] = sc;
    plan->ibase[t] = (uint32_t)ib;
    plan->icap[t] = ic;
    plan->tboff[t] = (uint32_t)bo;
    plan->meta[t] = RtMeta{(uint32_t)s, sc, (uint32_t)ib, ic, 0u, {0u, 0u, 0u}};
    s += sc;
    ib += ic;
    bo += nb[t] + 1u;
    plan->max_scap = std::max(plan->max_scap, sc);
    plan->max_icap = std::max(plan->max_icap, ic);
    plan->max_nb = std::max(plan->max_nb, nb[t]);
  }
  plan->stay_total = s;
  plan->inbox_total = ib;
  plan->boff_total = bo;
}

void rt_expected_counts(const Geom& g, uint64_t agents, std::vector<double>* expect) {
  expect->resize(g.ntiles);
  const double per_cell = (double)agents / ((double)g.dw * (double)g.dh);
  for (uint32_t t = 0; t < g.ntiles; ++t) {
    const int32_t tcx = (int32_t)(t % (uint32_t)g.ntx), tcy = (int32_t)(t / (uint32_t)g.ntx);
    const int32_t cols = std::min(g.tx, g.dw - tcx * g.tx);
    const int32_t rows = std::min(g.ty, g.dh - tcy * g.ty);
    (*expect)[t] = per_cell * (double)cols * (double)rows;
  }
}

namespace {

uint32_t round_up(double v, uint32_t q) {
  const double r = std::ceil(std::max(v, 1.0) / q) * q;
  return (uint32_t)std::min(r, 1.0e9);
}

}  // namespace

void rt_needs(const RtPlan& plan, const std::vector<uint32_t>& counts,
              const std::vector<RtRepInfo>& reps, std::vector<uint32_t>* own,
              std::vector<uint32_t>* staged) {
  const uint32_t nt = (uint32_t)plan.scap.size();
  own->resize(nt);
  staged->resize(nt);
  // LDS holds the agents actually present, with room for 64 steps of growth
  // (a tile past it defers to the dense launch once and the host re-plans);
  // the HBM capacities of the plan are far roomier.
  std::vector<double> want(nt);
  for (uint32_t t = 0; t < nt; ++t)
    want[t] = std::min<double>(plan.scap[t], std::ceil(1.3 * counts[t]) + 32.0);
  for (const RtRepInfo& ri : reps) {
    const Geom& g = ri.g;
    // rows staged above / below a tile: whole bins covering dd rows
    const double hrows = (double)g.hb * std::ceil((double)ri.dd / (double)g.hb);
    const double fx = std::min(1.0, (double)ri.dd / (double)g.tx);
    const double fy = std::min(1.0, hrows / (double)g.ty);
    for (uint32_t lt = 0; lt < g.ntiles; ++lt) {
      const uint32_t t = ri.tile0 + lt;
      const int32_t tcx = (int32_t)(lt % (uint32_t)g.ntx), tcy = (int32_t)(lt / (uint32_t)g.ntx);
      double st = want[t];
      for (int k = 0; k < 9; ++k) {
        const int32_t ax = tcx - 1 + k % 3, ay = tcy - 1 + k / 3;
        if (k == 4 || ax < 0 || ax >= g.ntx || ay < 0 || ay >= g.nty) continue;
        const double w = (ax != tcx ? fx : 1.0) * (ay != tcy ? fy : 1.0);
        st += w * want[ri.tile0 + (uint32_t)ay * (uint32_t)g.ntx + (uint32_t)ax];
      }
      (*own)[t] = (uint32_t)want[t];
      (*staged)[t] = (uint32_t)std::min(1.0e9, std::ceil(st * 1.1));
    }
  }
}

RtCaps rt_caps_for(const std::vector<RtRepInfo>& reps, uint32_t own, uint32_t staged) {
  RtCaps c;
  c.nb = 1;
  c.cols = 1;
  c.nbc = 2;
  for (const RtRepInfo& ri : reps) {
    const Geom& g = ri.g;
    c.nb = std::max(c.nb, g.tx * g.tr);
    c.cols = std::max(c.cols, std::min(kMaxStageCols, g.tx + 2 * ri.dd));
    // staged bins per column: the tile's tr plus whole bins covering dd
    // rows above and below
    c.nbc = std::max(c.nbc, g.tr + 2 * (int)((ri.dd + g.hb - 1) / g.hb) + 1);
  }
  c.own = (int)round_up(own, 64);
  c.runs = (int)round_up(staged, 64);
  // extras per step: agents that crossed into the staged rectangle, a few
  // percent of the tile's count; a quarter of the own capacity is ample
  c.extras = (int)std::min<uint32_t>(2048u, std::max<uint32_t>(64u, round_up(0.25 * own, 64)));
  // one own bird per thread up to 1024: a launch lasts as long as its
  // fullest tile, so larger classes get larger work-groups
  c.threads = c.own <= 512 ? 256 : (c.own <= 1024 ? 512 : 1024);
  return c;
}

size_t rt_lds_budget(int threads) {
  // A CU holds 16 of the step kernel's waves (about 100 VGPRs each): four
  // 256-thread or two 512-thread work-groups, each beside about 4 KB of
  // static LDS, in 160 KB.  Measured at 1M boids and on the 512-replica
  // ensemble: 28 KB for 256 threads (best of 24-40 KB), 76 KB for 512
  // (two fill the CU; best of 28-96 KB).  Class 0 never takes 1024-thread
  // work-groups: one per CU leaves the CU idle between its phases.
  return threads <= 256 ? (size_t)28 * 1024 : (size_t)(160 * 1024 / 2 - 4096);
}

void rt_classes(RtPlan* plan, const std::vector<uint32_t>& counts,
                const std::vector<RtRepInfo>& reps, size_t lds_normal, size_t lds_limit,
                RtClassPlan* out) {
  const uint32_t nt = (uint32_t)plan->scap.size();
  std::vector<uint32_t> own, staged;
  rt_needs(*plan, counts, reps, &own, &staged);
  out->dense.clear();
  std::fill(out->off, out->off + kRtMaxClasses + 1, 0u);
  out->fits = true;
  if (nt == 0) {
    out->ncls = 2;
    out->caps[0] = out->caps[1] = rt_caps_for(reps, 64, 64);
    return;
  }
  // class 0: the largest own-count quantile whose capacities fit the normal
  // LDS budget (lds_normal, or 0: what keeps a CU full for the class's
  // work-group size, rt_lds_budget)
  std::vector<uint32_t> so(own);
  std::sort(so.begin(), so.end());
  auto caps_at = [&](double qtl) {
    const uint32_t own_n = so[std::min<size_t>(nt - 1, (size_t)std::floor(qtl * (nt - 1)))];
    uint32_t st_n = 0;
    for (uint32_t t = 0; t < nt; ++t)
      if (own[t] <= own_n) st_n = std::max(st_n, staged[t]);
    return rt_caps_for(reps, own_n, st_n);
  };
  // class 0 keeps the work-group size of the median tile (at most 512)
  const int t0 = std::min(512, caps_at(0.5).threads);
  const size_t budget = lds_normal ? lds_normal : std::min(lds_limit, rt_lds_budget(t0));
  RtCaps cn{};
  for (double qtl : {1.0, 0.999, 0.995, 0.99, 0.98, 0.95, 0.9, 0.8, 0.5}) {
    cn = caps_at(qtl);
    if (cn.threads <= t0 && rt_smem_bytes(cn) <= budget) break;
  }
  uint32_t own_max = 0, st_max = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    own_max = std::max(own_max, own[t]);
    st_max = std::max(st_max, staged[t]);
  }
  // middle classes: capacities 1.6x the previous class's, while they leave
  // tiles uncovered and fit the LDS limit
  int nc = 1;
  out->caps[0] = cn;
  double f = 1.6;
  while (nc < kRtMaxClasses - 1) {
    const RtCaps& prev = out->caps[nc - 1];
    if ((uint32_t)prev.own >= own_max && (uint32_t)prev.runs >= st_max) break;
    const RtCaps c = rt_caps_for(reps, (uint32_t)(cn.own * f), (uint32_t)(cn.runs * f));
    if (rt_smem_bytes(c) > lds_limit) break;
    out->caps[nc++] = c;
    f *= 1.6;
  }
  // the last class: the fullest tile, and room for tiles that outgrow class
  // 0 (twice its capacities, or less where LDS cannot hold that), whose
  // extras may be most of their agents
  uint32_t own_x = std::max<uint32_t>(own_max, 2u * (uint32_t)cn.own);
  uint32_t st_x = std::max<uint32_t>(st_max, 2u * (uint32_t)cn.runs);
  RtCaps cl;
  for (;;) {
    cl = rt_caps_for(reps, own_x, st_x);
    cl.extras = std::min(2048, std::max(cl.extras, cl.own));
    if (rt_smem_bytes(cl) <= lds_limit || (own_x <= own_max && st_x <= st_max)) break;
    const uint32_t so2 = std::max<uint32_t>(64u, own_x / 8u), ss2 = std::max<uint32_t>(64u, st_x / 8u);
    own_x = own_x > own_max + so2 ? own_x - so2 : own_max;
    st_x = st_x > st_max + ss2 ? st_x - ss2 : st_max;
  }
  out->caps[nc++] = cl;
  out->ncls = nc;
  // each tile in the smallest class that holds it
  std::vector<std::vector<uint32_t>> lists(nc);
  for (uint32_t t = 0; t < nt; ++t) {
    int k = 0;
    while (k < nc - 1 && (own[t] > (uint32_t)out->caps[k].own ||
                          staged[t] > (uint32_t)out->caps[k].runs))
      ++k;
    plan->meta[t].cls = (uint32_t)k;
    if (k > 0) lists[k].push_back(t);
  }
  for (int k = 1; k < nc; ++k) {
    out->off[k] = (uint32_t)out->dense.size();
    out->dense.insert(out->dense.end(), lists[k].begin(), lists[k].end());
  }
  out->off[nc] = (uint32_t)out->dense.size();
  for (int k = 0; k < nc; ++k)
    if (rt_smem_bytes(out->caps[k]) > lds_limit) out->fits = false;
  out->fits = out->fits && cn.nb <= kRtMaxBins && cn.cols <= kMaxStageCols;
}

bool make_geom_rt(float w, float h, float d, int tor, uint64_t expected_agents, uint32_t tx,
                  uint32_t ty, uint32_t hb, Geom* g, std::string* why, bool rows_nearest) {
  if (!make_geom(w, h, d, tor, expected_agents, hb, g, why)) return false;
  const double cells = (double)g->dw * (double)g->dh;
  const double dens = expected_agents ? (double)expected_agents / cells : 0.1;
  if (!tx || !ty) {
    // tiles of about 448 agents, four times taller than wide
    const double tile_cells = std::max(16.0, 448.0 / std::max(dens, 1e-6));
    double wdt = std::sqrt(tile_cells / 4.0);
    int tl = 0;
    while ((1 << (tl + 1)) <= wdt && tl < 6) ++tl;
    tx = 1u << tl;
    // rows: the power of two at or below the wanted height, or the nearest
    // one (in ratio) when asked
    int yl = 0;
    while ((double)(1 << (yl + 1)) <= tile_cells / tx && yl < 12) ++yl;
    if (rows_nearest && yl < 12 && tile_cells / tx >= 1.5 * (double)(1 << yl)) ++yl;
    ty = 1u << yl;
  }
  if (!hb) {
    hb = 1u << g->hb_log2;
    while (hb > ty) hb >>= 1;
  }
  if ((tx & (tx - 1)) || (ty & (ty - 1)) || (hb & (hb - 1)) || hb > ty || tx > 128) {
    *why = "tile sizes must be powers of two with bin rows <= tile rows and tile columns <= 128";
    return false;
  }
  int a = 0;
  while ((1u << a) < tx) ++a;
  int b = 0;
  while ((1u << b) < hb) ++b;
  int c = 0;
  while ((1u << c) < ty) ++c;
  while (a > 0 && (1 << (a - 1)) >= g->dw) --a;
  while (c > b && (1 << (c - 1)) >= g->dh) --c;
  g->tx_log2 = a;
  g->hb_log2 = b;
  g->tr_log2 = c - b;
  g->tx = 1 << a;
  g->hb = 1 << b;
  g->tr = 1 << g->tr_log2;
  g->ty = g->tr * g->hb;
  g->ntx = (g->dw + g->tx - 1) / g->tx;
  g->nty = (g->dh + g->ty - 1) / g->ty;
  const uint64_t ntiles = (uint64_t)g->ntx * (uint64_t)g->nty;
  const uint64_t nbins = ntiles * (uint64_t)g->tx * (uint64_t)g->tr;
  if (nbins >= (1ull << 31) || g->tx * g->tr > kRtMaxBins) {
    *why = "tile-resident layout: too many bins";
    return false;
  }
  g->ntiles = (uint32_t)ntiles;
  g->nbins = (uint32_t)nbins;
  return true;
}

}  // namespace kg
