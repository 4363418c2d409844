// emu_check.cpp -- runs the device kernel phases on the CPU (kg_emu.cpp)
// under AddressSanitizer/UBSan and compares every run bit for bit with the
// oracle.  Exit status 0 = all scenarios identical and no sanitizer report.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "kg_geom.hpp"
#include "oracle.h"

extern "C" {
struct emu_field;
emu_field* emu_create(float w, float h, float d, int tor, uint32_t cap, uint32_t bin_rows,
                      int lds_cap);
void emu_destroy(emu_field* f);
void emu_options(emu_field* f, int force_global, uint64_t shuffle_seed);
void emu_set_gather(emu_field* f, int on);
void emu_gather_stats(emu_field* f, uint64_t* steps, uint64_t* fallbacks);
void emu_stats(emu_field* f, uint64_t* lds_tiles, uint64_t* global_tiles,
               uint64_t* chunked_tiles);
int emu_set(emu_field* f, const uint32_t* ids, const float* x, const float* y, const float* hx,
            const float* hy, uint32_t n);
int emu_lazy_update(emu_field* f);
int emu_flock_step(emu_field* f, const kg_flock_params* p, uint64_t step);
uint32_t emu_read(emu_field* f, uint32_t* ids, float* x, float* y, float* hx, float* hy);
emu_field* emu_create_rt(float w, float h, float d, int tor, uint32_t cap, uint32_t tx,
                         uint32_t ty, uint32_t hb);
int emu_rt_enter(emu_field* f, int caps_runs, int caps_own, int caps_extras, int tight);
int emu_rt_steps(emu_field* f, const kg_flock_params* p, uint64_t step0, uint32_t nsteps);
uint32_t emu_rt_flags(emu_field* f);
uint32_t emu_rt_read(emu_field* f, uint32_t* ids, float* x, float* y, float* hx, float* hy);
void emu_rt_counts(emu_field* f, uint64_t* lds, uint64_t* glob, uint64_t* deferred,
                   uint32_t* ndense, int* ncls);
}

struct RtScenario {
  const char* name;
  uint32_t n;
  float side, d;
  uint32_t steps;
  uint64_t seed;
  uint32_t tx, ty, hb;  // 0 = automatic
  int runs, own, extras; // LDS capacities, 0 = planned
  int tight;            // bit 0: capacities below the counts; bit 1: class split
  uint64_t shuffle;
  int layout;
  float radius;
  float jump;
};

struct Scenario {
  const char* name;
  uint32_t n;
  float side, d;
  int tor;
  uint32_t steps;
  uint64_t seed;
  uint32_t bin_rows;
  int lds_cap;
  int force_global;
  uint64_t shuffle;
  int layout;  // 0 uniform, 1 clustered, 2 hugging the far edges (padding cells),
               // 3 stacked birds and cell-boundary positions, 4 bands one
               // jump wide at both sides of the seams, heading across them
  float radius;
  float jump = 0.7f;
  int gather = 1;  // the gather rebinning after flock steps (the device default)
  int all_gathered = 0;  // every flock step must rebin by gather
};

static void make_positions(const Scenario& s, std::vector<float>& x, std::vector<float>& y,
                           std::vector<float>& hx, std::vector<float>& hy) {
  x.resize(s.n);
  y.resize(s.n);
  hx.resize(s.n);
  hy.resize(s.n);
  or_init_uniform(s.n, s.side, s.side, s.seed, 1, x.data(), y.data(), hx.data(), hy.data());
  std::mt19937_64 rng(s.seed * 31 + 7);
  if (s.layout == 1) {
    std::normal_distribution<double> nd(0.0, s.side / 40.0);
    std::uniform_real_distribution<double> ud(0.0, s.side);
    double cx[6], cy[6];
    for (int k = 0; k < 6; ++k) {
      cx[k] = ud(rng);
      cy[k] = ud(rng);
    }
    for (uint32_t i = 0; i < s.n; ++i) {
      if (i % 5 == 0) continue;  // 20 % stay uniform
      const int k = (int)(i % 6);
      x[i] = or_toroidal_transform((float)(cx[k] + nd(rng)), s.side);
      y[i] = or_toroidal_transform((float)(cy[k] + nd(rng)), s.side);
    }
  } else if (s.layout == 4) {
    std::uniform_real_distribution<double> ud(0.05, 1.0);
    for (uint32_t i = 0; i < s.n; ++i) {
      const float off = (float)(ud(rng) * s.jump);
      const float sgn = (i & 2u) ? 1.0f : -1.0f;  // toward the seam
      const float at = sgn > 0 ? s.side - off : off;
      if (i & 1u) {
        x[i] = at;
        hx[i] = sgn;
        hy[i] = 0.0f;
      } else {
        y[i] = at;
        hx[i] = 0.0f;
        hy[i] = sgn;
      }
    }
  } else if (s.layout == 3) {
    for (uint32_t i = 0; i < s.n; ++i) {
      const uint32_t r = i % 4;
      if (r == 1 || r == 2) {
        // exact duplicates of the group's first bird (distance 0, equal cells)
        x[i] = x[i - r];
        y[i] = y[i - r];
      } else if (r == 3) {
        // on a cell corner, or one ulp either side of it
        const float d = s.d;
        float bx = (float)(int)(x[i] / d) * d, by = (float)(int)(y[i] / d) * d;
        const int m = (int)((i / 4) % 3);
        if (m == 1) { bx = std::nextafter(bx, 0.0f); by = std::nextafter(by, s.side); }
        if (m == 2) { bx = std::nextafter(bx, s.side); by = std::nextafter(by, 0.0f); }
        x[i] = bx < s.side ? bx : 0.0f;
        y[i] = by < s.side ? by : 0.0f;
      }
    }
  } else if (s.layout == 2) {
    for (uint32_t i = 0; i < s.n; i += 3) {
      // the largest floats below the side: x/d may round up to ceil(w/d),
      // the padding column/row the reference's window never scans
      const int k = (int)(i % 7);
      float v = s.side;
      for (int j = 0; j <= k; ++j) v = std::nextafter(v, 0.0f);
      if (i % 2) x[i] = v;
      else y[i] = v;
    }
  }
}

static int run(const Scenario& s) {
  std::vector<float> x, y, hx, hy;
  make_positions(s, x, y, hx, hy);
  or_flock_params op{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, s.jump, s.radius, s.seed};
  or_flocker* fl = or_flocker_new(s.n, s.side, s.side, s.d, s.tor, x.data(), y.data(), hx.data(),
                                  hy.data(), &op);
  emu_field* e = emu_create(s.side, s.side, s.d, s.tor, s.n, s.bin_rows, s.lds_cap);
  if (!e) {
    std::printf("%-28s emu_create failed\n", s.name);
    return 1;
  }
  emu_options(e, s.force_global, s.shuffle);
  emu_set_gather(e, s.gather);
  std::vector<uint32_t> ids(s.n);
  for (uint32_t i = 0; i < s.n; ++i) ids[i] = i;
  emu_set(e, ids.data(), x.data(), y.data(), hx.data(), hy.data(), s.n);
  int rc = emu_lazy_update(e);
  kg_flock_params kp{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, s.jump, s.radius, 0u, s.seed};
  for (uint32_t k = 0; k < s.steps && rc == 0; ++k) {
    rc = emu_flock_step(e, &kp, k);
    if (rc == 0) rc = emu_lazy_update(e);
    or_flocker_step(fl);
  }
  std::vector<uint32_t> gid(s.n);
  std::vector<float> gx(s.n), gy(s.n), ghx(s.n), ghy(s.n);
  const uint32_t m = emu_read(e, gid.data(), gx.data(), gy.data(), ghx.data(), ghy.data());
  std::vector<float> rx(s.n), ry(s.n), rhx(s.n), rhy(s.n);
  or_flocker_get(fl, rx.data(), ry.data(), rhx.data(), rhy.data());
  uint64_t lds = 0, glob = 0, chunked = 0, gsteps = 0, gfalls = 0;
  emu_stats(e, &lds, &glob, &chunked);
  emu_gather_stats(e, &gsteps, &gfalls);
  uint32_t bad = 0;
  std::vector<int> seen(s.n, 0);
  for (uint32_t k = 0; k < m; ++k) {
    const uint32_t i = gid[k];
    if (i >= s.n || seen[i]++) {
      ++bad;
      continue;
    }
    if (std::memcmp(&gx[k], &rx[i], 4) || std::memcmp(&gy[k], &ry[i], 4) ||
        std::memcmp(&ghx[k], &rhx[i], 4) || std::memcmp(&ghy[k], &rhy[i], 4))
      ++bad;
  }
  if (m != s.n) bad += s.n;
  if (s.all_gathered && (gfalls != 0 || gsteps != s.steps)) bad += 1;
  std::printf("%-28s rc=%d agents=%u lds_tiles=%llu chunked=%llu global_tiles=%llu "
              "gathered=%llu/%llu mismatches=%u\n",
              s.name, rc, m, (unsigned long long)lds, (unsigned long long)chunked,
              (unsigned long long)glob, (unsigned long long)(gsteps - gfalls),
              (unsigned long long)gsteps, bad);
  emu_destroy(e);
  or_flocker_free(fl);
  return (rc != 0 || bad != 0) ? 1 : 0;
}

static int compare(const char* name, uint32_t n, const std::vector<uint32_t>& gid,
                   const std::vector<float>& gx, const std::vector<float>& gy,
                   const std::vector<float>& ghx, const std::vector<float>& ghy, uint32_t m,
                   or_flocker* fl, const char* extra) {
  std::vector<float> rx(n), ry(n), rhx(n), rhy(n);
  or_flocker_get(fl, rx.data(), ry.data(), rhx.data(), rhy.data());
  uint32_t bad = 0;
  std::vector<int> seen(n, 0);
  for (uint32_t k = 0; k < m; ++k) {
    const uint32_t i = gid[k];
    if (i >= n || seen[i]++) {
      ++bad;
      continue;
    }
    if (std::memcmp(&gx[k], &rx[i], 4) || std::memcmp(&gy[k], &ry[i], 4) ||
        std::memcmp(&ghx[k], &rhx[i], 4) || std::memcmp(&ghy[k], &rhy[i], 4))
      ++bad;
  }
  if (m != n) bad += n;
  std::printf("%-28s agents=%u %s mismatches=%u\n", name, m, extra, bad);
  return bad ? 1 : 0;
}

static int run_rt(const RtScenario& s) {
  Scenario base{s.name, s.n, s.side, s.d, 1, s.steps, s.seed, 0, 0, 0, 0, s.layout, s.radius};
  std::vector<float> x, y, hx, hy;
  make_positions(base, x, y, hx, hy);
  or_flock_params op{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, s.jump, s.radius, s.seed};
  or_flocker* fl = or_flocker_new(s.n, s.side, s.side, s.d, 1, x.data(), y.data(), hx.data(),
                                  hy.data(), &op);
  emu_field* e = emu_create_rt(s.side, s.side, s.d, 1, s.n, s.tx, s.ty, s.hb);
  if (!e) {
    std::printf("%-28s emu_create_rt failed\n", s.name);
    return 1;
  }
  emu_options(e, 0, s.shuffle);
  std::vector<uint32_t> ids(s.n);
  for (uint32_t i = 0; i < s.n; ++i) ids[i] = i;
  emu_set(e, ids.data(), x.data(), y.data(), hx.data(), hy.data(), s.n);
  int rc = emu_lazy_update(e);  // schedule.rs:349-351
  if (rc == 0) rc = emu_rt_enter(e, s.runs, s.own, s.extras, s.tight);
  kg_flock_params kp{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, s.jump, s.radius, 0u, s.seed};
  if (rc == 0) rc = emu_rt_steps(e, &kp, 0, s.steps);
  for (uint32_t k = 0; k < s.steps; ++k) or_flocker_step(fl);
  std::vector<uint32_t> gid(s.n);
  std::vector<float> gx(s.n), gy(s.n), ghx(s.n), ghy(s.n);
  const uint32_t m = emu_rt_read(e, gid.data(), gx.data(), gy.data(), ghx.data(), ghy.data());
  char extra[160];
  uint64_t nl = 0, ng = 0, ndf = 0;
  uint32_t nde = 0;
  int ncls = 0;
  emu_rt_counts(e, &nl, &ng, &ndf, &nde, &ncls);
  std::snprintf(extra, sizeof(extra),
                "rc=%d flags=0x%x classes=%d dense=%u lds=%llu global=%llu deferred=%llu", rc,
                emu_rt_flags(e), ncls, nde, (unsigned long long)nl, (unsigned long long)ng,
                (unsigned long long)ndf);
  const int bad = compare(s.name, s.n, gid, gx, gy, ghx, ghy, m, fl, extra);
  emu_destroy(e);
  or_flocker_free(fl);
  return (rc != 0 || bad) ? 1 : 0;
}

int main() {
  std::setvbuf(stdout, nullptr, _IOLBF, 0);
  const float D = 10.0f / 1.5f;
  const Scenario sc[] = {
      {"flockers-200", 1000, 200.0f, D, 1, 40, 0, 0, 0, 0, 0, 0, 10.0f, 0.7f, 1, 1},
      {"flockers-200-global", 1000, 200.0f, D, 1, 10, 0, 0, 0, 1, 0, 0, 10.0f},
      {"flockers-200-shuffled", 1000, 200.0f, D, 1, 20, 0, 0, 0, 0, 99, 0, 10.0f},
      {"flockers-200-hb16-cap64", 1000, 200.0f, D, 1, 10, 1, 16, 64, 0, 0, 0, 10.0f},
      {"sparse-2048", 10000, 2048.0f, D, 1, 6, 2, 0, 0, 0, 5, 0, 10.0f},
      {"sparse-2048-hb1-global", 10000, 2048.0f, D, 1, 6, 2, 1, 0, 1, 5, 0, 10.0f},
      {"clustered-hb2-global", 6000, 700.0f, D, 1, 4, 4, 2, 0, 1, 0, 1, 10.0f},
      {"dense-60", 3000, 60.0f, D, 1, 8, 3, 0, 0, 0, 0, 0, 10.0f},
      {"clustered", 6000, 700.0f, D, 1, 8, 4, 0, 0, 0, 0, 1, 10.0f},
      {"clustered-cap128", 6000, 700.0f, D, 1, 4, 4, 0, 128, 0, 3, 1, 10.0f},
      {"chunked-cap192", 6000, 700.0f, D, 1, 6, 4, 8, 192, 0, 3, 1, 10.0f},
      {"chunked-uniform-cap256", 8000, 300.0f, D, 1, 6, 11, 4, 256, 0, 0, 0, 10.0f},
      {"chunked-5x5-cap256", 4000, 160.0f, 4.0f, 1, 6, 6, 4, 256, 0, 0, 0, 10.0f},
      {"padding-edges", 3000, 200.0f, D, 1, 12, 5, 0, 0, 0, 0, 2, 10.0f},
      {"padding-edges-global", 3000, 200.0f, D, 1, 6, 5, 0, 0, 1, 0, 2, 10.0f},
      {"radius-5x5", 2500, 160.0f, 4.0f, 1, 6, 6, 0, 0, 0, 0, 0, 10.0f},
      {"fine-cells-global", 500, 10.0f, 0.5f, 1, 5, 7, 0, 0, 0, 0, 0, 10.0f},
      {"non-toroidal", 1500, 150.0f, D, 0, 6, 8, 0, 0, 0, 0, 0, 10.0f},
      {"partial-cells", 1200, 151.3f, 6.1f, 1, 8, 9, 0, 0, 0, 0, 0, 10.0f},
      // w = 1024, d = 1.1596829: the largest float below w discretizes to
      // ceil(w/d) = 883, the padding column/row no window scans
      {"padding-reachable", 4000, 1024.0f, 1.1596828699111938f, 1, 10, 10, 0, 0, 0, 0, 2, 2.0f},
      {"padding-reachable-global", 4000, 1024.0f, 1.1596828699111938f, 1, 4, 10, 0, 0, 1, 0, 2,
       2.0f},
      {"stacked-boundaries", 3000, 200.0f, D, 1, 6, 15, 0, 0, 0, 0, 3, 10.0f},
      {"stacked-boundaries-global", 3000, 200.0f, D, 1, 4, 15, 0, 0, 1, 0, 3, 10.0f},
      {"empty", 0, 200.0f, D, 1, 3, 16, 0, 0, 0, 0, 0, 10.0f},
      // the scatter and the bin sort every step
      {"flockers-200-no-gather", 1000, 200.0f, D, 1, 20, 0, 0, 0, 0, 0, 0, 10.0f, 0.7f, 0},
      {"clustered-no-gather", 6000, 700.0f, D, 1, 4, 4, 0, 0, 0, 3, 1, 10.0f, 0.7f, 0},
      // jumps of several cells: migrants from far tiles, through the inboxes
      {"big-jumps", 3000, 600.0f, D, 1, 10, 17, 0, 0, 0, 5, 0, 10.0f, 45.0f},
      // most birds change tile every step: outboxes past their capacity send
      // steps to the scatter and the bin sort
      {"outbox-overflow", 6000, 200.0f, D, 1, 6, 18, 0, 0, 0, 0, 0, 10.0f, 60.0f},
      // jumps longer than a tile: birds from beyond the neighbouring tiles
      // leave counts short, and those steps go to the scatter and the bin sort
      {"far-jumps", 3000, 1200.0f, D, 1, 6, 20, 0, 0, 0, 0, 0, 10.0f, 150.0f},
      {"non-toroidal-shuffled", 1500, 150.0f, D, 0, 6, 8, 0, 0, 0, 11, 0, 10.0f},
      // tiles of more than 2,048 birds: every step goes to the scatter and the bin sort
      {"dense-tiles", 40000, 700.0f, D, 1, 2, 7, 0, 0, 0, 3, 1, 10.0f},
      // a narrow last tile column: birds cross the seam between tile
      // columns 0 and n - 2 (20,480 / (10 / 1.5) cells: the 1M field's seam)
      {"seam-sliver", 4000, 20480.0f / 16.0f, D / 16.0f, 1, 8, 19, 0, 0, 0, 0, 4, 10.0f / 16.0f,
       0.7f / 16.0f, 1, 1},
      {"seam-sliver-shuffled", 4000, 20480.0f / 16.0f, D / 16.0f, 1, 4, 21, 0, 0, 0, 9, 4,
       10.0f / 16.0f, 0.7f / 16.0f, 1, 1},
  };
  int fails = 0;
  for (const Scenario& s : sc) fails += run(s);
  const RtScenario rs[] = {
      {"rt-flockers-200", 1000, 200.0f, D, 60, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 10.0f, 0.7f},
      {"rt-flockers-200-shuffled", 1000, 200.0f, D, 30, 0, 0, 0, 0, 0, 0, 0, 0, 77, 0, 10.0f, 0.7f},
      {"rt-small-tiles", 1500, 200.0f, D, 30, 1, 4, 8, 2, 0, 0, 0, 0, 0, 0, 10.0f, 0.7f},
      {"rt-sparse", 10000, 2048.0f, D, 10, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 10.0f, 0.7f},
      {"rt-dense", 1200, 60.0f, D, 3, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 10.0f, 0.7f},
      {"rt-clustered", 6000, 700.0f, D, 10, 4, 0, 0, 0, 0, 0, 0, 0, 5, 1, 10.0f, 0.7f},
      {"rt-clustered-global", 2000, 400.0f, D, 4, 4, 0, 0, 0, 96, 64, 8, 0, 0, 1, 10.0f, 0.7f},
      {"rt-tight-capacity", 2000, 300.0f, D, 8, 6, 0, 0, 0, 0, 0, 0, 1, 3, 0, 10.0f, 0.7f},
      {"rt-padding-edges", 3000, 200.0f, D, 12, 5, 0, 0, 0, 0, 0, 0, 0, 0, 2, 10.0f, 0.7f},
      {"rt-padding-reachable", 4000, 1024.0f, 1.1596828699111938f, 10, 10, 0, 0, 0, 0, 0, 0, 0,
       0, 2, 2.0f, 0.7f},
      {"rt-radius-5x5", 2500, 160.0f, 4.0f, 8, 6, 0, 0, 0, 0, 0, 0, 0, 0, 0, 10.0f, 0.7f},
      {"rt-big-jumps", 2000, 300.0f, D, 10, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0, 10.0f, 9.0f},
      // tiny normal-class LDS budget: most tiles run in the dense class
      {"rt-class-split", 1500, 200.0f, D, 30, 11, 0, 0, 0, 0, 0, 0, 2, 0, 0, 10.0f, 0.7f},
      {"rt-class-split-clustered", 6000, 700.0f, D, 10, 12, 0, 0, 0, 0, 0, 0, 2, 9, 1, 10.0f,
       0.7f},
      // normal-class capacities far below the counts: tiles defer to the
      // dense launch (past the deferral list's capacity they take the global path)
      {"rt-deferred", 3000, 300.0f, D, 12, 13, 0, 0, 0, 64, 24, 8, 4, 0, 1, 10.0f, 0.7f},
      {"rt-deferred-shuffled", 1500, 200.0f, D, 10, 14, 0, 0, 0, 96, 40, 16, 4, 21, 0, 10.0f,
       0.7f},
      {"rt-stacked-boundaries", 3000, 200.0f, D, 20, 15, 0, 0, 0, 0, 0, 0, 0, 0, 3, 10.0f, 0.7f},
      {"rt-empty", 0, 200.0f, D, 3, 16, 0, 0, 0, 0, 0, 0, 0, 0, 0, 10.0f, 0.7f},
  };
  for (const RtScenario& s : rs) fails += run_rt(s);
  std::printf("%d scenario(s) failed\n", fails);
  return fails ? 1 : 0;
}
