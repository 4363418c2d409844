// Tests of the C++ host interface (include/krabgpu.hpp).
//
// --cpu: Schedule::step semantics (src/engine/schedule.rs:347-413), the
//        simulate! loop (src/lib.rs:1158-1173) with host-only agents, and
//        the exploration front end's sweeps, gen_param, number text and
//        CSV (include/krabgpu_explore.hpp), and the serving loop's
//        admission and ticks with a stand-in ensemble; no device call.
// (no flag): additionally the reference's own Field2D tests
//        (tests/engine/field_2d.rs:30-247) through the C++ Field2D;
//        flockers runs through Schedule/simulate and through Flocker::run,
//        compared bit for bit with the oracle; Ensemble runs (one retired
//        on the way) against the same runs alone; parallel and sequential
//        sweeps against each other and against runs alone; and served runs
//        against the same runs alone.
// Prints one line per test and "<k> failed"; exit status 0 when none failed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "krabgpu.hpp"
#include "krabgpu_explore.hpp"
#include "krabgpu_serve.hpp"
#include "oracle.h"

using krabgpu::Real2D;

static int g_failed = 0;
static int g_checks_failed = 0;

#define EXPECT(c)                                                         \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::printf("    check failed at line %d: %s\n", __LINE__, #c);     \
      ++g_checks_failed;                                                  \
    }                                                                     \
  } while (0)

static void run_test(const char* name, const std::function<void()>& fn) {
  const int before = g_checks_failed;
  try {
    fn();
  } catch (const std::exception& e) {
    std::printf("    exception: %s\n", e.what());
    ++g_checks_failed;
  }
  const bool ok = g_checks_failed == before;
  if (!ok) ++g_failed;
  std::printf("%-44s %s\n", name, ok ? "ok" : "FAILED");
  std::fflush(stdout);
}

static std::vector<uint32_t> sorted(std::vector<uint32_t> v) {
  std::sort(v.begin(), v.end());
  return v;
}

// ---- Schedule semantics, host only ---------------------------------------

struct Log {
  std::vector<std::string> ev;
};

struct LogState : krabgpu::State {
  Log* log;
  uint64_t stop_after;
  explicit LogState(Log* l, uint64_t stop = ~0ull) : log(l), stop_after(stop) {}
  std::vector<std::pair<std::shared_ptr<krabgpu::Agent>, std::pair<float, int>>> agents;
  bool repeating = true;
  void init(krabgpu::Schedule& s) override {
    log->ev.push_back("init");
    for (auto& a : agents) {
      if (repeating)
        s.schedule_repeating(a.first, a.second.first, a.second.second);
      else
        s.schedule_once(a.first, a.second.first, a.second.second);
    }
  }
  void update(uint64_t step) override { log->ev.push_back("update" + std::to_string(step)); }
  void before_step(krabgpu::Schedule&) override { log->ev.push_back("before"); }
  void after_step(krabgpu::Schedule&) override { log->ev.push_back("after"); }
  bool end_condition(krabgpu::Schedule& s) override { return s.step_count >= stop_after; }
};

struct LogAgent : krabgpu::Agent {
  Log* log;
  std::string name;
  int stop_at = -1;  // is_stopped after this many steps
  int steps = 0;
  LogAgent(Log* l, std::string n, int stop = -1) : log(l), name(std::move(n)), stop_at(stop) {}
  void step(krabgpu::State&) override {
    ++steps;
    log->ev.push_back(name);
  }
  bool is_stopped(krabgpu::State&) override { return stop_at >= 0 && steps >= stop_at; }
};

static std::string join(const std::vector<std::string>& v) {
  std::string s;
  for (const auto& x : v) s += (s.empty() ? "" : " ") + x;
  return s;
}

// A stand-in ensemble for the serving loop's host logic: each schedule step
// advances a fake clock by 1 ms.
struct FakeEnsemble {
  explicit FakeEnsemble(double* c) : clock(c) {}
  double* clock;
  uint32_t next = 0;
  std::vector<uint32_t> retired;
  std::vector<uint32_t> admit(const std::vector<kg_run_desc>& runs) {
    std::vector<uint32_t> ids;
    for (size_t k = 0; k < runs.size(); ++k) ids.push_back(next++);
    return ids;
  }
  void run(uint32_t n) { *clock += 0.001 * n; }
  void sync() {}
  krabgpu::Agents read(uint32_t id) {
    krabgpu::Agents a;
    a.id.push_back(id);
    return a;
  }
  void retire(uint32_t id) { retired.push_back(id); }
};

static kg_run_desc fake_run(uint32_t agents) {
  kg_run_desc r{};
  r.agents = agents;
  return r;
}

static void cpu_tests() {
  run_test("serving: admission, ticks, queueing, idle", [] {
    double t = 0.0;
    FakeEnsemble ens(&t);
    krabgpu::Server<FakeEnsemble> srv(250, 64, true);
    srv.footprint = [](const kg_run_desc& r) { return (uint64_t)r.agents; };
    srv.clock = [&t] { return t; };
    srv.sleep = [&t](double s) { t += s; };
    const std::vector<krabgpu::Request> reqs{{0.0, fake_run(100), 10},
                                             {0.0, fake_run(100), 30},
                                             {0.005, fake_run(100), 5},  // waits for memory
                                             {1.0, fake_run(50), 3}};   // after an idle gap
    const auto done = srv.serve(reqs, ens);
    EXPECT(done.size() == 4);
    if (done.size() == 4) {
      EXPECT(done[0].index == 0 && done[1].index == 2 && done[2].index == 1 && done[3].index == 3);
      EXPECT(std::fabs(done[0].done - 0.010) < 1e-9);
      EXPECT(std::fabs(done[1].admitted - 0.010) < 1e-9 && std::fabs(done[1].done - 0.015) < 1e-9);
      EXPECT(std::fabs(done[2].done - 0.030) < 1e-9);
      EXPECT(done[3].admitted >= 1.0 && std::fabs(done[3].done - done[3].admitted - 0.003) < 1e-9);
      EXPECT(done[1].state.id.size() == 1 && done[1].state.id[0] == 2);
    }
    EXPECT(srv.ticks == 4 && srv.admissions == 3 && srv.peak_bytes == 200);
    EXPECT(ens.retired == (std::vector<uint32_t>{0, 2, 1, 3}));
    bool threw = false;
    try {
      krabgpu::Server<FakeEnsemble> small(99);
      small.footprint = srv.footprint;
      FakeEnsemble e2(&t);
      small.serve({{0.0, fake_run(100), 1}}, e2);
    } catch (const krabgpu::Error& e) {
      threw = e.code() == KG_E_ADMIT;
    }
    EXPECT(threw);
  });

  run_test("schedule: order, update timing", [] {
    Log log;
    LogState st(&log);
    // ascending (time, ordering); equal priorities in scheduling order
    st.agents = {{std::make_shared<LogAgent>(&log, "b"), {0.0f, 1}},
                 {std::make_shared<LogAgent>(&log, "a"), {0.0f, 0}},
                 {std::make_shared<LogAgent>(&log, "c"), {0.0f, 1}},
                 {std::make_shared<LogAgent>(&log, "late"), {1.0f, -5}}};
    krabgpu::simulate(st, 2);
    EXPECT(join(log.ev) ==
           "init update0 before a b c after update1 before late a b c after update2");
  });
  run_test("schedule: once, stopped agents, empty queue", [] {
    Log log;
    LogState st(&log);
    st.repeating = false;
    st.agents = {{std::make_shared<LogAgent>(&log, "x"), {0.0f, 0}}};
    krabgpu::simulate(st, 3);
    // an event scheduled once runs once; later steps find the queue empty
    EXPECT(join(log.ev) ==
           "init update0 before x after update1 before after update2 before after update3");
    Log log2;
    LogState st2(&log2);
    st2.agents = {{std::make_shared<LogAgent>(&log2, "s", 2), {0.0f, 0}}};
    krabgpu::simulate(st2, 3);
    EXPECT(join(log2.ev) ==
           "init update0 before s after update1 before s after update2 before after update3");
  });
  run_test("explore: exhaustive order, matched rules", [] {
    using krabgpu::ExploreMode;
    const krabgpu::Inputs in{{"agents", {100, 200}}, {"jump", {0.25, 0.5, 0.75}}};
    const auto c = krabgpu::configurations(in, ExploreMode::Exaustive);
    EXPECT(c.size() == 6);
    size_t k = 0;
    for (double a : {100.0, 200.0})
      for (double j : {0.25, 0.5, 0.75}) {
        EXPECT(c[k][0].second == a && c[k][1].second == j);
        ++k;
      }
    const auto m = krabgpu::configurations({{"a", {1, 2, 3, 4}}, {"b", {10, 20}}},
                                           ExploreMode::Matched);
    EXPECT(m.size() == 2 && m[1][0].second == 2 && m[1][1].second == 20);
    bool threw = false;
    try {
      krabgpu::configurations({{"a", {1}}, {"b", {10, 20}}}, ExploreMode::Matched);
    } catch (const krabgpu::Error&) {
      threw = true;
    }
    EXPECT(threw);
    threw = false;
    try {
      krabgpu::configurations({{"a", {}}}, ExploreMode::Exaustive);
    } catch (const krabgpu::Error&) {
      threw = true;
    }
    EXPECT(threw);
    EXPECT(krabgpu::configurations({}, ExploreMode::Exaustive).size() == 1);
  });
  run_test("explore: gen_param ranges", [] {
    std::mt19937_64 rng(5);
    const auto v = krabgpu::gen_param<int>(10, 0, 1000, rng);  // bounds swapped
    EXPECT(*std::min_element(v.begin(), v.end()) == 0 && *std::max_element(v.begin(), v.end()) == 9);
    const auto e = krabgpu::gen_param<int>(7, 7, 50, rng);     // [7, 8)
    EXPECT(std::all_of(e.begin(), e.end(), [](int x) { return x == 7; }));
    const auto f = krabgpu::gen_param<double>(2.0, 2.0, 100, rng);
    EXPECT(std::all_of(f.begin(), f.end(), [](double x) { return x >= 2.0 && x < 3.0; }));
    EXPECT(krabgpu::gen_param<float>(0.0f, 1.0f, 0, rng).empty());
  });
  run_test("explore: Rust number text", [] {
    const std::pair<double, const char*> dbg[] = {
        {1.0, "1.0"}, {0.1, "0.1"}, {0.0, "0.0"}, {-2.5, "-2.5"}, {1e16, "1e16"},
        {1.5e-7, "1.5e-7"}, {1e-4, "0.0001"}, {123456.789, "123456.789"},
        {9.99e15, "9990000000000000.0"}, {-2.5e20, "-2.5e20"}};
    for (const auto& t : dbg) EXPECT(krabgpu::rust_debug(t.first) == t.second);
    EXPECT(krabgpu::rust_debug(std::nan("")) == "NaN");
    EXPECT(krabgpu::rust_debug_f32(16777217.0f) == "16777216.0");
    const std::pair<float, const char*> disp[] = {
        {1.0f, "1"}, {0.1f, "0.1"}, {3.25f, "3.25"}, {1e20f, "100000000000000000000"},
        {1234.5678f, "1234.5677"}, {0.0f, "0"}};
    for (const auto& t : disp) EXPECT(krabgpu::rust_display_f32(t.first) == t.second);
  });
  run_test("explore: write_csv text", [] {
    std::vector<krabgpu::FrameRow> rows(2);
    for (uint32_t j = 0; j < 2; ++j) {
      rows[j].conf_num = 0;
      rows[j].conf_rep = j;
      rows[j].inputs = {{"agents", 100}, {"jump", 0.25}};
      rows[j].outputs = {{"polarisation", 0.5 + 0.125 * j}};
      rows[j].run_duration = 0.5f;
      rows[j].step_per_sec = 1000.0f * (float)(j + 1);
    }
    const std::string name = "/tmp/krabgpu_explore_test";
    krabgpu::write_csv(name, rows);
    FILE* f = std::fopen((name + ".csv").c_str(), "r");
    char line[256];
    std::vector<std::string> lines;
    while (f && std::fgets(line, sizeof(line), f)) lines.push_back(line);
    if (f) std::fclose(f);
    const auto cols = krabgpu::load_csv(name + ".csv", 7);
    std::remove((name + ".csv").c_str());
    EXPECT(cols.size() == 7 && cols[1] == (std::vector<double>{0, 1}));
    EXPECT(cols[4] == (std::vector<double>{0.5, 0.625}) && cols[6] == (std::vector<double>{1000, 2000}));
    EXPECT(lines.size() == 3);
    if (lines.size() == 3) {
      EXPECT(lines[0] == "Simulation,Run,agents,jump,polarisation,Run Duration,Step per sec\n");
      EXPECT(lines[1] == "0,0,100,0.25,0.5,0.5,1000\n");
      EXPECT(lines[2] == "0,1,100,0.25,0.625,0.5,2000\n");
    }
  });
  run_test("schedule: 200,000 agents", [] {
    struct Count : krabgpu::Agent {
      uint32_t* hits;
      explicit Count(uint32_t* h) : hits(h) {}
      void step(krabgpu::State&) override { ++*hits; }
    };
    struct Many : krabgpu::State {
      std::vector<uint32_t> hits = std::vector<uint32_t>(200000, 0);
      void init(krabgpu::Schedule& s) override {
        for (size_t k = 0; k < hits.size(); ++k)
          s.schedule_repeating(std::make_shared<Count>(&hits[k]), 0.0f, (int32_t)(k % 7));
      }
    } st;
    krabgpu::simulate(st, 3);
    EXPECT(std::all_of(st.hits.begin(), st.hits.end(), [](uint32_t h) { return h == 3; }));
  });
  run_test("simulate: end_condition and repetitions", [] {
    Log log;
    LogState st(&log, 2);
    st.agents = {{std::make_shared<LogAgent>(&log, "a"), {0.0f, 0}}};
    krabgpu::simulate(st, 10, 2);
    const std::string one = "init update0 before a after update1 before a after update2";
    EXPECT(join(log.ev) == one + " " + one);
  });
}

// ---- device tests ---------------------------------------------------------

static krabgpu::Field2D make_small(uint32_t cap = 16) {
  // tests/engine/field_2d.rs fixture: 10 x 10, discretization 0.5, toroidal
  return krabgpu::Field2D(10.0f, 10.0f, 0.5f, true, cap);
}

static kg_flock_params params(float radius, uint64_t seed, float jump = 0.7f) {
  kg_flock_params p;
  p.cohesion = p.avoidance = p.consistency = p.randomness = p.momentum = 1.0f;
  p.jump = jump;
  p.radius = radius;
  p.reserved = 0;
  p.seed = seed;
  return p;
}

static void field_tests() {
  run_test("field_2d_single_step (field_2d.rs:30-50)", [] {
    auto f = make_small(10);
    f.init_uniform(10, 7, 0);
    f.schedule_run(params(10.0f, 7), 0, 1);
    EXPECT(f.nagents() == 10);
    EXPECT(f.get_neighbors_within_distance({5.0f, 5.0f}, 10.0f).size() == 10);
    EXPECT(f.get_neighbors_within_relax_distance({5.0f, 5.0f}, 10.0f).size() == 10);
  });
  run_test("field_2d_neighbors (field_2d.rs:57-117)", [] {
    for (int fly = 5; fly < 10; ++fly) {
      auto f = make_small();
      f.set_object_location(1, {0.0f, 0.0f});
      f.set_object_location(2, {0.0f, 0.0f});
      f.lazy_update();
      EXPECT(f.nagents() == 2);
      EXPECT(f.get_neighbors_within_distance({5.0f, 5.0f}, 1.0f).empty());
      EXPECT(f.get_neighbors_within_relax_distance({5.0f, 5.0f}, 1.0f).empty());
      const Real2D b1{(float)fly, (float)fly};
      f.set_object_location(1, b1);
      f.set_object_location(2, {0.0f, 0.0f});
      f.lazy_update();
      EXPECT(f.get_neighbors_within_distance(b1, 1.0f) == std::vector<uint32_t>{1});
      EXPECT(f.get_neighbors_within_distance({0.0f, 0.0f}, 1.0f) == std::vector<uint32_t>{2});
      EXPECT(sorted(f.get_neighbors_within_distance({5.0f, 5.0f}, 10.0f)) ==
             (std::vector<uint32_t>{1, 2}));
      EXPECT(sorted(f.get_neighbors_within_relax_distance({5.0f, 5.0f}, 10.0f)) ==
             (std::vector<uint32_t>{1, 2}));
    }
  });
  run_test("field_2d_gets (field_2d.rs:124-178)", [] {
    auto f = make_small();
    f.set_object_location(1, {0.0f, 0.0f});
    f.set_object_location(2, {5.0f, 5.0f});
    f.set_object_location(3, {5.0f, 5.0f});
    f.lazy_update();
    EXPECT(f.nagents() == 3);
    EXPECT(sorted(f.get_objects({5.0f, 5.0f})) == (std::vector<uint32_t>{2, 3}));
    EXPECT(f.get_objects({10.0f, 0.0f}).empty());
    EXPECT(f.num_objects_at_location({5.0f, 5.0f}) == 2);
    EXPECT(f.num_objects_at_location({0.0f, 0.0f}) == 1);
    f.set_object_location(4, {0.0f, 0.0f});
    EXPECT(f.get_objects_unbuffered({0.0f, 0.0f}) == std::vector<uint32_t>{4});
    EXPECT(f.get_objects({0.0f, 0.0f}) == std::vector<uint32_t>{1});
    f.remove_object_location(4, {0.0f, 0.0f});
    EXPECT(f.get_objects_unbuffered({0.0f, 0.0f}).empty());
  });
  run_test("field_2d_bags (field_2d.rs:185-210)", [] {
    auto f = make_small();
    EXPECT(f.num_empty_bags() == 21 * 21);
    f.set_object_location(1, {0.0f, 0.0f});
    f.set_object_location(2, {0.0f, 0.0f});
    f.set_object_location(3, {4.0f, 4.0f});
    f.lazy_update();
    EXPECT(f.num_empty_bags() == 21 * 21 - 2);
  });
  run_test("field_2d_iter (field_2d.rs:217-247)", [] {
    auto f = make_small();
    f.set_object_location(1, {0.0f, 0.0f});
    f.set_object_location(2, {0.01f, 0.01f});
    f.set_object_location(3, {5.0f, 5.0f});
    const krabgpu::Agents wb = f.read_agents(1);
    for (size_t k = 0; k < wb.id.size(); ++k) {
      const auto o = f.get_objects_unbuffered({wb.x[k], wb.y[k]});
      EXPECT(std::find(o.begin(), o.end(), wb.id[k]) != o.end());
    }
    f.lazy_update();
    EXPECT(f.get_objects({0.0f, 0.0f}).size() == 2);
    const krabgpu::Agents rb = f.read_agents(0);
    EXPECT(sorted(rb.id) == (std::vector<uint32_t>{1, 2, 3}));
  });
  run_test("out-of-grid location throws", [] {
    auto f = make_small();
    int code = 0;
    try {
      f.set_object_location(1, {20.0f, 1.0f});
    } catch (const krabgpu::Error& e) {
      code = e.code();
    }
    EXPECT(code == KG_E_RANGE);
  });
}

// (4, n) state by id from a device buffer
static std::vector<float> by_id(const krabgpu::Agents& a, uint32_t n) {
  std::vector<float> s(4 * (size_t)n, NAN);
  for (size_t k = 0; k < a.id.size(); ++k) {
    const uint32_t i = a.id[k];
    if (i >= n) continue;
    s[i] = a.x[k];
    s[n + i] = a.y[k];
    s[2 * (size_t)n + i] = a.hx[k];
    s[3 * (size_t)n + i] = a.hy[k];
  }
  return s;
}

static std::vector<float> oracle_run(uint32_t n, float side, float d, uint64_t seed,
                                     uint32_t steps) {
  std::vector<float> x(n), y(n), hx(n), hy(n);
  or_init_uniform(n, side, side, seed, 1, x.data(), y.data(), hx.data(), hy.data());
  or_flock_params op{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0.7f, 10.0f, seed};
  or_flocker* fl = or_flocker_new(n, side, side, d, 1, x.data(), y.data(), hx.data(), hy.data(),
                                  &op);
  for (uint32_t k = 0; k < steps; ++k) or_flocker_step(fl);
  std::vector<float> s(4 * (size_t)n);
  or_flocker_get(fl, s.data(), s.data() + n, s.data() + 2 * (size_t)n, s.data() + 3 * (size_t)n);
  or_flocker_free(fl);
  return s;
}

static bool same_bits(const std::vector<float>& a, const std::vector<float>& b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * 4) == 0;
}

static void flockers_tests() {
  const float d = 10.0f / 1.5f;
  run_test("flockers: simulate through Schedule", [&] {
    const uint32_t n = 1000, steps = 12;
    krabgpu::Flocker fl({200.0f, 200.0f}, n, d, params(10.0f, 3));
    krabgpu::simulate(fl, steps);
    EXPECT(fl.step == steps);
    EXPECT(same_bits(by_id(fl.field1.read_agents(0), n), oracle_run(n, 200.0f, d, 3, steps)));
  });
  run_test("flockers: repetitions restart from init", [&] {
    const uint32_t n = 800, steps = 5;
    krabgpu::Flocker fl({180.0f, 180.0f}, n, d, params(10.0f, 4));
    krabgpu::simulate(fl, steps, 2);
    EXPECT(fl.field1.count(0) == n && fl.field1.count(1) == 0);
    EXPECT(same_bits(by_id(fl.field1.read_agents(0), n), oracle_run(n, 180.0f, d, 4, steps)));
  });
  run_test("flockers: Flocker::run equals simulate", [&] {
    const uint32_t n = 5000, steps = 40;
    krabgpu::Flocker a({450.0f, 450.0f}, n, d, params(10.0f, 5));
    a.run(steps);
    EXPECT(a.step == steps);
    EXPECT(same_bits(by_id(a.field1.read_agents(0), n), oracle_run(n, 450.0f, d, 5, steps)));
  });
  run_test("ensemble: runs equal the same runs alone", [&] {
    struct R {
      uint32_t n;
      float side;
      uint64_t seed;
    };
    const R rs[] = {{700, 170.0f, 11}, {3000, 350.0f, 12}, {1500, 240.0f, 13}};
    const uint32_t steps = 30;
    krabgpu::Ensemble ens;
    std::vector<kg_run_desc> descs;
    for (const R& r : rs)
      descs.push_back(krabgpu::Ensemble::run_desc({r.side, r.side}, r.n, d, params(10.0f, r.seed)));
    EXPECT(krabgpu::Ensemble::footprint(descs[0]) > 0);
    const std::vector<uint32_t> ids = ens.admit(descs);
    ens.run(steps);
    for (size_t k = 0; k < 3; ++k) {
      EXPECT(ens.steps(ids[k]) == steps);
      krabgpu::Flocker alone({rs[k].side, rs[k].side}, rs[k].n, d, params(10.0f, rs[k].seed));
      alone.run(steps);
      EXPECT(same_bits(by_id(ens.read(ids[k]), rs[k].n), by_id(alone.field1.read_agents(0), rs[k].n)));
    }
    ens.retire(ids[1]);
    ens.run(5);
    krabgpu::Flocker alone({rs[2].side, rs[2].side}, rs[2].n, d, params(10.0f, rs[2].seed));
    alone.run(steps + 5);
    EXPECT(same_bits(by_id(ens.read(ids[2]), rs[2].n), by_id(alone.field1.read_agents(0), rs[2].n)));
  });
  run_test("explore: parallel and sequential sweeps", [&] {
    const krabgpu::Inputs in{{"agents", {600, 1500}}, {"jump", {0.4, 0.9}}, {"cohesion", {0.5, 1.7}}};
    const std::vector<std::pair<std::string, krabgpu::Output>> outs{{"polarisation", krabgpu::polarisation}};
    krabgpu::ExploreOptions po;
    po.seed = 3;
    po.keep_state = true;
    po.hbm_budget = 900000;  // several batches
    krabgpu::ExploreOptions so = po;
    so.parallel = false;
    const uint32_t steps = 20;
    const auto par = krabgpu::explore(steps, 2, in, outs, krabgpu::ExploreMode::Exaustive, po);
    const auto seq = krabgpu::explore(steps, 2, in, outs, krabgpu::ExploreMode::Exaustive, so);
    EXPECT(par.size() == 16 && seq.size() == 16);
    for (size_t k = 0; k < par.size() && k < seq.size(); ++k) {
      EXPECT(par[k].conf_num == k / 2 && par[k].conf_rep == k % 2);
      EXPECT(seq[k].conf_rep == k % 2 + 1);
      const uint32_t n = (uint32_t)par[k].inputs[0].second;
      const auto sp = by_id(par[k].state, n);
      EXPECT(same_bits(sp, by_id(seq[k].state, n)));
      EXPECT(par[k].outputs[0].second == seq[k].outputs[0].second);
      const kg_run_desc d = krabgpu::flockers_run(par[k].inputs, (3ull << 32) | k);
      krabgpu::Flocker alone({d.width, d.height}, d.agents, d.discretization, d.params);
      alone.run(steps);
      EXPECT(same_bits(sp, by_id(alone.field1.read_agents(0), n)));
    }
  });
  run_test("serving: each run equals the run alone", [&] {
    struct R {
      uint32_t n;
      float side;
      uint64_t seed;
      uint32_t steps;
      double arrival;
    };
    const R rs[] = {{900, 190.0f, 21, 40, 0.0},  {2500, 320.0f, 22, 12, 0.0},
                    {600, 160.0f, 23, 25, 0.002}, {4000, 400.0f, 24, 30, 0.004},
                    {1200, 220.0f, 25, 7, 0.006}};
    std::vector<krabgpu::Request> reqs;
    uint64_t biggest = 0;
    for (const R& r : rs) {
      reqs.push_back({r.arrival,
                      krabgpu::Ensemble::run_desc({r.side, r.side}, r.n, d, params(10.0f, r.seed)),
                      r.steps});
      biggest = std::max(biggest, krabgpu::Ensemble::footprint(reqs.back().run));
    }
    krabgpu::Ensemble ens;
    krabgpu::Server<> srv(biggest * 2, 16, true);  // some requests wait for memory
    const auto done = srv.serve(reqs, ens);
    EXPECT(done.size() == 5);
    for (const auto& c : done) {
      const R& r = rs[c.index];
      krabgpu::Flocker alone({r.side, r.side}, r.n, d, params(10.0f, r.seed));
      alone.run(r.steps);
      EXPECT(same_bits(by_id(c.state, r.n), by_id(alone.field1.read_agents(0), r.n)));
      EXPECT(c.latency() >= 0.0);
    }
    EXPECT(srv.admissions >= 2);
  });
  run_test("flockers: no birds", [&] {
    krabgpu::Flocker fl({200.0f, 200.0f}, 0, d, params(10.0f, 6));
    krabgpu::simulate(fl, 3);
    EXPECT(fl.field1.count(0) == 0 && fl.step == 3);
    fl.run(3);
    EXPECT(fl.field1.count(0) == 0);
  });
}

int main(int argc, char** argv) {
  const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
  cpu_tests();
  if (!cpu_only) {
    field_tests();
    flockers_tests();
  }
  std::printf("%d failed\n", g_failed);
  return g_failed ? 1 : 0;
}
