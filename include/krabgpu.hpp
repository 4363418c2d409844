/* krabgpu.hpp -- C++ host interface over the C ABI in krabgpu.h, with the
 * reference's names: Field2D (src/engine/fields/field_2d.rs), the State and
 * Agent traits (src/engine/state.rs:8-30, src/engine/agent.rs:7-40),
 * Schedule (src/engine/schedule.rs:227-413), simulate! (src/lib.rs:
 * 1158-1173) and the flockers State (tests/model/flockers/state.rs).
 * Header-only; link libkrabgpu.so.  Errors throw krabgpu::Error where the
 * reference panics.
 *
 * The birds of a flockers model live on the device, so a Flocker schedules
 * one swarm agent that stands for all of them: its step is the agent loop of
 * Schedule::step (every Bird::step of that schedule step, schedule.rs:
 * 395-408) as one kg_flock_step, and Flocker::update is field1.lazy_update()
 * (state.rs:57-59).  The results are those of kg_schedule_run, which
 * Flocker::run uses to take the same steps in fewer launches. */
#ifndef KRABGPU_HPP
#define KRABGPU_HPP

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "krabgpu.h"

namespace krabgpu {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != KG_OK) throw Error(rc, std::string("krabgpu: ") + kg_last_error());
}

/* src/engine/location.rs:17-20 */
struct Real2D {
  float x, y;
};

/* The agents of one buffer, in storage order. */
struct Agents {
  std::vector<uint32_t> id;
  std::vector<float> x, y, hx, hy;
};

/* Field2D::new(w, h, d, toroidal) in HBM (field_2d.rs:304-322); objects are
 * bird ids carrying a position and a heading.  `capacity` bounds the agents
 * of each buffer. */
class Field2D {
 public:
  Field2D(float w, float h, float d, bool toroidal, uint32_t capacity) {
    kg_field_desc desc{w, h, d, toroidal ? 1 : 0, capacity, 0, 0, 0};
    check(kg_field_create(&desc, &f_));
  }
  ~Field2D() {
    if (f_) kg_field_destroy(f_);
  }
  Field2D(const Field2D&) = delete;
  Field2D& operator=(const Field2D&) = delete;
  Field2D(Field2D&& o) noexcept : f_(std::exchange(o.f_, nullptr)) {}
  Field2D& operator=(Field2D&& o) noexcept {
    if (this != &o) {
      if (f_) kg_field_destroy(f_);
      f_ = std::exchange(o.f_, nullptr);
    }
    return *this;
  }

  /* field_2d.rs:838-846: into the write buffer; a location outside the bag
   * grid throws (the reference panics). */
  void set_object_location(uint32_t id, Real2D loc, Real2D heading = {0.0f, 0.0f}) {
    check(kg_set_object_locations(f_, &id, &loc.x, &loc.y, &heading.x, &heading.y, 1));
  }
  /* Many set_object_location calls in one transfer. */
  void set_object_locations(const std::vector<uint32_t>& ids, const std::vector<float>& x,
                            const std::vector<float>& y, const std::vector<float>& hx,
                            const std::vector<float>& hy) {
    const size_t n = ids.size();
    if (x.size() != n || y.size() != n || hx.size() != n || hy.size() != n)
      throw Error(KG_E_INVALID, "krabgpu: set_object_locations: arrays of unequal length");
    check(kg_set_object_locations(f_, ids.data(), x.data(), y.data(), hx.data(), hy.data(), n));
  }
  /* field_2d.rs:885-898 */
  void remove_object_location(uint32_t id, Real2D loc) {
    check(kg_remove_object_location(f_, id, loc.x, loc.y));
  }
  /* Field::lazy_update (field_2d.rs:905-921) */
  void lazy_update() { check(kg_lazy_update(f_)); }

  /* field_2d.rs:386-440 */
  std::vector<uint32_t> get_neighbors_within_distance(Real2D loc, float dist) {
    return ids_call([&](uint32_t* p, size_t cap, size_t* n) {
      return kg_get_neighbors_within_distance(f_, loc.x, loc.y, dist, p, cap, n);
    });
  }
  /* field_2d.rs:472-516 */
  std::vector<uint32_t> get_neighbors_within_relax_distance(Real2D loc, float dist) {
    return ids_call([&](uint32_t* p, size_t cap, size_t* n) {
      return kg_get_neighbors_within_relax_distance(f_, loc.x, loc.y, dist, p, cap, n);
    });
  }
  /* field_2d.rs:544-582 */
  std::vector<uint32_t> get_objects(Real2D loc) {
    return ids_call([&](uint32_t* p, size_t cap, size_t* n) {
      return kg_get_objects(f_, loc.x, loc.y, 0, p, cap, n);
    });
  }
  std::vector<uint32_t> get_objects_unbuffered(Real2D loc) {
    return ids_call([&](uint32_t* p, size_t cap, size_t* n) {
      return kg_get_objects(f_, loc.x, loc.y, 1, p, cap, n);
    });
  }
  /* field_2d.rs:803-808 */
  size_t num_objects_at_location(Real2D loc) {
    size_t n = 0;
    check(kg_num_objects_at_location(f_, loc.x, loc.y, &n));
    return n;
  }
  /* field_2d.rs:727-739, count only */
  size_t num_empty_bags() {
    size_t n = 0;
    check(kg_num_empty_bags(f_, &n));
    return n;
  }
  /* field_2d.rs:727-762: not_discretize of every empty bag, bag index order */
  std::vector<Real2D> get_empty_bags() {
    std::vector<float> x(64), y(64);
    for (;;) {
      size_t n = 0;
      const int rc = kg_get_empty_bags(f_, x.data(), y.data(), x.size(), &n);
      if (rc == KG_E_CAPACITY) {
        x.resize(n);
        y.resize(n);
        continue;
      }
      check(rc);
      std::vector<Real2D> out(n);
      for (size_t i = 0; i < n; ++i) out[i] = Real2D{x[i], y[i]};
      return out;
    }
  }
  /* field_2d.rs:764-800 with a seeded draw: false (None) if no bag is empty */
  bool get_random_empty_bag(uint64_t seed, Real2D* out) {
    int found = 0;
    check(kg_get_random_empty_bag(f_, seed, &out->x, &out->y, &found));
    return found != 0;
  }
  /* field_2d.rs:619-726: closure(bag position, id, position, heading) per
   * agent, bag by bag */
  template <class F>
  void iter_objects(F&& closure, bool unbuffered = false) {
    auto tramp = [](void* user, float bx, float by, uint32_t id, float x, float y, float hx,
                    float hy) {
      (*static_cast<F*>(user))(Real2D{bx, by}, id, Real2D{x, y}, Real2D{hx, hy});
    };
    check(kg_iter_objects(f_, unbuffered ? 1 : 0, tramp, &closure));
  }
  template <class F>
  void iter_objects_unbuffered(F&& closure) {
    iter_objects(std::forward<F>(closure), true);
  }
  /* Field2D::nagents (field_2d.rs:277) */
  size_t nagents() {
    size_t n = 0;
    check(kg_nagents(f_, &n));
    return n;
  }
  /* Agents in the read (0) or write (1) buffer. */
  size_t count(int buffer) {
    size_t n = 0;
    check(kg_count(f_, buffer, &n));
    return n;
  }

  /* Flocker::init's placement (state.rs:41-53) drawn on the device. */
  void init_uniform(uint32_t n, uint64_t seed, int heading_mode = 1) {
    check(kg_init_uniform(f_, n, seed, heading_mode));
  }
  /* Bird::step for every agent of the read buffer at schedule step `step`. */
  void flock_step(const kg_flock_params& p, uint64_t step) { check(kg_flock_step(f_, &p, step)); }
  /* `nsteps` Schedule::step calls of a flockers state from `first_step`. */
  void schedule_run(const kg_flock_params& p, uint64_t first_step, uint32_t nsteps,
                    int flags = KG_RUN_GRAPH) {
    check(kg_schedule_run(f_, &p, first_step, nsteps, flags));
  }
  void sync() { check(kg_field_sync(f_)); }

  Agents read_agents(int buffer) {
    size_t n = 0;
    const int rc = kg_read_agents(f_, buffer, nullptr, nullptr, nullptr, nullptr, nullptr, 0, &n);
    if (rc != KG_OK && rc != KG_E_CAPACITY) check(rc);
    Agents a;
    a.id.resize(n);
    a.x.resize(n);
    a.y.resize(n);
    a.hx.resize(n);
    a.hy.resize(n);
    check(kg_read_agents(f_, buffer, a.id.data(), a.x.data(), a.y.data(), a.hx.data(),
                         a.hy.data(), n, &n));
    return a;
  }

  kg_field* handle() { return f_; }

 private:
  template <class Fn>
  std::vector<uint32_t> ids_call(Fn fn) {
    std::vector<uint32_t> out(64);
    for (;;) {
      size_t n = 0;
      const int rc = fn(out.data(), out.size(), &n);
      if (rc == KG_E_CAPACITY) {
        out.resize(std::max<size_t>(n, out.size() * 2));
        continue;
      }
      check(rc);
      out.resize(n);
      return out;
    }
  }

  kg_field* f_ = nullptr;
};

class Schedule;

/* src/engine/state.rs:8-30 */
class State {
 public:
  virtual ~State() = default;
  virtual void init(Schedule& schedule) = 0;
  virtual void reset() {}
  virtual void update(uint64_t step) { (void)step; }
  virtual void before_step(Schedule& schedule) { (void)schedule; }
  virtual void after_step(Schedule& schedule) { (void)schedule; }
  virtual bool end_condition(Schedule& schedule) {
    (void)schedule;
    return false;
  }
};

/* src/engine/agent.rs:7-40 */
class Agent {
 public:
  virtual ~Agent() = default;
  virtual void step(State& state) = 0;
  virtual bool is_stopped(State& state) {
    (void)state;
    return false;
  }
  virtual void before_step(State& state) { (void)state; }
  virtual void after_step(State& state) { (void)state; }
};

/* src/engine/schedule.rs:227-413.  Events pop in ascending (time,
 * ordering) (the priority order of src/engine/priority.rs:21-38); equal
 * priorities pop in the order the agents were first scheduled. */
class Schedule {
 public:
  uint64_t step_count = 0;  // the reference's Schedule::step field
  float time = 0.0f;

  /* schedule.rs:295-303 */
  bool schedule_repeating(std::shared_ptr<Agent> agent, float the_time, int32_t the_ordering) {
    push(Event{std::move(agent), the_time, the_ordering, next_id_++, true});
    return true;
  }
  /* schedule.rs:283-293 */
  void schedule_once(std::shared_ptr<Agent> agent, float the_time, int32_t the_ordering) {
    push(Event{std::move(agent), the_time, the_ordering, next_id_++, false});
  }
  size_t num_events() const { return events_.size(); }

  /* schedule.rs:347-413 */
  void step(State& state) {
    if (step_count == 0) state.update(step_count);
    state.before_step(*this);
    if (events_.empty()) {
      state.after_step(*this);
      step_count += 1;
      state.update(step_count);
      return;
    }
    time = events_.begin()->time;
    std::vector<Event> due;
    while (!events_.empty() && !(events_.begin()->time > time))
      due.push_back(std::move(events_.extract(events_.begin()).value()));
    for (Event& e : due) {
      e.agent->before_step(state);
      e.agent->step(state);
      e.agent->after_step(state);
      if (e.repeating && !e.agent->is_stopped(state)) {
        e.time += 1.0f;
        push(std::move(e));
      }
    }
    state.after_step(*this);
    step_count += 1;
    state.update(step_count);
  }

 private:
  struct Event {
    std::shared_ptr<Agent> agent;
    float time;
    int32_t ordering;
    uint64_t id;
    bool repeating;
  };
  struct Before {
    bool operator()(const Event& a, const Event& b) const {
      if (a.time != b.time) return a.time < b.time;
      if (a.ordering != b.ordering) return a.ordering < b.ordering;
      return a.id < b.id;
    }
  };
  void push(Event e) { events_.insert(std::move(e)); }

  std::set<Event, Before> events_;  // ids are unique, so no two events compare equal
  uint64_t next_id_ = 0;
};

/* simulate! without the terminal UI (src/lib.rs:1158-1173): `reps` times,
 * a fresh Schedule, state.init, then up to `steps` Schedule::step calls
 * until end_condition. */
inline void simulate(State& state, uint64_t steps, uint32_t reps = 1) {
  for (uint32_t r = 0; r < reps; ++r) {
    Schedule schedule;
    state.init(schedule);
    for (uint64_t i = 0; i < steps; ++i) {
      schedule.step(state);
      if (state.end_condition(schedule)) break;
    }
  }
}

/* The flockers State (tests/model/flockers/state.rs) with its birds on the
 * device: Flocker::new(dim, initial_flockers) plus the Bird::step weights,
 * jump, radius and Philox key in `params` and the cell size. */
class Flocker : public State {
 public:
  Flocker(Real2D dim, uint32_t initial_flockers, float discretization,
          const kg_flock_params& params, int heading_mode = 1)
      : field1(dim.x, dim.y, discretization, true, std::max<uint32_t>(initial_flockers, 1)),
        initial_flockers(initial_flockers),
        dim(dim),
        params(params),
        heading_mode(heading_mode) {}

  /* state.rs:40-54: the birds, and one swarm agent standing for them */
  void init(Schedule& schedule) override {
    step = 0;
    field1.init_uniform(initial_flockers, params.seed, heading_mode);
    if (initial_flockers > 0) schedule.schedule_repeating(std::make_shared<Swarm>(), 0.0f, 0);
  }
  /* state.rs:57-59 (the step number is kept for the swarm's random draws) */
  void update(uint64_t s) override {
    field1.lazy_update();
    step = s;
  }
  /* simulate(*this, steps) for this state alone -- a fresh Schedule, init,
   * `steps` Schedule::step calls -- taken by kg_schedule_run in one launch
   * per step where the field allows.  The birds end as simulate leaves them
   * (end_condition is not consulted). */
  void run(uint64_t steps) {
    Schedule schedule;
    init(schedule);
    uint64_t done = 0;
    while (done < steps) {
      const uint32_t chunk = (uint32_t)std::min<uint64_t>(steps - done, 1u << 30);
      field1.schedule_run(params, done, chunk);
      done += chunk;
    }
    step = steps;
  }

  Field2D field1;
  uint32_t initial_flockers;
  Real2D dim;
  kg_flock_params params;
  int heading_mode;
  uint64_t step = 0;

 private:
  struct Swarm : Agent {
    void step(State& state) override {
      Flocker* fl = dynamic_cast<Flocker*>(&state);
      if (!fl) throw Error(KG_E_STATE, "krabgpu: the flockers swarm stepped with another State");
      fl->field1.flock_step(fl->params, fl->step);
    }
  };
};

/* Many independent flockers runs advanced together on the device
 * (explore_parallel!'s runs, src/explore/model_exploration.rs:354-432): one
 * launch per schedule step moves every admitted run, and each run ends with
 * the birds it would have alone.  Runs join with admit and leave with
 * retire between run calls. */
class Ensemble {
 public:
  explicit Ensemble(const kg_tuning& tuning = kg_tuning{0, 0, 0, 0}) {
    check(kg_ensemble_create(&tuning, &e_));
  }
  ~Ensemble() {
    if (e_) kg_ensemble_destroy(e_);
  }
  Ensemble(const Ensemble&) = delete;
  Ensemble& operator=(const Ensemble&) = delete;

  /* A run: Flocker::new(dim, agents) with its cell size, parameters and
   * initial-placement key. */
  static kg_run_desc run_desc(Real2D dim, uint32_t agents, float discretization,
                              const kg_flock_params& params, int heading_mode = 1) {
    return kg_run_desc{dim.x, dim.y, discretization, 1, agents, heading_mode, params.seed, params};
  }
  static uint64_t footprint(const kg_run_desc& run, const kg_tuning& tuning = kg_tuning{0, 0, 0, 0}) {
    uint64_t b = 0;
    check(kg_run_footprint(&run, &tuning, &b));
    return b;
  }
  /* Returns the runs' ids; their schedule step 0 comes with the next run(). */
  std::vector<uint32_t> admit(const std::vector<kg_run_desc>& runs) {
    std::vector<uint32_t> ids(runs.size());
    check(kg_ensemble_admit(e_, runs.data(), (uint32_t)runs.size(), ids.data()));
    for (size_t k = 0; k < runs.size(); ++k) agents_.push_back({ids[k], runs[k].agents});
    return ids;
  }
  void retire(uint32_t id) {
    check(kg_ensemble_retire(e_, id));
    agents_.erase(std::remove_if(agents_.begin(), agents_.end(),
                                 [id](const std::pair<uint32_t, uint32_t>& a) { return a.first == id; }),
                  agents_.end());
  }
  /* Every run takes `nsteps` Schedule::step calls (asynchronous). */
  void run(uint32_t nsteps, int flags = KG_RUN_GRAPH) { check(kg_ensemble_run(e_, nsteps, flags)); }
  uint64_t steps(uint32_t id) {
    uint64_t s = 0;
    check(kg_ensemble_steps(e_, id, &s));
    return s;
  }
  Agents read(uint32_t id) {
    uint32_t m = 0;
    for (const auto& a : agents_)
      if (a.first == id) m = a.second;
    Agents a;
    a.id.resize(m);
    a.x.resize(m);
    a.y.resize(m);
    a.hx.resize(m);
    a.hy.resize(m);
    size_t n = 0;
    check(kg_ensemble_read(e_, id, a.id.data(), a.x.data(), a.y.data(), a.hx.data(), a.hy.data(),
                           m, &n));
    return a;
  }
  void sync() { check(kg_ensemble_sync(e_)); }

 private:
  kg_ensemble* e_ = nullptr;
  std::vector<std::pair<uint32_t, uint32_t>> agents_;  // (run id, birds)
};

}  // namespace krabgpu

#endif
