/* krabgpu_serve.hpp -- continuous-batching serving of flockers runs in C++
 * over krabgpu::Ensemble.
 *
 * A request is a run (kg_run_desc) with a number of schedule steps and an
 * arrival time.  Requests are admitted in arrival order while their device
 * footprint fits the budget; admitted runs share one ensemble, so each tick
 * advances all of them with one launch per schedule step.  A tick lasts until
 * the earliest active run completes (at most max_tick steps), so every run
 * stops at its own step count; completed runs are read and retired, which
 * frees their memory for the queue.  Requests arriving while others run join
 * at the next tick.  Each result is the bits the run has alone. */
#ifndef KRABGPU_SERVE_HPP
#define KRABGPU_SERVE_HPP

#include <algorithm>
#include <chrono>
#include <functional>
#include <map>
#include <thread>
#include <vector>

#include "krabgpu.hpp"

namespace krabgpu {

struct Request {
  double arrival;  // seconds from the start of the stream
  kg_run_desc run;
  uint32_t steps;
};

struct Completed {
  size_t index;  // position in the request list
  double arrival, admitted, done;
  uint32_t agents, steps;
  Agents state;  // when the server keeps states
  double latency() const { return done - arrival; }
};

/* Ens: admit(std::vector<kg_run_desc>) -> ids, run(n), sync(), read(id),
 * retire(id) -- krabgpu::Ensemble, or a stand-in for tests. */
template <class Ens = Ensemble>
class Server {
 public:
  explicit Server(uint64_t hbm_budget, uint32_t max_tick = 64, bool keep_state = false,
                  uint32_t admit_every = 1)
      : budget(hbm_budget),
        max_tick(std::max<uint32_t>(max_tick, 1)),
        keep_state(keep_state),
        admit_every(std::max<uint32_t>(admit_every, 1)) {}

  uint64_t budget;
  uint32_t max_tick;
  bool keep_state;
  uint32_t admit_every;  // while runs are active, admit at most every this many ticks
  std::function<uint64_t(const kg_run_desc&)> footprint = [](const kg_run_desc& r) {
    return Ensemble::footprint(r);
  };
  std::function<double()> clock = [] {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
  };
  std::function<void(double)> sleep = [](double s) {
    std::this_thread::sleep_for(std::chrono::duration<double>(s));
  };

  uint64_t ticks = 0, admissions = 0, peak_bytes = 0;
  double wall = 0.0;

  /* Runs the stream to completion; the Completed records in completion
   * order.  A request larger than the budget throws before anything runs. */
  std::vector<Completed> serve(const std::vector<Request>& requests, Ens& ens) {
    std::vector<size_t> order(requests.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return requests[a].arrival < requests[b].arrival; });
    std::vector<uint64_t> fp(requests.size());
    for (size_t k = 0; k < requests.size(); ++k) {
      fp[k] = footprint(requests[k].run);
      if (fp[k] > budget) throw Error(KG_E_ADMIT, "krabgpu: a request needs more than the budget");
    }
    struct Active {
      size_t index;
      uint32_t left;
      double admitted;
      uint64_t bytes;
    };
    std::map<uint32_t, Active> active;  // run id -> state
    std::vector<Completed> done;
    uint64_t used = 0;
    size_t head = 0;
    const double t0 = clock();
    int64_t last_admit = -(int64_t)admit_every;
    while (head < order.size() || !active.empty()) {
      double now = clock() - t0;
      std::vector<size_t> batch;
      const bool may_admit = active.empty() || (int64_t)ticks - last_admit >= (int64_t)admit_every;
      while (may_admit && head < order.size() && requests[order[head]].arrival <= now &&
             used + fp[order[head]] <= budget) {
        batch.push_back(order[head]);
        used += fp[order[head]];
        ++head;
      }
      peak_bytes = std::max(peak_bytes, used);
      if (!batch.empty()) {
        std::vector<kg_run_desc> runs;
        for (size_t i : batch) runs.push_back(requests[i].run);
        const std::vector<uint32_t> ids = ens.admit(runs);
        ++admissions;
        last_admit = (int64_t)ticks;
        for (size_t k = 0; k < batch.size(); ++k)
          active[ids[k]] = Active{batch[k], requests[batch[k]].steps, now, fp[batch[k]]};
      }
      if (active.empty()) {
        // idle until the next arrival
        const double wait = requests[order[head]].arrival - (clock() - t0);
        if (wait > 0) sleep(std::min(wait, 0.01));
        continue;
      }
      uint32_t tick = max_tick;
      for (const auto& a : active) tick = std::min(tick, a.second.left);
      if (tick > 0) ens.run(tick);
      ens.sync();
      ++ticks;
      now = clock() - t0;
      for (auto it = active.begin(); it != active.end();) {
        Active& a = it->second;
        if (a.left == tick) {
          const Request& r = requests[a.index];
          Completed c{a.index, r.arrival, a.admitted, now, r.run.agents, r.steps, Agents{}};
          if (keep_state) c.state = ens.read(it->first);
          ens.retire(it->first);
          used -= a.bytes;
          done.push_back(std::move(c));
          it = active.erase(it);
        } else {
          a.left -= tick;
          ++it;
        }
      }
    }
    wall = clock() - t0;
    return done;
  }
};

}  // namespace krabgpu

#endif
