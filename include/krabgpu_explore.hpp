/* krabgpu_explore.hpp -- model exploration of flockers in C++ over
 * krabgpu::Ensemble: gen_param! (src/lib.rs:1794-1816), ExploreMode
 * (src/lib.rs:481-486), build_configurations! (src/lib.rs:1724-1748),
 * explore_sequential! / explore_parallel! (src/explore/model_exploration.rs:
 * 231-432), the FrameRow of build_dataframe! (:451-540), write_csv
 * (src/lib.rs:1762-1773) and load_csv! (src/lib.rs:1878-1900).
 *
 * Differences from the reference, each forced by the device engine: run r
 * of a sweep uses the Philox keys (seed << 32) | r instead of an unseeded
 * thread RNG; the runs of one parallel batch share one timed region; the
 * state's end_condition is not consulted (flockers' is false). */
#ifndef KRABGPU_EXPLORE_HPP
#define KRABGPU_EXPLORE_HPP

#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "krabgpu.hpp"

namespace krabgpu {

enum class ExploreMode { Exaustive, Matched };

/* n values uniform in [min, max): bounds swapped when min > max, max raised
 * by one when equal; n = 0 gives none (the reference takes the original n). */
template <class T>
std::vector<T> gen_param(T lo, T hi, size_t n, std::mt19937_64& rng) {
  static_assert(std::is_arithmetic<T>::value, "gen_param needs a number type");
  if (lo > hi) std::swap(lo, hi);
  else if (lo == hi) hi = hi + T(1);
  std::vector<T> out;
  out.reserve(n);
  for (size_t k = 0; k < n; ++k) {
    if constexpr (std::is_integral<T>::value) {
      out.push_back(std::uniform_int_distribution<T>(lo, T(hi - 1))(rng));
    } else {
      T v = std::uniform_real_distribution<T>(lo, hi)(rng);
      if (!(v < hi)) v = std::nextafter(hi, lo);  // keep the interval half-open
      out.push_back(v);
    }
  }
  return out;
}

/* Row k: for every configuration, the index into input k (the first input
 * varies slowest).  An empty input throws (the reference divides by its
 * length). */
inline std::vector<std::vector<size_t>> build_configurations(const std::vector<size_t>& sizes) {
  size_t n_conf = 1;
  for (size_t s : sizes) {
    if (s == 0) throw Error(KG_E_INVALID, "krabgpu: exhaustive sweep over an empty input");
    n_conf *= s;
  }
  std::vector<std::vector<size_t>> rows;
  size_t rep = n_conf;
  for (size_t s : sizes) {
    rep /= s;
    std::vector<size_t> row(n_conf);
    for (size_t i = 0; i < n_conf; ++i) row[i] = (i / rep) % s;
    rows.push_back(std::move(row));
  }
  return rows;
}

/* A sweep's inputs, in order: (name, values).  Flockers takes agents,
 * width, height, discretization, cohesion, avoidance, consistency,
 * randomness, momentum, jump, radius, heading_mode. */
using Inputs = std::vector<std::pair<std::string, std::vector<double>>>;
using Conf = std::vector<std::pair<std::string, double>>;

inline std::vector<Conf> configurations(const Inputs& in, ExploreMode mode) {
  std::vector<Conf> out;
  if (mode == ExploreMode::Exaustive) {
    std::vector<size_t> sizes;
    for (const auto& p : in) sizes.push_back(p.second.size());
    const auto rows = build_configurations(sizes);
    const size_t n_conf = rows.empty() ? 1 : rows[0].size();
    for (size_t i = 0; i < n_conf; ++i) {
      Conf c;
      for (size_t j = 0; j < in.size(); ++j) c.push_back({in[j].first, in[j].second[rows[j][i]]});
      out.push_back(std::move(c));
    }
    return out;
  }
  // Matched: n_conf is the last input's length; a shorter input is out of range
  const size_t n_conf = in.empty() ? 1 : in.back().second.size();
  for (const auto& p : in)
    if (p.second.size() < n_conf)
      throw Error(KG_E_INVALID, "krabgpu: matched sweep: input " + p.first + " is too short");
  for (size_t i = 0; i < n_conf; ++i) {
    Conf c;
    for (const auto& p : in) c.push_back({p.first, p.second[i]});
    out.push_back(std::move(c));
  }
  return out;
}

/* The run of one configuration: Flocker::new((width, height), agents) and
 * the Bird::step constants (tests/model/flockers/bird.rs:12-17, 41), with
 * defaults 1,000 birds on 200 x 200 and cells of 10 / 1.5. */
inline kg_run_desc flockers_run(const Conf& c, uint64_t seed) {
  kg_run_desc r{200.0f, 200.0f, 10.0f / 1.5f, 1, 1000, 1, seed,
                kg_flock_params{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0.7f, 10.0f, 0, seed}};
  for (const auto& kv : c) {
    const std::string& k = kv.first;
    const float v = (float)kv.second;
    if (k == "agents") r.agents = (uint32_t)kv.second;
    else if (k == "width") r.width = v;
    else if (k == "height") r.height = v;
    else if (k == "discretization") r.discretization = v;
    else if (k == "cohesion") r.params.cohesion = v;
    else if (k == "avoidance") r.params.avoidance = v;
    else if (k == "consistency") r.params.consistency = v;
    else if (k == "randomness") r.params.randomness = v;
    else if (k == "momentum") r.params.momentum = v;
    else if (k == "jump") r.params.jump = v;
    else if (k == "radius") r.params.radius = v;
    else if (k == "heading_mode") r.heading_mode = (int32_t)kv.second;
    else throw Error(KG_E_INVALID, "krabgpu: flockers takes no input named " + k);
  }
  return r;
}

inline bool integer_input(const std::string& k) { return k == "agents" || k == "heading_mode"; }

/* ---- Rust's text for numbers (FrameRow::to_string uses {:?} for inputs and
 * outputs and to_string() for the two timings) ---- */
namespace detail {
template <class T>
std::string rust_float(T v, bool debug) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
  std::string sci(buf, r.ptr);  // shortest digits, e.g. "-1.25e+07"
  const bool neg = sci[0] == '-';
  const size_t e = sci.find('e');
  std::string mant = sci.substr(neg ? 1 : 0, e - (neg ? 1 : 0));
  const int exp10 = std::stoi(sci.substr(e + 1));
  std::string digits;
  for (char ch : mant)
    if (ch != '.') digits += ch;
  const double a = std::fabs((double)v);
  std::string out = neg ? "-" : "";
  if (debug && a != 0.0 && (a < 1e-4 || a >= 1e16)) {
    out += digits.substr(0, 1);
    if (digits.size() > 1) out += "." + digits.substr(1);
    return out + "e" + std::to_string(exp10);
  }
  if (a == 0.0) return out + (debug ? "0.0" : "0");
  std::string ip, fp;
  if (exp10 >= 0) {
    while ((int)digits.size() < exp10 + 1) digits += '0';
    ip = digits.substr(0, (size_t)exp10 + 1);
    fp = digits.substr((size_t)exp10 + 1);
  } else {
    ip = "0";
    fp = std::string((size_t)(-exp10 - 1), '0') + digits;
  }
  if (fp.empty()) return out + ip + (debug ? ".0" : "");
  return out + ip + "." + fp;
}
}  // namespace detail

inline std::string rust_debug(double v) { return detail::rust_float(v, true); }
inline std::string rust_debug_f32(float v) { return detail::rust_float(v, true); }
inline std::string rust_display_f32(float v) { return detail::rust_float(v, false); }

/* One run of a sweep (build_dataframe!'s FrameRow). */
struct FrameRow {
  uint32_t conf_num = 0, conf_rep = 0;
  Conf inputs;
  std::vector<std::pair<std::string, double>> outputs;
  float run_duration = 0.0f, step_per_sec = 0.0f;
  Agents state;  // final birds when requested

  std::vector<std::string> field_names() const {
    std::vector<std::string> v{"Simulation", "Run"};
    for (const auto& kv : inputs) v.push_back(kv.first);
    for (const auto& kv : outputs) v.push_back(kv.first);
    v.push_back("Run Duration");
    v.push_back("Step per sec");
    return v;
  }
  std::vector<std::string> to_strings() const {
    std::vector<std::string> v{std::to_string(conf_num), std::to_string(conf_rep)};
    for (const auto& kv : inputs)
      v.push_back(integer_input(kv.first) ? std::to_string((long long)kv.second)
                                          : rust_debug(kv.second));
    for (const auto& kv : outputs) v.push_back(rust_debug(kv.second));
    v.push_back(rust_display_f32(run_duration));
    v.push_back(rust_display_f32(step_per_sec));
    return v;
  }
};

/* write_csv(name, dataframe): "<name>.csv", a header and one record per run;
 * fields holding a comma, quote or line break are quoted. */
inline void write_csv(const std::string& name, const std::vector<FrameRow>& rows) {
  if (rows.empty()) throw Error(KG_E_INVALID, "krabgpu: write_csv of an empty dataframe");
  FILE* f = std::fopen((name + ".csv").c_str(), "w");
  if (!f) throw Error(KG_E_INVALID, "krabgpu: cannot open " + name + ".csv");
  auto put = [f](const std::vector<std::string>& rec) {
    for (size_t k = 0; k < rec.size(); ++k) {
      const std::string& s = rec[k];
      if (k) std::fputc(',', f);
      if (s.find_first_of(",\"\r\n") != std::string::npos) {
        std::string q = "\"";
        for (char ch : s) q += ch == '"' ? std::string("\"\"") : std::string(1, ch);
        std::fputs((q + "\"").c_str(), f);
      } else {
        std::fputs(s.c_str(), f);
      }
    }
    std::fputc('\n', f);
  };
  put(rows[0].field_names());
  for (const auto& r : rows) put(r.to_strings());
  std::fclose(f);
}

/* The birds in ascending id, so that outputs do not depend on the order a
 * buffer was read in. */
inline Agents sorted_by_id(const Agents& a) {
  std::vector<size_t> o(a.id.size());
  for (size_t k = 0; k < o.size(); ++k) o[k] = k;
  std::sort(o.begin(), o.end(), [&](size_t i, size_t j) { return a.id[i] < a.id[j]; });
  Agents s;
  for (size_t k : o) {
    s.id.push_back(a.id[k]);
    s.x.push_back(a.x[k]);
    s.y.push_back(a.y[k]);
    s.hx.push_back(a.hx[k]);
    s.hy.push_back(a.hy[k]);
  }
  return s;
}

/* load_csv!(file, x: T, ...) (src/lib.rs:1878-1900): the header row is
 * skipped and the first `ncols` fields of every record are parsed as numbers
 * (Rust's "NaN", "inf" and "-inf" included); one vector per column. */
inline std::vector<std::vector<double>> load_csv(const std::string& path, size_t ncols) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) throw Error(KG_E_INVALID, "krabgpu: cannot open " + path);
  std::vector<std::vector<double>> cols(ncols);
  std::string text;
  for (int ch; (ch = std::fgetc(f)) != EOF;) text += (char)ch;
  std::fclose(f);
  std::vector<std::string> rec;
  std::string field;
  bool quoted = false, header = true;
  auto end_record = [&] {
    rec.push_back(field);
    field.clear();
    if (!header) {
      if (rec.size() < ncols) throw Error(KG_E_INVALID, "krabgpu: load_csv: short record");
      for (size_t k = 0; k < ncols; ++k) {
        const std::string& t = rec[k];
        cols[k].push_back(t == "NaN" ? NAN : t == "inf" ? INFINITY : t == "-inf" ? -INFINITY
                                                                                 : std::stod(t));
      }
    }
    header = false;
    rec.clear();
  };
  for (size_t k = 0; k < text.size(); ++k) {
    const char ch = text[k];
    if (quoted) {
      if (ch == '"' && k + 1 < text.size() && text[k + 1] == '"') {
        field += '"';
        ++k;
      } else if (ch == '"') {
        quoted = false;
      } else {
        field += ch;
      }
    } else if (ch == '"') {
      quoted = true;
    } else if (ch == ',') {
      rec.push_back(field);
      field.clear();
    } else if (ch == '\n' || ch == '\r') {
      if (ch == '\r' && k + 1 < text.size() && text[k + 1] == '\n') ++k;
      end_record();
    } else {
      field += ch;
    }
  }
  if (!field.empty() || !rec.empty()) end_record();
  return cols;
}

/* Output of a run from its final birds (in ascending id). */
using Output = std::function<double(const Agents&, const kg_run_desc&)>;

/* Length of the mean unit heading (1 for an aligned flock). */
inline double polarisation(const Agents& a, const kg_run_desc&) {
  double sx = 0.0, sy = 0.0;
  size_t m = 0;
  for (size_t k = 0; k < a.hx.size(); ++k) {
    const double n = std::hypot((double)a.hx[k], (double)a.hy[k]);
    if (n > 0.0) {
      sx += a.hx[k] / n;
      sy += a.hy[k] / n;
      ++m;
    }
  }
  return m ? std::hypot(sx / (double)m, sy / (double)m) : 0.0;
}

struct ExploreOptions {
  bool parallel = true;               // explore_parallel! (false: explore_sequential!)
  uint64_t seed = 0;                  // run r uses Philox keys (seed << 32) | r
  uint64_t hbm_budget = 200ull << 30; // device bytes per parallel batch
  bool keep_state = false;
};

/* Every configuration `rep_conf` times for `nstep` schedule steps; one
 * FrameRow per run, configurations in order and repetitions within each
 * (numbered from 0 in parallel sweeps, from 1 in sequential ones, as the
 * reference numbers them). */
inline std::vector<FrameRow> explore(uint32_t nstep, uint32_t rep_conf, const Inputs& inputs,
                                     const std::vector<std::pair<std::string, Output>>& outputs,
                                     ExploreMode mode, const ExploreOptions& opt = {}) {
  const std::vector<Conf> confs = configurations(inputs, mode);
  std::vector<std::pair<uint32_t, uint32_t>> runs;
  std::vector<kg_run_desc> descs;
  for (uint32_t i = 0; i < confs.size(); ++i)
    for (uint32_t j = 0; j < rep_conf; ++j) {
      runs.push_back({i, j});
      descs.push_back(flockers_run(confs[i], (opt.seed << 32) | (uint64_t)(i * rep_conf + j)));
    }
  std::vector<FrameRow> rows(runs.size());
  auto finish = [&](size_t k, const Agents& read, double dt) {
    Agents st = sorted_by_id(read);
    FrameRow& r = rows[k];
    r.conf_num = runs[k].first;
    r.conf_rep = opt.parallel ? runs[k].second : runs[k].second + 1;
    r.inputs = confs[runs[k].first];
    for (const auto& o : outputs) r.outputs.push_back({o.first, o.second(st, descs[k])});
    r.run_duration = (float)dt;
    r.step_per_sec = (float)nstep / r.run_duration;
    if (opt.keep_state) r.state = std::move(st);
  };
  using clock = std::chrono::steady_clock;
  if (!opt.parallel) {
    for (size_t k = 0; k < descs.size(); ++k) {
      const kg_run_desc& d = descs[k];
      Flocker fl({d.width, d.height}, d.agents, d.discretization, d.params, d.heading_mode);
      Schedule schedule;
      fl.init(schedule);
      fl.field1.sync();
      const auto t0 = clock::now();
      if (nstep) fl.field1.schedule_run(d.params, 0, nstep);
      fl.field1.sync();
      const double dt = std::chrono::duration<double>(clock::now() - t0).count();
      finish(k, fl.field1.read_agents(nstep ? 0 : 1), dt);
    }
    return rows;
  }
  size_t k0 = 0;
  while (k0 < descs.size()) {
    // the next batch: consecutive runs whose footprints fit the budget
    size_t k1 = k0;
    uint64_t used = 0;
    while (k1 < descs.size()) {
      const uint64_t b = Ensemble::footprint(descs[k1]);
      if (b > opt.hbm_budget)
        throw Error(KG_E_ADMIT, "krabgpu: a run needs more than the sweep's device budget");
      if (k1 > k0 && used + b > opt.hbm_budget) break;
      used += b;
      ++k1;
    }
    Ensemble ens;
    const std::vector<kg_run_desc> batch(descs.begin() + (ptrdiff_t)k0,
                                         descs.begin() + (ptrdiff_t)k1);
    const std::vector<uint32_t> ids = ens.admit(batch);
    ens.sync();
    const auto t0 = clock::now();
    ens.run(nstep);
    ens.sync();
    const double dt = std::chrono::duration<double>(clock::now() - t0).count();
    for (size_t k = k0; k < k1; ++k) finish(k, ens.read(ids[k - k0]), dt);
    k0 = k1;
  }
  return rows;
}

}  // namespace krabgpu

#endif
