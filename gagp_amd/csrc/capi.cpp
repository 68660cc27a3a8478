// C ABI (include/gagp_hip.h) over the engine classes.
#include "../../include/gagp_hip.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>

#include "engine.hpp"
#include "serve.hpp"

using namespace gagp_amd;

struct gagp_session { std::unique_ptr<Session> s; };
struct gagp_batch { Batch b; std::vector<SearchResult> results; bool prepared = false, ran = false; };
struct gagp_server { Server s; bool ran = false; };
struct gagp_owner {
  explicit gagp_owner(const SearchSpec& spec) : o(spec) {}
  Owner o;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return GAGP_OK;
  } catch (const Unsupported& e) {
    g_last_error = e.what();
    return GAGP_ERR_UNSUPPORTED;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return GAGP_ERR_INVALID;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GAGP_ERR_DEVICE;
  }
}
int copy_out(const std::string& text, char* buf, size_t cap, size_t* needed) {
  if (needed) *needed = text.size() + 1;
  if (!buf || cap < text.size() + 1) {
    g_last_error = "buffer too small";
    return GAGP_ERR_INVALID;
  }
  std::memcpy(buf, text.c_str(), text.size() + 1);
  return GAGP_OK;
}
ProgramIn program_in(const gagp_program& p) {
  if (p.code_len <= 0 || !p.code) throw std::invalid_argument("empty program");
  ProgramIn out;
  for (int i = 0; i < p.code_len; ++i) {
    out.op.push_back(p.code[i].op);
    out.has_a.push_back(p.code[i].has_a);
    out.has_b.push_back(p.code[i].has_b);
    out.a.push_back(p.code[i].a);
    out.b.push_back(p.code[i].b);
    if (p.fuel) out.fuel.push_back(p.fuel[i]);
  }
  for (int i = 0; i < p.n_consts; ++i) {
    Val v{};
    std::memcpy(&v, &p.consts[i], sizeof v);
    out.consts.push_back(v);
  }
  out.n_locals = p.n_locals;
  return out;
}
}  // namespace

static_assert(sizeof(gagp_value) == sizeof(Val), "gagp_value matches the engine value");

extern "C" {

const char* gagp_last_error(void) { return g_last_error.c_str(); }

int gagp_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int gagp_session_create(const gagp_value* case_inputs, const uint8_t* input_set, int32_t n_cases, int32_t n_inputs,
                        const gagp_value* answers, const int64_t* list_arena, int64_t list_arena_len, int32_t fuel,
                        int32_t lanes, double penalty, gagp_policy policy, gagp_session** out) {
  return guarded([&] {
    if (!out || n_cases <= 0 || n_inputs < 0 || !answers || (n_inputs && !case_inputs) || fuel < 1 || lanes < 1)
      throw std::invalid_argument("invalid session arguments");
    if (policy.metric < 0 || policy.metric > 2 || policy.proximity < 0 || policy.proximity > 1)
      throw Unsupported("fitness metric outside the engine's scalar metrics");
    Cases c;
    c.n = n_cases;
    for (int k = 0; k < n_inputs; ++k) c.input_names.push_back("x" + std::to_string(k));
    c.input_types.assign(std::size_t(n_inputs), TAG_INVALID);
    c.inputs.resize(std::size_t(n_cases) * n_inputs);
    c.set.assign(std::size_t(n_cases) * n_inputs, 1);
    for (std::size_t i = 0; i < c.inputs.size(); ++i) {
      std::memcpy(&c.inputs[i], &case_inputs[i], sizeof(Val));
      if (input_set) c.set[i] = input_set[i];
    }
    c.expected.resize(std::size_t(n_cases));
    for (int i = 0; i < n_cases; ++i) std::memcpy(&c.expected[i], &answers[i], sizeof(Val));
    if (list_arena && list_arena_len > 0) c.arena.assign(list_arena, list_arena + list_arena_len);
    const Policy p{policy.metric, policy.proximity, policy.summary_count, policy.summary_weight, policy.summary_cap};
    auto* s = new gagp_session{std::make_unique<Session>(c, fuel, lanes, penalty, p)};
    *out = s;
  });
}

int gagp_session_eval(gagp_session* s, const gagp_program* programs, int32_t n, double* out_fitness, uint32_t* out_wrong,
                      uint32_t* out_counts) {
  return guarded([&] {
    if (!s || n < 0 || (n && (!programs || !out_fitness))) throw std::invalid_argument("invalid eval arguments");
    std::vector<ProgramIn> in;
    in.reserve(std::size_t(n));
    for (int i = 0; i < n; ++i) in.push_back(program_in(programs[i]));
    s->s->eval(in, out_fitness, out_wrong, out_counts);
  });
}

int gagp_session_eval_static(gagp_session* s, const gagp_program* programs, int32_t n, double* out_fitness, uint32_t* out_wrong,
                             uint32_t* out_counts) {
  return guarded([&] {
    if (!s || n < 0 || (n && (!programs || !out_fitness))) throw std::invalid_argument("invalid eval arguments");
    std::vector<ProgramIn> in;
    in.reserve(std::size_t(n));
    for (int i = 0; i < n; ++i) in.push_back(program_in(programs[i]));
    s->s->eval_static(in, out_fitness, out_wrong, out_counts);
  });
}

int gagp_session_eval_cases(gagp_session* s, const gagp_program* program, gagp_value* out_values, int32_t* out_errors) {
  return guarded([&] {
    if (!s || !program || !out_values || !out_errors) throw std::invalid_argument("invalid eval arguments");
    std::vector<Val> values;
    std::vector<int> errors;
    s->s->eval_cases(program_in(*program), values, errors);
    for (std::size_t i = 0; i < values.size(); ++i) {
      std::memcpy(&out_values[i], &values[i], sizeof(Val));
      out_errors[i] = errors[i];
    }
  });
}

float gagp_session_last_kernel_ms(gagp_session* s) { return s ? s->s->last_kernel_ms() : 0.f; }
void gagp_session_destroy(gagp_session* s) { delete s; }

int gagp_batch_create(gagp_batch** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null output");
    *out = new gagp_batch();
  });
}

int gagp_batch_add_init(gagp_batch* b, const char* task_path, const char* search_path, int32_t population,
                        int32_t generations, int32_t seed_is_set, uint64_t seed, const char* init_path,
                        int32_t* out_index) {
  return guarded([&] {
    if (!b || !task_path) throw std::invalid_argument("invalid batch arguments");
    if (b->prepared) throw std::invalid_argument("batch already prepared");
    SearchSpec spec;
    spec.task_path = task_path;
    if (search_path) spec.search_path = search_path;
    if (population > 0) spec.population = population;
    if (generations > 0) spec.generations = generations;
    if (seed_is_set) spec.seed = seed;
    if (init_path) spec.init_path = init_path;
    const int idx = b->b.add(spec);
    if (out_index) *out_index = idx;
  });
}

int gagp_batch_add(gagp_batch* b, const char* task_path, const char* search_path, int32_t population, int32_t generations,
                   int32_t seed_is_set, uint64_t seed, int32_t* out_index) {
  return gagp_batch_add_init(b, task_path, search_path, population, generations, seed_is_set, seed, nullptr, out_index);
}

int gagp_batch_planned_bytes(gagp_batch* b, int32_t index, uint64_t* out) {
  return guarded([&] {
    if (!b || index < 0 || std::size_t(index) >= b->b.size() || !out) throw std::invalid_argument("no such search");
    *out = b->b.planned_footprint(index);
  });
}

int gagp_batch_run(gagp_batch* b, int32_t use_graph) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    if (b->prepared || b->ran) throw std::invalid_argument("batch already ran");
    b->b.run(use_graph != 0);
    b->prepared = b->ran = true;
    b->results.clear();
    for (std::size_t i = 0; i < b->b.size(); ++i) b->results.push_back(b->b.result(int(i), false));
  });
}

int gagp_batch_prepare(gagp_batch* b) {
  return guarded([&] {
    if (!b || b->prepared) throw std::invalid_argument("batch already prepared");
    b->b.prepare();
    b->prepared = true;
  });
}

int gagp_batch_step(gagp_batch* b, int32_t rounds, int32_t use_graph) {
  return guarded([&] {
    if (!b || !b->prepared || b->ran) throw std::invalid_argument("batch not prepared");
    b->b.step(rounds, use_graph != 0);
  });
}

int gagp_batch_max_rounds(gagp_batch* b, int32_t* out) {
  return guarded([&] {
    if (!b || !b->prepared || !out) throw std::invalid_argument("batch not prepared");
    *out = b->b.max_rounds();
  });
}

int gagp_batch_finish(gagp_batch* b) {
  return guarded([&] {
    if (!b || !b->prepared || b->ran) throw std::invalid_argument("batch not prepared");
    b->b.finish();
    b->ran = true;
    b->results.clear();
    for (std::size_t i = 0; i < b->b.size(); ++i) b->results.push_back(b->b.result(int(i), false));
  });
}

int gagp_batch_counters(gagp_batch* b, int32_t index, uint64_t* out) {
  return guarded([&] {
    if (!b || !b->prepared || index < 0 || std::size_t(index) >= b->b.size() || !out) throw std::invalid_argument("no such search");
    const auto c = b->b.counters(index);
    for (std::size_t k = 0; k < 14 && k < c.size(); ++k) out[k] = c[k];
  });
}

int gagp_batch_result_json(gagp_batch* b, int32_t index, char* buf, size_t cap, size_t* needed) {
  std::string text;
  const int st = guarded([&] {
    if (!b || !b->ran || index < 0 || std::size_t(index) >= b->results.size()) throw std::invalid_argument("no such result");
    text = json::canonical(b->results[std::size_t(index)].result);
  });
  if (st != GAGP_OK) return st;
  return copy_out(text, buf, cap, needed);
}

int gagp_batch_evaluated_program_cases(gagp_batch* b, int32_t index, uint64_t* out) {
  return guarded([&] {
    if (!b || !b->ran || index < 0 || std::size_t(index) >= b->results.size() || !out) throw std::invalid_argument("no such result");
    *out = b->results[std::size_t(index)].evaluated_program_cases;
  });
}

int gagp_batch_set_profile(gagp_batch* b, int32_t on) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    b->b.set_profile(on != 0);
  });
}

int gagp_valu_probe(int32_t blocks, int32_t iterations, float* out_ms) {
  return guarded([&] {
    if (!out_ms) throw std::invalid_argument("null output");
    *out_ms = valu_probe_ms(blocks, iterations);
  });
}

int gagp_batch_set_autotune(gagp_batch* b, int32_t mode) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    b->b.set_autotune(mode);
  });
}

int gagp_batch_tuning_json(gagp_batch* b, char* buf, size_t cap, size_t* needed) {
  std::string text;
  const int st = guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    text = b->b.tuning_report();
  });
  if (st != GAGP_OK) return st;
  return copy_out(text, buf, cap, needed);
}

int gagp_batch_kernel_times_json(gagp_batch* b, char* buf, size_t cap, size_t* needed) {
  std::string text;
  const int st = guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    json::Value o = json::make_object();
    for (const auto& kv : b->b.kernel_times()) o.obj[kv.first] = json::make_number(kv.second);
    text = json::canonical(o);
  });
  if (st != GAGP_OK) return st;
  return copy_out(text, buf, cap, needed);
}

void gagp_batch_destroy(gagp_batch* b) { delete b; }


int gagp_server_create(uint64_t budget_bytes, int32_t poll_rounds, gagp_server** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("invalid output pointer");
    *out = new gagp_server{Server(std::size_t(budget_bytes), poll_rounds)};
  });
}

int gagp_server_set_max_live(gagp_server* s, int32_t max_live) {
  return guarded([&] {
    if (!s || s->ran || max_live < 0) throw std::invalid_argument("invalid server or limit");
    s->s.set_max_live(max_live);
  });
}

int gagp_server_submit(gagp_server* s, const char* task_path, const char* search_path, int32_t population, int32_t generations,
                       int32_t seed_is_set, uint64_t seed, int32_t early_exit, double arrival_s, int32_t* out_ticket) {
  return guarded([&] {
    if (!s || !task_path || !out_ticket || s->ran) throw std::invalid_argument("invalid server request");
    SearchSpec spec;
    spec.task_path = task_path;
    if (search_path) spec.search_path = search_path;
    if (population > 0) spec.population = population;
    if (generations > 0) spec.generations = generations;
    if (seed_is_set) spec.seed = seed;
    spec.early_exit = early_exit != 0;
    *out_ticket = s->s.submit(spec, arrival_s);
  });
}

int gagp_server_run(gagp_server* s) {
  return guarded([&] {
    if (!s || s->ran) throw std::invalid_argument("server already ran");
    s->ran = true;
    s->s.run();
  });
}

int gagp_server_outcome(gagp_server* s, int32_t ticket, double* times, int32_t* generations, int32_t* status) {
  return guarded([&] {
    if (!s || !s->ran || ticket < 0 || std::size_t(ticket) >= s->s.size()) throw std::invalid_argument("no such request");
    const ServeOutcome& o = s->s.outcome(ticket);
    if (times) { times[0] = o.arrival_s; times[1] = o.admit_s; times[2] = o.done_s; }
    if (generations) *generations = o.generations;
    if (status) *status = o.status;
  });
}

int gagp_server_result_json(gagp_server* s, int32_t ticket, char* buf, size_t cap, size_t* needed) {
  std::string text;
  const int st = guarded([&] {
    if (!s || !s->ran || ticket < 0 || std::size_t(ticket) >= s->s.size()) throw std::invalid_argument("no such request");
    const ServeOutcome& o = s->s.outcome(ticket);
    if (o.status != GAGP_OK) throw std::invalid_argument("request failed: " + o.message);
    text = o.result_json;
  });
  if (st != GAGP_OK) return st;
  return copy_out(text, buf, cap, needed);
}

void gagp_server_destroy(gagp_server* s) { delete s; }

int gagp_owner_create(const char* task_path, const char* search_path, int32_t population, int32_t seed_is_set, uint64_t seed,
                      const char* init_path, gagp_owner** out) {
  return guarded([&] {
    if (!task_path || !out) throw std::invalid_argument("invalid owner arguments");
    SearchSpec spec;
    spec.task_path = task_path;
    if (search_path) spec.search_path = search_path;
    if (population > 0) spec.population = population;
    if (seed_is_set) spec.seed = seed;
    if (init_path) spec.init_path = init_path;
    *out = new gagp_owner(spec);
  });
}

int gagp_owner_population_size(gagp_owner* o, int32_t* out) {
  return guarded([&] {
    if (!o || !out) throw std::invalid_argument("null owner");
    *out = o->o.population_size();
  });
}

int gagp_owner_batch_size(gagp_owner* o, int32_t* out) {
  return guarded([&] {
    if (!o || !out) throw std::invalid_argument("null owner");
    *out = o->o.batch_size();
  });
}

int gagp_owner_has_wrong(gagp_owner* o, int32_t* out) {
  return guarded([&] {
    if (!o || !out) throw std::invalid_argument("null owner");
    *out = o->o.has_wrong() ? 1 : 0;
  });
}

int gagp_owner_evaluate(gagp_owner* o, double* fitness, uint32_t* wrong, uint64_t* counts) {
  return guarded([&] {
    if (!o || !fitness || !wrong || !counts) throw std::invalid_argument("null owner argument");
    o->o.evaluate(fitness, wrong, counts);
  });
}

int gagp_owner_node_counts(gagp_owner* o, int32_t grammar_sites, uint32_t* out) {
  return guarded([&] {
    if (!o || !out) throw std::invalid_argument("null owner argument");
    o->o.node_counts(grammar_sites != 0, out);
  });
}

int gagp_owner_begin_steady(gagp_owner* o, int32_t batch) {
  return guarded([&] {
    if (!o) throw std::invalid_argument("null owner");
    o->o.begin_steady(batch);
  });
}

int gagp_owner_make_batch(gagp_owner* o, const double* selection, int32_t pressure, double p_mut, uint64_t seed,
                          int32_t* reuse) {
  return guarded([&] {
    if (!o || !selection || !reuse) throw std::invalid_argument("null owner argument");
    o->o.make_batch(selection, pressure, p_mut, seed, reuse);
  });
}

int gagp_owner_reproduce(gagp_owner* o, const double* selection, int32_t pressure, double p_mut, uint64_t seed,
                         int32_t elites) {
  return guarded([&] {
    if (!o || !selection) throw std::invalid_argument("null owner argument");
    o->o.reproduce(selection, pressure, p_mut, seed, elites);
  });
}

int gagp_owner_evaluate_batch(gagp_owner* o, double* fitness, uint32_t* wrong, uint64_t* counts) {
  return guarded([&] {
    if (!o || !fitness || !wrong || !counts) throw std::invalid_argument("null owner argument");
    o->o.evaluate_batch(fitness, wrong, counts);
  });
}

int gagp_owner_batch_nodes(gagp_owner* o, int32_t grammar_sites, uint32_t* out) {
  return guarded([&] {
    if (!o || !out) throw std::invalid_argument("null owner argument");
    o->o.batch_nodes(grammar_sites != 0, out);
  });
}

int gagp_owner_commit_batch(gagp_owner* o, const int32_t* pairs, int32_t n) {
  return guarded([&] {
    if (!o || n < 0 || (n > 0 && !pairs)) throw std::invalid_argument("invalid placements");
    o->o.commit_batch(pairs, n);
  });
}

int gagp_owner_export_member(gagp_owner* o, int32_t member, char* buf, size_t cap, size_t* needed) {
  std::string text;
  const int st = guarded([&] {
    if (!o) throw std::invalid_argument("null owner");
    text = json::canonical(o->o.export_member(member));
  });
  if (st != GAGP_OK) return st;
  return copy_out(text, buf, cap, needed);
}

void gagp_owner_destroy(gagp_owner* o) { delete o; }

}  // extern "C"
