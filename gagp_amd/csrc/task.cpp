// Task, case and search-configuration documents (see task.hpp).
#include "task.hpp"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

namespace gagp_amd {
namespace {
using json::Value;

std::string read_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::invalid_argument("cannot read " + path);
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}
void reject_unknown(const std::string& where, const Value& v, const std::set<std::string>& allowed) {
  if (v.kind != Value::Object) throw std::invalid_argument(where + " must be an object");
  for (const auto& kv : v.obj)
    if (!allowed.count(kv.first)) throw std::invalid_argument("unknown key " + where + "." + kv.first);
}
[[noreturn]] void search_fail(const std::string& m) { throw std::invalid_argument("search config: " + m); }
double number(const Value& v, const std::string& where) {
  if (v.kind != Value::Number || !std::isfinite(v.num)) search_fail(where + " must be a finite number");
  return v.num;
}
int integer(const Value& v, const std::string& where, int lo, int hi) {
  const double r = number(v, where);
  if (std::floor(r) != r || r < lo || r > hi)
    search_fail(where + " must be an integer in [" + std::to_string(lo) + "," + std::to_string(hi) + "]");
  return int(r);
}
double probability(const Value& v, const std::string& where) {
  const double r = number(v, where);
  if (r < 0 || r > 1) search_fail(where + " must be in [0,1]");
  return r;
}
std::string text(const Value& v, const std::string& where) {
  if (v.kind != Value::String) search_fail(where + " must be a string");
  return v.str;
}

// Typed value of a fitness case (codec.cpp:150-211): ints travel as JSON
// numbers and are taken through the double like the reference reader.
Val typed_value(const Value& v, std::vector<std::int64_t>& arena, std::uint8_t& type) {
  if (v.kind != Value::Object) throw std::invalid_argument("typed value must be an object");
  const std::string& t = v.at("type").as_string("type");
  const Value& raw = v.at("value");
  const auto integral = [](const Value& x) -> std::int64_t {
    if (x.kind != Value::Number) throw std::invalid_argument("int typed value requires numeric payload");
    const long long i = static_cast<long long>(x.num);
    if (double(i) != x.num) throw std::invalid_argument("int typed value must be integral");
    return i;
  };
  if (t == "int") { type = TAG_INT; return val_int(integral(raw)); }
  if (t == "bool") {
    if (raw.kind != Value::Bool) throw std::invalid_argument("bool typed value requires bool payload");
    type = TAG_BOOL;
    return val_bool(raw.b);
  }
  if (t == "int_list") {
    const auto& items = raw.as_array("int_list value");
    if (items.size() > 0xffff) throw std::invalid_argument("int_list longer than 65535 elements");
    const std::uint64_t offset = arena.size();
    for (const auto& x : items) arena.push_back(integral(x));
    Val out{};
    out.i = std::int64_t(offset * 65536 + items.size());
    out.tag = TAG_INT_LIST;
    type = TAG_INT_LIST;
    return out;
  }
  throw std::invalid_argument("unsupported case value type: " + t);
}
}  // namespace

Task load_task(const std::string& path) {
  const Value doc = json::parse(read_file(path));
  reject_unknown("task", doc, {"format", "name", "grammar", "train", "test", "fitness"});
  if (!doc.has("format") || doc.at("format").kind != Value::String || doc.at("format").str != "gagp-task-v1")
    throw std::invalid_argument("task " + path + ": format must be gagp-task-v1");
  const auto base = std::filesystem::path(path).parent_path();
  const auto resolve = [&](const char* key, bool required) -> std::string {
    if (!doc.has(key)) {
      if (required) throw std::invalid_argument("task " + path + ": missing " + key);
      return {};
    }
    const auto& v = doc.at(key);
    if (v.kind != Value::String || v.str.empty()) throw std::invalid_argument(std::string(key) + " must be a nonempty string");
    return (base / v.str).string();
  };
  Task t;
  t.name = doc.has("name") ? doc.at("name").as_string("name") : base.filename().string();
  t.grammar_path = resolve("grammar", true);
  t.train_path = resolve("train", true);
  t.test_path = resolve("test", false);
  if (!doc.has("fitness")) throw std::invalid_argument("task " + path + ": missing fitness");
  const auto& f = doc.at("fitness");
  reject_unknown("fitness", f, {"metric", "proximity", "case_cap", "summaries", "list_allocation_limit"});
  const auto& m = f.at("metric").as_string("fitness.metric");
  if (m == "absolute") t.policy.metric = METRIC_ABSOLUTE;
  else if (m == "exact") t.policy.metric = METRIC_EXACT;
  else throw std::invalid_argument("fitness metric outside the engine's scalar metrics: " + m);
  if (f.has("proximity")) {
    const auto& p = f.at("proximity").as_string("fitness.proximity");
    if (p == "capped") t.policy.proximity = PROX_CAPPED;
    else if (p == "discounted") t.policy.proximity = PROX_DISCOUNTED;
    else throw std::invalid_argument("fitness.proximity must be capped or discounted");
  }
  if (f.has("case_cap")) {
    const auto& c = f.at("case_cap");
    if (c.kind != Value::Number || !std::isfinite(c.num) || c.num <= 0)
      throw std::invalid_argument("fitness.case_cap must be a positive number");
    t.case_cap = c.num;
  }
  if (f.has("summaries") || f.has("list_allocation_limit"))
    throw std::invalid_argument("fitness summaries and list allocation limits are outside the engine's metrics");
  return t;
}

Cases parse_cases(const Value& doc) {
  if (doc.kind != Value::Object || !doc.has("format_version") || doc.at("format_version").kind != Value::String ||
      doc.at("format_version").str != "fitness-cases")
    throw std::invalid_argument("input JSON must include format_version=fitness-cases");
  const auto& rows = doc.at("cases").as_array("cases");
  if (rows.empty()) throw std::invalid_argument("cases must not be empty");
  std::set<std::string> names;
  for (const auto& r : rows) for (const auto& kv : r.at("inputs").obj) names.insert(kv.first);
  Cases c;
  c.input_names.assign(names.begin(), names.end());
  c.input_types.assign(c.input_names.size(), TAG_INVALID);
  c.n = int(rows.size());
  c.inputs.assign(std::size_t(c.n) * c.input_names.size(), val_invalid());
  c.set.assign(std::size_t(c.n) * c.input_names.size(), 0);
  for (int i = 0; i < c.n; ++i) {
    const auto& r = rows[i];
    if (r.kind != Value::Object || !r.has("inputs") || !r.has("expected"))
      throw std::invalid_argument("cases[i] must include inputs/expected");
    const auto& in = r.at("inputs");
    for (std::size_t k = 0; k < c.input_names.size(); ++k) {
      if (!in.has(c.input_names[k])) continue;
      std::uint8_t type = TAG_INVALID;
      c.inputs[std::size_t(i) * c.input_names.size() + k] = typed_value(in.at(c.input_names[k]), c.arena, type);
      c.set[std::size_t(i) * c.input_names.size() + k] = 1;
      if (c.input_types[k] == TAG_INVALID) c.input_types[k] = type;
      else if (c.input_types[k] != type) throw std::invalid_argument("case input has inconsistent types");
    }
    std::uint8_t t = TAG_INVALID;
    c.expected.push_back(typed_value(r.at("expected"), c.arena, t));
  }
  return c;
}
Cases load_cases(const std::string& path) { return parse_cases(json::parse(read_file(path))); }

SearchConfig default_search_config() { return SearchConfig{}; }

void apply_search(SearchConfig& s, const Value& doc) {
  reject_unknown("document", doc, {"format", "representation", "population_size", "generations", "seed", "selection",
      "elitism", "crossover", "mutation", "sites", "limits", "ranking", "fitness", "execution", "schedule"});
  if (!doc.has("format") || text(doc.at("format"), "format") != "gagp-search-config-v1")
    search_fail("format must be \"gagp-search-config-v1\"");
  if (doc.has("representation")) {
    const auto r = text(doc.at("representation"), "representation");
    if (r == "resident") s.resident = true;
    else if (r == "native_ast") s.resident = false;
    else search_fail("representation must be resident or native_ast");
  }
  if (doc.has("population_size")) s.population = integer(doc.at("population_size"), "population_size", 1, 1 << 20);
  if (doc.has("generations")) s.generations = integer(doc.at("generations"), "generations", 1, 1 << 20);
  if (doc.has("seed")) {
    const auto& v = doc.at("seed");
    if (v.kind == Value::String) {
      if (v.str.empty() || v.str.size() > 20 || v.str.find_first_not_of("0123456789") != std::string::npos)
        search_fail("seed string must be decimal digits");
      errno = 0;
      char* end = nullptr;
      const auto parsed = std::strtoull(v.str.c_str(), &end, 10);
      if (errno || *end) search_fail("seed out of range");
      s.seed = parsed;
    } else {
      const double r = number(v, "seed");
      if (std::floor(r) != r || r < 0 || r > 9007199254740992.0) search_fail("numeric seed must be an integer in [0,2^53]");
      s.seed = std::uint64_t(r);
    }
  }
  if (doc.has("selection")) {
    const auto& v = doc.at("selection");
    reject_unknown("selection", v, {"method", "size"});
    if (v.has("method") && text(v.at("method"), "selection.method") != "tournament") search_fail("selection.method must be tournament");
    if (v.has("size")) s.tournament = integer(v.at("size"), "selection.size", 1, 1 << 20);
  }
  if (doc.has("elitism")) s.elitism = integer(doc.at("elitism"), "elitism", 0, 1 << 20);
  // Variation members are explicit as a group.
  std::optional<double> cp;
  std::optional<int> ca, mm;
  std::optional<bool> up, co;
  if (doc.has("crossover")) {
    const auto& v = doc.at("crossover");
    reject_unknown("crossover", v, {"probability", "attempts"});
    if (v.has("probability")) cp = probability(v.at("probability"), "crossover.probability");
    if (v.has("attempts")) ca = integer(v.at("attempts"), "crossover.attempts", 1, 64);
  }
  if (doc.has("mutation")) {
    const auto& v = doc.at("mutation");
    reject_unknown("mutation", v, {"probability", "max_subtree_nodes"});
    if (v.has("probability")) s.mutation_probability = probability(v.at("probability"), "mutation.probability");
    if (v.has("max_subtree_nodes")) mm = integer(v.at("max_subtree_nodes"), "mutation.max_subtree_nodes", 0, 1024);
  }
  if (doc.has("sites")) {
    const auto& v = doc.at("sites");
    reject_unknown("sites", v, {"scope", "draw"});
    if (v.has("scope")) {
      const auto n = text(v.at("scope"), "sites.scope");
      if (n == "genome") up = false;
      else if (n == "uniform_phase") up = true;
      else search_fail("sites.scope must be genome or uniform_phase");
    }
    if (v.has("draw")) {
      const auto n = text(v.at("draw"), "sites.draw");
      if (n == "reservoir") co = false;
      else if (n == "counted") co = true;
      else search_fail("sites.draw must be reservoir or counted");
    }
  }
  if (cp || ca || mm || up || co) {
    s.crossover_probability = cp.value_or(1.0);
    s.crossover_attempts = ca.value_or(8);
    s.mutation_max_nodes = mm.value_or(0);
    s.uniform_phase = up.value_or(false);
    s.counted = co.value_or(false);
  }
  if (doc.has("limits")) {
    const auto& v = doc.at("limits");
    reject_unknown("limits", v, {"site_depth"});
    if (v.has("site_depth")) s.site_depth = integer(v.at("site_depth"), "limits.site_depth", 0, 15);
  }
  if (doc.has("ranking")) {
    const auto& v = doc.at("ranking");
    reject_unknown("ranking", v, {"tie_break"});
    if (v.has("tie_break")) {
      const auto n = text(v.at("tie_break"), "ranking.tie_break");
      if (n == "ast_nodes") s.grammar_sites = false;
      else if (n == "grammar_sites") s.grammar_sites = true;
      else search_fail("ranking.tie_break must be ast_nodes or grammar_sites");
    }
  }
  if (doc.has("execution")) {
    const auto& v = doc.at("execution");
    reject_unknown("execution", v, {"fuel"});
    if (v.has("fuel")) s.fuel = integer(v.at("fuel"), "execution.fuel", 1, 2147483647);
  }
  if (doc.has("fitness")) {
    const auto& v = doc.at("fitness");
    reject_unknown("fitness", v, {"penalty"});
    if (v.has("penalty")) {
      const double p = number(v.at("penalty"), "fitness.penalty");
      if (p < 0) search_fail("fitness.penalty must be nonnegative");
      s.penalty = p;
    }
  }
  if (doc.has("schedule")) {
    const auto& v = doc.at("schedule");
    reject_unknown("schedule", v, {"mode", "replacement_size", "batch"});
    Schedule mode = Schedule::Generational;
    int rs = 2, batch = 0;
    if (v.has("mode")) {
      const auto n = text(v.at("mode"), "schedule.mode");
      if (n == "generational") mode = Schedule::Generational;
      else if (n == "steady_state") mode = Schedule::SteadyState;
      else search_fail("schedule.mode must be generational or steady_state");
    }
    if (v.has("replacement_size")) rs = integer(v.at("replacement_size"), "schedule.replacement_size", 1, 64);
    if (v.has("batch")) batch = integer(v.at("batch"), "schedule.batch", 1, 1 << 20);
    if (mode != Schedule::SteadyState && batch) search_fail("batch must be >= 1 (or omitted) for steady_state and absent otherwise");
    if (mode != Schedule::SteadyState && rs != 2) search_fail("replacement size is a steady_state member");
    s.schedule = mode;
    s.replacement_size = rs;
    s.batch = batch;
  }
}

json::Value effective_search_json(const SearchConfig& s, double penalty, int fuel) {
  using json::make_number;
  using json::make_object;
  using json::make_string;
  Value out = make_object();
  out.obj["format"] = make_string("gagp-search-config-v1");
  out.obj["representation"] = make_string(s.resident ? "resident" : "native_ast");
  out.obj["population_size"] = make_number(s.population);
  out.obj["generations"] = make_number(s.generations);
  out.obj["seed"] = make_string(std::to_string(s.seed));
  Value sel = make_object();
  sel.obj["method"] = make_string("tournament");
  sel.obj["size"] = make_number(s.tournament);
  out.obj["selection"] = sel;
  out.obj["elitism"] = make_number(s.elitism);
  Value mut = make_object();
  mut.obj["probability"] = make_number(s.mutation_probability);
  out.obj["mutation"] = mut;
  Value fit = make_object();
  fit.obj["penalty"] = make_number(penalty);
  out.obj["fitness"] = fit;
  Value exe = make_object();
  exe.obj["fuel"] = make_number(fuel);
  out.obj["execution"] = exe;
  if (s.resident) {
    out.obj["mutation"].obj["max_subtree_nodes"] = make_number(s.mutation_max_nodes);
    Value x = make_object();
    x.obj["probability"] = make_number(s.crossover_probability);
    x.obj["attempts"] = make_number(s.crossover_attempts);
    out.obj["crossover"] = x;
    Value sites = make_object();
    sites.obj["scope"] = make_string(s.uniform_phase ? "uniform_phase" : "genome");
    sites.obj["draw"] = make_string(s.counted ? "counted" : "reservoir");
    out.obj["sites"] = sites;
    Value lim = make_object();
    lim.obj["site_depth"] = make_number(s.site_depth);
    out.obj["limits"] = lim;
    Value rk = make_object();
    rk.obj["tie_break"] = make_string(s.grammar_sites ? "grammar_sites" : "ast_nodes");
    out.obj["ranking"] = rk;
    if (s.schedule == Schedule::SteadyState) {
      Value sc = make_object();
      sc.obj["mode"] = make_string("steady_state");
      sc.obj["batch"] = make_number(s.batch ? s.batch : s.population);
      sc.obj["replacement_size"] = make_number(s.replacement_size);
      out.obj["schedule"] = sc;
    }
  }
  return out;
}

}  // namespace gagp_amd
