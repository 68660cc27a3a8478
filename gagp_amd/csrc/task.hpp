// Task (`gagp-task-v1`), fitness cases (`fitness-cases`) and search
// configuration (`gagp-search-config-v1`) documents, plus the effective
// search record written to result.json.
//
// Reference behaviour: cpp/src/cli/task.cpp:62-118 (task),
// cpp/src/cli/program_json.cpp:806-842 and codec.cpp:150-211 (cases),
// cpp/src/evolution/case_set.cpp:48-82 (input names sorted, locals 0..n-1),
// cpp/src/cli/search_config.cpp:60-255 and cpp/src/cli/run.cpp:36-49,285-303
// (built-in default search, overrides, effective record).
#pragma once
#include <cstdint>
#include <optional>
#include <string>
#include <vector>
#include "core.hpp"
#include "json.hpp"

namespace gagp_amd {

struct Task {
  std::string name, grammar_path, train_path, test_path;
  Policy policy{METRIC_ABSOLUTE, PROX_CAPPED, 0, 0.1, 100.0};
  double case_cap = 1000;
};
Task load_task(const std::string& path);

// A case set in evaluator layout: inputs sorted by name, SoA per input.
struct Cases {
  std::vector<std::string> input_names;   // sorted
  std::vector<std::uint8_t> input_types;  // per input
  int n = 0;
  std::vector<Val> inputs;                // [n][n_inputs]
  std::vector<std::uint8_t> set;          // [n][n_inputs] 1 if bound
  std::vector<Val> expected;              // [n]
  std::vector<std::int64_t> arena;        // IntList elements (views: offset*65536+len)
};
Cases load_cases(const std::string& path);
Cases parse_cases(const json::Value& doc);

enum class Schedule { Generational, SteadyState };
struct SearchConfig {
  bool resident = true;
  int population = 1024, generations = 64;
  std::uint64_t seed = 0;
  int tournament = 2;
  int elitism = 1;
  double crossover_probability = 0.7;
  int crossover_attempts = 8;
  double mutation_probability = 0.3;
  int mutation_max_nodes = 0;
  bool uniform_phase = true, counted = true;
  int site_depth = 7;
  bool grammar_sites = false;
  std::optional<double> penalty;
  std::optional<int> fuel;
  Schedule schedule = Schedule::SteadyState;
  int replacement_size = 2;
  int batch = 0;  // 0: population
};
// The built-in default search (configs/search/default.json) with a search
// document applied on top; an explicit variation member resets the whole
// variation policy to the document's values and the policy defaults.
SearchConfig default_search_config();
void apply_search(SearchConfig& s, const json::Value& doc);
json::Value effective_search_json(const SearchConfig& s, double penalty, int fuel);

}  // namespace gagp_amd
