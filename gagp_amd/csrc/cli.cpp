// `gagp run TASK.json [options]` and `gagp eval TASK.json PROGRAM.json` on
// the MI355X engine: the drop-in command line of the reference
// (cpp/src/cli/run.cpp:105-141,267-407, cpp/src/cli/gagp_main.cpp:23-53).
// Errors print one line and exit 2.
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"

using namespace gagp_amd;

namespace {
[[noreturn]] void usage_error(const std::string& m, const char* command = "run") {
  throw std::invalid_argument(m + " (see `gagp " + command + " --help`)");
}

const char* kRunUsage =
    "usage: gagp run TASK.json [options]\n"
    "\n"
    "Search for a program that solves TASK.json (grammar, train cases, optional test cases)\n"
    "on the MI355X engine.\n"
    "\n"
    "  --search FILE       search configuration (gagp-search-config-v1); default:\n"
    "                      configs/search/default.json, built in\n"
    "  --population N      population size (default 1024, or the search file's)\n"
    "  --generations N     generations (default 64, or the search file's)\n"
    "  --seed N            random seed (default 0)\n"
    "  --init FILE         initial population from `gagp run --save-all` (its first\n"
    "                      N programs, made for the task's grammar)\n"
    "  --engine gpu        the device engine (the only engine of this command)\n"
    "  --out DIR           write DIR/result.json\n"
    "  --save-all          also write DIR/population.json and DIR/history.json\n"
    "  --quiet             no progress output\n";

const char* kEvalUsage =
    "usage: gagp eval TASK.json PROGRAM.json [--split train|test]\n"
    "\n"
    "Score a program on the task's test cases (default) or train cases on the MI355X\n"
    "engine. PROGRAM.json is a result.json (its best program), a population.json (its\n"
    "first program) or a single program.\n";

// `gagp eval` (cpp/src/cli/run.cpp:372-407): one canonical verdict line.
int eval_command(const std::vector<std::string>& args) {
  std::vector<std::string> positional;
  std::string split = "test";
  for (std::size_t i = 0; i < args.size(); ++i) {
    if (args[i] == "--help" || args[i] == "-h") { std::cout << kEvalUsage; return 0; }
    if (args[i] == "--split") {
      if (i + 1 >= args.size()) usage_error("--split requires a value", "eval");
      split = args[++i];
    } else if (!args[i].empty() && args[i][0] == '-') usage_error("unknown option " + args[i], "eval");
    else positional.push_back(args[i]);
  }
  if (positional.size() != 2) usage_error("eval needs TASK.json and PROGRAM.json", "eval");
  if (split != "train" && split != "test") usage_error("--split must be train or test", "eval");
  std::cout << json::canonical(evaluate_program(positional[0], positional[1], split)) << '\n';
  return 0;
}

void write_json(const std::filesystem::path& p, const json::Value& v) {
  std::ofstream out(p);
  out << json::canonical(v) << '\n';
  if (!out) throw std::runtime_error("cannot write " + p.string());
}

int run_command(const std::vector<std::string>& args) {
  SearchSpec spec;
  std::string out_dir, engine = "gpu";
  bool save_all = false, quiet = false;
  for (std::size_t i = 0; i < args.size(); ++i) {
    const std::string& a = args[i];
    const auto value = [&]() -> const std::string& {
      if (i + 1 >= args.size()) usage_error(a + " requires a value");
      return args[++i];
    };
    const auto integer = [&](int minimum) {
      const std::string& t = value();
      std::size_t used = 0;
      long long v = 0;
      try { v = std::stoll(t, &used); } catch (const std::exception&) { used = 0; }
      if (used != t.size() || v < minimum || v > (1LL << 30)) usage_error(a + " needs an integer >= " + std::to_string(minimum));
      return int(v);
    };
    if (a == "--help" || a == "-h") { std::cout << kRunUsage; return 0; }
    if (a == "--search") spec.search_path = value();
    else if (a == "--out") out_dir = value();
    else if (a == "--engine") engine = value();
    else if (a == "--seed") {
      const std::string& t = value();
      if (t.empty() || t.find_first_not_of("0123456789") != std::string::npos) usage_error("--seed needs a nonnegative integer");
      spec.seed = std::stoull(t);
    } else if (a == "--population") spec.population = integer(2);
    else if (a == "--generations") spec.generations = integer(1);
    else if (a == "--save-all") save_all = true;
    else if (a == "--quiet") quiet = true;
    else if (a == "--init") spec.init_path = value();
    else if (!a.empty() && a[0] == '-') usage_error("unknown option " + a);
    else if (spec.task_path.empty()) spec.task_path = a;
    else usage_error("unexpected argument " + a);
  }
  if (spec.task_path.empty()) usage_error("missing TASK");
  if (engine == "cpu") throw Unsupported("--engine cpu: this command runs the device engine; the CPU engine is the reference's");
  if (engine != "gpu") usage_error("--engine must be gpu or cpu");
  Batch batch;
  batch.add(spec);
  batch.run(true);
  const SearchResult r = batch.result(0, save_all);
  const auto& res = r.result.obj;
  if (!quiet) {
    std::fprintf(stderr, "gagp run %s: engine=gpu population=%g generations=%g seed=%s cases=%g\n",
                 res.at("task").str.c_str(), res.at("population").num, res.at("generations").num, res.at("seed").str.c_str(),
                 res.at("train_cases").num);
    const auto& best = res.at("best").obj;
    std::fprintf(stderr, "best fitness %.6g, wrong %g of %g train cases, %g nodes; %.1f s\n", best.at("fitness").num,
                 best.at("wrong").num, res.at("train_cases").num, best.at("nodes").num, r.search_ms / 1000);
    if (best.at("text").str.size() <= 400) std::fprintf(stderr, "program: %s\n", best.at("text").str.c_str());
    if (res.count("test")) {
      const auto& t = res.at("test").obj;
      std::fprintf(stderr, "test: %g of %g correct%s\n", t.at("correct").num, t.at("cases").num, t.at("solved").b ? " (solved)" : "");
    }
  }
  if (!out_dir.empty()) {
    const std::filesystem::path out(out_dir);
    std::filesystem::create_directories(out);
    write_json(out / "result.json", r.result);
    if (save_all) {
      json::Value pop = json::make_object();
      pop.obj["format"] = json::make_string("gagp-population-v1");
      pop.obj["grammar_hash"] = json::make_string(r.grammar_hash);
      json::Value programs = json::make_array();
      programs.arr = r.population;
      pop.obj["programs"] = programs;
      write_json(out / "population.json", pop);
      json::Value hist = json::make_object();
      json::Value rows = json::make_array();
      rows.arr = r.lineage;
      hist.obj["history"] = rows;
      write_json(out / "history.json", hist);
    }
  }
  return 0;
}
}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> args(argv + 1, argv + argc);
  try {
    if (args.empty() || args[0] == "--help" || args[0] == "-h") {
      std::cout << "gagp - grammar-guided genetic programming on MI355X\n\n"
                   "  gagp run TASK.json [options]          search for a program\n"
                   "  gagp eval TASK.json PROGRAM.json      score a program (see `gagp eval --help`)\n";
      return args.empty() ? 2 : 0;
    }
    if (args[0] == "run") return run_command(std::vector<std::string>(args.begin() + 1, args.end()));
    if (args[0] == "eval") return eval_command(std::vector<std::string>(args.begin() + 1, args.end()));
    throw std::invalid_argument("unknown command " + args[0]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gagp: %s\n", e.what());
    return 2;
  }
}
