// Device-resident GAGP search engine: every search's population, keys,
// random state and history stay in HBM; one steady-state round is a fixed
// sequence of kernels over all searches of a batch (no host round trip), so a
// round can be captured once as a HIP graph and replayed.
#pragma once
#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <vector>
#include "export.hpp"
#include "grammar.hpp"
#include "json.hpp"
#include "phase.hpp"
#include "task.hpp"

namespace gagp_amd {

// Host description of one search (one `gagp run`).
struct SearchSpec {
  std::string task_path;
  std::string search_path;             // empty: built-in default
  std::optional<int> population, generations;
  std::optional<std::uint64_t> seed;
  // Stop at the first generation boundary whose best member has no wrong
  // train case; the result is that of a run with that many generations.
  bool early_exit = false;
  // Initial population: the first `population` programs of a
  // `gagp run --save-all` population document (empty: generated).
  std::string init_path;
};

struct HistoryRow {
  double best_fitness = 0, mean_fitness = 0;
  int best_wrong = 0, best_nodes = 0;
  double elapsed_ms = 0, eval_ms = 0, repro_ms = 0;
};

struct SearchResult {
  json::Value result;                  // result.json document
  std::vector<json::Value> population; // ranked final population (ast JSON)
  std::vector<json::Value> lineage;    // per-generation best (history.json)
  std::string grammar_hash;            // content hash of the search's grammar
  std::uint64_t evaluated_program_cases = 0;
  std::uint64_t evaluated_programs = 0;
  double search_ms = 0;
};

class Search;  // one admitted search (host + device state)

// A batch of searches stepped together, one kernel sequence per round.
class Batch {
 public:
  Batch();
  ~Batch();
  // Admits a search; throws Unsupported (exit 2) for grammars/tasks outside
  // the engine, std::invalid_argument for invalid documents.
  int add(const SearchSpec& spec);
  // Runs every search to completion; `use_graph` captures one round.
  void run(bool use_graph = true);
  // Split form of run(): allocate, build and score the initial populations;
  // then advance every live search by `rounds` steady-state rounds (a search
  // past its generation budget idles); then check for device errors.
  void prepare();
  void step(int rounds, bool use_graph = true);
  void finish();
  // Device counters of search `index` (see gagp_batch_counters).
  std::vector<std::uint64_t> counters(int index) const;
  double init_ms() const;
  SearchResult result(int index, bool save_all);
  std::size_t size() const;
  std::size_t device_bytes() const;
  // Per-kernel device time (ms) accumulated with HIP events, by kernel name.
  std::vector<std::pair<std::string, double>> kernel_times() const;
  void set_profile(bool on);
  // Evaluation launch autotuning: in the first round of a new round table,
  // the evaluation kernels are timed over that round's pending programs
  // under each candidate launch shape and the fastest shapes are kept for
  // the following rounds (cached per batch shape for the process). Results
  // are identical under every shape. Modes: 0 off (fixed shapes), 1 (the
  // default) when a round evaluates at least 2^22 program-cases, 2 always.
  void set_autotune(int mode);
  std::string tuning_report() const;  // JSON of the last tuning pass ("{}" if none)
  // Generation budget of the batch (max over searches).
  int max_rounds() const;

  // Continuous batching. add() parses a search on the host only;
  // planned_footprint() is the device memory admitting it will take (its
  // arena's chunks), computed without allocating; admit() allocates and
  // initializes it between rounds (it joins the round table at the next
  // step) -- if that fails, the search's device memory is returned and the
  // search never joins; drop() discards a search that was added but never
  // admitted; status() reports every search of the table; retire() drops a
  // finished search and returns its device memory (assemble its result
  // first).
  struct Status { int index, round, stopped_at, error; bool done; };
  std::size_t planned_footprint(int index) const;
  void admit(int index);
  // admit() for several searches with one initialization launch sequence.
  // A search whose allocation fails is undone alone; a device error in the
  // shared initialization undoes them all. Returns (index, message) of each
  // search not admitted.
  std::vector<std::pair<int, std::string>> admit_many(const std::vector<int>& indices);
  void drop(int index);
  std::vector<Status> status();
  void retire(int index);
  std::size_t footprint(int index) const;  // device bytes one admitted search holds

 private:
  friend class Owner;
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

// The in-process seam of the reference's resident steady-state loop: the
// owner adapter GpuPhaseOwner (cpp/src/evolution/repro/gpu/phase_population.hpp:15-59)
// that run_phase_population_steady (cpp/src/evolution/evolve.cpp:331-441)
// drives one round at a time (and run_phase_population, evolve.cpp:253-322,
// one generation at a time). The caller keeps fitness, ranks, selection
// values and the run generator and chooses the victims; the owner keeps the
// population on the device, makes and scores each batch, and commits the
// caller's placements, or replaces the whole population by reproduce().
class Owner {
 public:
  // Builds the search, its initial population (generated, or the spec's
  // init_path) lowered and scored on the device.
  explicit Owner(const SearchSpec& spec);
  ~Owner();
  int population_size() const;
  int batch_size() const;  // the search configuration's steady-state batch
  // Whether scores carry wrong counts (every metric but legacy): the
  // reference's evaluate returns none otherwise, and ranks by fitness alone.
  bool has_wrong() const;
  // Scores of the members: ranking-canonical fitness (canonicalizing again
  // changes nothing), wrong train cases, and per member {program cases,
  // errors, timeouts, fallbacks, unscored} (zero for a member that reused
  // its parent's result after reproduce()).
  void evaluate(double* fitness, unsigned* wrong, std::uint64_t* counts);
  // Generational schedule: P children from tournaments on `selection`,
  // crossover and mutation from `seed`, the `elites` members of highest
  // selection value (ties to the lower index) copied into the first
  // children, scored (a child equal to a member reuses its result), and the
  // population replaced by them.
  void reproduce(const double* selection, int pressure, double p_mut, std::uint64_t seed, int elites);
  // AST node counts, or grammar-site counts (`grammar_sites`), per member.
  void node_counts(bool grammar_sites, unsigned* out);
  void begin_steady(int batch);
  // Tournament parents drawn by `selection` (higher wins, earlier draw on
  // ties), crossover and mutation from `seed`, lowering, and the member
  // each child reuses (its row equals it) or -1.
  void make_batch(const double* selection, int pressure, double p_mut, std::uint64_t seed, int* reuse);
  // Scores of the batch's children (those with reuse -1; the others are
  // left to the caller), as evaluate() per member.
  void evaluate_batch(double* fitness, unsigned* wrong, std::uint64_t* counts);
  void batch_nodes(bool grammar_sites, unsigned* out);
  // (member, child) pairs: each member's row becomes that child's.
  void commit_batch(const int* pairs, int n);
  json::Value export_member(int member);

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

// Device milliseconds of the int32 VALU probe: `blocks` blocks of 256
// threads, each running `iters` iterations of 8 VALU operations.
float valu_probe_ms(int blocks, int iters);

// Size of the fixed device workspace of Session::eval_static, per GPU.
constexpr std::size_t kStaticScratchBytes = std::size_t(16) << 20;

// Evaluation session: the FitnessSessionGpu boundary (fixed case set,
// arbitrary verified bytecode programs).
struct ProgramIn {
  std::vector<std::uint8_t> op, has_a, has_b;
  std::vector<std::int32_t> a, b;
  std::vector<std::uint32_t> fuel;  // empty: unit schedule
  std::vector<Val> consts;
  int n_locals = 0;
};
class Session {
 public:
  Session(const Cases& cases, int fuel, int lanes, double penalty, Policy policy);
  ~Session();
  void eval(const std::vector<ProgramIn>& programs, double* fitness, std::uint32_t* wrong, std::uint32_t* counts);
  // eval() without any device allocation: the launch's program table, code,
  // pools and results live in a fixed per-GPU workspace of
  // kStaticScratchBytes (throws std::invalid_argument when a launch needs
  // more, see static_bytes). Calls that overlap on one GPU race on it. The
  // kernel is not timed (last_kernel_ms keeps its value).
  void eval_static(const std::vector<ProgramIn>& programs, double* fitness, std::uint32_t* wrong, std::uint32_t* counts);
  static std::size_t static_bytes(std::size_t programs, std::size_t instructions, std::size_t pool_entries);
  // Per-case outcome of one program: value and error code (ERR_NONE = ok).
  void eval_cases(const ProgramIn& program, std::vector<Val>& values, std::vector<int>& errors);
  float last_kernel_ms() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

// `gagp eval TASK.json PROGRAM.json --split train|test`: the verdict of one
// program (a run result's best, a population's first, or a bare ast-prefix-v2
// program, which must be a member of the task's grammar) on the split's
// cases with the grammar's fuel: cases, correct, wrong, runtime_errors,
// mean_loss, solved, task and split (cpp/src/cli/run.cpp:372-407).
json::Value evaluate_program(const std::string& task_path, const std::string& program_path, const std::string& split);

}  // namespace gagp_amd
