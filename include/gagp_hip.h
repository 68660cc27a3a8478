/*
 * gagp_hip.h -- C ABI of the MI355X GAGP engine (libgagp_hip.so).
 *
 * Plain C types only: pointers, sizes, status codes. Every function returns
 * GAGP_OK or an error status; gagp_last_error() returns the thread-local
 * message of the last failure. Handles are opaque; device buffers live inside
 * them. Reference interfaces each entry point replaces are cited per function
 * (paths relative to the hschi1106/gagp tree).
 */
#ifndef GAGP_HIP_H
#define GAGP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GAGP_OK = 0,
  GAGP_ERR_INVALID = 1,     /* invalid document, argument or program */
  GAGP_ERR_UNSUPPORTED = 2, /* outside the engine: the reference would decline to native_ast */
  GAGP_ERR_DEVICE = 3,      /* HIP runtime or device-side failure */
};

/* Runtime value, 16 bytes: payload (int64 or float bits), Bool flag, tag.
 * Tags: 0 Int, 1 Float, 2 Bool, 3 Char, 4 String, 5 IntList, 6 FloatList,
 * 7 StringList, 8 FallbackToken, 9 Invalid.
 * IntList values are views into the session's element arena:
 * payload = offset * 65536 + length.
 * Reference: Value, cpp/include/gagp/core/value.hpp:14-33. */
typedef struct gagp_value {
  int64_t i;
  uint8_t b;
  uint8_t tag;
  uint8_t pad[6];
} gagp_value;

/* Bytecode instruction. Reference: Instr, cpp/include/gagp/core/bytecode.hpp:15-21. */
typedef struct gagp_instr {
  uint8_t op;
  uint8_t has_a;
  uint8_t has_b;
  uint8_t pad;
  int32_t a;
  int32_t b;
} gagp_instr;

/* One bytecode program. Reference: BytecodeProgram,
 * cpp/include/gagp/core/bytecode.hpp:48-56 (without bounded-region segments). */
typedef struct gagp_program {
  const gagp_instr* code;
  int32_t code_len;
  const uint32_t* fuel; /* per-instruction charge, or NULL for the unit schedule */
  const gagp_value* consts;
  int32_t n_consts;
  int32_t n_locals;
} gagp_program;

/* Fitness policy. metric: 0 Legacy, 1 Absolute, 2 Exact; proximity: 0 Capped,
 * 1 Discounted. Reference: FitnessPolicy, cpp/include/gagp/core/fitness_policy.hpp:6-15. */
typedef struct gagp_policy {
  int32_t metric;
  int32_t proximity;
  int32_t summary_count;
  double summary_weight;
  double summary_cap;
} gagp_policy;

/* int32 VALU issue-rate probe: device milliseconds of one launch of
 * `blocks` blocks of 256 threads, each thread running `iterations`
 * iterations of 8 dependent-pair int32 add/xor operations in four
 * independent chains (8 * iterations * 256 * blocks lane operations).
 * Calibrates the VALU roofline; no reference counterpart. */
int gagp_valu_probe(int32_t blocks, int32_t iterations, float* out_ms);

/* Message of the calling thread's last failure ("" if none). */
const char* gagp_last_error(void);
/* Number of visible HIP devices (0 without a GPU). */
int gagp_device_count(void);

/* ---- evaluation boundary ---- */
typedef struct gagp_session gagp_session;

/* Uploads a case set once. case_inputs/input_set are [n_cases][n_inputs]
 * (input_set may be NULL: all bound); inputs are the evaluator locals
 * 0..n_inputs-1 (sorted input names). lanes is the canonical reduction width
 * (512 for `gagp run`).
 * Reference: FitnessSessionGpu::init, cpp/include/gagp/runtime/gpu/fitness_gpu.hpp:91-121,
 * cpp/src/runtime/gpu/fitness_gpu.cu:540-713. */
int gagp_session_create(const gagp_value* case_inputs, const uint8_t* input_set, int32_t n_cases, int32_t n_inputs,
                        const gagp_value* answers, const int64_t* list_arena, int64_t list_arena_len, int32_t fuel,
                        int32_t lanes, double penalty, gagp_policy policy, gagp_session** out);

/* Scores n programs on every case: fitness (canonical 48-bit total), wrong
 * counts and per-program counts[3] = {cases, errors, timeouts}.
 * Reference: FitnessSessionGpu::eval_programs, fitness_gpu.hpp:49-121,
 * fitness_gpu.cu:715-1790; CPU semantics cpp/src/runtime/cpu/fitness_cpu.cpp:57-176. */
int gagp_session_eval(gagp_session* s, const gagp_program* programs, int32_t n, double* out_fitness, uint32_t* out_wrong,
                      uint32_t* out_counts);

/* gagp_session_eval without any device allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference.
 * Same arguments, results and errors as gagp_session_eval; the launch's
 * program table, code, constant pools and per-program results are staged in
 * a fixed per-GPU device workspace of 16 MiB (a file-scope device array),
 * held by every GPU the library runs on until the process ends. Each
 * launch needs, in 256-byte aligned parts, 40 B per program (table),
 * 8 B per instruction, 16 B per constant-pool entry (at least one) and
 * 20 B per program of results; a launch that needs more fails with
 * GAGP_ERR_INVALID before anything reaches the device. Calls that overlap on
 * one GPU share its workspace and race; serialize them. The kernel is not
 * timed (gagp_session_last_kernel_ms keeps its previous value). */
int gagp_session_eval_static(gagp_session* s, const gagp_program* programs, int32_t n, double* out_fitness,
                             uint32_t* out_wrong, uint32_t* out_counts);

/* Per-case outcome of one program: out_values[c] and out_errors[c]
 * (-1 ok, 0 NameError, 1 TypeError, 2 ZeroDivisionError, 3 ValueError, 4 Timeout).
 * Reference: CpuExecutionSession::execute as used by the run judge,
 * cpp/src/cli/run.cpp:208-239. */
int gagp_session_eval_cases(gagp_session* s, const gagp_program* program, gagp_value* out_values, int32_t* out_errors);

/* Device time (ms) of the last gagp_session_eval kernel. */
float gagp_session_last_kernel_ms(gagp_session* s);
void gagp_session_destroy(gagp_session* s);

/* ---- search boundary (`gagp run TASK.json --engine gpu`) ---- */
typedef struct gagp_batch gagp_batch;

/* A batch of searches stepped together, one device kernel sequence per round.
 * Reference: evolve_population / run_phase_population_steady,
 * cpp/src/evolution/evolve.cpp:331-441,492-558; run_task_command,
 * cpp/src/cli/run.cpp:275-370. */
int gagp_batch_create(gagp_batch** out);

/* Admits one search. search_path may be NULL (built-in default search);
 * population/generations <= 0 and seed_is_set == 0 keep the document values.
 * *out_index receives the search's index in the batch.
 * Reference: the `gagp run` options, cpp/src/cli/run.cpp:105-141,281-303. */
int gagp_batch_add(gagp_batch* b, const char* task_path, const char* search_path, int32_t population, int32_t generations,
                   int32_t seed_is_set, uint64_t seed, int32_t* out_index);

/* As gagp_batch_add, with the initial population read from init_path: the
 * first `population` programs of a `gagp run --save-all` population.json
 * made for the task's grammar, each derived from the grammar and uploaded in
 * place of the generated population (GAGP_ERR_INVALID when a program is not
 * a member of the grammar, the grammar hash differs or the document has too
 * few programs; GAGP_ERR_UNSUPPORTED for grammar-population-v2 replays and
 * for programs the resident layout declines -- a site height above the
 * search's site_depth or more phase nodes than a row holds -- which the
 * reference then runs in its native representation).
 * Reference: `gagp run --init`, cpp/src/cli/run.cpp:122,143-173,305-307;
 * import_native_phase_layout, cpp/src/evolution/repro/gpu/phase_layout.hpp:92-260;
 * the resident layout's import checks, phase_population_cpu.cpp:229-235. */
int gagp_batch_add_init(gagp_batch* b, const char* task_path, const char* search_path, int32_t population,
                        int32_t generations, int32_t seed_is_set, uint64_t seed, const char* init_path,
                        int32_t* out_index);

/* Device bytes search `index` will hold once allocated (its arena's
 * chunks: population and child rows, bytecode, cases, history, keys),
 * computed on the host from its documents without allocating. This is the
 * footprint serving admits by. No reference counterpart: the reference
 * sizes its buffers inside the owner's constructor,
 * cpp/src/evolution/repro/gpu/phase_population.cu:75-79. */
int gagp_batch_planned_bytes(gagp_batch* b, int32_t index, uint64_t* out);

/* Runs every admitted search to completion (use_graph: capture one round as
 * a HIP graph and replay it). */
int gagp_batch_run(gagp_batch* b, int32_t use_graph);

/* Split form of gagp_batch_run: gagp_batch_prepare allocates every search
 * and scores its initial population; gagp_batch_step advances all live
 * searches by `rounds` steady-state rounds (searches past their generation
 * budget idle); gagp_batch_finish checks device errors and assembles the
 * results. */
int gagp_batch_prepare(gagp_batch* b);
int gagp_batch_step(gagp_batch* b, int32_t rounds, int32_t use_graph);
int gagp_batch_finish(gagp_batch* b);

/* Rounds the longest search of a prepared batch needs: P*G/B
 * (cpp/src/evolution/evolve.cpp:385). */
int gagp_batch_max_rounds(gagp_batch* b, int32_t* out);

/* Device counters of search `index`: out[14] = {evaluated programs,
 * evaluated program-cases, runtime errors, timeouts, VM instructions
 * executed (charged fuel), instructions and constant-pool entries of the
 * evaluated programs, and -- when the batch was prepared with profiling on
 * (gagp_batch_set_profile) -- children made, their nodes, their constants,
 * nodes and constants of the parent and donor rows they read, instructions
 * and pool entries lowered for them}.
 * Reference counterpart: the per-generation evaluation totals of
 * cpp/src/evolution/evolve.cpp:347-350,403-409. */
int gagp_batch_counters(gagp_batch* b, int32_t index, uint64_t* out);

/* result.json of search `index` (gagp-run-v1, canonical JSON). If buf is
 * NULL or cap is too small, *needed receives the size including the NUL.
 * Reference: cpp/src/cli/run.cpp:315-367. */
int gagp_batch_result_json(gagp_batch* b, int32_t index, char* buf, size_t cap, size_t* needed);

/* Evaluated program-case count of search `index` (non-reused children times
 * train cases, plus the initial population). Reference: evolve.cpp:586,600. */
int gagp_batch_evaluated_program_cases(gagp_batch* b, int32_t index, uint64_t* out);

/* Per-kernel device milliseconds when the batch ran with profiling on
 * (gagp_batch_set_profile before gagp_batch_run): a JSON object
 * {"k_eval": ms, ...}. */
int gagp_batch_set_profile(gagp_batch* b, int32_t on);
int gagp_batch_kernel_times_json(gagp_batch* b, char* buf, size_t cap, size_t* needed);

/* Evaluation launch autotuner. In the first round of a new round table,
 * between reproduction and placement, the evaluation kernels are timed over
 * that round's pending programs under each candidate launch shape -- 1, 2
 * or 4 pending programs per wavefront of the typed straight-line VM; 1, 2
 * or 4 64-case tiles per work item of the looping VM -- and the fastest
 * shapes (cached per batch shape for the process) are captured into the
 * round graph. Results are identical under every shape. mode: 0 off (fixed
 * shapes: one program per wavefront, one tile per item), 1 (the default)
 * when a round evaluates at least 2^22 program-cases, 2 always.
 * gagp_batch_tuning_json returns the last pass: {"key", "cached",
 * "typed_ms", "loop_ms", "fast_programs", "loop_group"}, with "skipped"
 * when the round was too small ("{}" before the first). No reference
 * counterpart: the reference picks its launch profile by heuristics,
 * cpp/src/runtime/gpu/fitness_gpu.cu:754-980. */
int gagp_batch_set_autotune(gagp_batch* b, int32_t mode);
int gagp_batch_tuning_json(gagp_batch* b, char* buf, size_t cap, size_t* needed);
void gagp_batch_destroy(gagp_batch* b);

/* ---- serving: a request loop with continuous batching ----
 * No counterpart in the reference, which runs one search per process; every
 * published result is the reference's `gagp run` result for that request
 * (cpp/src/cli/run.cpp:275-370), or, with early_exit, the result of the run
 * with as many generations as were completed when the search first had a
 * best member with no wrong train case. */
typedef struct gagp_server gagp_server;

/* budget_bytes: device memory the live searches may hold together;
 * poll_rounds: rounds stepped between completion checks. */
int gagp_server_create(uint64_t budget_bytes, int32_t poll_rounds, gagp_server** out);

/* At most max_live searches on the device at once (0, the default: only
 * the budget limits them; 1 serves one search at a time). */
int gagp_server_set_max_live(gagp_server* s, int32_t max_live);

/* Queues one request arriving `arrival_s` seconds after gagp_server_run
 * starts; arguments as gagp_batch_add. *out_ticket identifies it. */
int gagp_server_submit(gagp_server* s, const char* task_path, const char* search_path, int32_t population, int32_t generations,
                       int32_t seed_is_set, uint64_t seed, int32_t early_exit, double arrival_s, int32_t* out_ticket);

/* Serves every submitted request to completion in real time: a request is
 * admitted once it has arrived and its device footprint (computed from its
 * documents before anything is allocated) fits the budget beside the live
 * searches, stepped with all live searches in one kernel sequence per round,
 * and retired when done. A request that cannot be served (invalid or
 * unsupported documents, a footprint above the whole budget, a failed device
 * allocation) fails alone with its own status; the others are served. */
int gagp_server_run(gagp_server* s);

/* times[3] = {arrival, admission, completion} in seconds from the start of
 * gagp_server_run; *generations of the published result; *status is the
 * request's own status (GAGP_OK or an error, see gagp_last_error after
 * gagp_server_result_json). */
int gagp_server_outcome(gagp_server* s, int32_t ticket, double* times, int32_t* generations, int32_t* status);
int gagp_server_result_json(gagp_server* s, int32_t ticket, char* buf, size_t cap, size_t* needed);
void gagp_server_destroy(gagp_server* s);

/* ---- Owner seam: the reference's generation loop drives the device ----
 * One steady-state search whose population stays on the device while the
 * caller runs the generation loop: the caller keeps fitness, ranks,
 * selection values and the run generator, and picks each child's victim.
 * Reference: the owner adapter GpuPhaseOwner and its NativePhasePopulation
 * (cpp/src/evolution/repro/gpu/phase_population.hpp:15-59), driven by
 * run_phase_population_steady (cpp/src/evolution/evolve.cpp:331-441). */
typedef struct gagp_owner gagp_owner;

/* Builds the search of a task (search_path, population, seed and init_path
 * as gagp_batch_add_init), its initial population generated (or read),
 * lowered and scored on the device. Declines (GAGP_ERR_UNSUPPORTED) as
 * gagp_batch_add does. A steady-state search is driven by begin_steady,
 * make_batch, evaluate_batch and commit_batch; a generational one by
 * reproduce and evaluate.
 * Reference: GpuPhaseOwner construction, evolve.cpp:457-470. */
int gagp_owner_create(const char* task_path, const char* search_path, int32_t population, int32_t seed_is_set, uint64_t seed,
                      const char* init_path, gagp_owner** out);
/* population_size(); the configuration's steady-state batch B. */
int gagp_owner_population_size(gagp_owner* o, int32_t* out);
int gagp_owner_batch_size(gagp_owner* o, int32_t* out);
/* 1 when scores carry wrong counts (every metric but legacy; the reference's
 * evaluate returns an empty wrong vector otherwise and ranks by fitness). */
int gagp_owner_has_wrong(gagp_owner* o, int32_t* out);
/* The initial population's scores (evaluate_owner, evolve.cpp:341-350):
 * fitness[P] (ranking-canonical: canonicalize_fitness_for_ranking leaves
 * them unchanged), wrong[P], counts[P][5] = {program cases, errors,
 * timeouts, fallbacks, unscored}. */
int gagp_owner_evaluate(gagp_owner* o, double* fitness, uint32_t* wrong, uint64_t* counts);
/* node_counts(grammar_sites): AST nodes, or grammar-site nodes, per member. */
int gagp_owner_node_counts(gagp_owner* o, int32_t grammar_sites, uint32_t* out);
/* begin_steady(B): B must be the configuration's batch. */
int gagp_owner_begin_steady(gagp_owner* o, int32_t batch);
/* make_batch(selection[P], pressure, p_mut, seed) -> reuse[B]: tournaments
 * on the selection values, crossover and mutation from the seed word,
 * lowering; reuse[b] is the member child b equals, or -1. */
int gagp_owner_make_batch(gagp_owner* o, const double* selection, int32_t pressure, double p_mut, uint64_t seed,
                          int32_t* reuse);
/* Generational schedule: reproduce(selection[P], pressure, p_mut, seed,
 * elites) replaces the population by P children (tournaments on the
 * selection values; the `elites` members of highest value, ties to the
 * lower index, copied into the first children), scored on the device; a
 * child equal to a member reuses its result. gagp_owner_evaluate then
 * reports the new members (counts zero for reused ones).
 * Reference: NativePhasePopulation::reproduce, driven by run_phase_population
 * (cpp/src/evolution/evolve.cpp:253-322). */
int gagp_owner_reproduce(gagp_owner* o, const double* selection, int32_t pressure, double p_mut, uint64_t seed,
                         int32_t elites);
/* evaluate_batch: fitness[B], wrong[B], counts[B][5] of the children whose
 * reuse is -1 (the others are the caller's: their source's scores). */
int gagp_owner_evaluate_batch(gagp_owner* o, double* fitness, uint32_t* wrong, uint64_t* counts);
/* batch_nodes(grammar_sites): node counts of the batch's children. */
int gagp_owner_batch_nodes(gagp_owner* o, int32_t grammar_sites, uint32_t* out);
/* commit_batch: pairs[n][2] = (member, child); each member's row becomes
 * its child's. */
int gagp_owner_commit_batch(gagp_owner* o, const int32_t* pairs, int32_t n);
/* export_member(i): the member's program as ast-prefix-v2 JSON (buffer
 * protocol as gagp_batch_result_json). */
int gagp_owner_export_member(gagp_owner* o, int32_t member, char* buf, size_t cap, size_t* needed);
void gagp_owner_destroy(gagp_owner* o);

#ifdef __cplusplus
}
#endif
#endif /* GAGP_HIP_H */
