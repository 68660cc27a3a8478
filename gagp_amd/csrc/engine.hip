// Device-resident search engine for gfx950 (see engine.hpp).
//
// One steady-state round (cpp/src/evolution/evolve.cpp:331-441 with the owner
// bodies of cpp/src/evolution/repro/gpu/phase_population_cpu.cpp:314-394) is
// the kernel sequence
//   k_record  history row of generation r/per_epoch (best row, fitness vector)
//   k_vary    tournaments, crossover, mutation, compare, completed-result
//             reuse through the member table, lowering, pending lists
//             (wavefront/child)
//   k_eval    bytecode VM + canonical reduction (wavefront/program)
//   k_place   mt19937_64 inverse-tournament placement (block/search), with
//             k_clear emptying the member table beside it
//   k_commit  placed children overwrite their members; every member is
//             indexed in the next round's table (wavefront/member)
// over every search of a batch (grid.y = search). Nothing returns to the
// host between rounds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <fstream>
#include <functional>
#include <numeric>
#include <sstream>
#include <stdexcept>

#include "engine.hpp"
#include "population.hpp"
#include "phase_wave.hpp"
#include "vm.hpp"

namespace gagp_amd {

#define HIP_OK(expr)                                                                                 \
  do {                                                                                               \
    const hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
  } while (0)

namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kScoreCases = 4096;  // largest case count with an LDS score buffer per wavefront
constexpr int kEvalThreads = 64 * kWavesPerBlock;

struct Counters {
  // Evaluated programs and program-cases, runtime errors, timeouts, VM
  // instructions executed (charged fuel), instructions of the evaluated
  // programs and their constant-pool entries; when profiling, the children
  // made, their nodes and constants, the nodes and constants of the parent
  // and donor rows they read, and the instructions and pool entries lowered
  // for them.
  unsigned long long programs, program_cases, errors, timeouts, vm_instructions, code_words, pool_words;
  unsigned long long children, child_nodes, child_consts, parent_nodes, parent_consts, child_code_words, child_pool_words;
};

// Device state of one search. Every pointer is device memory owned by the
// host Search object.
// Totals of one program's tiles in the looping kernel (device atomics).
struct LoopAcc {
  unsigned long long sum;   // two's-complement int64 sum of the integral scores
  unsigned long long fuel;  // charged fuel
  unsigned wrong, errors, timeouts, fractional, tiles, pad;
};

struct DSearch {
  int active;
  int P, B, R, code_cap, pool_cap, stack_slots;
  int rounds, per_epoch;
  int fuel, lanes;
  double penalty;
  Policy policy;
  int pressure, replacement_size;
  double p_mut;
  OpPolicy pol;
  SlotInfo slot;
  int fixed_nodes, grammar_sites, has_wrong, expected_type, n_locals, site_depth;
  std::uint64_t seed;
  PhaseTables g;
  int stage_tables;  // the phase tables fit the reproduction kernel's LDS
  GAS const std::uint8_t* local_types;
  CaseData cd;
  GAS int *cur_origin, *cur_site, *cur_n, *cur_nc;
  GAS Val* cur_cval;
  GAS int *x_origin, *x_site, *x_n, *x_nc;
  GAS Val* x_cval;
  GAS int *t_origin, *t_site;
  GAS Val* t_cval;
  GAS int *s_origin, *s_site;
  GAS Val* s_cval;
  GAS double* fitness;
  GAS unsigned *wrong, *nodes;
  GAS OpStats* stats;
  GAS int* reuse;
  GAS double* c_fit;
  GAS unsigned *c_wrong, *c_nodes;
  GAS Ins* code;
  GAS Val* pool;
  GAS int *code_len, *pool_len, *code_jumps;
  GAS std::uint64_t* child_hash;
  // pending[2B]: children to score, general ones from the front of [0, B),
  // vm_fast ones from its back, vm_loop ones in [B, 2B); n_pending[3] counts each.
  GAS int *pending, *n_pending;
  GAS std::uint64_t* table_key;  // genotype hash | 1, 0 = empty
  GAS int* table;                // a member with that hash
  int table_mask;
  GAS std::uint64_t* mt;      // std::mt19937_64 state after its last twist
  GAS std::uint64_t* mt_left; // tempered words generated and not yet consumed
  GAS int* n_left;
  int place_lds;          // placement keys fit in LDS
  GAS std::uint64_t* member_hash;
  GAS std::uint64_t* round_seed;
  GAS int *round, *best, *placed;
  GAS int *hist_origin, *hist_n, *hist_nc;
  GAS Val* hist_cval;
  GAS double *hist_fit, *hist_all;
  GAS unsigned* hist_wrong;
  GAS Counters* counters;
  GAS int* error;
  GAS unsigned *e_errors, *e_timeouts;  // per evaluated slot
  GAS unsigned long long* e_fuel;
  int profile;  // tally reproduction row sizes too
  int generational, elites;  // generational schedule (B = P, one round per generation) and its elite count
  int early_exit;            // stop at the first generation boundary whose best member has no wrong case
  GAS int* stop;             // generations completed when it stopped early, else 0
  GAS int* elite;            // elite members of the current generation, best first
  int temp_slots;      // hidden-local slots per lane of the general VM (temporaries, grammar locals, input shadows)
  int input_shadow;    // the last input_shadow of them shadow stored inputs (statement input targets), else 0
  int walks;           // the grammar has traversals: lowering runs the kWalks instantiations
  int imported;        // the initial population is uploaded, not generated
  // Selection values of a caller-driven round (the owner seam): tournaments
  // compare them alone; null in the engine's own rounds.
  GAS const double* select;
  int general_scores;  // the general evaluation kernel has an LDS score buffer for this search
  GAS double* loop_scores;  // the looping kernel's case scores: C per program slot
  GAS LoopAcc* loop_acc;    // the looping kernel's per-program tile totals (zero between evaluations)
  int stage_cases;     // the typed evaluation kernel stages the case set in LDS
  int typed_tiles;     // typed programs go to k_eval_tiles (more than one 64-case tile, C <= kScoreCases)
};

__device__ inline CRow member_row(const DSearch& d, int m) {
  const std::size_t o = std::size_t(m) * d.R;
  return CRow{d.cur_origin + o, d.cur_site + o, d.cur_cval + o, d.cur_n[m], d.cur_nc[m]};
}
__device__ inline CRow child_row(const DSearch& d, int c) {
  const std::size_t o = std::size_t(c) * d.R;
  return CRow{d.x_origin + o, d.x_site + o, d.x_cval + o, d.x_n[c], d.x_nc[c]};
}
__device__ inline bool search_live(const DSearch& d) { return d.active && *d.round < d.rounds && !*d.error && !*d.stop; }
__device__ inline void raise(const DSearch& d, int code) { atomicCAS(d.error, 0, code); }

// One entry per launched block of a batch kernel: its search, its block
// index within that search's share and the share's size. Batches launch a
// flat grid of exactly the blocks their searches need (no padding to the
// largest search).
struct BlockMap { int search, block, blocks; };

constexpr int kVaryWaves = 4;  // wavefronts (rows) per block of the row kernels

// ---- initial population (phase_population_cpu.cpp:54-113) ----
// One wavefront per member runs the derivation uniformly (as mutation does),
// so members with different retry counts do not diverge within a wavefront.
__global__ __launch_bounds__(64 * kVaryWaves) void k_init_pop(DSearch* all, const BlockMap* map) {
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  const int m = bm.block * kVaryWaves + wave_index();
  if (!d.active || d.imported || m >= d.P) return;
  const int sd = d.site_depth;
  const int ramp = sd < 2 ? sd : 2 + m % (sd - 1);
  const std::size_t o = std::size_t(m) * d.R;
  GenResult gen{ST_INVALID, 0, 0, 0};
  const int room = d.slot.max_variable_nodes < kPhaseCapacity ? d.slot.max_variable_nodes : kPhaseCapacity;
  const PhaseView g = view_of(d.g);
  for (int logical = ramp, attempt = 0; gen.status != ST_OK; ++attempt) {
    if (room <= 0 || room > d.R) { raise(d, 101); return; }
    const std::uint64_t s = splitmix_mix(d.seed ^ splitmix_mix((std::uint64_t(m) * 8 + 0) * 64 + std::uint64_t(attempt)));
    WaveTasks pending;
    gen = generate_phase_with(g, g.root, 0, d.slot.max_depth, room, s, d.cur_origin + o, d.cur_cval + o, d.cur_site + o, logical,
                              pending);
    // A statement program may read an unassigned local: the next seed is tried.
    if (gen.status == ST_OK && d.slot.statement &&
        !locals_defined_w(g, CRow{d.cur_origin + o, d.cur_site + o, d.cur_cval + o, gen.nodes, gen.constants}))
      gen.status = ST_INVALID;
    if (gen.status != ST_OK && attempt % 4 == 3 && logical < sd) ++logical;
    if (attempt >= 256) { raise(d, 102); return; }
  }
  if (wlane() == 0) {
    d.cur_n[m] = gen.nodes;
    d.cur_nc[m] = gen.constants;
    d.placed[m] = -1;
  }
}

// ---- lowering: one wavefront per row ----
// kWalks: the lowering instantiation with traversals (grammars that have them).
// Returns the code flags it stores (code_jumps).
template <bool kWalks>
__device__ inline int lower_into_w(const PhaseView& g, const DSearch& d, const CRow& row, int slot, WalkStore<kWalks>& ws) {
  Ins* code = d.code + std::size_t(slot) * d.code_cap;
  Val* pool = d.pool + std::size_t(slot) * d.pool_cap;
  const LowerResult r =
      lower_phase_w<kWalks>(g, row, d.local_types, d.n_locals, d.expected_type, code, d.code_cap, pool, d.pool_cap, ws);
  const bool ok = r.status == ST_OK && r.stack_bound >= 1 && r.stack_bound <= d.stack_slots &&
                  r.locals - d.n_locals <= d.temp_slots - d.input_shadow;
  // bit 0: jumps; bit 1: typed straight-line (vm_fast); bit 2: Bool result; bit 3: typed looping (vm_loop)
  const int flags = r.jumps | (r.fast ? 2 : 0) | (r.result_type == TAG_BOOL ? 4 : 0) | (r.loop ? 8 : 0);
  if (wlane() == 0) {
    if (!ok) raise(d, 200 + r.status);
    d.code_len[slot] = ok ? r.code : 0;
    d.pool_len[slot] = ok ? r.constants : 0;
    d.code_jumps[slot] = flags;
  }
  return flags;
}
template <bool kWalks>
__global__ __launch_bounds__(64 * kVaryWaves) void k_lower_members(DSearch* all, const BlockMap* map) {
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  const int m = bm.block * kVaryWaves + wave_index();
  if (!d.active || (d.walks != 0) != kWalks || m >= d.P || *d.error) return;
  __shared__ WalkStore<kWalks> walk_lds[kVaryWaves];
  lower_into_w<kWalks>(view_of(d.g), d, member_row(d, m), m, walk_lds[wave_index()]);
}

// ---- member hash table (parent_fitness_cache.cuh:11-58) ----
// One entry per distinct genotype hash: identical members share a hash, and
// any one of them carries the result a child identical to them reuses (the
// rows are compared in full before a result is reused). Keys are hash | 1,
// 0 marks an empty slot. A round reads the table its previous round's
// commit built (the initialization builds the first); it is cleared while
// the previous round places its children.
__global__ void k_clear(DSearch* all, const BlockMap* map) {
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  const int i = bm.block * blockDim.x + threadIdx.x;
  if (!d.active || i > d.table_mask) return;
  d.table_key[i] = 0;
}
// Indexes member m under hash h (one thread).
__device__ inline void insert_member(const DSearch& d, int m, std::uint64_t h) {
  const std::uint64_t key = h | 1;
  for (int probe = 0; probe <= d.table_mask; ++probe) {
    const int slot = int((h + std::uint64_t(probe)) & std::uint64_t(d.table_mask));
    const unsigned long long seen = atomicCAS(reinterpret_cast<unsigned long long*>(d.table_key) + slot, 0ull, key);
    if (seen == 0) { d.table[slot] = m; return; }
    if (seen == key) return;  // a member with this hash is indexed
  }
  raise(d, 400);
}
__global__ void k_index(DSearch* all, const BlockMap* map) {
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  const int m = bm.block * blockDim.x + threadIdx.x;
  if (!search_live(d) || m >= d.P) return;
  insert_member(d, m, d.member_hash[m]);
}
// The member a child identical to it reuses, or -1: the table probe is
// uniform and the row comparison runs across the lanes (rows_equal's test).
__device__ inline int find_member_w(const DSearch& d, const CRow& child, std::uint64_t h) {
  const std::uint64_t key = h | 1;
  for (int probe = 0; probe <= d.table_mask; ++probe) {
    const int slot = int((h + std::uint64_t(probe)) & std::uint64_t(d.table_mask));
    const std::uint64_t k = d.table_key[slot];
    if (k == 0) return -1;
    if (k != key) continue;
    const int m = d.table[slot];
    if (d.member_hash[m] != h) return -1;
    const CRow member = member_row(d, m);
    if (member.n != child.n || member.nc != child.nc) return -1;
    bool mine = false;
    for (int i = wlane(); i < child.n; i += 64) mine = mine || member.origin[i] != child.origin[i];
    for (int i = wlane(); i < child.nc; i += 64) {
      const Val x = member.cval[i], y = child.cval[i];
      mine = mine || x.tag != y.tag || (x.tag == TAG_BOOL ? x.b != y.b : x.i != y.i);
    }
    return __ballot(mine) != 0 ? -1 : m;
  }
  return -1;
}
// The evaluation kernel that scores a program: 1 the typed straight-line
// path, 2 the typed looping path (both with an LDS score buffer, so only
// where the cases fit it), 0 the general VM.
__device__ inline int eval_class(int code_flags, int C) {
  if (C > kScoreCases) return 0;
  return (code_flags & 2) ? 1 : (code_flags & 8) ? 2 : 0;
}
// Pending list of a round: general children from the front of [0, B), the
// straight-line ones from its back, the looping ones in [B, 2B).
__device__ inline int pending_at(int B, int kind, int i) { return kind == 0 ? i : kind == 1 ? B - 1 - i : B + i; }

// ---- reproduction: one wavefront per child (phase_operators.cuh:156-265) ----
// kStaged: the block first copies the phase tables into LDS (searches whose
// tables fit); otherwise the operators read them from global memory.
template <bool kStaged, bool kWalks>
__global__ __launch_bounds__(64 * kVaryWaves) void k_vary(DSearch* all, const BlockMap* map) {
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  if (!search_live(d) || (d.stage_tables != 0) != kStaged || (d.walks != 0) != kWalks) return;  // uniform per block
  extern __shared__ __attribute__((aligned(16))) unsigned char vary_lds[];
  const PhaseView g = kStaged ? stage_phase_tables<64 * kVaryWaves>(d.g, vary_lds) : view_of(d.g);
  const int c = bm.block * kVaryWaves + wave_index();
  if (c >= d.B) return;
  const std::uint64_t seed = *d.round_seed;
  const Keys keys = d.select ? Keys{d.select, nullptr, nullptr, 0} : Keys{d.fitness, d.wrong, d.nodes, d.has_wrong};
  const std::size_t o = std::size_t(c) * d.R;
  Row x{d.x_origin + o, d.x_site + o, d.x_cval + o, 0, 0};
  OpStats st;
  if (d.generational && c < d.elites) {
    // Elite rows replace the first children after variation; a child's
    // draws depend only on (seed, child), so the elite children skip theirs.
    // An elite child is its source member and reuses its result.
    const int src = d.elite[c];
    copy_row_w(member_row(d, src), x);
    st = OpStats{src, 0, 0, 0, 0, 0, 0, src, -1};
    wave_fence();
  } else {
    crossover_w(g, d.slot, c, [&](int m) { return member_row(d, m); }, x, keys, d.P, d.pressure, seed, d.pol, st);
    wave_fence();
    Row tmp{d.t_origin + o, d.t_site + o, d.t_cval + o, 0, 0};
    Row spl{d.s_origin + o, d.s_site + o, d.s_cval + o, 0, 0};
    mutate_w(g, d.slot, c, x, tmp, spl, d.p_mut, seed, d.pol, st);
    wave_fence();
    if (!st.error) compare_w(member_row(d, st.parent), as_const(x), d.P, st);
  }
  const CRow child = as_const(x);
  const std::uint64_t h = st.error ? 0 : row_hash_w(child);
  if (wlane() == 0) {
    d.x_n[c] = x.n;
    d.x_nc[c] = x.nc;
    d.stats[c] = st;
    d.child_hash[c] = h;
    if (st.error) raise(d, 300 + st.error);
  }
  if (st.error) return;
  // A child identical to its parent, or to any member, takes that member's
  // completed result and is neither lowered nor evaluated. Results are
  // fixed before any placement of the round.
  const int r = st.reusable_parent >= 0 ? st.reusable_parent : find_member_w(d, child, h);
  if (r >= 0) {
    if (wlane() == 0) {
      d.code_len[c] = 0;
      d.pool_len[c] = 0;
      d.reuse[c] = r;
      d.c_fit[c] = d.fitness[r];
      d.c_wrong[c] = d.wrong[r];
      d.c_nodes[c] = d.nodes[r];
    }
    return;
  }
  __shared__ WalkStore<kWalks> walk_lds[kVaryWaves];
  const int flags = lower_into_w<kWalks>(g, d, child, c, walk_lds[wave_index()]);
  unsigned n;
  if (d.grammar_sites) {
    unsigned mine = 0;
    for (int i = wlane(); i < child.n; i += 64) mine += child.site[i] >= 0;
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off, 64);
    n = mine;
  } else {
    n = unsigned(d.fixed_nodes) + unsigned(child.n);
  }
  if (wlane() == 0) {
    d.reuse[c] = -1;
    d.c_nodes[c] = n;
    const int k = eval_class(flags, d.cd.C);
    d.pending[pending_at(d.B, k, atomicAdd(d.n_pending + k, 1))] = c;
  }
}

// ---- evaluation: one wavefront per program ----
__device__ inline double wave_sum_d(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}
__device__ inline unsigned wave_sum_u(unsigned x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}
struct EvalParams {
  CaseData cd;
  int fuel, lanes;
  double penalty;
  Policy policy;
};
struct EvalOut { double fitness; unsigned wrong, errors, timeouts; unsigned long long fuel; };

__device__ inline Val expected_of(const CaseData& cd, int c) {
  Val expected{};
  expected.i = cd.exp_val[c]; expected.tag = cd.exp_tag[c]; expected.b = cd.exp_b[c];
  return expected;
}
__device__ inline double lane_score(const EvalParams& e, const LaneResult& r, const Val& expected, bool& exact) {
  if (r.err != ERR_NONE) { exact = false; return -error_ceiling(e.policy, e.penalty); }
  Val actual{};
  actual.i = r.v; actual.tag = r.t; actual.b = r.b;
  return case_score(actual, expected, e.penalty, e.policy, exact);
}
__device__ inline double lane_score(const EvalParams& e, const LaneResult& r, int c, bool& exact) {
  return lane_score(e, r, expected_of(e.cd, c), exact);
}

// The case set of one search copied into a block's LDS for the typed path:
// per case, each input's payload (Bool as 0/1) and a bound mask, and the
// expected value. Every wavefront of a block scores programs of the same
// search, so the block stages it once.
struct StagedCases {
  LAS std::int64_t* value;      // [n_inputs][C]
  LAS std::uint8_t* bound;      // [C]: bit k = input k bound
  LAS std::int64_t* expected;   // [C] payload
  LAS std::uint8_t* exp_tag;    // [C]
  LAS std::uint8_t* exp_b;      // [C]
  bool on;                      // staged for this block
};
// Input rows the typed path reads: those bound by the program's locals and
// present in the case data, at most kFastInputs.
GHD inline int staged_inputs_of(int n_locals, int n_inputs) {
  return std::min(std::min(n_locals, n_inputs), kFastInputs);
}
GHD inline std::size_t staged_case_bytes(int C, int n_rows) { return std::size_t(C) * (8 * std::size_t(n_rows) + 8 + 3); }
template <int kThreads>
__device__ inline StagedCases stage_cases(const CaseData& cd, int n_locals, unsigned char* lds) {
  const int C = cd.C, n_inputs = staged_inputs_of(n_locals, cd.n_inputs);
  StagedCases sc;
  sc.on = true;
  sc.value = (LAS std::int64_t*)lds;
  sc.expected = sc.value + std::size_t(n_inputs) * C;
  sc.bound = (LAS std::uint8_t*)(sc.expected + C);
  sc.exp_tag = sc.bound + C;
  sc.exp_b = sc.exp_tag + C;
  for (int c = threadIdx.x; c < C; c += kThreads) {
    unsigned mask = 0;
    for (int k = 0; k < n_inputs; ++k) {
      const std::size_t at = std::size_t(k) * C + c;
      const std::uint8_t tag = cd.in_tag[at];
      if (tag != TAG_INVALID) mask |= 1u << k;
      sc.value[at] = tag == TAG_BOOL ? std::int64_t(cd.in_b[at] != 0) : cd.in_val[at];
    }
    sc.bound[c] = std::uint8_t(mask);
    sc.expected[c] = cd.exp_val[c];
    sc.exp_tag[c] = cd.exp_tag[c];
    sc.exp_b[c] = cd.exp_b[c];
  }
  __syncthreads();
  return sc;
}
__device__ inline FastInputs staged_inputs(const StagedCases& sc, int C, int n_inputs, int c, bool active) {
  FastInputs in{0, 0, 0, 0, 0};
  if (!active) return in;
  // Unstaged rows beyond n_inputs are unbound, as in fast_inputs.
  in.bound = sc.bound[c];
  if (n_inputs > 0) in.v0 = sc.value[c];
  if (n_inputs > 1) in.v1 = sc.value[std::size_t(C) + c];
  if (n_inputs > 2) in.v2 = sc.value[2 * std::size_t(C) + c];
  if (n_inputs > 3) in.v3 = sc.value[3 * std::size_t(C) + c];
  return in;
}

// Scores one program over all cases with the canonical reduction
// (fitness_cpu.cpp:97-171: L logical lanes over contiguous case chunks,
// 48-bit rounding after every addition, 32-lane groups then the total).
// When every case score is an integer of magnitude <= 2^36 (and C <= 4096)
// every partial sum is an exact integer below 2^48, the rounding is the
// identity and any summation order gives the canonical total.
// kFast: `pv` is certified for the typed straight-line VM (vm_fast);
// otherwise the general VM runs it with its stack in LDS.
// `scores` (this wavefront's LDS buffer of C doubles, or null) keeps every
// case score of the first pass, so the canonical order is replayed from it
// instead of by running the program again; the fast kernel always has it.
template <int kKind>
__device__ EvalOut eval_program(const EvalParams& e, const ProgView& pv, WaveStack& st, double* scores, int lane,
                                StagedCases staged = StagedCases{}) {
  static_assert(kKind == 0 || kKind == 1, "the looping path has its own kernel");
  constexpr bool kFast = kKind != 0;  // a typed path: scores in LDS, inputs untagged
  const int C = e.cd.C;
  double isum = 0;
  unsigned w = 0, errs = 0, touts = 0;
  unsigned long long fuel = 0;
  bool integral = C <= 4096;
  FastProgram fp{};
  if (kKind == 1) fp = fast_load(pv, lane);
  const int n_in = staged_inputs_of(pv.n_locals, e.cd.n_inputs);
  const auto typed = [&](const FastInputs& in, bool active) { return vm_fast(fp, in, active, e.fuel, pv.result_tag); };
  const auto run = [&](int c, bool active) {
    if (kFast) return typed(fast_inputs(e.cd, pv.n_locals, c, active), active);
    return vm_run(pv, e.cd, active ? c : 0, active, e.fuel, st, lane);
  };
  for (int tile = 0; tile < C; tile += 64) {
    const int c = tile + lane;
    const bool active = c < C;
    LaneResult r;
    Val expected{};
    if constexpr (kFast) {
      // One call site of the typed VM for both input sources.
      const FastInputs in =
          staged.on ? staged_inputs(staged, C, n_in, c, active) : fast_inputs(e.cd, pv.n_locals, c, active);
      r = typed(in, active);
      if (active) {
        if (staged.on) { expected.i = staged.expected[c]; expected.tag = staged.exp_tag[c]; expected.b = staged.exp_b[c]; }
        else expected = expected_of(e.cd, c);
      }
    } else {
      r = run(c, active);
      if (active) expected = expected_of(e.cd, c);
    }
    if (active) {
      bool exact = false;
      const double s = lane_score(e, r, expected, exact);
      w += !exact;
      errs += r.err != ERR_NONE;
      touts += r.err == ERR_TIMEOUT;
      fuel += unsigned(r.fuel_used);
      integral = integral && s == double(std::int64_t(s)) && s >= -68719476736.0 && s <= 68719476736.0;
      isum += s;
      if (scores) scores[c] = s;
    }
  }
  EvalOut out;
  out.wrong = wave_sum_u(w);
  out.errors = wave_sum_u(errs);
  out.timeouts = wave_sum_u(touts);
  for (int off = 32; off > 0; off >>= 1) fuel += __shfl_xor(fuel, off, 64);
  out.fuel = fuel;
  if (__ballot(!integral) == 0) {
    out.fitness = wave_sum_d(isum);
    return out;
  }
  // Canonical order: lane j owns logical lanes j, j+64, ... of L.
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // scores[] of other lanes
  __builtin_amdgcn_wave_barrier();
  const int L = e.lanes;
  double group_total = 0;
  for (int base = 0; base < L; base += 64) {
    const int l = base + lane;
    const bool own = l < L;
    const int begin = own ? int((std::int64_t(C) * l) / L) : 0;
    const int end = own ? int((std::int64_t(C) * (l + 1)) / L) : 0;
    int longest = end - begin;
    for (int off = 32; off > 0; off >>= 1) { const int y = __shfl_xor(longest, off, 64); longest = y > longest ? y : longest; }
    double local = 0;
    for (int s = 0; s < longest; ++s) {
      const int c = begin + s;
      const bool active = c < end;
      if constexpr (kFast) {
        if (active) local = canon48(local + scores[c]);
      } else if (scores) {
        if (active) local = canon48(local + scores[c]);
      } else {
        const LaneResult r = run(active ? c : 0, active);
        if (active) {
          bool exact = false;
          local = canon48(local + lane_score(e, r, c, exact));
        }
      }
    }
    // Two 32-lane groups per 64 logical lanes, summed in lane order.
    for (int half = 0; half < 2; ++half) {
      if (base + half * 32 >= L) break;
      double group = 0;
      for (int k = 0; k < 32; ++k) {
        const int src = half * 32 + k;
        const double v = __shfl(local, src, 64);
        if (base + src < L) group = canon48(group + v);
      }
      if (e.policy.metric == METRIC_LEGACY) {
        // Legacy sums every lane partial in order, with no group stage.
        for (int k = 0; k < 32; ++k) {
          const int src = half * 32 + k;
          const double v = __shfl(local, src, 64);
          if (base + src < L) group_total = canon48(group_total + v);
        }
      } else {
        group_total = canon48(group_total + group);
      }
    }
  }
  out.fitness = group_total;
  return out;
}

// The canonical total of C case scores held in `scores` (LDS), computed by
// one wavefront: L logical lanes over contiguous case chunks with 48-bit
// rounding, then 32-lane groups and the total (the legacy metric sums the
// lane partials directly), as eval_program's replay.
__device__ inline double canonical_from_scores(const EvalParams& e, int C, const double* scores, int lane) {
  const int L = e.lanes;
  double group_total = 0;
  for (int base = 0; base < L; base += 64) {
    const int l = base + lane;
    const bool own = l < L;
    const int begin = own ? int((std::int64_t(C) * l) / L) : 0;
    const int end = own ? int((std::int64_t(C) * (l + 1)) / L) : 0;
    int longest = end - begin;
    for (int off = 32; off > 0; off >>= 1) { const int y = __shfl_xor(longest, off, 64); longest = y > longest ? y : longest; }
    double local = 0;
    for (int k = 0; k < longest; ++k) {
      const int c = begin + k;
      if (c < end) local = canon48(local + scores[c]);
    }
    for (int half = 0; half < 2; ++half) {
      if (base + half * 32 >= L) break;
      double group = 0;
      for (int k = 0; k < 32; ++k) {
        const int src = half * 32 + k;
        const double v = __shfl(local, src, 64);
        if (base + src < L) {
          if (e.policy.metric == METRIC_LEGACY) group_total = canon48(group_total + v);
          else group = canon48(group + v);
        }
      }
      if (e.policy.metric != METRIC_LEGACY) group_total = canon48(group_total + group);
    }
  }
  return group_total;
}

// Items of the looping kernel: (program, group of `group` consecutive
// 64-case tiles) pairs, one wavefront each; blocks for n programs (at most
// 8192, which loop over the rest).
__host__ __device__ inline int loop_blocks(int C, int n, int group = 1) {
  const int tiles = (C + 63) / 64, groups = (tiles + group - 1) / group;
  const long long waves = (long long)n * groups;
  return int(std::min<long long>(8192, (waves + kWavesPerBlock - 1) / kWavesPerBlock));
}

// The looping path as a tile queue: every wavefront takes (program, tile
// group) items of its own, program-major, so no wavefront waits for another.
// A group is `group` consecutive 64-case tiles of one program, run one after
// another by the wavefront (the program's decoded code stays in registers).
// Scores go to the program's row of loop_scores (HBM). A program with one
// group is finished by its wavefront; otherwise each group adds its counts
// and its exact integer score sum to the program's LoopAcc with device
// atomics, and the wavefront that completes the last group forms the
// fitness: the sum, or the canonical replay from the scores when some tile
// had a non-integral score. It then clears the record for the program's next
// evaluation. LDS: per wavefront kLoopSlots slots of 64 lanes, independent of
// the case count.
// kGrouped false is the one-tile instantiation (group 1): its loop-carried
// totals fold away, which keeps five wavefronts per SIMD.
template <bool kGrouped>
__global__ __launch_bounds__(kEvalThreads) void k_eval_loop(DSearch* all, const BlockMap* map, int mode, int group_arg) {
  const int group = kGrouped ? group_arg : 1;
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  if (mode == 1 ? !search_live(d) : (!d.active || *d.error)) return;
  extern __shared__ unsigned char smem[];
  const int wave = wave_index(), lane = threadIdx.x % 64;
  const int C = d.cd.C, tiles = (C + 63) / 64, groups = (tiles + group - 1) / group;
  LAS std::int64_t* const slots = (LAS std::int64_t*)smem + std::size_t(wave) * kLoopSlots * 64;
  const EvalParams e{d.cd, d.fuel, d.lanes, d.penalty, d.policy};
  const int count = mode == 1 ? d.n_pending[2] : d.P;
  const long long items = (long long)count * groups;
  for (long long k = (long long)bm.block * kWavesPerBlock + wave; k < items; k += (long long)bm.blocks * kWavesPerBlock) {
    const int i = int(k / groups), first = int(k % groups) * group, last = min(tiles, first + group);
    const int slot = mode == 1 ? d.pending[pending_at(d.B, 2, i)] : i;
    const int pmode = d.code_jumps[slot];
    if (mode == 0 && eval_class(pmode, C) != 2) continue;  // another kernel scores it (uniform per wavefront)
    const ProgView pv{d.code + std::size_t(slot) * d.code_cap, d.pool + std::size_t(slot) * d.pool_cap, d.code_len[slot],
                      d.n_locals, pmode & 1, (pmode >> 1) & 1, d.pool_len[slot], (pmode & 4) ? TAG_BOOL : TAG_INT};
    const LoopProgram lp = loop_load(pv, lane);
    const int n_in = staged_inputs_of(pv.n_locals, e.cd.n_inputs);
    GAS double* const scores = d.loop_scores + std::size_t(slot) * C;
    long long total = 0;  // exact while no score is fractional
    bool frac = false;
    unsigned tw = 0, te = 0, tt = 0;
    unsigned long long tf = 0;
    const int n_tiles = kGrouped ? last - first : 1;  // first < tiles: one tile when ungrouped
    for (int q = 0; q < n_tiles; ++q) {
      // Cases in the search's visiting order (similar loop lengths per
      // tile); scores are kept by case index, so the totals are unchanged.
      const int at = (first + q) * 64 + lane;
      const bool active = at < C;
      const int c = active ? e.cd.order[at] : 0;
      const LaneResult r = vm_loop(lp, fast_inputs(e.cd, pv.n_locals, c, active), n_in, e.cd.arena, slots, lane, active, e.fuel,
                                   pv.result_tag);
      double sc = 0;
      bool exact = true;
      if (active) {
        sc = lane_score(e, r, expected_of(e.cd, c), exact);
        scores[c] = sc;
      }
      // Integral scores below 2^36 in magnitude: every partial sum is exact.
      const bool fractional =
          __ballot(active && !(sc == double(std::int64_t(sc)) && sc >= -68719476736.0 && sc <= 68719476736.0)) != 0;
      const double ws = wave_sum_d(sc);
      frac = frac || fractional;
      if (!frac) total += (long long)ws;
      tw += wave_sum_u(active && !exact);
      te += wave_sum_u(active && r.err != ERR_NONE);
      tt += wave_sum_u(active && r.err == ERR_TIMEOUT);
      unsigned long long f = active ? unsigned(r.fuel_used) : 0u;
      for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off, 64);
      tf += f;
    }
    if (frac) total = 0;
    if (groups > 1) {
      GAS LoopAcc* const acc = d.loop_acc + slot;
      __threadfence();  // this group's scores before its arrival
      unsigned prev = 0;
      if (lane == 0) {
        atomicAdd(&acc->sum, (unsigned long long)total);
        atomicAdd(&acc->fuel, tf);
        atomicAdd(&acc->wrong, tw);
        atomicAdd(&acc->errors, te);
        atomicAdd(&acc->timeouts, tt);
        if (frac) atomicOr(&acc->fractional, 1u);
        __threadfence();
        prev = atomicAdd(&acc->tiles, 1u);
      }
      if (__shfl(prev, 0, 64) != unsigned(groups - 1)) continue;  // another group finishes the program
      __threadfence();  // every group's scores and counts after the last arrival
      unsigned long long v[2] = {0, 0};
      unsigned u[4] = {0, 0, 0, 0};
      if (lane == 0) {
        v[0] = atomicAdd(&acc->sum, 0ull); v[1] = atomicAdd(&acc->fuel, 0ull);
        u[0] = atomicAdd(&acc->wrong, 0u); u[1] = atomicAdd(&acc->errors, 0u);
        u[2] = atomicAdd(&acc->timeouts, 0u); u[3] = atomicAdd(&acc->fractional, 0u);
        // cleared for the program's next evaluation
        atomicExch(&acc->sum, 0ull); atomicExch(&acc->fuel, 0ull); atomicExch(&acc->wrong, 0u);
        atomicExch(&acc->errors, 0u); atomicExch(&acc->timeouts, 0u); atomicExch(&acc->fractional, 0u);
        atomicExch(&acc->tiles, 0u);
      }
      total = (long long)__shfl(v[0], 0, 64);
      tf = __shfl(v[1], 0, 64);
      tw = __shfl(u[0], 0, 64); te = __shfl(u[1], 0, 64); tt = __shfl(u[2], 0, 64);
      frac = __shfl(u[3], 0, 64) != 0;
    } else {
      wave_fence();  // the canonical replay reads the other lanes' scores
    }
    const double fitness = frac ? canonical_from_scores(e, C, scores, lane) : double(total);
    if (lane == 0) {
      // Ranking rounds every published fitness to a 40-bit mantissa.
      if (mode == 1) { d.c_fit[slot] = canon40(fitness); d.c_wrong[slot] = tw; }
      else { d.fitness[slot] = canon40(fitness); d.wrong[slot] = tw; }
      d.e_errors[slot] = te;
      d.e_timeouts[slot] = tt;
      d.e_fuel[slot] = tf;
    }
  }
}

// A wavefront's LDS stack of `slots` operand slots followed by `temps`
// hidden-local slots, 10 bytes per lane and slot; the last `shadow` of them
// shadow stored inputs.
__device__ inline WaveStack stack_of(unsigned char* smem, int slots, int temps, int wave, int shadow) {
  const int all = slots + temps;
  unsigned char* base = smem + std::size_t(wave) * all * 64 * 10;
  return WaveStack{reinterpret_cast<std::int64_t*>(base), base + std::size_t(all) * 64 * 8,
                   base + std::size_t(all) * 64 * 9, slots, temps, shadow};
}

// mode 0: score members 0..P-1 (initial population), mode 1: pending
// children. kFast: the programs certified for vm_fast (pending entries from
// the back of the list); otherwise the rest (from the front).
template <int kKind>
__global__ __launch_bounds__(kEvalThreads) void k_eval(DSearch* all, const BlockMap* map, int mode) {
  constexpr bool kFast = kKind != 0;
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  if (mode == 1 ? !search_live(d) : (!d.active || *d.error)) return;
  if (kFast && d.typed_tiles) return;  // k_eval_tiles scores them
  extern __shared__ unsigned char smem[];
  const int wave = wave_index(), lane = threadIdx.x % 64;
  WaveStack st = stack_of(smem, kFast ? 0 : d.stack_slots, kFast ? 0 : d.temp_slots, wave, kFast ? 0 : d.input_shadow);
  // Score buffers follow the stacks of the block's wavefronts.
  double* scores =
      (kFast ? d.cd.C <= kScoreCases : d.general_scores != 0)
          ? reinterpret_cast<double*>(smem + std::size_t(kWavesPerBlock) * (kFast ? 0 : d.stack_slots + d.temp_slots) * 64 * 10) +
                std::size_t(wave) * d.cd.C
          : nullptr;
  const EvalParams e{d.cd, d.fuel, d.lanes, d.penalty, d.policy};
  // The typed kernel stages the case set after the score buffers.
  StagedCases staged{};
  const bool stage = kFast && d.stage_cases;
  if (stage)
    staged = stage_cases<kEvalThreads>(d.cd, d.n_locals,
                                       smem + std::size_t(kWavesPerBlock) * 8 * std::size_t(d.cd.C));
  const int count = mode == 1 ? d.n_pending[kKind] : d.P;
  for (int i = bm.block * kWavesPerBlock + wave; i < count; i += bm.blocks * kWavesPerBlock) {
    const int slot = mode == 1 ? d.pending[pending_at(d.B, kKind, i)] : i;
    const int pmode = d.code_jumps[slot];
    if (mode == 0 && eval_class(pmode, d.cd.C) != kKind) continue;  // another kernel scores it
    const ProgView pv{d.code + std::size_t(slot) * d.code_cap, d.pool + std::size_t(slot) * d.pool_cap, d.code_len[slot],
                      d.n_locals, pmode & 1, (pmode >> 1) & 1, d.pool_len[slot], (pmode & 4) ? TAG_BOOL : TAG_INT};
    const EvalOut r = eval_program<kKind>(e, pv, st, scores, lane, staged);
    if (lane == 0) {
      // Ranking rounds every published fitness to a 40-bit mantissa.
      if (mode == 1) { d.c_fit[slot] = canon40(r.fitness); d.c_wrong[slot] = r.wrong; }
      else { d.fitness[slot] = canon40(r.fitness); d.wrong[slot] = r.wrong; }
      // Per-program tallies, summed once per launch by k_tally.
      d.e_errors[slot] = r.errors;
      d.e_timeouts[slot] = r.timeouts;
      d.e_fuel[slot] = r.fuel;
    }
  }
}

// The typed straight-line VM with one block per program: the block's four
// wavefronts take the program's 64-case tiles in turn (tile w, w + 4, ...),
// so a program waits for a quarter of the tile chain one wavefront would
// run. Each wavefront reads its tiles' inputs from the case arrays. The
// case scores go to the block's LDS; the totals of the four wavefronts are
// combined in LDS -- exactly, when every score is an integer of magnitude
// <= 2^36 (all partial sums are integers below 2^48) -- and otherwise the
// canonical replay reads the scores, as in eval_program. mode as k_eval.
__global__ __launch_bounds__(kEvalThreads) void k_eval_tiles(DSearch* all, const BlockMap* map, int mode) {
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  if (mode == 1 ? !search_live(d) : (!d.active || *d.error)) return;
  if (!d.typed_tiles) return;  // uniform per block
  extern __shared__ unsigned char smem[];
  double* const scores = reinterpret_cast<double*>(smem);  // [C]
  __shared__ double part_sum[kWavesPerBlock];
  __shared__ unsigned part_count[kWavesPerBlock][4];  // wrong, errors, timeouts, integral
  __shared__ unsigned long long part_fuel[kWavesPerBlock];
  const int wave = wave_index(), lane = threadIdx.x % 64;
  const EvalParams e{d.cd, d.fuel, d.lanes, d.penalty, d.policy};
  const int C = d.cd.C;
  const int count = mode == 1 ? d.n_pending[1] : d.P;
  for (int i = bm.block; i < count; i += bm.blocks) {  // one program per block, uniform
    const int slot = mode == 1 ? d.pending[pending_at(d.B, 1, i)] : i;
    const int pmode = d.code_jumps[slot];
    if (mode == 0 && eval_class(pmode, C) != 1) continue;
    const ProgView pv{d.code + std::size_t(slot) * d.code_cap, d.pool + std::size_t(slot) * d.pool_cap, d.code_len[slot],
                      d.n_locals, pmode & 1, (pmode >> 1) & 1, d.pool_len[slot], (pmode & 4) ? TAG_BOOL : TAG_INT};
    const FastProgram fp = fast_load(pv, lane);
    double isum = 0;
    unsigned w = 0, errs = 0, touts = 0;
    unsigned long long fuel = 0;
    bool integral = true;  // C <= kScoreCases <= 4096
    for (int tile = wave * 64; tile < C; tile += kWavesPerBlock * 64) {
      const int c = tile + lane;
      const bool active = c < C;
      const LaneResult r = vm_fast(fp, fast_inputs(e.cd, pv.n_locals, c, active), active, e.fuel, pv.result_tag);
      if (active) {
        bool exact = false;
        const double sc = lane_score(e, r, expected_of(e.cd, c), exact);
        w += !exact;
        errs += r.err != ERR_NONE;
        touts += r.err == ERR_TIMEOUT;
        fuel += unsigned(r.fuel_used);
        integral = integral && sc == double(std::int64_t(sc)) && sc >= -68719476736.0 && sc <= 68719476736.0;
        isum += sc;
        scores[c] = sc;
      }
    }
    const double ws = wave_sum_d(isum);
    const unsigned ww = wave_sum_u(w), we = wave_sum_u(errs), wt = wave_sum_u(touts);
    for (int off = 32; off > 0; off >>= 1) fuel += __shfl_xor(fuel, off, 64);
    const bool wint = __ballot(!integral) == 0;
    if (lane == 0) {
      part_sum[wave] = ws;
      part_count[wave][0] = ww;
      part_count[wave][1] = we;
      part_count[wave][2] = wt;
      part_count[wave][3] = wint;
      part_fuel[wave] = fuel;
    }
    __syncthreads();  // every wavefront's scores and totals
    if (wave == 0) {
      double total = 0;
      unsigned tw = 0, te = 0, tt = 0;
      unsigned long long tf = 0;
      bool all_integral = true;
      for (int k = 0; k < kWavesPerBlock; ++k) {
        total += part_sum[k];
        tw += part_count[k][0];
        te += part_count[k][1];
        tt += part_count[k][2];
        all_integral = all_integral && part_count[k][3] != 0;
        tf += part_fuel[k];
      }
      const double fitness = all_integral ? total : canonical_from_scores(e, C, scores, lane);
      if (lane == 0) {
        // Ranking rounds every published fitness to a 40-bit mantissa.
        if (mode == 1) { d.c_fit[slot] = canon40(fitness); d.c_wrong[slot] = tw; }
        else { d.fitness[slot] = canon40(fitness); d.wrong[slot] = tw; }
        d.e_errors[slot] = te;
        d.e_timeouts[slot] = tt;
        d.e_fuel[slot] = tf;
      }
    }
    __syncthreads();  // before the next program reuses the scores
  }
}

// ---- counters of one evaluation launch and, when profiling, of the
// reproduction that fed it (mode as k_eval) ----
constexpr int kTallyBlocks = 16;  // blocks per search, one atomic per counter each
__global__ __launch_bounds__(256) void k_tally(DSearch* all, int mode) {
  const DSearch& d = all[blockIdx.y];
  if (mode == 1 ? !search_live(d) : (!d.active || *d.error)) return;
  const int front = mode == 1 ? d.n_pending[0] : d.P, back = mode == 1 ? d.n_pending[1] : 0;
  const int count = mode == 1 ? front + back + d.n_pending[2] : d.P;
  const int first = int(blockIdx.x) * 256 + int(threadIdx.x), stride = kTallyBlocks * 256;
  unsigned long long v[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = first; i < count; i += stride) {
    const int slot = mode != 1        ? i
                     : i < front        ? d.pending[i]
                     : i < front + back ? d.pending[d.B - 1 - (i - front)]
                                        : d.pending[d.B + (i - front - back)];
    v[0] += d.e_errors[slot];
    v[1] += d.e_timeouts[slot];
    v[2] += d.e_fuel[slot];
    v[3] += unsigned(d.code_len[slot]);
    v[9] += unsigned(d.pool_len[slot]);
  }
  if (mode == 1 && d.profile)
    for (int c = first; c < d.B; c += stride) {
      const OpStats st = d.stats[c];
      v[4] += unsigned(d.x_n[c]);
      v[5] += unsigned(d.x_nc[c]);
      if (st.parent >= 0) { v[6] += unsigned(d.cur_n[st.parent]); v[7] += unsigned(d.cur_nc[st.parent]); }
      if (st.crossover_attempted && st.donor >= 0) { v[6] += unsigned(d.cur_n[st.donor]); v[7] += unsigned(d.cur_nc[st.donor]); }
      v[8] += unsigned(d.code_len[c]);
      v[10] += unsigned(d.pool_len[c]);
    }
  __shared__ unsigned long long part[11][256];
  for (int k = 0; k < 11; ++k) part[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int span = 128; span > 0; span >>= 1) {
    if (int(threadIdx.x) < span)
      for (int k = 0; k < 11; ++k) part[k][threadIdx.x] += part[k][threadIdx.x + span];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    Counters* c = d.counters;
    if (blockIdx.x == 0) {
      atomicAdd(&c->programs, (unsigned long long)count);
      atomicAdd(&c->program_cases, (unsigned long long)count * unsigned(d.cd.C));
      if (mode == 1) atomicAdd(&c->children, (unsigned long long)d.B);
    }
    atomicAdd(&c->errors, part[0][0]);
    atomicAdd(&c->timeouts, part[1][0]);
    atomicAdd(&c->vm_instructions, part[2][0]);
    atomicAdd(&c->code_words, part[3][0]);
    atomicAdd(&c->pool_words, part[9][0]);
    if (mode == 1 && d.profile) {
      atomicAdd(&c->child_nodes, part[4][0]);
      atomicAdd(&c->child_consts, part[5][0]);
      atomicAdd(&c->parent_nodes, part[6][0]);
      atomicAdd(&c->parent_consts, part[7][0]);
      atomicAdd(&c->child_code_words, part[8][0]);
      atomicAdd(&c->child_pool_words, part[10][0]);
    }
  }
}

// ---- std::mt19937_64 on the device, generated 312 outputs at a time ----
__device__ inline std::uint64_t mt_temper(std::uint64_t y) {
  y ^= (y >> 29) & 0x5555555555555555ULL;
  y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
  y ^= (y << 37) & 0xFFF7EEE000000000ULL;
  return y ^ (y >> 43);
}
__device__ inline std::uint64_t mt_next_word(std::uint64_t cur, std::uint64_t nxt, std::uint64_t far) {
  const std::uint64_t x = (cur & 0xFFFFFFFF80000000ULL) | (nxt & 0x7FFFFFFFULL);
  return far ^ (x >> 1) ^ ((x & 1ULL) ? 0xB5026F5AA96619E9ULL : 0ULL);
}
// In-place twist by a block of kThreads >= 64 threads: element i < 156 reads
// only old words; 156 <= i < 311 reads the new word i-156; 311 reads new
// words 0 and 155. The thread count is a template constant so the loops do
// not re-read the dispatch packet.
template <int kThreads>
__device__ inline void mt_twist(std::uint64_t* mt) {
  static_assert(kThreads >= 64, "three words per thread at most");
  const int t = threadIdx.x;
  std::uint64_t v[3];
  int n = 0;
  for (int i = t; i < 156; i += kThreads) v[n++] = mt_next_word(mt[i], mt[i + 1], mt[i + 156]);
  __syncthreads();
  n = 0;
  for (int i = t; i < 156; i += kThreads) mt[i] = v[n++];
  __syncthreads();
  n = 0;
  for (int i = 156 + t; i < 311; i += kThreads) v[n++] = mt_next_word(mt[i], mt[i + 1], mt[i - 156]);
  __syncthreads();
  n = 0;
  for (int i = 156 + t; i < 311; i += kThreads) mt[i] = v[n++];
  __syncthreads();
  if (t == 0) mt[311] = mt_next_word(mt[311], mt[0], mt[155]);
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_seed(DSearch* all) {
  const DSearch& d = all[blockIdx.x];
  if (!d.active) return;
  __shared__ std::uint64_t mt[312];
  if (threadIdx.x == 0) {
    mt[0] = d.seed;
    for (int i = 1; i < 312; ++i) mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + std::uint64_t(i);
  }
  __syncthreads();
  mt_twist<256>(mt);
  for (int i = threadIdx.x; i < 312; i += 256) d.mt[i] = mt[i];
  // Word 0 seeds round 0; words 1..311 wait for the first placement.
  for (int i = 1 + threadIdx.x; i < 312; i += 256) d.mt_left[i - 1] = mt_temper(mt[i]);
  if (threadIdx.x == 0) {
    *d.round_seed = mt_temper(mt[0]);
    *d.n_left = 311;
    *d.round = 0;
  }
}

// Node counts, best member, placements reset (after the initial evaluation).
__global__ void k_init_keys(DSearch* all) {
  const DSearch& d = all[blockIdx.x];
  if (!d.active || *d.error) return;
  for (int m = threadIdx.x; m < d.P; m += blockDim.x) {
    const CRow r = member_row(d, m);
    unsigned n = d.grammar_sites ? 0u : unsigned(d.fixed_nodes);
    if (d.grammar_sites) { for (int i = 0; i < r.n; ++i) n += r.site[i] >= 0; }
    else n += unsigned(r.n);
    d.nodes[m] = n;
    d.member_hash[m] = row_hash(r);
  }
  __syncthreads();
  __shared__ int best[256];
  const Keys keys{d.fitness, d.wrong, d.nodes, d.has_wrong};
  int b = -1;
  for (int m = threadIdx.x; m < d.P; m += blockDim.x)
    if (b < 0 || better(keys, m, b)) b = m;
  best[threadIdx.x] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    int x = -1;
    for (int t = 0; t < int(blockDim.x); ++t)
      if (best[t] >= 0 && (x < 0 || better(keys, best[t], x))) x = best[t];
    *d.best = x;
  }
}

// ---- history (evolve.cpp:364-372) ----
__global__ void k_record(DSearch* all) {
  const DSearch& d = all[blockIdx.x];
  if (!search_live(d)) return;
  const int r = *d.round;
  if (r % d.per_epoch) return;
  const int gen = r / d.per_epoch;
  const int b = *d.best;
  // Early exit: the state after `gen` generations is exactly that of a run
  // with `gen` generations, whose history has no row `gen`.
  if (d.early_exit && gen > 0 && d.has_wrong && d.wrong[b] == 0) {
    if (threadIdx.x == 0) *d.stop = gen;
    return;
  }
  const CRow row = member_row(d, b);
  const std::size_t o = std::size_t(gen) * d.R;
  for (int i = threadIdx.x; i < row.n; i += blockDim.x) d.hist_origin[o + i] = row.origin[i];
  for (int i = threadIdx.x; i < row.nc; i += blockDim.x) d.hist_cval[o + i] = row.cval[i];
  for (int m = threadIdx.x; m < d.P; m += blockDim.x) d.hist_all[std::size_t(gen) * d.P + m] = d.fitness[m];
  if (threadIdx.x == 0) {
    d.hist_n[gen] = row.n;
    d.hist_nc[gen] = row.nc;
    d.hist_fit[gen] = d.fitness[b];
    d.hist_wrong[gen] = d.wrong[b];
  }
}

// Round, early-stop generation and error of every search of the table.
__global__ void k_status(DSearch* all, int n, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DSearch& d = all[i];
  out[3 * i + 0] = *d.round;
  out[3 * i + 1] = *d.stop;
  out[3 * i + 2] = *d.error;
}

// ---- placement (evolve.cpp:378-422) ----
// The reference draws victims one group of `replacement_size` words at a
// time. Here one wavefront examines up to W consecutive groups at once
// against the keys as they stand at the start of the chunk, then commits the
// longest prefix of groups whose outcome cannot differ from the sequential
// order. A chunk ends before the first group that drew a member replaced by
// an earlier group of the chunk, after the first group whose child becomes
// the new best (the redraw test and the best tracking change there), and
// after the group that places the last child. Group 0 is always committed.
constexpr int kRing = 1024;        // tempered generator words, a power of two
constexpr int kPlaceWords = 512;   // words examined per chunk at most

template <bool kLds>
struct PlaceKeys {
  // kLds: fitness and (wrong << 12 | nodes) staged in LDS; otherwise the
  // member arrays in HBM, read past the L1 so every lane sees the stores of
  // the previous chunk. Fields are copied out of the search state so that
  // stores to the keys never force it to be re-read.
  GAS double* fitness;
  GAS unsigned* wrong;
  GAS unsigned* nodes;
  int has_wrong;
  double* f;
  unsigned* wn;
  __device__ double fit(int m) const {
    if (kLds) return f[m];
    const unsigned long long bits = __hip_atomic_load((GAS unsigned long long*)(fitness + m), __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)bits);
  }
  __device__ std::uint64_t order(int m) const {
    if (kLds) return wn[m];
    const unsigned w = __hip_atomic_load(wrong + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned n = __hip_atomic_load(nodes + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return std::uint64_t(w) << 32 | n;
  }
  __device__ std::uint64_t pack(unsigned w, unsigned n) const { return kLds ? std::uint64_t(w << 12 | n) : std::uint64_t(w) << 32 | n; }
  // `better` (fitness desc, wrong asc, nodes asc, index asc) of a key against member b.
  __device__ bool key_better(double fa, std::uint64_t oa, int a, int b) const {
    const double fb = fit(b);
    if (fa != fb) return fa > fb;
    if (has_wrong) { const std::uint64_t ob = order(b); if (oa != ob) return oa < ob; }
    return a < b;
  }
  __device__ bool better(int a, int b) const { return key_better(fit(a), has_wrong ? order(a) : 0, a, b); }
  __device__ void store(int m, double fv, unsigned w, unsigned n) const {
    if (kLds) { f[m] = fv; wn[m] = w << 12 | n; }
    fitness[m] = fv;
    wrong[m] = w;
    nodes[m] = n;
  }
};

// The block's four wavefronts stage the keys and refill the generator ring
// together; wavefront 0 alone runs the chunks (their LDS ordering needs only
// workgroup fences, not barriers). Children's keys are prefetched into two
// 64-child register windows one window ahead of use.
constexpr int kPlaceThreads = 256;
template <bool kLds>
__global__ __launch_bounds__(kPlaceThreads) void k_place(DSearch* all) {
  const DSearch& d = all[blockIdx.x];
  if (!search_live(d) || d.generational || d.place_lds != int(kLds)) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char place_smem[];
  __shared__ std::uint64_t ring[kRing];
  __shared__ unsigned ridx[kRing];
  // Generator state, double-buffered: a twist reads one buffer and writes
  // the other (`cur` is the buffer holding the current state).
  __shared__ std::uint64_t mt[2][312];
  __shared__ int s_head, s_b, s_best;
  const int t = threadIdx.x, lane = t & 63;
  const bool leader = wave_index() == 0;
  const int P = d.P, B = d.B, rs = d.replacement_size, has_wrong = d.has_wrong;
  const std::uint64_t P64 = std::uint64_t(P);
  GAS const double* const c_fit = d.c_fit;
  GAS const unsigned* const c_wrong = d.c_wrong;
  GAS const unsigned* const c_nodes = d.c_nodes;
  GAS int* const placed = d.placed;
  PlaceKeys<kLds> keys{d.fitness, d.wrong, d.nodes, has_wrong, reinterpret_cast<double*>(place_smem),
                       reinterpret_cast<unsigned*>(place_smem + (kLds ? 8 * std::size_t(P) : 0))};
  unsigned* mark = reinterpret_cast<unsigned*>(place_smem + (kLds ? 12 * std::size_t(P) : 0));
  // x mod P by a Barrett step: q = mulhi(x, floor((2^64-1)/P)) is
  // floor(x/P) or one less, so one conditional subtraction is exact.
  const std::uint64_t barrett = ~0ull / P64;
  const auto mod_p = [&](std::uint64_t x) {
    std::uint64_t r = x - __umul64hi(x, barrett) * P64;
    return unsigned(r >= P64 ? r - P64 : r);
  };
  constexpr int kStage = 8;  // loads in flight per thread while staging keys
  for (int m0 = 0; m0 < P; m0 += kPlaceThreads * kStage) {
    double f[kStage];
    unsigned w[kStage], n[kStage];
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
      const int m = m0 + k * kPlaceThreads + t;
      if (kLds && m < P) { f[k] = keys.fitness[m]; w[k] = keys.wrong[m]; n[k] = keys.nodes[m]; }
    }
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
      const int m = m0 + k * kPlaceThreads + t;
      if (m < P) {
        if (kLds) { keys.f[m] = f[k]; keys.wn[m] = has_wrong ? w[k] << 12 | n[k] : 0u; }
        mark[m] = 0;
      }
    }
  }
  int cur = 0;
  for (int i = t; i < 312; i += kPlaceThreads) mt[0][i] = d.mt[i];
  int head = 0, tail = *d.n_left;
  for (int i = t; i < tail; i += kPlaceThreads) {
    const std::uint64_t x = d.mt_left[i];
    ring[i] = x;
    ridx[i] = mod_p(x);
  }
  __syncthreads();
  // Every thread, uniformly: the next state into the other buffer and its
  // 312 tempered words at the ring's tail. Words 0..155 read only the old
  // state; 156..310 read the new word i-156, and 311 the new words 0 and 155.
  const auto twist = [&] {
    const int nxt = cur ^ 1;
    const auto put = [&](int i, std::uint64_t w) {
      mt[nxt][i] = w;
      const std::uint64_t x = mt_temper(w);
      ring[(tail + i) & (kRing - 1)] = x;
      ridx[(tail + i) & (kRing - 1)] = mod_p(x);
    };
    for (int i = t; i < 156; i += kPlaceThreads) put(i, mt_next_word(mt[cur][i], mt[cur][i + 1], mt[cur][i + 156]));
    __syncthreads();
    for (int i = 156 + t; i < 312; i += kPlaceThreads)
      put(i, mt_next_word(mt[cur][i], i == 311 ? mt[nxt][0] : mt[cur][i + 1], mt[nxt][i - 156]));
    __syncthreads();
    cur = nxt;
    tail += 312;
  };
  const auto refill = [&](int need) {
    while (tail - head < need) twist();
  };
  const int W = rs < kPlaceWords / 64 ? 64 : kPlaceWords / rs;
  const std::uint64_t below = (1ull << lane) - 1;
  int b = 0, best = *d.best;
  double best_f = keys.fit(best);
  std::uint64_t best_o = has_wrong ? keys.order(best) : 0;
  unsigned chunk = 0;
  const unsigned chunk_limit = 4u * unsigned(B) + 4096u;
  // Child-key windows [win, win+64) and [win+64, win+128), lane l holding child win+l.
  int win = 0;
  double f0 = 0, f1 = 0;
  unsigned w0 = 0, n0 = 0, w1 = 0, n1 = 0;
  if (leader) {
    if (lane < B) { f0 = c_fit[lane]; w0 = c_wrong[lane]; n0 = c_nodes[lane]; }
    if (64 + lane < B) { f1 = c_fit[64 + lane]; w1 = c_wrong[64 + lane]; n1 = c_nodes[64 + lane]; }
  }
  while (b < B) {
    // The ring is filled with whole twists (at least kRing - 311 unconsumed
    // words, above one chunk's W * rs <= kPlaceWords), and the leader runs
    // chunks until fewer words than a chunk's are left, so the block meets
    // at a barrier once per fill rather than once per chunk.
    while (tail - head + 312 <= kRing) twist();
    bool stop = false;
    while (leader && !stop && b < B && tail - head >= W * rs) {
      stop = ++chunk > chunk_limit;
      if (stop && lane == 0) raise(d, 500);
      while (!stop && b >= win + 64) {  // slide the windows
        f0 = f1; w0 = w1; n0 = n1;
        win += 64;
        const int next = win + 64 + lane;
        f1 = 0; w1 = 0; n1 = 0;
        if (next < B) { f1 = c_fit[next]; w1 = c_wrong[next]; n1 = c_nodes[next]; }
      }
      if (!stop) {
        const bool active = lane < W;
        const int base = head + lane * rs;
        int worst = -1;
        // The first four draws stay in registers for the conflict test
        // (named registers: a dynamically indexed array would spill).
        int d0 = 0, d1 = 0, d2 = 0, d3 = 0;
        if (active) {
          worst = int(ridx[base & (kRing - 1)]);
          d0 = worst;
          for (int i = 1; i < rs; ++i) {
            const int c = int(ridx[(base + i) & (kRing - 1)]);
            d1 = i == 1 ? c : d1;
            d2 = i == 2 ? c : d2;
            d3 = i == 3 ? c : d3;
            if (keys.better(worst, c)) worst = c;
          }
        }
        const bool place = active && worst != best;
        const std::uint64_t pmask = __ballot(place);
        const int rank = __popcll(pmask & below);
        const int child = b + rank;
        const bool fills = place && child < B;
        const int k = child - win;  // 0 <= k < 128
        // Both windows are read by every lane before the select: a shuffle
        // under a divergent branch reads 0 from the lanes that branch masks off.
        const double a0 = __shfl(f0, k & 63, 64), a1 = __shfl(f1, k & 63, 64);
        const int b0 = __shfl(int(w0), k & 63, 64), b1 = __shfl(int(w1), k & 63, 64);
        const int c0 = __shfl(int(n0), k & 63, 64), c1 = __shfl(int(n1), k & 63, 64);
        const double cf = k < 64 ? a0 : a1;
        const unsigned cw = unsigned(k < 64 ? b0 : b1);
        const unsigned cn = unsigned(k < 64 ? c0 : c1);
        if (place) atomicMax(&mark[worst], chunk * 64u + unsigned(63 - lane));
        // LDS operations of one wavefront complete in issue order, so the
        // mark reads below see every lane's maximum without a wait between
        // them; keys in HBM need the fence.
        if constexpr (!kLds) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        bool conflict = false;
        if (active)
          for (int i = 0; i < rs; ++i) {
            const int drawn = i == 0 ? d0 : i == 1 ? d1 : i == 2 ? d2 : i == 3 ? d3 : int(ridx[(base + i) & (kRing - 1)]);
            const unsigned m = mark[drawn];
            conflict = conflict || ((m >> 6) == chunk && int(63u - (m & 63u)) < lane);
          }
        // `better` of the child (at the victim's index) against the best
        // member, whose keys stay in registers: the best is never a victim,
        // so they change only when a child becomes the best.
        const std::uint64_t co = keys.pack(cw, cn);
        const bool raises = fills && (cf != best_f ? cf > best_f : has_wrong && co != best_o ? co < best_o : worst < best);
        int cut = W;
        const std::uint64_t last = __ballot(fills && child == B - 1);
        if (last) cut = min(cut, __ffsll((long long)last));
        const std::uint64_t clash = __ballot(active && conflict);
        if (clash) cut = min(cut, __ffsll((long long)clash) - 1);
        const std::uint64_t lifts = __ballot(raises);
        const int lifter = lifts ? __ffsll((long long)lifts) - 1 : 64;
        if (lifter < cut) cut = lifter + 1;
        if (fills && lane < cut) {
          keys.store(worst, cf, cw, cn);
          placed[worst] = child;
        }
        if (lifter < cut) {
          best = __shfl(worst, lifter, 64);
          best_f = __shfl(cf, lifter, 64);
          best_o = __shfl(co, lifter, 64);
        }
        b += __popcll(cut >= 64 ? pmask : pmask & ((1ull << cut) - 1));
        head += cut * rs;
        if constexpr (!kLds) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (leader && lane == 0) { s_head = head; s_b = stop ? B : b; s_best = best; }
    __syncthreads();
    head = s_head;
    b = s_b;
    best = s_best;
    __syncthreads();
  }
  refill(1);
  const std::uint64_t next_seed = ring[head & (kRing - 1)];
  ++head;
  const int left = tail - head;
  for (int i = t; i < left; i += kPlaceThreads) d.mt_left[i] = ring[(head + i) & (kRing - 1)];
  for (int i = t; i < 312; i += kPlaceThreads) d.mt[i] = mt[cur][i];
  if (t == 0) {
    *d.n_left = left;
    *d.round_seed = next_seed;
    *d.best = best;
    d.n_pending[0] = 0;
    d.n_pending[1] = 0;
    d.n_pending[2] = 0;
    atomicAdd(d.round, 1);
  }
}

// ---- generational schedule (evolve.cpp:253-322; phase_population_cpu.cpp:269-312) ----
// Elites (the top `elites` members by the ranking order, best first; the
// best alone when elitism is 0) and the generation's seed, one generator
// word drawn after the history row as in the reference's loop.
__global__ __launch_bounds__(256) void k_gen_begin(DSearch* all) {
  const DSearch& d = all[blockIdx.x];
  if (!d.generational || !search_live(d)) return;
  const int t = threadIdx.x, P = d.P, want = d.elites > 0 ? d.elites : 1;
  const Keys keys{d.fitness, d.wrong, d.nodes, d.has_wrong};
  __shared__ unsigned taken[16384 / 32];
  __shared__ int cand[256];
  __shared__ std::uint64_t mt[312];
  for (int i = t; i < (P + 31) / 32; i += 256) taken[i] = 0;
  __syncthreads();
  for (int k = 0; k < want; ++k) {
    int b = -1;
    for (int m = t; m < P; m += 256)
      if (!((taken[m >> 5] >> (m & 31)) & 1u) && (b < 0 || better(keys, m, b))) b = m;
    cand[t] = b;
    __syncthreads();
    for (int span = 128; span > 0; span >>= 1) {
      if (t < span && cand[t + span] >= 0 && (cand[t] < 0 || better(keys, cand[t + span], cand[t]))) cand[t] = cand[t + span];
      __syncthreads();
    }
    const int pick = cand[0];
    if (t == 0) {
      if (k < d.elites) d.elite[k] = pick;
      if (k == 0) *d.best = pick;
      taken[pick >> 5] |= 1u << (pick & 31);
    }
    __syncthreads();
  }
  // One word of the run generator: the oldest unconsumed word, refilled by a
  // twist when none is left. Word 0 seeds generation 0 (stored by k_seed).
  if (*d.round == 0) return;
  int left = *d.n_left;
  if (left == 0) {
    for (int i = t; i < 312; i += 256) mt[i] = d.mt[i];
    __syncthreads();
    mt_twist<256>(mt);
    for (int i = t; i < 312; i += 256) { d.mt[i] = mt[i]; d.mt_left[i] = mt_temper(mt[i]); }
    left = 312;
    __syncthreads();
  }
  const std::uint64_t word = d.mt_left[0];
  std::uint64_t keep[4];
  for (int j = 0; j < 4; ++j) {
    const int i = t + 256 * j + 1;
    keep[j] = i < left ? d.mt_left[i] : 0;
  }
  __syncthreads();
  for (int j = 0; j < 4; ++j) {
    const int i = t + 256 * j + 1;
    if (i < left) d.mt_left[i - 1] = keep[j];
  }
  if (t == 0) {
    *d.round_seed = word;
    *d.n_left = left - 1;
  }
}
// Every child becomes the member of its index (keys now; rows in k_commit).
__global__ void k_gen_place(DSearch* all, const BlockMap* map) {
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  const int m = bm.block * blockDim.x + threadIdx.x;
  if (!d.generational || !search_live(d) || m >= d.P) return;
  d.fitness[m] = d.c_fit[m];
  d.wrong[m] = d.c_wrong[m];
  d.nodes[m] = d.c_nodes[m];
  d.placed[m] = m;
}
__global__ void k_gen_end(DSearch* all) {
  const DSearch& d = all[blockIdx.x];
  if (!d.generational || !search_live(d) || threadIdx.x) return;
  d.n_pending[0] = 0;
  d.n_pending[1] = 0;
  d.n_pending[2] = 0;
  atomicAdd(d.round, 1);
}

// One wavefront per member: a placed child's row is copied across the lanes.
__global__ __launch_bounds__(64 * kVaryWaves) void k_commit(DSearch* all, const BlockMap* map) {
  const BlockMap bm = map[blockIdx.x];
  const DSearch& d = all[bm.search];
  const int m = __builtin_amdgcn_readfirstlane(bm.block * kVaryWaves + wave_index());
  if (!d.active || m >= d.P) return;
  const int c = d.placed[m];
  if (c >= 0) {
    const CRow src = child_row(d, c);
    const std::size_t o = std::size_t(m) * d.R;
    GAS int* const origin = d.cur_origin + o;
    GAS int* const site = d.cur_site + o;
    GAS Val* const cval = d.cur_cval + o;
    for (int i = wlane(); i < src.n; i += 64) { origin[i] = src.origin[i]; site[i] = src.site[i]; }
    for (int i = wlane(); i < src.nc; i += 64) cval[i] = src.cval[i];
  }
  if (wlane() == 0) {
    const std::uint64_t h = c >= 0 ? d.child_hash[c] : d.member_hash[m];
    if (c >= 0) {
      d.cur_n[m] = d.x_n[c];
      d.cur_nc[m] = d.x_nc[c];
      d.member_hash[m] = h;
      d.placed[m] = -1;
    }
    // The next round's member table (cleared during placement).
    if (search_live(d)) insert_member(d, m, h);
  }
}

// ---- evaluation session kernel (external bytecode) ----
struct DProg { GAS const Ins* code; GAS const Val* pool; int len, has_jumps, n_locals, fast, n_pool, result_tag; };
static_assert(sizeof(DProg) == 40, "program table entry size is part of the workspace bound");
__device__ inline ProgView view_of(const DProg& p) {
  return ProgView{p.code, p.pool, p.len, p.n_locals, p.has_jumps, p.fast, p.n_pool, p.result_tag};
}
// kFast: scores the programs certified for vm_fast (LDS: a score buffer of C
// doubles per wavefront); otherwise the rest (LDS: the general stacks).
template <bool kFast>
__global__ __launch_bounds__(kEvalThreads) void k_eval_programs(const DProg* progs, int n, EvalParams e, int slots, int temps,
                                                                 int shadow, double* fitness, unsigned* wrong, unsigned* counts) {
  extern __shared__ unsigned char smem[];
  const int wave = wave_index(), lane = threadIdx.x % 64;
  WaveStack st = stack_of(smem, kFast ? 0 : slots, kFast ? 0 : temps, wave, kFast ? 0 : shadow);
  double* scores = kFast ? reinterpret_cast<double*>(smem) + std::size_t(wave) * e.cd.C : nullptr;
  for (int i = blockIdx.x * kWavesPerBlock + wave; i < n; i += gridDim.x * kWavesPerBlock) {
    const DProg p = progs[i];
    if ((p.fast != 0) != kFast) continue;
    const EvalOut r = eval_program<kFast ? 1 : 0>(e, view_of(p), st, scores, lane);
    if (lane == 0) {
      fitness[i] = r.fitness;
      wrong[i] = e.policy.metric == METRIC_LEGACY ? 0 : r.wrong;
      counts[3 * i + 0] = unsigned(e.cd.C);
      counts[3 * i + 1] = r.errors;
      counts[3 * i + 2] = r.timeouts;
    }
  }
}
template <bool kFast>
__global__ __launch_bounds__(kEvalThreads) void k_eval_cases(DProg p, CaseData cd, int fuel, int slots, int temps, int shadow,
                                                              Val* values, int* errors) {
  extern __shared__ unsigned char smem[];
  const int wave = wave_index(), lane = threadIdx.x % 64;
  WaveStack st = stack_of(smem, kFast ? 0 : slots, kFast ? 0 : temps, wave, kFast ? 0 : shadow);
  const ProgView pv = view_of(p);
  FastProgram fp{};
  if (kFast) fp = fast_load(pv, lane);
  for (int tile = (blockIdx.x * kWavesPerBlock + wave) * 64; tile < cd.C; tile += gridDim.x * kWavesPerBlock * 64) {
    const int c = tile + lane;
    const bool active = c < cd.C;
    const LaneResult r = kFast ? vm_fast(fp, fast_inputs(cd, pv.n_locals, c, active), active, fuel, pv.result_tag)
                               : vm_run(pv, cd, active ? c : 0, active, fuel, st, lane);
    if (active) {
      Val v{};
      v.i = r.v; v.tag = r.t; v.b = r.b;
      values[c] = v;
      errors[c] = r.err;
    }
  }
}

// ---- device buffers ----
// Bump allocator over device chunks that double in size (1 MiB up to
// 256 MiB): a search's arrays take a handful of hipMalloc calls in all.
// Arrays are 256-byte aligned. A counting arena (`dry`) runs the same chunk
// arithmetic without touching the device, so the footprint of a search is
// known before any of its memory is allocated. The arena owns its chunks:
// it can be moved but not copied, and release() (or destruction) returns
// them to the device.
struct DeviceArena {
  std::vector<void*> blocks;
  std::size_t bytes = 0;     // bytes handed out
  std::size_t reserved = 0;  // bytes of device chunks held (the footprint)
  bool dry = false;
  char* cursor = nullptr;
  std::size_t left = 0, next_chunk = kFirstChunk;
  static constexpr std::size_t kFirstChunk = std::size_t(1) << 20;

  DeviceArena() = default;
  explicit DeviceArena(bool count_only) : dry(count_only) {}
  DeviceArena(const DeviceArena&) = delete;
  DeviceArena& operator=(const DeviceArena&) = delete;
  DeviceArena(DeviceArena&& o) noexcept { take(o); }
  DeviceArena& operator=(DeviceArena&& o) noexcept {
    if (this != &o) {
      release();
      take(o);
    }
    return *this;
  }
  ~DeviceArena() { release(); }

  void release() {
    for (void* p : blocks) (void)hipFree(p);
    blocks.clear();
    bytes = reserved = 0;
    cursor = nullptr;
    left = 0;
    next_chunk = kFirstChunk;
  }
  template <class T>
  GAS T* alloc(std::size_t n) {
    const std::size_t sz = (std::max<std::size_t>(n, 1) * sizeof(T) + 255) & ~std::size_t(255);
    if (sz > left) {
      const std::size_t chunk = std::max(sz, next_chunk);
      void* p = nullptr;
      if (!dry) {
        HIP_OK(hipMalloc(&p, chunk));
        blocks.push_back(p);
      }
      cursor = static_cast<char*>(p);
      left = chunk;
      reserved += chunk;
      next_chunk = std::min(next_chunk * 2, std::size_t(256) << 20);
    }
    void* p = cursor;
    if (!dry) cursor += sz;
    left -= sz;
    bytes += sz;
    return gbl(static_cast<T*>(p));
  }
  template <class T>
  GAS T* upload(const std::vector<T>& v) {
    GAS T* p = alloc<T>(v.size());
    if (!dry && !v.empty()) HIP_OK(hipMemcpy((void*)p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  }
  // Clears `n` elements (skipped when counting).
  template <class T>
  void zero(GAS T* p, std::size_t n) {
    if (!dry && n) HIP_OK(hipMemset((void*)p, 0, n * sizeof(T)));
  }

 private:
  void take(DeviceArena& o) {
    blocks = std::move(o.blocks);
    bytes = o.bytes;
    reserved = o.reserved;
    dry = o.dry;
    cursor = o.cursor;
    left = o.left;
    next_chunk = o.next_chunk;
    o.blocks.clear();
    o.bytes = o.reserved = 0;
    o.cursor = nullptr;
    o.left = 0;
    o.next_chunk = kFirstChunk;
  }
};

struct DeviceCases {
  CaseData cd{};
  void upload(DeviceArena& a, const Cases& c) {
    const int C = c.n, K = int(c.input_names.size());
    std::vector<std::int64_t> v(std::size_t(K) * C);
    std::vector<std::uint8_t> t(std::size_t(K) * C, TAG_INVALID), b(std::size_t(K) * C, 0);
    for (int i = 0; i < C; ++i)
      for (int k = 0; k < K; ++k) {
        const std::size_t src = std::size_t(i) * K + k, dst = std::size_t(k) * C + i;
        if (!c.set[src]) continue;
        v[dst] = c.inputs[src].i;
        t[dst] = c.inputs[src].tag;
        b[dst] = c.inputs[src].b;
      }
    std::vector<std::int64_t> ev(C);
    std::vector<std::uint8_t> et(C), eb(C);
    for (int i = 0; i < C; ++i) { ev[i] = c.expected[i].i; et[i] = c.expected[i].tag; eb[i] = c.expected[i].b; }
    std::vector<std::int64_t> arena = c.arena;
    if (arena.empty()) arena.push_back(0);
    cd.in_val = a.upload(v);
    cd.in_tag = a.upload(t);
    cd.in_b = a.upload(b);
    cd.exp_val = a.upload(ev);
    cd.exp_tag = a.upload(et);
    cd.exp_b = a.upload(eb);
    cd.arena = a.upload(arena);
    // Visiting order of the looping path: cases by descending total list
    // length (stable), so a tile's lanes run loops of similar length.
    const std::size_t n = std::size_t(C);
    std::vector<int> order(n);
    std::vector<std::int64_t> work(n, 0);
    for (int i = 0; i < C; ++i) {
      order[std::size_t(i)] = i;
      for (int k = 0; k < K; ++k) {
        const std::size_t src = std::size_t(i) * K + k;
        if (c.set[src] && c.inputs[src].tag == TAG_INT_LIST) work[std::size_t(i)] += std::int64_t(std::uint64_t(c.inputs[src].i) & 0xffff);
      }
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[std::size_t(x)] > work[std::size_t(y)]; });
    cd.order = a.upload(order);
    cd.C = C;
    cd.n_inputs = K;
  }
};

int next_pow2(int x) { int p = 1; while (p < x) p <<= 1; return p; }

}  // namespace


// ====================== host side ======================

// One admitted search: host documents, tables and the device state it owns.
struct SearchState {
  SearchSpec spec;
  Task task;
  Grammar grammar;
  Cases train;
  SearchConfig cfg;
  Layout layout;
  PhaseGrammar phase;
  int P = 0, B = 0, R = 0, G = 0, rounds = 0, per_epoch = 0;
  int stepped = 0;  // rounds the host has stepped this search (it idles past `rounds`)
  int fuel = 0;
  double penalty = 0;
  bool may_jump = false;   // programs may go to the looping path: if/and/or, traversals, statements, lets
  bool statement = false;  // a Program entry: statement programs with grammar locals
  // The export name table: the names the first initial member uses, in
  // declaration order (the layout prototype of phase_population_cpu.cpp:195),
  // or an imported first program's own table; and the first binder id of an
  // exported program (-1: the generated population's numbering).
  std::vector<std::string> layout_names;
  int binder_base = -1;
  // Rows of an imported initial population, held until they are uploaded.
  ImportedPopulation imported;
  bool has_import = false;
  DeviceArena arena;
  DSearch d{};
  double init_ms = 0;
  bool configured = false, retired = false;

  explicit SearchState(const SearchSpec& s) : spec(s) {
    task = load_task(s.task_path);
    grammar = load_grammar_file(task.grammar_path);
    train = load_cases(task.train_path);
    cfg = default_search_config();
    if (!s.search_path.empty()) {
      std::ifstream in(s.search_path);
      if (!in) throw std::invalid_argument("search config: cannot read " + s.search_path);
      std::ostringstream text;
      text << in.rdbuf();
      apply_search(cfg, json::parse(text.str()));
    }
    if (s.population) cfg.population = *s.population;
    if (s.generations) cfg.generations = *s.generations;
    if (s.seed) cfg.seed = *s.seed;
    if (!cfg.resident) throw Unsupported("native_ast representation (host reproduction)");
    if (cfg.site_depth < 1 && s.init_path.empty())
      throw Unsupported("site_depth 0 (initial population from the native generator)");
    if (cfg.population < 2 || cfg.population > 16384) throw Unsupported("population outside 2..16384");
    if (cfg.elitism < 0 || cfg.elitism >= cfg.population) throw std::invalid_argument("elite_count must be >= 0 and < population_size");
    P = cfg.population;
    B = cfg.schedule == Schedule::Generational ? P : (cfg.batch ? cfg.batch : P);
    if (B > P || P % B) throw std::invalid_argument("search config: steady_state batch must divide the population");
    G = cfg.generations;
    rounds = int((std::int64_t(P) * G) / B);
    per_epoch = P / B;
    penalty = cfg.penalty.value_or(task.case_cap);
    fuel = cfg.fuel.value_or(int(grammar.fuel));
    // Inputs must exactly cover the grammar inputs.
    {
      std::vector<std::string> want;
      for (const auto& in : grammar.inputs) want.push_back(in.name);
      std::sort(want.begin(), want.end());
      if (want != train.input_names) throw std::invalid_argument("every fitness case must supply every compiled grammar input");
      for (std::size_t k = 0; k < train.input_names.size(); ++k)
        for (const auto& in : grammar.inputs)
          if (in.name == train.input_names[k] && in.type != train.input_types[k])
            throw std::invalid_argument("case input type disagrees with the grammar");
    }
    layout = whole_program_layout(grammar);
    phase = compile_phase(grammar, grammar.entry, grammar.max_depth, unsigned(cfg.site_depth), train.input_names);
    R = std::min<int>(kPhaseCapacity, int(layout.max_variable_nodes));
    for (const auto& x : phase.exprs)
      may_jump = may_jump || x.node.kind == TK_IF || x.node.kind == TK_AND || x.node.kind == TK_OR || x.node.kind == TK_TRAVERSE ||
                 x.node.kind == TK_FOR_RANGE;
    statement = grammar.nts[grammar.entry].cat == Cat::Program;
    // Hidden locals (statements, lets) also make programs for the looping
    // path, with or without jumps, and so does straight-line code beyond the
    // straight-line VM's limits: one instruction per node plus RETURN, and a
    // stack of at most (arity - 1) pending operands per level plus one.
    int arity = 1;
    for (const auto& x : phase.exprs) arity = std::max(arity, x.node.arity);
    may_jump = may_jump || statement || phase.has_let || R + 1 > kFastCode ||
               (arity - 1) * int(layout.slot_max_depth) + 1 > kFastStack;
    if (!s.init_path.empty()) {
      imported = load_population(s.init_path, grammar, phase, unsigned(cfg.site_depth), R, std::size_t(P));
      has_import = true;
      layout_names = imported.layout_names;
      binder_base = imported.binder_base;
    }
  }

  // Allocates every device array of the search in its arena.
  void allocate() { d = build(arena); }
  // Device bytes the search holds once allocated (its arena's chunks),
  // computed without allocating.
  std::size_t planned_bytes() const {
    DeviceArena count(true);
    (void)build(count);
    return count.reserved;
  }

  DSearch build(DeviceArena& a) const {
    DSearch x{};
    x.active = 1;
    x.P = P; x.B = B; x.R = R;
    // A traversal lowers to at most 64 instructions and 8 constants of its
    // own besides those of its nodes, and keeps 9 temporaries live.
    const int walks = phase.traversals;
    // A for_range lowers to 23 instructions and 4 constants of its own, and
    // takes at least three nodes (itself, its bound and its body).
    const int loops = statement ? (R + 2) / 3 : 0;
    x.code_cap = 4 * R + 8 + 64 * walks + 23 * loops;
    // The general VM schedules instruction indices below 8191.
    if (x.code_cap > 8190) throw Unsupported("lowered programs longer than 8190 instructions");
    x.pool_cap = 2 * R + 8 + 8 * walks + 4 * loops;
    x.stack_slots = std::min(kMaxStack, 2 * int(layout.slot_max_depth) + 2);
    // Statement programs lower through the hidden-locals instantiation too.
    x.walks = walks > 0 || statement || phase.has_let;
    x.imported = has_import ? 1 : 0;
    x.temp_slots = statement || phase.has_let ? std::min(64 - int(train.input_names.size()), int(hidden_local_bound(grammar)))
                   : walks  ? std::max(0, std::min(64 - int(train.input_names.size()), 9 * std::min(walks, 8)))
                            : 0;
    x.input_shadow = phase.input_writes ? int(train.input_names.size()) : 0;
    x.temp_slots += x.input_shadow;
    x.rounds = rounds; x.per_epoch = per_epoch;
    x.fuel = fuel; x.lanes = 512; x.penalty = penalty; x.policy = task.policy;
    x.pressure = std::clamp(cfg.tournament, 1, P);
    x.replacement_size = cfg.replacement_size;
    x.p_mut = cfg.mutation_probability;
    x.pol = OpPolicy{cfg.counted, cfg.uniform_phase, cfg.crossover_probability, cfg.crossover_attempts, cfg.mutation_max_nodes};
    x.slot = SlotInfo{phase.root, 0, int(layout.slot_max_depth), cfg.site_depth, int(layout.max_variable_nodes), R, statement ? 1 : 0};
    x.fixed_nodes = int(layout.fixed_nodes);
    x.grammar_sites = cfg.grammar_sites;
    x.has_wrong = task.policy.metric != METRIC_LEGACY;
    x.expected_type = statement ? kStatementType : grammar.nts[grammar.entry].type;
    x.n_locals = int(train.input_names.size());
    x.site_depth = cfg.site_depth;
    x.seed = cfg.seed;
    x.g = PhaseTables{a.upload(phase.exprs), int(phase.exprs.size()), a.upload(phase.prods), int(phase.prods.size()),
                    a.upload(phase.nts), int(phase.nts.size()), a.upload(phase.domains), int(phase.domains.size()),
                    a.upload(phase.values), int(phase.values.size()), a.upload(phase.minimum), phase.depth_count,
                    phase.logical_count, phase.site_depth ? 1 : 0, phase.root};
    x.local_types = a.upload(train.input_types);
    DeviceCases dc;
    dc.upload(a, train);
    x.cd = dc.cd;
    if (x.cd.C <= kScoreCases) {
      x.loop_scores = a.alloc<double>(std::size_t(std::max(P, B)) * std::size_t(x.cd.C));
      x.loop_acc = a.upload(std::vector<LoopAcc>(std::size_t(std::max(P, B))));
    }
    const std::size_t rows = std::size_t(P) * R, crows = std::size_t(B) * R;
    x.cur_origin = a.alloc<int>(rows); x.cur_site = a.alloc<int>(rows); x.cur_cval = a.alloc<Val>(rows);
    x.cur_n = a.alloc<int>(P); x.cur_nc = a.alloc<int>(P);
    x.x_origin = a.alloc<int>(crows); x.x_site = a.alloc<int>(crows); x.x_cval = a.alloc<Val>(crows);
    x.x_n = a.alloc<int>(B); x.x_nc = a.alloc<int>(B);
    x.t_origin = a.alloc<int>(crows); x.t_site = a.alloc<int>(crows); x.t_cval = a.alloc<Val>(crows);
    x.s_origin = a.alloc<int>(crows); x.s_site = a.alloc<int>(crows); x.s_cval = a.alloc<Val>(crows);
    x.fitness = a.alloc<double>(P); x.wrong = a.alloc<unsigned>(P); x.nodes = a.alloc<unsigned>(P);
    x.stats = a.alloc<OpStats>(B); x.reuse = a.alloc<int>(B);
    x.c_fit = a.alloc<double>(B); x.c_wrong = a.alloc<unsigned>(B); x.c_nodes = a.alloc<unsigned>(B);
    const int code_rows = std::max(P, B);
    x.code = a.alloc<Ins>(std::size_t(code_rows) * x.code_cap);
    x.pool = a.alloc<Val>(std::size_t(code_rows) * x.pool_cap);
    x.code_len = a.alloc<int>(code_rows); x.pool_len = a.alloc<int>(code_rows); x.code_jumps = a.alloc<int>(code_rows);
    x.child_hash = a.alloc<std::uint64_t>(B);
    x.pending = a.alloc<int>(2 * std::size_t(B)); x.n_pending = a.alloc<int>(3);
    const int tsize = next_pow2(2 * P);
    x.table_key = a.alloc<std::uint64_t>(tsize); x.table = a.alloc<int>(tsize); x.table_mask = tsize - 1;
    x.mt = a.alloc<std::uint64_t>(312); x.mt_left = a.alloc<std::uint64_t>(kRing); x.n_left = a.alloc<int>(1);
    x.round_seed = a.alloc<std::uint64_t>(1);
    x.member_hash = a.alloc<std::uint64_t>(P);
    // Placement keys (fitness, wrong << 12 | nodes) and victim marks in LDS.
    x.place_lds = P <= 8192 && train.n < (1 << 20) && int(layout.fixed_nodes) + R < 4096;
    x.round = a.alloc<int>(1); x.best = a.alloc<int>(1); x.placed = a.alloc<int>(P);
    x.hist_origin = a.alloc<int>(std::size_t(G) * R); x.hist_cval = a.alloc<Val>(std::size_t(G) * R);
    x.hist_n = a.alloc<int>(G); x.hist_nc = a.alloc<int>(G);
    x.hist_fit = a.alloc<double>(G); x.hist_wrong = a.alloc<unsigned>(G);
    x.hist_all = a.alloc<double>(std::size_t(G) * P);
    x.counters = a.alloc<Counters>(1);
    x.generational = cfg.schedule == Schedule::Generational;
    x.elites = x.generational ? cfg.elitism : 0;
    x.elite = a.alloc<int>(x.generational ? std::size_t(P) : 1);  // any elite count below P
    x.early_exit = spec.early_exit ? 1 : 0;
    x.stop = a.alloc<int>(1);
    x.e_errors = a.alloc<unsigned>(code_rows); x.e_timeouts = a.alloc<unsigned>(code_rows);
    x.e_fuel = a.alloc<unsigned long long>(code_rows);
    x.error = a.alloc<int>(1);
    a.zero(x.n_pending, 3);
    a.zero(x.counters, 1);
    a.zero(x.error, 1);
    a.zero(x.round, 1);
    a.zero(x.stop, 1);
    return x;
  }

  template <class T>
  std::vector<T> down(GAS const T* p, std::size_t n) const {
    std::vector<T> v(n);
    if (n) HIP_OK(hipMemcpy(v.data(), (const void*)p, n * sizeof(T), hipMemcpyDeviceToHost));
    return v;
  }
  int device_error() const { return down(d.error, 1)[0]; }
};

struct KernelTimer {
  bool on = false;
  std::map<std::string, double> ms;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  template <class F>
  void time(const char* name, hipStream_t s, F&& launch) {
    if (!on) { launch(); return; }
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    HIP_OK(hipEventRecord(a, s));
    launch();
    HIP_OK(hipEventRecord(b, s));
    pending.push_back({name, {a, b}});
  }
  void collect() {
    for (auto& p : pending) {
      HIP_OK(hipEventSynchronize(p.second.second));
      float t = 0;
      HIP_OK(hipEventElapsedTime(&t, p.second.first, p.second.second));
      ms[p.first] += t;
      (void)hipEventDestroy(p.second.first);
      (void)hipEventDestroy(p.second.second);
    }
    pending.clear();
  }
};

// ---- evaluation launch autotuner ----
// Launch shapes of the evaluation kernels that leave every result unchanged:
// pending programs per wavefront of the typed straight-line kernel (its
// grid, and so how many blocks stage the case set), and tiles per work item
// of the looping kernel (cases per item = 64 * loop_group).
struct TuneChoice {
  int fast_programs = 1;
  int loop_group = 1;
};
// Choices measured so far in this process, by batch shape.
std::mutex tune_mutex;
std::map<std::string, TuneChoice> tune_cache;

struct Batch::Impl {
  std::vector<std::unique_ptr<SearchState>> searches;  // every search ever added (stable indices)
  std::vector<int> live;                                // searches of the round table, in table order
  bool dirty = true;                                    // table, grids and graph need rebuilding
  int* d_status = nullptr;                              // k_status output, 3 ints per table entry
  std::size_t status_cap = 0;
  DSearch* d_all = nullptr;
  hipStream_t stream = nullptr;
  KernelTimer timer;
  int max_P = 0, max_B = 0, max_slots = 0, max_table = 0, max_rounds = 0;
  std::size_t place_lds_bytes = 0, place_hbm_bytes = 0;  // dynamic LDS of the two placement variants
  double run_ms = 0, init_ms = 0;
  bool graph_ready = false;
  hipGraphExec_t graph_exec = nullptr;
  // Side streams of a round's parallel branches (default priority, like
  // `stream`) and the events that fork them from `stream` and join them
  // back; captured, they become parallel branches of the round graph.
  static constexpr int kSides = 2;
  hipStream_t side[kSides] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kSides] = {};
  ~Impl() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (d_all) (void)hipFree(d_all);
    if (d_status) (void)hipFree(d_status);
    for (int i = 0; i < kSides; ++i) {
      if (side[i]) (void)hipStreamDestroy(side[i]);
      if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void make_streams() {
    if (!stream) HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    if (fork_ev) return;
    HIP_OK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    for (int i = 0; i < kSides; ++i) {
      HIP_OK(hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking));
      HIP_OK(hipEventCreateWithFlags(&join_ev[i], hipEventDisableTiming));
    }
  }
  // Runs parts[0] on `stream` and parts[1..] on side streams forked from it,
  // then joins them: the parts must touch disjoint data. With the kernel
  // timer on, every part runs on `stream` in order, so each launch's events
  // bracket that launch alone.
  void branches(const std::vector<std::function<void(hipStream_t)>>& parts) {
    if (timer.on) {
      for (const auto& part : parts) part(stream);
      return;
    }
    HIP_OK(hipEventRecord(fork_ev, stream));
    for (std::size_t i = 1; i < parts.size(); ++i) HIP_OK(hipStreamWaitEvent(side[i - 1], fork_ev, 0));
    parts[0](stream);
    for (std::size_t i = 1; i < parts.size(); ++i) {
      parts[i](side[i - 1]);
      HIP_OK(hipEventRecord(join_ev[i - 1], side[i - 1]));
      HIP_OK(hipStreamWaitEvent(stream, join_ev[i - 1], 0));
    }
  }
  std::size_t eval_fast_lds = 0, eval_general_lds = 0, eval_tiles_lds = 0;  // dynamic LDS of the evaluation kernels
  static constexpr std::size_t kLoopLds = std::size_t(kWavesPerBlock) * kLoopSlots * 64 * 8;  // the looping slots
  std::size_t eval_loop_lds = 0;
  std::size_t vary_lds_bytes = 0;  // phase-table copy of the staged reproduction kernel
  bool vary_global = false;        // some search reproduces from global tables
  bool any_generational = false;

  // Flat grids: every search contributes exactly its own blocks.
  struct Grid { GAS BlockMap* map = nullptr; int blocks = 0; };
  DeviceArena map_arena;
  Grid g_init, g_lower, g_eval_init, g_vary_staged, g_vary_global, g_vary_staged_w, g_vary_global_w, g_table, g_member,
      g_eval, g_eval_tiles, g_eval_general, g_eval_loop, g_member_w;
  // Evaluation launch shapes of the round (the autotuner's choice, or the
  // fixed default: one program per wavefront, one 64-case tile per item).
  TuneChoice tune;
  int autotune = 1;  // 0 off, 1 above kTuneWork, 2 always
  bool seam = false;  // rounds driven through the owner seam
  bool tuned = false;
  std::string tune_report;  // JSON of the last tuning pass
  bool any_walks(const std::vector<int>& ids) const {
    for (int id : ids)
      if (searches[std::size_t(id)]->d.walks) return true;
    return false;
  }
  // Some grammar can lower to jumps (if/and/or or traversals): the looping
  // evaluation kernel has work.
  bool any_jumps(const std::vector<int>& ids) const {
    for (int id : ids)
      if (searches[std::size_t(id)]->may_jump) return true;
    return false;
  }
  bool loop_kernel = false;  // any_jumps(live)
  template <class Blocks>
  Grid build_grid(DeviceArena& arena, const std::vector<int>& table, Blocks blocks_of) {
    std::vector<BlockMap> v;
    for (std::size_t i = 0; i < table.size(); ++i) {
      const int nb = blocks_of(*searches[std::size_t(table[i])]);
      for (int j = 0; j < nb; ++j) v.push_back(BlockMap{int(i), j, nb});
    }
    Grid g;
    g.blocks = int(v.size());
    if (!v.empty()) g.map = arena.upload(v);
    return g;
  }
  static int up(int x, int k) { return (x + k - 1) / k; }
  void build_round_grids() {
    map_arena.release();
    const auto vary = [&](bool staged, bool walks) {
      return build_grid(map_arena, live, [&](const SearchState& s) {
        return (s.d.stage_tables != 0) == staged && (s.d.walks != 0) == walks ? up(s.B, kVaryWaves) : 0;
      });
    };
    g_vary_staged = vary(true, false);
    g_vary_global = vary(false, false);
    g_vary_staged_w = vary(true, true);
    g_vary_global_w = vary(false, true);
    g_table = build_grid(map_arena, live, [&](const SearchState& s) { return up(s.d.table_mask + 1, 256); });
    g_member = build_grid(map_arena, live, [&](const SearchState& s) { return up(s.P, 128); });
    // one wavefront per child / member
    g_member_w = build_grid(map_arena, live, [&](const SearchState& s) { return up(s.P, kVaryWaves); });
    g_eval = typed_grid(map_arena, tune.fast_programs);
    g_eval_tiles = tiles_grid(map_arena, live, false);
    g_eval_general = eval_grid(map_arena, 1);
    g_eval_loop = loop_grid(map_arena, tune.loop_group);
  }
  // Evaluation grids of the live searches: `per_wave` pending programs per
  // wavefront of the typed and general kernels; `group` tiles per work item
  // of the looping kernel.
  Grid eval_grid(DeviceArena& arena, int per_wave) {
    return build_grid(arena, live, [&](const SearchState& s) { return std::min(4096, up(s.B, kWavesPerBlock * per_wave)); });
  }
  // The typed kernels: k_eval<1> for searches whose cases fit one tile,
  // k_eval_tiles (one block per program) for the rest.
  Grid typed_grid(DeviceArena& arena, int per_wave) {
    return build_grid(arena, live, [&](const SearchState& s) {
      return s.d.typed_tiles ? 0 : std::min(4096, up(s.B, kWavesPerBlock * per_wave));
    });
  }
  Grid tiles_grid(DeviceArena& arena, const std::vector<int>& table, bool members) {
    return build_grid(arena, table, [&](const SearchState& s) { return s.d.typed_tiles ? std::min(4096, members ? s.P : s.B) : 0; });
  }
  Grid loop_grid(DeviceArena& arena, int group) {
    return build_grid(arena, live, [&](const SearchState& s) { return s.may_jump ? loop_blocks(s.d.cd.C, s.B, group) : 0; });
  }

  // Allocates one search and fixes its kernel variants and LDS needs.
  void configure(SearchState& s) {
    constexpr std::size_t kLdsLimit = 160 * 1024, kStageLimit = 64 * 1024;
    s.configured = true;
    s.allocate();
    s.d.profile = timer.on ? 1 : 0;
    s.d.stage_tables = phase_tables_bytes(s.d.g) <= kStageLimit;
    const std::size_t stacks = std::size_t(kWavesPerBlock) * (s.d.stack_slots + s.d.temp_slots) * 64 * 10;
    const std::size_t score_buf = std::size_t(kWavesPerBlock) * 8 * s.d.cd.C;
    s.d.general_scores = s.d.cd.C <= kScoreCases && stacks + score_buf <= kLdsLimit;
    const std::size_t cases = staged_case_bytes(s.d.cd.C, staged_inputs_of(s.d.n_locals, s.d.cd.n_inputs));
    s.d.stage_cases = s.d.cd.C <= kScoreCases && cases <= 48 * 1024 && score_buf + cases + 16 <= kLdsLimit;
    s.d.typed_tiles = s.d.cd.C > 64 && s.d.cd.C <= kScoreCases;
  }
  // LDS sizes and kernel attributes for a table of searches.
  void size_kernels(const std::vector<int>& table) {
    eval_fast_lds = eval_general_lds = eval_tiles_lds = vary_lds_bytes = place_lds_bytes = place_hbm_bytes = 0;
    eval_loop_lds = kLoopLds;
    vary_global = any_generational = false;
    max_P = max_B = max_slots = max_table = 0;
    for (int id : table) {
      const SearchState& x = *searches[std::size_t(id)];
      const std::size_t stacks = std::size_t(kWavesPerBlock) * (x.d.stack_slots + x.d.temp_slots) * 64 * 10;
      const std::size_t score_buf = std::size_t(kWavesPerBlock) * 8 * x.d.cd.C;
      eval_general_lds = std::max(eval_general_lds, stacks + (x.d.general_scores ? score_buf : 0));
      const std::size_t cases = x.d.stage_cases ? staged_case_bytes(x.d.cd.C, staged_inputs_of(x.d.n_locals, x.d.cd.n_inputs)) : 0;
      if (x.d.cd.C <= kScoreCases) {
        eval_fast_lds = std::max(eval_fast_lds, score_buf + cases);
        eval_loop_lds = kLoopLds;
      }
      if (x.d.typed_tiles) eval_tiles_lds = std::max(eval_tiles_lds, std::size_t(8) * x.d.cd.C);
      if (x.d.stage_tables) vary_lds_bytes = std::max(vary_lds_bytes, phase_tables_bytes(x.d.g));
      else vary_global = true;
      if (x.d.place_lds) place_lds_bytes = std::max(place_lds_bytes, std::size_t(16) * x.P);
      else place_hbm_bytes = std::max(place_hbm_bytes, std::size_t(4) * x.P);
      any_generational = any_generational || x.d.generational;
      max_P = std::max(max_P, x.P);
      max_B = std::max(max_B, x.B);
      max_slots = std::max(max_slots, x.d.stack_slots);
      max_table = std::max(max_table, x.d.table_mask + 1);
    }
    const auto attr = [](const void* k, std::size_t bytes) {
      if (bytes > 65536) HIP_OK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
    };
    attr(reinterpret_cast<const void*>(&k_vary<true, false>), vary_lds_bytes);
    attr(reinterpret_cast<const void*>(&k_vary<true, true>), vary_lds_bytes);
    attr(reinterpret_cast<const void*>(&k_eval<0>), eval_general_lds);
    attr(reinterpret_cast<const void*>(&k_eval<1>), eval_fast_lds);
    attr(reinterpret_cast<const void*>(&k_eval_tiles), eval_tiles_lds);
    attr(reinterpret_cast<const void*>(&k_eval_loop<false>), eval_loop_lds);
    attr(reinterpret_cast<const void*>(&k_eval_loop<true>), eval_loop_lds);
    attr(reinterpret_cast<const void*>(&k_place<true>), place_lds_bytes);
    attr(reinterpret_cast<const void*>(&k_place<false>), place_hbm_bytes);
  }
  // Builds and scores the initial populations of `ids` (not yet in the
  // round table) on the stream, through a table of their own.
  void init_searches(const std::vector<int>& ids) {
    if (ids.empty()) return;
    std::vector<DSearch> host;
    for (int id : ids) host.push_back(searches[std::size_t(id)]->d);
    const int n = int(ids.size());
    DeviceArena tmp;
    GAS DSearch* table = tmp.upload(host);
    DSearch* t = (DSearch*)table;
    size_kernels(ids);
    const Grid gi = build_grid(tmp, ids, [&](const SearchState& x) { return up(x.P, kVaryWaves); });
    const Grid gl = build_grid(tmp, ids, [&](const SearchState& x) { return up(x.P, kVaryWaves); });
    const Grid ge = build_grid(tmp, ids, [&](const SearchState& x) { return std::min(4096, up(x.P, kWavesPerBlock)); });
    // Imported populations: their rows go up in place of k_init_pop's
    // (which skips those searches); the staging vectors live until the
    // stream is synchronized below.
    std::vector<std::vector<int>> up_i;
    std::vector<std::vector<Val>> up_v;
    for (int id : ids) {
      SearchState& x = *searches[std::size_t(id)];
      if (!x.has_import) continue;
      const std::size_t P = std::size_t(x.P), R = std::size_t(x.R);
      if (x.imported.rows.size() != P) throw std::logic_error("imported population size");
      std::vector<int> origin(P * R, 0), site(P * R, -1), n(P), nc(P), placed(P, -1);
      std::vector<Val> cval(P * R, Val{});
      for (std::size_t m = 0; m < P; ++m) {
        const PhaseRow& r = x.imported.rows[m];
        if (r.origin.size() > R || r.cval.size() > R) throw std::logic_error("imported row capacity");
        std::copy(r.origin.begin(), r.origin.end(), origin.begin() + std::ptrdiff_t(m * R));
        std::copy(r.site.begin(), r.site.end(), site.begin() + std::ptrdiff_t(m * R));
        std::copy(r.cval.begin(), r.cval.end(), cval.begin() + std::ptrdiff_t(m * R));
        n[m] = int(r.origin.size());
        nc[m] = int(r.cval.size());
      }
      const auto put = [&](GAS void* to, const void* from, std::size_t bytes) {
        HIP_OK(hipMemcpyAsync((void*)to, from, bytes, hipMemcpyHostToDevice, stream));
      };
      put(x.d.cur_origin, origin.data(), origin.size() * sizeof(int));
      put(x.d.cur_site, site.data(), site.size() * sizeof(int));
      put(x.d.cur_cval, cval.data(), cval.size() * sizeof(Val));
      put(x.d.cur_n, n.data(), n.size() * sizeof(int));
      put(x.d.cur_nc, nc.data(), nc.size() * sizeof(int));
      put(x.d.placed, placed.data(), placed.size() * sizeof(int));
      for (auto* v : {&origin, &site, &n, &nc, &placed}) up_i.push_back(std::move(*v));
      up_v.push_back(std::move(cval));
    }
    timer.time("k_init_pop", stream,
               [&] { hipLaunchKernelGGL(k_init_pop, dim3(gi.blocks), dim3(64 * kVaryWaves), 0, stream, t, gi.map); });
    timer.time("k_lower_members", stream,
               [&] { hipLaunchKernelGGL(k_lower_members<false>, dim3(gl.blocks), dim3(64 * kVaryWaves), 0, stream, t, gl.map); });
    if (any_walks(ids))
      timer.time("k_lower_members_walks", stream,
                 [&] { hipLaunchKernelGGL(k_lower_members<true>, dim3(gl.blocks), dim3(64 * kVaryWaves), 0, stream, t, gl.map); });
    timer.time("k_eval_init", stream,
               [&] { hipLaunchKernelGGL(k_eval<1>, dim3(ge.blocks), dim3(kEvalThreads), eval_fast_lds, stream, t, ge.map, 0); });
    const Grid gt = tiles_grid(tmp, ids, true);
    if (gt.blocks)
      timer.time("k_eval_init_tiles", stream, [&] {
        hipLaunchKernelGGL(k_eval_tiles, dim3(gt.blocks), dim3(kEvalThreads), eval_tiles_lds, stream, t, gt.map, 0);
      });
    if (any_jumps(ids))
      timer.time("k_eval_init_loop", stream, [&] {
        const Grid gp = build_grid(tmp, ids, [&](const SearchState& x) { return loop_blocks(x.d.cd.C, x.P); });
        hipLaunchKernelGGL(k_eval_loop<false>, dim3(gp.blocks), dim3(kEvalThreads), eval_loop_lds, stream, t, gp.map, 0, 1);
      });
    timer.time("k_eval_init_general", stream, [&] {
      hipLaunchKernelGGL(k_eval<0>, dim3(ge.blocks), dim3(kEvalThreads), eval_general_lds, stream, t, ge.map, 0);
    });
    timer.time("k_tally_init", stream, [&] { hipLaunchKernelGGL(k_tally, dim3(kTallyBlocks, n), dim3(256), 0, stream, t, 0); });
    timer.time("k_init_keys", stream, [&] { hipLaunchKernelGGL(k_init_keys, dim3(n), dim3(256), 0, stream, t); });
    timer.time("k_seed", stream, [&] { hipLaunchKernelGGL(k_seed, dim3(n), dim3(256), 0, stream, t); });
    // The first round's member table (later rounds' are built by k_commit).
    const Grid gt_clear = build_grid(tmp, ids, [&](const SearchState& x) { return up(x.d.table_mask + 1, 256); });
    const Grid gt_index = build_grid(tmp, ids, [&](const SearchState& x) { return up(x.P, 128); });
    timer.time("k_clear_init", stream, [&] { hipLaunchKernelGGL(k_clear, dim3(gt_clear.blocks), dim3(256), 0, stream, t, gt_clear.map); });
    timer.time("k_index_init", stream, [&] { hipLaunchKernelGGL(k_index, dim3(gt_index.blocks), dim3(128), 0, stream, t, gt_index.map); });
    HIP_OK(hipStreamSynchronize(stream));
    timer.collect();
    for (int id : ids) {
      SearchState& x = *searches[std::size_t(id)];
      if (const int e = x.device_error()) throw std::runtime_error("device initialization error " + std::to_string(e));
      if (x.has_import) {
        x.imported.rows.clear();
        x.imported.rows.shrink_to_fit();
        continue;  // the first imported program's names are the layout's
      }
      const int n0 = x.down(x.d.cur_n, 1)[0];
      x.layout_names = names_used(x.grammar, x.down(x.d.cur_origin, std::size_t(n0)));
    }
  }
  // Round table, grids, kernel sizes and graph for the current live set.
  // Batch shape for the tuning cache: per kernel, log2 buckets of the
  // searches' programs per round and cases, summed over the live searches.
  std::string shape_key() const {
    const auto bucket = [](double x) { return x <= 1 ? 0 : int(std::floor(std::log2(x))); };
    double fast_programs = 0, fast_cases = 0, loop_items = 0, loop_cases = 0;
    int loops = 0;
    for (int id : live) {
      const SearchState& x = *searches[std::size_t(id)];
      fast_programs += x.B;
      fast_cases += double(x.B) * x.d.cd.C;
      if (x.may_jump) {
        loop_items += double(x.B) * ((x.d.cd.C + 63) / 64);
        loop_cases += double(x.B) * x.d.cd.C;
        ++loops;
      }
    }
    std::ostringstream k;
    k << "typed:" << bucket(fast_programs) << "/" << bucket(fast_cases / std::max(1.0, fast_programs)) << " loop:" << loops << "/"
      << bucket(loop_items) << "/" << bucket(loop_cases / std::max(1.0, loop_items * 64) * 64);
    return k.str();
  }
  // Evaluation work of one round (program-cases over the live searches):
  // below kTuneWork, or with fewer than kTuneRounds rounds left to run,
  // the tuning pass would cost more than it can gain, and the fixed shapes
  // run.
  static constexpr double kTuneWork = double(1 << 22);
  static constexpr int kTuneRounds = 128;
  int rounds_left() const {
    int left = 0;
    for (int id : live) left = std::max(left, searches[std::size_t(id)]->rounds - searches[std::size_t(id)]->stepped);
    return left;
  }
  double round_work() const {
    double w = 0;
    for (int id : live) w += double(searches[std::size_t(id)]->B) * searches[std::size_t(id)]->d.cd.C;
    return w;
  }
  // Measures the evaluation kernels over the current round's pending
  // programs under each candidate launch shape and keeps the fastest.
  // Re-running an evaluation kernel over the same pending list rewrites the
  // same results (each program's fitness, wrong count and tallies; the
  // looping kernel's per-program totals are cleared by the program's last
  // item), so tuning changes no search. Call between enqueue_front and
  // enqueue_eval of a round, with the stream idle.
  void tune_launches() {
    const std::string key = shape_key();
    if (autotune == 1 && (round_work() < kTuneWork || rounds_left() < kTuneRounds)) {
      apply_tuning(TuneChoice{});
      std::ostringstream r;
      r << "{\"key\":\"" << key << "\",\"cached\":false,\"skipped\":\"fewer than " << kTuneWork
        << " program-cases per round or " << kTuneRounds << " rounds left\",\"fast_programs\":1,\"loop_group\":1}";
      tune_report = r.str();
      return;
    }
    {
      std::lock_guard<std::mutex> lock(tune_mutex);
      auto it = tune_cache.find(key);
      if (it != tune_cache.end()) {
        apply_tuning(it->second);
        std::ostringstream r;
        r << "{\"key\":\"" << key << "\",\"cached\":true,\"fast_programs\":" << tune.fast_programs
          << ",\"loop_group\":" << tune.loop_group << "}";
        tune_report = r.str();
        return;
      }
    }
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    // The first candidate of each kernel runs twice (its first launch warms
    // code and caches), every other candidate once.
    int reps = 2;
    const auto measure = [&](auto launch) {
      float best = 1e30f;
      for (int rep = 0; rep < reps; ++rep) {
        HIP_OK(hipEventRecord(a, stream));
        launch();
        HIP_OK(hipEventRecord(b, stream));
        HIP_OK(hipEventSynchronize(b));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        best = std::min(best, ms);
      }
      return best;
    };
    std::ostringstream r;
    r << "{\"key\":\"" << key << "\",\"cached\":false";
    TuneChoice choice;
    DeviceArena tmp;
    // A candidate replaces the default only when at least 3% faster; the
    // candidates grow the work per wavefront, so once one is 15% slower
    // than the best the larger ones are not tried.
    const auto pick = [&](const char* name, const std::vector<int>& values, auto time_of) {
      r << ",\"" << name << "_ms\":{";
      int chosen = values[0];
      float chosen_ms = 0;
      for (std::size_t i = 0; i < values.size(); ++i) {
        reps = i == 0 ? 2 : 1;
        const float ms = time_of(values[i]);
        r << (i ? "," : "") << "\"" << values[i] << "\":" << ms;
        if (i == 0) chosen_ms = ms;
        else if (ms < 0.97f * chosen_ms) { chosen = values[i]; chosen_ms = ms; }
        else if (ms > 1.15f * chosen_ms) break;
      }
      r << "}";
      return chosen;
    };
    choice.fast_programs = pick("typed", {1, 2, 4}, [&](int k) {
      tmp.release();
      const Grid g = typed_grid(tmp, k);
      return g.blocks ? measure([&] {
        hipLaunchKernelGGL(k_eval<1>, dim3(g.blocks), dim3(kEvalThreads), eval_fast_lds, stream, d_all, g.map, 1);
      }) : 0.f;
    });
    if (loop_kernel) {
      choice.loop_group = pick("loop", {1, 2, 4}, [&](int group) {
        tmp.release();
        const Grid g = loop_grid(tmp, group);
        return g.blocks ? measure([&] {
          if (group == 1)
            hipLaunchKernelGGL(k_eval_loop<false>, dim3(g.blocks), dim3(kEvalThreads), eval_loop_lds, stream, d_all, g.map, 1, 1);
          else
            hipLaunchKernelGGL(k_eval_loop<true>, dim3(g.blocks), dim3(kEvalThreads), eval_loop_lds, stream, d_all, g.map, 1, group);
        }) : 0.f;
      });
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(stream));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    r << ",\"fast_programs\":" << choice.fast_programs << ",\"loop_group\":" << choice.loop_group << "}";
    tune_report = r.str();
    {
      std::lock_guard<std::mutex> lock(tune_mutex);
      tune_cache[key] = choice;
    }
    apply_tuning(choice);
  }
  // New launch shapes: grids rebuilt, the round graph captured again.
  void apply_tuning(const TuneChoice& c) {
    tune = c;
    tuned = true;
    if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
    graph_ready = false;
    build_round_grids();
  }

  static constexpr int kCaptureAfter = 8;
  int table_rounds = 0;  // rounds stepped since the round table was formed
  void rebuild() {
    if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
    graph_ready = false;
    tuned = false;
    table_rounds = 0;
    if (d_all) { (void)hipFree(d_all); d_all = nullptr; }
    max_rounds = 0;
    if (!live.empty()) {
      std::vector<DSearch> host;
      for (int id : live) {
        host.push_back(searches[std::size_t(id)]->d);
        max_rounds = std::max(max_rounds, searches[std::size_t(id)]->rounds);
      }
      HIP_OK(hipMalloc(&d_all, sizeof(DSearch) * host.size()));
      HIP_OK(hipMemcpy(d_all, host.data(), sizeof(DSearch) * host.size(), hipMemcpyHostToDevice));
      size_kernels(live);
      build_round_grids();
      loop_kernel = any_jumps(live);
      if (status_cap < live.size()) {
        if (d_status) (void)hipFree(d_status);
        status_cap = std::max<std::size_t>(64, 2 * live.size());
        HIP_OK(hipMalloc(&d_status, sizeof(int) * 3 * status_cap));
      }
    }
    dirty = false;
  }

  template <class Kernel, class... Args>
  void launch_on(hipStream_t s, const char* name, const Grid& g, Kernel kernel, int threads, std::size_t lds, Args... args) {
    if (g.blocks)
      timer.time(name, s, [&] { hipLaunchKernelGGL(kernel, dim3(g.blocks), dim3(threads), lds, s, d_all, g.map, args...); });
  }
  template <class Kernel, class... Args>
  void launch(const char* name, const Grid& g, Kernel kernel, int threads, std::size_t lds, Args... args) {
    launch_on(stream, name, g, kernel, threads, lds, args...);
  }
  void enqueue_round() {
    enqueue_front();
    enqueue_eval();
    enqueue_back();
  }
  // A round in three parts: reproduction up to the pending lists, the
  // evaluation kernels, then tallies, placement and commit.
  void enqueue_front() {
    const int n = int(live.size());
    // Caller-driven rounds (the owner seam) bring their own elites, seed and
    // history.
    if (any_generational && !seam)
      timer.time("k_gen_begin", stream, [&] { hipLaunchKernelGGL(k_gen_begin, dim3(n), dim3(256), 0, stream, d_all); });
    if (!seam) timer.time("k_record", stream, [&] { hipLaunchKernelGGL(k_record, dim3(n), dim3(256), 0, stream, d_all); });
    // Three independent branches: the reproduction of searches without and
    // with traversal walks (disjoint searches) and the genotype index of the
    // current members (read by nothing reproduction writes); the pending
    // lists need all three.
    branches({[&](hipStream_t s) {
                launch_on(s, "k_vary", g_vary_staged, k_vary<true, false>, 64 * kVaryWaves, vary_lds_bytes);
                launch_on(s, "k_vary_global", g_vary_global, k_vary<false, false>, 64 * kVaryWaves, 0);
              },
              [&](hipStream_t s) {
                launch_on(s, "k_vary_walks", g_vary_staged_w, k_vary<true, true>, 64 * kVaryWaves, vary_lds_bytes);
                launch_on(s, "k_vary_global_walks", g_vary_global_w, k_vary<false, true>, 64 * kVaryWaves, 0);
              }});
  }
  void enqueue_eval() {
    // The three evaluation kernels score disjoint pending programs.
    branches({[&](hipStream_t s) {
                if (loop_kernel && tune.loop_group == 1)
                  launch_on(s, "k_eval_loop", g_eval_loop, k_eval_loop<false>, kEvalThreads, eval_loop_lds, 1, 1);
                else if (loop_kernel)
                  launch_on(s, "k_eval_loop", g_eval_loop, k_eval_loop<true>, kEvalThreads, eval_loop_lds, 1, tune.loop_group);
              },
              [&](hipStream_t s) {
                launch_on(s, "k_eval", g_eval, k_eval<1>, kEvalThreads, eval_fast_lds, 1);
                launch_on(s, "k_eval_tiles", g_eval_tiles, k_eval_tiles, kEvalThreads, eval_tiles_lds, 1);
              },
              [&](hipStream_t s) { launch_on(s, "k_eval_general", g_eval_general, k_eval<0>, kEvalThreads, eval_general_lds, 1); }});
  }
  void enqueue_back() {
    const int n = int(live.size());
    timer.time("k_tally", stream, [&] { hipLaunchKernelGGL(k_tally, dim3(kTallyBlocks, n), dim3(256), 0, stream, d_all, 1); });
    // Placement and the clearing of the member table k_commit refills are
    // independent.
    branches({[&](hipStream_t s) {
                if (place_lds_bytes)
                  timer.time("k_place", s, [&] {
                    hipLaunchKernelGGL(k_place<true>, dim3(n), dim3(kPlaceThreads), place_lds_bytes, s, d_all);
                  });
                if (place_hbm_bytes)
                  timer.time("k_place_hbm", s, [&] {
                    hipLaunchKernelGGL(k_place<false>, dim3(n), dim3(kPlaceThreads), place_hbm_bytes, s, d_all);
                  });
              },
              [&](hipStream_t s) { launch_on(s, "k_clear", g_table, k_clear, 256, 0); }});
    if (any_generational) launch("k_gen_place", g_member, k_gen_place, 128, 0);
    launch("k_commit", g_member_w, k_commit, 64 * kVaryWaves, 0);
    if (any_generational) timer.time("k_gen_end", stream, [&] { hipLaunchKernelGGL(k_gen_end, dim3(n), dim3(64), 0, stream, d_all); });
  }
};

Batch::Batch() : impl_(std::make_unique<Impl>()) {}
Batch::~Batch() = default;
std::size_t Batch::size() const { return impl_->searches.size(); }
void Batch::set_profile(bool on) { impl_->timer.on = on; }
void Batch::set_autotune(int mode) {
  if (mode < 0 || mode > 2) throw std::invalid_argument("autotune mode must be 0, 1 or 2");
  impl_->autotune = mode;
  if (!mode && impl_->tuned) impl_->apply_tuning(TuneChoice{});
  impl_->tuned = false;
}
std::string Batch::tuning_report() const { return impl_->tune_report.empty() ? "{}" : impl_->tune_report; }
int Batch::max_rounds() const { return impl_->max_rounds; }
std::size_t Batch::device_bytes() const {
  std::size_t b = 0;
  for (const auto& s : impl_->searches) b += s->arena.reserved;
  return b;
}
std::vector<std::pair<std::string, double>> Batch::kernel_times() const {
  return {impl_->timer.ms.begin(), impl_->timer.ms.end()};
}

int Batch::add(const SearchSpec& spec) {
  auto s = std::make_unique<SearchState>(spec);
  impl_->searches.push_back(std::move(s));
  return int(impl_->searches.size() - 1);
}

void Batch::prepare() {
  auto& I = *impl_;
  const auto t0 = std::chrono::steady_clock::now();
  I.make_streams();
  std::vector<int> fresh;
  for (std::size_t i = 0; i < I.searches.size(); ++i) {
    if (I.searches[i]->configured) continue;
    I.configure(*I.searches[i]);
    fresh.push_back(int(i));
  }
  I.init_searches(fresh);
  I.live.insert(I.live.end(), fresh.begin(), fresh.end());
  I.rebuild();
  I.init_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

std::size_t Batch::planned_footprint(int index) const { return impl_->searches.at(std::size_t(index))->planned_bytes(); }

void Batch::admit(int index) {
  auto& I = *impl_;
  SearchState& s = *I.searches.at(std::size_t(index));
  if (s.configured || s.retired) throw std::invalid_argument("search already admitted");
  I.make_streams();
  try {
    I.configure(s);
    I.init_searches({index});
  } catch (...) {
    // Undo the partial admission: nothing of the search stays on the
    // device, and the round table never saw it.
    (void)hipStreamSynchronize(I.stream);
    s.arena.release();
    s.d.active = 0;
    s.retired = true;
    throw;
  }
  I.live.push_back(index);
  I.dirty = true;
}

std::vector<std::pair<int, std::string>> Batch::admit_many(const std::vector<int>& indices) {
  auto& I = *impl_;
  std::vector<std::pair<int, std::string>> failed;
  std::vector<int> ready;
  I.make_streams();
  const auto undo = [&](SearchState& s) {
    s.arena.release();
    s.d.active = 0;
    s.retired = true;
  };
  for (int index : indices) {
    SearchState& s = *I.searches.at(std::size_t(index));
    if (s.configured || s.retired) throw std::invalid_argument("search already admitted");
    try {
      I.configure(s);  // allocation: a failure here is this search's alone
      ready.push_back(index);
    } catch (const std::exception& e) {
      (void)hipStreamSynchronize(I.stream);
      undo(s);
      failed.push_back({index, e.what()});
    }
  }
  try {
    I.init_searches(ready);  // one launch sequence for all of them
  } catch (const std::exception& e) {
    (void)hipStreamSynchronize(I.stream);
    for (int index : ready) {
      undo(*I.searches[std::size_t(index)]);
      failed.push_back({index, e.what()});
    }
    ready.clear();
  }
  for (int index : ready) I.live.push_back(index);
  if (!ready.empty()) I.dirty = true;
  return failed;
}

void Batch::drop(int index) {
  SearchState& s = *impl_->searches.at(std::size_t(index));
  if (s.configured && !s.retired) throw std::invalid_argument("search is admitted: retire it");
  s.arena.release();
  s.retired = true;
}

void Batch::retire(int index) {
  auto& I = *impl_;
  const auto it = std::find(I.live.begin(), I.live.end(), index);
  if (it == I.live.end()) throw std::invalid_argument("search is not in the batch");
  I.live.erase(it);
  I.dirty = true;
  SearchState& s = *I.searches.at(std::size_t(index));
  s.arena.release();  // the stream is idle between steps: the memory is free to return
  s.d.active = 0;
  s.retired = true;
}

std::vector<Batch::Status> Batch::status() {
  auto& I = *impl_;
  if (I.dirty) I.rebuild();
  std::vector<Status> out;
  if (I.live.empty()) return out;
  const int n = int(I.live.size());
  hipLaunchKernelGGL(k_status, dim3((n + 63) / 64), dim3(64), 0, I.stream, I.d_all, n, I.d_status);
  HIP_OK(hipGetLastError());
  std::vector<int> v(std::size_t(3 * n));
  HIP_OK(hipMemcpyAsync(v.data(), I.d_status, sizeof(int) * v.size(), hipMemcpyDeviceToHost, I.stream));
  HIP_OK(hipStreamSynchronize(I.stream));
  for (int i = 0; i < n; ++i) {
    const SearchState& x = *I.searches[std::size_t(I.live[std::size_t(i)])];
    Status st{I.live[std::size_t(i)], v[std::size_t(3 * i)], v[std::size_t(3 * i + 1)], v[std::size_t(3 * i + 2)], false};
    st.done = st.error != 0 || st.stopped_at != 0 || st.round >= x.rounds;
    out.push_back(st);
  }
  return out;
}

std::size_t Batch::footprint(int index) const { return impl_->searches.at(std::size_t(index))->arena.reserved; }

void Batch::step(int rounds, bool use_graph) {
  auto& I = *impl_;
  if (I.dirty) I.rebuild();
  if (I.live.empty() || rounds <= 0) return;
  // Every live search has been stepped `rounds` more when this returns (the
  // tuning rule below reads the count from before this step).
  struct Stepped {
    Impl& impl;
    int n;
    ~Stepped() {
      for (int id : impl.live) impl.searches[std::size_t(id)]->stepped += n;
    }
  } stepped{I, rounds};
  const auto t0 = std::chrono::steady_clock::now();
  if (I.autotune && !I.tuned) {
    // The first round of a new table: its pending programs are measured
    // under every candidate shape between reproduction and placement.
    I.enqueue_front();
    HIP_OK(hipStreamSynchronize(I.stream));
    I.tune_launches();
    I.enqueue_eval();
    I.enqueue_back();
    HIP_OK(hipStreamSynchronize(I.stream));
    I.timer.collect();
    if (--rounds == 0) {
      I.run_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      return;
    }
  }
  // A new round table runs on stream launches until it has stepped
  // kCaptureAfter rounds: capturing and instantiating a graph costs more
  // than the launches of a table that changes every few rounds (serving).
  const bool replay = I.graph_ready || I.table_rounds + rounds >= Impl::kCaptureAfter;
  I.table_rounds += rounds;
  if (use_graph && !I.timer.on && replay) {
    if (!I.graph_ready) {
      hipGraph_t graph;
      HIP_OK(hipStreamBeginCapture(I.stream, hipStreamCaptureModeThreadLocal));
      I.enqueue_round();
      HIP_OK(hipStreamEndCapture(I.stream, &graph));
      HIP_OK(hipGraphInstantiate(&I.graph_exec, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
      I.graph_ready = true;
    }
    for (int r = 0; r < rounds; ++r) HIP_OK(hipGraphLaunch(I.graph_exec, I.stream));
    HIP_OK(hipStreamSynchronize(I.stream));
  } else {
    for (int r = 0; r < rounds; ++r) {
      I.enqueue_round();
      if (I.timer.on && (r % 8 == 7)) I.timer.collect();
    }
    HIP_OK(hipStreamSynchronize(I.stream));
    I.timer.collect();
  }
  I.run_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Batch::finish() {
  for (int id : impl_->live) {
    if (const int e = impl_->searches[std::size_t(id)]->device_error()) throw std::runtime_error("device search error " + std::to_string(e));
  }
}

void Batch::run(bool use_graph) {
  prepare();
  step(impl_->max_rounds, use_graph);
  finish();
}

std::vector<std::uint64_t> Batch::counters(int index) const {
  const SearchState& s = *impl_->searches.at(std::size_t(index));
  const Counters c = s.down(s.d.counters, 1)[0];
  return {c.programs,     c.program_cases, c.errors,       c.timeouts,      c.vm_instructions,  c.code_words,
          c.pool_words,   c.children,      c.child_nodes,  c.child_consts,  c.parent_nodes,     c.parent_consts,
          c.child_code_words, c.child_pool_words};
}
double Batch::init_ms() const { return impl_->init_ms; }

// ---- the owner seam (caller-driven steady-state rounds) ----
struct Owner::Impl {
  Batch batch;
  SearchState* s = nullptr;
  double* select = nullptr;  // device selection values, P doubles
  bool made = false;         // a batch is made and not yet committed
  bool scored = false;
  bool reproduced = false;   // the members are the children of reproduce()
  std::vector<int> reuse;
  ~Impl() {
    if (select) (void)hipFree(select);
  }
};

Owner::Owner(const SearchSpec& spec) : impl_(std::make_unique<Impl>()) {
  Impl& O = *impl_;
  O.batch.add(spec);
  O.s = O.batch.impl_->searches[0].get();
  O.batch.set_autotune(0);
  O.batch.impl_->seam = true;
  O.batch.prepare();
  HIP_OK(hipMalloc(&O.select, sizeof(double) * std::size_t(O.s->P)));
  O.s->d.select = (GAS const double*)O.select;
  O.batch.impl_->rebuild();  // the round table with the selection array
  if (const int e = O.s->device_error()) throw std::runtime_error("device initialization error " + std::to_string(e));
}
Owner::~Owner() = default;
int Owner::population_size() const { return impl_->s->P; }
int Owner::batch_size() const { return impl_->s->B; }
bool Owner::has_wrong() const { return impl_->s->d.has_wrong != 0; }

namespace {
// {program cases, errors, timeouts, fallbacks, unscored} of evaluated slots.
void slot_counts(const SearchState& s, int n, const std::vector<int>* reuse, std::uint64_t* counts) {
  const std::vector<unsigned> errors = s.down(s.d.e_errors, std::size_t(n)), timeouts = s.down(s.d.e_timeouts, std::size_t(n));
  for (int i = 0; i < n; ++i) {
    std::uint64_t* c = counts + 5 * std::size_t(i);
    const bool scored = !reuse || (*reuse)[std::size_t(i)] < 0;
    c[0] = scored ? std::uint64_t(s.d.cd.C) : 0;
    c[1] = scored ? errors[std::size_t(i)] : 0;
    c[2] = scored ? timeouts[std::size_t(i)] : 0;
    c[3] = 0;
    c[4] = 0;
  }
}
// Node or grammar-site counts of n rows (counts, site rows of stride R).
void row_counts(const SearchState& s, bool grammar_sites, GAS const int* n_rows, GAS const int* sites, int n, unsigned* out) {
  const std::vector<int> len = s.down(n_rows, std::size_t(n));
  if (!grammar_sites) {
    for (int i = 0; i < n; ++i) out[i] = unsigned(s.d.fixed_nodes) + unsigned(len[std::size_t(i)]);
    return;
  }
  const std::vector<int> site = s.down(sites, std::size_t(n) * std::size_t(s.R));
  for (int i = 0; i < n; ++i) {
    unsigned k = 0;
    for (int j = 0; j < len[std::size_t(i)]; ++j) k += site[std::size_t(i) * std::size_t(s.R) + std::size_t(j)] >= 0;
    out[i] = k;
  }
}
}  // namespace

void Owner::evaluate(double* fitness, unsigned* wrong, std::uint64_t* counts) {
  const SearchState& s = *impl_->s;
  const std::vector<double> f = s.down(s.d.fitness, std::size_t(s.P));
  const std::vector<unsigned> w = s.down(s.d.wrong, std::size_t(s.P));
  std::copy(f.begin(), f.end(), fitness);
  std::copy(w.begin(), w.end(), wrong);
  // After reproduce(), member m is child m: reused children were not scored.
  slot_counts(s, s.P, impl_->reproduced ? &impl_->reuse : nullptr, counts);
}

void Owner::reproduce(const double* selection, int pressure, double p_mut, std::uint64_t seed, int elites) {
  Impl& O = *impl_;
  auto& I = *O.batch.impl_;
  SearchState& s = *O.s;
  if (!s.d.generational) throw std::invalid_argument("reproduce() is the generational schedule's; steady-state searches make batches");
  if (pressure < 1 || pressure > s.P) throw std::invalid_argument("tournament pressure outside 1..P");
  if (!(p_mut >= 0 && p_mut <= 1)) throw std::invalid_argument("mutation probability outside [0, 1]");
  if (elites < 0 || elites >= s.P) throw std::invalid_argument("elite count outside 0..P-1");
  for (int m = 0; m < s.P; ++m)
    if (std::isnan(selection[m])) throw std::invalid_argument("selection values must not be NaN");
  if (pressure != s.d.pressure || p_mut != s.d.p_mut || elites != s.d.elites) {
    s.d.pressure = pressure;
    s.d.p_mut = p_mut;
    s.d.elites = elites;
    I.rebuild();
  }
  // Elites: highest selection value, ties to the lowest index.
  std::vector<int> order(std::size_t(s.P));
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return selection[a] > selection[b]; });
  HIP_OK(hipMemcpyAsync((void*)s.d.elite, order.data(), sizeof(int) * std::size_t(elites), hipMemcpyHostToDevice, I.stream));
  HIP_OK(hipMemcpyAsync(O.select, selection, sizeof(double) * std::size_t(s.P), hipMemcpyHostToDevice, I.stream));
  HIP_OK(hipMemcpyAsync((void*)s.d.round_seed, &seed, sizeof(seed), hipMemcpyHostToDevice, I.stream));
  I.enqueue_front();
  I.enqueue_eval();
  I.launch("k_gen_place", I.g_member, k_gen_place, 128, 0);
  I.launch("k_clear", I.g_table, k_clear, 256, 0);
  I.launch("k_commit", I.g_member_w, k_commit, 64 * kVaryWaves, 0);
  const int zero[3] = {0, 0, 0};
  HIP_OK(hipMemcpyAsync((void*)s.d.n_pending, zero, sizeof(zero), hipMemcpyHostToDevice, I.stream));
  HIP_OK(hipStreamSynchronize(I.stream));
  if (const int e = s.device_error()) throw std::runtime_error("device reproduction error " + std::to_string(e));
  O.reuse = s.down(s.d.reuse, std::size_t(s.P));
  O.reproduced = true;
}

void Owner::node_counts(bool grammar_sites, unsigned* out) {
  const SearchState& s = *impl_->s;
  row_counts(s, grammar_sites, s.d.cur_n, s.d.cur_site, s.P, out);
}

void Owner::begin_steady(int batch) {
  if (batch != impl_->s->B)
    throw std::invalid_argument("steady batch " + std::to_string(batch) + " differs from the search configuration's " +
                                std::to_string(impl_->s->B));
}

void Owner::make_batch(const double* selection, int pressure, double p_mut, std::uint64_t seed, int* reuse) {
  Impl& O = *impl_;
  auto& I = *O.batch.impl_;
  SearchState& s = *O.s;
  if (s.d.generational) throw std::invalid_argument("make_batch() is the steady-state schedule's; generational searches reproduce");
  if (O.made) throw std::invalid_argument("the previous batch is not committed");
  if (pressure < 1 || pressure > s.P) throw std::invalid_argument("tournament pressure outside 1..P");
  if (!(p_mut >= 0 && p_mut <= 1)) throw std::invalid_argument("mutation probability outside [0, 1]");
  if (pressure != s.d.pressure || p_mut != s.d.p_mut) {
    s.d.pressure = pressure;
    s.d.p_mut = p_mut;
    I.rebuild();
  }
  HIP_OK(hipMemcpyAsync(O.select, selection, sizeof(double) * std::size_t(s.P), hipMemcpyHostToDevice, I.stream));
  HIP_OK(hipMemcpyAsync((void*)s.d.round_seed, &seed, sizeof(seed), hipMemcpyHostToDevice, I.stream));
  I.enqueue_front();
  HIP_OK(hipStreamSynchronize(I.stream));
  if (const int e = s.device_error()) throw std::runtime_error("device reproduction error " + std::to_string(e));
  O.reuse = s.down(s.d.reuse, std::size_t(s.B));
  std::copy(O.reuse.begin(), O.reuse.end(), reuse);
  O.made = true;
  O.scored = false;
}

void Owner::evaluate_batch(double* fitness, unsigned* wrong, std::uint64_t* counts) {
  Impl& O = *impl_;
  auto& I = *O.batch.impl_;
  const SearchState& s = *O.s;
  if (!O.made || O.scored) throw std::invalid_argument("no batch to evaluate");
  I.enqueue_eval();
  HIP_OK(hipStreamSynchronize(I.stream));
  if (const int e = s.device_error()) throw std::runtime_error("device evaluation error " + std::to_string(e));
  const std::vector<double> f = s.down(s.d.c_fit, std::size_t(s.B));
  const std::vector<unsigned> w = s.down(s.d.c_wrong, std::size_t(s.B));
  std::copy(f.begin(), f.end(), fitness);
  std::copy(w.begin(), w.end(), wrong);
  slot_counts(s, s.B, &O.reuse, counts);
  O.scored = true;
}

void Owner::batch_nodes(bool grammar_sites, unsigned* out) {
  const SearchState& s = *impl_->s;
  if (!impl_->made) throw std::invalid_argument("no batch made");
  row_counts(s, grammar_sites, s.d.x_n, s.d.x_site, s.B, out);
}

void Owner::commit_batch(const int* pairs, int n) {
  Impl& O = *impl_;
  auto& I = *O.batch.impl_;
  const SearchState& s = *O.s;
  if (!O.made || !O.scored) throw std::invalid_argument("commit before the batch is made and evaluated");
  std::vector<int> placed(std::size_t(s.P), -1);
  for (int k = 0; k < n; ++k) {
    const int m = pairs[2 * k], c = pairs[2 * k + 1];
    if (m < 0 || m >= s.P || c < 0 || c >= s.B) throw std::invalid_argument("placement outside the population or batch");
    placed[std::size_t(m)] = c;
  }
  const int zero[3] = {0, 0, 0};
  HIP_OK(hipMemcpyAsync((void*)s.d.placed, placed.data(), sizeof(int) * placed.size(), hipMemcpyHostToDevice, I.stream));
  I.launch("k_clear", I.g_table, k_clear, 256, 0);
  I.launch("k_commit", I.g_member_w, k_commit, 64 * kVaryWaves, 0);
  HIP_OK(hipMemcpyAsync((void*)s.d.n_pending, zero, sizeof(zero), hipMemcpyHostToDevice, I.stream));
  HIP_OK(hipStreamSynchronize(I.stream));
  if (const int e = s.device_error()) throw std::runtime_error("device commit error " + std::to_string(e));
  O.made = O.scored = false;
}

namespace {
Ast export_row(const SearchState& s, const int* origin, const Val* cval, int n, int nc);
}

json::Value Owner::export_member(int member) {
  const SearchState& s = *impl_->s;
  if (member < 0 || member >= s.P) throw std::invalid_argument("member outside the population");
  const std::size_t o = std::size_t(member) * std::size_t(s.R);
  const int n = s.down(s.d.cur_n + member, 1)[0], nc = s.down(s.d.cur_nc + member, 1)[0];
  const std::vector<int> origin = s.down(s.d.cur_origin + o, std::size_t(n));
  const std::vector<Val> cval = s.down(s.d.cur_cval + o, std::size_t(nc));
  return ast_json(export_row(s, origin.data(), cval.data(), n, nc));
}

namespace {
Ast export_row(const SearchState& s, const int* origin, const Val* cval, int n, int nc) {
  return export_whole_program(s.grammar, s.phase, s.layout_names, std::vector<int>(origin, origin + n),
                              std::vector<Val>(cval, cval + nc), s.binder_base);
}

// The run's success criterion on `test`, the held-out cases of a run or the
// split `gagp eval` names, with `fuel` per case (cpp/src/cli/run.cpp:179-239).
json::Value judge(const SearchState& s, const Cases& test, int fuel, const std::vector<int>& origin, const std::vector<Val>& cval) {
  // Lower the best phase exactly as the engine does, then score on the device.
  std::vector<Ins> code(std::size_t(s.d.code_cap));
  std::vector<Val> pool(std::size_t(s.d.pool_cap));
  std::vector<int> sites(origin.size(), -1);
  PhaseView gv{s.phase.exprs.data(), int(s.phase.exprs.size()), s.phase.prods.data(), int(s.phase.prods.size()),
               s.phase.nts.data(), int(s.phase.nts.size()), s.phase.domains.data(), int(s.phase.domains.size()),
               s.phase.values.data(), int(s.phase.values.size()), s.phase.minimum.data(), s.phase.depth_count,
               s.phase.logical_count, s.phase.site_depth ? 1 : 0, s.phase.root};
  const CRow row{origin.data(), sites.data(), cval.data(), int(origin.size()), int(cval.size())};
  std::vector<std::uint8_t> local_types = s.train.input_types;
  if (test.input_names != s.train.input_names) throw std::invalid_argument("test cases must bind the train inputs");
  const LowerResult lr = s.d.walks ? lower_phase<true>(gv, row, local_types.data(), int(local_types.size()), s.d.expected_type,
                                                       code.data(), int(code.size()), pool.data(), int(pool.size()))
                                    : lower_phase<false>(gv, row, local_types.data(), int(local_types.size()), s.d.expected_type,
                                                         code.data(), int(code.size()), pool.data(), int(pool.size()));
  if (lr.status != ST_OK) throw std::runtime_error("lowering of the best program failed");
  ProgramIn p;
  for (int i = 0; i < lr.code; ++i) {
    const Ins& x = code[i];
    p.op.push_back(x.op);
    const bool a = x.op == OP_PUSH_CONST || x.op == OP_LOAD || x.op == OP_STORE || x.op == OP_JMP || x.op == OP_JMP_IF_FALSE ||
                   x.op == OP_JMP_IF_TRUE || x.op == OP_CALL_BUILTIN;
    p.has_a.push_back(a);
    p.has_b.push_back(x.op == OP_CALL_BUILTIN);
    p.a.push_back(x.a);
    p.b.push_back(x.argc);
    p.fuel.push_back(x.fuel);
  }
  p.consts.assign(pool.begin(), pool.begin() + lr.constants);
  p.n_locals = lr.locals;
  Session session(test, fuel, 512, s.penalty, s.task.policy);
  std::vector<Val> values;
  std::vector<int> errors;
  session.eval_cases(p, values, errors);
  std::size_t correct = 0, errs = 0;
  double loss = 0;
  const bool absolute = s.task.policy.metric == METRIC_ABSOLUTE;
  for (int i = 0; i < test.n; ++i) {
    bool ok = false, error = false;
    double l = absolute ? s.task.case_cap : 1;
    const Val& want = test.expected[i];
    if (errors[i] != ERR_NONE) error = true;
    else if (want.tag == TAG_INT) {
      if (values[i].tag != TAG_INT) error = true;
      else {
        ok = values[i].i == want.i;
        if (absolute) l = std::min(s.task.case_cap, integer_gap(values[i].i, want.i));
      }
    } else if (want.tag == TAG_INT_LIST) {
      error = true;  // list results are outside the engine's values
    } else {
      bool exact = false;
      const Policy exact_policy{METRIC_EXACT, PROX_CAPPED, 0, 0.1, 100};
      (void)case_score(values[i], want, 1, exact_policy, exact);
      ok = exact;
    }
    if (ok) l = 0;
    correct += ok;
    errs += error;
    loss += l;
  }
  const double n = double(test.n);
  json::Value v = json::make_object();
  v.obj["cases"] = json::make_number(n);
  v.obj["correct"] = json::make_number(double(correct));
  v.obj["wrong"] = json::make_number(n - double(correct));
  v.obj["runtime_errors"] = json::make_number(double(errs));
  v.obj["mean_loss"] = json::make_number(loss / n);
  v.obj["solved"] = json::make_bool(correct == std::size_t(test.n));
  return v;
}
}  // namespace

json::Value evaluate_program(const std::string& task_path, const std::string& program_path, const std::string& split) {
  if (split != "train" && split != "test") throw std::invalid_argument("--split must be train or test");
  SearchSpec spec;
  spec.task_path = task_path;
  SearchState s(spec);
  const std::string& cases_path = split == "train" ? s.task.train_path : s.task.test_path;
  if (cases_path.empty()) throw std::invalid_argument("task has no " + split + " split");
  std::ifstream in(program_path, std::ios::binary);
  if (!in) throw std::invalid_argument("cannot read " + program_path);
  std::ostringstream text;
  text << in.rdbuf();
  const json::Value doc = json::parse(text.str());
  // The program must be a member of the task's grammar; its row lowers as
  // the engine lowers a member. Lowering capacities come from the search's
  // shapes, computed without allocating.
  const PhaseRow row = import_row(s.grammar, s.phase, decode_ast(program_of_document(doc)), 0, s.R);
  DeviceArena shapes(true);
  s.d = s.build(shapes);
  json::Value v = judge(s, load_cases(cases_path), int(s.grammar.fuel), row.origin, row.cval);
  v.obj["task"] = json::make_string(s.task.name);
  v.obj["split"] = json::make_string(split);
  return v;
}

SearchResult Batch::result(int index, bool save_all) {
  auto& I = *impl_;
  SearchState& s = *I.searches.at(std::size_t(index));
  const DSearch& d = s.d;
  SearchResult out;
  // A search that stopped early after G generations publishes what a run
  // with G generations publishes.
  const int stopped = s.down(d.stop, 1)[0];
  const int G = stopped ? stopped : s.G;
  SearchConfig published = s.cfg;
  published.generations = G;
  const auto fitness = s.down(d.fitness, std::size_t(s.P));
  const auto wrong = s.down(d.wrong, std::size_t(s.P));
  const auto nodes = s.down(d.nodes, std::size_t(s.P));
  // The front of the strict order over the final keys (for the generational
  // schedule the device's best index predates the last replacement).
  const Keys final_keys{fitness.data(), wrong.data(), nodes.data(), d.has_wrong};
  int best = 0;
  for (int m = 1; m < s.P; ++m)
    if (better(final_keys, m, best)) best = m;
  const auto cur_n = s.down(d.cur_n, std::size_t(s.P));
  const auto cur_nc = s.down(d.cur_nc, std::size_t(s.P));
  const Counters counters = s.down(d.counters, 1)[0];
  out.evaluated_program_cases = counters.program_cases;
  out.evaluated_programs = counters.programs;
  // History rows.
  const auto hn = s.down(d.hist_n, std::size_t(G));
  const auto hnc = s.down(d.hist_nc, std::size_t(G));
  const auto ho = s.down(d.hist_origin, std::size_t(G) * s.R);
  const auto hc = s.down(d.hist_cval, std::size_t(G) * s.R);
  const auto hf = s.down(d.hist_fit, std::size_t(G));
  const auto hw = s.down(d.hist_wrong, std::size_t(G));
  const auto ha = s.down(d.hist_all, std::size_t(G) * s.P);
  const double per_gen_ms = G ? I.run_ms / G : 0;
  json::Value history = json::make_array();
  for (int gen = 0; gen < G; ++gen) {
    const Ast ast = export_row(s, &ho[std::size_t(gen) * s.R], &hc[std::size_t(gen) * s.R], hn[gen], hnc[gen]);
    long double sum = 0;
    for (int m = 0; m < s.P; ++m) sum += ha[std::size_t(gen) * s.P + m];
    json::Value h = json::make_object();
    h.obj["generation"] = json::make_number(gen);
    h.obj["best_fitness"] = json::make_number(hf[gen]);
    h.obj["mean_fitness"] = json::make_number(double(sum / s.P));
    h.obj["best_wrong"] = json::make_number(hw[gen]);
    h.obj["best_nodes"] = json::make_number(double(ast.nodes.size()));
    h.obj["elapsed_ms"] = json::make_number(I.init_ms + per_gen_ms * (gen + 1));
    h.obj["eval_ms"] = json::make_number(per_gen_ms);
    h.obj["repro_ms"] = json::make_number(0);
    history.arr.push_back(h);
    if (save_all) {
      json::Value l = json::make_object();
      l.obj["generation"] = json::make_number(gen);
      l.obj["fitness"] = json::make_number(hf[gen]);
      l.obj["program"] = ast_json(ast);
      out.lineage.push_back(l);
    }
  }
  out.grammar_hash = s.grammar.content_hash;
  // Best member (the strict order's front after the last round).
  const std::vector<int> bo = s.down(d.cur_origin + std::size_t(best) * s.R, std::size_t(cur_n[best]));
  const std::vector<Val> bc = s.down(d.cur_cval + std::size_t(best) * s.R, std::size_t(cur_nc[best]));
  const Ast best_ast = export_whole_program(s.grammar, s.phase, s.layout_names, bo, bc, s.binder_base);
  json::Value bj = json::make_object();
  bj.obj["fitness"] = json::make_number(fitness[best]);
  bj.obj["wrong"] = json::make_number(wrong[best]);
  bj.obj["nodes"] = json::make_number(double(best_ast.nodes.size()));
  bj.obj["text"] = json::make_string(program_text(best_ast));
  bj.obj["program"] = ast_json(best_ast);
  if (save_all) {
    const Keys keys{fitness.data(), wrong.data(), nodes.data(), d.has_wrong};
    std::vector<int> order(std::size_t(s.P));
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return better(keys, a, b); });
    const auto all_o = s.down(d.cur_origin, std::size_t(s.P) * s.R);
    const auto all_c = s.down(d.cur_cval, std::size_t(s.P) * s.R);
    for (int m : order)
      out.population.push_back(ast_json(export_row(s, &all_o[std::size_t(m) * s.R], &all_c[std::size_t(m) * s.R], cur_n[m], cur_nc[m])));
  }
  json::Value r = json::make_object();
  r.obj["format"] = json::make_string("gagp-run-v1");
  r.obj["task"] = json::make_string(s.task.name);
  r.obj["engine"] = json::make_string("gpu");
  r.obj["population"] = json::make_number(s.P);
  r.obj["generations"] = json::make_number(G);
  r.obj["seed"] = json::make_string(std::to_string(s.cfg.seed));
  r.obj["train_cases"] = json::make_number(s.train.n);
  r.obj["search"] = effective_search_json(published, s.penalty, s.fuel);
  r.obj["best"] = bj;
  r.obj["history"] = history;
  json::Value tm = json::make_object();
  tm.obj["total"] = json::make_number(I.init_ms + I.run_ms);
  tm.obj["search"] = json::make_number(I.init_ms + I.run_ms);
  tm.obj["initialization"] = json::make_number(I.init_ms);
  r.obj["time_ms"] = tm;
  if (!s.task.test_path.empty()) r.obj["test"] = judge(s, load_cases(s.task.test_path), s.fuel, bo, bc);
  out.result = r;
  out.search_ms = I.init_ms + I.run_ms;
  return out;
}

// ====================== VALU issue-rate probe ======================
// Eight int32 VALU operations per iteration in four independent
// add/xor chains, which no closed form removes; the final test keeps every
// chain live without a store per iteration.
__global__ __launch_bounds__(256) void k_valu_probe(unsigned* out, int iters, unsigned seed) {
  unsigned a = threadIdx.x ^ seed, b = blockIdx.x, c = a + 1u, d = b + 3u, e = a ^ 5u, f = b ^ 7u, g = a + b, h = a - b;
  for (int i = 0; i < iters; ++i) {
    a += b; c += d; e += f; g += h;
    b ^= a; d ^= c; f ^= e; h ^= g;
  }
  if ((a ^ c ^ e ^ g) == 0x9e3779b9u) out[blockIdx.x] = b ^ d ^ f ^ h;
}

float valu_probe_ms(int blocks, int iters) {
  if (blocks <= 0 || iters <= 0) throw std::invalid_argument("probe needs positive blocks and iterations");
  DeviceArena tmp;
  unsigned* out = (unsigned*)tmp.alloc<unsigned>(std::size_t(blocks));
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, nullptr, out, iters, 1u);  // warm
  HIP_OK(hipEventRecord(a, nullptr));
  hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, nullptr, out, iters, 2u);
  HIP_OK(hipEventRecord(b, nullptr));
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventSynchronize(b));
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

// ====================== evaluation session ======================
namespace {
// Static stack heights of external bytecode (the verifier's stack-shape
// pass, cpp/src/runtime/bytecode_verify.cpp): every instruction gets one
// height; joins must agree; operand/constant/jump domains are checked.
std::vector<Ins> verify_program(const ProgramIn& p, int n_inputs, int& max_stack, bool& jumps, bool* input_stores = nullptr) {
  const int n = int(p.op.size());
  if (n == 0) throw std::invalid_argument("empty program");
  if (p.a.size() != std::size_t(n) || p.b.size() != std::size_t(n) || p.has_a.size() != std::size_t(n) ||
      p.has_b.size() != std::size_t(n) || (!p.fuel.empty() && p.fuel.size() != std::size_t(n)))
    throw std::invalid_argument("program arrays disagree in length");
  std::vector<int> height(std::size_t(n) + 1, -1);
  std::vector<int> work{0};
  height[0] = 0;
  max_stack = 0;
  jumps = false;
  // A NEG of a constant that is not a number is a guaranteed runtime error
  // (the verifier's abstract types, bytecode_verify.cpp:629-637): nothing
  // falls through it. Only the pushed constant's own path reaches it when no
  // jump targets the NEG.
  std::vector<char> targeted(std::size_t(n) + 1, 0);
  for (int i = 0; i < n; ++i)
    if ((p.op[i] == OP_JMP || p.op[i] == OP_JMP_IF_FALSE || p.op[i] == OP_JMP_IF_TRUE) && p.has_a[i] && p.a[i] >= 0 &&
        p.a[i] <= n)
      targeted[std::size_t(p.a[i])] = 1;
  const auto failing_neg = [&](int ip) {
    if (ip == 0 || targeted[std::size_t(ip)] || p.op[ip - 1] != OP_PUSH_CONST || !p.has_a[ip - 1] || p.a[ip - 1] < 0 ||
        p.a[ip - 1] >= int(p.consts.size()))
      return false;
    const int tag = p.consts[std::size_t(p.a[ip - 1])].tag;
    return tag != TAG_INT && tag != TAG_FLOAT;
  };
  const auto flow = [&](int to, int h) {
    if (to < 0 || to > n) throw std::invalid_argument("jump target out of range");
    if (height[to] < 0) { height[to] = h; work.push_back(to); }
    else if (height[to] != h) throw std::invalid_argument("stack heights disagree at a join");
  };
  while (!work.empty()) {
    const int ip = work.back();
    work.pop_back();
    if (ip == n) continue;
    const int h = height[ip];
    const int op = p.op[ip];
    int pop = 0, push = 0;
    bool falls = true;
    switch (op) {
      case OP_PUSH_CONST:
        if (!p.has_a[ip] || p.a[ip] < 0 || p.a[ip] >= int(p.consts.size())) throw std::invalid_argument("const index out of range");
        push = 1; break;
      case OP_LOAD:
        if (!p.has_a[ip] || p.a[ip] < 0 || p.a[ip] >= 64) throw std::invalid_argument("local index out of range");
        push = 1; break;
      case OP_STORE:
        if (!p.has_a[ip] || p.a[ip] < 0 || p.a[ip] >= 64) throw std::invalid_argument("local index out of range");
        // A store into an input slot: the evaluators shadow it per execution.
        if (p.a[ip] < n_inputs && input_stores) *input_stores = true;
        pop = 1; break;
      case OP_NEG: pop = 1; push = 1; falls = !failing_neg(ip); break;
      case OP_NOT: case OP_CHECK_INT: case OP_CHECK_LIST: pop = 1; push = 1; break;
      case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_MOD:
      case OP_LT: case OP_LE: case OP_GT: case OP_GE: case OP_EQ: case OP_NE: pop = 2; push = 1; break;
      case OP_JMP: jumps = true; if (!p.has_a[ip]) throw std::invalid_argument("jump target missing"); flow(p.a[ip], h); falls = false; break;
      case OP_JMP_IF_FALSE: case OP_JMP_IF_TRUE:
        jumps = true;
        if (!p.has_a[ip]) throw std::invalid_argument("jump target missing");
        if (h < 1) throw std::invalid_argument("stack underflow");
        flow(p.a[ip], h - 1);
        pop = 1; break;
      case OP_CALL_BUILTIN:
        if (!p.has_b[ip] || p.b[ip] < 0 || p.b[ip] > 8) throw std::invalid_argument("invalid builtin argc");
        pop = p.b[ip]; push = 1; break;
      case OP_RETURN: if (h < 1) throw std::invalid_argument("return requires value on stack"); falls = false; break;
      default: throw Unsupported("opcode outside the device VM: " + std::to_string(op));
    }
    if (h < pop) throw std::invalid_argument("stack underflow");
    const int after = h - pop + push;
    max_stack = std::max(max_stack, after);
    if (after > kMaxStack) throw Unsupported("stack deeper than the device VM");
    if (falls) flow(ip + 1, after);
  }
  std::vector<Ins> out(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) {
    const std::uint32_t f = p.fuel.empty() ? 1u : p.fuel[i];
    if (f > 255) throw Unsupported("instruction fuel charges above 255");
    const int h = height[i] < 0 ? 0 : height[i];  // unreachable instructions never run
    out[i] = Ins{p.op[i], std::uint8_t(p.b[i]), std::uint8_t(h), std::uint8_t(f), p.a[i]};
  }
  return out;
}
}  // namespace

// Static tags of straight-line external bytecode: the program is certified
// for vm_fast when every operand tag is known and Int or Bool and every
// instruction is one vm_fast restates with checks that cannot fail (the same
// rules as the lowering's fast_operation). `input_tags[k]` is the one tag
// input k has wherever it is bound (TAG_INVALID when the cases disagree).
struct FastCert { bool fast = false; int result_tag = TAG_INT; };
FastCert certify_fast(const std::vector<Ins>& code, const std::vector<Val>& consts, const std::vector<std::uint8_t>& input_tags,
                      int n_locals, int max_stack, bool jumps, int C) {
  FastCert cert;
  if (jumps || max_stack > kFastStack || code.size() > std::size_t(kFastCode) || consts.size() > std::size_t(kFastPool) ||
      n_locals > kFastInputs || C > kScoreCases)
    return cert;
  std::vector<std::uint8_t> tags;
  for (const Ins& ins : code) {
    const int h = int(tags.size());
    if (ins.sp != h) return cert;
    const std::uint8_t* args = h >= ins.argc ? tags.data() + (h - ins.argc) : nullptr;
    switch (ins.op) {
      case OP_PUSH_CONST: {
        const int t = consts[std::size_t(ins.a)].tag;
        if (!fast_type(t)) return cert;
        tags.push_back(std::uint8_t(t));
        break;
      }
      case OP_LOAD: {
        if (ins.a >= n_locals || ins.a >= int(input_tags.size()) || !fast_type(input_tags[std::size_t(ins.a)])) return cert;
        tags.push_back(input_tags[std::size_t(ins.a)]);
        break;
      }
      case OP_RETURN:
        if (h < 1 || !fast_type(tags.back())) return cert;
        cert.fast = true;
        cert.result_tag = tags.back();
        return cert;
      case OP_CALL_BUILTIN: {
        if (!args || !fast_operation(OP_CALL_BUILTIN, ins.a, ins.argc, args[0], ins.argc > 1 ? args[1] : 0, ins.argc > 2 ? args[2] : 0))
          return cert;
        tags.resize(std::size_t(h - ins.argc));
        tags.push_back(TAG_INT);
        break;
      }
      default: {
        const int arity = ins.op == OP_NEG || ins.op == OP_NOT || ins.op == OP_CHECK_INT ? 1 : 2;
        if (h < arity) return cert;
        const std::uint8_t* a = tags.data() + (h - arity);
        if (!fast_operation(ins.op, 0, arity, a[0], arity > 1 ? a[1] : 0, 0)) return cert;
        const bool compare = ins.op >= OP_LT && ins.op <= OP_NE;
        const std::uint8_t rt = compare ? TAG_BOOL : ins.op == OP_NOT ? TAG_BOOL : TAG_INT;
        tags.resize(std::size_t(h - arity));
        tags.push_back(rt);
        break;
      }
    }
  }
  return cert;  // falls off the end: the general VM reports it
}

namespace {
// Host packing of one evaluation launch: verified code and pools laid out
// back to back, with each program's offsets and certificate.
struct PackedPrograms {
  std::vector<Ins> code;
  std::vector<Val> pool;
  std::vector<std::size_t> code_off, pool_off;
  std::vector<int> lens, jumps, locals, fast, pool_n, result;
  int slots = 1, temps = 0, shadow = 0;
  bool any_fast = false, any_general = false;
};

}  // namespace

struct Session::Impl {
  DeviceArena arena;
  DeviceCases cases;
  std::vector<std::uint8_t> input_tags;
  std::size_t arena_len = 0;
  int fuel, lanes;
  double penalty;
  Policy policy;
  float last_ms = 0;
  PackedPrograms pack(const std::vector<ProgramIn>& programs) const;
  std::vector<DProg> program_table(const PackedPrograms& k, GAS const Ins* code, GAS const Val* pool) const;
  void launch(const PackedPrograms& k, const DProg* progs, int n, double* fit, unsigned* wrong, unsigned* counts) const;
};

namespace {
// Container values reach the device only as views inside the case arena.
void check_view(const Val& v, std::size_t arena_len) {
  if (v.tag == TAG_INT_LIST) {
    const std::uint64_t u = std::uint64_t(v.i), off = u >> 16, len = u & 0xffff;
    if (off > arena_len || len > arena_len - off) throw std::invalid_argument("IntList view outside the list arena");
  } else if (v.tag == TAG_STRING || v.tag == TAG_FLOAT_LIST || v.tag == TAG_STRING_LIST || v.tag == TAG_CHAR) {
    throw Unsupported("string, char and non-Int list values are outside the device VM");
  } else if (v.tag > TAG_INVALID) {
    throw std::invalid_argument("unknown value tag");
  }
}
}  // namespace

Session::Session(const Cases& c, int fuel, int lanes, double penalty, Policy policy) : impl_(std::make_unique<Impl>()) {
  for (std::size_t i = 0; i < c.inputs.size(); ++i)
    if (c.set.empty() || c.set[i]) check_view(c.inputs[i], c.arena.size());
  for (const auto& v : c.expected) check_view(v, c.arena.size());
  impl_->arena_len = c.arena.size();
  impl_->cases.upload(impl_->arena, c);
  const std::size_t K = c.input_names.size();
  impl_->input_tags.assign(K, TAG_INVALID);
  std::vector<bool> seen(K, false), mixed(K, false);
  for (int i = 0; i < c.n; ++i)
    for (std::size_t k = 0; k < K; ++k) {
      const std::size_t at = std::size_t(i) * K + k;
      if (!c.set.empty() && !c.set[at]) continue;
      const std::uint8_t t = c.inputs[at].tag;
      if (!seen[k]) { seen[k] = true; impl_->input_tags[k] = t; }
      else if (impl_->input_tags[k] != t) mixed[k] = true;
    }
  for (std::size_t k = 0; k < K; ++k)
    if (mixed[k]) impl_->input_tags[k] = TAG_INVALID;
  impl_->fuel = fuel;
  impl_->lanes = lanes;
  impl_->penalty = penalty;
  impl_->policy = policy;
}
Session::~Session() = default;
float Session::last_kernel_ms() const { return impl_->last_ms; }

PackedPrograms Session::Impl::pack(const std::vector<ProgramIn>& programs) const {
  const auto& I = *this;
  PackedPrograms k;
  bool stores = false;  // some program stores into an input
  for (const auto& p : programs) {
    k.temps = std::max(k.temps, std::min(64, p.n_locals) - I.cases.cd.n_inputs);
    int ms = 0;
    bool j = false;
    const auto ins = verify_program(p, I.cases.cd.n_inputs, ms, j, &stores);
    for (const auto& v : p.consts) check_view(v, I.arena_len);
    const FastCert cert = certify_fast(ins, p.consts, I.input_tags, p.n_locals, ms, j, I.cases.cd.C);
    k.fast.push_back(cert.fast);
    k.result.push_back(cert.result_tag);
    k.pool_n.push_back(int(p.consts.size()));
    k.slots = std::max(k.slots, ms + 1);
    k.code_off.push_back(k.code.size());
    k.pool_off.push_back(k.pool.size());
    k.code.insert(k.code.end(), ins.begin(), ins.end());
    k.pool.insert(k.pool.end(), p.consts.begin(), p.consts.end());
    k.lens.push_back(int(ins.size()));
    k.jumps.push_back(j);
    k.locals.push_back(p.n_locals);
  }
  if (k.pool.empty()) k.pool.push_back(val_invalid());
  k.shadow = stores ? I.cases.cd.n_inputs : 0;
  k.temps += k.shadow;
  k.any_fast = std::count(k.fast.begin(), k.fast.end(), 1) > 0;
  k.any_general = std::count(k.fast.begin(), k.fast.end(), 0) > 0;
  return k;
}

std::vector<DProg> Session::Impl::program_table(const PackedPrograms& k, GAS const Ins* code, GAS const Val* pool) const {
  std::vector<DProg> progs(k.lens.size());
  for (std::size_t i = 0; i < progs.size(); ++i)
    progs[i] = DProg{code + k.code_off[i], pool + k.pool_off[i], k.lens[i], k.jumps[i], k.locals[i], k.fast[i], k.pool_n[i], k.result[i]};
  return progs;
}

// Both evaluation kernels over a packed launch on the null stream.
void Session::Impl::launch(const PackedPrograms& k, const DProg* progs, int n, double* fit, unsigned* wrong, unsigned* counts) const {
  const auto& I = *this;
  const EvalParams e{I.cases.cd, I.fuel, I.lanes, I.penalty, I.policy};
  const std::size_t lds = std::size_t(kWavesPerBlock) * (k.slots + k.temps) * 64 * 10;
  const std::size_t fast_lds = std::size_t(kWavesPerBlock) * 8 * std::size_t(I.cases.cd.C);
  if (k.any_general && lds > 65536)
    HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_eval_programs<false>), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(lds)));
  if (k.any_fast && fast_lds > 65536)
    HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_eval_programs<true>), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(fast_lds)));
  const int blocks = std::min(4096, (n + kWavesPerBlock - 1) / kWavesPerBlock);
  if (k.any_fast)
    hipLaunchKernelGGL(k_eval_programs<true>, dim3(blocks), dim3(kEvalThreads), fast_lds, nullptr, progs, n, e, k.slots, k.temps,
                       k.shadow, fit, wrong, counts);
  if (k.any_general)
    hipLaunchKernelGGL(k_eval_programs<false>, dim3(blocks), dim3(kEvalThreads), lds, nullptr, progs, n, e, k.slots, k.temps,
                       k.shadow, fit, wrong, counts);
  HIP_OK(hipGetLastError());
}

void Session::eval(const std::vector<ProgramIn>& programs, double* fitness, std::uint32_t* wrong, std::uint32_t* counts) {
  auto& I = *impl_;
  const int n = int(programs.size());
  if (!n) return;
  const PackedPrograms k = I.pack(programs);
  DeviceArena tmp;
  auto dcode = tmp.upload(k.code);
  auto dpool = tmp.upload(k.pool);
  auto dprogs = tmp.upload(I.program_table(k, dcode, dpool));
  auto dfit = tmp.alloc<double>(std::size_t(n));
  auto dwrong = tmp.alloc<unsigned>(std::size_t(n));
  auto dcounts = tmp.alloc<unsigned>(std::size_t(n) * 3);
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  HIP_OK(hipEventRecord(a, nullptr));
  I.launch(k, (const DProg*)dprogs, n, (double*)dfit, (unsigned*)dwrong, (unsigned*)dcounts);
  HIP_OK(hipEventRecord(b, nullptr));
  HIP_OK(hipEventSynchronize(b));
  HIP_OK(hipEventElapsedTime(&I.last_ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIP_OK(hipMemcpy(fitness, (const void*)dfit, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (wrong) HIP_OK(hipMemcpy(wrong, (const void*)dwrong, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
  if (counts) HIP_OK(hipMemcpy(counts, (const void*)dcounts, sizeof(unsigned) * 3 * n, hipMemcpyDeviceToHost));
}

// ---- allocation-free evaluation ----
// A fixed device workspace, one copy per GPU, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc). It holds one launch's program table,
// code, constant pools and per-program results. Every call on a GPU shares
// that GPU's copy: calls that overlap on one GPU race, and the workspace keeps
// the last call's contents (the kernels write every result they report, so
// nothing needs clearing).
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
std::mutex scratch_mutex;
std::map<int, unsigned char*> scratch_address;  // device -> its copy of `scratch`

// The current device's copy of the workspace, resolved once per device with
// that device current. Host code uses only this address, never the symbol.
unsigned char* scratch_on_current_device() {
  int device = 0;
  HIP_OK(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(scratch_mutex);
  auto it = scratch_address.find(device);
  if (it != scratch_address.end()) return it->second;
  void* p = nullptr;
  HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw std::runtime_error("no device address for the evaluation workspace");
  scratch_address[device] = static_cast<unsigned char*>(p);
  return static_cast<unsigned char*>(p);
}

std::size_t align256(std::size_t x) { return (x + 255) & ~std::size_t(255); }
}  // namespace

std::size_t Session::static_bytes(std::size_t n, std::size_t code, std::size_t pool) {
  return align256(n * sizeof(DProg)) + align256(code * sizeof(Ins)) + align256(pool * sizeof(Val)) +
         align256(n * sizeof(double)) + align256(n * sizeof(unsigned)) + align256(3 * n * sizeof(unsigned));
}

void Session::eval_static(const std::vector<ProgramIn>& programs, double* fitness, std::uint32_t* wrong, std::uint32_t* counts) {
  const auto& I = *impl_;
  const int n = int(programs.size());
  if (!n) return;
  const PackedPrograms k = I.pack(programs);
  const std::size_t need = static_bytes(std::size_t(n), k.code.size(), k.pool.size());
  if (need > kStaticScratchBytes)
    throw std::invalid_argument("programs need " + std::to_string(need) + " bytes of evaluation workspace; the allocation-free "
                                "entry point holds at most " + std::to_string(kStaticScratchBytes));
  unsigned char* base = scratch_on_current_device();
  std::size_t at = 0;
  const auto carve = [&](std::size_t bytes) { unsigned char* p = base + at; at += align256(bytes); return p; };
  auto* dprogs = reinterpret_cast<DProg*>(carve(std::size_t(n) * sizeof(DProg)));
  auto* dcode = reinterpret_cast<Ins*>(carve(k.code.size() * sizeof(Ins)));
  auto* dpool = reinterpret_cast<Val*>(carve(k.pool.size() * sizeof(Val)));
  auto* dfit = reinterpret_cast<double*>(carve(std::size_t(n) * sizeof(double)));
  auto* dwrong = reinterpret_cast<unsigned*>(carve(std::size_t(n) * sizeof(unsigned)));
  auto* dcounts = reinterpret_cast<unsigned*>(carve(std::size_t(3) * n * sizeof(unsigned)));
  const auto progs = I.program_table(k, gbl(static_cast<const Ins*>(dcode)), gbl(static_cast<const Val*>(dpool)));
  HIP_OK(hipMemcpy(dcode, k.code.data(), k.code.size() * sizeof(Ins), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dpool, k.pool.data(), k.pool.size() * sizeof(Val), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dprogs, progs.data(), progs.size() * sizeof(DProg), hipMemcpyHostToDevice));
  I.launch(k, dprogs, n, dfit, dwrong, dcounts);
  HIP_OK(hipMemcpy(fitness, dfit, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (wrong) HIP_OK(hipMemcpy(wrong, dwrong, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
  if (counts) HIP_OK(hipMemcpy(counts, dcounts, sizeof(unsigned) * 3 * n, hipMemcpyDeviceToHost));
}

void Session::eval_cases(const ProgramIn& program, std::vector<Val>& values, std::vector<int>& errors) {
  auto& I = *impl_;
  int ms = 0;
  bool j = false;
  bool stores = false;
  const auto ins = verify_program(program, I.cases.cd.n_inputs, ms, j, &stores);
  for (const auto& v : program.consts) check_view(v, I.arena_len);
  const int shadow = stores ? I.cases.cd.n_inputs : 0;
  const int slots = ms + 1, temps = std::max(0, std::min(64, program.n_locals) - I.cases.cd.n_inputs) + shadow;
  DeviceArena tmp;
  std::vector<Val> consts = program.consts;
  if (consts.empty()) consts.push_back(val_invalid());
  const FastCert cert = certify_fast(ins, program.consts, I.input_tags, program.n_locals, ms, j, I.cases.cd.C);
  const DProg p{tmp.upload(ins), tmp.upload(consts), int(ins.size()), j, program.n_locals, cert.fast, int(program.consts.size()),
                cert.result_tag};
  const int C = I.cases.cd.C;
  auto dv = tmp.alloc<Val>(std::size_t(C));
  auto de = tmp.alloc<int>(std::size_t(C));
  const int blocks = std::max(1, std::min(1024, (C + 255) / 256));
  if (cert.fast) {
    hipLaunchKernelGGL(k_eval_cases<true>, dim3(blocks), dim3(kEvalThreads), 0, nullptr, p, I.cases.cd, I.fuel, slots, temps, shadow,
                       dv, de);
  } else {
    const std::size_t lds = std::size_t(kWavesPerBlock) * (slots + temps) * 64 * 10;
    if (lds > 65536)
      HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_eval_cases<false>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 int(lds)));
    hipLaunchKernelGGL(k_eval_cases<false>, dim3(blocks), dim3(kEvalThreads), lds, nullptr, p, I.cases.cd, I.fuel, slots, temps, shadow,
                       dv, de);
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipDeviceSynchronize());
  values.resize(std::size_t(C));
  errors.resize(std::size_t(C));
  HIP_OK(hipMemcpy(values.data(), dv, sizeof(Val) * C, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(errors.data(), de, sizeof(int) * C, hipMemcpyDeviceToHost));
}

}  // namespace gagp_amd
