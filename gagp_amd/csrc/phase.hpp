// Per-child operators of the resident phase population on this engine's row
// layout: a tree row stores, per node, only its origin (phase expression id)
// and its packed variation site; node fields (kind, opcode, arity, fuel, type)
// are read from the expression table, constants live in a per-row value list
// in prefix order.
//
// The draw order and every accept/reject rule restate the reference bodies
// (the wavefront operators that use these pieces are in phase_wave.hpp):
//   generate_phase        cpp/src/evolution/repro/gpu/device/phase_generate.cuh:125-270
//   tournament, sites     cpp/src/evolution/repro/gpu/phase_operators.cuh:46-83
//   locals_defined        .../device/phase_compile.cuh:170-200
//   compile_phase_tree    .../device/phase_compile.cuh:206-895 (expressions,
//                         traversals and statements)
#pragma once
#include "core.hpp"
#include "grammar.hpp"

namespace gagp_amd {

// Semantic phase capacity (nodes per phase row) used by the operator
// formulas; physical rows are sized to the grammar's node budget, which the
// operators never exceed (crossover and mutation are bounded by it first).
constexpr int kPhaseCapacity = 1024;
constexpr int kPending = 192;

enum Status : int { ST_OK = 0, ST_INVALID = 1, ST_UNSUPPORTED = 2, ST_CAPACITY = 3 };

// The phase tables as the operators read them. On the device the pointers
// are built inside the kernel (from PhaseTables, or from an LDS copy), so the
// compiler sees which memory they address.
struct PhaseView {
  const PExpr* expr; int n_expr;
  const PProd* prod; int n_prod;
  const PNt* nt; int n_nt;
  const PDomain* dom; int n_dom;
  const Val* values; int n_values;
  const unsigned* minimum; int depth_count, logical_count; int site_depth;
  int root;
};
// Device storage of the phase tables (global memory), kept in the search state.
struct PhaseTables {
  GAS const PExpr* expr; int n_expr;
  GAS const PProd* prod; int n_prod;
  GAS const PNt* nt; int n_nt;
  GAS const PDomain* dom; int n_dom;
  GAS const Val* values; int n_values;
  GAS const unsigned* minimum; int depth_count, logical_count; int site_depth;
  int root;
};
GHD inline PhaseView view_of(const PhaseTables& t) {
  return PhaseView{t.expr,   t.n_expr,  t.prod,    t.n_prod,      t.nt,          t.n_nt,         t.dom, t.n_dom,
                   t.values, t.n_values, t.minimum, t.depth_count, t.logical_count, t.site_depth, t.root};
}
// LDS bytes of a table copy: each array padded to 16 bytes.
GHD inline std::size_t table_bytes(std::size_t bytes) { return (bytes + 15) & ~std::size_t(15); }
GHD inline std::size_t phase_tables_bytes(const PhaseTables& t) {
  return table_bytes(sizeof(PExpr) * std::size_t(t.n_expr)) + table_bytes(sizeof(PProd) * std::size_t(t.n_prod)) +
         table_bytes(sizeof(PNt) * std::size_t(t.n_nt)) + table_bytes(sizeof(PDomain) * std::size_t(t.n_dom)) +
         table_bytes(sizeof(Val) * std::size_t(t.n_values)) +
         table_bytes(sizeof(unsigned) * std::size_t(t.n_expr) * std::size_t(t.depth_count) * std::size_t(t.logical_count));
}

// Site packing: -1 for a non-site node, else entry expression | parent scope << 17.
GHD inline int site_pack(int entry, int parent_scope) { return entry | (parent_scope << 17); }
GHD inline int site_entry(int s) { return s & 0x1ffff; }
GHD inline int site_parent_scope(int s) { return (s >> 17) & 0xff; }
GHD inline int site_nt(const PhaseView& g, int s) { return g.expr[site_entry(s)].target; }
GHD inline int site_scope(const PhaseView& g, int s) {
  const PExpr& e = g.expr[site_entry(s)];
  return e.scope_count | (g.nt[e.target].fixed ? kFixedSite : 0);
}

// A row view: origins, sites, constants of one phase tree.
struct Row {
  int* origin;
  int* site;
  Val* cval;
  int n, nc;  // node and constant counts (n == 0: invalid row)
};
struct CRow {
  const int* origin;
  const int* site;
  const Val* cval;
  int n, nc;
};
GHD inline CRow as_const(const Row& r) { return {r.origin, r.site, r.cval, r.n, r.nc}; }

struct GenResult { int status, nodes, constants, steps; };

GHD inline unsigned g_minimum(const PhaseView& g, int e, int d, int l) {
  return (e < 0 || e >= g.n_expr || d < 0 || d >= g.depth_count || l < 0 || l >= g.logical_count)
             ? 0xffffffffu : g.minimum[(unsigned(e) * g.depth_count + d) * g.logical_count + l];
}
GHD inline int child_scope(const PExpr& x, int i, int scope) { return x.child_scope[i] >= 0 ? x.child_scope[i] : scope; }

// A pending derivation task: expression, remaining depth, scope count,
// remaining logical (site) depth.
struct GenTask { int e, d, sc, l; };
// The pending stack of a derivation on one thread.
struct LocalTasks {
  GenTask t[kPending];
  GHD GenTask get(int k) const { return t[k]; }
  GHD void set(int k, const GenTask& x) { t[k] = x; }
};

// The production a reference expands to: among the nonterminal's
// productions whose minimum cost at (depth d, logical depth l) fits
// `available`, weights divided by the largest eligible weight unless the
// table is normalized statically, one unit draw scaled by their sum and
// scanned in production order. Returns the chosen expression, or one of the
// codes below.
constexpr int kChooseNone = -1;      // no production selected
constexpr int kChooseInvalid = -2;   // an eligible production names no expression
constexpr int kChooseCapacity = -3;  // no eligible positive weight
GHD inline int choose_production_in_order(const PhaseView& g, const PNt& nt, int d, int l, unsigned available,
                                          SplitMix& random) {
  double maximum = nt.normalized ? 1 : 0, total = 0;
  if (!nt.normalized)
    for (int i = 0; i < nt.production_count; ++i) {
      const PProd p = g.prod[nt.production_offset + i];
      if (g_minimum(g, p.expression, d, l) <= available && p.weight > maximum) maximum = p.weight;
    }
  if (!(maximum > 0)) return kChooseCapacity;
  for (int i = 0; i < nt.production_count; ++i) {
    const PProd p = g.prod[nt.production_offset + i];
    if (g_minimum(g, p.expression, d, l) <= available) total += nt.normalized ? p.weight : p.weight / maximum;
  }
  double choice = random.unit() * total;
  int chosen = kChooseNone;
  for (int i = 0; i < nt.production_count; ++i) {
    const PProd p = g.prod[nt.production_offset + i];
    if (g_minimum(g, p.expression, d, l) > available) continue;
    if (p.expression < 0 || p.expression >= g.n_expr) return kChooseInvalid;
    chosen = p.expression;
    choice -= nt.normalized ? p.weight : p.weight / maximum;
    if (choice < 0) break;
  }
  return chosen;
}
// A derivation's pending-stack type selects how it chooses productions: one
// thread scans them in order; a wavefront's stack (WaveTasks) has its own
// overload with one production per lane and the same result.
template <class Tasks>
GHD inline int choose_production(const Tasks&, const PhaseView& g, const PNt& nt, int d, int l, unsigned available,
                                 SplitMix& random) {
  return choose_production_in_order(g, nt, d, l, available, random);
}

// Top-down derivation with a minimum-cost reservation and a pending stack.
// Writes origins (and sites when `sites` is non-null) and constants.
template <class Tasks>
GHD inline GenResult generate_phase_with(const PhaseView& g, int root_expression, int scope_count, int depth, int capacity,
                                         std::uint64_t seed, int* origins, Val* constants, int* sites, int logical_depth,
                                         Tasks& pending) {
  GenResult r{ST_INVALID, 0, 0, 0};
  if (root_expression < 0 || root_expression >= g.n_expr || scope_count < 0 || scope_count > 15 || depth <= 0 ||
      depth >= g.depth_count || depth > 64 || logical_depth < 0 || logical_depth >= g.logical_count || capacity <= 0 ||
      capacity > kPhaseCapacity || g.n_expr > 131071)
    return r;
  int count = 1;
  pending.set(0, GenTask{root_expression, depth, scope_count, logical_depth});
  unsigned reserved = g_minimum(g, root_expression, depth, logical_depth);
  if (reserved > unsigned(capacity)) { r.status = ST_CAPACITY; return r; }
  int next_site = -1;
  SplitMix random(seed);
  constexpr int unselected = 131071;
  while (count) {
    if (++r.steps > 8192) { r.status = ST_CAPACITY; return r; }
    const GenTask task = pending.get(--count);
    const unsigned required = g_minimum(g, task.e, task.d, task.l);
    if (required > reserved) return r;
    reserved -= required;
    const PExpr& x = g.expr[task.e];
    if (x.kind == PK_REFERENCE) {
      if (x.target < 0 || x.target >= g.n_nt || x.scope_count < 0 || x.scope_count > 8) return r;
      if (sites) {
        if (next_site >= 0) return r;
        next_site = site_pack(task.e, task.sc);
      }
      const PNt nt = g.nt[x.target];
      if (nt.scope_count != x.scope_count || nt.production_offset < 0 || nt.production_count <= 0 ||
          nt.production_offset > g.n_prod || nt.production_count > g.n_prod - nt.production_offset)
        return r;
      if (g.site_depth && task.l == 0) { r.status = ST_CAPACITY; return r; }
      GenTask next{unselected, task.d, x.scope_count, task.l - (g.site_depth ? 1 : 0)};
      if (reserved > unsigned(capacity - r.nodes)) { r.status = ST_CAPACITY; return r; }
      const unsigned available = unsigned(capacity - r.nodes) - reserved;
      const int chosen = choose_production(pending, g, nt, task.d, next.l, available, random);
      if (chosen == kChooseCapacity) { r.status = ST_CAPACITY; return r; }
      if (chosen < 0 || count >= kPending) return r;
      next.e = chosen;
      reserved += g_minimum(g, next.e, next.d, next.l);
      pending.set(count++, next);
      continue;
    }
    if (x.kind == PK_ALIAS) return r;  // templates are not admitted by the front end
    if (task.d <= 0 || r.nodes >= capacity) { r.status = ST_CAPACITY; return r; }
    const int arity = x.node.arity;
    if (arity < 0 || arity > kMaxArity) return r;
    if (x.kind == PK_BOUND) {
      if (arity != 0 || x.node.operand < 0 || x.node.operand >= 16) return r;
    } else if (x.kind == PK_CONSTANT) {
      if (x.target < 0 || x.target >= g.n_dom || arity != 0) return r;
      const PDomain d = g.dom[x.target];
      Val v;
      if (d.integer_range) {
        if (d.minimum > d.maximum) return r;
        v = val_int(random.integer(d.minimum, d.maximum));
      } else {
        if (d.value_offset < 0 || d.value_count <= 0 || d.value_offset > g.n_values || d.value_count > g.n_values - d.value_offset)
          return r;
        v = g.values[d.value_offset + random.bounded(std::uint64_t(d.value_count))];
      }
      if (v.tag == TAG_INVALID || v.tag == TAG_FALLBACK) return r;
      constants[r.constants++] = v;
    } else if (x.kind != PK_OPERATION && x.kind != PK_INPUT) return r;
    if (sites) { sites[r.nodes] = next_site; next_site = -1; }
    origins[r.nodes++] = task.e;
    for (int i = arity - 1; i >= 0; --i) {
      if (count >= kPending) { r.status = ST_CAPACITY; return r; }
      const int child = x.children[i];
      const unsigned need = g_minimum(g, child, task.d - 1, task.l);
      if (need > unsigned(capacity - r.nodes) || reserved > unsigned(capacity - r.nodes) - need) {
        r.status = ST_CAPACITY;
        return r;
      }
      GenTask next = task;
      next.e = child;
      --next.d;
      next.sc = child_scope(x, i, task.sc);
      pending.set(count++, next);
      reserved += need;
    }
  }
  if (reserved || !r.nodes) return r;
  r.status = ST_OK;
  return r;
}
GHD inline GenResult generate_phase(const PhaseView& g, int root_expression, int scope_count, int depth, int capacity,
                                    std::uint64_t seed, int* origins, Val* constants, int* sites, int logical_depth) {
  LocalTasks pending;
  return generate_phase_with(g, root_expression, scope_count, depth, capacity, seed, origins, constants, sites, logical_depth,
                             pending);
}

GHD inline int arity_of(const PhaseView& g, int origin) { return g.expr[origin].node.arity; }
GHD inline bool is_constant(const PhaseView& g, int origin) { return g.expr[origin].kind == PK_CONSTANT; }

struct OpStats { int parent, crossover_attempted, crossover, crossover_rejected, mutation, changed, error, reusable_parent, donor; };
struct OpPolicy {
  int counted_sites;
  int uniform_phase;
  double crossover_probability;
  int crossover_attempts;
  int mutation_max_nodes;
};

// Member ranking keys: fitness (desc), wrong (asc), nodes (asc), index (asc).
struct Keys {
  const double* fitness;
  const unsigned* wrong;
  const unsigned* nodes;
  int has_wrong;
};
// Strict total order of the steady-state schedule (cpp/src/evolution/evolve.cpp:358-362).
GHD inline bool better(const Keys& k, int a, int b) {
  if (k.fitness[a] != k.fitness[b]) return k.fitness[a] > k.fitness[b];
  if (k.has_wrong) {
    if (k.wrong[a] != k.wrong[b]) return k.wrong[a] < k.wrong[b];
    if (k.nodes[a] != k.nodes[b]) return k.nodes[a] < k.nodes[b];
  }
  return a < b;
}
// Tournament on selection values. With wrong counts the selection value is
// P - rank under `better`, so comparing values is comparing by `better`;
// without, the value is the fitness and ties keep the earlier draw.
GHD inline bool selected_over(const Keys& k, int candidate, int best) {
  return k.has_wrong ? better(k, candidate, best) : k.fitness[candidate] > k.fitness[best];
}
GHD inline int tournament(const Keys& k, int population, int pressure, SplitMix& random) {
  int best = int(random.bounded(std::uint64_t(population)));
  for (int i = 1; i < pressure; ++i) {
    const int c = int(random.bounded(std::uint64_t(population)));
    if (selected_over(k, c, best)) best = c;
  }
  return best;
}

GHD inline bool site_matches(const PhaseView& g, const CRow& t, int i, int nt, int scope) {
  const int s = t.site[i];
  if (s < 0) return false;
  const int sc = site_scope(g, s);
  if (sc & kFixedSite) return false;
  return nt < 0 || (site_nt(g, s) == nt && sc == scope);
}

GHD inline void copy_row(const CRow& from, Row to, int* to_n, int* to_nc) {
  for (int i = 0; i < from.n; ++i) { to.origin[i] = from.origin[i]; to.site[i] = from.site[i]; }
  for (int i = 0; i < from.nc; ++i) to.cval[i] = from.cval[i];
  *to_n = from.n;
  *to_nc = from.nc;
}

struct SlotInfo {
  int root, scope, max_depth, logical_depth;
  int max_variable_nodes;
  int row_capacity;  // physical row size
  int statement;     // the phase is a statement program: variation checks definite assignment
};

// Definite assignment of grammar locals (phase_compile.cuh:170-200): a local
// is readable after an assign to it, inside a for_range body that names it,
// and after an if_stmt only when both branches assigned it; a for_range
// defines nothing after its body. `OriginAt(i)` gives node i's origin in
// prefix order; `Frames` holds the open nodes (at most 64).
struct DefFrame { int origin, seen; unsigned saved, branch; };
template <class OriginAt, class Frames>
GHD inline bool locals_defined_with(const PhaseView& g, int n, OriginAt origin_at, Frames& frames) {
  int depth = 0;
  unsigned mask = 0;
  const auto bit = [](int operand) { return operand >= 0 && operand < 16 ? 1u << operand : 0u; };
  for (int i = 0; i < n; ++i) {
    const int o = origin_at(i);
    const PNode& node = g.expr[o].node;
    if (node.kind == TK_VAR && !(mask & bit(node.operand))) return false;
    if (node.arity > 0) {
      if (depth == 64) return false;
      frames.set(depth++, DefFrame{o, 0, mask, 0});
      continue;
    }
    while (depth) {
      DefFrame f = frames.get(depth - 1);
      const PNode& owner = g.expr[f.origin].node;
      ++f.seen;
      if (owner.kind == TK_FOR_RANGE && f.seen == 1) {
        f.saved = mask;
        mask |= bit(owner.operand);
      } else if (owner.kind == TK_IF && owner.type == kStatementType) {
        if (f.seen == 1) f.saved = mask;
        else if (f.seen == 2) { f.branch = mask; mask = f.saved; }
        else mask &= f.branch;
      }
      if (f.seen < owner.arity) { frames.set(depth - 1, f); break; }
      if (owner.kind == TK_ASSIGN) mask |= bit(owner.operand);
      else if (owner.kind == TK_FOR_RANGE) mask = f.saved;
      --depth;
    }
  }
  return depth == 0;
}
struct LocalDefFrames {
  DefFrame f[64];
  GHD DefFrame get(int k) const { return f[k]; }
  GHD void set(int k, const DefFrame& x) { f[k] = x; }
};
GHD inline bool locals_defined(const PhaseView& g, const CRow& t) {
  LocalDefFrames frames;
  const auto origin_at = [&](int i) { return t.origin[i]; };
  return locals_defined_with(g, t.n, origin_at, frames);
}


// Genotype hash: a wrapping sum of one mixed term per node (origin and
// position) and per constant (position, tag and the payload that equality
// compares), so a wavefront can add the terms in any order. Only a lookup
// key: equality is always checked in full.
GHD inline std::uint64_t row_hash_node(int i, int origin) {
  return splitmix_mix(std::uint64_t(std::uint32_t(origin)) << 24 ^ std::uint64_t(i));
}
GHD inline std::uint64_t row_hash_constant(int i, const Val& v) {
  const std::uint64_t payload = v.tag == TAG_BOOL ? std::uint64_t(v.b) : std::uint64_t(v.i);
  return splitmix_mix(splitmix_mix(payload) ^ (std::uint64_t(i) << 8 | v.tag) ^ 0x5851f42d4c957f2dULL);
}
GHD inline std::uint64_t row_hash_counts(int n, int nc) { return splitmix_mix(std::uint64_t(std::uint32_t(n)) << 32 | std::uint32_t(nc)); }
GHD inline std::uint64_t row_hash(const CRow& r) {
  std::uint64_t h = row_hash_counts(r.n, r.nc);
  for (int i = 0; i < r.n; ++i) h += row_hash_node(i, r.origin[i]);
  for (int i = 0; i < r.nc; ++i) h += row_hash_constant(i, r.cval[i]);
  return h;
}
// Genotype identity used for completed-result reuse (parent_fitness_cache.cuh:11-58).
GHD inline bool rows_equal(const CRow& a, const CRow& b) {
  if (a.n != b.n || a.nc != b.nc) return false;
  for (int i = 0; i < a.n; ++i) if (a.origin[i] != b.origin[i]) return false;
  for (int i = 0; i < a.nc; ++i) {
    const Val x = a.cval[i], y = b.cval[i];
    if (x.tag != y.tag || (x.tag == TAG_BOOL ? x.b != y.b : x.i != y.i)) return false;
  }
  return true;
}

// ---- lowering (expression phases) ----
// Instruction of this engine's VM: opcode, builtin argc, static stack height
// before the instruction, fuel charge, operand.
struct Ins {
  std::uint8_t op;
  std::uint8_t argc;
  std::uint8_t sp;
  std::uint8_t fuel;
  std::int32_t a;
};
static_assert(sizeof(Ins) == 8, "Ins is 8 bytes");
constexpr int kMaxStack = 64;

// Limits of the typed straight-line VM path (vm_fast in vm.hpp).
constexpr int kFastStack = 16, kFastCode = 128, kFastPool = 128, kFastInputs = 4;
// Limits of the typed looping VM path (vm_loop): hidden locals (one
// traversal's temporaries; sequential traversals reuse them), code and
// constants held in registers.
constexpr int kLoopStack = 6, kLoopLocals = 9, kLoopCode = 256, kLoopPool = 256;
constexpr int kLoopSlots = kLoopStack + kLoopLocals;  // LDS slots per lane of the looping path
// Tags of a traversal's fixed instruction sequences, held in the argc byte
// of their first instruction (a LOAD or STORE, whose argc is otherwise 0);
// the low bits give the direction (1 forward, 2 reverse).
constexpr int kWalkTest = 0x80, kWalkNext = 0x90;
// Likewise a for_range loop's 6-instruction test/bind sequence at its head
// (LOAD counter, LOAD bound, LT, JIF end, LOAD counter, STORE variable) and
// its 5-instruction advance (LOAD counter, PUSH 1, ADD, STORE counter, JMP head).
constexpr int kRangeTest = 0xA0, kRangeNext = 0xB0;
// A LOAD or PUSH_CONST whose next instruction is a binary operation taking it
// as its second operand, entered by no jump: the pair runs as one step.
constexpr int kFuse = 0xC0;
GHD inline bool fusable_binary(int op) {
  return op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_MOD || (op >= OP_LT && op <= OP_NE);
}
// `fast`: every instruction and operand is certified for the typed
// straight-line path (no jumps, every value statically Int or Bool, integer
// builtins only, within the kFast* limits).
// `locals`: evaluator locals plus the hidden temporaries of traversals.
// `loop`: the same certificate for programs with jumps, hidden locals or
// IntList views (length and indexing only), for the typed looping path.
struct LowerResult { int status, code, constants, stack_bound, result_type, jumps, fast, locals, loop; };
GHD inline bool fast_type(int tag) { return tag == TAG_INT || tag == TAG_BOOL; }
GHD inline bool loop_type(int tag) { return fast_type(tag) || tag == TAG_INT_LIST; }
// a0..a2: operand tags (unused ones ignored); passed by value so the callers'
// frames stay in registers.
GHD inline bool fast_operation(int opcode, int builtin, int arity, int a0, int a1, int a2) {
  const auto all_int = [&] {
    return arity >= 1 && arity <= 3 && a0 == TAG_INT && (arity < 2 || a1 == TAG_INT) && (arity < 3 || a2 == TAG_INT);
  };
  switch (opcode) {
    case OP_NEG: case OP_CHECK_INT: return a0 == TAG_INT;
    case OP_NOT: return a0 == TAG_BOOL;
    case OP_ADD: case OP_SUB: case OP_MUL: case OP_MOD: case OP_LT: case OP_LE: case OP_GT: case OP_GE: return all_int();
    case OP_EQ: case OP_NE: return a0 == a1 && fast_type(a0);
    case OP_CALL_BUILTIN:
      switch (builtin) {
        case B_ABS: return arity == 1 && all_int();
        case B_MIN: case B_MAX: case B_IDIV0: case B_IMOD0: case B_TRUNC_DIV0: case B_TRUNC_MOD0: return arity == 2 && all_int();
        case B_CLIP: return arity == 3 && all_int();
        default: return false;
      }
    default: return false;
  }
}
// Operations of the looping path: those of the straight-line path plus the
// length and indexing of an IntList view.
GHD inline bool loop_operation(int opcode, int builtin, int arity, int a0, int a1, int a2) {
  if (opcode == OP_CALL_BUILTIN) {
    if (builtin == B_LEN) return arity == 1 && a0 == TAG_INT_LIST;
    if (builtin == B_INDEX || builtin == B_INDEX_NONNEG || builtin == B_INDEX_CLAMPED)
      return arity == 2 && a0 == TAG_INT_LIST && a1 == TAG_INT;
  }
  return fast_operation(opcode, builtin, arity, a0, a1, a2);
}

// An open operation of the post-order walk: its node, operands seen so far,
// pending jump sites, the stack height at its start and its operand types.
struct Frame { int origin, seen, first_jump, second_jump, base_stack; std::uint8_t args[3]; };
struct LocalFrames {
  using Def = LocalDefFrames;  // frames of the definite-assignment check
  Frame f[kMaxStack];
  GHD Frame get(int k) const { return f[k]; }
  GHD void set(int k, const Frame& x) { f[k] = x; }
};

// Lowering state of traversals (compile_phase_tree's temporaries, binder
// environment and open traversals, phase_compile.cuh:241-276): hidden
// temporaries are allocated after the evaluator locals and recycled LIFO
// when their traversal completes; binder numbers map to the temporaries
// holding element, index and accumulator inside a body.
struct Walk {
  int live_begin, xs, start, begin, limit, length, counter, element, index, state, loop_head, done_jump;
  int saved_slot[3], saved_type[3];
};
// An open for_range (phase_compile.cuh:251): its bound and counter
// temporaries, loop head, exit jump, the pool index of its constant 1 and
// the live temporaries at its start.
struct Loop { int bound, counter, head, end_jump, one, live_begin; };
// An open let (phase_compile.cuh:268): its binder number, value slot, the
// live temporaries at its start and the binder's outer slot and type.
struct LetScope { int index, slot, live_begin, saved_slot, saved_type; };
// The lowering sets every field before use (the record may live in LDS, one
// per wavefront, where all lanes store the same values).
template <bool kWalks> struct WalkStore {};
template <> struct WalkStore<true> {
  int live_count, free_count, walk_depth, loop_depth, let_depth;
  unsigned named_mask;  // grammar locals given a slot (on first use)
  signed char live[64], freed[64];
  signed char env_slot[16];
  unsigned char env_type[16];
  signed char named[16];
  Walk walks[8];
  Loop loops[16];
  LetScope lets[16];
};

// Post-order lowering of an expression phase plus the program terminator,
// with the reference's jump shapes for if/and/or. Constants are emitted in
// first-use order (one pool entry per constant node). `OriginAt(i)` gives the
// origin of node i in prefix order. kWalks is the instantiation with hidden
// locals. It adds lexical reads and the traversal lowering of
// phase_compile.cuh:444-542,596-637,730-740 without source fuel profiles
// (each emitted instruction costs one, no operand is aliased and begin is a
// stored temporary), `let` (its value slot allocated before the value
// expression, phase_compile.cuh:662-667,802-815), and statement programs (expected type
// kStatementType, phase_compile.cuh:589-594,660-667,766-860): grammar locals
// take a slot on first use, for_range keeps a bound and a counter temporary,
// and no terminator follows the code.
template <class Frames, class OriginAt, bool kWalks = false>
GHD inline LowerResult lower_phase_with(const PhaseView& g, const CRow& t, OriginAt origin_at, const std::uint8_t* local_types,
                                        int local_count, int expected_type, Ins* code, int code_capacity, Val* pool,
                                        int pool_capacity, Frames& frames, [[maybe_unused]] WalkStore<kWalks>& ws) {
  LowerResult out{ST_INVALID, 0, 0, 0, TAG_INVALID, 0, 0, local_count, 0};
  if (t.n <= 0 || t.n > 4096) return out;
  if constexpr (kWalks) {
    ws.live_count = ws.free_count = ws.walk_depth = ws.loop_depth = ws.let_depth = 0;
    ws.named_mask = 0;
    for (int i = 0; i < 16; ++i) { ws.env_slot[i] = -1; ws.env_type[i] = TAG_INVALID; }
  }
  bool fast = local_count <= kFastInputs;
  bool typed = local_count <= kFastInputs;  // every value Int, Bool or an IntList view; loop_operation only
  int depth = 0, cursor = 0, stack = 0, cconst = 0;
  int completed = TAG_INVALID;
  bool have_value = false;
  int last_op = -1;  // the previous instruction's opcode and argc (fusion)
  int last_argc = 0;
  const auto emit = [&](int op, int a, int argc, unsigned fuel) -> bool {
    if (out.code >= code_capacity) { out.status = ST_CAPACITY; return false; }
    if (fuel > 255 || stack < 0 || stack > kMaxStack) { out.status = ST_UNSUPPORTED; return false; }
    code[out.code++] = Ins{std::uint8_t(op), std::uint8_t(argc), std::uint8_t(stack), std::uint8_t(fuel), a};
    out.jumps |= op == OP_JMP || op == OP_JMP_IF_FALSE || op == OP_JMP_IF_TRUE;
    last_op = op;
    last_argc = argc;
    return true;
  };
  const auto grow = [&](int by) { stack += by; if (stack > out.stack_bound) out.stack_bound = stack; };
  // The last position a jump targets when it is the next instruction's: a
  // binary operation there is not fused with the instruction before it.
  int target_at = -1;
  const auto patch = [&](int jump) {
    code[jump].a = out.code;
    target_at = out.code;
  };
  // Traversal emission helpers: one unit of fuel per instruction.
  const auto put = [&](int op, int a, int argc, int delta) -> bool {
    if (!emit(op, a, argc, 1)) return false;
    if (delta > 0) grow(delta);
    else stack += delta;
    return true;
  };
  const auto put_load = [&](int slot) { return put(OP_LOAD, slot, 0, 1); };
  const auto put_store = [&](int slot) { return put(OP_STORE, slot, 0, -1); };
  const auto put_int = [&](std::int64_t v) -> bool {
    if (out.constants >= pool_capacity) { out.status = ST_CAPACITY; return false; }
    pool[out.constants] = val_int(v);
    if (!put(OP_PUSH_CONST, out.constants, 0, 1)) return false;
    ++out.constants;
    return true;
  };
  const auto put_jump = [&](int op, int delta) -> int {
    const int at = out.code;
    return put(op, 0, 0, delta) ? at : -1;
  };
  const auto temporary = [&]() -> int {
    int slot = -1;
    if constexpr (kWalks) {
      if (ws.free_count) slot = ws.freed[--ws.free_count];
      else if (out.locals >= 64) out.status = ST_CAPACITY;
      else slot = out.locals++;
      if (slot >= 0) ws.live[ws.live_count++] = static_cast<signed char>(slot);
    }
    return slot;
  };
  // The slot of grammar local `operand`, allocated on first use and never
  // freed; a negative operand names evaluator input slot -1 - operand.
  const auto variable_slot = [&](int operand) -> int {
    int slot = -1;
    if constexpr (kWalks) {
      if (operand < 0) return -1 - operand < local_count ? -1 - operand : -1;
      if (operand >= 16) return -1;
      if (!(ws.named_mask >> operand & 1u)) {
        if (out.locals >= 64) { out.status = ST_CAPACITY; return -1; }
        ws.named[operand] = static_cast<signed char>(out.locals++);
        ws.named_mask |= 1u << operand;
      }
      slot = ws.named[operand];
    }
    return slot;
  };
  const bool statement = expected_type == kStatementType;
  int returned = TAG_INVALID;  // the value type of the program's returns
  if (statement) {
    if (!kWalks) { out.status = ST_UNSUPPORTED; return out; }
    typename Frames::Def def;
    if (!locals_defined_with(g, t.n, origin_at, def)) return out;
    fast = false;
  }
  // Code after child `c` of traversal `w` (source, start, [begin, end,]
  // initial accumulator, body) completed (walk_after without aliases).
  const auto walk_after = [&](Walk& w, const PNode& node, int c) -> bool {
    if constexpr (kWalks) {
      const bool ranged = node.arity == 6, reverse = node.opcode == 2;
      if (c == 0 && !put_store(w.xs)) return false;
      if (c == 1 && !put_store(w.start)) return false;
      if (ranged && c == 2 && !put_store(w.begin)) return false;
      if (ranged && c == 3 && !put_store(w.limit)) return false;
      if (c == (ranged ? 3 : 1)) {
        const int checked[3] = {w.start, w.begin, w.limit};
        for (int role = 0; role < (ranged ? 3 : 1); ++role)
          if (!put_load(checked[role]) || !put(OP_CHECK_INT, 0, 0, 0) || !put_store(checked[role])) return false;
        if (!put_load(w.xs) || !put(OP_CALL_BUILTIN, B_LEN, 1, 0) || !put_store(w.length)) return false;
        if (ranged) {
          for (int k = 0; k < 2; ++k) {  // clamp begin, then end, to [0, length]
            const int slot = k ? w.limit : w.begin;
            if (!put_load(slot) || !put_int(0) || !put(OP_LT, 0, 0, -1)) return false;
            const int nonnegative = put_jump(OP_JMP_IF_FALSE, -1);
            if (nonnegative < 0 || !put_int(0) || !put_store(slot)) return false;
            const int clamped = put_jump(OP_JMP, 0);
            if (clamped < 0) return false;
            patch(nonnegative);
            if (!put_load(slot) || !put_load(w.length) || !put(OP_GT, 0, 0, -1)) return false;
            const int above = put_jump(OP_JMP_IF_FALSE, -1);
            if (above < 0 || !put_load(w.length) || !put_store(slot)) return false;
            patch(clamped);
            patch(above);
          }
        } else {
          if (!put_int(0) || !put_store(w.begin)) return false;
          if (!put_load(w.length) || !put_store(w.limit)) return false;
        }
      }
      if (c == node.arity - 2) {
        if (!put_store(w.state)) return false;
        if (!put_load(reverse ? w.limit : w.begin)) return false;
        if (reverse && (!put_int(1) || !put(OP_SUB, 0, 0, -1))) return false;
        if (!put_store(w.counter)) return false;
        w.loop_head = out.code;
        if (!put_load(w.counter) || !put_load(reverse ? w.begin : w.limit) || !put(reverse ? OP_GE : OP_LT, 0, 0, -1))
          return false;
        // The loop head starts the fixed 12-instruction test/read/bind/index
        // sequence: tagged in the otherwise unused argc byte of its LOAD for
        // the looping VM (kWalkTest | direction).
        code[w.loop_head].argc = std::uint8_t(kWalkTest | (reverse ? 2 : 1));
        w.done_jump = put_jump(OP_JMP_IF_FALSE, -1);
        if (w.done_jump < 0) return false;
        if (!put_load(w.xs) || !put_load(w.counter) || !put(OP_CALL_BUILTIN, B_INDEX, 2, -1)) return false;
        if (!put_store(w.element)) return false;
        if (!put_load(w.start) || !put_load(w.counter) || !put(OP_ADD, 0, 0, -1) || !put_store(w.index)) return false;
        const int slots[3] = {w.element, w.index, w.state};
        const int types[3] = {TAG_INT, TAG_INT, node.type};
        for (int k = 0; k < 3; ++k) {
          const int number = node.operand + k;
          w.saved_slot[k] = ws.env_slot[number];
          w.saved_type[k] = ws.env_type[number];
          ws.env_slot[number] = static_cast<signed char>(slots[k]);
          ws.env_type[number] = static_cast<unsigned char>(types[k]);
        }
      }
      if (c == node.arity - 1) {
        for (int k = 2; k >= 0; --k) {
          ws.env_slot[node.operand + k] = static_cast<signed char>(w.saved_slot[k]);
          ws.env_type[node.operand + k] = static_cast<unsigned char>(w.saved_type[k]);
        }
        const int update = out.code;
        if (!put_store(w.state)) return false;
        if (!put_load(w.counter) || !put_int(1) || !put(reverse ? OP_SUB : OP_ADD, 0, 0, -1) || !put_store(w.counter))
          return false;
        if (!emit(OP_JMP, w.loop_head, 0, 1)) return false;
        // The fixed 6-instruction update/advance/repeat sequence.
        code[update].argc = std::uint8_t(kWalkNext | (reverse ? 2 : 1));
        patch(w.done_jump);
        if (!put_load(w.state)) return false;
      }
      return true;
    }
    return false;
  };
  for (;;) {
    if (!have_value) {
      if (cursor >= t.n) return out;
      const int o = origin_at(cursor++);
      const PExpr& x = g.expr[o];
      const PNode& node = x.node;
      if (node.kind == TK_CONSTANT) {
        if (cconst >= t.nc) return out;
        const Val& src = t.cval[cconst++];
        Val v{};  // fields only: the padding never leaves the row
        v.i = src.i;
        v.b = src.b;
        v.tag = src.tag;
        if (v.tag == TAG_INVALID || v.tag == TAG_FALLBACK) { out.status = ST_UNSUPPORTED; return out; }
        if (out.constants >= pool_capacity) { out.status = ST_CAPACITY; return out; }
        pool[out.constants] = v;
        fast = fast && fast_type(v.tag);
        typed = typed && fast_type(v.tag);
        if (!emit(OP_PUSH_CONST, out.constants, 0, node.fuel)) return out;
        ++out.constants;
        grow(1);
        completed = v.tag;
        have_value = true;
      } else if (kWalks && node.kind == TK_BOUND) {
        if constexpr (kWalks) {
          const int b = node.operand;
          if (node.arity != 0 || b < 0 || b >= 16 || ws.env_slot[b] < 0) return out;
          completed = ws.env_type[b];
          if (node.type != TAG_INVALID && node.type != completed) return out;
          if (!emit(OP_LOAD, ws.env_slot[b], 0, node.fuel)) return out;
          grow(1);
          fast = false;
          typed = typed && loop_type(completed);
          have_value = true;
        }
      } else if (kWalks && node.kind == TK_TRAVERSE) {
        if constexpr (kWalks) {
          if (node.arity != 4 && node.arity != 6) return out;
          if (node.operand < 0 || node.operand + 2 >= 16) return out;
          if (depth == 64 || ws.walk_depth == 8) { out.status = ST_CAPACITY; return out; }
          Walk& w = ws.walks[ws.walk_depth++];
          w.live_begin = ws.live_count;
          // Temporaries in the native order; all stored, none aliased.
          w.xs = temporary();
          w.start = temporary();
          w.begin = temporary();
          w.limit = temporary();
          w.length = temporary();
          w.counter = temporary();
          w.element = temporary();
          w.index = temporary();
          w.state = temporary();
          if (w.xs < 0 || w.start < 0 || w.begin < 0 || w.limit < 0 || w.length < 0 || w.counter < 0 || w.element < 0 ||
              w.index < 0 || w.state < 0)
            return out;
          fast = false;
          frames.set(depth++, Frame{o, 0, -1, -1, stack, {TAG_INVALID, TAG_INVALID, TAG_INVALID}});
          continue;
        }
      } else if (kWalks && node.kind == TK_VAR) {
        // A read of a grammar local.
        if (node.arity != 0 || node.operand < 0) return out;
        if (node.type == TAG_INVALID || node.type == TAG_FALLBACK || node.type == kStatementType) {
          out.status = ST_UNSUPPORTED;
          return out;
        }
        const int slot = variable_slot(node.operand);
        if (slot < 0 || !emit(OP_LOAD, slot, 0, node.fuel)) return out;
        grow(1);
        completed = node.type;
        typed = typed && loop_type(completed);
        have_value = true;
      } else if (kWalks && node.kind == TK_NIL) {
        if (node.arity != 0) return out;
        completed = kStatementType;
        have_value = true;
      } else if (kWalks && (node.kind == TK_PROGRAM || node.kind == TK_ASSIGN || node.kind == TK_RETURN ||
                            node.kind == TK_BLOCK || node.kind == TK_FOR_RANGE || node.kind == TK_LET)) {
        if constexpr (kWalks) {
          const bool pair = node.kind == TK_BLOCK || node.kind == TK_FOR_RANGE || node.kind == TK_LET;
          if (node.arity != (pair ? 2 : 1)) return out;
          if (depth == 64) { out.status = ST_CAPACITY; return out; }
          if (node.kind == TK_FOR_RANGE) {
            // The bound temporary precedes the bound expression's own.
            if (ws.loop_depth == 16) { out.status = ST_CAPACITY; return out; }
            Loop& loop = ws.loops[ws.loop_depth++];
            loop.live_begin = ws.live_count;
            loop.bound = temporary();
            if (loop.bound < 0) return out;
          } else if (node.kind == TK_LET) {
            // The value slot precedes the value expression's temporaries.
            if (node.operand < 0 || node.operand >= 16) return out;
            if (ws.let_depth == 16) { out.status = ST_CAPACITY; return out; }
            LetScope& scope = ws.lets[ws.let_depth++];
            scope.index = node.operand;
            scope.live_begin = ws.live_count;
            scope.slot = temporary();
            if (scope.slot < 0) return out;
            fast = false;  // its STORE is outside the straight-line VM, read or not
          }
          frames.set(depth++, Frame{o, 0, -1, -1, stack, {TAG_INVALID, TAG_INVALID, TAG_INVALID}});
          continue;
        }
      } else if (node.kind == TK_LOCAL) {
        if (node.operand < 0 || node.operand >= local_count) return out;
        completed = local_types[node.operand];
        if (completed == TAG_INVALID || completed == TAG_FALLBACK) { out.status = ST_UNSUPPORTED; return out; }
        fast = fast && fast_type(completed);
        typed = typed && loop_type(completed);
        if (!emit(OP_LOAD, node.operand, 0, node.fuel)) return out;
        grow(1);
        have_value = true;
      } else {
        if (node.arity < 1 || node.arity > 3) return out;
        if (node.kind == TK_IF) { if (node.arity != 3) return out; }
        else if (node.kind == TK_AND || node.kind == TK_OR) { if (node.arity != 2) return out; }
        else if (node.kind != TK_OPERATION) { out.status = ST_UNSUPPORTED; return out; }
        if (depth == 64) { out.status = ST_CAPACITY; return out; }
        frames.set(depth++, Frame{o, 0, -1, -1, stack, {TAG_INVALID, TAG_INVALID, TAG_INVALID}});
        continue;
      }
    }
    if (depth == 0) {
      if (cursor != t.n || stack != (statement ? 0 : 1) || completed != expected_type) return out;
      if (out.stack_bound > kMaxStack) { out.status = ST_CAPACITY; return out; }
      // program terminator (whole-program expression entry); a statement
      // program that runs off its end fails with ValueError
      if (!statement && !emit(OP_RETURN, 0, 0, 1)) return out;
      out.status = ST_OK;
      // A statement program's value is that of its returns.
      out.result_type = statement ? returned : completed;
      out.fast = fast && !out.jumps && fast_type(completed) && out.stack_bound <= kFastStack && out.code <= kFastCode &&
                 out.constants <= kFastPool;
      out.loop = !out.fast && typed && fast_type(out.result_type) && out.stack_bound <= kLoopStack &&
                 out.code <= kLoopCode && out.constants <= kLoopPool && out.locals - local_count <= kLoopLocals;
      return out;
    }
    Frame f = frames.get(depth - 1);
    const PNode& node = g.expr[f.origin].node;
    // Constant indices keep the frame in registers.
    if (f.seen == 0) f.args[0] = std::uint8_t(completed);
    else if (f.seen == 1) f.args[1] = std::uint8_t(completed);
    else if (f.seen == 2) f.args[2] = std::uint8_t(completed);
    ++f.seen;
    if constexpr (kWalks) {
      if (node.kind == TK_TRAVERSE) {
        Walk& w = ws.walks[ws.walk_depth - 1];
        const int c = f.seen - 1;
        const int want = c == 0 ? TAG_INT_LIST : c >= node.arity - 2 ? node.type : TAG_INT;
        if (completed != want || !walk_after(w, node, c)) return out;
        if (f.seen < node.arity) { frames.set(depth - 1, f); have_value = false; continue; }
        while (ws.live_count > w.live_begin) ws.freed[ws.free_count++] = ws.live[--ws.live_count];
        --ws.walk_depth;
        completed = node.type;
        --depth;
        continue;
      }
    }
    const bool branch = node.kind == TK_IF || node.kind == TK_AND || node.kind == TK_OR;
    if (branch && f.seen == 1) {
      if (completed != TAG_BOOL) return out;
      f.first_jump = out.code;
      if (!emit(node.kind == TK_OR ? OP_JMP_IF_TRUE : OP_JMP_IF_FALSE, 0, 0, node.fuel)) return out;
      --stack;
    } else if (node.kind == TK_IF && f.seen == 2) {
      f.second_jump = out.code;
      if (!emit(OP_JMP, 0, 0, node.merge_fuel)) return out;
      patch(f.first_jump);
      stack = f.base_stack;
    }
    if constexpr (kWalks) {
      if (node.kind == TK_FOR_RANGE && f.seen == 1) {
        // Validate the stored bound (a non-Int or negative bound reaches
        // NEG(true), a runtime error), then count from zero; every
        // instruction costs one.
        Loop& loop = ws.loops[ws.loop_depth - 1];
        if (completed == kStatementType) return out;
        // On an Int bound IS_INT is always true and NEG(true) is reached only
        // by a negative bound; the looping path then raises TypeError for the
        // NEG that argc 1 marks.
        typed = typed && completed == TAG_INT;
        if (!put_store(loop.bound)) return out;
        if (!put_load(loop.bound) || !put(OP_CALL_BUILTIN, B_IS_INT, 1, 0)) return out;
        const int bad = put_jump(OP_JMP_IF_FALSE, -1);
        if (bad < 0 || !put_load(loop.bound) || !put_int(0) || !put(OP_LT, 0, 0, -1)) return out;
        const int valid = put_jump(OP_JMP_IF_FALSE, -1);
        if (valid < 0) return out;
        patch(bad);
        if (out.constants >= pool_capacity) { out.status = ST_CAPACITY; return out; }
        pool[out.constants] = val_bool(true);
        if (!put(OP_PUSH_CONST, out.constants, 0, 1)) return out;
        ++out.constants;
        if (!put(OP_NEG, 0, 1, 0)) return out;
        stack = f.base_stack;
        patch(valid);
        if (out.constants + 2 > pool_capacity) { out.status = ST_CAPACITY; return out; }
        const int zero = out.constants++;
        pool[zero] = val_int(0);
        loop.one = out.constants++;
        pool[loop.one] = val_int(1);
        loop.counter = temporary();
        const int user = variable_slot(node.operand);
        if (loop.counter < 0 || user < 0) return out;
        if (!put(OP_PUSH_CONST, zero, 0, 1) || !put_store(loop.counter)) return out;
        loop.head = out.code;
        if (!put_load(loop.counter) || !put_load(loop.bound) || !put(OP_LT, 0, 0, -1)) return out;
        loop.end_jump = put_jump(OP_JMP_IF_FALSE, -1);
        if (loop.end_jump < 0 || !put_load(loop.counter) || !put_store(user)) return out;
        code[loop.head].argc = std::uint8_t(kRangeTest);
      }
    }
    if constexpr (kWalks) {
      if (node.kind == TK_LET && f.seen == 1) {
        // The value is stored, then the body reads it through the binder.
        LetScope& scope = ws.lets[ws.let_depth - 1];
        if (completed == kStatementType) return out;
        if (!emit(OP_STORE, scope.slot, 0, node.fuel)) return out;
        stack = f.base_stack;
        scope.saved_slot = ws.env_slot[scope.index];
        scope.saved_type = ws.env_type[scope.index];
        ws.env_slot[scope.index] = static_cast<signed char>(scope.slot);
        ws.env_type[scope.index] = static_cast<unsigned char>(completed);
      }
    }
    if (f.seen < node.arity) { frames.set(depth - 1, f); have_value = false; continue; }
    if (kWalks && node.kind == TK_LET) {
      if constexpr (kWalks) {
        if (completed == kStatementType || (node.type != TAG_INVALID && node.type != completed)) return out;
        const LetScope scope = ws.lets[--ws.let_depth];
        ws.env_slot[scope.index] = static_cast<signed char>(scope.saved_slot);
        ws.env_type[scope.index] = static_cast<unsigned char>(scope.saved_type);
        while (ws.live_count > scope.live_begin) ws.freed[ws.free_count++] = ws.live[--ws.live_count];
      }
    } else if (kWalks && node.kind == TK_PROGRAM) {
      if (completed != kStatementType) return out;
    } else if (kWalks && node.kind == TK_BLOCK) {
      if (f.args[0] != kStatementType || f.args[1] != kStatementType) return out;
    } else if (kWalks && (node.kind == TK_ASSIGN || node.kind == TK_RETURN)) {
      // The grammar fixes the stored or returned type; an assigned local
      // takes its slot after its value expression.
      if (completed != node.type || completed == kStatementType) return out;
      if (node.kind == TK_ASSIGN) {
        const int slot = variable_slot(node.operand);
        if (slot < 0 || !emit(OP_STORE, slot, 0, node.fuel)) return out;
      } else {
        if (!emit(OP_RETURN, 0, 0, node.fuel)) return out;
        typed = typed && (returned == TAG_INVALID || returned == completed);
        returned = completed;
      }
      stack = f.base_stack;
      completed = kStatementType;
    } else if (kWalks && node.kind == TK_FOR_RANGE) {
      if constexpr (kWalks) {
        if (completed != kStatementType) return out;
        const Loop loop = ws.loops[--ws.loop_depth];
        const int advance = out.code;
        if (!put_load(loop.counter) || !put(OP_PUSH_CONST, loop.one, 0, 1) || !put(OP_ADD, 0, 0, -1) ||
            !put_store(loop.counter) || !emit(OP_JMP, loop.head, 0, 1))
          return out;
        code[advance].argc = std::uint8_t(kRangeNext);
        patch(loop.end_jump);
        stack = f.base_stack;
        while (ws.live_count > loop.live_begin) ws.freed[ws.free_count++] = ws.live[--ws.live_count];
      }
    } else if (node.kind == TK_IF) {
      if (f.args[1] != f.args[2]) return out;
      patch(f.second_jump);
    } else if (node.kind == TK_AND || node.kind == TK_OR) {
      if (completed != TAG_BOOL) return out;
      if (!emit(OP_NOT, 0, 0, node.fuel) || !emit(OP_NOT, 0, 0, node.fuel)) return out;
      const int finish = out.code;
      if (!emit(OP_JMP, 0, 0, node.fuel)) return out;
      patch(f.first_jump);
      stack = f.base_stack;
      // push the short-circuit constant
      if (out.constants >= pool_capacity) { out.status = ST_CAPACITY; return out; }
      pool[out.constants] = val_bool(node.kind == TK_OR);
      if (!emit(OP_PUSH_CONST, out.constants, 0, node.fuel)) return out;
      ++out.constants;
      grow(1);
      patch(finish);
    } else {
      // operation result type: the certified Int/Bool rules, else the grammar type
      int rt = TAG_INVALID;
      const std::uint8_t* args = f.args;
      switch (node.opcode) {
        case OP_NEG: case OP_CHECK_INT: rt = args[0] == TAG_INT ? TAG_INT : TAG_INVALID; break;
        case OP_NOT: rt = args[0] == TAG_BOOL ? TAG_BOOL : TAG_INVALID; break;
        case OP_ADD: case OP_SUB: case OP_MUL: case OP_MOD:
          rt = args[0] == TAG_INT && args[1] == TAG_INT ? TAG_INT : TAG_INVALID; break;
        case OP_LT: case OP_LE: case OP_GT: case OP_GE:
          rt = args[0] == TAG_INT && args[1] == TAG_INT ? TAG_BOOL : TAG_INVALID; break;
        case OP_EQ: case OP_NE:
          rt = args[0] == args[1] && (args[0] == TAG_INT || args[0] == TAG_BOOL) ? TAG_BOOL : TAG_INVALID; break;
        default: break;
      }
      if (rt == TAG_INVALID) rt = node.type;
      else if (node.type != TAG_INVALID && node.type != rt) return out;
      if (rt == TAG_INVALID || rt == TAG_FALLBACK) { out.status = ST_UNSUPPORTED; return out; }
      fast = fast && fast_operation(node.opcode, node.operand, node.arity, args[0], args[1], args[2]) && fast_type(rt);
      typed = typed && loop_operation(node.opcode, node.operand, node.arity, args[0], args[1], args[2]) && fast_type(rt);
      // A binary operation right after the LOAD or PUSH of its second
      // operand, which no jump enters between, fuses with it (kFuse).
      if (node.arity == 2 && fusable_binary(node.opcode) && (last_op == OP_LOAD || last_op == OP_PUSH_CONST) &&
          last_argc == 0 && target_at != out.code && out.code > 0)
        code[out.code - 1].argc = std::uint8_t(kFuse);
      // operands are on the stack: the instruction sees the stack with them
      if (node.opcode == OP_CALL_BUILTIN) { if (!emit(OP_CALL_BUILTIN, node.operand, node.arity, node.fuel)) return out; }
      else if (!emit(node.opcode, 0, 0, node.fuel)) return out;
      stack -= node.arity - 1;
      completed = rt;
    }
    --depth;
  }
}
template <bool kWalks = false>
GHD inline LowerResult lower_phase(const PhaseView& g, const CRow& t, const std::uint8_t* local_types, int local_count,
                                   int expected_type, Ins* code, int code_capacity, Val* pool, int pool_capacity) {
  LocalFrames frames;
  WalkStore<kWalks> ws;
  const auto origin_at = [&](int i) { return t.origin[i]; };
  return lower_phase_with<LocalFrames, decltype(origin_at), kWalks>(g, t, origin_at, local_types, local_count, expected_type, code,
                                                                    code_capacity, pool, pool_capacity, frames, ws);
}

}  // namespace gagp_amd
