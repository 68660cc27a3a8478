// Test driver (CPU only): lowers seeded random programs of a task's grammar
// with the engine's host lowering (the kWalks instantiation, which the
// device kernels share), scores the bytecode on the task's train cases with
// the C oracle VM, and prints per program one JSON line with the score and
// the exported AST. tests/test_host_lowering.py compares those lines with the
// reference's own evaluation (`gagp eval --split train`) of the same ASTs.
//
//   host_lowering TASK.json N
//   host_lowering TASK.json N defs   (statement tasks: one draw per program,
//                                     its definite-assignment verdict and AST)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "export.hpp"
#include "grammar.hpp"
#include "phase.hpp"
#include "task.hpp"

extern "C" {
// oracle/vm_oracle.c
typedef struct { std::uint8_t op, has_a, has_b, pad; std::int32_t a, b; } OIns;
typedef struct {
  const OIns* code; std::int32_t code_len;
  const std::uint32_t* fuel;
  const gagp_amd::Val* consts; std::int32_t n_consts;
  std::int32_t n_locals;
} OProg;
void orc_eval_fitness(const OProg* programs, int n_programs, const gagp_amd::Val* inputs, const std::uint8_t* set, int n_cases,
                      int n_inputs, const gagp_amd::Val* answers, const std::int64_t* arena, int fuel, double penalty, int lanes,
                      int metric, int proximity, double* fitness, std::uint32_t* wrong, std::uint32_t* counts);
}

using namespace gagp_amd;

namespace {
// The fixed sequences the lowering tags for the looping VM (the argc byte of
// their first instruction), checked against the code: the VM runs each as one
// step, so the pattern must hold and no jump may enter the sequence after its
// first instruction. Returns the number of violations; `tags` counts each kind.
int tag_violations(const std::vector<Ins>& code, int n, int (&tags)[5]) {
  std::vector<char> target(std::size_t(n) + 1, 0);
  for (int i = 0; i < n; ++i)
    if (code[i].op == OP_JMP || code[i].op == OP_JMP_IF_FALSE || code[i].op == OP_JMP_IF_TRUE) target.at(std::size_t(code[i].a)) = 1;
  const auto ops = [&](int at, std::initializer_list<int> want) {
    int k = 0;
    for (const int op : want) {
      if (at + k >= n || code[at + k].op != op || (k > 0 && target[std::size_t(at + k)])) return false;
      ++k;
    }
    return true;
  };
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    const int tag = code[i].op == OP_CALL_BUILTIN || code[i].op == OP_NEG ? 0 : code[i].argc & 0xf0;
    if (tag == kFuse) {
      ++tags[0];
      bad += !((code[i].op == OP_LOAD || code[i].op == OP_PUSH_CONST) && i + 1 < n && fusable_binary(code[i + 1].op) &&
               !target[std::size_t(i + 1)]);
    } else if (tag == kRangeTest) {
      ++tags[1];
      bad += !ops(i, {OP_LOAD, OP_LOAD, OP_LT, OP_JMP_IF_FALSE, OP_LOAD, OP_STORE});
    } else if (tag == kRangeNext) {
      ++tags[2];
      bad += !(ops(i, {OP_LOAD, OP_PUSH_CONST, OP_ADD, OP_STORE, OP_JMP}) && code[i + 4].a >= 0 && code[code[i + 4].a].argc == kRangeTest);
    } else if (tag == kWalkTest) {
      ++tags[3];
      const bool reverse = (code[i].argc & 3) == 2;
      bad += !ops(i, {OP_LOAD, OP_LOAD, reverse ? OP_GE : OP_LT, OP_JMP_IF_FALSE, OP_LOAD, OP_LOAD, OP_CALL_BUILTIN, OP_STORE,
                      OP_LOAD, OP_LOAD, OP_ADD, OP_STORE});
    } else if (tag == kWalkNext) {
      ++tags[4];
      const bool reverse = (code[i].argc & 3) == 2;
      bad += !ops(i, {OP_STORE, OP_LOAD, OP_PUSH_CONST, reverse ? OP_SUB : OP_ADD, OP_STORE, OP_JMP});
    } else if (code[i].argc != 0 && code[i].op != OP_CALL_BUILTIN && !(code[i].op == OP_NEG && code[i].argc == 1)) {
      ++bad;  // an unknown tag
    }
  }
  return bad;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) { std::fprintf(stderr, "usage: host_lowering TASK.json N [defs]\n"); return 2; }
  const bool defs = argc == 4 && std::string(argv[3]) == "defs";
  const Task task = load_task(argv[1]);
  const int n = std::atoi(argv[2]);
  const Grammar g = load_grammar_file(task.grammar_path);
  const Cases cases = load_cases(task.train_path);
  const Layout layout = whole_program_layout(g);
  const PhaseGrammar ph = compile_phase(g, g.entry, g.max_depth, 7, cases.input_names);
  const PhaseView gv{ph.exprs.data(), int(ph.exprs.size()), ph.prods.data(),  int(ph.prods.size()),  ph.nts.data(),
                     int(ph.nts.size()), ph.domains.data(), int(ph.domains.size()), ph.values.data(), int(ph.values.size()),
                     ph.minimum.data(), ph.depth_count,     ph.logical_count,     ph.site_depth ? 1 : 0, ph.root};
  const int R = std::min<int>(kPhaseCapacity, int(layout.max_variable_nodes));
  const std::size_t cap = std::size_t(R);
  std::vector<int> origin(cap), site(cap);
  std::vector<Val> cval(cap);
  // The engine's per-program capacities.
  const bool statement = g.nts[g.entry].cat == Cat::Program;
  const int loops = statement ? (R + 2) / 3 : 0;
  std::vector<Ins> code(std::size_t(4 * R + 8 + 64 * ph.traversals + 23 * loops));
  std::vector<Val> pool(std::size_t(2 * R + 8 + 8 * ph.traversals + 4 * loops));
  for (int i = 0; defs && i < n; ++i) {
    const GenResult gen = generate_phase(gv, gv.root, 0, int(layout.slot_max_depth), R, 0x51ed2701ull * std::uint64_t(i + 1),
                                         origin.data(), cval.data(), site.data(), 2 + i % 6);
    if (gen.status != ST_OK) continue;
    const std::vector<int> row_origin(origin.begin(), origin.begin() + gen.nodes);
    const Ast ast = export_whole_program(g, ph, names_used(g, row_origin), row_origin,
                                         std::vector<Val>(cval.begin(), cval.begin() + gen.constants));
    json::Value line = json::make_object();
    line.obj["i"] = json::make_number(i);
    line.obj["defined"] = json::make_bool(locals_defined(gv, CRow{origin.data(), site.data(), cval.data(), gen.nodes, gen.constants}));
    line.obj["text"] = json::make_string(program_text(ast));
    line.obj["program"] = ast_json(ast);
    std::printf("%s\n", json::canonical(line).c_str());
  }
  for (int i = 0; !defs && i < n; ++i) {
    GenResult gen{ST_INVALID, 0, 0, 0};
    for (int attempt = 0; gen.status != ST_OK && attempt < 64; ++attempt)
    {
      gen = generate_phase(gv, gv.root, 0, int(layout.slot_max_depth), R, 0x9e3779b97f4a7c15ull * std::uint64_t(i + 1) + attempt,
                           origin.data(), cval.data(), site.data(), 2 + i % 6);
      // A statement program must read only assigned locals.
      if (gen.status == ST_OK && statement && !locals_defined(gv, CRow{origin.data(), site.data(), cval.data(), gen.nodes, gen.constants}))
        gen.status = ST_INVALID;
    }
    if (gen.status != ST_OK) continue;
    const CRow row{origin.data(), site.data(), cval.data(), gen.nodes, gen.constants};
    const LowerResult lr = lower_phase<true>(gv, row, cases.input_types.data(), int(cases.input_types.size()),
                                             statement ? kStatementType : int(g.nts[g.entry].type), code.data(), int(code.size()),
                                             pool.data(), int(pool.size()));
    if (lr.status != ST_OK) { std::printf("{\"i\":%d,\"lowering_status\":%d}\n", i, lr.status); continue; }
    std::vector<OIns> oc;
    std::vector<std::uint32_t> fuel;
    for (int k = 0; k < lr.code; ++k) {
      const Ins& x = code[std::size_t(k)];
      const bool has_a = x.op == OP_PUSH_CONST || x.op == OP_LOAD || x.op == OP_STORE || x.op == OP_JMP ||
                         x.op == OP_JMP_IF_FALSE || x.op == OP_JMP_IF_TRUE || x.op == OP_CALL_BUILTIN;
      oc.push_back(OIns{x.op, std::uint8_t(has_a), std::uint8_t(x.op == OP_CALL_BUILTIN), 0, x.a, x.argc});
      fuel.push_back(x.fuel);
    }
    const OProg p{oc.data(), std::int32_t(oc.size()), fuel.data(), pool.data(), lr.constants, lr.locals};
    double fit = 0;
    std::uint32_t wrong = 0, counts[3] = {0, 0, 0};
    orc_eval_fitness(&p, 1, cases.inputs.data(), cases.set.data(), cases.n, int(cases.input_names.size()), cases.expected.data(),
                     cases.arena.data(), int(g.fuel), task.case_cap, 512, task.policy.metric, task.policy.proximity, &fit, &wrong,
                     counts);
    const std::vector<int> row_origin(origin.begin(), origin.begin() + gen.nodes);
    const Ast ast = export_whole_program(g, ph, names_used(g, row_origin), row_origin,
                                         std::vector<Val>(cval.begin(), cval.begin() + gen.constants));
    json::Value line = json::make_object();
    line.obj["i"] = json::make_number(i);
    line.obj["fitness"] = json::make_number(fit);
    line.obj["wrong"] = json::make_number(wrong);
    line.obj["runtime_errors"] = json::make_number(counts[1]);
    line.obj["timeouts"] = json::make_number(counts[2]);
    line.obj["code"] = json::make_number(lr.code);
    line.obj["locals"] = json::make_number(lr.locals);
    line.obj["stack"] = json::make_number(lr.stack_bound);
    line.obj["loop"] = json::make_number(lr.loop);
    int tags[5] = {0, 0, 0, 0, 0};
    line.obj["tag_violations"] = json::make_number(tag_violations(code, lr.code, tags));
    json::Value t = json::make_array();
    for (const int k : tags) t.arr.push_back(json::make_number(k));
    line.obj["tags"] = t;
    line.obj["text"] = json::make_string(program_text(ast));
    line.obj["program"] = ast_json(ast);
    std::printf("%s\n", json::canonical(line).c_str());
  }
  return 0;
}
