// Test infrastructure only: writes known-answer vectors of the reference's
// fitness reduction. Links the reference CPU engine's own objects (built
// from its sources by ref.mk into oracle/_ref/obj) and calls its
// gagp::eval_fitness_cpu (cpp/src/runtime/cpu/fitness_cpu.cpp:57-176) on
// seeded scenarios: straight-line programs over one Int input with runtime
// errors, Bool results and timeouts, scored under absolute capped (integral
// and fractional caps), absolute discounted and exact policies, case counts
// from 1 to 4097 and reduction widths 1, 32, 100 and 512. The scenario of
// cpp/tests/parity/test_fitness_cpu_gpu_parity.cpp:2681-2695 (1024 cases,
// constant programs -1..15, 512 lanes, discounted) is included as written
// there. Prints one JSON document; fitness values are C99 hex floats.
//
//   oracle/_ref/ref_fitness > tests/golden/fitness/reduction.json
#include <cinttypes>
#include <cstdio>
#include <string>
#include <vector>

#include "gagp/core/bytecode.hpp"
#include "gagp/core/fitness_policy.hpp"
#include "gagp/runtime/cpu/fitness_cpu.hpp"

namespace {

struct Mix {
  std::uint64_t s;
  std::uint64_t next() {
    std::uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  int range(int lo, int hi) { return lo + int(next() % std::uint64_t(hi - lo + 1)); }
};

using gagp::Instr;
using gagp::Opcode;
using gagp::Value;

Instr ins(Opcode op, int a = 0, bool has_a = false) {
  Instr i;
  i.op = op;
  i.a = a;
  i.has_a = has_a;
  return i;
}

// One random program over input 0 (x): x + k, x*x - k, k mod x (an error
// when x is 0), x > k (a Bool), the constant k, or an endless loop (a
// timeout).
gagp::BytecodeProgram random_program(Mix& r) {
  gagp::BytecodeProgram p;
  p.n_locals = 1;
  const int k = r.range(-60, 60);
  p.consts.push_back(Value::from_int(k));
  switch (r.range(0, 9)) {
    case 0: case 1: case 2:
      p.code = {ins(Opcode::Load, 0, true), ins(Opcode::PushConst, 0, true), ins(Opcode::Add), ins(Opcode::Return)};
      break;
    case 3: case 4:
      p.code = {ins(Opcode::Load, 0, true), ins(Opcode::Load, 0, true), ins(Opcode::Mul), ins(Opcode::PushConst, 0, true),
                ins(Opcode::Sub), ins(Opcode::Return)};
      break;
    case 5: case 6:
      p.code = {ins(Opcode::PushConst, 0, true), ins(Opcode::Load, 0, true), ins(Opcode::Mod), ins(Opcode::Return)};
      break;
    case 7:
      p.code = {ins(Opcode::Load, 0, true), ins(Opcode::PushConst, 0, true), ins(Opcode::Gt), ins(Opcode::Return)};
      break;
    case 8:
      p.code = {ins(Opcode::PushConst, 0, true), ins(Opcode::Return)};
      break;
    default:
      p.code = {ins(Opcode::Jmp, 0, true)};
      break;
  }
  return p;
}

struct Scenario {
  std::string name;
  int cases, lanes;
  gagp::FitnessPolicy policy;
  double penalty;
  int fuel;
  bool parity_case;  // the reference parity test's own inputs
};

void print_program(const gagp::BytecodeProgram& p) {
  std::printf("{\"n_locals\":%d,\"consts\":[", p.n_locals);
  for (std::size_t i = 0; i < p.consts.size(); ++i) std::printf("%s%" PRId64, i ? "," : "", p.consts[i].i);
  std::printf("],\"code\":[");
  for (std::size_t i = 0; i < p.code.size(); ++i)
    std::printf("%s[%d,%s]", i ? "," : "", int(p.code[i].op), p.code[i].has_a ? std::to_string(p.code[i].a).c_str() : "null");
  std::printf("]}");
}

}  // namespace

int main() {
  using M = gagp::FitnessMetric;
  using X = gagp::FitnessProximity;
  const gagp::FitnessPolicy capped{M::Absolute, X::Capped}, disc{M::Absolute, X::Discounted}, exact{M::Exact};
  const std::vector<Scenario> scenarios = {
      {"parity_512_lane_discounted", 1024, 512, disc, 1000, 1000, true},
      {"c1_l1_disc", 1, 1, disc, 1000, 100, false},
      {"c1_l512_disc", 1, 512, disc, 1000, 100, false},
      {"c5_cap2p5", 5, 512, capped, 2.5, 100, false},
      {"c64_disc", 64, 512, disc, 1000, 100, false},
      {"c65_disc", 65, 512, disc, 1000, 100, false},
      {"c200_cap7p25", 200, 512, capped, 7.25, 100, false},
      {"c200_l32_disc", 200, 32, disc, 1000, 100, false},
      {"c511_disc", 511, 512, disc, 1000, 100, false},
      {"c512_disc", 512, 512, disc, 1000, 100, false},
      {"c513_disc", 513, 512, disc, 1000, 100, false},
      {"c1100_disc", 1100, 512, disc, 1000, 100, false},
      {"c1100_cap2p5", 1100, 512, capped, 2.5, 100, false},
      {"c1100_l100_disc", 1100, 100, disc, 1000, 100, false},
      {"c2048_disc", 2048, 512, disc, 1000, 60, false},
      {"c4097_disc", 4097, 512, disc, 1000, 60, false},
      {"c4097_cap7p25", 4097, 512, capped, 7.25, 60, false},
      {"c300_exact", 300, 512, exact, 1000, 100, false},
      {"c1024_exact", 1024, 512, exact, 1000, 100, false},
      {"c700_abs1000", 700, 512, capped, 1000, 100, false},
  };
  Mix r{0x5eedf00dULL};
  std::printf("{\"source\":\"gagp::eval_fitness_cpu, cpp/src/runtime/cpu/fitness_cpu.cpp:57-176\",\"scenarios\":[\n");
  for (std::size_t si = 0; si < scenarios.size(); ++si) {
    const Scenario& s = scenarios[si];
    std::vector<gagp::CaseBindings> cases(std::size_t(s.cases));
    std::vector<Value> answers;
    std::vector<gagp::BytecodeProgram> programs;
    std::vector<std::int64_t> xs;
    if (s.parity_case) {
      for (int i = 0; i < s.cases; ++i) answers.push_back(Value::from_int((i * 17) % 51 - 1));
      for (int i = -1; i < 16; ++i) {
        gagp::BytecodeProgram p;
        p.consts.push_back(Value::from_int(i));
        p.code = {ins(Opcode::PushConst, 0, true), ins(Opcode::Return)};
        programs.push_back(p);
      }
    } else {
      for (int c = 0; c < s.cases; ++c) {
        const std::int64_t x = r.range(-50, 50);
        xs.push_back(x);
        cases[std::size_t(c)].push_back(gagp::InputBinding{0, Value::from_int(x)});
        answers.push_back(Value::from_int(r.range(0, 3) == 0 ? x + 7 : r.range(-200, 200)));
      }
      for (int i = 0; i < 12; ++i) programs.push_back(random_program(r));
    }
    std::vector<unsigned> wrong;
    std::vector<gagp::CpuFitnessCaseCounts> counts;
    const auto fitness = gagp::eval_fitness_cpu(programs, cases, answers, s.fuel, s.penalty, s.lanes, s.policy, &wrong, &counts);
    std::printf("%s{\"name\":\"%s\",\"cases\":%d,\"lanes\":%d,\"metric\":%d,\"proximity\":%d,\"penalty\":%.17g,\"fuel\":%d,",
                si ? ",\n" : "", s.name.c_str(), s.cases, s.lanes, int(s.policy.metric), int(s.policy.proximity), s.penalty,
                s.fuel);
    std::printf("\"inputs\":[");
    for (std::size_t c = 0; c < xs.size(); ++c) std::printf("%s%" PRId64, c ? "," : "", xs[c]);
    std::printf("],\"answers\":[");
    for (std::size_t c = 0; c < answers.size(); ++c) std::printf("%s%" PRId64, c ? "," : "", answers[c].i);
    std::printf("],\"programs\":[");
    for (std::size_t p = 0; p < programs.size(); ++p) {
      if (p) std::printf(",");
      print_program(programs[p]);
    }
    std::printf("],\"fitness\":[");
    for (std::size_t p = 0; p < fitness.size(); ++p) std::printf("%s\"%a\"", p ? "," : "", fitness[p]);
    std::printf("],\"wrong\":[");
    for (std::size_t p = 0; p < wrong.size(); ++p) std::printf("%s%u", p ? "," : "", wrong[p]);
    std::printf("],\"counts\":[");
    for (std::size_t p = 0; p < counts.size(); ++p)
      std::printf("%s[%u,%u,%u]", p ? "," : "", counts[p].cases, counts[p].errors, counts[p].timeouts);
    std::printf("]}");
  }
  std::printf("\n]}\n");
  return 0;
}
