// Device bytecode VM: one wavefront (64 lanes) per program, one case per lane
// per tile. The instruction stream is wave-uniform: straight-line programs run
// ip in lockstep; programs with jumps use min-ip scheduling (each step runs the
// smallest pending ip, for the lanes that are at it), so every fetched
// instruction and its static stack height are scalar values. The operand
// stack lives in LDS, slot-major, one 8-byte payload and one tag byte per
// lane and slot, so every stack access is conflict-free.
//
// Instruction semantics restate cpp/src/runtime/cpu/execute_bytecode_cpu.cpp:
// 194-614 (fuel charged and checked first, the first error ends the case)
// and the scalar builtins of cpp/src/runtime/cpu/builtins_cpu.cpp:130-217,
// 340-375,493-500.
#pragma once
#include "core.hpp"
#include "phase.hpp"

namespace gagp_amd {

struct ProgView {
  GAS const Ins* code;
  GAS const Val* pool;
  int len;
  int n_locals;
  int has_jumps;
  int fast;      // certified for vm_fast (phase.hpp fast_operation)
  int n_pool;    // constant-pool entries
  int result_tag;  // static result type of a fast program
};

struct CaseData {
  GAS const std::int64_t* in_val;  // [n_inputs][C]
  GAS const std::uint8_t* in_tag;  // [n_inputs][C] (TAG_INVALID: unbound)
  GAS const std::uint8_t* in_b;    // [n_inputs][C]
  GAS const std::int64_t* exp_val; // [C]
  GAS const std::uint8_t* exp_tag; // [C]
  GAS const std::uint8_t* exp_b;   // [C]
  GAS const std::int64_t* arena;   // IntList elements
  GAS const int* order;            // [C] a permutation of the cases (visiting order of the looping path)
  int C, n_inputs;
};

// Per-wave LDS stack, slot-major: payload[slot*64 + lane], tag[...], flag[...].
// Slots [slots, slots + temps) hold the hidden locals after the evaluator
// inputs (local index n_inputs + k in slot slots + k).
struct WaveStack {
  std::int64_t* v;
  std::uint8_t* t;
  std::uint8_t* b;
  int slots;
  int temps;
  // The last `shadow` hidden slots shadow the evaluator inputs of programs
  // that store into them (statement targets): a store lands there and later
  // loads of that input read it; 0 when inputs are read-only.
  int shadow;
};

#if defined(__HIPCC__)
__device__ inline int wave_min(int x) {
  for (int off = 32; off > 0; off >>= 1) {
    const int y = __shfl_xor(x, off, 64);
    x = y < x ? y : x;
  }
  return x;
}
#endif

// Smallest value of x over the wave for 0 <= x < 2^kBits, by a bitwise
// search with ballots (scalar compares, no LDS round trips): bit b of the
// minimum is 0 iff some lane lies in [m, m + 2^b - 1] for the bits m above it.
#if defined(__HIPCC__)
template <int kBits>
__device__ inline int wave_min_bits(int x) {
  int m = 0;
#pragma unroll
  for (int b = kBits - 1; b >= 0; --b)
    if (__ballot(x <= (m | ((1 << b) - 1))) == 0) m |= 1 << b;
  return m;
}
#endif

GHD inline double fbits(std::int64_t i) { union { std::int64_t i; double f; } u{i}; return u.f; }

struct LaneResult {
  std::int64_t v;
  std::uint8_t t, b;
  int err;
  int fuel_used;  // charged fuel (instructions executed under the unit schedule)
};

#if defined(__HIPCC__)
// Runs `p` for case `c` on this lane (active = lane has a case). Returns the
// result or error of the case.
__device__ inline LaneResult vm_run(const ProgView& p, const CaseData& cd, int c, bool active, int fuel0, WaveStack& st,
                                    int lane) {
  LaneResult r{0, TAG_INVALID, 0, active ? ERR_NONE : ERR_VALUE, 0};
  bool done = !active;
  int fuel = fuel0;
  int ip = 0;
  int uncharged = 0;
  // Every iteration runs at least one live lane's instruction, which charges
  // fuel or counts toward the zero-cost cycle limit; this bound is never
  // reached by a correct schedule and ends the wave with ValueError otherwise.
  const long long iteration_limit = 64LL * ((long long)fuel0 + 2LL * p.len + 2) + 64;
  long long iterations = 0;
  for (int k = 0; k < st.temps; ++k) st.t[(st.slots + k) * 64 + lane] = TAG_INVALID;  // unassigned
  for (;;) {
    // Converged live lanes (always so in straight-line code) share their ip:
    // take it from the first live lane (finished lanes keep a stale ip);
    // otherwise take the smallest one (ip <= len < 8191, the code capacity).
    const unsigned long long live = __ballot(!done);
    if (live == 0) break;
    int at = __builtin_amdgcn_readlane(ip, __ffsll((long long)live) - 1);
    if (p.has_jumps && __ballot(!done && ip != at) != 0)
      at = __builtin_amdgcn_readfirstlane(wave_min_bits<13>(done ? 8191 : (ip < 8191 ? ip : 8190)));
    if (++iterations > iteration_limit) {
      if (!done) { r.err = ERR_VALUE; done = true; }
      break;
    }
    at = __builtin_amdgcn_readfirstlane(at);  // uniform: keeps the fetch scalar
    const bool mine = !done && ip == at;
    if (at >= p.len) {  // ran off the end without RETURN
      if (mine) { r.err = ERR_VALUE; done = true; }
      continue;
    }
    const Ins ins = p.code[at];
    const int sp = ins.sp;
    if (mine) {
      const int cost = ins.fuel;
      if (fuel < cost) { r.err = ERR_TIMEOUT; done = true; }
      else {
        fuel -= cost;
        if (cost) uncharged = 0;
        else if (++uncharged > p.len) { r.err = ERR_VALUE; done = true; }
      }
    }
    if (!mine || done) continue;
    ++ip;
    switch (ins.op) {
      case OP_PUSH_CONST: {
        const Val v = p.pool[ins.a];
        st.v[(sp) * 64 + lane] = v.i; st.t[(sp) * 64 + lane] = v.tag; st.b[(sp) * 64 + lane] = v.b;
        break;
      }
      case OP_LOAD: {
        if (ins.a >= cd.n_inputs) {  // a hidden local
          const int k = ins.a - cd.n_inputs;
          const int at_t = (st.slots + k) * 64 + lane;
          if (k >= st.temps - st.shadow || st.t[at_t] == TAG_INVALID) { r.err = ERR_NAME; done = true; break; }
          st.v[(sp) * 64 + lane] = st.v[at_t]; st.t[(sp) * 64 + lane] = st.t[at_t]; st.b[(sp) * 64 + lane] = st.b[at_t];
          break;
        }
        if (ins.a < st.shadow) {  // a stored input reads its shadow
          const int at_s = (st.slots + st.temps - st.shadow + ins.a) * 64 + lane;
          if (st.t[at_s] != TAG_INVALID) {
            st.v[(sp) * 64 + lane] = st.v[at_s]; st.t[(sp) * 64 + lane] = st.t[at_s]; st.b[(sp) * 64 + lane] = st.b[at_s];
            break;
          }
        }
        const std::size_t at_in = std::size_t(ins.a) * cd.C + c;
        const std::uint8_t tag = cd.in_tag[at_in];
        if (tag == TAG_INVALID) { r.err = ERR_NAME; done = true; break; }
        st.v[(sp) * 64 + lane] = cd.in_val[at_in]; st.t[(sp) * 64 + lane] = tag; st.b[(sp) * 64 + lane] = cd.in_b[at_in];
        break;
      }
      case OP_STORE: {  // a hidden local, or an input's shadow
        const int k = ins.a < cd.n_inputs ? st.temps - st.shadow + ins.a : ins.a - cd.n_inputs;
        if (ins.a < 0 || (ins.a < cd.n_inputs ? ins.a >= st.shadow : k >= st.temps - st.shadow)) {
          r.err = ERR_VALUE; done = true; break;
        }
        const int at_t = (st.slots + k) * 64 + lane;
        st.v[at_t] = st.v[(sp - 1) * 64 + lane]; st.t[at_t] = st.t[(sp - 1) * 64 + lane]; st.b[at_t] = st.b[(sp - 1) * 64 + lane];
        break;
      }
      case OP_CHECK_INT:
        if (st.t[(sp - 1) * 64 + lane] != TAG_INT) { r.err = ERR_TYPE; done = true; }
        break;
      case OP_CHECK_LIST: {
        const std::uint8_t t = st.t[(sp - 1) * 64 + lane];
        if (t != TAG_INT_LIST && t != TAG_FLOAT_LIST && t != TAG_STRING_LIST) { r.err = ERR_TYPE; done = true; }
        break;
      }
      case OP_NEG: {
        const std::uint8_t t = st.t[(sp - 1) * 64 + lane];
        if (t == TAG_INT) st.v[(sp - 1) * 64 + lane] = wrap_neg(st.v[(sp - 1) * 64 + lane]);
        else if (t == TAG_FLOAT) {
          union { std::int64_t i; double f; } u{st.v[(sp - 1) * 64 + lane]};
          st.v[(sp - 1) * 64 + lane] = val_float(canon32(-u.f)).i;
        } else { r.err = ERR_TYPE; done = true; }
        break;
      }
      case OP_NOT:
        if (st.t[(sp - 1) * 64 + lane] != TAG_BOOL) { r.err = ERR_TYPE; done = true; }
        else st.b[(sp - 1) * 64 + lane] = !st.b[(sp - 1) * 64 + lane];
        break;
      case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_MOD: {
        const std::uint8_t ta = st.t[(sp - 2) * 64 + lane], tb = st.t[(sp - 1) * 64 + lane];
        const std::int64_t a = st.v[(sp - 2) * 64 + lane], b = st.v[(sp - 1) * 64 + lane];
        if (ta == TAG_INT && tb == TAG_INT) {
          if ((ins.op == OP_DIV || ins.op == OP_MOD) && b == 0) { r.err = ERR_ZERO_DIV; done = true; break; }
          std::int64_t out;
          if (ins.op == OP_ADD) out = wrap_add(a, b);
          else if (ins.op == OP_SUB) out = wrap_sub(a, b);
          else if (ins.op == OP_MUL) out = wrap_mul(a, b);
          else if (ins.op == OP_MOD) out = py_mod(a, b);
          else {
            st.v[(sp - 2) * 64 + lane] = val_float(canon32(double(a) / double(b))).i;
            st.t[(sp - 2) * 64 + lane] = TAG_FLOAT;
            break;
          }
          st.v[(sp - 2) * 64 + lane] = out;
        } else if (ta == TAG_FLOAT && tb == TAG_FLOAT) {
          union { std::int64_t i; double f; } x{a}, y{b};
          if ((ins.op == OP_DIV || ins.op == OP_MOD) && y.f == 0.0) { r.err = ERR_ZERO_DIV; done = true; break; }
          double out;
          if (ins.op == OP_ADD) out = x.f + y.f;
          else if (ins.op == OP_SUB) out = x.f - y.f;
          else if (ins.op == OP_MUL) out = x.f * y.f;
          else if (ins.op == OP_DIV) out = x.f / y.f;
          else {
            double m = fmod(x.f, y.f);
            if (m == 0.0) m = copysign(0.0, y.f);
            else if (signbit(m) != signbit(y.f)) m += y.f;
            out = m;
          }
          st.v[(sp - 2) * 64 + lane] = val_float(canon32(out)).i;
        } else { r.err = ERR_TYPE; done = true; }
        break;
      }
      case OP_LT: case OP_LE: case OP_GT: case OP_GE: case OP_EQ: case OP_NE: {
        const std::uint8_t ta = st.t[(sp - 2) * 64 + lane], tb = st.t[(sp - 1) * 64 + lane];
        const std::int64_t a = st.v[(sp - 2) * 64 + lane], b = st.v[(sp - 1) * 64 + lane];
        bool out = false;
        if (ins.op == OP_EQ || ins.op == OP_NE) {
          if (ta != tb) { r.err = ERR_TYPE; done = true; break; }
          bool eq;
          if (ta == TAG_BOOL) eq = st.b[(sp - 2) * 64 + lane] == st.b[(sp - 1) * 64 + lane];
          else if (ta == TAG_FLOAT) { union { std::int64_t i; double f; } x{a}, y{b}; eq = x.f == y.f; }
          else if (ta == TAG_INT || ta == TAG_CHAR || ta == TAG_STRING || ta == TAG_INT_LIST || ta == TAG_FLOAT_LIST ||
                   ta == TAG_STRING_LIST) eq = a == b;
          else { r.err = ERR_TYPE; done = true; break; }
          out = ins.op == OP_EQ ? eq : !eq;
        } else if (ta == TAG_INT && tb == TAG_INT) {
          out = ins.op == OP_LT ? a < b : ins.op == OP_LE ? a <= b : ins.op == OP_GT ? a > b : a >= b;
        } else if (ta == TAG_FLOAT && tb == TAG_FLOAT) {
          union { std::int64_t i; double f; } x{a}, y{b};
          out = ins.op == OP_LT ? x.f < y.f : ins.op == OP_LE ? x.f <= y.f : ins.op == OP_GT ? x.f > y.f : x.f >= y.f;
        } else { r.err = ERR_TYPE; done = true; break; }
        st.v[(sp - 2) * 64 + lane] = 0; st.t[(sp - 2) * 64 + lane] = TAG_BOOL; st.b[(sp - 2) * 64 + lane] = out;
        break;
      }
      case OP_JMP:
        ip = ins.a;
        break;
      case OP_JMP_IF_FALSE: case OP_JMP_IF_TRUE: {
        if (st.t[(sp - 1) * 64 + lane] != TAG_BOOL) { r.err = ERR_TYPE; done = true; break; }
        const bool cond = st.b[(sp - 1) * 64 + lane] != 0;
        if ((ins.op == OP_JMP_IF_FALSE && !cond) || (ins.op == OP_JMP_IF_TRUE && cond)) ip = ins.a;
        break;
      }
      case OP_RETURN:
        r.v = st.v[(sp - 1) * 64 + lane]; r.t = st.t[(sp - 1) * 64 + lane]; r.b = st.b[(sp - 1) * 64 + lane];
        done = true;
        break;
      case OP_CALL_BUILTIN: {
        const int argc = ins.argc, base = sp - argc;
        std::int64_t out = 0;
        std::uint8_t out_t = TAG_INT, out_b = 0;
        const auto ti = [&](int k) { return st.t[(base + k) * 64 + lane]; };
        const auto vi = [&](int k) { return st.v[(base + k) * 64 + lane]; };
        switch (ins.a) {
          case B_ABS:
            if (argc != 1 || !is_numeric(ti(0))) { r.err = ERR_TYPE; break; }
            if (ti(0) == TAG_FLOAT) {
              const double x = fbits(vi(0));
              out = val_float(canon32(x < 0 ? -x : x)).i;
              out_t = TAG_FLOAT;
            } else out = vi(0) < 0 ? wrap_neg(vi(0)) : vi(0);
            break;
          case B_MIN: case B_MAX:
            if (argc != 2 || !is_numeric(ti(0)) || ti(0) != ti(1)) { r.err = ERR_TYPE; break; }
            if (ti(0) == TAG_FLOAT) {
              const double a = fbits(vi(0)), b = fbits(vi(1));
              out = val_float(canon32(ins.a == B_MIN ? (a <= b ? a : b) : (a >= b ? a : b))).i;
              out_t = TAG_FLOAT;
            } else out = ins.a == B_MIN ? (vi(0) < vi(1) ? vi(0) : vi(1)) : (vi(0) > vi(1) ? vi(0) : vi(1));
            break;
          case B_CLIP: {
            if (argc != 3 || !is_numeric(ti(0)) || ti(0) != ti(1) || ti(0) != ti(2)) { r.err = ERR_TYPE; break; }
            if (ti(0) == TAG_FLOAT) {
              const double x = fbits(vi(0)), lo = fbits(vi(1)), hi = fbits(vi(2));
              const double m = x > lo ? x : lo;  // min(max(x, lo), hi)
              out = val_float(canon32(m < hi ? m : hi)).i;
              out_t = TAG_FLOAT;
              break;
            }
            const std::int64_t m = vi(0) > vi(1) ? vi(0) : vi(1);
            out = m < vi(2) ? m : vi(2);
            break;
          }
          case B_IDIV0: case B_IMOD0: case B_TRUNC_DIV0: case B_TRUNC_MOD0: {
            if (argc != 2 || ti(0) != TAG_INT || ti(1) != TAG_INT) { r.err = ERR_TYPE; break; }
            const std::int64_t a = vi(0), b = vi(1);
            if (b == 0) out = 0;
            else if (ins.a == B_TRUNC_DIV0) out = trunc_div0(a, b);
            else if (ins.a == B_TRUNC_MOD0) out = trunc_mod0(a, b);
            else if (ins.a == B_IDIV0) out = wrap_div(a, b);
            else out = py_mod(a, b);
            break;
          }
          case B_LEN: {
            const std::uint8_t t = ti(0);
            if (argc != 1 || !(t == TAG_INT_LIST || t == TAG_FLOAT_LIST || t == TAG_STRING_LIST)) { r.err = ERR_TYPE; break; }
            out = std::int64_t(std::uint64_t(vi(0)) & 0xffff);
            break;
          }
          case B_INDEX: case B_INDEX_NONNEG: case B_INDEX_CLAMPED: {
            if (argc != 2 || ti(0) != TAG_INT_LIST || ti(1) != TAG_INT) { r.err = ERR_TYPE; break; }
            const std::int64_t n = std::int64_t(std::uint64_t(vi(0)) & 0xffff);
            const std::uint64_t off = std::uint64_t(vi(0)) >> 16;
            std::int64_t j = vi(1);
            if (ins.a == B_INDEX_CLAMPED) {
              if (n == 0) { out = 0; break; }
              j = j < 0 ? 0 : (j >= n ? n - 1 : j);
            } else {
              if (ins.a == B_INDEX_NONNEG && j < 0) { r.err = ERR_VALUE; break; }
              if (j < 0) j += n;
              if (j < 0 || j >= n) { r.err = ERR_VALUE; break; }
            }
            out = cd.arena[off + std::uint64_t(j)];
            break;
          }
          case B_IS_INT:
            if (argc != 1) { r.err = ERR_TYPE; break; }
            out_t = TAG_BOOL;
            out_b = ti(0) == TAG_INT;
            break;
          default: r.err = ERR_NAME; break;
        }
        if (r.err != ERR_NONE) { done = true; break; }
        st.v[(base) * 64 + lane] = out; st.t[(base) * 64 + lane] = out_t; st.b[(base) * 64 + lane] = out_b;
        break;
      }
      default:
        r.err = ERR_VALUE; done = true;
        break;
    }
  }
  r.fuel_used = fuel0 - fuel;
  return r;
}

// ---- typed straight-line path ----
// A certified program (no jumps, every value statically Int or Bool, integer
// builtins only) needs no run-time tags: values are int64 (Bool as 0/1), the
// stack is a register array indexed by the instruction's static height (the
// index is uniform, so the compiler addresses the array with indexed register
// moves), and every live lane executes the same instruction. Instruction j is
// held by lane j % 64 of register group j / 64 and read with readlane; so are
// the constants. Fuel is charged uniformly: a timeout ends every live lane at
// the same instruction, as the per-lane charge of vm_run would.
struct FastProgram {
  int w0[2], w1[2];  // instruction: op | argc << 8 | sp << 16 | fuel << 24, operand
  int c0[2], c1[2];  // constant payload, low and high words
  int len;
};
__device__ inline FastProgram fast_load(const ProgView& p, int lane) {
  FastProgram f;
  f.len = p.len;
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int j = g * 64 + lane;
    int w0 = 0, w1 = 0;
    if (j < p.len) {
      const Ins ins = p.code[j];
      w0 = int(ins.op) | int(ins.argc) << 8 | int(ins.sp) << 16 | int(unsigned(ins.fuel) << 24);
      w1 = ins.a;
    }
    std::int64_t v = 0;
    if (j < p.n_pool) {
      const Val x = p.pool[j];
      v = x.tag == TAG_BOOL ? std::int64_t(x.b != 0) : x.i;
    }
    f.w0[g] = w0;
    f.w1[g] = w1;
    f.c0[g] = int(std::uint32_t(std::uint64_t(v)));
    f.c1[g] = int(std::uint32_t(std::uint64_t(v) >> 32));
  }
  return f;
}
__device__ inline int fast_pick(const int (&v)[2], int k) {
  const int lo = __builtin_amdgcn_readlane(v[0], k & 63), hi = __builtin_amdgcn_readlane(v[1], k & 63);
  return k < 64 ? lo : hi;
}
// Inputs of this lane's case: payloads (Bool as 0/1) and the bound mask.
// (Separate fields and a uniform select keep them in registers; an array
// indexed by the operand would live in scratch.)
static_assert(kFastInputs == 4, "four input registers");
struct FastInputs {
  std::int64_t v0, v1, v2, v3;
  unsigned bound;
  __device__ std::int64_t pick(int k) const { return k == 0 ? v0 : k == 1 ? v1 : k == 2 ? v2 : v3; }
  // A store into input k (uniform): later loads read it, and it is bound.
  __device__ void set(int k, std::int64_t v) {
    if (k == 0) v0 = v;
    else if (k == 1) v1 = v;
    else if (k == 2) v2 = v;
    else v3 = v;
    bound |= 1u << (k & 31);
  }
};
__device__ inline FastInputs fast_inputs(const CaseData& cd, int n_inputs, int c, bool active) {
  FastInputs in{0, 0, 0, 0, 0};
  std::int64_t v[kFastInputs];
#pragma unroll
  for (int k = 0; k < kFastInputs; ++k) {
    v[k] = 0;
    if (k < n_inputs && k < cd.n_inputs && active) {
      const std::size_t at = std::size_t(k) * cd.C + c;
      const std::uint8_t tag = cd.in_tag[at];
      if (tag != TAG_INVALID) in.bound |= 1u << k;
      v[k] = tag == TAG_BOOL ? std::int64_t(cd.in_b[at] != 0) : cd.in_val[at];
    }
  }
  in.v0 = v[0];
  in.v1 = v[1];
  in.v2 = v[2];
  in.v3 = v[3];
  return in;
}
__device__ inline LaneResult vm_fast(const FastProgram& f, const FastInputs& in, bool active, int fuel0, int result_tag) {
  LaneResult r{0, TAG_INVALID, 0, active ? ERR_NONE : ERR_VALUE, 0};
  bool done = !active;
  std::int64_t s[kFastStack];
#pragma unroll
  for (int k = 0; k < kFastStack; ++k) s[k] = 0;
  int fuel = fuel0;
  const auto finish = [&](int err) {
    if (!done) { r.err = err; r.fuel_used = fuel0 - fuel; done = true; }
  };
  for (int ip = 0; ip < f.len; ++ip) {
    if (__ballot(!done) == 0) break;
    const int w = fast_pick(f.w0, ip), a = fast_pick(f.w1, ip);
    const int op = w & 0xff, argc = (w >> 8) & 0xff, sp = (w >> 16) & 0xff, cost = int(unsigned(w) >> 24);
    if (fuel < cost) { finish(ERR_TIMEOUT); break; }
    fuel -= cost;
    // Heights are certified within the stack; the masks keep every index in range.
    const int top = (sp - 1) & (kFastStack - 1), below = (sp - 2) & (kFastStack - 1);
    // MOD and the four protected-division builtins share one 64-bit
    // truncating division: x = y*q + r, with y == -1 giving q = -x wrapped,
    // r = 0 (never the trapping INT64_MIN / -1).
    const bool divide = op == OP_MOD || (op == OP_CALL_BUILTIN && (a == B_IDIV0 || a == B_IMOD0 || a == B_TRUNC_DIV0 ||
                                                                   a == B_TRUNC_MOD0));
    if (divide) {
      const int at = op == OP_MOD ? below : (sp - argc) & (kFastStack - 1);
      const std::int64_t x = s[at], y = s[(at + 1) & (kFastStack - 1)];
      const std::int64_t safe = (y == 0 || y == -1) ? 1 : y;
      std::int64_t q = x / safe, rem = x - q * safe;
      if (y == -1) { q = wrap_neg(x); rem = 0; }
      const std::int64_t floor_mod = (rem != 0 && ((rem < 0) != (y < 0))) ? rem + y : rem;  // Python %
      std::int64_t out;
      if (op == OP_MOD) {
        if (y == 0) finish(ERR_ZERO_DIV);
        out = y == 0 ? 0 : floor_mod;
      } else if (a == B_IDIV0) out = y == 0 ? 0 : q;                                           // wrap_div
      else if (a == B_IMOD0) out = y == 0 ? 0 : floor_mod;                                     // py_mod
      else if (a == B_TRUNC_DIV0) out = (y == 0 || (x == INT64_MIN && y == -1)) ? 0 : q;       // trunc_div0
      else out = (y == 0 || (x == INT64_MIN && y == -1)) ? 0 : rem;                            // trunc_mod0
      s[at] = out;
      continue;
    }
    switch (op) {
      case OP_PUSH_CONST: {
        const std::uint32_t lo = std::uint32_t(fast_pick(f.c0, a)), hi = std::uint32_t(fast_pick(f.c1, a));
        s[sp & (kFastStack - 1)] = std::int64_t(std::uint64_t(hi) << 32 | lo);
        break;
      }
      case OP_LOAD:
        if (!((in.bound >> (a & 31)) & 1u)) finish(ERR_NAME);
        s[sp & (kFastStack - 1)] = in.pick(a);
        break;
      case OP_CHECK_INT: break;
      case OP_NEG: s[top] = wrap_neg(s[top]); break;
      case OP_NOT: s[top] = s[top] == 0; break;
      case OP_ADD: s[below] = wrap_add(s[below], s[top]); break;
      case OP_SUB: s[below] = wrap_sub(s[below], s[top]); break;
      case OP_MUL: s[below] = wrap_mul(s[below], s[top]); break;
      case OP_LT: s[below] = s[below] < s[top]; break;
      case OP_LE: s[below] = s[below] <= s[top]; break;
      case OP_GT: s[below] = s[below] > s[top]; break;
      case OP_GE: s[below] = s[below] >= s[top]; break;
      case OP_EQ: s[below] = s[below] == s[top]; break;
      case OP_NE: s[below] = s[below] != s[top]; break;
      case OP_CALL_BUILTIN: {
        const int base = (sp - argc) & (kFastStack - 1);
        const std::int64_t x = s[base], y = s[(base + 1) & (kFastStack - 1)];
        std::int64_t out = 0;
        switch (a) {
          case B_ABS: out = x < 0 ? wrap_neg(x) : x; break;
          case B_MIN: out = x < y ? x : y; break;
          case B_MAX: out = x > y ? x : y; break;
          case B_CLIP: {
            const std::int64_t hi = s[(base + 2) & (kFastStack - 1)], m = x > y ? x : y;
            out = m < hi ? m : hi;
            break;
          }
          default: finish(ERR_NAME); break;
        }
        s[base] = out;
        break;
      }
      case OP_RETURN: {
        if (!done) {
          if (result_tag == TAG_BOOL) { r.t = TAG_BOOL; r.b = s[top] != 0; r.v = 0; }
          else { r.t = TAG_INT; r.v = s[top]; }
          r.fuel_used = fuel0 - fuel;
          done = true;
        }
        ip = f.len;
        break;
      }
      default: finish(ERR_VALUE); break;
    }
  }
  finish(ERR_VALUE);  // ran off the end without RETURN
  return r;
}

// ---- typed looping path ----
// A program certified like the straight-line path but with jumps, hidden
// locals or IntList views (phase.hpp loop_operation): values are untagged
// int64 (Bool as 0/1, an IntList as its view payload), the operand stack and
// the hidden locals are 8-byte LDS slots per lane (conflict-free, addressed
// by the instruction's static height and operand), and code and constants
// sit in four register groups read with readlane. Lanes follow their own branches under the
// reference's per-case semantics: fuel is charged per lane, and the wave
// executes the smallest pending instruction index for the lanes that are at
// it. That index is the previous one plus one except after a jump or when a
// lane finished, where a wave minimum recomputes it.
struct LoopProgram {
  int w0[4], w1[4];  // instruction: op | argc << 8 | sp << 16 | fuel << 24, operand
  int c0[4], c1[4];  // constant payload, low and high words
  int len;
};
__device__ inline LoopProgram loop_load(const ProgView& p, int lane) {
  LoopProgram f;
  f.len = p.len;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int j = g * 64 + lane;
    int w0 = 0, w1 = 0;
    if (j < p.len) {
      const Ins ins = p.code[j];
      w0 = int(ins.op) | int(ins.argc) << 8 | int(ins.sp) << 16 | int(unsigned(ins.fuel) << 24);
      w1 = ins.a;
    }
    std::int64_t v = 0;
    if (j < p.n_pool) {
      const Val x = p.pool[j];
      v = x.tag == TAG_BOOL ? std::int64_t(x.b != 0) : x.i;
    }
    f.w0[g] = w0;
    f.w1[g] = w1;
    f.c0[g] = int(std::uint32_t(std::uint64_t(v)));
    f.c1[g] = int(std::uint32_t(std::uint64_t(v) >> 32));
  }
  return f;
}
// The binary operations of the looping path on untagged operands; `zero`:
// MOD by zero (ZeroDivisionError). MOD is Python's, by the shared truncating
// division of vm_fast.
__device__ inline std::int64_t loop_binary(int op, std::int64_t x, std::int64_t y, bool& zero) {
  switch (op) {
    case OP_ADD: return wrap_add(x, y);
    case OP_SUB: return wrap_sub(x, y);
    case OP_MUL: return wrap_mul(x, y);
    case OP_LT: return x < y;
    case OP_LE: return x <= y;
    case OP_GT: return x > y;
    case OP_GE: return x >= y;
    case OP_EQ: return x == y;
    case OP_NE: return x != y;
    default: {  // OP_MOD
      const std::int64_t safe = (y == 0 || y == -1) ? 1 : y;
      const std::int64_t rem = y == -1 ? 0 : x - (x / safe) * safe;
      zero = y == 0;
      return y == 0 ? 0 : ((rem != 0 && ((rem < 0) != (y < 0))) ? rem + y : rem);
    }
  }
}
__device__ inline int loop_pick(const int (&v)[4], int k) {
  const int x0 = __builtin_amdgcn_readlane(v[0], k & 63), x1 = __builtin_amdgcn_readlane(v[1], k & 63);
  const int x2 = __builtin_amdgcn_readlane(v[2], k & 63), x3 = __builtin_amdgcn_readlane(v[3], k & 63);
  return k < 128 ? (k < 64 ? x0 : x1) : (k < 192 ? x2 : x3);
}
// `slots`: this wave's LDS area of kLoopSlots untagged slots, slot-major
// (slot k of lane j at k * 64 + j): the operand stack, then the hidden locals.
__device__ inline LaneResult vm_loop(const LoopProgram& f, FastInputs in, int n_inputs, GAS const std::int64_t* arena,
                                     LAS std::int64_t* slots, int lane, bool active, int fuel0, int result_tag) {
  // `in` is this lane's copy: statement targets may store into inputs.
  LaneResult r{0, TAG_INVALID, 0, active ? ERR_NONE : ERR_VALUE, 0};
  bool done = !active;
  LAS std::int64_t* const sl = slots + lane;
  // Heights are certified below kLoopStack; the clamp keeps every slot in range.
  const auto S = [&](int k) -> LAS std::int64_t& { return sl[(k < 0 ? 0 : k < kLoopStack ? k : kLoopStack - 1) * 64]; };
  const auto L = [&](int k) -> LAS std::int64_t& { return sl[(kLoopStack + (k < 0 ? 0 : k < kLoopLocals ? k : 0)) * 64]; };
  unsigned assigned = 0;  // hidden locals stored so far
  int fuel = fuel0, ip = 0, at = 0;
  // The top of this lane's operand stack lives in a register; the LDS
  // slots hold the elements below it (slot k: element k, for k < height - 1).
  // Only the lanes that execute a step change either.
  std::int64_t tos = 0;
  // Every step that finishes a lane, jumps or runs a fixed sequence sets
  // `resched`; any other step leaves the live lanes as they were, at + 1.
  bool resched = true;
  // Every step runs one instruction of at least one live lane, which charges
  // fuel, so this bound is never reached by a certified program. A wave that
  // never stops must jump, so the bound is checked when rescheduling.
  const long long limit = 64LL * ((long long)fuel0 + 2LL * f.len + 2) + 64;
  for (long long step = 0;; ++step) {
    if (resched) {
      // Converged live lanes (the common case) share their ip; otherwise
      // take the smallest one.
      const unsigned long long live = __ballot(!done);
      if (live == 0) break;
      if (step > limit) {
        if (!done) { r.err = ERR_VALUE; r.fuel_used = fuel0 - fuel; done = true; }
        break;
      }
      const int lead = __builtin_amdgcn_readlane(ip, __ffsll((long long)live) - 1);
      if (__ballot(!done && ip != lead) == 0) at = lead;
      else at = __builtin_amdgcn_readfirstlane(wave_min_bits<9>(done ? 511 : (ip < 511 ? ip : 510)));  // ip <= len <= 256
      resched = false;
    }
    // `at` is wave-uniform on every path; restating that keeps it (and the
    // decode below) in scalar registers across the loop's merges.
    at = __builtin_amdgcn_readfirstlane(at);
    const bool mine = !done && ip == at;
    if (at >= f.len) {  // ran off the end without RETURN
      if (mine) { r.err = ERR_VALUE; r.fuel_used = fuel0 - fuel; done = true; }
      resched = true;
      continue;
    }
    // The decoded fields are scalar values; readfirstlane states that for
    // the compiler, which then indexes the register arrays with them.
    const int w = __builtin_amdgcn_readfirstlane(loop_pick(f.w0, at)), a = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at));
    const int op = w & 0xff, argc = (w >> 8) & 0xff, sp = (w >> 16) & 0xff, cost = int(unsigned(w) >> 24);
    if ((argc & 0x80) && (argc & 0xf0) != kFuse) {
      // A traversal's or for_range loop's fixed sequence as one step
      // (phase.hpp kWalkTest, kWalkNext, kRangeTest, kRangeNext). Every
      // instruction of it costs one unit of fuel and none
      // can fail otherwise, so a lane with fewer units than the sequence
      // charges times out inside it, with no fuel left.
      const bool reverse = (argc & 3) == 2;
      const int k_counter = a - n_inputs;
      const int tag = argc & 0xf0;
      if (tag == kRangeTest || tag == kRangeNext) {
        // A for_range step (phase.hpp kRangeTest, kRangeNext): unit costs
        // and no failure but fuel, as for the traversal sequences.
        const bool test = tag == kRangeTest;
        const int k_bound = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 1)) - n_inputs;
        const int exit_at = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 3));
        const int k_var = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 5)) - n_inputs;
        const int head = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 4));
        const std::int64_t counter = L(k_counter);
        const bool more = test && counter < L(k_bound);
        const int need = test ? (more ? 6 : 4) : 5;
        const bool starve = mine && fuel < need;
        const bool run = mine && !starve;
        if (run) {
          if (more && k_var < 0) {
            in.set(k_var + n_inputs, counter);  // an input target (uniform)
          } else if (more) {
            L(k_var) = counter;
            assigned |= 1u << k_var;
          } else if (!test) {
            L(k_counter) = wrap_add(counter, 1);
          }
          fuel -= need;
          ip = test ? (more ? at + 6 : exit_at) : head;
        }
        if (starve) {
          fuel = 0;
          r.err = ERR_TIMEOUT;
          r.fuel_used = fuel0;
          done = true;
        }
      } else if (tag == kWalkTest) {
        // LOAD counter, LOAD bound, LT/GE, JIF done, LOAD xs, LOAD counter,
        // index, STORE element, LOAD start, LOAD counter, ADD, STORE index.
        const int k_bound = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 1)) - n_inputs;
        const int done_at = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 3));
        const int k_xs = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 4)) - n_inputs;
        const int k_element = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 7)) - n_inputs;
        const int k_start = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 8)) - n_inputs;
        const int k_index = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 11)) - n_inputs;
        const std::int64_t counter = L(k_counter), bound = L(k_bound);
        const bool more = reverse ? counter >= bound : counter < bound;
        const int need = more ? 12 : 4;
        const bool starve = mine && fuel < need;
        const bool run = mine && !starve;
        // The counter lies in [0, len): the element read is in bounds.
        const std::int64_t xs = L(k_xs);
        const bool read = run && more;
        const std::int64_t element = arena[read ? (std::uint64_t(xs) >> 16) + std::uint64_t(counter) : 0];
        if (read) {
          L(k_element) = element;
          L(k_index) = wrap_add(L(k_start), counter);
          assigned |= (1u << k_element) | (1u << k_index);
        }
        if (run) {
          fuel -= need;
          ip = more ? at + 12 : done_at;
        }
        if (starve) {
          fuel = 0;
          r.err = ERR_TIMEOUT;
          r.fuel_used = fuel0;
          done = true;
        }
      } else {
        // STORE state, LOAD counter, PUSH 1, ADD/SUB, STORE counter, JMP head.
        const int k_state = k_counter;
        const int k_step = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 1)) - n_inputs;
        const int head = __builtin_amdgcn_readfirstlane(loop_pick(f.w1, at + 5));
        const bool starve = mine && fuel < 6;
        const bool run = mine && !starve;
        const std::int64_t deeper = S(sp - 2);  // the top once the state is stored
        if (run) {
          L(k_state) = tos;
          assigned |= 1u << k_state;
          const std::int64_t c = L(k_step);
          L(k_step) = reverse ? wrap_sub(c, 1) : wrap_add(c, 1);
          fuel -= 6;
          ip = head;
          tos = deeper;
        }
        if (starve) {
          fuel = 0;
          r.err = ERR_TIMEOUT;
          r.fuel_used = fuel0;
          done = true;
        }
      }
      resched = true;
      continue;
    }
    if ((argc & 0xf0) == kFuse) {
      // A LOAD or PUSH_CONST and the binary operation after it (phase.hpp
      // kFuse) as one step, when every lane at `at` has the fuel of both;
      // otherwise the first runs alone below.
      const int w2 = __builtin_amdgcn_readfirstlane(loop_pick(f.w0, at + 1));
      const int op2 = w2 & 0xff, cost2 = int(unsigned(w2) >> 24);
      if (__ballot(mine && fuel < cost + cost2) == 0) {
        std::int64_t y;
        bool bound = true;
        if (op == OP_PUSH_CONST) {
          const std::uint32_t lo = std::uint32_t(loop_pick(f.c0, a)), hi = std::uint32_t(loop_pick(f.c1, a));
          y = std::int64_t(std::uint64_t(hi) << 32 | lo);
        } else if (a < n_inputs) {  // uniform branch
          y = in.pick(a);
          bound = (in.bound >> (a & 31)) & 1u;
        } else {
          const int k = a - n_inputs < kLoopLocals ? a - n_inputs : 0;
          y = L(k);
          bound = (assigned >> k) & 1u;
        }
        bool zero = false;
        const std::int64_t v = loop_binary(op2, tos, y, zero);
        int err2 = ERR_NONE;
        if (mine) {
          if (!bound) {
            err2 = ERR_NAME;
            fuel -= cost;
          } else {
            fuel -= cost + cost2;
            if (zero) err2 = ERR_ZERO_DIV;
            tos = v;
            ip = at + 2;
          }
        }
        if (mine && err2 != ERR_NONE) {
          r.err = err2;
          r.fuel_used = fuel0 - fuel;
          done = true;
        }
        resched = true;
        continue;
      }
    }
    // Every lane computes the instruction; the lanes at `at` with fuel
    // commit it (predication instead of divergent branches).
    const bool starve = mine && fuel < cost;
    const bool run = mine && !starve;
    int err = starve ? ERR_TIMEOUT : ERR_NONE;
    bool ret = false;
    int next = ip + 1;
    const int below = sp - 2, push = sp;
    // Binary operations read both operands before any store.
    switch (op) {
      case OP_PUSH_CONST: {
        const std::uint32_t lo = std::uint32_t(loop_pick(f.c0, a)), hi = std::uint32_t(loop_pick(f.c1, a));
        if (run) {
          if (push > 0) S(push - 1) = tos;
          tos = std::int64_t(std::uint64_t(hi) << 32 | lo);
        }
        break;
      }
      case OP_LOAD: {
        std::int64_t v;
        bool bound;
        if (a < n_inputs) {  // uniform branch
          v = in.pick(a);
          bound = (in.bound >> (a & 31)) & 1u;
        } else {
          const int k = a - n_inputs < kLoopLocals ? a - n_inputs : 0;
          v = L(k);
          bound = (assigned >> k) & 1u;
        }
        if (run && !bound) err = ERR_NAME;
        if (run) {
          if (push > 0) S(push - 1) = tos;
          tos = v;
        }
        break;
      }
      case OP_STORE: {
        const std::int64_t deeper = S(below);  // the top after the pop
        if (a < n_inputs) {  // an input target (uniform branch)
          if (run) { in.set(a, tos); tos = deeper; }
          break;
        }
        const int k = a - n_inputs < kLoopLocals ? a - n_inputs : 0;
        if (run) { L(k) = tos; assigned |= 1u << k; tos = deeper; }
        break;
      }
      case OP_JMP: next = a; break;
      case OP_JMP_IF_FALSE: case OP_JMP_IF_TRUE: {
        const std::int64_t deeper = S(below);  // the condition is popped
        next = (tos != 0) == (op == OP_JMP_IF_TRUE) ? a : ip + 1;
        if (run) tos = deeper;
        break;
      }
      case OP_CHECK_INT: break;
      // argc 1 marks a NEG of a Bool (the for_range bound check): TypeError.
      case OP_NEG:
        if (run && argc == 1) err = ERR_TYPE;
        else if (run) tos = wrap_neg(tos);
        break;
      case OP_NOT: if (run) tos = std::int64_t(tos == 0); break;
      case OP_ADD: case OP_SUB: case OP_MUL: case OP_MOD: case OP_LT: case OP_LE: case OP_GT: case OP_GE: case OP_EQ:
      case OP_NE: {
        bool zero = false;
        const std::int64_t v = loop_binary(op, S(below), tos, zero);
        if (run && zero) err = ERR_ZERO_DIV;
        if (run) tos = v;
        break;
      }
      case OP_CALL_BUILTIN: {
        // The last argument is the register top; the others are below it.
        const int base = sp - argc;
        const std::int64_t x = argc == 1 ? tos : S(base), y = argc == 2 ? tos : S(base + 1);
        std::int64_t out = 0;
        switch (a) {
          case B_ABS: out = x < 0 ? wrap_neg(x) : x; break;
          case B_MIN: out = x < y ? x : y; break;
          case B_MAX: out = x > y ? x : y; break;
          case B_CLIP: {
            const std::int64_t hi = argc == 3 ? tos : S(base + 2), m = x > y ? x : y;
            out = m < hi ? m : hi;
            break;
          }
          case B_IDIV0: case B_IMOD0: case B_TRUNC_DIV0: case B_TRUNC_MOD0: {
            const std::int64_t safe = (y == 0 || y == -1) ? 1 : y;
            std::int64_t q = x / safe, rem = x - q * safe;
            if (y == -1) { q = wrap_neg(x); rem = 0; }
            const std::int64_t floor_mod = (rem != 0 && ((rem < 0) != (y < 0))) ? rem + y : rem;
            if (a == B_IDIV0) out = y == 0 ? 0 : q;
            else if (a == B_IMOD0) out = y == 0 ? 0 : floor_mod;
            else if (a == B_TRUNC_DIV0) out = (y == 0 || (x == INT64_MIN && y == -1)) ? 0 : q;
            else out = (y == 0 || (x == INT64_MIN && y == -1)) ? 0 : rem;
            break;
          }
          case B_LEN: out = std::int64_t(std::uint64_t(x) & 0xffff); break;
          case B_IS_INT: out = 1; break;  // certified on an Int operand
          case B_INDEX: case B_INDEX_NONNEG: case B_INDEX_CLAMPED: {  // as in vm_run
            const std::int64_t n = std::int64_t(std::uint64_t(x) & 0xffff);
            const std::uint64_t off = std::uint64_t(x) >> 16;
            std::int64_t j = y;
            bool ok = true, zero = false;
            if (a == B_INDEX_CLAMPED) {
              zero = n == 0;
              j = j < 0 ? 0 : (j >= n ? n - 1 : j);
            } else {
              if (a == B_INDEX_NONNEG && j < 0) ok = false;
              if (j < 0) j += n;
              if (j < 0 || j >= n) ok = false;
            }
            if (run && !ok) err = ERR_VALUE;
            // Only running lanes with a valid index read the arena; the rest
            // read its first element (the arena is never empty).
            const bool read = run && ok && !zero;
            const std::int64_t v = arena[read ? off + std::uint64_t(j) : 0];
            out = read ? v : 0;
            break;
          }
          default: if (run) err = ERR_NAME; break;
        }
        if (run) tos = out;
        break;
      }
      case OP_RETURN: ret = true; break;
      default: if (run) err = ERR_VALUE; break;
    }
    if (run) {
      ip = next;
      fuel -= cost;
    }
    const bool finished = mine && (err != ERR_NONE || ret);
    if (finished) {
      r.err = err;
      if (err == ERR_NONE) {
        const std::int64_t v = tos;
        if (result_tag == TAG_BOOL) { r.t = TAG_BOOL; r.b = v != 0; r.v = 0; }
        else { r.t = TAG_INT; r.v = v; }
      }
      r.fuel_used = fuel0 - fuel;
      done = true;
    }
    // The next instruction index: after a jump, or when a lane at this index
    // finished, recompute the minimum; otherwise every lane that ran is at
    // at + 1 and every other live lane was already past `at`.
    const bool jump = op == OP_JMP || op == OP_JMP_IF_FALSE || op == OP_JMP_IF_TRUE;
    if (jump || __ballot(finished) != 0) resched = true;
    else ++at;
  }
  return r;
}
#endif

}  // namespace gagp_amd
