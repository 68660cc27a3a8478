/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, single-threaded restatement of the reference's CPU evaluation
 * path. tests/ use it to verify the HIP engine and bench.py times it as the
 * CPU baseline. Nothing in the product (gagp_amd/) links, loads or calls it.
 *
 * What it restates (reference paths relative to the hschi1106/gagp tree):
 *   orc_run            cpp/src/runtime/cpu/execute_bytecode_cpu.cpp:152-625
 *                      (fuel charged before each instruction, first error ends
 *                      the case; BOUNDED_REGION is not restated)
 *   scalar semantics   cpp/include/gagp/core/value_semantics.hpp:33-191
 *   builtins (scalar, list length/index)
 *                      cpp/src/runtime/cpu/builtins_cpu.cpp:130-375,493-500
 *   orc_case_score     cpp/include/gagp/core/fitness_policy.hpp:28-59 and
 *                      cpp/src/runtime/cpu/fitness_cpu.cpp:19-42 (scalar metrics)
 *   orc_eval_fitness   cpp/src/runtime/cpu/fitness_cpu.cpp:57-176 (L logical
 *                      lanes, 48-bit rounding, 32-lane groups)
 *   orc_canon          fitness_cpu.cpp:44-53 (48 bits), selection.cpp:10-20 (40)
 *   orc_splitmix_*     cpp/src/evolution/repro/gpu/device/grammar_random.cuh:10-51
 *   orc_mt19937_64     std::mt19937_64 as specified by [rand.predef]
 *
 * How it is pinned (tests/test_oracle.py): the reference's bytecode fixture
 * files (cpp/tests/fixtures/runtime/*.json, committed as golden vectors), the
 * reference's RNG known answers, the C++ standard's mt19937_64 10000th-output
 * value, closed-form fitness totals, and golden runs of the reference engine.
 *
 * Value layout (16 bytes): int64 payload / float bits, Bool flag, tag.
 * IntList values are views: payload = offset * 65536 + length into the
 * caller's int64 element arena.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

typedef struct { int64_t i; uint8_t b; uint8_t tag; uint8_t pad[6]; } OVal;
typedef struct { uint8_t op, has_a, has_b, pad; int32_t a, b; } OIns;
typedef struct {
  const OIns* code; int32_t code_len;
  const uint32_t* fuel; /* NULL: unit schedule */
  const OVal* consts; int32_t n_consts;
  int32_t n_locals;
} OProg;

enum { T_INT = 0, T_FLOAT = 1, T_BOOL = 2, T_CHAR = 3, T_STRING = 4, T_INT_LIST = 5, T_FLOAT_LIST = 6,
       T_STRING_LIST = 7, T_FALLBACK = 8, T_INVALID = 9 };
enum { E_NONE = -1, E_NAME = 0, E_TYPE = 1, E_ZERO_DIV = 2, E_VALUE = 3, E_TIMEOUT = 4 };

static double f_of(OVal v) { double d; memcpy(&d, &v.i, 8); return d; }
static OVal mk_int(int64_t x) { OVal v; memset(&v, 0, sizeof v); v.i = x; v.tag = T_INT; return v; }
static OVal mk_bool(int x) { OVal v; memset(&v, 0, sizeof v); v.b = x ? 1 : 0; v.tag = T_BOOL; return v; }
static OVal mk_float(double d) { OVal v; memset(&v, 0, sizeof v); memcpy(&v.i, &d, 8); v.tag = T_FLOAT; return v; }
static int numeric(OVal v) { return v.tag == T_INT || v.tag == T_FLOAT; }
static double num(OVal v) { return v.tag == T_FLOAT ? f_of(v) : (double)v.i; }
static int is_list(OVal v) { return v.tag == T_INT_LIST || v.tag == T_FLOAT_LIST || v.tag == T_STRING_LIST; }

static double canon_bits(double value, int bits) {
  if (!isfinite(value) || value == 0.0) return value == 0.0 ? 0.0 : value;
  int e = 0;
  const double m = frexp(value, &e);
  const long long q = llround(ldexp(m, bits));
  return ldexp((double)q, e - bits);
}
double orc_canon(double v, int bits) { return canon_bits(v, bits); }

static int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
static int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
static int64_t wmul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
static int64_t wneg(int64_t a) { return (int64_t)(0ULL - (uint64_t)a); }
static int64_t pymod(int64_t a, int64_t b) {
  if (b == -1) return 0;
  int64_t r = a % b;
  if (r != 0 && ((r < 0) != (b < 0))) r += b;
  return r;
}
static double pyfmod(double a, double b) {
  double r = fmod(a, b);
  if (r == 0.0) return copysign(0.0, b);
  if (signbit(r) != signbit(b)) r += b;
  return r;
}

/* Comparison (value_semantics.hpp:131-191). Returns error code or E_NONE. */
static int compare(uint8_t op, OVal a, OVal b, int* out) {
  if (op == 14 || op == 15) {
    int eq;
    if (a.tag != b.tag) return E_TYPE;
    if (a.tag == T_BOOL) eq = a.b == b.b;
    else if (a.tag == T_FLOAT) eq = f_of(a) == f_of(b);
    else if (a.tag == T_INT || a.tag == T_CHAR || a.tag == T_STRING || is_list(a)) eq = a.i == b.i;
    else return E_TYPE;
    *out = op == 14 ? eq : !eq;
    return E_NONE;
  }
  if (a.tag == T_INT && b.tag == T_INT) {
    *out = op == 10 ? a.i < b.i : op == 11 ? a.i <= b.i : op == 12 ? a.i > b.i : a.i >= b.i;
    return E_NONE;
  }
  if (numeric(a) && numeric(b) && a.tag == b.tag) {
    const double x = num(a), y = num(b);
    *out = op == 10 ? x < y : op == 11 ? x <= y : op == 12 ? x > y : x >= y;
    return E_NONE;
  }
  return E_TYPE;
}

static int list_len(OVal v) { return (int)((uint64_t)v.i & 0xffff); }
static int64_t list_at(OVal v, const int64_t* arena, int j) { return arena[((uint64_t)v.i >> 16) + (uint64_t)j]; }

/* Builtins with scalar and list-view arguments (builtins_cpu.cpp). */
static int builtin(int id, const OVal* a, int argc, const int64_t* arena, OVal* out) {
  switch (id) {
    case 0: /* abs */
      if (argc != 1 || !numeric(a[0])) return E_TYPE;
      *out = a[0].tag == T_FLOAT ? mk_float(canon_bits(f_of(a[0]) < 0 ? -f_of(a[0]) : f_of(a[0]), 32))
                                 : mk_int(a[0].i < 0 ? wneg(a[0].i) : a[0].i);
      return E_NONE;
    case 1: case 2: { /* min, max */
      if (argc != 2 || !numeric(a[0]) || !numeric(a[1]) || a[0].tag != a[1].tag) return E_TYPE;
      if (a[0].tag == T_FLOAT) {
        const double x = f_of(a[0]), y = f_of(a[1]);
        *out = mk_float(canon_bits(id == 1 ? (x <= y ? x : y) : (x >= y ? x : y), 32));
      } else *out = mk_int(id == 1 ? (a[0].i < a[1].i ? a[0].i : a[1].i) : (a[0].i > a[1].i ? a[0].i : a[1].i));
      return E_NONE;
    }
    case 3: { /* clip */
      if (argc != 3 || !numeric(a[0]) || !numeric(a[1]) || !numeric(a[2])) return E_TYPE;
      if (a[0].tag != a[1].tag || a[0].tag != a[2].tag) return E_TYPE;
      if (a[0].tag == T_FLOAT) {
        double x = f_of(a[0]), lo = f_of(a[1]), hi = f_of(a[2]);
        double m = x > lo ? x : lo;
        *out = mk_float(canon_bits(m < hi ? m : hi, 32));
      } else {
        int64_t m = a[0].i > a[1].i ? a[0].i : a[1].i;
        *out = mk_int(m < a[2].i ? m : a[2].i);
      }
      return E_NONE;
    }
    case 4: /* len */
      if (argc != 1 || !is_list(a[0])) return E_TYPE;
      *out = mk_int(list_len(a[0]));
      return E_NONE;
    case 7: case 30: case 31: { /* index, index_nonnegative, index_clamped_or_zero */
      if (argc != 2) return E_TYPE;
      if (a[0].tag != T_INT_LIST) return E_TYPE;
      if (a[1].tag != T_INT) return E_TYPE;
      const int64_t n = list_len(a[0]);
      int64_t j;
      if (id == 31) {
        if (n == 0) { *out = mk_int(0); return E_NONE; }
        j = a[1].i < 0 ? 0 : (a[1].i >= n ? n - 1 : a[1].i);
      } else {
        if (id == 30 && a[1].i < 0) return E_VALUE;
        j = a[1].i;
        if (j < 0) j += n;
        if (j < 0 || j >= n) return E_VALUE;
      }
      *out = mk_int(list_at(a[0], arena, (int)j));
      return E_NONE;
    }
    case 12: /* is_int */
      if (argc != 1) return E_TYPE;
      *out = mk_bool(a[0].tag == T_INT);
      return E_NONE;
    case 13: case 14: case 28: case 29: { /* idiv0, imod0, trunc_div0, trunc_mod0 */
      if (argc != 2 || a[0].tag != T_INT || a[1].tag != T_INT) return E_TYPE;
      const int64_t x = a[0].i, y = a[1].i;
      if (y == 0) { *out = mk_int(0); return E_NONE; }
      if (id == 28 || id == 29) {
        const int ovf = x == INT64_MIN && y == -1;
        *out = mk_int(ovf ? 0 : (id == 28 ? x / y : x % y));
      } else if (id == 13) *out = mk_int(y == -1 ? wneg(x) : x / y);
      else *out = mk_int(pymod(x, y));
      return E_NONE;
    }
    default: return E_NAME; /* unknown or payload-constructing builtin */
  }
}

#define ORC_MAX_STACK 256
#define ORC_MAX_LOCALS 256

/* Runs one program on one case. `inputs`/`set` preset locals 0..n_inputs-1.
 * Returns E_NONE with *result, or an error code. *fuel_left receives the
 * remaining fuel. */
int orc_run(const OProg* p, const OVal* inputs, const uint8_t* set, int n_inputs, const int64_t* arena, int fuel,
            OVal* result, int* fuel_left) {
  OVal stack[ORC_MAX_STACK];
  OVal locals[ORC_MAX_LOCALS];
  uint8_t lset[ORC_MAX_LOCALS];
  int sp = 0, ip = 0, err = E_NONE;
  size_t uncharged = 0;
  if (p->n_locals > ORC_MAX_LOCALS || p->n_locals < 0) return E_VALUE;
  memset(lset, 0, sizeof lset);
  for (int k = 0; k < n_inputs; ++k)
    if (k < p->n_locals && (!set || set[k])) { locals[k] = inputs[k]; lset[k] = 1; }
  while (ip < p->code_len) {
    if (p->fuel) {
      const uint32_t cost = p->fuel[ip];
      if (cost > 0x7fffffffu) { err = E_VALUE; break; }
      if (fuel < (int)cost) { err = E_TIMEOUT; break; }
      fuel -= (int)cost;
      if (cost) uncharged = 0;
      else if (++uncharged > (size_t)p->code_len) { err = E_VALUE; break; }
    } else {
      if (fuel <= 0) { err = E_TIMEOUT; break; }
      fuel -= 1;
    }
    const OIns ins = p->code[ip++];
    if (sp >= ORC_MAX_STACK - 1) { err = E_VALUE; break; }
    switch (ins.op) {
      case 0: /* PUSH_CONST */
        if (!ins.has_a || ins.a < 0 || ins.a >= p->n_consts) { err = E_VALUE; goto done; }
        stack[sp++] = p->consts[ins.a];
        break;
      case 1: /* LOAD */
        if (!ins.has_a || ins.a < 0 || ins.a >= p->n_locals) { err = E_NAME; goto done; }
        if (!lset[ins.a]) { err = E_NAME; goto done; }
        stack[sp++] = locals[ins.a];
        break;
      case 2: /* STORE */
        if (!ins.has_a || ins.a < 0 || ins.a >= p->n_locals) { err = E_NAME; goto done; }
        if (!sp) { err = E_VALUE; goto done; }
        locals[ins.a] = stack[--sp];
        lset[ins.a] = 1;
        break;
      case 21: /* CHECK_LIST */
        if (!sp) { err = E_VALUE; goto done; }
        if (!is_list(stack[sp - 1])) { err = E_TYPE; goto done; }
        break;
      case 22: /* CHECK_INT */
        if (!sp) { err = E_VALUE; goto done; }
        if (stack[sp - 1].tag != T_INT) { err = E_TYPE; goto done; }
        break;
      case 3: case 4: { /* NEG, NOT */
        if (!sp) { err = E_VALUE; goto done; }
        const OVal x = stack[--sp];
        if (ins.op == 3) {
          if (!numeric(x)) { err = E_TYPE; goto done; }
          stack[sp++] = x.tag == T_FLOAT ? mk_float(canon_bits(-f_of(x), 32)) : mk_int(wneg(x.i));
        } else {
          if (x.tag != T_BOOL) { err = E_TYPE; goto done; }
          stack[sp++] = mk_bool(!x.b);
        }
        break;
      }
      case 5: case 6: case 7: case 8: case 9: { /* ADD SUB MUL DIV MOD */
        if (sp < 2) { err = E_VALUE; goto done; }
        const OVal b = stack[--sp], a = stack[--sp];
        if (!numeric(a) || !numeric(b) || a.tag != b.tag) { err = E_TYPE; goto done; }
        const int fl = a.tag == T_FLOAT;
        const double x = num(a), y = num(b);
        if ((ins.op == 8 || ins.op == 9) && y == 0.0) { err = E_ZERO_DIV; goto done; }
        if (ins.op == 5) stack[sp++] = fl ? mk_float(canon_bits(x + y, 32)) : mk_int(wadd(a.i, b.i));
        else if (ins.op == 6) stack[sp++] = fl ? mk_float(canon_bits(x - y, 32)) : mk_int(wsub(a.i, b.i));
        else if (ins.op == 7) stack[sp++] = fl ? mk_float(canon_bits(x * y, 32)) : mk_int(wmul(a.i, b.i));
        else if (ins.op == 8) stack[sp++] = mk_float(canon_bits(x / y, 32));
        else stack[sp++] = fl ? mk_float(canon_bits(pyfmod(x, y), 32)) : mk_int(pymod(a.i, b.i));
        break;
      }
      case 10: case 11: case 12: case 13: case 14: case 15: {
        if (sp < 2) { err = E_VALUE; goto done; }
        const OVal b = stack[--sp], a = stack[--sp];
        int r = 0;
        const int e = compare(ins.op, a, b, &r);
        if (e != E_NONE) { err = e; goto done; }
        stack[sp++] = mk_bool(r);
        break;
      }
      case 16: /* JMP */
        if (!ins.has_a || ins.a < 0 || ins.a > p->code_len) { err = E_VALUE; goto done; }
        ip = ins.a;
        break;
      case 17: case 18: { /* JMP_IF_FALSE, JMP_IF_TRUE */
        if (!sp) { err = E_VALUE; goto done; }
        if (!ins.has_a || ins.a < 0 || ins.a > p->code_len) { err = E_VALUE; goto done; }
        const OVal c = stack[--sp];
        if (c.tag != T_BOOL) { err = E_TYPE; goto done; }
        if ((ins.op == 17 && !c.b) || (ins.op == 18 && c.b)) ip = ins.a;
        break;
      }
      case 19: { /* CALL_BUILTIN */
        const int argc = ins.has_b ? ins.b : -1;
        if (argc < 0) { err = E_TYPE; goto done; }
        if (sp < argc) { err = E_VALUE; goto done; }
        OVal out;
        const int e = builtin(ins.has_a ? ins.a : -1, stack + sp - argc, argc, arena, &out);
        sp -= argc;
        if (e != E_NONE) { err = e; goto done; }
        stack[sp++] = out;
        break;
      }
      case 20: /* RETURN */
        if (!sp) { err = E_VALUE; goto done; }
        *result = stack[sp - 1];
        *fuel_left = fuel;
        return E_NONE;
      default: err = E_VALUE; goto done;
    }
  }
  if (err == E_NONE) err = E_VALUE; /* program finished without return */
done:
  *fuel_left = fuel;
  return err;
}

/* Per-case score and exactness for the scalar metrics. metric: 0 legacy,
 * 1 absolute, 2 exact; proximity: 0 capped, 1 discounted. */
static double ceiling(int metric, int proximity, double penalty) {
  return (metric == 2 || (metric != 0 && proximity == 1)) ? 1.0 : fabs(penalty);
}
double orc_case_score(OVal actual, OVal expected, double penalty, int metric, int proximity, int* exact) {
  int ex;
  if (actual.tag == T_INT && expected.tag == T_INT) ex = actual.i == expected.i;
  else if (numeric(actual) && numeric(expected)) ex = isfinite(num(actual)) && isfinite(num(expected)) && num(actual) == num(expected);
  else if (actual.tag != expected.tag || actual.tag == T_INVALID || actual.tag == T_FALLBACK) ex = 0;
  else ex = actual.tag == T_BOOL ? actual.b == expected.b : actual.i == expected.i;
  *exact = ex;
  if (metric == 2) return ex ? 0.0 : -1.0;
  if (metric == 0) {
    const double mag = fabs(penalty);
    if (numeric(actual) && numeric(expected)) {
      const double d = num(actual) - num(expected);
      if (!isfinite(num(actual)) || !isfinite(num(expected)) || !isfinite(d)) return -mag;
      return -fmin(fabs(d), mag);
    }
    if (numeric(expected) || actual.tag != expected.tag) return -mag;
    if (actual.tag == T_BOOL) return actual.b == expected.b ? 1.0 : 0.0;
    if (actual.tag == T_CHAR || actual.tag == T_STRING || is_list(actual)) return actual.i == expected.i ? 1.0 : 0.0;
    return -mag;
  }
  if (!numeric(actual) || !numeric(expected) || !isfinite(num(actual)) || !isfinite(num(expected)))
    return -ceiling(metric, proximity, penalty);
  double err;
  if (actual.tag == T_INT && expected.tag == T_INT)
    err = (double)(actual.i >= expected.i ? (uint64_t)actual.i - (uint64_t)expected.i : (uint64_t)expected.i - (uint64_t)actual.i);
  else err = fabs(num(actual) - num(expected));
  if (!isfinite(err)) return -ceiling(metric, proximity, penalty);
  return proximity == 1 ? -err / (1.0 + err) : -fmin(err, fabs(penalty));
}

/* eval_fitness_cpu restated: programs x cases, L logical lanes, canonical
 * reduction. counts[p*3+0..2] = cases, errors, timeouts. */
void orc_eval_fitness(const OProg* programs, int n_programs, const OVal* inputs, const uint8_t* set, int n_cases, int n_inputs,
                      const OVal* answers, const int64_t* arena, int fuel, double penalty, int lanes, int metric, int proximity,
                      double* fitness, uint32_t* wrong, uint32_t* counts) {
  if (lanes < 1) lanes = 1;
  double* partial = (double*)malloc(sizeof(double) * (size_t)lanes);
  for (int p = 0; p < n_programs; ++p) {
    uint32_t w = 0, errors = 0, timeouts = 0;
    for (int l = 0; l < lanes; ++l) {
      const size_t begin = ((size_t)n_cases * (size_t)l) / (size_t)lanes;
      const size_t end = ((size_t)n_cases * (size_t)(l + 1)) / (size_t)lanes;
      double local = 0.0;
      for (size_t c = begin; c < end; ++c) {
        OVal out;
        int left = 0;
        const int e = orc_run(&programs[p], inputs + c * (size_t)n_inputs, set ? set + c * (size_t)n_inputs : NULL, n_inputs,
                              arena, fuel, &out, &left);
        if (e != E_NONE) {
          ++errors;
          if (e == E_TIMEOUT) ++timeouts;
          ++w;
          local = canon_bits(local - ceiling(metric, proximity, penalty), 48);
          continue;
        }
        int ex = 0;
        local = canon_bits(local + orc_case_score(out, answers[c], penalty, metric, proximity, &ex), 48);
        if (!ex) ++w;
      }
      partial[l] = local;
    }
    double score = 0.0;
    if (metric == 0) {
      for (int l = 0; l < lanes; ++l) score = canon_bits(score + partial[l], 48);
    } else {
      for (int base = 0; base < lanes; base += 32) {
        double group = 0.0;
        for (int l = base; l < base + 32 && l < lanes; ++l) group = canon_bits(group + partial[l], 48);
        score = canon_bits(score + group, 48);
      }
    }
    fitness[p] = score;
    if (wrong) wrong[p] = metric == 0 ? 0 : w;
    if (counts) { counts[p * 3 + 0] = (uint32_t)n_cases; counts[p * 3 + 1] = errors; counts[p * 3 + 2] = timeouts; }
  }
  free(partial);
}

/* SplitMix64 stream: writes n draws of next(). */
void orc_splitmix_next(uint64_t seed, uint64_t* out, int n) {
  uint64_t s = seed;
  for (int k = 0; k < n; ++k) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    out[k] = z ^ (z >> 31);
  }
}
/* bounded(n) draws: rejection below (2^64 - n) mod n. */
void orc_splitmix_bounded(uint64_t seed, uint64_t bound, uint64_t* out, int n) {
  uint64_t s = seed;
  const uint64_t threshold = (0ULL - bound) % bound;
  for (int k = 0; k < n; ++k) {
    for (;;) {
      uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      z ^= z >> 31;
      if (z >= threshold) { out[k] = z % bound; break; }
    }
  }
}

/* std::mt19937_64: n outputs from seed. */
void orc_mt19937_64(uint64_t seed, uint64_t* out, int n) {
  enum { N = 312, M = 156 };
  uint64_t mt[N];
  int idx = N;
  mt[0] = seed;
  for (int i = 1; i < N; ++i) mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
  for (int k = 0; k < n; ++k) {
    if (idx >= N) {
      for (int i = 0; i < N; ++i) {
        const uint64_t x = (mt[i] & 0xFFFFFFFF80000000ULL) | (mt[(i + 1) % N] & 0x7FFFFFFFULL);
        uint64_t xa = x >> 1;
        if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
        mt[i] = mt[(i + M) % N] ^ xa;
      }
      idx = 0;
    }
    uint64_t y = mt[idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    out[k] = y;
  }
}
