// Value model, opcodes and scalar semantics shared by host and device code.
//
// These are the execution rules of the GAGP bytecode ISA (reference:
// spec/bytecode_isa.md, spec/builtins_base.md) restated for this engine:
//   * Value: cpp/include/gagp/core/value.hpp:14-241 (tag numbering kept, since
//     tags are part of the C-ABI value layout and of constant identity),
//   * int64 wraparound arithmetic, Python modulo, comparisons:
//     cpp/include/gagp/core/value_semantics.hpp:91-191,
//   * per-case scoring: cpp/include/gagp/core/fitness_policy.hpp:28-59,
//   * 48-bit accumulator rounding: cpp/src/runtime/cpu/fitness_cpu.cpp:44-53,
//   * 40-bit ranking rounding: cpp/src/evolution/selection.cpp:10-20.
#pragma once
#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GHD __host__ __device__
#else
#define GHD
#endif

// Device memory referenced from structs: in the device pass the pointers are
// typed as global memory, so loads and stores through them are global (not
// flat) instructions, which do not serialize with LDS traffic.
#if defined(__HIP_DEVICE_COMPILE__)
#define GAS __attribute__((address_space(1)))
#define LAS __attribute__((address_space(3)))
#else
#define GAS
#define LAS
#endif

namespace gagp_amd {

template <class T>
GHD inline GAS T* gbl(T* p) { return (GAS T*)p; }

enum Tag : std::uint8_t {
  TAG_INT = 0, TAG_FLOAT = 1, TAG_BOOL = 2, TAG_CHAR = 3, TAG_STRING = 4,
  TAG_INT_LIST = 5, TAG_FLOAT_LIST = 6, TAG_STRING_LIST = 7, TAG_FALLBACK = 8, TAG_INVALID = 9,
};
// Completion marker of a statement phase; never a runtime tag.
constexpr std::uint8_t TAG_STATEMENT = 0xF0;

// 16-byte value: payload word, Bool flag, tag (reference field order).
struct alignas(16) Val {
  std::int64_t i;
  std::uint8_t b;
  std::uint8_t tag;
  std::uint8_t pad[6];
};
static_assert(sizeof(Val) == 16, "Val is 16 bytes");

GHD inline Val val_int(std::int64_t v) { Val x{}; x.i = v; x.b = 0; x.tag = TAG_INT; return x; }
GHD inline Val val_bool(bool v) { Val x{}; x.i = 0; x.b = v ? 1 : 0; x.tag = TAG_BOOL; return x; }
GHD inline Val val_invalid() { Val x{}; x.i = 0; x.b = 0; x.tag = TAG_INVALID; return x; }
GHD inline double val_f(const Val& v) { union { std::int64_t i; double f; } u{v.i}; return u.f; }
GHD inline Val val_float(double f) { Val x{}; union { double f; std::int64_t i; } u{f}; x.i = u.i; x.tag = TAG_FLOAT; return x; }

// Opcode numbers are fixed by the ISA (spec/bytecode_isa.md "Numeric opcode values").
enum Op : std::uint8_t {
  OP_PUSH_CONST = 0, OP_LOAD = 1, OP_STORE = 2, OP_NEG = 3, OP_NOT = 4, OP_ADD = 5, OP_SUB = 6,
  OP_MUL = 7, OP_DIV = 8, OP_MOD = 9, OP_LT = 10, OP_LE = 11, OP_GT = 12, OP_GE = 13, OP_EQ = 14,
  OP_NE = 15, OP_JMP = 16, OP_JMP_IF_FALSE = 17, OP_JMP_IF_TRUE = 18, OP_CALL_BUILTIN = 19,
  OP_RETURN = 20, OP_CHECK_LIST = 21, OP_CHECK_INT = 22, OP_EMPTY_LIST = 23,
  OP_EMPTY_LIST_LIKE = 24, OP_BOUNDED_REGION = 28,
};

// Builtin IDs (spec/bytecode_isa.md "Builtin IDs").
enum Builtin : int {
  B_ABS = 0, B_MIN = 1, B_MAX = 2, B_CLIP = 3, B_LEN = 4, B_CONCAT = 5, B_SLICE = 6, B_INDEX = 7,
  B_APPEND = 8, B_REVERSE = 9, B_FIND = 10, B_CONTAINS = 11, B_IS_INT = 12, B_IDIV0 = 13,
  B_IMOD0 = 14, B_PREPEND = 15, B_TRUNC_DIV0 = 28, B_TRUNC_MOD0 = 29, B_INDEX_NONNEG = 30,
  B_INDEX_CLAMPED = 31,
};

// Runtime error codes, in the reference's enumeration order (errors.hpp:7-13).
enum Err : int { ERR_NONE = -1, ERR_NAME = 0, ERR_TYPE = 1, ERR_ZERO_DIV = 2, ERR_VALUE = 3, ERR_TIMEOUT = 4 };

// Fitness metrics and proximity (fitness_policy.hpp:4-15).
enum Metric : int { METRIC_LEGACY = 0, METRIC_ABSOLUTE = 1, METRIC_EXACT = 2, METRIC_LEVENSHTEIN = 3, METRIC_TUPLE = 4 };
enum Proximity : int { PROX_CAPPED = 0, PROX_DISCOUNTED = 1 };

struct Policy {
  int metric;
  int proximity;
  int summary_count;
  double summary_weight;
  double summary_cap;
};

// ---- int64 semantics (two's complement wraparound) ----
GHD inline std::int64_t wrap_add(std::int64_t a, std::int64_t b) { return (std::int64_t)((std::uint64_t)a + (std::uint64_t)b); }
GHD inline std::int64_t wrap_sub(std::int64_t a, std::int64_t b) { return (std::int64_t)((std::uint64_t)a - (std::uint64_t)b); }
GHD inline std::int64_t wrap_mul(std::int64_t a, std::int64_t b) { return (std::int64_t)((std::uint64_t)a * (std::uint64_t)b); }
GHD inline std::int64_t wrap_neg(std::int64_t a) { return (std::int64_t)(0ULL - (std::uint64_t)a); }
// Python modulo; b == -1 gives 0 so INT64_MIN % -1 never traps. b != 0.
GHD inline std::int64_t py_mod(std::int64_t a, std::int64_t b) {
  if (b == -1) return 0;
  std::int64_t r = a % b;
  if (r != 0 && ((r < 0) != (b < 0))) r += b;
  return r;
}
// idiv0 for a nonzero divisor: C++ truncation, the overflow pair wrapping to
// INT64_MIN (value_semantics.hpp:114-117, builtins_cpu.cpp:198-217).
GHD inline std::int64_t wrap_div(std::int64_t a, std::int64_t b) { return b == -1 ? wrap_neg(a) : a / b; }
GHD inline std::int64_t trunc_div0(std::int64_t a, std::int64_t b) {
  return (b == 0 || (a == INT64_MIN && b == -1)) ? 0 : a / b;
}
GHD inline std::int64_t trunc_mod0(std::int64_t a, std::int64_t b) {
  return (b == 0 || (a == INT64_MIN && b == -1)) ? 0 : a % b;
}

// Round `value` to a `bits`-bit mantissa, ties away from zero (frexp/llround/ldexp).
// Exact bit manipulation so host and device agree without libm.
GHD inline double round_mantissa(double value, int bits) {
  union { double f; std::uint64_t u; } x{value};
  const int biased = int((x.u >> 52) & 0x7ff);
  if (biased == 0x7ff || value == 0.0) return value == 0.0 ? 0.0 : value;  // inf/nan pass, -0 -> +0
  // Significand bits beyond the kept `bits` (53-bit normal significand incl. hidden bit).
  int drop = 53 - bits;
  if (biased == 0) {
    // Subnormal: frexp normalizes, so the kept width is relative to the leading bit.
    std::uint64_t m = x.u & ((1ULL << 52) - 1);
    int lead = 63 - __builtin_clzll(m);  // position of the leading one (< 52)
    drop = (lead + 1) - bits;
    if (drop <= 0) return value;
    const std::uint64_t sign = x.u & (1ULL << 63);
    std::uint64_t half = 1ULL << (drop - 1);
    std::uint64_t q = (m + half) >> drop;  // ties away from zero on the magnitude
    std::uint64_t r = q << drop;
    union { std::uint64_t u; double f; } y{sign | r};
    return y.f;
  }
  const std::uint64_t sign = x.u & (1ULL << 63);
  std::uint64_t m = (x.u & ((1ULL << 52) - 1)) | (1ULL << 52);
  std::uint64_t half = 1ULL << (drop - 1);
  std::uint64_t q = (m + half) >> drop;
  std::uint64_t r = q << drop;  // may carry into bit 53
  int e = biased;
  if (r >> 53) { r >>= 1; ++e; }
  if (e >= 0x7ff) { union { std::uint64_t u; double f; } inf{sign | (0x7ffULL << 52)}; return inf.f; }
  union { std::uint64_t u; double f; } y{sign | ((std::uint64_t)e << 52) | (r & ((1ULL << 52) - 1))};
  return y.f;
}
GHD inline double canon48(double v) { return round_mantissa(v, 48); }
GHD inline double canon40(double v) { return round_mantissa(v, 40); }
GHD inline double canon32(double v) { return round_mantissa(v, 32); }

GHD inline double fabs_(double x) { return x < 0 ? -x : x; }
GHD inline bool finite_(double x) { union { double f; std::uint64_t u; } y{x}; return ((y.u >> 52) & 0x7ff) != 0x7ff; }
GHD inline double fmin_(double a, double b) { return a < b ? a : b; }

GHD inline double error_ceiling(const Policy& p, double penalty) {
  if (p.metric == METRIC_TUPLE) return 1.0 + p.summary_count * p.summary_weight;
  return (p.metric == METRIC_EXACT || (p.metric != METRIC_LEGACY && p.proximity == PROX_DISCOUNTED)) ? 1.0 : fabs_(penalty);
}
GHD inline double integer_gap(std::int64_t a, std::int64_t b) {
  return double(a >= b ? (std::uint64_t)a - (std::uint64_t)b : (std::uint64_t)b - (std::uint64_t)a);
}
GHD inline bool is_numeric(std::uint8_t t) { return t == TAG_INT || t == TAG_FLOAT; }
GHD inline double num_of(const Val& v) { return v.tag == TAG_FLOAT ? val_f(v) : double(v.i); }

// Scalar (non-sequence) case score for the Absolute / Exact metrics and the
// exactness flag used for wrong counts (fitness_policy.hpp:37-59,
// fitness_cpu.cpp:19-23). Legacy scoring is value_semantics.hpp:193-237.
GHD inline double case_score(const Val& actual, const Val& expected, double penalty, const Policy& p, bool& exact) {
  // exactness under the Exact metric
  {
    bool e;
    if (actual.tag == TAG_INT && expected.tag == TAG_INT) e = actual.i == expected.i;
    else if (is_numeric(actual.tag) && is_numeric(expected.tag)) {
      const double a = num_of(actual), b = num_of(expected);
      e = finite_(a) && finite_(b) && a == b;
    } else if (actual.tag != expected.tag || actual.tag == TAG_INVALID || actual.tag == TAG_FALLBACK) e = false;
    else e = actual.tag == TAG_BOOL ? actual.b == expected.b : actual.i == expected.i;
    exact = e;
  }
  if (p.metric == METRIC_EXACT) return exact ? 0.0 : -1.0;
  if (p.metric == METRIC_LEGACY) {
    const double mag = fabs_(penalty);
    if (is_numeric(actual.tag) && is_numeric(expected.tag)) {
      const double a = num_of(actual), b = num_of(expected), d = a - b;
      if (!finite_(a) || !finite_(b) || !finite_(d)) return -mag;
      return -fmin_(fabs_(d), mag);
    }
    if (is_numeric(expected.tag)) return -mag;
    if (actual.tag != expected.tag) return -mag;
    if (actual.tag == TAG_BOOL) return actual.b == expected.b ? 1.0 : 0.0;
    if (actual.tag == TAG_CHAR || actual.tag == TAG_STRING || actual.tag == TAG_INT_LIST ||
        actual.tag == TAG_FLOAT_LIST || actual.tag == TAG_STRING_LIST) return actual.i == expected.i ? 1.0 : 0.0;
    return -mag;
  }
  // Absolute (scalar path)
  if (!is_numeric(actual.tag) || !is_numeric(expected.tag)) return -error_ceiling(p, penalty);
  const double a = num_of(actual), b = num_of(expected);
  if (!finite_(a) || !finite_(b)) return -error_ceiling(p, penalty);
  const double err = (actual.tag == TAG_INT && expected.tag == TAG_INT) ? integer_gap(actual.i, expected.i) : fabs_(a - b);
  if (!finite_(err)) return -error_ceiling(p, penalty);
  return p.proximity == PROX_DISCOUNTED ? -err / (1.0 + err) : -fmin_(err, fabs_(penalty));
}

// ---- SplitMix64 draws (cpp/src/evolution/repro/gpu/device/grammar_random.cuh:10-51) ----
struct SplitMix {
  std::uint64_t s;
  GHD explicit SplitMix(std::uint64_t seed) : s(seed) {}
  GHD std::uint64_t next() {
    std::uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  // Rejection below (-n mod n): uniform over [0, n). n > 0.
  GHD std::uint64_t bounded(std::uint64_t n) {
    const std::uint64_t threshold = (0ULL - n) % n;
    for (;;) { const std::uint64_t v = next(); if (v >= threshold) return v % n; }
  }
  // Uniform over [lo, hi] through the sign-flipped unsigned width.
  GHD std::int64_t integer(std::int64_t lo, std::int64_t hi) {
    const std::uint64_t sign = 1ULL << 63;
    const std::uint64_t first = (std::uint64_t)lo ^ sign, last = (std::uint64_t)hi ^ sign;
    const std::uint64_t width = last - first + 1;
    const std::uint64_t bits = (first + (width ? bounded(width) : next())) ^ sign;
    if (bits < sign) return (std::int64_t)bits;
    return INT64_MIN + (std::int64_t)(bits - sign);
  }
  // 53-bit uniform in [0,1).
  GHD double unit() { return double(next() >> 11) * 0x1.0p-53; }
};
GHD inline std::uint64_t splitmix_mix(std::uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

}  // namespace gagp_amd
