// Minimal strict JSON reader and canonical writer for the task, grammar,
// fitness-case, search-config and result documents.
//
// Canonical output follows the reference writer
// (cpp/src/evolution/grammar/definition.cpp:339-362): object keys in byte
// order, no whitespace, numbers printed with 17 significant digits in the
// shortest "%g" form, strings escaped for quote, backslash and controls.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace gagp_amd::json {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object };
  Kind kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  bool has(const std::string& k) const { return kind == Object && obj.count(k); }
  const Value& at(const std::string& k) const {
    auto it = obj.find(k);
    if (kind != Object || it == obj.end()) throw std::invalid_argument("missing JSON field: " + k);
    return it->second;
  }
  const std::string& as_string(const char* what) const {
    if (kind != String) throw std::invalid_argument(std::string(what) + " must be a string");
    return str;
  }
  const std::vector<Value>& as_array(const char* what) const {
    if (kind != Array) throw std::invalid_argument(std::string(what) + " must be an array");
    return arr;
  }
  double as_number(const char* what) const {
    if (kind != Number) throw std::invalid_argument(std::string(what) + " must be a number");
    return num;
  }
};

inline Value make_string(std::string s) { Value v; v.kind = Value::String; v.str = std::move(s); return v; }
inline Value make_number(double d) { Value v; v.kind = Value::Number; v.num = d; return v; }
inline Value make_bool(bool b) { Value v; v.kind = Value::Bool; v.b = b; return v; }
inline Value make_array() { Value v; v.kind = Value::Array; return v; }
inline Value make_object() { Value v; v.kind = Value::Object; return v; }

class Parser {
 public:
  explicit Parser(const std::string& text) : t_(text) {}
  Value parse() {
    Value v = value(0);
    ws();
    if (p_ != t_.size()) fail("trailing characters");
    return v;
  }

 private:
  [[noreturn]] void fail(const char* why) const {
    throw std::invalid_argument(std::string("JSON parse error: ") + why + " at offset " + std::to_string(p_));
  }
  void ws() { while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\r' || t_[p_] == '\t')) ++p_; }
  Value value(int depth) {
    if (depth > 512) fail("nesting too deep");
    ws();
    if (p_ >= t_.size()) fail("unexpected end");
    const char c = t_[p_];
    if (c == '{') return object(depth);
    if (c == '[') return array(depth);
    if (c == '"') return make_string(string());
    if (c == 't') { word("true"); return make_bool(true); }
    if (c == 'f') { word("false"); return make_bool(false); }
    if (c == 'n') { word("null"); return Value{}; }
    return number();
  }
  void word(const char* w) {
    const size_t n = std::strlen(w);
    if (t_.compare(p_, n, w) != 0) fail("bad literal");
    p_ += n;
  }
  Value object(int depth) {
    Value v = make_object();
    ++p_;
    ws();
    if (p_ < t_.size() && t_[p_] == '}') { ++p_; return v; }
    for (;;) {
      ws();
      if (p_ >= t_.size() || t_[p_] != '"') fail("expected key");
      std::string k = string();
      ws();
      if (p_ >= t_.size() || t_[p_] != ':') fail("expected ':'");
      ++p_;
      Value item = value(depth + 1);
      if (!v.obj.emplace(k, std::move(item)).second) fail("duplicate key");
      ws();
      if (p_ < t_.size() && t_[p_] == ',') { ++p_; continue; }
      if (p_ < t_.size() && t_[p_] == '}') { ++p_; return v; }
      fail("expected ',' or '}'");
    }
  }
  Value array(int depth) {
    Value v = make_array();
    ++p_;
    ws();
    if (p_ < t_.size() && t_[p_] == ']') { ++p_; return v; }
    for (;;) {
      v.arr.push_back(value(depth + 1));
      ws();
      if (p_ < t_.size() && t_[p_] == ',') { ++p_; continue; }
      if (p_ < t_.size() && t_[p_] == ']') { ++p_; return v; }
      fail("expected ',' or ']'");
    }
  }
  static void put_utf8(std::string& out, unsigned cp) {
    if (cp < 0x80) out += char(cp);
    else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) { out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
    else { out += char(0xF0 | (cp >> 18)); out += char(0x80 | ((cp >> 12) & 0x3F)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
  }
  unsigned hex4() {
    if (p_ + 4 > t_.size()) fail("bad escape");
    unsigned v = 0;
    for (int i = 0; i < 4; ++i) {
      const char c = t_[p_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= c - '0';
      else if (c >= 'a' && c <= 'f') v |= c - 'a' + 10;
      else if (c >= 'A' && c <= 'F') v |= c - 'A' + 10;
      else fail("bad hex escape");
    }
    return v;
  }
  std::string string() {
    ++p_;
    std::string out;
    for (;;) {
      if (p_ >= t_.size()) fail("unterminated string");
      const char c = t_[p_++];
      if (c == '"') return out;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c != '\\') { out += c; continue; }
      if (p_ >= t_.size()) fail("bad escape");
      const char e = t_[p_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (p_ + 2 > t_.size() || t_[p_] != '\\' || t_[p_ + 1] != 'u') fail("lone surrogate");
            p_ += 2;
            const unsigned lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) fail("bad surrogate pair");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) fail("lone surrogate");
          put_utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
  }
  Value number() {
    const size_t start = p_;
    if (p_ < t_.size() && t_[p_] == '-') ++p_;
    if (p_ >= t_.size() || !(t_[p_] >= '0' && t_[p_] <= '9')) fail("bad number");
    if (t_[p_] == '0') ++p_;
    else while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') ++p_;
    if (p_ < t_.size() && t_[p_] == '.') {
      ++p_;
      if (p_ >= t_.size() || !(t_[p_] >= '0' && t_[p_] <= '9')) fail("bad fraction");
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') ++p_;
    }
    if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
      ++p_;
      if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) ++p_;
      if (p_ >= t_.size() || !(t_[p_] >= '0' && t_[p_] <= '9')) fail("bad exponent");
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') ++p_;
    }
    const std::string s = t_.substr(start, p_ - start);
    const double d = std::strtod(s.c_str(), nullptr);
    if (!std::isfinite(d)) fail("non-finite number");
    return make_number(d);
  }
  const std::string& t_;
  size_t p_ = 0;
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

inline std::string number_text(double d) {
  if (!std::isfinite(d)) throw std::invalid_argument("non-finite JSON number");
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.17g", d);
  return buf;
}
inline void quoted(std::string& out, const std::string& s) {
  out += '"';
  for (unsigned char c : s) {
    if (c == '"') out += "\\\"";
    else if (c == '\\') out += "\\\\";
    else if (c < 0x20) {
      switch (c) {
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        default: { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", c); out += b; }
      }
    } else out += char(c);
  }
  out += '"';
}
inline void emit(std::string& out, const Value& v) {
  switch (v.kind) {
    case Value::Null: out += "null"; break;
    case Value::Bool: out += v.b ? "true" : "false"; break;
    case Value::Number: out += number_text(v.num); break;
    case Value::String: quoted(out, v.str); break;
    case Value::Array:
      out += '[';
      for (size_t i = 0; i < v.arr.size(); ++i) { if (i) out += ','; emit(out, v.arr[i]); }
      out += ']';
      break;
    case Value::Object: {
      out += '{';
      bool first = true;
      for (const auto& kv : v.obj) {
        if (!first) out += ',';
        first = false;
        quoted(out, kv.first);
        out += ':';
        emit(out, kv.second);
      }
      out += '}';
      break;
    }
  }
}
inline std::string canonical(const Value& v) { std::string s; emit(s, v); return s; }

}  // namespace gagp_amd::json
