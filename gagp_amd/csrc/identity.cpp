// Grammar identity (see identity.hpp).
#include "identity.hpp"

#include <algorithm>
#include <array>
#include <cstdint>
#include <set>
#include <stdexcept>

namespace gagp_amd {
namespace {

// The first n primes.
std::vector<std::uint32_t> primes(int n) {
  std::vector<std::uint32_t> out;
  for (std::uint32_t k = 2; int(out.size()) < n; ++k) {
    bool prime = true;
    for (std::uint32_t p : out) {
      if (p * p > k) break;
      if (k % p == 0) { prime = false; break; }
    }
    if (prime) out.push_back(k);
  }
  return out;
}

// floor(x^(1/root)) for x < 2^110 and root 2 or 3, by bisection in 128-bit
// integers.
unsigned __int128 integer_root(unsigned __int128 x, int root) {
  unsigned __int128 lo = 0, hi = (unsigned __int128)1 << (root == 2 ? 56 : 38);
  while (lo < hi) {
    const unsigned __int128 mid = (lo + hi + 1) / 2;
    unsigned __int128 power = mid * mid;
    if (root == 3) power *= mid;
    if (power <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// FIPS 180-4 section 4.2.2 and 5.3.3: the first 32 bits of the fractional
// parts of the cube roots of the first 64 primes (round constants) and of
// the square roots of the first 8 primes (initial hash value), computed
// exactly: floor(frac(p^(1/r)) * 2^32) is the low 32 bits of
// floor((p * 2^(32 r))^(1/r)).
struct Sha256Constants {
  std::array<std::uint32_t, 64> k{};
  std::array<std::uint32_t, 8> h0{};
  Sha256Constants() {
    const auto p = primes(64);
    for (int i = 0; i < 64; ++i) k[std::size_t(i)] = std::uint32_t(integer_root((unsigned __int128)p[std::size_t(i)] << 96, 3));
    for (int i = 0; i < 8; ++i) h0[std::size_t(i)] = std::uint32_t(integer_root((unsigned __int128)p[std::size_t(i)] << 64, 2));
  }
};

const Sha256Constants& sha_constants() {
  static const Sha256Constants c;
  return c;
}

inline std::uint32_t rotr(std::uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// One 64-byte block (FIPS 180-4 section 6.2.2).
void sha256_block(std::array<std::uint32_t, 8>& h, const unsigned char* block) {
  const auto& k = sha_constants().k;
  std::uint32_t w[64];
  for (int t = 0; t < 16; ++t)
    w[t] = std::uint32_t(block[4 * t]) << 24 | std::uint32_t(block[4 * t + 1]) << 16 | std::uint32_t(block[4 * t + 2]) << 8 |
           std::uint32_t(block[4 * t + 3]);
  for (int t = 16; t < 64; ++t) {
    const std::uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
    const std::uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
    w[t] = s1 + w[t - 7] + s0 + w[t - 16];
  }
  std::array<std::uint32_t, 8> v = h;  // a..h
  for (int t = 0; t < 64; ++t) {
    const std::uint32_t e = v[4], a = v[0];
    const std::uint32_t t1 = v[7] + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & v[5]) ^ (~e & v[6])) + k[std::size_t(t)] + w[t];
    const std::uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & v[1]) ^ (a & v[2]) ^ (v[1] & v[2]));
    for (int j = 7; j > 0; --j) v[std::size_t(j)] = v[std::size_t(j - 1)];
    v[4] += t1;
    v[0] = t1 + t2;
  }
  for (int j = 0; j < 8; ++j) h[std::size_t(j)] += v[std::size_t(j)];
}

using json::Value;

std::string id_of(const Value& v) {
  if (v.kind != Value::Object || !v.has("id") || v.at("id").kind != Value::String)
    throw std::invalid_argument("grammar resource without an id");
  return v.at("id").str;
}

// Sorts a resource array by id (ids must be distinct).
void sort_by_id(Value& arr) {
  if (arr.kind != Value::Array) throw std::invalid_argument("expected grammar array");
  std::set<std::string> seen;
  for (const auto& v : arr.arr)
    if (!seen.insert(id_of(v)).second) throw std::invalid_argument("duplicate grammar ID: " + id_of(v));
  std::sort(arr.arr.begin(), arr.arr.end(), [](const Value& a, const Value& b) { return id_of(a) < id_of(b); });
}

}  // namespace

std::string sha256_hex(const std::string& bytes) {
  std::array<std::uint32_t, 8> h = sha_constants().h0;
  const std::size_t n = bytes.size();
  std::size_t at = 0;
  for (; n - at >= 64; at += 64) sha256_block(h, reinterpret_cast<const unsigned char*>(bytes.data()) + at);
  // Padding: 0x80, zeros, the 64-bit big-endian bit length, to a multiple of
  // 64 bytes (one or two final blocks).
  unsigned char tail[128] = {};
  const std::size_t rest = n - at;
  std::copy(bytes.begin() + std::ptrdiff_t(at), bytes.end(), tail);
  tail[rest] = 0x80;
  const std::size_t len = rest + 9 <= 64 ? 64 : 128;
  const std::uint64_t bits = std::uint64_t(n) * 8;
  for (int i = 0; i < 8; ++i) tail[len - 1 - std::size_t(i)] = static_cast<unsigned char>(bits >> (8 * i));
  for (std::size_t b = 0; b < len; b += 64) sha256_block(h, tail + b);
  static const char* hex = "0123456789abcdef";
  std::string out;
  for (std::uint32_t word : h)
    for (int s = 28; s >= 0; s -= 4) out += hex[(word >> s) & 15];
  return out;
}

std::string grammar_content_hash(const Value& grammar_document) {
  if (grammar_document.kind != Value::Object) throw std::invalid_argument("grammar must be an object");
  if (grammar_document.has("imports") && !grammar_document.at("imports").as_array("imports").empty())
    throw std::invalid_argument("grammar imports are outside this engine");
  // The source as the resolver keeps it: id-sorted resources, no imports.
  Value source = grammar_document;
  source.obj.erase("imports");
  if (source.has("nonterminals")) {
    Value& nts = source.obj.at("nonterminals");
    sort_by_id(nts);
    for (auto& nt : nts.arr)
      if (nt.has("alternatives")) sort_by_id(nt.obj.at("alternatives"));
  }
  if (source.has("templates")) {
    Value& ts = source.obj.at("templates");
    sort_by_id(ts);
    for (auto& t : ts.arr)
      if (t.has("holes")) sort_by_id(t.obj.at("holes"));
  }
  Value out = json::make_object();
  out.obj["format_version"] = json::make_string("grammar-definition-v2");
  out.obj["catalog_version"] = json::make_string("gagp-primitives-v3");
  out.obj["normalization_version"] = json::make_string("1");
  out.obj["resolved"] = json::make_bool(true);
  for (const char* field : {"entry", "search_limits", "execution_limits"}) out.obj[field] = grammar_document.at(field);
  out.obj["inputs"] = grammar_document.has("inputs") ? grammar_document.at("inputs") : json::make_array();
  out.obj["locals"] = grammar_document.has("locals") ? grammar_document.at("locals") : json::make_array();
  for (const char* category : {"nonterminals", "templates"}) {
    Value resources = json::make_array();
    if (source.has(category))
      for (Value r : source.at(category).arr) {
        r.obj.erase("operation");
        resources.arr.push_back(std::move(r));
      }
    out.obj[category] = std::move(resources);
  }
  Value content = json::make_array();
  content.arr.push_back(source);
  out.obj["source_content"] = std::move(content);
  return sha256_hex(json::canonical(out));
}

}  // namespace gagp_amd
