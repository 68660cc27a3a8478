// Program import (see population.hpp).
#include "population.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <tuple>

#include "phase.hpp"

namespace gagp_amd {
namespace {
using json::Value;

[[noreturn]] void invalid(const std::string& m) { throw std::invalid_argument(m); }

int int_field(const Value& row, const char* key, const char* what) {
  if (!row.has(key)) invalid(std::string("AST ") + what + " missing field: " + key);
  const Value& v = row.at(key);
  if (v.kind != Value::Number || std::floor(v.num) != v.num || std::fabs(v.num) > 2147483647.0)
    invalid(std::string("AST ") + what + " field " + key + " must be an integer");
  return int(v.num);
}

// Typed constant JSON ({"type": "int", "value": 3}); the engine's values are
// Int and Bool (the writer is typed_value_json in export.cpp).
Val typed_constant(const Value& v) {
  if (v.kind != Value::Object || !v.has("type") || !v.has("value")) invalid("AST constant must be a typed value object");
  const std::string& type = v.at("type").as_string("AST constant type");
  const Value& x = v.at("value");
  if (type == "int") {
    if (x.kind != Value::Number || std::floor(x.num) != x.num) invalid("AST int constant must be an integer");
    // JSON numbers are read as doubles: beyond 2^53 the value is not exact.
    if (std::fabs(x.num) > 9007199254740992.0) throw Unsupported("AST int constant beyond 2^53");
    return val_int(std::int64_t(x.num));
  }
  if (type == "bool") {
    if (x.kind != Value::Bool) invalid("AST bool constant must be true or false");
    return val_bool(x.b);
  }
  throw Unsupported("AST constant of type " + type + " (the engine's values are int and bool)");
}

bool same_value(const Val& a, const Val& b) {
  if (a.tag != b.tag) return false;
  if (a.tag == TAG_INT) return a.i == b.i;
  if (a.tag == TAG_BOOL) return a.b == b.b;
  return false;
}

// First-match membership of one AST against the compiled grammar.
class Matcher {
 public:
  Matcher(const Grammar& g, const Ast& ast) : g_(g), ast_(ast) {}

  std::vector<int> run() {
    const std::size_t n = ast_.nodes.size();
    if (!n || n > g_.max_nodes) invalid("materialized node budget exceeded or empty AST");
    // Prefix shape and depth (membership.cpp:121-131).
    std::vector<int> pending{1};
    for (const AstNode& node : ast_.nodes) {
      while (!pending.empty() && pending.back() == 0) pending.pop_back();
      if (pending.empty()) invalid("trailing prefix nodes");
      if (pending.size() > g_.max_depth) invalid("materialized depth budget exceeded");
      --pending.back();
      const int arity = node_prefix_arity(node.kind);
      if (arity) pending.push_back(arity);
    }
    for (int c : pending)
      if (c) invalid("missing prefix children");
    end_.assign(n, 0);
    {
      std::vector<std::size_t> sizes;
      for (std::size_t j = n; j-- > 0;) {
        std::size_t size = 1;
        for (int c = 0; c < node_prefix_arity(ast_.nodes[j].kind); ++c) {
          size += sizes.back();
          sizes.pop_back();
        }
        sizes.push_back(size);
        end_[j] = j + size;
      }
    }
    const CNt& entry = g_.nts[g_.entry];
    std::size_t start = 0, stop = n;
    if (entry.cat == Cat::Expression) {
      const auto& x = ast_.nodes;
      if (n < 5 || x[0].kind != NK_PROGRAM || x[1].kind != NK_BLOCK_CONS || x[2].kind != NK_RETURN || x.back().kind != NK_BLOCK_NIL ||
          end_[3] != n - 1)
        invalid("expression entry requires its single-return envelope");
      start = 3;
      stop = n - 1;
    }
    env_.assign(entry.scope.size(), -1);
    if (!nonterminal(g_.entry, start)) invalid("AST cannot be derived from the grammar entry");
    // The witness: the successful decisions, walked again from the entry.
    origin_.assign(n, -1);
    env_.assign(entry.scope.size(), -1);
    derive_nt(g_.entry, start);
    std::vector<int> out(origin_.begin() + std::ptrdiff_t(start), origin_.begin() + std::ptrdiff_t(stop));
    for (int o : out)
      if (o < 0) invalid("AST node outside the entry derivation");
    return out;
  }

 private:
  using Key = std::tuple<std::uint32_t, std::size_t, std::vector<int>>;

  const std::string* name_at(int i) const {
    return i >= 0 && std::size_t(i) < ast_.names.size() ? &ast_.names[std::size_t(i)] : nullptr;
  }
  bool constant_in(std::uint32_t domain, const AstNode& node) const {
    if (node.i0 < 0 || std::size_t(node.i0) >= ast_.consts.size()) return false;
    const Val& v = ast_.consts[std::size_t(node.i0)];
    const CDomain& d = g_.domains.at(domain);
    if (d.integer_range) return v.tag == TAG_INT && v.i >= d.minimum && v.i <= d.maximum;
    for (const Val& w : d.values)
      if (same_value(v, w)) return true;
    return false;
  }
  bool reversed_at(std::size_t index, bool* reverse) const {
    for (const TraversalSpec& t : ast_.traversal_specs)
      if (std::size_t(t.node_index) == index) { *reverse = t.reverse; return true; }
    return false;
  }
  // The node-local contract of an expression (membership.cpp:448-488).
  bool shallow(const CExpr& e, std::size_t index) const {
    const AstNode& node = ast_.nodes[index];
    switch (e.kind) {
      case EKind::Constant: return node.kind == NK_CONST && constant_in(e.target, node);
      case EKind::Bound:
        return node.kind == NK_REGION_VAR && e.target < env_.size() && env_[e.target] >= 0 && node.i0 == env_[e.target];
      case EKind::Input: {
        const std::string* name = name_at(node.i0);
        return node.kind == NK_VAR && name && *name == g_.inputs.at(e.target).name;
      }
      case EKind::Local: {
        const std::string* name = name_at(node.i0);
        return node.kind == NK_VAR && name && *name == g_.locals.at(e.target).name;
      }
      case EKind::Primitive:
        if (node.kind != e.node_kind) return false;
        if (e.direction) {
          bool reverse = false;
          if (!reversed_at(index, &reverse) || reverse != (e.direction == 2)) return false;
        }
        break;
      case EKind::Control:
        if (node.kind != e.node_kind) return false;
        if (e.local >= 0 || e.input >= 0) {
          const std::string* name = name_at(node.i0);
          const std::string& want = e.local >= 0 ? g_.locals.at(std::size_t(e.local)).name : g_.inputs.at(std::size_t(e.input)).name;
          if (!name || *name != want) return false;
        }
        break;
      default: return false;
    }
    return node_prefix_arity(node.kind) == int(e.children.size());
  }
  // A let or traversal body sees the binders its region declares
  // (membership.cpp:89-107).
  void extend(const CExpr& e, std::size_t index, std::size_t argument) {
    if (e.region_argument < 0 || std::size_t(e.region_argument) != argument) return;
    for (const LexicalRegion& r : ast_.lexical_regions)
      if (std::size_t(r.node_index) == index) {
        for (const LexicalBinding& b : r.bindings) env_.push_back(b.id);
        return;
      }
    invalid("AST is missing the lexical region of node " + std::to_string(index));
  }
  bool nonterminal(std::uint32_t nt, std::size_t index) {
    Key key{nt, index, env_};
    const auto found = decisions_.find(key);
    if (found != decisions_.end()) return found->second >= 0;  // -1: the same match in progress
    const auto it = decisions_.emplace(std::move(key), -1).first;
    for (auto p : g_.nts[nt].productions)
      if (expression(g_.prods[p].expression, index)) {
        it->second = int(p);
        return true;
      }
    decisions_.erase(it);
    return false;
  }
  bool expression(std::uint32_t id, std::size_t index) {
    if (++steps_ > (1u << 22)) invalid("grammar membership exceeds 4194304 steps");
    const CExpr& e = g_.exprs[id];
    if (e.kind == EKind::Reference) {
      std::vector<int> saved = std::move(env_);
      env_.clear();
      for (auto m : e.scope_mapping) env_.push_back(saved.at(m));
      const bool ok = nonterminal(e.target, index);
      env_ = std::move(saved);
      return ok;
    }
    if (!shallow(e, index)) return false;
    std::size_t child = index + 1;
    for (std::size_t a = 0; a < e.children.size(); ++a) {
      const std::vector<int> saved = env_;
      extend(e, index, a);
      const bool ok = expression(e.children[a], child);
      env_ = saved;
      if (!ok) return false;
      child = end_[child];
    }
    return child == end_[index];
  }
  void derive_nt(std::uint32_t nt, std::size_t index) {
    const int p = decisions_.at(Key{nt, index, env_});
    derive_expr(g_.prods.at(std::size_t(p)).expression, index);
  }
  void derive_expr(std::uint32_t id, std::size_t index) {
    const CExpr& e = g_.exprs[id];
    if (e.kind == EKind::Reference) {
      std::vector<int> saved = std::move(env_);
      env_.clear();
      for (auto m : e.scope_mapping) env_.push_back(saved.at(m));
      derive_nt(e.target, index);
      env_ = std::move(saved);
      return;
    }
    origin_[index] = int(id);
    std::size_t child = index + 1;
    for (std::size_t a = 0; a < e.children.size(); ++a) {
      const std::vector<int> saved = env_;
      extend(e, index, a);
      derive_expr(e.children[a], child);
      env_ = saved;
      child = end_[child];
    }
  }

  const Grammar& g_;
  const Ast& ast_;
  std::vector<std::size_t> end_;
  std::map<Key, int> decisions_;
  std::vector<int> env_;
  std::vector<int> origin_;
  std::size_t steps_ = 0;
};

// Sites from origins: the reference expression whose nonterminal choice
// starts each node, packed with its parent scope (phase_variation.cuh:23-77).
bool annotate(const PhaseGrammar& p, const std::vector<int>& origin, std::vector<int>& site) {
  struct Task { int e, sc, site; };
  const int n = int(origin.size());
  if (n <= 0 || n > kPhaseCapacity || p.root < 0) return false;
  site.assign(origin.size(), -1);
  std::vector<Task> pending{{p.root, 0, -1}};
  int cursor = 0, steps = 0;
  while (!pending.empty()) {
    if (++steps > 8192 || cursor >= n) return false;
    const Task t = pending.back();
    pending.pop_back();
    if (t.e < 0 || t.e >= int(p.exprs.size())) return false;
    const PExpr& x = p.exprs[std::size_t(t.e)];
    if (x.kind == PK_REFERENCE) {
      if (t.site >= 0 || x.target < 0 || x.target >= int(p.nts.size()) || x.scope_count < 0 || x.scope_count > 8 ||
          int(pending.size()) >= kPending)
        return false;
      const PNt& nt = p.nts[std::size_t(x.target)];
      if (nt.scope_count != x.scope_count || nt.production_count <= 0) return false;
      int production = -1;
      for (int i = 0; i < nt.production_count && production < 0; ++i)
        if (p.prods[std::size_t(nt.production_offset + i)].expression == origin[std::size_t(cursor)])
          production = p.prods[std::size_t(nt.production_offset + i)].expression;
      if (production < 0) return false;
      pending.push_back({production, x.scope_count, site_pack(t.e, t.sc)});
      continue;
    }
    if (x.kind == PK_ALIAS || origin[std::size_t(cursor)] != t.e || x.node.arity < 0 || x.node.arity > kMaxArity) return false;
    site[std::size_t(cursor++)] = t.site;
    for (int i = x.node.arity - 1; i >= 0; --i) {
      if (int(pending.size()) >= kPending) return false;
      pending.push_back({x.children[i], child_scope(x, i, t.sc), -1});
    }
  }
  return cursor == n;
}

// The most sites on one root-to-leaf path (phase_variation.cuh:127-142).
int site_height(const PhaseGrammar& p, const std::vector<int>& origin, const std::vector<int>& site) {
  int pending[64], incoming[64];
  int depth = 1, maximum = 0;
  pending[0] = 1;
  incoming[0] = 0;
  for (std::size_t i = 0; i < origin.size(); ++i) {
    while (depth && !pending[depth - 1]) --depth;
    if (!depth) return -1;
    const int arity = p.exprs.at(std::size_t(origin[i])).node.arity;
    const int value = incoming[depth - 1] + (site[i] >= 0 ? 1 : 0);
    maximum = std::max(maximum, value);
    --pending[depth - 1];
    if (arity) {
      if (depth == 64) return -1;
      pending[depth] = arity;
      incoming[depth++] = value;
    }
  }
  while (depth && !pending[depth - 1]) --depth;
  return depth ? -1 : maximum;
}

std::string read_text(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) invalid("cannot read " + path);
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}
}  // namespace

Ast decode_ast(const Value& doc) {
  if (doc.kind != Value::Object) invalid("AST JSON must be an object");
  if (!doc.has("version") || doc.at("version").kind != Value::String) invalid("AST JSON missing string field: version");
  if (doc.at("version").str != "ast-prefix-v2") invalid("AST JSON version must be ast-prefix-v2");
  for (const char* legacy : {"linear_rec_binders", "asgp_dc_binders", "asgp_dp1d_specs", "asgp_dp2d_specs"})
    if (doc.has(legacy)) invalid(std::string("AST legacy sidecar field is not valid in ast-prefix-v2: ") + legacy);
  if (doc.has("fuel_specs") && !doc.at("fuel_specs").as_array("AST fuel_specs").empty())
    throw Unsupported("AST fuel_specs (per-node fuel charges)");
  if (doc.has("bounded_region_specs") && !doc.at("bounded_region_specs").as_array("AST bounded_region_specs").empty())
    throw Unsupported("AST bounded regions");
  Ast ast;
  if (!doc.has("nodes") || doc.at("nodes").kind != Value::Array) invalid("AST JSON missing array field: nodes");
  for (const Value& row : doc.at("nodes").arr) {
    if (row.kind != Value::Object) invalid("AST node must be an object");
    const int kind = int_field(row, "kind", "node");
    if (!node_known(kind)) invalid("AST node kind " + std::to_string(kind) + " is outside the engine's node set");
    ast.nodes.push_back({kind, int_field(row, "i0", "node"), int_field(row, "i1", "node")});
  }
  if (!doc.has("names") || doc.at("names").kind != Value::Array) invalid("AST JSON missing array field: names");
  for (const Value& v : doc.at("names").arr) ast.names.push_back(v.as_string("AST names"));
  if (!doc.has("consts") || doc.at("consts").kind != Value::Array) invalid("AST JSON missing array field: consts");
  for (const Value& v : doc.at("consts").arr) ast.consts.push_back(typed_constant(v));
  if (doc.has("lexical_regions")) {
    for (const Value& row : doc.at("lexical_regions").as_array("AST lexical_regions")) {
      if (row.kind != Value::Object) invalid("AST lexical_regions item must be an object");
      LexicalRegion r;
      r.node_index = int_field(row, "node_index", "lexical_regions");
      r.body_argument = int_field(row, "body_argument", "lexical_regions");
      if (!row.has("bindings")) invalid("AST lexical_regions.bindings must be an array");
      for (const Value& b : row.at("bindings").as_array("AST lexical_regions.bindings")) {
        if (b.kind != Value::Object || !b.has("type")) invalid("AST lexical_regions.bindings missing field: type");
        const int id = int_field(b, "id", "lexical_regions.bindings");
        if (id < 0) invalid("AST lexical binder ids must be nonnegative");
        r.bindings.push_back({id, parse_type_name(b.at("type").as_string("AST lexical binding type"))});
      }
      if (r.node_index < 0 || std::size_t(r.node_index) >= ast.nodes.size()) invalid("AST lexical region outside the program");
      ast.lexical_regions.push_back(std::move(r));
    }
  }
  if (doc.has("traversal_specs")) {
    for (const Value& row : doc.at("traversal_specs").as_array("AST traversal_specs")) {
      if (row.kind != Value::Object) invalid("AST traversal_specs item must be an object");
      const int at = int_field(row, "node_index", "traversal_specs");
      if (!row.has("direction") || row.at("direction").kind != Value::String)
        invalid("AST traversal_specs.direction must be a string");
      const std::string& d = row.at("direction").str;
      if (d != "forward" && d != "reverse") invalid("AST traversal_specs.direction must be forward or reverse");
      if (at < 0 || std::size_t(at) >= ast.nodes.size()) invalid("AST traversal spec outside the program");
      ast.traversal_specs.push_back({at, d == "reverse"});
    }
  }
  // Table indexes of name and constant operands.
  for (const AstNode& n : ast.nodes) {
    if (node_i0_is_name(n.kind) && (n.i0 < 0 || std::size_t(n.i0) >= ast.names.size())) invalid("AST name index out of range");
    if (node_i0_is_constant(n.kind) && (n.i0 < 0 || std::size_t(n.i0) >= ast.consts.size()))
      invalid("AST constant index out of range");
  }
  return ast;
}

const Value& program_of_document(const Value& doc) {
  if (doc.has("best")) return doc.at("best").at("program");
  if (doc.has("programs")) {
    const auto& programs = doc.at("programs").as_array("programs");
    if (programs.empty()) invalid("population document has no programs");
    return programs[0];
  }
  return doc;
}

std::vector<int> derive_origins(const Grammar& g, const Ast& ast) { return Matcher(g, ast).run(); }

PhaseRow import_row(const Grammar& g, const PhaseGrammar& phase, const Ast& ast, unsigned site_depth, int capacity) {
  PhaseRow row;
  row.origin = derive_origins(g, ast);
  // The reference's resident layout declines these two (its native
  // representation then runs them, phase_population_cpu.cpp:229-235), so
  // they are declines here too, not invalid input.
  if (int(row.origin.size()) > capacity)
    throw Unsupported("imported program has " + std::to_string(row.origin.size()) +
                      " phase nodes, above the row capacity " + std::to_string(capacity));
  const std::size_t start = g.nts[g.entry].cat == Cat::Expression ? 3 : 0;
  for (std::size_t i = 0; i < row.origin.size(); ++i) {
    const AstNode& n = ast.nodes[start + i];
    if (n.kind == NK_CONST) row.cval.push_back(ast.consts.at(std::size_t(n.i0)));
  }
  if (!annotate(phase, row.origin, row.site) ||
      (site_depth && phase.site_depth && site_height(phase, row.origin, row.site) > int(site_depth)))
    throw Unsupported("import site annotation failed");
  // The native verifier rejects a statement program that reads a local
  // before every path assigns it (undefined_local).
  if (g.nts[g.entry].cat == Cat::Program) {
    const PhaseView gv{phase.exprs.data(), int(phase.exprs.size()), phase.prods.data(), int(phase.prods.size()),
                       phase.nts.data(), int(phase.nts.size()), phase.domains.data(), int(phase.domains.size()),
                       phase.values.data(), int(phase.values.size()), phase.minimum.data(), phase.depth_count,
                       phase.logical_count, phase.site_depth ? 1 : 0, phase.root};
    if (!locals_defined(gv, CRow{row.origin.data(), row.site.data(), row.cval.data(), int(row.origin.size()), int(row.cval.size())}))
      invalid("invalid AST (undefined_local): a grammar local is read before it is assigned");
  }
  return row;
}

ImportedPopulation load_population(const std::string& path, const Grammar& g, const PhaseGrammar& phase, unsigned site_depth,
                                   int capacity, std::size_t count) {
  const Value doc = json::parse(read_text(path));
  if (doc.has("format_version") && doc.at("format_version").kind == Value::String &&
      doc.at("format_version").str == "grammar-population-v2")
    throw Unsupported(path + ": grammar-population-v2 documents (replays of `gagp grammar generate`)");
  if (doc.kind != Value::Object || !doc.has("programs"))
    invalid(path + ": expected a population document with a programs array");
  if (doc.has("grammar_hash")) {
    const Value& h = doc.at("grammar_hash");
    if (h.kind != Value::String || h.str != g.content_hash) invalid(path + ": population was made for a different grammar");
  }
  const auto& programs = doc.at("programs").as_array("programs");
  ImportedPopulation out;
  for (std::size_t i = 0; i < std::min(count, programs.size()); ++i) {
    const Ast ast = decode_ast(programs[i]);
    out.rows.push_back(import_row(g, phase, ast, site_depth, capacity));
    if (i == 0) {
      out.layout_names = ast.names;
      for (const LexicalRegion& r : ast.lexical_regions)
        for (const LexicalBinding& b : r.bindings) out.binder_base = std::max(out.binder_base, b.id + 1);
    }
  }
  if (out.rows.size() < count)
    invalid(path + ": has " + std::to_string(out.rows.size()) + " programs, the run needs " + std::to_string(count));
  return out;
}

}  // namespace gagp_amd
