// Grammar front end (host). See grammar.hpp for the reference behaviour
// each step follows.
#include "grammar.hpp"
#include "identity.hpp"

#include <algorithm>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <sstream>

namespace gagp_amd {
namespace {
using json::Value;

struct NodeInfo { int kind; const char* name; int arity; int i0; };  // i0: 0 unused, 1 name, 2 constant
// Serialized names and prefix arities of the node kinds this engine emits
// (cpp/src/evolution/node_descriptor.cpp:24-85).
const NodeInfo kNodes[] = {
    {NK_PROGRAM, "program", 1, 0}, {NK_BLOCK_NIL, "block_nil", 0, 0}, {NK_BLOCK_CONS, "block_cons", 2, 0},
    {NK_ASSIGN, "assign", 1, 1}, {NK_IF_STMT, "if_stmt", 3, 0}, {NK_FOR_RANGE, "for_range", 2, 1},
    {NK_RETURN, "return", 1, 0}, {NK_CONST, "const", 0, 2}, {NK_VAR, "var", 0, 1}, {NK_NEG, "neg", 1, 0},
    {NK_NOT, "not", 1, 0}, {NK_ADD, "add", 2, 0}, {NK_SUB, "sub", 2, 0}, {NK_MUL, "mul", 2, 0},
    {NK_DIV, "div", 2, 0}, {NK_MOD, "mod", 2, 0}, {NK_LT, "lt", 2, 0}, {NK_LE, "le", 2, 0},
    {NK_GT, "gt", 2, 0}, {NK_GE, "ge", 2, 0}, {NK_EQ, "eq", 2, 0}, {NK_NE, "ne", 2, 0},
    {NK_AND, "and", 2, 0}, {NK_OR, "or", 2, 0}, {NK_IF_EXPR, "if_expr", 3, 0},
    {NK_CALL_ABS, "call_abs", 1, 0}, {NK_CALL_MIN, "call_min", 2, 0}, {NK_CALL_MAX, "call_max", 2, 0},
    {NK_CALL_CLIP, "call_clip", 3, 0}, {NK_CALL_IDIV0, "call_idiv0", 2, 0}, {NK_CALL_IMOD0, "call_imod0", 2, 0},
    {NK_CALL_LEN, "call_len", 1, 0}, {NK_CALL_INDEX, "call_index", 2, 0}, {NK_BOUND_VAR, "bound_var", 0, 1},
    {NK_LET_REGION, "let_region", 2, 0}, {NK_TRAVERSE, "traverse", 4, 0}, {NK_TRAVERSE_RANGE, "traverse_range", 6, 0},
    {NK_REGION_VAR, "region_var", 0, 0}, {NK_CHECK_INT, "check_int", 1, 0}, {NK_CHECK_LIST, "check_list", 1, 0},
    {NK_CALL_TRUNC_DIV0, "call_trunc_div0", 2, 0}, {NK_CALL_TRUNC_MOD0, "call_trunc_mod0", 2, 0},
    {NK_CALL_INDEX_NONNEG, "call_index_nonnegative", 2, 0}, {NK_CALL_INDEX_CLAMPED, "call_index_clamped_or_zero", 2, 0},
};
const NodeInfo& node_info(int kind) {
  for (const auto& n : kNodes) if (n.kind == kind) return n;
  throw std::logic_error("unknown node kind " + std::to_string(kind));
}

void keys(const Value& v, std::initializer_list<const char*> allowed, const char* what) {
  if (v.kind != Value::Object) throw std::invalid_argument(std::string(what) + " must be an object");
  const std::set<std::string> names(allowed.begin(), allowed.end());
  for (const auto& kv : v.obj)
    if (!names.count(kv.first)) throw std::invalid_argument(std::string(what) + ": unknown key " + kv.first);
}
std::uint32_t positive(const Value& v, std::uint32_t maximum, const char* what) {
  if (v.kind != Value::Number || v.num < 1 || v.num > maximum || std::floor(v.num) != v.num)
    throw std::invalid_argument(std::string(what) + " must be an integer in 1.." + std::to_string(maximum));
  return std::uint32_t(v.num);
}
std::string identifier(const Value& v) {
  const std::string id = v.at("id").as_string("id");
  const auto letter = [](char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; };
  if (id.empty() || !letter(id[0]) || !std::all_of(id.begin(), id.end(), [&](char c) {
        return letter(c) || (c >= '0' && c <= '9') || c == '.' || c == '-'; }))
    throw std::invalid_argument("invalid grammar identifier: " + id);
  return id;
}
std::vector<Binding> scope_of(const Value& v) {
  std::vector<Binding> out;
  if (v.kind == Value::Null) return out;
  for (const auto& item : v.as_array("scope")) {
    keys(item, {"name", "type"}, "binding");
    Binding b{item.at("name").as_string("name"), parse_type_name(item.at("type").as_string("type"))};
    for (const auto& o : out) if (o.name == b.name) throw std::invalid_argument("duplicate binding: " + b.name);
    out.push_back(b);
  }
  return out;
}
Cat category_of(const Value& v) {
  if (!v.has("category")) return Cat::Expression;
  const auto& c = v.at("category").as_string("category");
  if (c == "Expression") return Cat::Expression;
  if (c == "Program") return Cat::Program;
  if (c == "Block") return Cat::Block;
  if (c == "Statement") return Cat::Statement;
  throw std::invalid_argument("unknown category " + c);
}
std::int64_t int_literal(const Value& v) {
  const auto& text = v.as_string("Int constant (decimal string)");
  if (text.empty()) throw std::invalid_argument("Int constant must be a canonical decimal string");
  errno = 0;
  char* end = nullptr;
  const long long r = std::strtoll(text.c_str(), &end, 10);
  if (errno || end != text.c_str() + text.size() || std::to_string(r) != text)
    throw std::invalid_argument("Int constant must be canonical signed 64-bit decimal string");
  return r;
}

struct Sig { std::string op; std::vector<std::uint8_t> args; std::uint8_t result; };
Sig parse_signature(const std::string& key) {
  Sig s;
  const auto open = key.find('('), close = key.find(")->");
  if (open == std::string::npos || close == std::string::npos || close < open)
    throw std::invalid_argument("unknown exact primitive overload: " + key);
  s.op = key.substr(0, open);
  std::string list = key.substr(open + 1, close - open - 1);
  std::stringstream ss(list);
  std::string item;
  while (!list.empty() && std::getline(ss, item, ',')) s.args.push_back(parse_type_name(item));
  s.result = parse_type_name(key.substr(close + 3));
  return s;
}

// The admitted slice of the primitive catalog (cpp/src/evolution/grammar/catalog.cpp:72-180):
// Int/Bool scalar operations, list length and indexing.
void resolve_primitive(const Sig& s, CExpr& e) {
  const auto I = TAG_INT, B = TAG_BOOL, L = TAG_INT_LIST;
  const auto& a = s.args;
  const auto is = [&](std::initializer_list<std::uint8_t> want, std::uint8_t result) {
    return std::vector<std::uint8_t>(want) == a && s.result == result;
  };
  struct Bin { const char* op; int nk; int builtin; };
  static const Bin ints[] = {{"add", NK_ADD, -1}, {"sub", NK_SUB, -1}, {"mul", NK_MUL, -1}, {"mod", NK_MOD, -1},
      {"min", NK_CALL_MIN, B_MIN}, {"max", NK_CALL_MAX, B_MAX}, {"idiv0", NK_CALL_IDIV0, B_IDIV0},
      {"imod0", NK_CALL_IMOD0, B_IMOD0}, {"trunc_div0", NK_CALL_TRUNC_DIV0, B_TRUNC_DIV0},
      {"trunc_mod0", NK_CALL_TRUNC_MOD0, B_TRUNC_MOD0}};
  for (const auto& b : ints)
    if (s.op == b.op && is({I, I}, I)) { e.node_kind = b.nk; e.builtin = b.builtin; return; }
  static const Bin cmps[] = {{"lt", NK_LT, -1}, {"le", NK_LE, -1}, {"gt", NK_GT, -1}, {"ge", NK_GE, -1}};
  for (const auto& b : cmps)
    if (s.op == b.op && is({I, I}, B)) { e.node_kind = b.nk; return; }
  if ((s.op == "eq" || s.op == "ne") && a.size() == 2 && a[0] == a[1] && (a[0] == I || a[0] == B) && s.result == B) {
    e.node_kind = s.op == "eq" ? NK_EQ : NK_NE; return;
  }
  if (s.op == "neg" && is({I}, I)) { e.node_kind = NK_NEG; return; }
  if (s.op == "abs" && is({I}, I)) { e.node_kind = NK_CALL_ABS; e.builtin = B_ABS; return; }
  if (s.op == "clip" && is({I, I, I}, I)) { e.node_kind = NK_CALL_CLIP; e.builtin = B_CLIP; return; }
  if (s.op == "check_int" && is({I}, I)) { e.node_kind = NK_CHECK_INT; return; }
  if (s.op == "not" && is({B}, B)) { e.node_kind = NK_NOT; return; }
  if (s.op == "and" && is({B, B}, B)) { e.node_kind = NK_AND; return; }
  if (s.op == "or" && is({B, B}, B)) { e.node_kind = NK_OR; return; }
  if (s.op == "if" && a.size() == 3 && a[0] == B && a[1] == a[2] && a[1] == s.result && (s.result == I || s.result == B)) {
    e.node_kind = NK_IF_EXPR; return;
  }
  if (s.op == "len" && is({L}, I)) { e.node_kind = NK_CALL_LEN; e.builtin = B_LEN; return; }
  if (s.op == "index" && is({L, I}, I)) { e.node_kind = NK_CALL_INDEX; e.builtin = B_INDEX; return; }
  if (s.op == "index_nonnegative" && is({L, I}, I)) { e.node_kind = NK_CALL_INDEX_NONNEG; e.builtin = B_INDEX_NONNEG; return; }
  if (s.op == "index_clamped_or_zero" && is({L, I}, I)) { e.node_kind = NK_CALL_INDEX_CLAMPED; e.builtin = B_INDEX_CLAMPED; return; }
  throw Unsupported("primitive outside the engine's admitted catalog: " + s.op);
}

// A lexical binding name: an identifier not already visible
// (cpp/src/evolution/grammar/compiled.cpp:118-125).
void add_binding(std::vector<Binding>& scope, const std::string& name, std::uint8_t type) {
  const auto letter = [](char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; };
  if (name.empty() || !letter(name.front()) ||
      !std::all_of(name.begin(), name.end(), [&](char c) { return letter(c) || (c >= '0' && c <= '9'); }))
    throw std::invalid_argument("invalid lexical binding name");
  for (const auto& b : scope)
    if (b.name == name) throw std::invalid_argument("duplicate/shadowed lexical binding: " + name);
  scope.push_back({name, type});
}

// Traversals of the catalog (catalog.cpp:134-150) over an IntList source with
// an Int or Bool accumulator: traverse[_reverse](xs, start, init, body) and
// traverse_range[_reverse](xs, start, begin, end, init, body). The body
// argument binds element (Int), index (Int) and accumulator.
bool resolve_traversal(const Sig& s, CExpr& e) {
  const bool ranged = s.op == "traverse_range" || s.op == "traverse_range_reverse";
  if (!ranged && s.op != "traverse" && s.op != "traverse_reverse") return false;
  const auto& a = s.args;
  const std::size_t arity = ranged ? 6 : 4;
  if (a.size() != arity || a[arity - 1] != s.result || a[arity - 2] != s.result)
    throw std::invalid_argument("unknown exact primitive overload: " + s.op);
  for (std::size_t i = 1; i + 2 < arity; ++i)
    if (a[i] != TAG_INT) throw std::invalid_argument("unknown exact primitive overload: " + s.op);
  if (a[0] != TAG_INT_LIST) throw Unsupported("traversal over a non-IntList sequence");
  if (s.result != TAG_INT && s.result != TAG_BOOL) throw Unsupported("traversal accumulator outside Int/Bool");
  e.node_kind = ranged ? NK_TRAVERSE_RANGE : NK_TRAVERSE;
  e.direction = s.op == "traverse" || s.op == "traverse_range" ? 1 : 2;
  e.region_argument = int(arity - 1);
  e.region_bindings = {{"element", TAG_INT}, {"index", TAG_INT}, {"accumulator", s.result}};
  return true;
}

// let(T,U)->U (catalog.cpp:91-92): the body argument binds the value.
bool resolve_let(const Sig& s, CExpr& e) {
  if (s.op != "let") return false;
  if (s.args.size() != 2 || s.args[1] != s.result) throw std::invalid_argument("unknown exact primitive overload: let");
  const auto value = [](std::uint8_t t) { return t == TAG_INT || t == TAG_BOOL || t == TAG_INT_LIST; };
  if (!value(s.args[0]) || (s.result != TAG_INT && s.result != TAG_BOOL)) throw Unsupported("let outside Int/Bool values");
  e.node_kind = NK_LET_REGION;
  e.region_argument = 1;
  e.region_bindings = {{"value", s.args[0]}};
  return true;
}

class Compiler {
 public:
  Grammar compile(const Value& doc) {
    keys(doc, {"format_version", "entry", "inputs", "locals", "search_limits", "execution_limits", "nonterminals",
               "templates", "imports"}, "grammar definition");
    if (doc.at("format_version").as_string("format_version") != "grammar-definition-v2")
      throw std::invalid_argument("grammar format_version must be grammar-definition-v2");
    if (doc.has("imports") && !doc.at("imports").as_array("imports").empty()) throw Unsupported("grammar imports");
    if (doc.has("templates") && !doc.at("templates").as_array("templates").empty()) throw Unsupported("grammar templates");
    const auto& search = doc.at("search_limits");
    keys(search, {"max_nodes", "max_depth"}, "search_limits");
    g_.max_nodes = positive(search.at("max_nodes"), 65536, "max_nodes");
    g_.max_depth = positive(search.at("max_depth"), 256, "max_depth");
    const auto& exec = doc.at("execution_limits");
    keys(exec, {"fuel"}, "execution_limits");
    g_.fuel = positive(exec.at("fuel"), 0x7fffffff, "fuel");
    g_.inputs = scope_of(doc.has("inputs") ? doc.at("inputs") : Value{});
    g_.locals = scope_of(doc.has("locals") ? doc.at("locals") : Value{});
    if (g_.locals.size() > 16) throw Unsupported("more than 16 grammar locals");
    for (const auto& l : g_.locals)
      for (const auto& in : g_.inputs)
        if (l.name == in.name) throw std::invalid_argument("local declaration conflicts with input name");
    // Resolution order: nonterminals and their alternatives sorted by id.
    std::vector<const Value*> rules;
    for (const auto& r : doc.at("nonterminals").as_array("nonterminals")) rules.push_back(&r);
    if (rules.empty() || rules.size() > 4096) throw std::invalid_argument("grammar requires 1..4096 nonterminals");
    std::stable_sort(rules.begin(), rules.end(), [](const Value* a, const Value* b) { return identifier(*a) < identifier(*b); });
    for (const Value* r : rules) {
      keys(*r, {"id", "type", "scope", "alternatives", "category", "variation"}, "nonterminal");
      CNt nt;
      nt.id = identifier(*r);
      if (ids_.count(nt.id)) throw std::invalid_argument("duplicate nonterminal id " + nt.id);
      nt.type = parse_type_name(r->at("type").as_string("type"));
      nt.cat = category_of(*r);
      if (r->has("variation")) {
        if (r->at("variation").kind != Value::Bool) throw std::invalid_argument("nonterminal variation must be Boolean");
        nt.variation = r->at("variation").b;
      }
      nt.scope = scope_of(r->at("scope"));
      ids_.emplace(nt.id, std::uint32_t(g_.nts.size()));
      g_.nts.push_back(std::move(nt));
    }
    const auto& entry = doc.at("entry");
    keys(entry, {"nonterminal", "type", "category"}, "entry");
    g_.entry = reference(entry.at("nonterminal").as_string("entry nonterminal"));
    if (g_.nts[g_.entry].type != parse_type_name(entry.at("type").as_string("entry type")))
      throw std::invalid_argument("entry result type mismatch");
    if (category_of(entry) != g_.nts[g_.entry].cat ||
        (g_.nts[g_.entry].cat != Cat::Expression && g_.nts[g_.entry].cat != Cat::Program))
      throw std::invalid_argument("entry requires matching Expression or Program category");
    if (!g_.nts[g_.entry].scope.empty()) throw std::invalid_argument("entry nonterminal requires unavailable lexical bindings");
    for (std::size_t i = 0; i < rules.size(); ++i) {
      std::vector<const Value*> alts;
      for (const auto& a : rules[i]->at("alternatives").as_array("alternatives")) alts.push_back(&a);
      if (alts.empty()) throw std::invalid_argument("nonterminal has no alternatives");
      std::stable_sort(alts.begin(), alts.end(), [](const Value* a, const Value* b) { return identifier(*a) < identifier(*b); });
      std::set<std::string> seen;
      for (const Value* a : alts) {
        keys(*a, {"id", "weight", "expression"}, "alternative");
        const std::string aid = identifier(*a);
        if (!seen.insert(aid).second) throw std::invalid_argument("duplicate alternative id " + aid);
        const auto& w = a->at("weight");
        if (w.kind != Value::Number || !std::isfinite(w.num) || w.num <= 0)
          throw std::invalid_argument("production weight must be positive and finite");
        CProd p;
        p.nt = std::uint32_t(i);
        p.weight = w.num;
        p.stable_id = g_.nts[i].id + "/" + aid;
        p.expression = expression(a->at("expression"), g_.nts[i].scope, 0);
        const auto& e = g_.exprs[p.expression];
        if (e.type != g_.nts[i].type || e.cat != g_.nts[i].cat)
          throw std::invalid_argument("production result type mismatch: " + p.stable_id);
        g_.nts[i].productions.push_back(std::uint32_t(g_.prods.size()));
        g_.prods.push_back(std::move(p));
      }
    }
    costs();
    return std::move(g_);
  }

 private:
  std::uint32_t reference(const std::string& id) const {
    const auto it = ids_.find(id);
    if (it == ids_.end()) throw std::invalid_argument("unknown nonterminal: " + id);
    return it->second;
  }
  static std::uint32_t binding(const std::vector<Binding>& scope, const std::string& name) {
    for (std::size_t i = 0; i < scope.size(); ++i) if (scope[i].name == name) return std::uint32_t(i);
    throw std::invalid_argument("unknown binding: " + name);
  }
  std::uint32_t append(CExpr e) {
    if (g_.exprs.size() >= 65536) throw std::invalid_argument("grammar expression capacity exceeded");
    g_.exprs.push_back(std::move(e));
    return std::uint32_t(g_.exprs.size() - 1);
  }
  CDomain domain(const Value& v) {
    keys(v, {"type", "values", "range"}, "constant domain");
    CDomain d;
    d.type = parse_type_name(v.at("type").as_string("constant type"));
    if (v.has("values") == v.has("range")) throw std::invalid_argument("constant domain needs exactly one of values/range");
    if (v.has("range")) {
      if (d.type != TAG_INT) throw Unsupported("non-Int constant range");
      const auto& r = v.at("range").as_array("range");
      if (r.size() != 2) throw std::invalid_argument("range needs two endpoints");
      d.integer_range = true;
      d.minimum = int_literal(r[0]);
      d.maximum = int_literal(r[1]);
      if (d.minimum > d.maximum) throw std::invalid_argument("range endpoints must be ascending");
      return d;
    }
    const auto& vals = v.at("values").as_array("values");
    if (vals.empty()) throw std::invalid_argument("constant values must be nonempty");
    for (const auto& x : vals) {
      if (d.type == TAG_INT) d.values.push_back(val_int(int_literal(x)));
      else if (d.type == TAG_BOOL) {
        if (x.kind != Value::Bool) throw std::invalid_argument("Bool constant must be boolean");
        d.values.push_back(val_bool(x.b));
      } else throw Unsupported(std::string("constant domain of type ") + type_name(d.type));
    }
    return d;
  }
  std::uint32_t expression(const Value& v, const std::vector<Binding>& env, unsigned depth) {
    if (depth >= 256) throw std::invalid_argument("grammar expression capacity exceeded");
    if (v.kind != Value::Object) throw std::invalid_argument("expression must be an object");
    CExpr e;
    if (v.has("constant")) {
      keys(v, {"constant"}, "constant expression");
      e.kind = EKind::Constant;
      e.target = std::uint32_t(g_.domains.size());
      g_.domains.push_back(domain(v.at("constant")));
      e.type = g_.domains.back().type;
      e.node_kind = NK_CONST;
    } else if (v.has("ref")) {
      keys(v, {"ref"}, "nonterminal reference");
      e.kind = EKind::Reference;
      e.target = reference(v.at("ref").as_string("ref"));
      const auto& t = g_.nts[e.target];
      e.type = t.type;
      e.cat = t.cat;
      for (const auto& req : t.scope) {
        const auto idx = binding(env, req.name);
        if (env[idx].type != req.type) throw std::invalid_argument("reference scope type mismatch");
        e.scope_mapping.push_back(idx);
      }
    } else if (v.has("input") || v.has("bound")) {
      const bool input = v.has("input");
      keys(v, {input ? "input" : "bound"}, "binding expression");
      const auto& scope = input ? g_.inputs : env;
      e.kind = input ? EKind::Input : EKind::Bound;
      e.target = binding(scope, v.at(input ? "input" : "bound").as_string("binding name"));
      e.type = scope[e.target].type;
      e.node_kind = input ? NK_VAR : NK_REGION_VAR;
    } else if (v.has("signature")) {
      keys(v, {"signature", "args", "bind", "fuel_events", "resource_charge"}, "primitive expression");
      if (v.has("fuel_events")) throw Unsupported("source fuel profiles (fuel_events)");
      if (v.has("resource_charge")) throw Unsupported("resource charges");
      const Sig s = parse_signature(v.at("signature").as_string("signature"));
      e.kind = EKind::Primitive;
      if (!resolve_traversal(s, e) && !resolve_let(s, e)) resolve_primitive(s, e);
      e.type = s.result;
      const auto& args = v.at("args").as_array("args");
      if (args.size() != s.args.size()) throw std::invalid_argument("primitive argument count mismatch");
      // Region arguments see the enclosing scope plus the names bound there
      // (compiled.cpp:664-690: no shadowing, one name per region binding).
      std::vector<Binding> body = env;
      if (v.has("bind")) {
        const auto& bind = v.at("bind");
        if (bind.kind != Value::Object) throw std::invalid_argument("bind must be an object keyed by region argument");
        for (const auto& item : bind.obj) {
          if (e.region_argument < 0 || std::to_string(e.region_argument) != item.first)
            throw std::invalid_argument("binding names target a non-region argument");
          const auto& names = item.second.as_array("region binding names");
          if (names.size() != e.region_bindings.size()) throw std::invalid_argument("region binding arity mismatch");
          for (std::size_t k = 0; k < names.size(); ++k) {
            const std::string name = names[k].as_string("region binding name");
            add_binding(body, name, e.region_bindings[k].type);
            e.region_bindings[k].name = name;
          }
        }
      }
      if (e.region_argument >= 0 && body.size() == env.size())
        throw std::invalid_argument("primitive requires explicit region binding names");
      for (std::size_t i = 0; i < args.size(); ++i) {
        const auto child = expression(args[i], int(i) == e.region_argument ? body : env, depth + 1);
        if (g_.exprs[child].type != s.args[i] || g_.exprs[child].cat != Cat::Expression)
          throw std::invalid_argument("primitive argument type mismatch");
        e.children.push_back(child);
      }
    } else if (v.has("local")) {
      keys(v, {"local", "fuel_events", "resource_charge"}, "local expression");
      if (v.has("fuel_events") || v.has("resource_charge")) throw Unsupported("fuel profiles and resource charges");
      e.kind = EKind::Local;
      e.target = binding(g_.locals, v.at("local").as_string("local"));
      e.type = g_.locals[e.target].type;
      e.node_kind = NK_VAR;
    } else if (v.has("control")) {
      keys(v, {"control", "type", "args", "name", "input_name", "fuel_events", "resource_charge"}, "control expression");
      if (v.has("fuel_events") || v.has("resource_charge")) throw Unsupported("fuel profiles and resource charges");
      control(v, e, env, depth);
    } else if (v.has("template") || v.has("hole") || v.has("structured")) {
      throw Unsupported("templates and structured regions");
    } else {
      throw std::invalid_argument("unknown expression form");
    }
    return append(std::move(e));
  }
  // Structural controls (catalog.cpp control signatures; compiled.cpp:498-528):
  // program(Block)->Program, block_nil()->Block, block_cons(Statement,Block)->Block,
  // if_stmt(Bool,Block,Block)->Statement, for_range(Int,Block)->Statement and
  // assign(T)->Statement / return(T)->Statement per value type. Structural
  // children carry the enclosing result type; assign and for_range name a
  // declared local of the signature's value type.
  void control(const Value& v, CExpr& e, const std::vector<Binding>& env, unsigned depth) {
    const std::string key = v.at("control").as_string("control");
    const auto open = key.find('('), close = key.find(")->");
    if (open == std::string::npos || close == std::string::npos || close < open)
      throw std::invalid_argument("unknown control signature: " + key);
    const std::string op = key.substr(0, open), list = key.substr(open + 1, close - open - 1), result = key.substr(close + 3);
    std::vector<std::string> args;
    {
      std::stringstream ss(list);
      std::string item;
      while (!list.empty() && std::getline(ss, item, ',')) args.push_back(item);
    }
    struct Arg { Cat cat; std::uint8_t type; };  // type: TAG_INVALID for a structural argument
    std::vector<Arg> want;
    Cat cat;
    bool named = false;
    const auto is_value = [](const std::string& t) {
      try { parse_type_name(t); return true; } catch (const std::invalid_argument&) { return false; }
    };
    if (op == "program" && args == std::vector<std::string>{"Block"} && result == "Program") {
      e.node_kind = NK_PROGRAM; cat = Cat::Program; want = {{Cat::Block, TAG_INVALID}};
    } else if (op == "block_nil" && args.empty() && result == "Block") {
      e.node_kind = NK_BLOCK_NIL; cat = Cat::Block;
    } else if (op == "block_cons" && args == std::vector<std::string>{"Statement", "Block"} && result == "Block") {
      e.node_kind = NK_BLOCK_CONS; cat = Cat::Block; want = {{Cat::Statement, TAG_INVALID}, {Cat::Block, TAG_INVALID}};
    } else if (op == "if_stmt" && args == std::vector<std::string>{"Bool", "Block", "Block"} && result == "Statement") {
      e.node_kind = NK_IF_STMT; cat = Cat::Statement;
      want = {{Cat::Expression, TAG_BOOL}, {Cat::Block, TAG_INVALID}, {Cat::Block, TAG_INVALID}};
    } else if (op == "for_range" && args == std::vector<std::string>{"Int", "Block"} && result == "Statement") {
      e.node_kind = NK_FOR_RANGE; cat = Cat::Statement; named = true;
      want = {{Cat::Expression, TAG_INT}, {Cat::Block, TAG_INVALID}};
    } else if ((op == "assign" || op == "return") && args.size() == 1 && is_value(args[0]) && result == "Statement") {
      e.node_kind = op == "assign" ? NK_ASSIGN : NK_RETURN; cat = Cat::Statement; named = op == "assign";
      want = {{Cat::Expression, parse_type_name(args[0])}};
    } else {
      throw std::invalid_argument("unknown control signature: " + key);
    }
    e.kind = EKind::Control;
    e.cat = cat;
    e.type = parse_type_name(v.at("type").as_string("control type"));
    if (named) {
      if (v.has("name") + v.has("input_name") != 1)
        throw std::invalid_argument("control requires exactly one name or input_name target");
      if (v.has("input_name")) {
        // A write to an evaluator input: its slot, for this execution only.
        e.input = int(binding(g_.inputs, v.at("input_name").as_string("input_name")));
        if (g_.inputs[std::size_t(e.input)].type != want[0].type) throw std::invalid_argument("control target input type mismatch");
      } else {
        e.local = int(binding(g_.locals, v.at("name").as_string("name")));
        if (g_.locals[std::size_t(e.local)].type != want[0].type) throw std::invalid_argument("control target local type mismatch");
      }
    } else if (v.has("name") || v.has("input_name")) {
      throw std::invalid_argument("control does not accept a target name");
    }
    if (e.node_kind == NK_RETURN && want[0].type != e.type)
      throw std::invalid_argument("return type disagrees with enclosing result contract");
    const auto& children = v.at("args").as_array("args");
    if (children.size() != want.size()) throw std::invalid_argument("control argument count mismatch");
    for (std::size_t i = 0; i < children.size(); ++i) {
      const auto child = expression(children[i], env, depth + 1);
      const auto& c = g_.exprs[child];
      const std::uint8_t expected = want[i].cat == Cat::Expression ? want[i].type : e.type;
      if (c.cat != want[i].cat || c.type != expected) throw std::invalid_argument("control argument category or type mismatch");
      e.children.push_back(child);
    }
  }
  // minimum_expression_nodes without templates (budget.cpp:19-55).
  std::uint32_t cost(std::uint32_t id, std::uint32_t depth) const {
    const auto& e = g_.exprs[id];
    if (e.kind == EKind::Reference) return g_.nts[e.target].minimum_by_depth[depth];
    if (!depth) return kNoId;
    std::uint64_t r = 1;
    for (auto c : e.children) {
      const auto n = cost(c, depth - 1);
      if (n == kNoId) return kNoId;
      r += n;
      if (r > g_.max_nodes) return kNoId;
    }
    return std::uint32_t(r);
  }
  void costs() {
    const auto depths = g_.max_depth;
    for (auto& nt : g_.nts) nt.minimum_by_depth.assign(depths + 1, kNoId);
    for (std::uint32_t d = 1; d <= depths; ++d) {
      bool changed;
      do {
        changed = false;
        for (const auto& p : g_.prods) {
          const auto n = cost(p.expression, d);
          auto& t = g_.nts[p.nt].minimum_by_depth[d];
          if (n < t) { t = n; changed = true; }
        }
      } while (changed);
    }
    for (const auto& nt : g_.nts) {
      std::uint32_t best = kNoId;
      for (auto n : nt.minimum_by_depth) best = std::min(best, n);
      if (best == kNoId) throw std::invalid_argument("unproductive nonterminal or no derivation within search budget: " + nt.id);
    }
  }
  Grammar g_;
  std::map<std::string, std::uint32_t> ids_;
};

std::string read_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::invalid_argument("cannot read " + path);
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}
}  // namespace

const char* node_serialized_name(int kind) { return node_info(kind).name; }
int node_prefix_arity(int kind) { return node_info(kind).arity; }
bool node_known(int kind) {
  for (const auto& n : kNodes) if (n.kind == kind) return true;
  return false;
}
bool node_i0_is_name(int kind) { return node_info(kind).i0 == 1; }
bool node_i0_is_constant(int kind) { return node_info(kind).i0 == 2; }

std::uint8_t parse_type_name(const std::string& n) {
  static const char* names[] = {"Int", "Float", "Bool", "Char", "String", "IntList", "FloatList", "StringList"};
  for (int i = 0; i < 8; ++i) if (n == names[i]) return std::uint8_t(i);
  throw std::invalid_argument("unknown grammar value type: " + n);
}
const char* type_name(std::uint8_t t) {
  static const char* names[] = {"Int", "Float", "Bool", "Char", "String", "IntList", "FloatList", "StringList"};
  return t < 8 ? names[t] : "Invalid";
}

Grammar load_grammar_text(const std::string& text) {
  const json::Value doc = json::parse(text);
  Grammar g = Compiler{}.compile(doc);
  g.content_hash = grammar_content_hash(doc);
  return g;
}
Grammar load_grammar_file(const std::string& path) { return load_grammar_text(read_file(path)); }

PhaseGrammar compile_phase(const Grammar& g, unsigned root, unsigned max_depth, unsigned site_depth,
                           const std::vector<std::string>& input_order) {
  const auto fail = [](const std::string& why) -> void { throw Unsupported("GPU phase construction: " + why); };
  if (root >= g.nts.size() || !max_depth || max_depth > 64 || max_depth > g.max_depth || site_depth > 15) fail("invalid root/depth");
  PhaseGrammar out;
  out.depth_count = int(max_depth + 1);
  out.site_depth = site_depth > 0;
  out.exprs.resize(g.exprs.size());
  out.nts.resize(g.nts.size());
  out.domains.resize(g.domains.size());
  std::vector<bool> seen_nt(g.nts.size()), seen_expr(g.exprs.size()), seen_domain(g.domains.size());
  using Env = std::vector<int>;
  std::vector<Env> nt_env(g.nts.size()), expr_env(g.exprs.size());
  std::function<void(unsigned, const Env&)> visit_expr, visit_nt;
  bool has_let = false;
  visit_nt = [&](unsigned id, const Env& env) {
    if (seen_nt[id]) {
      if (nt_env[id] != env) fail("nonterminal entered under two lexical environments");
      return;
    }
    seen_nt[id] = true;
    nt_env[id] = env;
    const auto& nt = g.nts[id];
    if (nt.scope.size() > 8) fail("nonterminal capability");
    auto& t = out.nts[id];
    t.scope_count = int(nt.scope.size());
    t.production_offset = int(out.prods.size());
    for (auto p : nt.productions) {
      out.prods.push_back({int(g.prods[p].expression), g.prods[p].weight});
      ++t.production_count;
    }
    if (!t.production_count) fail("no production");
    if (env.size() != nt.scope.size()) fail("nonterminal environment size");
    for (auto p : nt.productions) visit_expr(g.prods[p].expression, env);
  };
  visit_expr = [&](unsigned id, const Env& env) {
    if (seen_expr[id]) {
      if (expr_env[id] != env) fail("expression entered under two lexical environments");
      return;
    }
    seen_expr[id] = true;
    expr_env[id] = env;
    const auto& e = g.exprs[id];
    auto& x = out.exprs[id];
    if (env.size() > 15) fail("lexical environment capacity");
    // A statement completes without a value; its type is the enclosing result.
    x.node.type = e.cat == Cat::Expression ? e.type : kStatementType;
    if (e.kind == EKind::Reference) {
      Env inner;
      for (auto m : e.scope_mapping) inner.push_back(env.at(m));
      x.kind = PK_REFERENCE;
      x.target = int(e.target);
      x.scope_count = int(e.scope_mapping.size());
      visit_nt(e.target, inner);
      return;
    }
    const bool traversal = e.kind == EKind::Primitive && (e.node_kind == NK_TRAVERSE || e.node_kind == NK_TRAVERSE_RANGE);
    const bool let = e.kind == EKind::Primitive && e.node_kind == NK_LET_REGION;
    if (e.children.size() > (traversal ? std::size_t(kMaxArity) : 3)) fail("arity capacity");
    x.node.arity = int(e.children.size());
    std::vector<Env> child_env(e.children.size(), env);
    for (std::size_t i = 0; i < e.children.size(); ++i) x.children[i] = int(e.children[i]);
    if (e.kind == EKind::Constant) {
      x.kind = PK_CONSTANT;
      x.target = int(e.target);
      x.native_kind = NK_CONST;
      x.node.kind = TK_CONSTANT;
      if (!seen_domain[e.target]) {
        seen_domain[e.target] = true;
        const auto& d = g.domains[e.target];
        auto& pd = out.domains[e.target];
        pd.integer_range = d.integer_range;
        pd.minimum = d.minimum;
        pd.maximum = d.maximum;
        if (!d.integer_range) {
          pd.value_offset = int(out.values.size());
          pd.value_count = int(d.values.size());
          out.values.insert(out.values.end(), d.values.begin(), d.values.end());
        }
      }
    } else if (e.kind == EKind::Input) {
      const auto& name = g.inputs.at(e.target).name;
      const auto it = std::find(input_order.begin(), input_order.end(), name);
      if (it == input_order.end()) fail("input missing from evaluator order");
      const int position = int(it - input_order.begin());
      if (position >= 64) fail("input local capacity");
      x.kind = PK_INPUT;
      x.target = position;
      x.native_kind = NK_VAR;
      x.node.kind = TK_LOCAL;
      x.node.operand = position;
    } else if (e.kind == EKind::Local) {
      // A read of a grammar local (phase_grammar.hpp:300-303).
      if (e.target >= 16) fail("grammar local capacity");
      x.kind = PK_OPERATION;
      x.native_kind = NK_VAR;
      x.node.kind = TK_VAR;
      x.node.operand = int(e.target);
      out.statements = true;
    } else if (e.kind == EKind::Control) {
      x.kind = PK_OPERATION;
      x.native_kind = e.node_kind;
      switch (e.node_kind) {
        case NK_PROGRAM: x.node.kind = TK_PROGRAM; break;
        case NK_BLOCK_CONS: x.node.kind = TK_BLOCK; break;
        case NK_BLOCK_NIL: x.node.kind = TK_NIL; break;
        case NK_ASSIGN: x.node.kind = TK_ASSIGN; break;
        case NK_RETURN: x.node.kind = TK_RETURN; break;
        case NK_IF_STMT: x.node.kind = TK_IF; break;
        case NK_FOR_RANGE: x.node.kind = TK_FOR_RANGE; break;
        default: fail("control capability");
      }
      if (e.local >= 0) {
        if (e.local >= 16) fail("grammar local capacity");
        x.node.operand = e.local;
      } else if (e.input >= 0) {
        // An input target: -1 - its evaluator slot (phase_compile.cuh:281-283).
        const auto it = std::find(input_order.begin(), input_order.end(), g.inputs.at(std::size_t(e.input)).name);
        if (it == input_order.end()) fail("input missing from evaluator order");
        x.node.operand = -1 - int(it - input_order.begin());
        out.input_writes = true;
      }
      // Assign/Return carry the exact stored or returned value type.
      if (e.node_kind == NK_ASSIGN || e.node_kind == NK_RETURN) x.node.type = g.exprs[e.children.at(0)].type;
      out.statements = true;
    } else if (e.kind == EKind::Bound) {
      // A lexical read: the binder number the environment gives the name.
      if (e.target >= env.size()) fail("bound reference outside its environment");
      x.kind = PK_BOUND;
      x.target = int(e.target);
      x.native_kind = NK_REGION_VAR;
      x.node.kind = TK_BOUND;
      x.node.operand = env[e.target];
    } else if (let) {
      // The body (argument 1) binds the value to the binder number above
      // every visible one (phase_grammar.hpp let binders).
      const int binder = env.empty() ? 0 : *std::max_element(env.begin(), env.end()) + 1;
      if (binder >= 16 || e.region_argument != 1 || e.region_bindings.size() != 1) fail("let capability");
      x.kind = PK_OPERATION;
      x.native_kind = NK_LET_REGION;
      x.node.kind = TK_LET;
      x.node.operand = binder;
      child_env.back().push_back(binder);
      out.lets = true;
      out.has_let = has_let = true;
    } else if (traversal) {
      // Element, index and accumulator bind the three binder numbers above
      // every visible one for the body (the last argument); the node keeps
      // the first of them and the direction.
      const int binder = env.empty() ? 0 : *std::max_element(env.begin(), env.end()) + 1;
      if (binder + 2 >= 16 || e.region_argument != int(e.children.size()) - 1 || e.region_bindings.size() != 3)
        fail("traversal capability");
      x.kind = PK_OPERATION;
      x.native_kind = e.node_kind;
      x.node.kind = TK_TRAVERSE;
      x.node.operand = binder;
      x.node.opcode = e.direction;
      for (int k = 0; k < 3; ++k) child_env.back().push_back(binder + k);
      out.lets = true;
    } else if (e.kind == EKind::Primitive) {
      x.kind = PK_OPERATION;
      x.native_kind = e.node_kind;
      switch (e.node_kind) {
        case NK_IF_EXPR: x.node.kind = TK_IF; break;
        case NK_AND: x.node.kind = TK_AND; break;
        case NK_OR: x.node.kind = TK_OR; break;
        default:
          x.node.kind = TK_OPERATION;
          if (e.builtin >= 0) { x.node.opcode = OP_CALL_BUILTIN; x.node.operand = e.builtin; }
          else {
            switch (e.node_kind) {
              case NK_ADD: x.node.opcode = OP_ADD; break;
              case NK_SUB: x.node.opcode = OP_SUB; break;
              case NK_MUL: x.node.opcode = OP_MUL; break;
              case NK_DIV: x.node.opcode = OP_DIV; break;
              case NK_MOD: x.node.opcode = OP_MOD; break;
              case NK_NEG: x.node.opcode = OP_NEG; break;
              case NK_NOT: x.node.opcode = OP_NOT; break;
              case NK_LT: x.node.opcode = OP_LT; break;
              case NK_LE: x.node.opcode = OP_LE; break;
              case NK_GT: x.node.opcode = OP_GT; break;
              case NK_GE: x.node.opcode = OP_GE; break;
              case NK_EQ: x.node.opcode = OP_EQ; break;
              case NK_NE: x.node.opcode = OP_NE; break;
              case NK_CHECK_INT: x.node.opcode = OP_CHECK_INT; break;
              case NK_CHECK_LIST: x.node.opcode = OP_CHECK_LIST; break;
              default: fail("unsupported primitive opcode");
            }
          }
      }
    } else {
      fail("expression kind requires native fallback");
    }
    for (std::size_t i = 0; i < e.children.size(); ++i) {
      if (child_env[i].size() > 15) fail("lexical environment capacity");
      x.child_scope[i] = static_cast<signed char>(child_env[i].size());
      visit_expr(e.children[i], child_env[i]);
    }
  };
  {  // the root environment numbers its scope 0..n-1
    Env root_env;
    for (std::size_t i = 0; i < g.nts[root].scope.size(); ++i) root_env.push_back(int(i));
    visit_nt(root, root_env);
  }
  // Exported programs number their traversal binders after those of the
  // first member and of the prototype it was exported against, so the
  // engine admits traversals only where every program has the same number
  // of them.
  {
    const auto is_traversal = [&](const CExpr& e) {
      return e.kind == EKind::Primitive && (e.node_kind == NK_TRAVERSE || e.node_kind == NK_TRAVERSE_RANGE);
    };
    const auto is_let = [&](const CExpr& e) { return e.kind == EKind::Primitive && e.node_kind == NK_LET_REGION; };
    // Nonterminals that can derive a traversal (fixpoint over references).
    std::vector<char> reach(g.nts.size(), 0);
    std::function<bool(unsigned)> expr_reach = [&](unsigned id) -> bool {
      const auto& e = g.exprs[id];
      if (e.kind == EKind::Reference) return reach[e.target] != 0;
      if (is_traversal(e) || is_let(e)) return true;
      for (auto c : e.children) if (expr_reach(c)) return true;
      return false;
    };
    for (bool changed = true; changed;) {
      changed = false;
      for (std::size_t n = 0; n < g.nts.size(); ++n) {
        if (reach[n]) continue;
        for (auto p : g.nts[n].productions)
          if (expr_reach(g.prods[p].expression)) { reach[n] = 1; changed = true; break; }
      }
    }
    std::vector<int> nt_count(g.nts.size(), -2);  // -2 unvisited, -3 in progress
    int w_traversal = 3, w_let = 1;  // binder numbers first, then traversals
    std::function<int(unsigned)> count_expr, count_nt;
    count_expr = [&](unsigned id) -> int {
      const auto& e = g.exprs[id];
      if (e.kind == EKind::Reference) return count_nt(e.target);
      int n = is_traversal(e) ? w_traversal : is_let(e) ? w_let : 0;
      for (auto c : e.children) {
        const int k = count_expr(c);
        if (k < 0) return -1;
        n += k;
      }
      return n;
    };
    count_nt = [&](unsigned id) -> int {
      if (!reach[id]) return 0;
      if (nt_count[id] == -3) return -1;  // a traversal under recursion: no fixed count
      if (nt_count[id] != -2) return nt_count[id];
      nt_count[id] = -3;
      int n = -2;
      for (auto p : g.nts[id].productions) {
        const int k = count_expr(g.prods[p].expression);
        if (k < 0 || (n != -2 && k != n)) { n = -1; break; }
        n = k;
      }
      nt_count[id] = n < 0 ? -1 : n;
      return nt_count[id];
    };
    const int k = count_nt(root);
    if (k < 0) fail("traversal or let count that varies between programs");
    out.binder_ids = k;
    std::fill(nt_count.begin(), nt_count.end(), -2);
    w_traversal = 1;
    w_let = 0;
    out.traversals = count_nt(root);
    // The reference's traversal lowering consults facts about let-bound
    // values (phase_compile.cuh:254-266), which this lowering does not restate.
    if (has_let && out.traversals > 0) fail("let together with traversals");
  }
  // Each production root must resolve to a distinct node (site annotation
  // recovers a production from its first node's origin).
  std::vector<bool> resolved(out.exprs.size());
  for (const auto& p : out.prods) {
    const int e = p.expression;
    if (out.exprs[e].kind == PK_REFERENCE) fail("root aliases require native fallback");
    if (resolved[e]) fail("productions sharing a root node require native fallback");
    resolved[e] = true;
  }
  const std::size_t D = out.depth_count;
  PExpr entry;
  entry.kind = PK_REFERENCE;
  entry.target = int(root);
  entry.scope_count = int(g.nts[root].scope.size());
  entry.node.type = g.nts[root].type;
  out.root = int(out.exprs.size());
  out.exprs.push_back(entry);
  if (!site_depth) {
    out.logical_count = 1;
    out.minimum.assign(out.exprs.size() * D, kNoId);
    // minimum_expression_nodes (budget.cpp) over the compiled grammar
    std::function<std::uint32_t(std::uint32_t, std::uint32_t)> cost = [&](std::uint32_t id, std::uint32_t d) -> std::uint32_t {
      const auto& e = g.exprs[id];
      if (e.kind == EKind::Reference) return g.nts[e.target].minimum_by_depth.at(d);
      if (!d) return kNoId;
      std::uint64_t r = 1;
      for (auto c : e.children) {
        const auto n = cost(c, d - 1);
        if (n == kNoId) return kNoId;
        r += n;
        if (r > g.max_nodes) return kNoId;
      }
      return std::uint32_t(r);
    };
    for (std::size_t i = 0; i < seen_expr.size(); ++i)
      if (seen_expr[i]) for (unsigned d = 0; d <= max_depth; ++d) out.minimum[i * D + d] = cost(std::uint32_t(i), d);
    for (unsigned d = 0; d <= max_depth; ++d) out.minimum[std::size_t(out.root) * D + d] = g.nts[root].minimum_by_depth.at(d);
  } else {
    out.logical_count = int(site_depth + 1);
    const std::size_t L = out.logical_count;
    const std::size_t size = out.exprs.size() * D * L;
    if (size > 4u * 1024 * 1024) fail("site depth table capacity");
    out.minimum.assign(size, kNoId);
    std::vector<char> computed(size);
    std::function<unsigned(int, int, int)> cost = [&](int id, int d, int l) -> unsigned {
      const std::size_t off = (std::size_t(id) * D + d) * L + l;
      if (computed[off]) return out.minimum[off];
      computed[off] = 1;
      const auto& e = out.exprs[id];
      unsigned result = kNoId;
      if (e.kind == PK_REFERENCE) {
        if (l) {
          const auto& nt = out.nts[e.target];
          for (int i = 0; i < nt.production_count; ++i)
            result = std::min(result, cost(out.prods[nt.production_offset + i].expression, d, l - 1));
        }
      } else if (d > 0) {
        std::uint64_t total = 1;
        for (int i = 0; i < e.node.arity; ++i) total += cost(e.children[i], d - 1, l);
        result = total >= kNoId ? kNoId : unsigned(total);
      }
      out.minimum[off] = result;
      return result;
    };
    for (std::size_t i = 0; i < out.exprs.size(); ++i)
      if (int(i) == out.root || seen_expr[i])
        for (unsigned d = 0; d <= max_depth; ++d)
          for (unsigned l = 0; l <= site_depth; ++l) cost(int(i), int(d), int(l));
  }
  // Variation flags (phase_variation.hpp:6-14).
  for (std::size_t i = 0; i < out.nts.size(); ++i)
    if (out.nts[i].production_count) out.nts[i].fixed = !g.nts[i].variation;
  // Static weight normalization (phase_grammar.hpp:48-67).
  const std::size_t DL = std::size_t(out.depth_count) * out.logical_count;
  for (auto& nt : out.nts) {
    if (!nt.production_count) continue;
    double maximum = 0;
    for (int i = 0; i < nt.production_count; ++i) maximum = std::max(maximum, out.prods[nt.production_offset + i].weight);
    bool proven = maximum > 0;
    for (std::size_t k = 0; k < DL && proven; ++k) {
      unsigned cheapest = kNoId, max_weight_cheapest = kNoId;
      for (int i = 0; i < nt.production_count; ++i) {
        const auto& p = out.prods[nt.production_offset + i];
        const unsigned c = out.minimum.at(std::size_t(p.expression) * DL + k);
        cheapest = std::min(cheapest, c);
        if (p.weight == maximum) max_weight_cheapest = std::min(max_weight_cheapest, c);
      }
      proven = cheapest == max_weight_cheapest;
    }
    if (!proven) continue;
    for (int i = 0; i < nt.production_count; ++i) out.prods[nt.production_offset + i].weight /= maximum;
    nt.normalized = 1;
  }
  return out;
}

unsigned hidden_local_bound(const Grammar& g) {
  const std::size_t levels = std::size_t(g.max_depth) + 1;
  std::vector<int> memo(g.exprs.size() * levels, -1);
  std::vector<char> open(g.exprs.size() * levels, 0);
  // peak(e, d): temporaries live at once inside expression e placed with d
  // levels of depth left; a reference takes its alternative's place.
  std::function<int(std::uint32_t, int)> peak = [&](std::uint32_t id, int d) -> int {
    if (d <= 0) return 0;
    const std::size_t key = std::size_t(id) * levels + std::size_t(d);
    if (memo[key] >= 0) return memo[key];
    if (open[key]) return 64;  // a reference cycle that spends no depth: no finite bound
    open[key] = 1;
    const CExpr& e = g.exprs[id];
    int v = 0;
    if (e.kind == EKind::Reference) {
      for (const auto p : g.nts[e.target].productions) v = std::max(v, peak(g.prods[p].expression, d));
    } else if (e.kind == EKind::Control && e.node_kind == NK_FOR_RANGE) {
      // The bound temporary is live over the bound expression; the counter
      // joins it over the body.
      v = std::max(1 + peak(e.children.at(0), d - 1), 2 + peak(e.children.at(1), d - 1));
    } else {
      for (const auto c : e.children) v = std::max(v, peak(c, d - 1));
      if (e.direction != 0) v += 9;                // a traversal's temporaries
      if (e.node_kind == NK_LET_REGION) v += 1;    // a let's value slot
    }
    open[key] = 0;
    return memo[key] = std::min(v, 64);
  };
  int v = 0;
  for (const auto p : g.nts[g.entry].productions) v = std::max(v, peak(g.prods[p].expression, int(g.max_depth)));
  return unsigned(std::min<std::size_t>(64, g.locals.size() + std::size_t(v)));
}

Layout whole_program_layout(const Grammar& g) {
  Layout l;
  const auto& entry = g.nts[g.entry];
  if (entry.cat == Cat::Program) {
    // A Program entry already contains its native envelope: the phase is the
    // whole program, at the full depth budget, with no terminator.
    if (g.max_nodes > 2048 || g.max_depth > 64) throw Unsupported("GPU genotype capacity requires fallback");
    if (entry.minimum_by_depth.at(g.max_depth) > g.max_nodes) throw Unsupported("generation request has no feasible derivation");
    l.fixed_nodes = 0;
    l.slot_max_depth = g.max_depth;
    l.max_variable_nodes = g.max_nodes;
    l.entry_nt = g.entry;
    l.program_terminator = false;
    return l;
  }
  if (entry.cat != Cat::Expression) throw Unsupported("entry category");
  if (g.max_depth <= 3 || g.max_nodes <= 4)
    throw Unsupported("expression entry needs four envelope nodes and three prefix levels");
  if (g.max_nodes > 2048 || g.max_depth > 64) throw Unsupported("GPU genotype capacity requires fallback");
  if (entry.minimum_by_depth.at(g.max_depth - 3) > g.max_nodes - 4)
    throw Unsupported("generation request has no feasible derivation");
  l.fixed_nodes = 4;
  l.slot_max_depth = g.max_depth - 3;
  l.max_variable_nodes = g.max_nodes - 4;
  l.entry_nt = g.entry;
  l.program_terminator = true;
  return l;
}

}  // namespace gagp_amd
