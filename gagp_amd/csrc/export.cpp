// Genotype export (see export.hpp).
#include "export.hpp"

#include <algorithm>
#include <functional>
#include <stdexcept>

namespace gagp_amd {

namespace {
// The grammar name a node reads or writes (an input, a local read, or the
// target of assign/for_range), else null.
const std::string* node_name(const Grammar& g, const CExpr& e) {
  if (e.kind == EKind::Input) return &g.inputs.at(e.target).name;
  if (e.kind == EKind::Local) return &g.locals.at(e.target).name;
  if (e.kind == EKind::Control && e.local >= 0) return &g.locals.at(std::size_t(e.local)).name;
  if (e.kind == EKind::Control && e.input >= 0) return &g.inputs.at(std::size_t(e.input)).name;
  return nullptr;
}
}  // namespace

std::vector<std::string> names_used(const Grammar& g, const std::vector<int>& origins) {
  const std::size_t ni = g.inputs.size();
  std::vector<char> used(ni + g.locals.size(), 0);
  for (const int o : origins) {
    if (o < 0 || o >= int(g.exprs.size())) throw std::runtime_error("unknown expression origin");
    const CExpr& e = g.exprs[std::size_t(o)];
    if (e.kind == EKind::Input) used.at(e.target) = 1;
    else if (e.kind == EKind::Local) used.at(ni + e.target) = 1;
    else if (e.kind == EKind::Control && e.local >= 0) used.at(ni + std::size_t(e.local)) = 1;
    else if (e.kind == EKind::Control && e.input >= 0) used.at(std::size_t(e.input)) = 1;
  }
  std::vector<std::string> out;
  for (std::size_t k = 0; k < used.size(); ++k)
    if (used[k]) out.push_back(k < ni ? g.inputs[k].name : g.locals[k - ni].name);
  return out;
}

Ast export_whole_program(const Grammar& g, const PhaseGrammar& phase, const std::vector<std::string>& layout_names,
                         const std::vector<int>& origins, const std::vector<Val>& consts, int binder_base) {
  Ast ast;
  const bool envelope = g.nts[g.entry].cat == Cat::Expression;
  if (envelope) {
    ast.nodes.push_back({NK_PROGRAM, 0, 0});
    ast.nodes.push_back({NK_BLOCK_CONS, 0, 0});
    ast.nodes.push_back({NK_RETURN, 0, 0});
  }
  const int base = int(ast.nodes.size());
  // Name table: the layout names the program uses (their order), then its
  // other names in first-use order; nodes are patched once it is complete.
  std::vector<std::string> table = layout_names;
  std::vector<int> name_of_node(origins.size(), -1);
  std::size_t next_const = 0, j = 0;
  // Binder ids follow those of the layout prototype (phase_layout.hpp:334-342):
  // a generated population's first member holds ids above those of the
  // prototype it was exported against, twice the fixed count in all.
  int next_binder = binder_base >= 0 ? binder_base : 2 * phase.binder_ids;
  // Prefix-order emission; `env` maps binder numbers to public ids.
  std::function<void(const std::vector<int>&)> emit = [&](const std::vector<int>& env) {
    if (j >= origins.size()) throw std::runtime_error("invalid gene shape");
    const std::size_t at = j++;
    const int o = origins[at];
    if (o < 0 || o >= int(g.exprs.size())) throw std::runtime_error("unknown expression origin");
    const CExpr& e = g.exprs[std::size_t(o)];
    const PExpr& m = phase.exprs[std::size_t(o)];
    AstNode n{m.native_kind, 0, 0};
    int walk = -1, let_binder = -1;
    if (const std::string* name = node_name(g, e)) {
      auto it = std::find(table.begin(), table.end(), *name);
      if (it == table.end()) it = table.insert(table.end(), *name);
      name_of_node[at] = int(it - table.begin());
    }
    if (e.kind == EKind::Constant) {
      if (next_const >= consts.size()) throw std::runtime_error("constant index");
      n.i0 = int(ast.consts.size());
      ast.consts.push_back(consts[next_const++]);
    } else if (e.kind == EKind::Bound) {
      const int b = m.node.operand;
      if (b < 0 || b >= 16 || env[std::size_t(b)] < 0) throw std::runtime_error("lexical index");
      n.i0 = env[std::size_t(b)];
    } else if (m.node.kind == TK_LET) {
      // Recorded before its subtrees (phase_layout.hpp:400-404).
      let_binder = next_binder++;
      ast.lexical_regions.push_back(LexicalRegion{int(ast.nodes.size()), 1, {{let_binder, e.region_bindings.at(0).type}}});
    } else if (m.node.kind == TK_TRAVERSE) {
      // Recorded before its subtrees.
      walk = next_binder;
      LexicalRegion region{int(ast.nodes.size()), e.region_argument, {}};
      for (int k = 0; k < 3; ++k) region.bindings.push_back({next_binder++, e.region_bindings[std::size_t(k)].type});
      ast.lexical_regions.push_back(region);
      ast.traversal_specs.push_back({int(ast.nodes.size()), e.direction == 2});
    }
    ast.nodes.push_back(n);
    for (int c = 0; c < m.node.arity; ++c) {
      if (let_binder >= 0 && c == 1) {
        std::vector<int> inner = env;
        inner[std::size_t(m.node.operand)] = let_binder;
        emit(inner);
      } else if (walk >= 0 && c == m.node.arity - 1) {
        std::vector<int> inner = env;
        for (int k = 0; k < 3; ++k) inner[std::size_t(m.node.operand + k)] = walk + k;
        emit(inner);
      } else {
        emit(env);
      }
    }
  };
  emit(std::vector<int>(16, -1));
  if (j != origins.size()) throw std::runtime_error("invalid gene shape");
  if (envelope) ast.nodes.push_back({NK_BLOCK_NIL, 0, 0});
  // Stable compaction to the names in use.
  std::vector<int> used(table.size(), 0), remap(table.size(), -1);
  for (const int k : name_of_node)
    if (k >= 0) used[std::size_t(k)] = 1;
  for (std::size_t i = 0; i < table.size(); ++i)
    if (used[i]) { remap[i] = int(ast.names.size()); ast.names.push_back(table[i]); }
  for (std::size_t k = 0; k < origins.size(); ++k)
    if (name_of_node[k] >= 0) ast.nodes[std::size_t(base) + k].i0 = remap[std::size_t(name_of_node[k])];
  return ast;
}

namespace {
json::Value typed_value_json(const Val& v) {
  json::Value o = json::make_object();
  if (v.tag == TAG_INT) { o.obj["type"] = json::make_string("int"); o.obj["value"] = json::make_number(double(v.i)); }
  else if (v.tag == TAG_BOOL) { o.obj["type"] = json::make_string("bool"); o.obj["value"] = json::make_bool(v.b != 0); }
  else throw std::runtime_error("AST constant type outside the engine's scalar values");
  return o;
}
}  // namespace

json::Value ast_json(const Ast& ast) {
  json::Value o = json::make_object();
  o.obj["version"] = json::make_string("ast-prefix-v2");
  json::Value nodes = json::make_array();
  for (const auto& n : ast.nodes) {
    json::Value x = json::make_object();
    x.obj["kind"] = json::make_number(n.kind);
    x.obj["i0"] = json::make_number(n.i0);
    x.obj["i1"] = json::make_number(n.i1);
    nodes.arr.push_back(x);
  }
  o.obj["nodes"] = nodes;
  json::Value names = json::make_array();
  for (const auto& n : ast.names) names.arr.push_back(json::make_string(n));
  o.obj["names"] = names;
  json::Value consts = json::make_array();
  for (const auto& c : ast.consts) consts.arr.push_back(typed_value_json(c));
  o.obj["consts"] = consts;
  if (!ast.lexical_regions.empty()) {
    json::Value regions = json::make_array();
    for (const auto& r : ast.lexical_regions) {
      json::Value x = json::make_object();
      x.obj["node_index"] = json::make_number(r.node_index);
      x.obj["body_argument"] = json::make_number(r.body_argument);
      json::Value bs = json::make_array();
      for (const auto& b : r.bindings) {
        json::Value y = json::make_object();
        y.obj["id"] = json::make_number(b.id);
        y.obj["type"] = json::make_string(type_name(b.type));
        bs.arr.push_back(y);
      }
      x.obj["bindings"] = bs;
      regions.arr.push_back(x);
    }
    o.obj["lexical_regions"] = regions;
  }
  if (!ast.traversal_specs.empty()) {
    json::Value specs = json::make_array();
    for (const auto& t : ast.traversal_specs) {
      json::Value x = json::make_object();
      x.obj["node_index"] = json::make_number(t.node_index);
      x.obj["direction"] = json::make_string(t.reverse ? "reverse" : "forward");
      specs.arr.push_back(x);
    }
    o.obj["traversal_specs"] = specs;
  }
  return o;
}

namespace {
struct Printer {
  const Ast& ast;
  std::size_t pos = 0;
  void node(std::string& out) {
    if (pos >= ast.nodes.size()) throw std::runtime_error("truncated AST program");
    const std::size_t index = pos;
    const AstNode n = ast.nodes[pos++];
    switch (n.kind) {
      case NK_CONST: {
        const Val& v = ast.consts.at(n.i0);
        if (v.tag == TAG_INT) out += std::to_string(v.i);
        else if (v.tag == TAG_BOOL) out += v.b ? "true" : "false";
        else throw std::runtime_error("constant text outside the engine's scalar values");
        return;
      }
      case NK_VAR: case NK_BOUND_VAR: out += ast.names.at(n.i0); return;
      case NK_REGION_VAR: out += "v" + std::to_string(n.i0); return;
      case NK_PROGRAM: node(out); return;
      case NK_BLOCK_NIL: out += "(block)"; return;
      case NK_BLOCK_CONS: {
        out += "(block";
        do {
          out += ' ';
          node(out);
        } while (ast.nodes.at(pos).kind == NK_BLOCK_CONS && ++pos);
        ++pos;
        out += ')';
        return;
      }
      default: break;
    }
    std::string name = node_serialized_name(n.kind);
    if (name.compare(0, 5, "call_") == 0) name = name.substr(5);
    out += '(';
    out += name;
    if (node_i0_is_name(n.kind)) out += " " + ast.names.at(n.i0);
    // A binder argument shows its bound variables first: `[v6 v7 v8] body`.
    const LexicalRegion* binder = nullptr;
    for (const auto& r : ast.lexical_regions)
      if (r.node_index == int(index)) binder = &r;
    const int arity = node_prefix_arity(n.kind);
    for (int i = 0; i < arity; ++i) {
      out += ' ';
      if (binder && i == binder->body_argument) {
        out += '[';
        for (std::size_t b = 0; b < binder->bindings.size(); ++b) out += (b ? " v" : "v") + std::to_string(binder->bindings[b].id);
        out += "] ";
      }
      node(out);
    }
    out += ')';
  }
};
}  // namespace

std::string program_text(const Ast& ast) {
  Printer p{ast};
  std::string out;
  const auto& n = ast.nodes;
  if (n.size() >= 4 && n[0].kind == NK_PROGRAM && n[1].kind == NK_BLOCK_CONS && n[2].kind == NK_RETURN &&
      n.back().kind == NK_BLOCK_NIL) {
    p.pos = 3;
    p.node(out);
    if (p.pos == n.size() - 1) return out;
    out.clear();
  }
  p.pos = 0;
  p.node(out);
  return out;
}

}  // namespace gagp_amd
