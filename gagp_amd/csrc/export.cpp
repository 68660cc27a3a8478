// Genotype export (see export.hpp).
#include "export.hpp"

#include <stdexcept>

namespace gagp_amd {

Ast export_whole_program(const Grammar& g, const PhaseGrammar& phase, const std::vector<int>& origins,
                         const std::vector<Val>& consts) {
  Ast ast;
  ast.nodes.push_back({NK_PROGRAM, 0, 0});
  ast.nodes.push_back({NK_BLOCK_CONS, 0, 0});
  ast.nodes.push_back({NK_RETURN, 0, 0});
  // Prototype name table: grammar inputs then locals; compacted to used names.
  std::vector<std::string> table;
  for (const auto& in : g.inputs) table.push_back(in.name);
  for (const auto& lo : g.locals) table.push_back(lo.name);
  std::vector<int> used(table.size(), 0);
  std::vector<int> name_of_node(origins.size(), -1);
  std::size_t next_const = 0;
  for (std::size_t j = 0; j < origins.size(); ++j) {
    const int o = origins[j];
    if (o < 0 || o >= int(g.exprs.size())) throw std::runtime_error("unknown expression origin");
    const CExpr& e = g.exprs[o];
    const PExpr& m = phase.exprs[o];
    AstNode n{m.native_kind, 0, 0};
    if (e.kind == EKind::Constant) {
      if (next_const >= consts.size()) throw std::runtime_error("constant index");
      n.i0 = int(ast.consts.size());
      ast.consts.push_back(consts[next_const++]);
    } else if (e.kind == EKind::Input) {
      const int idx = int(e.target);  // grammar declaration order = prototype table order
      used[idx] = 1;
      name_of_node[j] = idx;
    }
    ast.nodes.push_back(n);
  }
  ast.nodes.push_back({NK_BLOCK_NIL, 0, 0});
  std::vector<int> remap(table.size(), -1);
  for (std::size_t i = 0; i < table.size(); ++i)
    if (used[i]) { remap[i] = int(ast.names.size()); ast.names.push_back(table[i]); }
  for (std::size_t j = 0; j < origins.size(); ++j)
    if (name_of_node[j] >= 0) ast.nodes[3 + j].i0 = remap[name_of_node[j]];
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
  return o;
}

namespace {
struct Printer {
  const Ast& ast;
  std::size_t pos = 0;
  void node(std::string& out) {
    if (pos >= ast.nodes.size()) throw std::runtime_error("truncated AST program");
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
    const int arity = node_prefix_arity(n.kind);
    for (int i = 0; i < arity; ++i) {
      out += ' ';
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
