// Genotype row -> native AST (ast-prefix-v2 JSON) and program text.
//
// Reference behaviour: export_native_phase_layout + compact_genome_tables
// (cpp/src/evolution/repro/gpu/phase_layout.hpp:316-475,
// cpp/src/evolution/repro/pack.cpp:11-77): the expression-entry envelope
// PROGRAM, BLOCK_CONS, RETURN, <phase>, BLOCK_NIL; one constant per CONST node
// in prefix order. Writers: write_ast_json (cpp/src/cli/program_json.cpp:211-232)
// and ProgramPrinter (program_json.cpp:876-946).
//
// Traversal binders take fresh public ids in prefix order, above every id of
// the layout prototype (phase_layout.hpp:330-340,388-397). A search's layout
// is its first initial member, itself exported against a natively generated
// prototype that numbers its binders from 0; with K traversals in every
// program those hold ids 0..3K-1 and 3K..6K-1, so exported programs number
// theirs from 6K.
#pragma once
#include <string>
#include <vector>
#include "grammar.hpp"
#include "json.hpp"

namespace gagp_amd {

struct AstNode { int kind, i0, i1; };
// A traversal's bindings (lexical_regions) and direction (traversal_specs).
struct LexicalBinding { int id; std::uint8_t type; };
struct LexicalRegion { int node_index, body_argument; std::vector<LexicalBinding> bindings; };
struct TraversalSpec { int node_index; bool reverse; };
struct Ast {
  std::vector<AstNode> nodes;
  std::vector<std::string> names;
  std::vector<Val> consts;
  std::vector<LexicalRegion> lexical_regions;
  std::vector<TraversalSpec> traversal_specs;
};

// Names of the grammar that the program reads or writes, in declaration order
// (inputs, then locals). Applied to the first initial member this is the name
// table of the search's layout prototype: the reference's generator lists
// every declared name (cpp/src/evolution/grammar/generate.cpp:32-35) and the
// member's export compacts that to the names it uses.
std::vector<std::string> names_used(const Grammar& g, const std::vector<int>& origins);

// `origins`/`consts` are one phase row; `phase` is its expression table.
// `layout_names` is the name table of the search's layout prototype (its
// first initial member's used names in declaration order): an exported
// program lists those it uses in that order, then its other names in
// first-use order (phase_layout.hpp name_index + pack.cpp compaction). An
// Expression entry is wrapped in the PROGRAM, BLOCK_CONS, RETURN ... BLOCK_NIL
// envelope; a Program entry is the phase itself. `binder_base` is the first
// binder id the program's lets and traversals take: one above every id of the
// layout prototype, which for a generated population is twice the grammar's
// binder count (-1 selects that).
Ast export_whole_program(const Grammar& g, const PhaseGrammar& phase, const std::vector<std::string>& layout_names,
                         const std::vector<int>& origins, const std::vector<Val>& consts, int binder_base = -1);
json::Value ast_json(const Ast& ast);
std::string program_text(const Ast& ast);

}  // namespace gagp_amd
