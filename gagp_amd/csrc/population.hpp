// Imported programs: ast-prefix-v2 JSON -> native AST -> the grammar
// expression each node derives from -> one phase row (origins, sites,
// constants) of the resident engine.
//
// Reference behaviour followed:
//   * decode_ast_json (cpp/src/cli/program_json.cpp:506-640);
//   * canonical membership: the first production, in nonterminal order,
//     whose expression matches the subtree, memoized per (nonterminal, node,
//     lexical environment), with the node, depth and envelope checks of
//     Matcher::run (cpp/src/evolution/grammar/membership.cpp:115-170,
//     331-350,448-488);
//   * the whole-program phase span and gene of import_native_phase_layout
//     (cpp/src/evolution/repro/gpu/phase_layout.hpp:92-260);
//   * site annotation and the site-depth check of an imported row
//     (cpp/src/evolution/repro/gpu/device/phase_variation.cuh:23-77,127-142,
//     cpp/src/evolution/repro/gpu/phase_population_cpu.cpp:220-238);
//   * load_population (cpp/src/cli/run.cpp:143-173).
#pragma once
#include <cstddef>
#include <string>
#include <vector>
#include "export.hpp"
#include "grammar.hpp"
#include "json.hpp"

namespace gagp_amd {

// One ast-prefix-v2 program document. Constants outside Int and Bool and
// fuel annotations are outside the engine's values and raise Unsupported.
Ast decode_ast(const json::Value& doc);

// The program a `gagp eval` document names: a run result's best program, a
// population's first program, or a bare program (cpp/src/cli/run.cpp:388-393).
const json::Value& program_of_document(const json::Value& doc);

// The compiled-grammar expression of every node of the entry derivation, in
// prefix order: nodes 3..n-2 of an Expression entry (inside its PROGRAM,
// BLOCK_CONS, RETURN ... BLOCK_NIL envelope), every node of a Program entry.
// Throws std::invalid_argument when the grammar does not derive the program.
std::vector<int> derive_origins(const Grammar& g, const Ast& ast);

struct PhaseRow {
  std::vector<int> origin, site;
  std::vector<Val> cval;
};

// The phase row of a program: its derivation origins, the site of every node
// a nonterminal choice starts (annotated from the origins), and its
// constants in prefix order. `site_depth` 0 skips the site-depth bound;
// `capacity` is the row capacity of the search. A statement program must
// assign every local before it reads it (the verifier's undefined_local).
PhaseRow import_row(const Grammar& g, const PhaseGrammar& phase, const Ast& ast, unsigned site_depth, int capacity);

struct ImportedPopulation {
  std::vector<PhaseRow> rows;
  // Export layout of the search: the first program's name table, and one
  // above every lexical binder id it holds (phase_layout.hpp:326-342).
  std::vector<std::string> layout_names;
  int binder_base = 0;
};

// The first `count` programs of a `gagp run --save-all` population document
// made for this grammar.
ImportedPopulation load_population(const std::string& path, const Grammar& g, const PhaseGrammar& phase, unsigned site_depth,
                                   int capacity, std::size_t count);

}  // namespace gagp_amd
