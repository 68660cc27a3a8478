// Test driver (CPU only): imports the programs of a population document with
// the engine's import path (AST decoding, grammar membership, site
// annotation) and exports each phase row again, printing one JSON line per
// program. tests/test_import.py checks that every export reproduces the
// program it came from and that the annotated sites are those the device's
// generator writes.
//
//   import_driver TASK.json POPULATION.json COUNT SITE_DEPTH
//     one line per program: {"i", "nodes", "origins", "sites", "same"}
//   import_driver TASK.json POPULATION.json COUNT SITE_DEPTH load
//     load_population as the engine calls it: {"rows", "names", "binder_base"}
//   import_driver TASK.json - N SITE_DEPTH generate
//     N generated rows of the grammar with their sites, re-annotated from
//     the origins: {"i", "nodes", "same_sites"}
//   import_driver TASK.json POPULATION.json COUNT SITE_DEPTH each
//     each program imported on its own: {"i", "error"} (null when admitted)
// Errors print {"error": "..."} and exit 2.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "export.hpp"
#include "grammar.hpp"
#include "phase.hpp"
#include "population.hpp"
#include "task.hpp"

using namespace gagp_amd;

namespace {
std::string ints(const std::vector<int>& v) {
  std::string s = "[";
  for (std::size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + "]";
}
std::string quoted(const std::string& s) {
  json::Value v = json::make_string(s);
  return json::canonical(v);
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: import_driver TASK.json POPULATION.json COUNT SITE_DEPTH [load|generate]\n");
    return 2;
  }
  const std::string mode = argc > 5 ? argv[5] : "";
  try {
    const Task task = load_task(argv[1]);
    const Grammar g = load_grammar_file(task.grammar_path);
    const Cases train = load_cases(task.train_path);
    const unsigned site_depth = unsigned(std::atoi(argv[4]));
    const PhaseGrammar ph = compile_phase(g, g.entry, g.max_depth, site_depth, train.input_names);
    const Layout layout = whole_program_layout(g);
    const int capacity = std::min<int>(kPhaseCapacity, int(layout.max_variable_nodes));
    const int count = std::atoi(argv[3]);
    if (mode == "load") {
      const ImportedPopulation pop = load_population(argv[2], g, ph, site_depth, capacity, std::size_t(count));
      std::string names = "[";
      for (std::size_t i = 0; i < pop.layout_names.size(); ++i) names += (i ? "," : "") + quoted(pop.layout_names[i]);
      std::printf("{\"rows\":%zu,\"names\":%s],\"binder_base\":%d}\n", pop.rows.size(), names.c_str(), pop.binder_base);
      return 0;
    }
    if (mode == "generate") {
      PhaseView gv{ph.exprs.data(), int(ph.exprs.size()), ph.prods.data(), int(ph.prods.size()), ph.nts.data(), int(ph.nts.size()),
                   ph.domains.data(), int(ph.domains.size()), ph.values.data(), int(ph.values.size()), ph.minimum.data(),
                   ph.depth_count, ph.logical_count, ph.site_depth ? 1 : 0, ph.root};
      const std::size_t cap = std::size_t(capacity);
      std::vector<int> origin(cap), site(cap);
      std::vector<Val> cval(cap);
      for (int i = 0; i < count; ++i) {
        const GenResult r = generate_phase(gv, ph.root, 0, int(layout.slot_max_depth), capacity, 0x9e3779b97f4a7c15ull * (i + 1),
                                           origin.data(), cval.data(), site.data(), int(site_depth));
        if (r.status != ST_OK) continue;
        // Statement rows that read an unassigned local are not members (the
        // engine's generator draws again).
        if (g.nts[g.entry].cat == Cat::Program &&
            !locals_defined(gv, CRow{origin.data(), site.data(), cval.data(), r.nodes, r.constants}))
          continue;
        // The sites the annotation recovers from the origins alone, through
        // the import path's export -> import round trip.
        const std::vector<int> o(origin.begin(), origin.begin() + r.nodes);
        const std::vector<Val> c(cval.begin(), cval.begin() + r.constants);
        const Ast ast = export_whole_program(g, ph, names_used(g, o), o, c);
        const PhaseRow row = import_row(g, ph, ast, site_depth, capacity);
        const bool same = row.origin == o && std::vector<int>(site.begin(), site.begin() + r.nodes) == row.site;
        std::printf("{\"i\":%d,\"nodes\":%d,\"same_sites\":%s}\n", i, r.nodes, same ? "true" : "false");
      }
      return 0;
    }
    std::ifstream in(argv[2], std::ios::binary);
    std::ostringstream text;
    text << in.rdbuf();
    const json::Value doc = json::parse(text.str());
    const auto& programs = doc.at("programs").as_array("programs");
    if (mode == "each") {
      for (int i = 0; i < count && std::size_t(i) < programs.size(); ++i) {
        std::string error = "null";
        try {
          (void)import_row(g, ph, decode_ast(programs[std::size_t(i)]), site_depth, capacity);
        } catch (const std::exception& e) {
          error = quoted(e.what());
        }
        std::printf("{\"i\":%d,\"error\":%s}\n", i, error.c_str());
      }
      return 0;
    }
    for (int i = 0; i < count && std::size_t(i) < programs.size(); ++i) {
      const Ast ast = decode_ast(programs[std::size_t(i)]);
      const PhaseRow row = import_row(g, ph, ast, site_depth, capacity);
      // Export against the program's own names and first binder id: the
      // result must be the document itself.
      int base = -1;
      for (const auto& r : ast.lexical_regions)
        for (const auto& b : r.bindings) base = base < 0 ? b.id : std::min(base, b.id);
      const Ast again = export_whole_program(g, ph, ast.names, row.origin, row.cval, base < 0 ? 0 : base);
      const bool same = json::canonical(ast_json(again)) == json::canonical(programs[std::size_t(i)]);
      std::printf("{\"i\":%d,\"nodes\":%zu,\"origins\":%s,\"sites\":%s,\"same\":%s}\n", i, ast.nodes.size(),
                  ints(row.origin).c_str(), ints(row.site).c_str(), same ? "true" : "false");
    }
  } catch (const std::exception& e) {
    std::printf("{\"error\":%s,\"unsupported\":%s}\n", quoted(e.what()).c_str(),
                dynamic_cast<const Unsupported*>(&e) ? "true" : "false");
    return 2;
  }
  return 0;
}
