// Grammar front end: `grammar-definition-v2` (the subset this engine
// admits) -> compiled grammar -> device phase tables.
//
// Reference behaviour followed:
//   * resolution sorts nonterminals and alternatives by stable id
//     (cpp/src/evolution/grammar/definition.cpp:104-109,434-464);
//   * compilation numbers nonterminals in that order, productions in
//     nonterminal/alternative order and expressions in post-order (children
//     first) (cpp/src/evolution/grammar/compiled.cpp:142-215,433-700);
//   * nonterminal minimum node tables by depth (compiled.cpp:711-731,
//     budget.cpp:19-55);
//   * the phase profile: expression table, production ranges, site-depth
//     minimum table and static weight normalization
//     (cpp/src/evolution/repro/gpu/phase_grammar.hpp:48-67,102-531).
// Constructs outside the admitted subset raise `Unsupported` with a reason.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include "core.hpp"
#include "json.hpp"

namespace gagp_amd {

struct Unsupported : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

// Native AST node kinds; the numbers are the ast-prefix-v2 serialization
// (cpp/include/gagp/evolution/ast_program.hpp:26-93).
enum NodeKind : int {
  NK_PROGRAM = 0, NK_BLOCK_NIL = 1, NK_BLOCK_CONS = 2, NK_ASSIGN = 3, NK_IF_STMT = 4, NK_FOR_RANGE = 5,
  NK_RETURN = 6, NK_CONST = 7, NK_VAR = 8, NK_NEG = 9, NK_NOT = 10, NK_ADD = 11, NK_SUB = 12, NK_MUL = 13,
  NK_DIV = 14, NK_MOD = 15, NK_LT = 16, NK_LE = 17, NK_GT = 18, NK_GE = 19, NK_EQ = 20, NK_NE = 21,
  NK_AND = 22, NK_OR = 23, NK_IF_EXPR = 24, NK_CALL_ABS = 25, NK_CALL_MIN = 26, NK_CALL_MAX = 27,
  NK_CALL_CLIP = 28, NK_CALL_IDIV0 = 29, NK_CALL_IMOD0 = 30, NK_CALL_LEN = 31, NK_CALL_INDEX = 34,
  NK_BOUND_VAR = 52, NK_LET_REGION = 71, NK_TRAVERSE = 72, NK_TRAVERSE_RANGE = 73, NK_REGION_VAR = 74,
  NK_CHECK_INT = 75, NK_CHECK_LIST = 76, NK_CALL_TRUNC_DIV0 = 78, NK_CALL_TRUNC_MOD0 = 79,
  NK_CALL_INDEX_NONNEG = 80, NK_CALL_INDEX_CLAMPED = 81,
};
const char* node_serialized_name(int kind);
int node_prefix_arity(int kind);
bool node_known(int kind);  // a kind of the table above
bool node_i0_is_name(int kind);
bool node_i0_is_constant(int kind);

enum class Cat { Program, Block, Statement, Expression };
enum class EKind { Primitive, Reference, Constant, Input, Bound, Control, Local };

struct Binding { std::string name; std::uint8_t type; };

struct CExpr {
  EKind kind = EKind::Primitive;
  std::uint8_t type = TAG_INVALID;
  Cat cat = Cat::Expression;
  std::uint32_t target = 0;  // nonterminal / domain / input / binding / local
  std::vector<std::uint32_t> children;
  std::vector<std::uint32_t> scope_mapping;
  int node_kind = -1;        // lowering node of a primitive
  int builtin = -1;          // builtin id of a builtin primitive
  // A traversal primitive: 1 forward, 2 reverse; its body argument and the
  // names and types it binds there (element, index, accumulator).
  int direction = 0;
  int region_argument = -1;
  std::vector<Binding> region_bindings;
  // A named control (assign, for_range): its target grammar local, or the
  // evaluator input it writes (input_name).
  int local = -1;
  int input = -1;
};
struct CNt {
  std::string id;
  std::uint8_t type = TAG_INVALID;
  Cat cat = Cat::Expression;
  bool variation = true;
  std::vector<Binding> scope;
  std::vector<std::uint32_t> productions;
  std::vector<std::uint32_t> minimum_by_depth;
};
struct CProd { std::uint32_t nt = 0; double weight = 0; std::uint32_t expression = 0; std::string stable_id; };
struct CDomain {
  std::uint8_t type = TAG_INVALID;
  std::vector<Val> values;
  bool integer_range = false;
  std::int64_t minimum = 0, maximum = 0;
};

struct Grammar {
  std::uint32_t max_nodes = 0, max_depth = 0, fuel = 0;
  std::vector<Binding> inputs, locals;
  std::vector<CNt> nts;
  std::vector<CProd> prods;
  std::vector<CExpr> exprs;
  std::vector<CDomain> domains;
  std::uint32_t entry = 0;
  std::string content_hash;  // population.json grammar_hash (identity.hpp)
};

Grammar load_grammar_file(const std::string& path);
Grammar load_grammar_text(const std::string& text);
std::uint8_t parse_type_name(const std::string& name);
const char* type_name(std::uint8_t tag);

// ---- device phase tables (POD, uploaded once per search) ----
enum PKind : int { PK_OPERATION = 0, PK_CONSTANT = 1, PK_BOUND = 2, PK_REFERENCE = 3, PK_INPUT = 4, PK_ALIAS = 5 };
enum TKind : int {
  TK_CONSTANT = 0, TK_LOCAL, TK_OPERATION, TK_IF, TK_AND, TK_OR, TK_PROGRAM, TK_BLOCK, TK_NIL,
  TK_ASSIGN, TK_RETURN, TK_FOR_RANGE, TK_VAR, TK_LET, TK_BOUND, TK_REGION, TK_TRAVERSE
};
constexpr int kMaxArity = 8;
// Completion type of a statement node (never a runtime value tag).
constexpr int kStatementType = 0xF0;
constexpr int kFixedSite = 1 << 30;
constexpr unsigned kNoId = 0xffffffffu;

// Node template of a phase expression; a tree node is its expression id
// (origin) plus, for a constant, the constant value.
struct PNode {
  int kind = TK_CONSTANT;
  int opcode = OP_PUSH_CONST;
  int operand = 0;
  int arity = 0;
  unsigned fuel = 1, merge_fuel = 1;
  int type = TAG_INVALID;
};
struct PExpr {
  int kind = PK_OPERATION;
  PNode node;
  int native_kind = -1;
  int target = -1;
  int children[kMaxArity] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int scope_count = 0;
  signed char child_scope[kMaxArity] = {-1, -1, -1, -1, -1, -1, -1, -1};
};
struct PProd { int expression; double weight; };
struct PNt { int production_offset = 0, production_count = 0, scope_count = 0, normalized = 0, fixed = 0; };
struct PDomain { int value_offset = 0, value_count = 0, integer_range = 0; std::int64_t minimum = 0, maximum = 0; };

struct PhaseGrammar {
  std::vector<PExpr> exprs;
  std::vector<PProd> prods;
  std::vector<PNt> nts;
  std::vector<PDomain> domains;
  std::vector<Val> values;
  std::vector<unsigned> minimum;
  int root = -1, depth_count = 0, logical_count = 1;
  bool site_depth = false;
  bool statements = false, lets = false;
  bool input_writes = false;  // a control writes an evaluator input
  bool has_let = false;       // let bindings (lexical binders with one value slot)
  // Traversal nodes in every program of the entry nonterminal (the grammar
  // admits traversals only where that count is the same for all programs).
  int traversals = 0;
  // Lexical binder numbers of every program (3 per traversal, 1 per let),
  // the same for all programs of the grammar.
  int binder_ids = 0;
};

// compile_phase_grammar + variation flags + static normalization.
PhaseGrammar compile_phase(const Grammar& g, unsigned root_nt, unsigned max_depth, unsigned site_depth,
                           const std::vector<std::string>& input_order);

// Whole-program layout of an expression-entry grammar: the program is
// PROGRAM(BLOCK_CONS(RETURN(<phase>), BLOCK_NIL)) and the phase is the entry
// nonterminal's derivation (cpp/src/evolution/grammar/generate.cpp:43-59,
// cpp/src/evolution/repro/gpu/phase_layout.hpp:120-140).
struct Layout {
  unsigned fixed_nodes = 0;
  unsigned slot_max_depth = 0;
  unsigned max_variable_nodes = 0;
  unsigned entry_nt = 0;
  bool program_terminator = true;
};
Layout whole_program_layout(const Grammar& g);

// An upper bound on the hidden locals (slots after the evaluator inputs) a
// program of the entry nonterminal keeps live at once: every declared
// grammar local, plus the deepest nesting of let value slots, for_range bound and counter
// temporaries and traversal temporaries within the depth budget.
unsigned hidden_local_bound(const Grammar& g);

}  // namespace gagp_amd
