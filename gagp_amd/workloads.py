"""Seeded generators of the benchmark search workloads.

Each workload is a directory with a ``gagp-task-v1`` task, a
``grammar-definition-v2`` grammar and ``fitness-cases`` train/test files, the
same documents ``gagp run`` reads (reference formats: spec/task.md,
spec/grammar_definition.md, spec/fitness_cases.md). Case data come from a
SplitMix64 stream seeded with the workload id, so every file is reproducible.

Workloads:
  poly_c1   y = x^3 - 2x + 1, train x = -10..10 (21), test 40 x ~ U{-100..100};
            grammar x, constants {1,2}, add/sub/mul; 31 nodes, depth 8.
  quintic   seeded quintic, coefficients ~ U{-5..5}; 256 train / 512 test,
            x ~ U{-50..50}; add/sub/mul, constants 0..9; 64 nodes.
  mixed     ``mixed_searches(n)``: polynomials of degree 2..5 with
            log-uniform population and case counts (the batched workload);
            with ``lists=True`` every second search is a list reduction.
  lists     ``list_task(kind)``: a reduction of an Int list (sum, max, count
            of elements > k, sum of the even elements), lengths 0..64 and
            values ~ U{-1000..1000}; grammar: one traversal of the list whose
            body nonterminal combines element, index and accumulator with
            add/sub/max/mod, comparisons and if.
"""
from __future__ import annotations

import json
import math
import os

MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Vigna) with the rejection-sampled bounded draw."""

    def __init__(self, seed: int):
        self.s = seed & MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def bounded(self, n: int) -> int:
        threshold = ((1 << 64) - n) % n
        while True:
            v = self.next()
            if v >= threshold:
                return v % n

    def uniform_int(self, lo: int, hi: int) -> int:
        return lo + self.bounded(hi - lo + 1)

    def unit(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53


def _int_case(x: int, y: int) -> dict:
    return {"inputs": {"x": {"type": "int", "value": x}}, "expected": {"type": "int", "value": y}}


def _write(path: str, doc) -> None:
    with open(path, "w") as f:
        json.dump(doc, f, separators=(",", ":"))
        f.write("\n")


def polynomial_grammar(constants, max_nodes: int, max_depth: int, fuel: int = 1000) -> dict:
    return {
        "format_version": "grammar-definition-v2",
        "entry": {"nonterminal": "Expr", "type": "Int"},
        "inputs": [{"name": "x", "type": "Int"}],
        "search_limits": {"max_nodes": max_nodes, "max_depth": max_depth},
        "execution_limits": {"fuel": fuel},
        "nonterminals": [{
            "id": "Expr", "type": "Int", "scope": [],
            "alternatives": [
                {"id": "x", "weight": 3, "expression": {"input": "x"}},
                {"id": "small", "weight": 2,
                 "expression": {"constant": {"type": "Int", "values": [str(c) for c in constants]}}},
                {"id": "add", "weight": 1, "expression": {"signature": "add(Int,Int)->Int",
                                                          "args": [{"ref": "Expr"}, {"ref": "Expr"}]}},
                {"id": "sub", "weight": 1, "expression": {"signature": "sub(Int,Int)->Int",
                                                          "args": [{"ref": "Expr"}, {"ref": "Expr"}]}},
                {"id": "mul", "weight": 1, "expression": {"signature": "mul(Int,Int)->Int",
                                                          "args": [{"ref": "Expr"}, {"ref": "Expr"}]}},
            ]}],
    }


def write_task(directory: str, name: str, grammar: dict, train, test, case_cap: float = 1000) -> str:
    os.makedirs(directory, exist_ok=True)
    _write(os.path.join(directory, "grammar.json"), grammar)
    _write(os.path.join(directory, "train.cases.json"), {"format_version": "fitness-cases", "cases": train})
    _write(os.path.join(directory, "test.cases.json"), {"format_version": "fitness-cases", "cases": test})
    task = {"format": "gagp-task-v1", "name": name, "grammar": "grammar.json", "train": "train.cases.json",
            "test": "test.cases.json", "fitness": {"metric": "absolute", "case_cap": case_cap}}
    path = os.path.join(directory, "task.json")
    _write(path, task)
    return path


def poly_c1(directory: str) -> str:
    rng = SplitMix64(1)
    f = lambda x: x ** 3 - 2 * x + 1
    train = [_int_case(x, f(x)) for x in range(-10, 11)]
    test = []
    for _ in range(40):
        x = rng.uniform_int(-100, 100)
        test.append(_int_case(x, f(x)))
    return write_task(directory, "poly_c1", polynomial_grammar([1, 2], 31, 8), train, test)


def _poly_cases(rng: SplitMix64, coeffs, n: int, lo: int, hi: int):
    out = []
    for _ in range(n):
        x = rng.uniform_int(lo, hi)
        out.append(_int_case(x, sum(c * x ** k for k, c in enumerate(coeffs))))
    return out


def quintic(directory: str, n_train: int = 256, n_test: int = 512, seed: int = 2) -> str:
    rng = SplitMix64(seed)
    coeffs = [rng.uniform_int(-5, 5) for _ in range(6)]
    if coeffs[5] == 0:
        coeffs[5] = 1  # keep the degree five
    train = _poly_cases(rng, coeffs, n_train, -50, 50)
    test = _poly_cases(rng, coeffs, n_test, -50, 50)
    return write_task(directory, "quintic", polynomial_grammar(list(range(10)), 64, 12), train, test)


def _const(values) -> dict:
    return {"constant": {"type": "Int", "values": [str(v) for v in values]}}


def _op(signature: str, *args) -> dict:
    return {"signature": signature, "args": list(args)}


def list_grammar(constants, max_nodes: int = 40, max_depth: int = 11, fuel: int = 4096, with_mod: bool = True,
                 fold: str = "traverse", nested: bool = False) -> dict:
    """One traversal of the list ``xs`` (``fold``: traverse,
    traverse_reverse, traverse_range or traverse_range_reverse; the ranged
    ones take begin ~ {0, 1, -1} and end ~ {64, 3, -2}); its body
    ``List.Step`` sees the element, its index and the accumulator. With
    ``nested`` the initial accumulator is itself a forward traversal
    (``List.Inner``), so every program has two."""
    step, cond = {"ref": "List.Step"}, {"ref": "List.Cond"}
    alternatives = [
        ("add", _op("add(Int,Int)->Int", step, step)),
        ("sub", _op("sub(Int,Int)->Int", step, step)),
        ("max", _op("max(Int,Int)->Int", step, step)),
        ("if", _op("if(Bool,Int,Int)->Int", cond, step, step)),
        ("element", {"bound": "element"}),
        ("index", {"bound": "index"}),
        ("acc", {"bound": "acc"}),
        ("const", _const(constants)),
    ]
    if with_mod:
        alternatives.append(("mod", _op("mod(Int,Int)->Int", step, step)))
    scope = [{"name": "element", "type": "Int"}, {"name": "index", "type": "Int"}, {"name": "acc", "type": "Int"}]
    ranged = fold.startswith("traverse_range")
    bounds = [_const([0, 1, -1]), _const([64, 3, -2])] if ranged else []
    body = "5" if ranged else "3"
    init = ([{"id": "inner", "weight": 1, "expression": {
                "signature": "traverse(IntList,Int,Int,Int)->Int",
                "args": [{"input": "xs"}, _const([0]), _const([0]), {"ref": "List.Inner"}],
                "bind": {"3": ["e2", "i2", "a2"]}}}] if nested else
            [{"id": "const", "weight": 1, "expression": _const(constants)}])
    extra = []
    if nested:
        inner = [{"name": "e2", "type": "Int"}, {"name": "i2", "type": "Int"}, {"name": "a2", "type": "Int"}]
        extra = [{"id": "List.Inner", "type": "Int", "scope": inner, "alternatives": [
            {"id": "add", "weight": 1, "expression": _op("add(Int,Int)->Int", {"bound": "a2"}, {"bound": "e2"})},
            {"id": "max", "weight": 1, "expression": _op("max(Int,Int)->Int", {"bound": "a2"}, {"bound": "e2"})},
            {"id": "index", "weight": 1, "expression": {"bound": "i2"}}]}]
    return {
        "format_version": "grammar-definition-v2",
        "entry": {"nonterminal": "List.Reduce", "type": "Int"},
        "inputs": [{"name": "xs", "type": "IntList"}],
        "search_limits": {"max_nodes": max_nodes, "max_depth": max_depth},
        "execution_limits": {"fuel": fuel},
        "nonterminals": [
            {"id": "List.Reduce", "type": "Int", "scope": [], "alternatives": [
                {"id": "fold", "weight": 1, "expression": {
                    "signature": "%s(IntList,Int,%sInt,Int)->Int" % (fold, "Int,Int," if ranged else ""),
                    "args": [{"input": "xs"}, _const([0])] + bounds + [{"ref": "List.Init"}, step],
                    "bind": {body: ["element", "index", "acc"]}}}]},
            {"id": "List.Init", "type": "Int", "scope": [], "alternatives": init},
            {"id": "List.Step", "type": "Int", "scope": scope, "alternatives": [
                {"id": name, "weight": 1, "expression": e} for name, e in alternatives]},
            {"id": "List.Cond", "type": "Bool", "scope": scope, "alternatives": [
                {"id": "gt", "weight": 1, "expression": _op("gt(Int,Int)->Bool", step, step)},
                {"id": "eq", "weight": 1, "expression": _op("eq(Int,Int)->Bool", step, step)}]},
        ] + extra,
    }


def statement_grammar(constants, max_nodes: int = 60, max_depth: int = 12, fuel: int = 4096, with_temp: bool = False,
                      with_traverse: bool = False) -> dict:
    """A statement program over the list ``xs``: it sets the accumulator
    ``acc``, then runs statements (assignments to ``acc``, a ``for_range``
    loop over ``i``, ``if_stmt``) and ends with a ``return``. Reads of ``i``
    outside a loop body fail definite assignment and are never generated.
    ``with_temp`` adds a second local ``t``; ``with_traverse`` makes the
    first assignment a forward traversal of ``xs`` (exactly one per
    program)."""
    i_ = {"ref": "St.Int"}
    b_ = {"ref": "St.Bool"}
    body = {"ref": "St.Body"}
    xs = {"input": "xs"}
    local = lambda name: {"local": name}  # noqa: E731
    assign = lambda name, arg: {"control": "assign(Int)->Statement", "type": "Int", "name": name, "args": [arg]}  # noqa: E731
    cons = lambda a, b: {"control": "block_cons(Statement,Block)->Block", "type": "Int", "args": [a, b]}  # noqa: E731
    nil = {"control": "block_nil()->Block", "type": "Int", "args": []}
    ints = [
        ("acc", local("acc")),
        ("i", local("i")),
        ("elem", _op("index(IntList,Int)->Int", xs, local("i"))),
        ("len", _op("len(IntList)->Int", xs)),
        ("const", _const(constants)),
        ("add", _op("add(Int,Int)->Int", i_, i_)),
        ("sub", _op("sub(Int,Int)->Int", i_, i_)),
        ("mod", _op("mod(Int,Int)->Int", i_, i_)),
    ]
    stmts = [
        ("acc", 3, assign("acc", i_)),
        ("for", 2, {"control": "for_range(Int,Block)->Statement", "type": "Int", "name": "i",
                    "args": [{"ref": "St.Bound"}, body]}),
        ("if", 2, {"control": "if_stmt(Bool,Block,Block)->Statement", "type": "Int", "args": [b_, body, body]}),
    ]
    locals_ = [{"name": "acc", "type": "Int"}, {"name": "i", "type": "Int"}]
    if with_temp:
        ints.append(("t", local("t")))
        stmts.append(("t", 1, assign("t", i_)))
        locals_.append({"name": "t", "type": "Int"})
    first = i_
    extra = []
    if with_traverse:
        first = {"signature": "traverse(IntList,Int,Int,Int)->Int",
                 "args": [xs, _const([0]), _const([0]), {"ref": "St.Fold"}], "bind": {"3": ["e", "k", "a"]}}
        scope = [{"name": "e", "type": "Int"}, {"name": "k", "type": "Int"}, {"name": "a", "type": "Int"}]
        fold = {"ref": "St.Fold"}
        extra = [{"id": "St.Fold", "type": "Int", "scope": scope, "alternatives": [
            {"id": "add", "weight": 2, "expression": _op("add(Int,Int)->Int", fold, fold)},
            {"id": "mod", "weight": 1, "expression": _op("mod(Int,Int)->Int", fold, fold)},
            {"id": "e", "weight": 2, "expression": {"bound": "e"}},
            {"id": "k", "weight": 1, "expression": {"bound": "k"}},
            {"id": "a", "weight": 2, "expression": {"bound": "a"}},
            {"id": "const", "weight": 1, "expression": _const(constants)}]}]
    return {
        "format_version": "grammar-definition-v2",
        "entry": {"category": "Program", "nonterminal": "St.Program", "type": "Int"},
        "inputs": [{"name": "xs", "type": "IntList"}],
        "locals": locals_,
        "search_limits": {"max_nodes": max_nodes, "max_depth": max_depth},
        "execution_limits": {"fuel": fuel},
        "nonterminals": [
            {"id": "St.Program", "category": "Program", "type": "Int", "scope": [], "variation": False, "alternatives": [
                {"id": "program", "weight": 1, "expression": {
                    "control": "program(Block)->Program", "type": "Int",
                    "args": [cons(assign("acc", first), {"ref": "St.Tail"})]}}]},
            {"id": "St.Tail", "category": "Block", "type": "Int", "scope": [], "alternatives": [
                {"id": "more", "weight": 3, "expression": cons({"ref": "St.Stmt"}, {"ref": "St.Tail"})},
                {"id": "final", "weight": 1, "expression": cons(
                    {"control": "return(Int)->Statement", "type": "Int", "args": [i_]}, nil)}]},
            {"id": "St.Body", "category": "Block", "type": "Int", "scope": [], "alternatives": [
                {"id": "nil", "weight": 2, "expression": nil},
                {"id": "cons", "weight": 2, "expression": cons({"ref": "St.Stmt"}, body)}]},
            {"id": "St.Stmt", "category": "Statement", "type": "Int", "scope": [], "alternatives": [
                {"id": name, "weight": w, "expression": e} for name, w, e in stmts]},
            {"id": "St.Bound", "type": "Int", "scope": [], "alternatives": [
                {"id": "len", "weight": 3, "expression": _op("len(IntList)->Int", xs)},
                {"id": "const", "weight": 1, "expression": _const([0, 2, 3])}]},
            {"id": "St.Int", "type": "Int", "scope": [], "alternatives": [
                {"id": name, "weight": 1, "expression": e} for name, e in ints]},
            {"id": "St.Bool", "type": "Bool", "scope": [], "alternatives": [
                {"id": "eq", "weight": 2, "expression": _op("eq(Int,Int)->Bool", i_, i_)},
                {"id": "lt", "weight": 1, "expression": _op("lt(Int,Int)->Bool", i_, i_)}]},
        ] + extra,
    }


def input_target_grammar(max_nodes: int = 48, max_depth: int = 10, fuel: int = 600) -> dict:
    """Statement programs over an Int input ``x`` that also write ``x``: an
    ``assign`` and a ``for_range`` with ``input_name`` targets besides the
    locals ``acc`` and ``i``."""
    i_, b_, body = {"ref": "In.Int"}, {"ref": "In.Bool"}, {"ref": "In.Body"}
    cons = lambda a, b: {"control": "block_cons(Statement,Block)->Block", "type": "Int", "args": [a, b]}  # noqa: E731
    nil = {"control": "block_nil()->Block", "type": "Int", "args": []}
    stmts = [
        ("acc", 3, {"control": "assign(Int)->Statement", "type": "Int", "name": "acc", "args": [i_]}),
        ("x", 1, {"control": "assign(Int)->Statement", "type": "Int", "input_name": "x", "args": [i_]}),
        ("for", 1, {"control": "for_range(Int,Block)->Statement", "type": "Int", "name": "i", "args": [_const([0, 1, 2, 3]), body]}),
        ("for_x", 1, {"control": "for_range(Int,Block)->Statement", "type": "Int", "input_name": "x",
                      "args": [_const([0, 1, 2, 3]), body]}),
        ("if", 1, {"control": "if_stmt(Bool,Block,Block)->Statement", "type": "Int", "args": [b_, body, body]}),
    ]
    ints = [("x", 3, {"input": "x"}), ("acc", 2, {"local": "acc"}), ("i", 1, {"local": "i"}), ("const", 2, _const([0, 1, 2])),
            ("add", 1, _op("add(Int,Int)->Int", i_, i_)), ("sub", 1, _op("sub(Int,Int)->Int", i_, i_)),
            ("mul", 1, _op("mul(Int,Int)->Int", i_, i_))]
    return {
        "format_version": "grammar-definition-v2",
        "entry": {"category": "Program", "nonterminal": "In.Program", "type": "Int"},
        "inputs": [{"name": "x", "type": "Int"}],
        "locals": [{"name": "acc", "type": "Int"}, {"name": "i", "type": "Int"}],
        "search_limits": {"max_nodes": max_nodes, "max_depth": max_depth},
        "execution_limits": {"fuel": fuel},
        "nonterminals": [
            {"id": "In.Program", "category": "Program", "type": "Int", "scope": [], "variation": False, "alternatives": [
                {"id": "program", "weight": 1, "expression": {
                    "control": "program(Block)->Program", "type": "Int",
                    "args": [cons({"control": "assign(Int)->Statement", "type": "Int", "name": "acc", "args": [i_]},
                                  {"ref": "In.Tail"})]}}]},
            {"id": "In.Tail", "category": "Block", "type": "Int", "scope": [], "alternatives": [
                {"id": "more", "weight": 3, "expression": cons({"ref": "In.Stmt"}, {"ref": "In.Tail"})},
                {"id": "final", "weight": 1, "expression": cons(
                    {"control": "return(Int)->Statement", "type": "Int", "args": [i_]}, nil)}]},
            {"id": "In.Body", "category": "Block", "type": "Int", "scope": [], "alternatives": [
                {"id": "nil", "weight": 2, "expression": nil},
                {"id": "cons", "weight": 2, "expression": cons({"ref": "In.Stmt"}, body)}]},
            {"id": "In.Stmt", "category": "Statement", "type": "Int", "scope": [], "alternatives": [
                {"id": name, "weight": w, "expression": e} for name, w, e in stmts]},
            {"id": "In.Int", "type": "Int", "scope": [], "alternatives": [
                {"id": name, "weight": w, "expression": e} for name, w, e in ints]},
            {"id": "In.Bool", "type": "Bool", "scope": [], "alternatives": [
                {"id": "lt", "weight": 1, "expression": _op("lt(Int,Int)->Bool", i_, i_)},
                {"id": "eq", "weight": 1, "expression": _op("eq(Int,Int)->Bool", i_, i_)}]},
        ],
    }


def input_target_task(directory: str, seed: int = 5) -> str:
    """y = x^2 - 2x + 3 over x ~ U{-20..20} (32 train / 40 test cases),
    searched with input_target_grammar."""
    rng = SplitMix64(seed)
    f = lambda x: x * x - 2 * x + 3  # noqa: E731
    train = [_int_case(x, f(x)) for x in (rng.uniform_int(-20, 20) for _ in range(32))]
    test = [_int_case(x, f(x)) for x in (rng.uniform_int(-20, 20) for _ in range(40))]
    return write_task(directory, "stmt_input", input_target_grammar(), train, test)


LIST_KINDS = ("sum", "max", "count_gt", "sum_even")


def _list_cases(rng: SplitMix64, kind: str, k: int, n: int, max_len: int):
    cases = []
    for _ in range(n):
        xs = [rng.uniform_int(-1000, 1000) for _ in range(rng.uniform_int(0, max_len))]
        if kind == "sum":
            y = sum(xs)
        elif kind == "max":
            y = max(xs) if xs else -1001
        elif kind == "count_gt":
            y = sum(1 for x in xs if x > k)
        else:
            y = sum(x for x in xs if x % 2 == 0)
        cases.append({"inputs": {"xs": {"type": "int_list", "value": xs}}, "expected": {"type": "int", "value": y}})
    return cases


def list_task(directory: str, kind: str, seed: int, n_train: int = 200, n_test: int = 400, max_len: int = 64,
              fold: str = "traverse", nested: bool = False, fuel: int = 4096, **statement) -> str:
    """A list reduction (``LIST_KINDS``); count_gt draws k ~ U{-100..100},
    the max of an empty list is -1001. ``fold`` "statement" writes a
    statement program grammar (``statement`` holds its options)."""
    rng = SplitMix64(seed)
    k = rng.uniform_int(-100, 100)
    train = _list_cases(rng, kind, k, n_train, max_len)
    test = _list_cases(rng, kind, k, n_test, max_len)
    constants = {"sum": [0, 1, 2], "max": [0, 1, -1001], "count_gt": [0, 1, k], "sum_even": [0, 1, 2]}[kind]
    if fold == "statement":
        grammar = statement_grammar(constants, fuel=fuel, **statement)
    else:
        grammar = list_grammar(constants, fuel=fuel, fold=fold, nested=nested)
    return write_task(directory, "list_%s" % kind, grammar, train, test)


def sum_even(directory: str, seed: int = 3, fold: str = "traverse") -> str:
    """The loop-heavy search: sum of the even elements of an Int list,
    200 train / 400 test cases, lengths 0..50, values ~ U{-1000..1000};
    ``fold`` "statement" searches statement programs (a for_range loop over
    the list, an accumulator local, if_stmt) instead of a traversal."""
    return list_task(directory, "sum_even", seed, 200, 400, 50, fold=fold)


def log_uniform_int(rng: SplitMix64, lo: int, hi: int) -> int:
    v = math.exp(math.log(lo) + rng.unit() * (math.log(hi) - math.log(lo)))
    return max(lo, min(hi, int(round(v))))


def mixed_searches(directory: str, n: int = 64, seed: int = 4, p_range=(256, 4096), c_range=(16, 1024), lists: bool = False):
    """Independent polynomial searches of degree 2..5 with log-uniform
    population and train-case counts; with ``lists`` every odd-numbered one
    is a list reduction (sum, max or count > k; lengths 0..64) instead.
    Returns [(task_path, population)]."""
    rng = SplitMix64(seed)
    out = []
    for i in range(n):
        if lists and i % 2 == 1:
            kind = LIST_KINDS[rng.bounded(3)]
            population = log_uniform_int(rng, *p_range)
            cases = log_uniform_int(rng, *c_range)
            sub = SplitMix64(rng.next())
            k = sub.uniform_int(-100, 100)
            train = _list_cases(sub, kind, k, cases, 64)
            test = _list_cases(sub, kind, k, 64, 64)
            constants = {"sum": [0, 1, 2], "max": [0, 1, -1001], "count_gt": [0, 1, k]}[kind]
            task = write_task(os.path.join(directory, "s%03d" % i), "mixed_%03d_%s" % (i, kind), list_grammar(constants),
                              train, test)
            out.append((task, population))
            continue
        degree = 2 + rng.bounded(4)
        coeffs = [rng.uniform_int(-5, 5) for _ in range(degree + 1)]
        if coeffs[-1] == 0:
            coeffs[-1] = 1
        population = log_uniform_int(rng, *p_range)
        cases = log_uniform_int(rng, *c_range)
        train = _poly_cases(rng, coeffs, cases, -50, 50)
        test = _poly_cases(rng, coeffs, 64, -50, 50)
        task = write_task(os.path.join(directory, "s%03d" % i), "mixed_%03d" % i,
                          polynomial_grammar(list(range(10)), 64, 12), train, test)
        out.append((task, population))
    return out


def served_requests(directory: str, n: int = 60, seed: int = 5, rate: float = 20.0, p_range=(256, 16384),
                    c_range=(16, 2048), g_range=(32, 256), lists: bool = False):
    """A Poisson request stream over mixed polynomial searches: arrival
    offsets with exponential gaps of mean 1/rate seconds, generation caps
    uniform in g_range. Returns [(task_path, population, generations,
    seed, arrival_s)]."""
    tasks = mixed_searches(directory, n=n, seed=seed, p_range=p_range, c_range=c_range, lists=lists)
    rng = SplitMix64(seed * 7919 + 1)
    t, out = 0.0, []
    for i, (task, population) in enumerate(tasks):
        out.append((task, population, rng.uniform_int(*g_range), 100 + i, t))
        t += -math.log(1.0 - rng.unit()) / rate
    return out

