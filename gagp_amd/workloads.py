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
            log-uniform population and case counts (the batched workload).
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


def log_uniform_int(rng: SplitMix64, lo: int, hi: int) -> int:
    v = math.exp(math.log(lo) + rng.unit() * (math.log(hi) - math.log(lo)))
    return max(lo, min(hi, int(round(v))))


def mixed_searches(directory: str, n: int = 64, seed: int = 4, p_range=(256, 4096), c_range=(16, 1024)):
    """Independent polynomial searches of degree 2..5 with log-uniform
    population and train-case counts. Returns [(task_path, population)]."""
    rng = SplitMix64(seed)
    out = []
    for i in range(n):
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
                    c_range=(16, 2048), g_range=(32, 256)):
    """A Poisson request stream over mixed polynomial searches: arrival
    offsets with exponential gaps of mean 1/rate seconds, generation caps
    uniform in g_range. Returns [(task_path, population, generations,
    seed, arrival_s)]."""
    tasks = mixed_searches(directory, n=n, seed=seed, p_range=p_range, c_range=c_range)
    rng = SplitMix64(seed * 7919 + 1)
    t, out = 0.0, []
    for i, (task, population) in enumerate(tasks):
        out.append((task, population, rng.uniform_int(*g_range), 100 + i, t))
        t += -math.log(1.0 - rng.unit()) / rate
    return out

