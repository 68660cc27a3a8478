"""Traversal lowering on the host (CPU): tests/host_lowering.cpp lowers seeded
random programs of the list tasks with the engine's lowering (the same
template the device kernels instantiate) and scores them with the C oracle
VM; the reference's own evaluation of the exported ASTs (`gagp eval --split
train`, kept in tests/golden/traversal/) must agree on fitness, wrong cases
and runtime errors, program by program. The tasks cover forward, reverse,
ranged (with clamping) and nested traversals, statement programs (grammar
locals, for_range, if_stmt, return, with and without a traversal), and
low-fuel tasks whose long lists time out, which pins the per-instruction fuel
schedule."""
import json
import os
import subprocess

import pytest

from _util import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "gagp_amd", "_build")
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
TASKS = sorted(f[:-5] for f in os.listdir(os.path.join(GOLDEN, "traversal")) if f.endswith(".json"))
STATEMENT_TASKS = sorted(f[:-5] for f in os.listdir(os.path.join(GOLDEN, "definite")) if f.endswith(".json"))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    objs = [os.path.join(BUILD, n) for n in ("grammar.o", "task.o", "export.o", "identity.o")]
    missing = [p for p in objs + [ORACLE] if not os.path.exists(p)]
    assert not missing, "build the engine and the oracle first (__graft_entry__.build()): missing %s" % missing
    exe = str(tmp_path_factory.mktemp("driver") / "host_lowering")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "gagp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host_lowering.cpp")] + objs +
                   [ORACLE, "-Wl,-rpath," + os.path.dirname(ORACLE), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("task", TASKS)
def test_lowering_matches_reference_evaluation(driver, task):
    golden = json.load(open(os.path.join(GOLDEN, "traversal", task + ".json")))
    out = subprocess.run([driver, os.path.join(GOLDEN, golden["task"]), str(golden["n"])], check=True, capture_output=True,
                         text=True, cwd=ROOT).stdout.splitlines()
    rows = [json.loads(line) for line in out]
    assert all("lowering_status" not in r for r in rows)
    assert [r["i"] for r in rows] == [p["i"] for p in golden["programs"]]
    for r, p in zip(rows, golden["programs"]):
        assert r["text"] == p["text"]
        # The reference reports the mean loss; the fitness is its negated total.
        assert r["fitness"] == pytest.approx(-p["mean_loss"] * p["cases"], rel=1e-12, abs=1e-9), r["text"]
        assert r["wrong"] == p["wrong"], r["text"]
        assert r["runtime_errors"] == p["runtime_errors"], r["text"]
    # The looping VM's single-step sequences: every tag on its pattern, no jump
    # into a sequence after its first instruction.
    assert all(r["tag_violations"] == 0 for r in rows)
    tags = [sum(r["tags"][k] for r in rows) for k in range(5)]  # fuse, range test/next, walk test/next
    assert tags[0] > 0
    assert (tags[1] > 0 and tags[1] == tags[2]) if task.startswith("stmt_sum") else True
    assert tags[3] == tags[4] and (tags[3] > 0) == (task not in ("stmt_sum_even", "stmt_lowfuel", "stmt_input", "let-sum"))
    if task in ("list_lowfuel", "list_range_nested", "stmt_lowfuel"):
        # The fuel schedule is only pinned where cases actually run out of fuel.
        assert sum(r["timeouts"] for r in rows) > 100


@pytest.mark.parametrize("task", STATEMENT_TASKS)
def test_definite_assignment_matches_reference_verifier(driver, task):
    # Random statement programs drawn once each (no retry): the engine's
    # definite-assignment check must accept exactly the ASTs the reference's
    # verifier accepts (it rejects a read of an unassigned local).
    golden = json.load(open(os.path.join(GOLDEN, "definite", task + ".json")))
    out = subprocess.run([driver, os.path.join(GOLDEN, golden["task"]), str(golden["n"]), "defs"], check=True,
                         capture_output=True, text=True, cwd=ROOT).stdout.splitlines()
    rows = [json.loads(line) for line in out]
    assert [r["i"] for r in rows] == [p["i"] for p in golden["programs"]]
    for r, p in zip(rows, golden["programs"]):
        assert r["text"] == p["text"]
        assert r["defined"] == p["accepted"], r["text"]
    # Both verdicts occur.
    assert 0 < sum(p["accepted"] for p in golden["programs"]) < len(rows)
