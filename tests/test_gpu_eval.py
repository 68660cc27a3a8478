"""`gagp eval` on the device against the reference's own `gagp eval`
(tests/golden/eval.json): the verdict (cases, correct, wrong, runtime
errors, mean loss, solved, task, split) of a run result's best program, a
population's first program and a bare program must be identical, over
straight-line, traversal, let and statement programs, low-fuel tasks whose
cases time out, and exact and discounted fitness tasks."""
import json
import os
import subprocess

import pytest

import gagp_amd
from _util import GOLDEN

pytestmark = pytest.mark.gpu
MANIFEST = json.load(open(os.path.join(GOLDEN, "runs", "manifest.json")))
CASES = json.load(open(os.path.join(GOLDEN, "eval.json")))["cases"]


def gagp_eval(*args):
    return subprocess.run([gagp_amd.CLI_PATH, "eval"] + [str(a) for a in args], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("case", CASES, ids=["%s-%s-%s" % (c["run"], c["document"][:-5], c["split"]) for c in CASES])
def test_eval_matches_reference(case):
    print("eval", case["run"], case["document"], case["split"], flush=True)
    task = os.path.join(GOLDEN, MANIFEST[case["run"]]["task"])
    r = gagp_eval(task, os.path.join(GOLDEN, "runs", case["run"], case["document"]), "--split", case["split"])
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout) == case["verdict"]


def test_eval_bare_program_and_default_split(tmp_path):
    # A bare program document; without --split the test cases are scored.
    case = next(c for c in CASES if c["document"] == "result.json" and c["split"] == "test")
    result = json.load(open(os.path.join(GOLDEN, "runs", case["run"], "result.json")))
    bare = tmp_path / "program.json"
    bare.write_text(json.dumps(result["best"]["program"]))
    r = gagp_eval(os.path.join(GOLDEN, MANIFEST[case["run"]]["task"]), bare)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout) == case["verdict"]


def test_eval_rejects_with_exit_2(tmp_path):
    poly = os.path.join(GOLDEN, "tasks", "polynomial", "task.json")
    result = os.path.join(GOLDEN, "runs", "polynomial_p1024_g64_s0", "result.json")
    # usage errors
    for args in ([poly], [poly, result, "--split", "valid"], [poly, result, "--frobnicate"]):
        r = gagp_eval(*args)
        assert r.returncode == 2 and "gagp:" in r.stderr, args
    # programs the task's grammar does not derive: a constant outside the
    # polynomial grammar's {1, 2, 3}, and a quintic program deeper than its
    # depth budget (the reference rejects both in grammar membership)
    program = json.load(open(result))["best"]["program"]
    for c in program["consts"]:
        c["value"] = 7
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps(program))
    r = gagp_eval(poly, bad)
    assert r.returncode == 2 and "cannot be derived" in r.stderr, r.stderr
    r = gagp_eval(poly, os.path.join(GOLDEN, "runs", "quintic_p512_g16_s3", "result.json"))
    assert r.returncode == 2 and "depth budget exceeded" in r.stderr, r.stderr
