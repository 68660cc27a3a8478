"""Non-integral fitness: the oracle's and the device's case scores and
canonical reduction against the reference's own eval_fitness_cpu.

tests/golden/fitness/reduction.json holds the reference's outputs
(oracle/ref_fitness.cpp links the reference CPU engine's objects and calls
gagp::eval_fitness_cpu, cpp/src/runtime/cpu/fitness_cpu.cpp:57-176) for
seeded scenarios: discounted proximity, fractional caps and ceilings, the
exact metric, 1 to 4097 cases and reduction widths 1, 32, 100 and 512 --
every sum goes through the 48-bit canonical order, and many totals are not
integers. The CPU tests pin the oracle with them; the GPU test holds the
device session (gagp_session_eval) to the same bits."""
import json
import os

import pytest

from _util import GOLDEN, oracle_fitness

DOC = json.load(open(os.path.join(GOLDEN, "fitness", "reduction.json")))
SCENARIOS = DOC["scenarios"]


def programs_of(s):
    return [{"code": [(op, a, None) for op, a in p["code"]], "consts": p["consts"], "n_locals": p["n_locals"]}
            for p in s["programs"]]


def inputs_of(s):
    return [[x] for x in s["inputs"]] if s["inputs"] else [[] for _ in s["answers"]]


def expected(s):
    return [float.fromhex(x) for x in s["fitness"]], s["wrong"], [tuple(c) for c in s["counts"]]


def test_vectors_exercise_the_non_integral_order():
    # Most totals are not integers, several scenarios put more than one case
    # on a reduction lane, and the widths differ from 512.
    fractional = sum(1 for s in SCENARIOS for f in expected(s)[0] if f != int(f))
    assert fractional >= 100
    assert any(s["cases"] > s["lanes"] for s in SCENARIOS)
    assert {s["lanes"] for s in SCENARIOS} >= {1, 32, 100, 512}
    assert {(s["metric"], s["proximity"]) for s in SCENARIOS} == {(1, 0), (1, 1), (2, 0)}


def test_discounted_known_answer_of_the_reference():
    # cpp/tests/runtime/test_fitness_contract.cpp:277-280: a program
    # returning 3 against 0 scores -3/4; a runtime error (a constant index
    # out of range) scores the discounted ceiling -1. One case, fuel 100,
    # penalty 5, one reduction lane.
    returns3 = {"code": [(0, 0, None), (20, None, None)], "consts": [3], "n_locals": 0}
    bad = {"code": [(0, 99, None)], "consts": [], "n_locals": 0}
    fit, wrong, _ = oracle_fitness([returns3, bad], [[]], [0], fuel=100, penalty=5.0, lanes=1, metric=1, proximity=1)
    assert fit == [-0.75, -1.0] and wrong == [1, 1]


@pytest.mark.parametrize("s", SCENARIOS, ids=lambda s: s["name"])
def test_oracle_matches_reference_reduction(s):
    fit, wrong, counts = oracle_fitness(programs_of(s), inputs_of(s), s["answers"], fuel=s["fuel"], penalty=s["penalty"],
                                        lanes=s["lanes"], metric=s["metric"], proximity=s["proximity"])
    want_fit, want_wrong, want_counts = expected(s)
    assert [x.hex() for x in fit] == [x.hex() for x in want_fit]
    assert wrong == want_wrong
    assert counts == want_counts


@pytest.mark.gpu
@pytest.mark.parametrize("s", SCENARIOS, ids=lambda s: s["name"])
def test_device_session_matches_reference_reduction(s):
    import gagp_amd
    print("reduction", s["name"], flush=True)
    session = gagp_amd.Session(inputs_of(s), s["answers"], fuel=s["fuel"], lanes=s["lanes"], penalty=s["penalty"],
                               metric=s["metric"], proximity=s["proximity"])
    fit, wrong, counts = session.eval(programs_of(s))
    want_fit, want_wrong, want_counts = expected(s)
    assert [x.hex() for x in fit] == [x.hex() for x in want_fit]
    assert wrong == want_wrong
    assert counts == want_counts
