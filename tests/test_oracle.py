"""Pins the oracle (oracle/vm_oracle.c) before the parity tests of the device
engine rely on it: reference bytecode fixtures, reference RNG known answers,
the C++ standard's mt19937_64 value, closed-form fitness totals and
hand-derived edge cases."""
import ctypes
import math
import struct

import pytest

from _util import E_NAMES, case_inputs, load_fixture_scenarios, oracle, oracle_fitness, oracle_run

SCENARIOS, SKIPPED = load_fixture_scenarios()


def test_fixture_coverage():
    # The admitted value model covers most reference scenarios; the rest use
    # string/char payloads or list construction.
    assert len(SCENARIOS) >= 18
    assert sum(len(s["cases"]) for s in SCENARIOS) >= 35


@pytest.mark.parametrize("scenario", SCENARIOS, ids=lambda s: s["file"] + ":" + s["intent"][:40])
def test_oracle_matches_reference_fixture(scenario):
    p = scenario["program"]
    for case in scenario["cases"]:
        vals, arena = case_inputs(case, p["n_locals"])
        err, out, _ = oracle_run(p, vals, arena, fuel=scenario["fuel"])
        if "expected_error" in case:
            assert E_NAMES[err] == case["expected_error"], case
            continue
        assert err == -1, (case, E_NAMES[err])
        exp = case["expected"]
        if exp["type"] == "int":
            assert (out.tag, out.i) == (0, int(exp["value"]))
        elif exp["type"] == "bool":
            assert (out.tag, out.b) == (2, int(exp["value"]))
        else:
            got = struct.unpack("<d", struct.pack("<q", out.i))[0]
            assert out.tag == 1 and got == float(exp["value"])


def test_splitmix_known_answers():
    # GrammarRandom(0) (cpp/tests/evolution/test_grammar_compiled.cpp:94-96).
    out = (ctypes.c_uint64 * 3)()
    oracle().orc_splitmix_next(0, out, 3)
    assert list(out) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_splitmix_bounded_is_rejection_sampling():
    # bounded(n) skips draws below (2^64 - n) mod n; for n = 3 that is 1.
    raw = (ctypes.c_uint64 * 64)()
    oracle().orc_splitmix_next(7, raw, 64)
    expect, threshold = [], (2 ** 64 - 3) % 3
    for v in raw:
        if v >= threshold:
            expect.append(v % 3)
    got = (ctypes.c_uint64 * len(expect))()
    oracle().orc_splitmix_bounded(7, 3, got, len(expect))
    assert list(got) == expect


def test_mt19937_64_standard_value():
    # [rand.predef]: the 10000th output of a default-constructed mt19937_64.
    out = (ctypes.c_uint64 * 10000)()
    oracle().orc_mt19937_64(5489, out, 10000)
    assert out[9999] == 9981545732273789042


@pytest.mark.parametrize("bits", [32, 40, 48])
def test_canonical_rounding_closed_form(bits):
    L = oracle()
    for v in [1.0, -3.0, 12345.0, 2.0 ** 60, -(2.0 ** 70)]:
        assert L.orc_canon(v, bits) == v  # already representable
    quantum = 2.0 ** (1 - bits)  # spacing of the kept mantissa on [1, 2)
    assert L.orc_canon(1.0 + 3 * quantum / 4, bits) == 1.0 + quantum
    assert L.orc_canon(1.0 + quantum / 4, bits) == 1.0
    assert L.orc_canon(-0.0, bits) == 0.0 and math.copysign(1, L.orc_canon(-0.0, bits)) == 1
    # ties round away from zero
    assert L.orc_canon(0.5 + 2.0 ** -(bits + 1), bits) == 0.5 + 2.0 ** -bits
    assert L.orc_canon(-(0.5 + 2.0 ** -(bits + 1)), bits) == -(0.5 + 2.0 ** -bits)


def _const_program(value):
    return {"code": [(0, 0, None), (20, None, None)], "consts": [value], "n_locals": 1}


def test_fitness_closed_form_totals():
    # A constant program scored against y = x: sum |c - x| capped at the
    # penalty; wrong = cases not exactly equal.
    xs = list(range(-30, 31))
    inputs = [[x] for x in xs]
    for lanes in (1, 17, 32, 33, 64, 512):
        for c in (0, 5, 2000):
            fit, wrong, counts = oracle_fitness([_const_program(c)], inputs, xs, penalty=1000.0, lanes=lanes)
            assert fit[0] == -sum(min(abs(c - x), 1000) for x in xs)
            assert wrong[0] == sum(1 for x in xs if x != c)
            assert counts[0] == (len(xs), 0, 0)


def test_fitness_errors_and_timeouts_cost_the_penalty():
    xs = list(range(10))
    inputs = [[x] for x in xs]
    # x mod x: ZeroDivisionError on x == 0 only.
    mod = {"code": [(1, 0, None), (1, 0, None), (9, None, None), (20, None, None)], "consts": [], "n_locals": 1}
    fit, wrong, counts = oracle_fitness([mod], inputs, [0] * 10, penalty=64.0, lanes=512)
    assert fit[0] == -64.0 and wrong[0] == 1 and counts[0] == (10, 1, 0)
    # Fuel 2 cannot reach RETURN (three instructions): every case times out.
    fit, wrong, counts = oracle_fitness([mod], inputs, [0] * 10, fuel=2, penalty=63.0)
    assert fit[0] == -630.0 and wrong[0] == 10 and counts[0] == (10, 10, 10)


def test_int64_protected_division_edges():
    lo = -(2 ** 63)
    prog = lambda bid: {"code": [(1, 0, None), (1, 1, None), (19, bid, 2), (20, None, None)], "consts": [], "n_locals": 2}
    for bid, a, b, want in [(13, lo, -1, lo), (14, lo, -1, 0), (28, lo, -1, 0), (29, lo, -1, 0), (13, 7, 0, 0),
                            (14, -7, 2, 1), (13, -7, 2, -3), (28, -7, 2, -3), (29, -7, 2, -1)]:
        err, out, _ = oracle_run(prog(bid), [a, b])
        assert err == -1 and out.i == want, (bid, a, b)
