"""Serving (continuous batching): requests arriving over time share the
device and each publishes exactly the reference engine's result.json for its
request; with early exit, the result of the reference run with as many
generations as the search completed when it first had a best member with no
wrong train case."""
import json
import os

import pytest

import gagp_amd
from _util import GOLDEN, strip_timing

pytestmark = pytest.mark.gpu
MANIFEST = json.load(open(os.path.join(GOLDEN, "runs", "manifest.json")))


def golden(name):
    return strip_timing(json.load(open(os.path.join(GOLDEN, "runs", name, "result.json"))))


def submit(server, name, arrival, early_exit=False, generations=None):
    m = MANIFEST[name]
    search = os.path.join(GOLDEN, m["search"]) if "search" in m else None
    return server.submit(os.path.join(GOLDEN, m["task"]), search, m["population"], generations or m["generations"],
                         m["seed"], early_exit, arrival)


def test_staggered_requests_match_reference():
    print("serve: staggered requests", flush=True)
    server = gagp_amd.Server(budget_bytes=8 << 30, poll_rounds=2)
    plain = ["polynomial_p512_g32_s7", "quintic_p512_g16_s3", "mixed_0", "mixed_3", "polynomial_gen_e1_p256_g12_s3",
             "polynomial_steady_b128_rs3_p512_g16_s4", "polynomial_p100_g10_s5", "list_count_gt_rev_p300_g8_s6",
             "list_range_nested_p256_g8_s7", "mixedl_1", "stmt_count_gt_p300_g8_s6", "stmt_lowfuel_p400_g8_s8"]
    tickets = {name: submit(server, name, 0.004 * i) for i, name in enumerate(plain)}
    # Early exit: the 64-generation requests stop where the reference's
    # shorter runs end.
    early = {"polynomial_p1024_g18_s0": submit(server, "polynomial_p1024_g64_s0", 0.002, early_exit=True),
             "poly_c1_p1024_g10_s0": submit(server, "poly_c1_p1024_g64_s0", 0.011, early_exit=True)}
    server.run()
    for name, t in list(tickets.items()) + list(early.items()):
        o = server.outcome(t)
        assert o["status"] == 0, (name, o)
        assert o["arrival"] <= o["admit"] <= o["done"], (name, o)
        assert strip_timing(server.result(t)) == golden(name), name
    assert server.outcome(early["polynomial_p1024_g18_s0"])["generations"] == 18
    assert server.outcome(early["poly_c1_p1024_g10_s0"])["generations"] == 10


def planned(names):
    b = gagp_amd.Batch()
    out = []
    for name in names:
        m = MANIFEST[name]
        search = os.path.join(GOLDEN, m["search"]) if "search" in m else None
        out.append(b.planned_bytes(b.add(os.path.join(GOLDEN, m["task"]), search, m["population"], m["generations"],
                                         m["seed"])))
    return out


@pytest.mark.parametrize("limit", ["footprint", "max_live"])
def test_admission_limits(limit):
    # Admission by planned footprint: a budget equal to the largest request's
    # footprint never lets the live searches' planned bytes exceed it, so the
    # largest search runs alone (the two small ones may share the device).
    # A limit of one live search serializes admission. Results are unchanged.
    print("serve: admission limit", limit, flush=True)
    names = ["mixed_1", "mixed_2", "polynomial_p100_g10_s5"]
    sizes = planned(names)
    budget = max(sizes)
    if limit == "footprint":
        server = gagp_amd.Server(budget_bytes=budget, poll_rounds=3)
    else:
        server = gagp_amd.Server(budget_bytes=8 << 30, poll_rounds=3, max_live=1)
    tickets = [submit(server, n, 0.0) for n in names]
    server.run()
    outcomes = [server.outcome(t) for t in tickets]
    for name, t, o in zip(names, tickets, outcomes):
        assert o["status"] == 0
        assert strip_timing(server.result(t)) == golden(name), name
    if limit == "footprint":
        # at every admission, the searches live at that moment fit the budget
        for o in outcomes:
            live = [s for s, x in zip(sizes, outcomes) if x["admit"] <= o["admit"] < x["done"]]
            assert sum(live) <= budget, (o, outcomes, sizes)
        big = outcomes[sizes.index(budget)]
        for o in outcomes:
            if o is not big:
                assert o["done"] <= big["admit"] or o["admit"] >= big["done"], (big, o)
    else:
        # admission waits for the previous search to retire
        for before, after in zip(outcomes, outcomes[1:]):
            assert after["admit"] >= before["done"]


def test_unsupported_request_fails_alone():
    print("serve: unsupported request", flush=True)
    server = gagp_amd.Server(poll_rounds=2)
    bad = server.submit(os.path.join(GOLDEN, "tasks", "fuel-profile", "task.json"), arrival=0.0)
    good = submit(server, "polynomial_p100_g10_s5", 0.0)
    server.run()
    assert server.outcome(bad)["status"] == gagp_amd.GAGP_ERR_UNSUPPORTED
    assert strip_timing(server.result(good)) == golden("polynomial_p100_g10_s5")


def hip_free_bytes():
    """Free device memory as the HIP runtime the engine links reports it
    (hipMemGetInfo of the same libamdhip64 instance)."""
    import ctypes
    gagp_amd.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_retired_searches_return_their_device_memory():
    # Searches admitted and retired one after another (two live at a time)
    # give every byte back: free device memory after the run equals free
    # memory before it, while the searches held far more between them.
    print("serve: device memory returned on retirement", flush=True)
    names = ["quintic_p512_g16_s3", "mixed_0", "mixed_3", "polynomial_p512_g32_s7", "list_sum_even_p512_g10_s5",
             "stmt_sum_even_p512_g10_s5", "mixedl_1", "polynomial_p1024_g18_s0"]
    warm = gagp_amd.Server(poll_rounds=2)  # runtime state of a first run (code objects, streams)
    submit(warm, "polynomial_p100_g10_s5", 0.0)
    warm.run()
    before = hip_free_bytes()
    server = gagp_amd.Server(budget_bytes=8 << 30, poll_rounds=2, max_live=2)
    tickets = [submit(server, n, 0.0) for n in names]
    server.run()
    after = hip_free_bytes()
    for name, t in zip(names, tickets):
        assert server.outcome(t)["status"] == 0, name
        assert strip_timing(server.result(t)) == golden(name), name
    held = sum(planned(names))
    assert held >= 32 << 20  # the searches held far more than the tolerance below
    assert before - after < 8 << 20, (before, after, held)
