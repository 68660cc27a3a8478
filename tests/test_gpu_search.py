"""Whole-search parity on the device: `gagp run --engine gpu` against the
reference CPU engine's golden result.json (best program, fitness vector
entries, test verdict, effective search and the per-generation history must
be identical; only timing fields may differ)."""
import json
import os
import subprocess

import pytest

import gagp_amd
from _util import GOLDEN, strip_timing

pytestmark = pytest.mark.gpu
MANIFEST = json.load(open(os.path.join(GOLDEN, "runs", "manifest.json")))
# the full-size single searches
FULL = ("quintic_p4096_g128_s0", "list_sum_even_p8192_g64_s0", "stmt_sum_even_p8192_g64_s0")
SMALL = sorted(n for n in MANIFEST if n not in FULL)


def golden(name):
    return json.load(open(os.path.join(GOLDEN, "runs", name, "result.json")))


def add(batch, name):
    m = MANIFEST[name]
    search = os.path.join(GOLDEN, m["search"]) if "search" in m else None
    init = os.path.join(GOLDEN, m["init"]) if "init" in m else None
    return batch.add(os.path.join(GOLDEN, m["task"]), search, m["population"], m["generations"], m["seed"], init=init)


@pytest.mark.parametrize("name", SMALL)
def test_search_matches_reference(name):
    print("search", name, flush=True)
    b = gagp_amd.Batch()
    add(b, name)
    b.run()
    assert strip_timing(b.result(0)) == strip_timing(golden(name))


@pytest.mark.parametrize("name", FULL)
def test_full_size_search_matches_reference(name):
    print("search", name, flush=True)
    b = gagp_amd.Batch()
    add(b, name)
    b.run()
    assert strip_timing(b.result(0)) == strip_timing(golden(name))


def test_batched_searches_match_alone():
    # Every search in one batch (uneven populations, case counts and
    # generations) publishes exactly its stand-alone result.
    print("batch of", len(SMALL), flush=True)
    b = gagp_amd.Batch()
    index = {name: add(b, name) for name in SMALL}
    b.run()
    for name, i in index.items():
        assert strip_timing(b.result(i)) == strip_timing(golden(name)), name


def test_autotuned_launch_shapes_leave_results_unchanged():
    # The evaluation launch autotuner (on by default) times candidate shapes
    # by re-running the evaluation kernels over a round's pending programs;
    # with it off the fixed shapes run. Both publish every search's
    # reference result, and a second batch of the same shape reuses the
    # cached choice.
    names = ["mixedl_0", "mixedl_1", "mixed_2", "list_count_gt_rev_p300_g8_s6", "stmt_count_gt_p300_g8_s6", "quintic_p512_g16_s3"]
    print("autotune batch of", len(names), flush=True)
    results = {}
    for on in (0, 2, 2):
        b = gagp_amd.Batch()
        b.set_autotune(on)
        index = {name: add(b, name) for name in names}
        b.run()
        for name, i in index.items():
            assert strip_timing(b.result(i)) == strip_timing(golden(name)), (on, name)
        results.setdefault(on, []).append(b.tuning())
    assert results[0] == [{}]
    first, second = results[2]
    assert not first["cached"] and second["cached"]
    assert first["key"] == second["key"]
    assert first["fast_programs"] in (1, 2, 4) and first["loop_group"] in (1, 2, 4)
    assert set(first["typed_ms"]) <= {"1", "2", "4"} and "1" in first["typed_ms"]
    assert (second["fast_programs"], second["loop_group"]) == (first["fast_programs"], first["loop_group"])
    assert set(first["loop_ms"]) <= {"1", "2", "4"} and "1" in first["loop_ms"]


def test_graph_and_stream_execution_agree():
    name = "polynomial_p256_g20_s123"
    results = []
    for use_graph in (True, False):
        b = gagp_amd.Batch()
        add(b, name)
        b.run(use_graph=use_graph)
        results.append(strip_timing(b.result(0)))
    assert results[0] == results[1] == strip_timing(golden(name))


@pytest.mark.parametrize("name", sorted(n for n in MANIFEST if MANIFEST[n]["save_all"]))
def test_cli_drop_in(tmp_path, name):
    # The command line, with every final member exported (population.json):
    # polynomial, traversal and statement grammars.
    print("cli", name, flush=True)
    m = MANIFEST[name]
    out = tmp_path / "run"
    cmd = [gagp_amd.CLI_PATH, "run", os.path.join(GOLDEN, m["task"]), "--engine", "gpu", "--population", str(m["population"]),
           "--generations", str(m["generations"]), "--seed", str(m["seed"]), "--out", str(out), "--save-all", "--quiet"]
    if "search" in m:
        cmd += ["--search", os.path.join(GOLDEN, m["search"])]
    if "init" in m:
        cmd += ["--init", os.path.join(GOLDEN, m["init"])]
    subprocess.run(cmd, check=True, timeout=300)
    mine = json.load(open(out / "result.json"))
    assert strip_timing(mine) == strip_timing(golden(name))
    # population.json (format, grammar content hash, ranked programs) and
    # history.json (every generation's best program and fitness) as the
    # reference engine wrote them, whole
    for doc in ("population.json", "history.json"):
        ref = json.load(open(os.path.join(GOLDEN, "runs", name, doc)))
        assert json.load(open(out / doc)) == ref, doc


def test_cli_rejects_unsupported_with_exit_2(tmp_path):
    r = subprocess.run([gagp_amd.CLI_PATH, "run", os.path.join(GOLDEN, "tasks", "fuel-profile", "task.json"), "--quiet"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "gagp:" in r.stderr


def test_init_that_the_resident_layout_declines_is_unsupported():
    # The reference's resident layout declines this population (a site height
    # above site_depth) and runs it in its native representation; through the
    # C ABI the engine declines it as GAGP_ERR_UNSUPPORTED, for that reason.
    declined = json.load(open(os.path.join(GOLDEN, "declined.json")))["quintic_wide_deep"]
    print("declined init", declined["init"], flush=True)
    b = gagp_amd.Batch()
    with pytest.raises(gagp_amd.GagpError) as e:
        b.add(os.path.join(GOLDEN, declined["task"]), None, declined["population"], 2, 6,
              os.path.join(GOLDEN, declined["init"]))
        b.prepare()
    assert e.value.status == 2 and "import site annotation failed" in str(e.value)
    assert "import site annotation failed" in declined["resident_declined"]
