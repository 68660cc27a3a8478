"""Population import on the host (CPU): tests/import_driver.cpp runs the
engine's import path (ast-prefix-v2 decoding, first-match grammar
membership, site annotation) on population documents the reference wrote
(`gagp run --save-all`, kept under tests/golden/runs/) and exports every
imported row again. The export must reproduce each program exactly, and the
sites annotated from a generated row's origins must be the ones the
generator wrote. The device runs of those imported populations are pinned
against the reference's `gagp run --init` in tests/test_gpu_search.py."""
import json
import os
import subprocess

import pytest

from _util import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "gagp_amd", "_build")
MANIFEST = json.load(open(os.path.join(GOLDEN, "runs", "manifest.json")))
SAVED = sorted(n for n in MANIFEST if os.path.exists(os.path.join(GOLDEN, "runs", n, "population.json")))
# Populations the reference's resident layout declines, with its recorded reason
DECLINED = json.load(open(os.path.join(GOLDEN, "declined.json")))
SITE_DEPTH = 7  # the default search's limits.site_depth


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    objs = [os.path.join(BUILD, n) for n in ("grammar.o", "task.o", "export.o", "identity.o", "population.o")]
    missing = [p for p in objs if not os.path.exists(p)]
    assert not missing, "build the engine first (__graft_entry__.build()): missing %s" % missing
    exe = str(tmp_path_factory.mktemp("import") / "import_driver")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "gagp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "import_driver.cpp")] + objs + ["-o", exe], check=True)
    return exe


def run(driver, *args):
    r = subprocess.run([driver] + [str(a) for a in args], capture_output=True, text=True, cwd=ROOT)
    lines = [json.loads(line) for line in r.stdout.splitlines()]
    return r.returncode, lines


def task_of(name):
    return os.path.join(GOLDEN, MANIFEST[name]["task"])


@pytest.mark.parametrize("name", SAVED)
def test_saved_population_round_trips(driver, name):
    pop = os.path.join(GOLDEN, "runs", name, "population.json")
    n = len(json.load(open(pop))["programs"])
    rc, rows = run(driver, task_of(name), pop, n, SITE_DEPTH)
    assert rc == 0, rows[-1:]
    assert len(rows) == n
    assert all(r["same"] for r in rows), [r["i"] for r in rows if not r["same"]][:5]
    # One origin per phase node: the whole program less the PROGRAM,
    # BLOCK_CONS, RETURN ... BLOCK_NIL envelope of an expression entry.
    entry_program = "stmt" in MANIFEST[name]["task"]
    for r in rows:
        assert len(r["origins"]) == r["nodes"] - (0 if entry_program else 4)
        assert len(r["sites"]) == len(r["origins"]) and r["sites"][0] >= 0


@pytest.mark.parametrize("task", ["polynomial", "poly_c1", "quintic", "list_sum_even", "list_range_nested", "let-sum",
                                  "stmt_count_gt", "stmt_input"])
def test_generated_rows_annotate_back(driver, task):
    # Rows from the device's generator (run on the host): exported, then
    # imported, they give back the same origins and the generator's sites.
    # (statement rows that read an unassigned local are skipped, as the
    # engine's generator skips them)
    rc, rows = run(driver, os.path.join(GOLDEN, "tasks", task, "task.json"), "-", 1000, SITE_DEPTH, "generate")
    assert rc == 0, rows[-1:]
    assert len(rows) > 100
    assert all(r["same_sites"] for r in rows), [r["i"] for r in rows if not r["same_sites"]][:5]


@pytest.mark.parametrize("name", ["polynomial_p1024_g64_s0", "list_sum_even_p512_g10_s5", "let_sum_p256_g8_s1"])
def test_load_population_takes_the_first_programs(driver, name):
    pop = os.path.join(GOLDEN, "runs", name, "population.json")
    programs = json.load(open(pop))["programs"]
    rc, rows = run(driver, task_of(name), pop, 100, SITE_DEPTH, "load")
    assert rc == 0, rows
    (row,) = rows
    assert row["rows"] == 100
    # The export layout is the first program's: its name table, and binder
    # ids above every one it holds.
    first = programs[0]
    assert row["names"] == first["names"]
    ids = [b["id"] for r in first.get("lexical_regions", []) for b in r["bindings"]]
    assert row["binder_base"] == (max(ids) + 1 if ids else 0)


def test_load_population_errors(driver, tmp_path):
    pop = os.path.join(GOLDEN, "runs", "polynomial_p1024_g64_s0", "population.json")
    poly = task_of("polynomial_p1024_g64_s0")
    # too few programs for the run
    rc, rows = run(driver, poly, pop, 2000, SITE_DEPTH, "load")
    assert rc == 2 and "has 1024 programs, the run needs 2000" in rows[0]["error"]
    # made for another grammar
    rc, rows = run(driver, os.path.join(GOLDEN, "tasks", "quintic", "task.json"), pop, 10, SITE_DEPTH, "load")
    assert rc == 2 and "different grammar" in rows[0]["error"]
    doc = json.load(open(pop))
    # a constant outside the grammar's domain: no derivation
    bad = json.loads(json.dumps(doc))
    prog = bad["programs"][0]
    for c in prog["consts"]:
        c["value"] = 99
    path = tmp_path / "bad.json"
    path.write_text(json.dumps(bad))
    rc, rows = run(driver, poly, path, 10, SITE_DEPTH, "load")
    assert rc == 2 and "cannot be derived" in rows[0]["error"]
    # without a grammar hash the programs are still checked one by one
    del bad["grammar_hash"]
    path.write_text(json.dumps(bad))
    rc, rows = run(driver, poly, path, 10, SITE_DEPTH, "load")
    assert rc == 2 and "cannot be derived" in rows[0]["error"]
    # a program deeper than the grammar's depth budget
    deep = json.loads(json.dumps(doc))
    prog = deep["programs"][0]
    neg = [{"kind": 9, "i0": 0, "i1": 0}] * 70  # NEG chain
    prog["nodes"] = prog["nodes"][:3] + neg + prog["nodes"][3:]
    path.write_text(json.dumps(deep))
    rc, rows = run(driver, poly, path, 10, SITE_DEPTH, "load")
    assert rc == 2 and "budget exceeded" in rows[0]["error"]
    # the replay format of `gagp grammar generate` is declined
    path.write_text(json.dumps({"format_version": "grammar-population-v2", "programs": []}))
    rc, rows = run(driver, poly, path, 10, SITE_DEPTH, "load")
    assert rc == 2 and "grammar-population-v2" in rows[0]["error"]
    # a document without programs
    path.write_text(json.dumps({"format": "gagp-population-v1"}))
    rc, rows = run(driver, poly, path, 10, SITE_DEPTH, "load")
    assert rc == 2 and "programs array" in rows[0]["error"]


@pytest.fixture(scope="module")
def lowering_driver(tmp_path_factory):
    objs = [os.path.join(BUILD, n) for n in ("grammar.o", "task.o", "export.o", "identity.o")]
    oracle = os.path.join(ROOT, "oracle", "liboracle.so")
    missing = [p for p in objs + [oracle] if not os.path.exists(p)]
    assert not missing, "build the engine and the oracle first (__graft_entry__.build()): missing %s" % missing
    exe = str(tmp_path_factory.mktemp("lowering") / "host_lowering")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "gagp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host_lowering.cpp")] + objs +
                   [oracle, "-Wl,-rpath," + os.path.dirname(oracle), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("task", sorted(f[:-5] for f in os.listdir(os.path.join(GOLDEN, "definite")) if f.endswith(".json")))
def test_import_rejects_undefined_local_reads(driver, lowering_driver, task, tmp_path):
    # Random statement programs, some reading a local before it is
    # assigned: the import admits exactly those the reference's verifier
    # accepts (tests/golden/definite/, from `gagp eval`).
    golden = json.load(open(os.path.join(GOLDEN, "definite", task + ".json")))
    task_path = os.path.join(GOLDEN, golden["task"])
    lines = subprocess.run([lowering_driver, task_path, str(golden["n"]), "defs"], check=True, capture_output=True,
                           text=True, cwd=ROOT).stdout.splitlines()
    programs = [json.loads(line) for line in lines]
    assert [p["i"] for p in programs] == [p["i"] for p in golden["programs"]]
    doc = tmp_path / "programs.json"
    doc.write_text(json.dumps({"programs": [p["program"] for p in programs]}))
    rc, rows = run(driver, task_path, doc, len(programs), SITE_DEPTH, "each")
    assert rc == 0 and len(rows) == len(programs)
    for row, want in zip(rows, golden["programs"]):
        if want["accepted"]:
            assert row["error"] is None, (want["text"], row["error"])
        else:
            assert row["error"] and "undefined_local" in row["error"], (want["text"], row["error"])


@pytest.mark.parametrize("name", sorted(DECLINED))
def test_import_declines_what_the_reference_resident_layout_declines(driver, name):
    # Programs of a site height above the search's site_depth: the
    # reference's resident layout declines them (and runs them in its native
    # representation); the engine declines them as unsupported, for the same
    # reason, rather than as invalid input.
    d = DECLINED[name]
    assert d["resident_declined"], name
    rc, rows = run(driver, os.path.join(GOLDEN, d["task"]), os.path.join(GOLDEN, d["init"]), d["population"], SITE_DEPTH,
                   "load")
    assert rc == 2 and rows[0]["unsupported"] is True, rows
    assert rows[0]["error"] in d["resident_declined"], (rows[0]["error"], d["resident_declined"])
