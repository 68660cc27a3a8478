"""Writes the golden fixtures of the parity tests.

Runs in the build container only: it executes the reference's own CPU engine
(`gagp run --engine cpu`, built from the reference sources by oracle/ref.mk
into oracle/_ref/gagp) on the committed tasks and keeps each result.json
under tests/golden/runs/<name>/. It also copies the reference's bytecode
fixture files (cpp/tests/fixtures/runtime/*.json: programs, cases and
expected values or error names) into tests/golden/bytecode/.

It also writes tests/golden/eval.json, the reference's `gagp eval` verdicts
on golden run documents, and tests/golden/traversal/<task>.json: the reference's
`gagp eval --split train` of the programs tests/host_lowering.cpp exports for
the list tasks (seeded random traversal and statement programs), which
tests/test_host_lowering.py compares with the engine's lowering run on the
oracle VM, and tests/golden/definite/<task>.json: whether the reference's
verifier accepts random statement programs, some of which read unassigned
locals.

    python tests/golden/make_golden.py            (from the repository root)
    python tests/golden/make_golden.py --only 'list_'   (runs whose name matches)
"""
import argparse
import json
import os
import random
import re
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from gagp_amd import workloads  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "gagp")
REFERENCE_FIXTURES = "/root/reference/cpp/tests/fixtures/runtime"
GOLDEN = os.path.join(ROOT, "tests", "golden")
TASKS = os.path.join(GOLDEN, "tasks")

# name: (task dir, population, generations, seed, save_all[, search config under tests/golden/search])
RUNS = {
    "polynomial_p1024_g64_s0": ("polynomial", 1024, 64, 0, True),
    "polynomial_p512_g32_s7": ("polynomial", 512, 32, 7, False),
    "polynomial_p256_g20_s123": ("polynomial", 256, 20, 123, False),
    "polynomial_p100_g10_s5": ("polynomial", 100, 10, 5, False),
    "poly_c1_p1024_g64_s0": ("poly_c1", 1024, 64, 0, True),
    "quintic_p512_g16_s3": ("quintic", 512, 16, 3, False),
    "quintic_p4096_g128_s0": ("quintic", 4096, 128, 0, False),
    # rows up to 128 nodes (walks over more than one 64-node tile) and
    # tournaments of 40 (more draws than a wavefront's lanes)
    "quintic_wide_p512_g16_s5": ("quintic_wide", 512, 16, 5, True),
    "polynomial_t40_p256_g10_s8": ("polynomial", 256, 10, 8, False, "tournament40.json"),
    # generational elitism of 40 elites
    "polynomial_gen_e40_p256_g10_s7": ("polynomial", 256, 10, 7, False, "generational_e40.json"),
    # populations above 8192 keep the placement keys in HBM
    "quintic_p16384_g3_s11": ("quintic", 16384, 3, 11, False),
    "polynomial_p12000_g4_s2": ("polynomial", 12000, 4, 2, False),
    # generational schedule (elites, reservoir sites, pressure 3) and a steady batch below P with 3-member inverse tournaments
    "polynomial_gen_e1_p256_g12_s3": ("polynomial", 256, 12, 3, False, "generational_e1.json"),
    "poly_c1_gen_e3r_p512_g24_s9": ("poly_c1", 512, 24, 9, True, "generational_e3_reservoir.json"),
    "quintic_gen_e1_p1024_g16_s1": ("quintic", 1024, 16, 1, False, "generational_e1.json"),
    "polynomial_steady_b128_rs3_p512_g16_s4": ("polynomial", 512, 16, 4, False, "steady_batch_rs3.json"),
    # the runs an early-exit request stops at: the first generation >= 1 whose
    # recorded best has no wrong train case (18 and 10 in the two histories)
    "polynomial_p1024_g18_s0": ("polynomial", 1024, 18, 0, False),
    "poly_c1_p1024_g10_s0": ("poly_c1", 1024, 10, 0, False),
    # traversal grammars: the reference's list-sum example, the loop-heavy
    # sum of even elements at its full size (P 8192, G 64, 200 cases) and
    # smaller runs over the reverse, ranged, nested and low-fuel variants
    "list_sum_example_p256_g8_s1": ("list-sum", 256, 8, 1, False),
    "list_sum_even_p8192_g64_s0": ("list_sum_even", 8192, 64, 0, False),
    "list_sum_even_p512_g10_s5": ("list_sum_even", 512, 10, 5, True),
    "list_count_gt_rev_p300_g8_s6": ("list_count_gt_rev", 300, 8, 6, False),
    "list_range_nested_p256_g8_s7": ("list_range_nested", 256, 8, 7, False),
    "list_lowfuel_p400_g8_s8": ("list_lowfuel", 400, 8, 8, False),
    "list_max_range_p200_g6_s9": ("list_max_range", 200, 6, 9, False),
    "list_count_gt_rev_gen_e1_p256_g8_s2": ("list_count_gt_rev", 256, 8, 2, False, "generational_e1.json"),
    # statement programs: the loop-heavy sum of even elements as a for_range
    # loop with an accumulator local (full size and small), two locals with a
    # fixed traversal, and a low-fuel variant whose long loops time out
    "stmt_sum_even_p8192_g64_s0": ("stmt_sum_even", 8192, 64, 0, False),
    "stmt_sum_even_p512_g10_s5": ("stmt_sum_even", 512, 10, 5, True),
    "stmt_count_gt_p300_g8_s6": ("stmt_count_gt", 300, 8, 6, False),
    "stmt_lowfuel_p400_g8_s8": ("stmt_lowfuel", 400, 8, 8, False),
    "stmt_count_gt_gen_e1_p256_g8_s2": ("stmt_count_gt", 256, 8, 2, False, "generational_e1.json"),
    # statement programs that write their Int input (assign and for_range with input_name targets)
    "stmt_input_p300_g10_s3": ("stmt_input", 300, 10, 3, True),
    # let bindings: one let per program whose body reads the bound value
    "let_sum_p256_g8_s1": ("let-sum", 256, 8, 1, True),
    "let_sum_p512_g12_s4": ("let-sum", 512, 12, 4, False),
}
# Initial populations read back (`gagp run --init`): the first P programs of
# a saved population.json of the same task, over straight-line, traversal,
# let and statement grammars (input targets too), a generational schedule and
# site_depth 0, where the search grows no initial population of its own.
# name: (task dir, source run, population, generations, seed, save_all[, search])
INIT_RUNS = {
    "init_polynomial_p512_g16_s3": ("polynomial", "polynomial_p1024_g64_s0", 512, 16, 3, True),
    "init_poly_c1_p1024_g12_s5": ("poly_c1", "poly_c1_p1024_g64_s0", 1024, 12, 5, False),
    "init_polynomial_sd0_p512_g16_s3": ("polynomial", "polynomial_p1024_g64_s0", 512, 16, 3, False, "site_depth0.json"),
    "init_poly_c1_gen_e3r_p512_g12_s4": ("poly_c1", "poly_c1_gen_e3r_p512_g24_s9", 512, 12, 4, False,
                                         "generational_e3_reservoir.json"),
    "init_list_sum_even_p512_g8_s2": ("list_sum_even", "list_sum_even_p512_g10_s5", 512, 8, 2, True),
    "init_let_sum_p256_g8_s3": ("let-sum", "let_sum_p256_g8_s1", 256, 8, 3, True),
    "init_stmt_sum_even_p400_g8_s1": ("stmt_sum_even", "stmt_sum_even_p512_g10_s5", 400, 8, 1, True),
    "init_stmt_input_p300_g8_s2": ("stmt_input", "stmt_input_p300_g10_s3", 300, 8, 2, False),
    # rows above one 64-node tile: programs of 70-124 expression nodes
    # (init/quintic_wide_big.json, from wide_population())
    "init_quintic_wide_big_p512_g8_s6": ("quintic_wide", "init/quintic_wide_big.json", 512, 8, 6, False),
}
# Populations the reference's resident layout declines (it then runs them in
# its native representation, which this engine does not take): the engine
# must decline them too, for the reason the reference records.
# name: (task dir, population document, population, generations, seed)
DECLINED_RUNS = {
    "quintic_wide_deep": ("quintic_wide", "init/quintic_wide_deep.json", 32, 2, 6),
}
# Non-integral fitness: discounted proximity (-e/(1+e) per case), a
# fractional case cap and the exact metric, over straight-line, general-VM,
# traversal and statement programs, including case counts above 64 (several
# tiles per program) and above the 512 reduction lanes (several cases per
# lane), so every case-score sum goes through the canonical 48-bit order.
FITNESS_RUNS = {
    "fit_quintic_disc_p512_g16_s3": ("fitness/quintic_disc", 512, 16, 3, False),
    "fit_quintic600_disc_p256_g12_s4": ("fitness/quintic600_disc", 256, 12, 4, False),
    "fit_polynomial_cap2p5_p256_g12_s2": ("fitness/polynomial_cap2p5", 256, 12, 2, True),
    "fit_polynomial_exact_p256_g12_s4": ("fitness/polynomial_exact", 256, 12, 4, False),
    "fit_poly_c1_disc_gen_e3r_p256_g10_s6": ("fitness/poly_c1_disc", 256, 10, 6, False, "generational_e3_reservoir.json"),
    "fit_list_sum_even_disc_p512_g10_s5": ("fitness/list_sum_even_disc", 512, 10, 5, True),
    "fit_list1100_disc_p256_g6_s7": ("fitness/list1100_disc", 256, 6, 7, False),
    "fit_list_count_gt_rev_exact_p300_g8_s6": ("fitness/list_count_gt_rev_exact", 300, 8, 6, False),
    "fit_stmt_sum_even_disc_p512_g8_s2": ("fitness/stmt_sum_even_disc", 512, 8, 2, False),
    "fit_stmt_count_gt_cap7p25_p300_g8_s3": ("fitness/stmt_count_gt_cap7p25", 300, 8, 3, False),
    "fit_let_sum_disc_p256_g8_s1": ("fitness/let_sum_disc", 256, 8, 1, False),
}
# (variant dir, base task dir, fitness member)
FITNESS_VARIANTS = [
    ("quintic_disc", "quintic", {"metric": "absolute", "proximity": "discounted"}),
    ("polynomial_cap2p5", "polynomial", {"metric": "absolute", "case_cap": 2.5}),
    ("polynomial_exact", "polynomial", {"metric": "exact"}),
    ("poly_c1_disc", "poly_c1", {"metric": "absolute", "proximity": "discounted"}),
    ("list_sum_even_disc", "list_sum_even", {"metric": "absolute", "proximity": "discounted"}),
    ("list_count_gt_rev_exact", "list_count_gt_rev", {"metric": "exact"}),
    ("stmt_sum_even_disc", "stmt_sum_even", {"metric": "absolute", "proximity": "discounted"}),
    ("stmt_count_gt_cap7p25", "stmt_count_gt", {"metric": "absolute", "case_cap": 7.25}),
    ("let_sum_disc", "let-sum", {"metric": "absolute", "proximity": "discounted"}),
]


def wide_task():
    """The quintic task with a 128-node, depth-14 budget (grammar
    search_limits), so rows exceed one 64-node tile."""
    src, dst = os.path.join(TASKS, "quintic"), os.path.join(TASKS, "quintic_wide")
    os.makedirs(dst, exist_ok=True)
    for f in os.listdir(src):  # the case files
        shutil.copy(os.path.join(src, f), os.path.join(dst, f))
    task = json.load(open(os.path.join(src, "task.json")))
    task["name"] = "quintic_wide"
    grammar = json.load(open(os.path.join(src, task["grammar"])))
    grammar["search_limits"] = {"max_nodes": 128, "max_depth": 14}
    json.dump(task, open(os.path.join(dst, "task.json"), "w"), indent=1)
    json.dump(grammar, open(os.path.join(dst, task["grammar"]), "w"), indent=1)


def wide_population():
    """512 programs of 70-124 expression nodes for the wide task, so rows
    span two 64-node tiles: random add/sub/mul trees over x and constants of
    the wide run's saved programs, at most 7 levels deep (every node of this
    grammar is a site, and the resident layout admits a site height up to the
    search's site_depth, 7). The reference's own --init run decides whether
    its resident layout takes them (its result records no resident_declined)."""
    src = json.load(open(os.path.join(GOLDEN, "runs", "quintic_wide_p512_g16_s5", "population.json")))
    progs = src["programs"]
    values = sorted({c["value"] for p in progs for c in p["consts"]})
    rnd = random.Random(12)

    def tree(level, nodes, consts):
        if level < 7 and rnd.random() < 0.88:
            nodes.append({"i0": 0, "i1": 0, "kind": rnd.choice([11, 12, 13])})  # add, sub, mul
            tree(level + 1, nodes, consts)
            tree(level + 1, nodes, consts)
        elif rnd.random() < 0.5:
            nodes.append({"i0": 0, "i1": 0, "kind": 8})  # x
        else:
            nodes.append({"i0": len(consts), "i1": 0, "kind": 7})
            consts.append({"type": "int", "value": rnd.choice(values)})

    out = []
    while len(out) < 512:
        body, consts = [], []
        tree(1, body, consts)
        if 70 <= len(body) <= 124:
            nodes = progs[0]["nodes"][:3] + body + progs[0]["nodes"][-1:]  # PROGRAM, BLOCK_CONS, RETURN ... BLOCK_NIL
            out.append({"consts": consts, "names": progs[0]["names"], "nodes": nodes, "version": progs[0]["version"]})
    doc = {"format": src["format"], "grammar_hash": src["grammar_hash"], "programs": out}
    os.makedirs(os.path.join(GOLDEN, "init"), exist_ok=True)
    json.dump(doc, open(os.path.join(GOLDEN, "init", "quintic_wide_big.json"), "w"), separators=(",", ":"))


def deep_population():
    """The wide programs one level deeper (add(program, x)): a site height
    of 8, above the search's site_depth of 7."""
    doc = json.load(open(os.path.join(GOLDEN, "init", "quintic_wide_big.json")))
    fits = [p for p in doc["programs"] if len(p["nodes"]) - 4 <= 118]  # room in the 128-node budget
    programs = []
    for k in range(32):
        p = json.loads(json.dumps(fits[k % len(fits)]))
        body = p["nodes"][3:-1]
        p["nodes"] = p["nodes"][:3] + [{"i0": 0, "i1": 0, "kind": 11}] + body + [{"i0": 0, "i1": 0, "kind": 8}] + p["nodes"][-1:]
        programs.append(p)
    doc["programs"] = programs
    json.dump(doc, open(os.path.join(GOLDEN, "init", "quintic_wide_deep.json"), "w"), separators=(",", ":"))


def fitness_tasks():
    """The non-integral fitness tasks: variants that reuse a base task's
    grammar and cases with another fitness member, and two tasks of their
    own with 600 and 1100 train cases."""
    root = os.path.join(TASKS, "fitness")
    for name, base, fitness in FITNESS_VARIANTS:
        base_task = json.load(open(os.path.join(TASKS, base, "task.json")))
        task = {"format": "gagp-task-v1", "name": name}
        for key in ("grammar", "train", "test"):
            task[key] = os.path.join("..", "..", base, base_task[key])
        task["fitness"] = fitness
        os.makedirs(os.path.join(root, name), exist_ok=True)
        with open(os.path.join(root, name, "task.json"), "w") as f:
            json.dump(task, f, sort_keys=True)
    disc = {"metric": "absolute", "proximity": "discounted"}
    for name, path in (("quintic600_disc", workloads.quintic(os.path.join(root, "quintic600_disc"), 600, 64, 7)),
                       ("list1100_disc", workloads.list_task(os.path.join(root, "list1100_disc"), "sum", 17, 1100, 64,
                                                             40))):
        task = json.load(open(path))
        task["fitness"] = disc
        with open(path, "w") as f:
            json.dump(task, f, sort_keys=True)


# Tasks whose random traversal programs are checked against `gagp eval`.
STATEMENT_TASKS = ["stmt_sum_even", "stmt_count_gt", "stmt_lowfuel", "stmt_input"]
TRAVERSAL_TASKS = ["list-sum", "list_count_gt_rev", "list_range_nested", "list_lowfuel", "list_max_range",
                   "list_sum_even", "let-sum"] + STATEMENT_TASKS


def list_tasks():
    workloads.sum_even(os.path.join(TASKS, "list_sum_even"), 3)
    workloads.list_task(os.path.join(TASKS, "list_count_gt_rev"), "count_gt", 8, 64, 32, 40, fold="traverse_reverse")
    workloads.list_task(os.path.join(TASKS, "list_range_nested"), "count_gt", 13, 64, 32, 40, fold="traverse_range_reverse",
                        nested=True, fuel=420)
    workloads.list_task(os.path.join(TASKS, "list_lowfuel"), "sum_even", 12, 64, 32, 40, fuel=300)
    workloads.list_task(os.path.join(TASKS, "list_max_range"), "max", 9, 48, 32, 40, fold="traverse_range")
    workloads.sum_even(os.path.join(TASKS, "stmt_sum_even"), 3, fold="statement")
    workloads.list_task(os.path.join(TASKS, "stmt_count_gt"), "count_gt", 22, 64, 32, 40, fold="statement", with_temp=True,
                        with_traverse=True)
    workloads.list_task(os.path.join(TASKS, "stmt_lowfuel"), "sum_even", 23, 64, 32, 40, fold="statement", fuel=300)
    workloads.input_target_task(os.path.join(TASKS, "stmt_input"), 5)


def traversal_goldens(n=40):
    """Builds tests/host_lowering.cpp, exports n programs per task and keeps
    the reference's evaluation of each."""
    out_dir = os.path.join(GOLDEN, "traversal")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        driver = build_driver(tmp)
        for task_dir in TRAVERSAL_TASKS:
            task = os.path.join(TASKS, task_dir, "task.json")
            lines = subprocess.run([driver, task, str(n)], check=True, capture_output=True, text=True,
                                   cwd=ROOT).stdout.splitlines()
            programs = []
            for line in lines:
                row = json.loads(line)
                path = os.path.join(tmp, "p.json")
                with open(path, "w") as f:
                    json.dump({"best": {"program": row["program"]}}, f)
                r = subprocess.run([REF, "eval", task, path, "--split", "train"], check=True, capture_output=True, text=True)
                e = json.loads(r.stdout)
                programs.append({"i": row["i"], "text": row["text"], "cases": e["cases"], "mean_loss": e["mean_loss"],
                                 "wrong": e["wrong"], "runtime_errors": e["runtime_errors"]})
            with open(os.path.join(out_dir, task_dir + ".json"), "w") as f:
                json.dump({"task": os.path.relpath(task, GOLDEN), "n": n, "programs": programs}, f, indent=1)
            print("traversal golden", task_dir, len(programs))


def definite_assignment_goldens(n=150):
    """For the statement tasks: one random draw per program (no retry), and
    whether the reference's verifier accepts the exported AST (`gagp eval`
    rejects a read of an unassigned local with exit 2, undefined_local)."""
    out_dir = os.path.join(GOLDEN, "definite")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        driver = build_driver(tmp)
        for task_dir in STATEMENT_TASKS:
            task = os.path.join(TASKS, task_dir, "task.json")
            lines = subprocess.run([driver, task, str(n), "defs"], check=True, capture_output=True, text=True,
                                   cwd=ROOT).stdout.splitlines()
            programs = []
            for line in lines:
                row = json.loads(line)
                path = os.path.join(tmp, "p.json")
                with open(path, "w") as f:
                    json.dump({"best": {"program": row["program"]}}, f)
                r = subprocess.run([REF, "eval", task, path, "--split", "train"], capture_output=True, text=True)
                if r.returncode not in (0, 2) or (r.returncode == 2 and "undefined_local" not in r.stderr):
                    raise RuntimeError("unexpected eval outcome: %s" % r.stderr)
                programs.append({"i": row["i"], "text": row["text"], "accepted": r.returncode == 0})
            with open(os.path.join(out_dir, task_dir + ".json"), "w") as f:
                json.dump({"task": os.path.relpath(task, GOLDEN), "n": n, "programs": programs}, f, indent=1)
            print("definite assignment golden", task_dir, len(programs), sum(p["accepted"] for p in programs))


# `gagp eval` of golden run documents: (run, document, split). A result.json
# names its best program, a population.json its first.
EVAL_CASES = [
    ("polynomial_p1024_g64_s0", "result.json", "test"),
    ("polynomial_p1024_g64_s0", "population.json", "train"),
    ("quintic_p512_g16_s3", "result.json", "train"),
    ("poly_c1_gen_e3r_p512_g24_s9", "population.json", "test"),
    ("list_sum_even_p512_g10_s5", "result.json", "test"),
    ("list_sum_even_p512_g10_s5", "population.json", "train"),
    ("list_lowfuel_p400_g8_s8", "result.json", "train"),
    ("list_range_nested_p256_g8_s7", "result.json", "test"),
    ("let_sum_p256_g8_s1", "population.json", "test"),
    ("stmt_sum_even_p512_g10_s5", "population.json", "test"),
    ("stmt_input_p300_g10_s3", "result.json", "train"),
    ("stmt_lowfuel_p400_g8_s8", "result.json", "test"),
    ("fit_polynomial_exact_p256_g12_s4", "result.json", "test"),
    ("fit_list_sum_even_disc_p512_g10_s5", "population.json", "train"),
    ("fit_stmt_count_gt_cap7p25_p300_g8_s3", "result.json", "train"),
]


def eval_goldens():
    """The reference's `gagp eval` verdict on each EVAL_CASES document."""
    manifest = json.load(open(os.path.join(GOLDEN, "runs", "manifest.json")))
    cases = []
    for run, doc, split in EVAL_CASES:
        task = os.path.join(GOLDEN, manifest[run]["task"])
        path = os.path.join(GOLDEN, "runs", run, doc)
        r = subprocess.run([REF, "eval", task, path, "--split", split], check=True, capture_output=True, text=True)
        cases.append({"run": run, "document": doc, "split": split, "verdict": json.loads(r.stdout)})
    with open(os.path.join(GOLDEN, "eval.json"), "w") as f:
        json.dump({"cases": cases}, f, indent=1)
    print("eval goldens", len(cases))


def build_driver(tmp):
    driver = os.path.join(tmp, "host_lowering")
    build = os.path.join(ROOT, "gagp_amd", "_build")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "gagp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host_lowering.cpp"), os.path.join(build, "grammar.o"),
                    os.path.join(build, "task.o"), os.path.join(build, "export.o"), os.path.join(build, "identity.o"),
                    os.path.join(ROOT, "oracle", "liboracle.so"), "-Wl,-rpath," + os.path.join(ROOT, "oracle"),
                    "-o", driver], check=True)
    return driver


def run_reference(task, population, generations, seed, out, save_all, search=None, init=None):
    cmd = [REF, "run", task, "--engine", "cpu", "--population", str(population), "--generations", str(generations),
           "--seed", str(seed), "--out", out, "--quiet"]
    if search:
        cmd += ["--search", search]
    if init:
        cmd += ["--init", init]
    if save_all:
        cmd.append("--save-all")
    subprocess.run(cmd, check=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None, help="regenerate only the runs whose name matches this regex")
    args = ap.parse_args()
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    workloads.poly_c1(os.path.join(TASKS, "poly_c1"))
    workloads.quintic(os.path.join(TASKS, "quintic"))
    wide_task()
    list_tasks()
    fitness_tasks()
    mixed = workloads.mixed_searches(os.path.join(TASKS, "mixed"), n=8, seed=4, p_range=(64, 512), c_range=(16, 256))
    mixedl = workloads.mixed_searches(os.path.join(TASKS, "mixedl"), n=6, seed=6, p_range=(64, 512), c_range=(16, 256),
                                      lists=True)
    runs = dict(RUNS)
    runs.update(FITNESS_RUNS)
    for i, (task, population) in enumerate(mixed):
        runs["mixed_%d" % i] = (os.path.relpath(os.path.dirname(task), TASKS), population, 8, 11 + i, False)
    for i, (task, population) in enumerate(mixedl):
        runs["mixedl_%d" % i] = (os.path.relpath(os.path.dirname(task), TASKS), population, 8, 31 + i, False)
    manifest_path = os.path.join(GOLDEN, "runs", "manifest.json")
    manifest = json.load(open(manifest_path)) if args.only and os.path.exists(manifest_path) else {}
    for name, spec in runs.items():
        if args.only and not re.search(args.only, name):
            continue
        task_dir, population, generations, seed, save_all = spec[:5]
        search = os.path.join(GOLDEN, "search", spec[5]) if len(spec) > 5 else None
        out = os.path.join(GOLDEN, "runs", name)
        os.makedirs(out, exist_ok=True)
        task = os.path.join(TASKS, task_dir, "task.json")
        run_reference(task, population, generations, seed, out, save_all, search)
        manifest[name] = {"task": os.path.relpath(task, GOLDEN), "population": population, "generations": generations,
                          "seed": seed, "save_all": save_all}
        if search:
            manifest[name]["search"] = os.path.relpath(search, GOLDEN)
        print("golden", name)
    # after the runs whose population documents they read
    for name, spec in INIT_RUNS.items():
        if args.only and not re.search(args.only, name):
            continue
        task_dir, source, population, generations, seed, save_all = spec[:6]
        search = os.path.join(GOLDEN, "search", spec[6]) if len(spec) > 6 else None
        if source.endswith(".json"):  # a population document of its own
            if source == "init/quintic_wide_big.json":
                wide_population()
            init = os.path.join(GOLDEN, source)
        else:
            init = os.path.join(GOLDEN, "runs", source, "population.json")
        out = os.path.join(GOLDEN, "runs", name)
        os.makedirs(out, exist_ok=True)
        task = os.path.join(TASKS, task_dir, "task.json")
        run_reference(task, population, generations, seed, out, save_all, search, init)
        manifest[name] = {"task": os.path.relpath(task, GOLDEN), "population": population, "generations": generations,
                          "seed": seed, "save_all": save_all, "init": os.path.relpath(init, GOLDEN)}
        if search:
            manifest[name]["search"] = os.path.relpath(search, GOLDEN)
        print("golden", name)
    with open(manifest_path, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    declined_path = os.path.join(GOLDEN, "declined.json")
    declined = json.load(open(declined_path)) if os.path.exists(declined_path) else {}
    for name, (task_dir, source, population, generations, seed) in DECLINED_RUNS.items():
        if args.only and not re.search(args.only, name):
            continue
        if source == "init/quintic_wide_deep.json":
            deep_population()
        task = os.path.join(TASKS, task_dir, "task.json")
        with tempfile.TemporaryDirectory() as out:
            run_reference(task, population, generations, seed, out, False, None, os.path.join(GOLDEN, source))
            result = json.load(open(os.path.join(out, "result.json")))
        declined[name] = {"task": os.path.relpath(task, GOLDEN), "init": source, "population": population,
                          "resident_declined": result.get("resident_declined")}
        print("declined", name, result.get("resident_declined"))
    with open(declined_path, "w") as f:
        json.dump(declined, f, indent=1, sort_keys=True)
    if not args.only or re.search(args.only, "traversal"):
        traversal_goldens()
    if not args.only or re.search(args.only, "definite"):
        definite_assignment_goldens()
    if not args.only or re.search(args.only, "eval"):
        eval_goldens()
    if os.path.isdir(REFERENCE_FIXTURES):
        dst = os.path.join(GOLDEN, "bytecode")
        os.makedirs(dst, exist_ok=True)
        for name in sorted(os.listdir(REFERENCE_FIXTURES)):
            if name.endswith(".json"):
                shutil.copy(os.path.join(REFERENCE_FIXTURES, name), os.path.join(dst, name))


if __name__ == "__main__":
    main()
