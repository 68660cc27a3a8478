"""Benchmark of the device GAGP engine on one MI355X.

Workload: the quintic symbolic-regression search (seeded quintic with
coefficients in [-5,5]; 256 train / 512 test cases, x in [-50,50];
add/sub/mul + constants 0..9; 64 nodes; population 4096, 128 generations;
absolute error; the default resident steady-state search). A step is one
generation (one steady-state round of B = P children: selection, crossover,
mutation, lowering, VM scoring of every non-reused child on every case,
placement), with the population resident in HBM.

Prints one JSON line: program-case evaluations per second over K timed steps
(after W warmup steps), the roofline block of the kernel that takes the most
time per step (its algorithmic bytes per launch over its average launch
time, measured with HIP events on the engine's stream), per-kernel times and
bytes, and the reference CPU engine's rate on the same search timed on this
host's cores (one thread).
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import gagp_amd  # noqa: E402

TASK = os.path.join(ROOT, "tests", "golden", "tasks", "quintic", "task.json")
POPULATION, GENERATIONS, TRAIN_CASES, INPUTS = 4096, 128, 256, 1
HBM_PEAK_GBS = 8000.0
# Graph replays (the engine's default) or plain stream launches (--no-graph,
# for tools that trace every dispatch).
GRAPH = True
REF_GAGP = os.path.join(ROOT, "oracle", "_ref", "gagp")
TRAFFIC = os.path.join(ROOT, "profiles", "quintic_p4096_hbm_traffic.md")
ROUND_KERNELS = ("k_record", "k_vary", "k_vary_global", "k_clear", "k_eval", "k_eval_tiles", "k_eval_general",
                 "k_tally", "k_place", "k_place_hbm", "k_commit")


def algorithmic_bytes(c, steps):
    """Compulsory HBM bytes per launch of each round kernel, from the device
    counters of `steps` rounds. Rows: 8 B per node (origin + site), 16 B per
    constant; code: 8 B per instruction, 16 B per pool entry; keys: 16 B
    (fitness, wrong, nodes) per member or child."""
    per = {k: v / steps for k, v in c.items()}
    case_bytes = TRAIN_CASES * (INPUTS + 1) * 10  # int64 payload + tag + flag per input and per answer
    row = lambda nodes, consts: 8 * nodes + 16 * consts  # noqa: E731
    return {
        "k_eval": 8 * per["code_words"] + 16 * per["pool_words"] + case_bytes + 28 * per["programs"],
        "k_eval_tiles": 8 * per["code_words"] + 16 * per["pool_words"] + case_bytes + 28 * per["programs"],
        # rows read and written, code, per-child keys, hash and stats, the
        # member-table probe (key, member, member hash) and both rows of every
        # child whose hash matches a member (compared in full before reuse)
        "k_vary": row(per["parent_nodes"], per["parent_consts"]) + row(per["child_nodes"], per["child_consts"])
        + 8 * per["child_code_words"] + 16 * per["child_pool_words"] + 60 * per["children"] + 20 * per["children"]
        + 2 * max(0.0, per["children"] - per["programs"]) * (
            (4 * per["child_nodes"] + 16 * per["child_consts"]) / max(1.0, per["children"])),
        "k_place": 16 * POPULATION + 16 * per["children"] + 20 * per["children"],
        # placed rows copied, placements, and every member's hash indexed
        # (key and member) in the next round's table
        "k_commit": row(per["child_nodes"], per["child_consts"]) * 2 + 4 * POPULATION + 8 * POPULATION + 12 * POPULATION,
    }


def profile(warm, steps):
    """Per-kernel average launch times (HIP events around every launch on the
    engine's stream) and counters over `steps` rounds of the same search."""
    b = gagp_amd.Batch()
    b.add(TASK, None, POPULATION, GENERATIONS, 0)
    b.set_profile(True)
    b.prepare()
    b.step(warm, use_graph=False)
    t0, c0 = b.kernel_times(), b.counters(0)
    b.step(steps, use_graph=False)
    t1, c1 = b.kernel_times(), b.counters(0)
    ms = {k: (t1[k] - t0.get(k, 0.0)) / steps for k in t1 if k in ROUND_KERNELS}
    counts = {k: c1[k] - c0[k] for k in c1}
    return ms, counts


# The engine's timer names and the kernel symbols rocprofv3 records.
PROFILED_NAME = {"k_vary": "k_vary<true, false>", "k_vary_global": "k_vary<false, false>",
                 "k_vary_walks": "k_vary<true, true>", "k_vary_global_walks": "k_vary<false, true>",
                 "k_eval": "k_eval<1>", "k_eval_general": "k_eval<0>", "k_place": "k_place<true>",
                 "k_place_hbm": "k_place<false>"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE
    and WRITE_SIZE summary of this workload (None when absent)."""
    if not os.path.exists(TRAFFIC):
        return None
    name = PROFILED_NAME.get(kernel, kernel)
    for line in open(TRAFFIC):
        m = re.match(r"\|\s*%s\s*\|\s*([0-9.]+)\s*\|" % re.escape(name), line)
        if m:
            return float(m.group(1))
    return None


VALU_PROFILE = os.path.join(ROOT, "profiles", "valu_lds_counters.md")


def profiled_valu(kernel):
    """Per-launch VALU wave-instructions, lane utilization and LDS conflicts
    of `kernel` in the quintic section of the committed counter profile
    (None when absent)."""
    if not os.path.exists(VALU_PROFILE):
        return None
    name = PROFILED_NAME.get(kernel, kernel)
    section = None
    for line in open(VALU_PROFILE):
        if line.startswith("## "):
            section = line
        if not section or "Quintic" not in section:
            continue
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 9 and cells[0] == name:
            return {"valu_wave_instructions": float(cells[3]), "lane_utilization": cells[6], "lds_conflicts": float(cells[8])}
    return None


def valu_roofline(kernel, avg_ms, peak_lane_ops):
    """VALU lane operations per second of the dominant kernel: its VALU
    wave-instructions per launch (counter profile) x 64 lanes over its live
    average launch time, against the probe's measured int32 peak."""
    prof = profiled_valu(kernel)
    if not prof or not peak_lane_ops:
        return None
    achieved = prof["valu_wave_instructions"] * 64 / (avg_ms * 1e-3)
    return {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 1), "peak": round(peak_lane_ops, 1),
            "unit": "lane-ops/s", "frac": round(achieved / peak_lane_ops, 5), "lane_utilization": prof["lane_utilization"],
            "lds_bank_conflicts_per_access": prof["lds_conflicts"],
            "source": "VALU wave-instructions per launch from profiles/valu_lds_counters.md; launch time measured here"}


def roofline(ms, counts, steps):
    nbytes = algorithmic_bytes(counts, steps)
    dominant = max((k for k in ms if k in nbytes), key=lambda k: ms[k])
    avg = ms[dominant]
    achieved = nbytes[dominant] / (avg * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(dominant),
            "kernel": dominant, "avg_launch_ms": round(avg, 4), "bytes_per_launch": int(nbytes[dominant])}
    kernels = {k: {"ms": round(v, 4), "bytes": int(nbytes[k]) if k in nbytes else None,
                   "gbs": round(nbytes[k] / (v * 1e-3) / 1e9, 2) if k in nbytes and v > 0 else None}
               for k, v in sorted(ms.items(), key=lambda kv: -kv[1])}
    # the typed straight-line kernel the search's programs ran in
    vm_kernel = "k_eval_tiles" if "k_eval_tiles" in ms else "k_eval"
    vm = {"vm_instructions_per_launch": counts["vm_instructions"] / steps,
          "vm_kernel": vm_kernel,
          "vm_instructions_per_s": round(counts["vm_instructions"] / steps / (ms[vm_kernel] * 1e-3), 1),
          "programs_per_launch": counts["programs"] / steps,
          "program_cases_per_s_in_vm_kernel": round(counts["program_cases"] / steps / (ms[vm_kernel] * 1e-3), 1)}
    return roof, kernels, vm


def cpu_baseline(generations):
    """The reference CPU engine (built from its sources by oracle/ref.mk) on
    the same search for its first `generations` generations, one thread on
    this host; the device runs the same search to count the program-cases
    the reference evaluates and to confirm identical results."""
    if not os.path.exists(REF_GAGP):
        return None
    with tempfile.TemporaryDirectory() as out:
        t0 = time.perf_counter()
        subprocess.run(pinned([REF_GAGP, "run", TASK, "--engine", "cpu", "--population", str(POPULATION), "--generations",
                               str(generations), "--seed", "0", "--out", out, "--quiet"]), check=True)
        wall = time.perf_counter() - t0
        ref = json.load(open(os.path.join(out, "result.json")))
    b = gagp_amd.Batch()
    b.add(TASK, None, POPULATION, generations, 0)
    b.run(use_graph=GRAPH)
    mine = b.result(0)
    same = mine["best"] == ref["best"] and [h["best_fitness"] for h in mine["history"]] == \
        [h["best_fitness"] for h in ref["history"]]
    cases = b.evaluated_program_cases(0)
    search_s = ref["time_ms"]["search"] / 1e3
    return {"value": round(cases / search_s, 1), "unit": "program-case evals/s", "cores": 1, "kind": "reference",
            "pinned": "taskset -c 0" if shutil.which("taskset") else None, "host": host_cores(),
            "sample": "reference CPU engine (gagp run --engine cpu, built from its sources), quintic P%d seed 0, "
                      "%d generations: %d program-cases in %.2f s of search time (%.2f s wall); best program and "
                      "history identical to the device run: %s" % (POPULATION, generations, cases, search_s, wall, same)}


def host_cores():
    """Host CPU description: cores this process may use and the model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": len(os.sched_getaffinity(0)), "model": model}


def pinned(cmd, core=0):
    """`cmd` pinned to one host core with taskset (when present)."""
    return (["taskset", "-c", str(core)] if shutil.which("taskset") else []) + cmd


def valu_peak():
    """Measured int32 VALU issue rate of this device: the best of a few
    launch sizes of the library's probe kernel (8 VALU operations per
    iteration per lane), in lane operations per second."""
    best = 0.0
    for blocks in (2048, 4096, 8192):
        iters = 8192
        ms = gagp_amd.valu_probe(blocks, iters)
        best = max(best, 8.0 * iters * 256 * blocks / (ms * 1e-3))
    return best


def end_to_end(name, task, population, generations, seed, with_cpu):
    """One whole search end to end: in-process on the device (prepare,
    every round, result), the `gagp run --engine gpu` command line (process
    start to exit), and the reference CPU engine on one pinned core with the
    same arguments, whose result must be identical."""
    b = gagp_amd.Batch()
    b.add(task, None, population, generations, seed)
    t0 = time.perf_counter()
    b.run(use_graph=GRAPH)
    device_s = time.perf_counter() - t0
    mine = strip_timing(b.result(0))
    cases = b.evaluated_program_cases(0)
    out = {"workload": name, "population": population, "generations": generations, "seed": seed,
           "device_search_s": round(device_s, 4), "program_case_evals_per_s": round(cases / device_s, 1)}
    with tempfile.TemporaryDirectory() as tmp:
        t0 = time.perf_counter()
        subprocess.run([gagp_amd.CLI_PATH, "run", task, "--engine", "gpu", "--population", str(population), "--generations",
                        str(generations), "--seed", str(seed), "--out", tmp, "--quiet"], check=True)
        out["cli_wall_s"] = round(time.perf_counter() - t0, 3)
        out["cli_identical"] = strip_timing(json.load(open(os.path.join(tmp, "result.json")))) == mine
        if with_cpu and os.path.exists(REF_GAGP):
            t0 = time.perf_counter()
            subprocess.run(pinned([REF_GAGP, "run", task, "--engine", "cpu", "--population", str(population), "--generations",
                                   str(generations), "--seed", str(seed), "--out", tmp, "--quiet"]), check=True)
            ref = json.load(open(os.path.join(tmp, "result.json")))
            out["reference_cpu"] = {"wall_s": round(time.perf_counter() - t0, 3),
                                    "search_s": round(ref["time_ms"]["search"] / 1e3, 4), "cores": 1,
                                    "identical": strip_timing(ref) == mine}
    return out


# The loop-heavy search in its two grammar forms: one traversal per program,
# and statement programs (a for_range loop, an accumulator local, if_stmt).
LOOP_FORMS = {
    "traversal": ("list_sum_even", "one traversal per program with mod/eq/if/add/sub/max in its body"),
    "statement": ("stmt_sum_even", "statement programs: for_range over the list, accumulator local, if_stmt, "
                                   "mod/eq/lt/add/sub, index and len"),
}


def loop_heavy(form, cpu_generations):
    """The loop-heavy search: the sum of the even elements of an Int list
    (200 train / 400 test cases, lengths 0-50, values in [-1000,1000]) in the
    grammar `form` of LOOP_FORMS, population 8192, 64 generations, seed 0, on
    the device, checked against the reference's golden result; and the
    reference CPU engine on the first `cpu_generations` generations of the
    same search (one thread), or None."""
    task_dir, grammar = LOOP_FORMS[form]
    task = os.path.join(ROOT, "tests", "golden", "tasks", task_dir, "task.json")
    golden = os.path.join(ROOT, "tests", "golden", "runs", task_dir + "_p8192_g64_s0", "result.json")
    b = gagp_amd.Batch()
    b.add(task, None, 8192, 64, 0)
    t0 = time.perf_counter()
    b.prepare()
    t1 = time.perf_counter()
    b.step(b.max_rounds(), use_graph=GRAPH)
    b.finish()
    t2 = time.perf_counter()
    cases = b.evaluated_program_cases(0)
    same = strip_timing(b.result(0)) == strip_timing(json.load(open(golden)))
    out = {"workload": "sum of the even elements of an Int list, population 8192, 64 generations, 200 train cases; "
                       + grammar,
           "value": round(cases / (t2 - t0), 1), "unit": "program-case evals/s", "seconds": round(t2 - t0, 3),
           "prepare_s": round(t1 - t0, 3), "ms_per_generation": round((t2 - t1) * 1e3 / 64, 3),
           "identical_to_reference": same, "cpu_reference": None}
    if cpu_generations and os.path.exists(REF_GAGP):
        with tempfile.TemporaryDirectory() as tmp:
            subprocess.run(pinned([REF_GAGP, "run", task, "--engine", "cpu", "--population", "8192", "--generations",
                                   str(cpu_generations), "--seed", "0", "--out", tmp, "--quiet"]), check=True)
            ref = json.load(open(os.path.join(tmp, "result.json")))
        c = gagp_amd.Batch()
        c.add(task, None, 8192, cpu_generations, 0)
        c.run(use_graph=GRAPH)
        ref_cases = c.evaluated_program_cases(0)
        search_s = ref["time_ms"]["search"] / 1e3
        out["cpu_reference"] = {
            "value": round(ref_cases / search_s, 1), "unit": "program-case evals/s", "cores": 1, "kind": "reference",
            "sample": "first %d generations: %d program-cases in %.2f s of search time" % (cpu_generations, ref_cases, search_s)}
        out["vs_cpu_reference"] = round(out["value"] / out["cpu_reference"]["value"], 1)
    return out


def percentile(values, q):
    v = sorted(values)
    return v[min(len(v) - 1, int(round(q / 100.0 * (len(v) - 1))))]


def batching(n, generations, run_alone=True):
    """Searches/s and completion latency of a mixed workload run one search at
    a time and as one batch: n seeded searches, half polynomials of degree 2-5
    (coefficients in [-5,5], x in [-50,50]) and half list reductions (sum,
    max or count of elements > k; lengths 0-64, values in [-1000,1000]),
    population log-uniform in 256-4096, train cases log-uniform in 16-1024,
    `generations` generations, all submitted at t=0. Latency = completion
    time of each search from t=0."""
    from gagp_amd import workloads

    def strip(r):
        return {k: r[k] for k in ("search", "best", "test")} | {"history": [
            {k: v for k, v in h.items() if k in ("best_fitness", "best_nodes", "best_wrong", "mean_fitness")}
            for h in r["history"]]}

    with tempfile.TemporaryDirectory() as tmp:
        tasks = workloads.mixed_searches(tmp, n=n, seed=4, p_range=(256, 4096), c_range=(16, 1024), lists=True)
        alone, lat_alone, cases = [], [], 0
        t0 = time.perf_counter()
        for i, (task, population) in enumerate(tasks if run_alone else []):
            b = gagp_amd.Batch()
            b.add(task, None, population, generations, 11 + i)
            b.run(use_graph=GRAPH)
            alone.append(strip(b.result(0)))
            lat_alone.append(time.perf_counter() - t0)
        alone_s = time.perf_counter() - t0
        # The batch with the evaluation launch shapes fixed (autotuner off):
        # the gain of the tuned shapes below is measured against it.
        t0 = time.perf_counter()
        f = gagp_amd.Batch()
        f.set_autotune(0)
        for i, (task, population) in enumerate(tasks):
            f.add(task, None, population, generations, 11 + i)
        f.run(use_graph=GRAPH)
        fixed = [strip(f.result(i)) for i in range(n)]
        fixed_s = time.perf_counter() - t0
        if not run_alone:
            alone = fixed  # identity checks against the fixed-shape batch instead
        t0 = time.perf_counter()
        b = gagp_amd.Batch()
        for i, (task, population) in enumerate(tasks):
            b.add(task, None, population, generations, 11 + i)
        t1 = time.perf_counter()
        b.prepare()
        t2 = time.perf_counter()
        b.step(b.max_rounds(), use_graph=GRAPH)
        b.finish()
        t3 = time.perf_counter()
        together = [strip(b.result(i)) for i in range(n)]
        batched_s = time.perf_counter() - t0
        phases = {"load_s": round(t1 - t0, 4), "prepare_s": round(t2 - t1, 4), "rounds_s": round(t3 - t2, 4),
                  "results_s": round(batched_s - (t3 - t0), 4)}
        # The batch with launch-shape tuning forced (the default skips it
        # for tables with fewer than 128 rounds left): its choices and its
        # net gain over the fixed shapes, measurement included.
        t0 = time.perf_counter()
        a = gagp_amd.Batch()
        a.set_autotune(2)
        for i, (task, population) in enumerate(tasks):
            a.add(task, None, population, generations, 11 + i)
        a.run(use_graph=GRAPH)
        forced = [strip(a.result(i)) for i in range(n)]
        forced_s = time.perf_counter() - t0
        tuning = a.tuning()
        cases = sum(b.evaluated_program_cases(i) for i in range(n))
        # The same batch once more with every launch bracketed by HIP events
        # on the engine's stream (stream launches, no graph): device
        # milliseconds per kernel over the whole batched run.
        p = gagp_amd.Batch()
        for i, (task, population) in enumerate(tasks):
            p.add(task, None, population, generations, 11 + i)
        p.set_profile(True)
        p.prepare()
        p.step(p.max_rounds(), use_graph=False)
        p.finish()
        kernel_ms = {k: round(v, 3) for k, v in sorted(p.kernel_times().items(), key=lambda kv: -kv[1])}
        profiled_same = sum(strip(p.result(i)) == x for i, x in enumerate(alone))
    same = sum(x == y for x, y in zip(alone, together))
    one = None if not run_alone else {"searches_per_s": round(n / alone_s, 3), "seconds": round(alone_s, 3),
                                      "p50_latency_s": round(percentile(lat_alone, 50), 3),
                                      "p99_latency_s": round(percentile(lat_alone, 99), 3)}
    return {"searches": n, "generations": generations,
            "one_at_a_time": one,
            "batched": {"searches_per_s": round(n / batched_s, 3), "seconds": round(batched_s, 3),
                        "p50_latency_s": round(batched_s, 3), "p99_latency_s": round(batched_s, 3),
                        "program_case_evals_per_s": round(cases / batched_s, 1), "phases": phases},
            "speedup": round(alone_s / batched_s, 2) if run_alone else None, "identical_to_alone": "%d/%d" % (same, n),
            "batched_kernel_ms": kernel_ms, "profiled_identical_to_alone": "%d/%d" % (profiled_same, n),
            "default_tuning": b.tuning(),
            "autotune": {"forced_tuning": tuning, "fixed_shapes_s": round(fixed_s, 3), "tuned_s": round(forced_s, 3),
                         "gain": round(fixed_s / forced_s, 3),
                         "fixed_identical_to_alone": "%d/%d" % (sum(x == y for x, y in zip(alone, fixed)), n),
                         "tuned_identical_to_alone": "%d/%d" % (sum(x == y for x, y in zip(alone, forced)), n)}}


def serve_stream(requests, max_live):
    """Serves `requests` [(task, population, generations, seed, arrival_s)]
    with early exit: max_live 0 is continuous batching, 1 one search at a
    time. Returns the summary, the canonical result documents and the
    outcomes."""
    n = len(requests)
    server = gagp_amd.Server(budget_bytes=64 << 30, poll_rounds=2, max_live=max_live)
    tickets = [server.submit(task, None, population, generations, seed, True, arrival)
               for task, population, generations, seed, arrival in requests]
    t0 = time.perf_counter()
    server.run()
    makespan = time.perf_counter() - t0
    outcomes = [server.outcome(t) for t in tickets]
    results = [json.dumps(strip_timing(server.result(t)), sort_keys=True) for t in tickets]
    lat = [o["done"] - o["arrival"] for o in outcomes]
    return {"searches_per_s": round(n / makespan, 3), "makespan_s": round(makespan, 3),
            "p50_latency_s": round(percentile(lat, 50), 4), "p99_latency_s": round(percentile(lat, 99), 4),
            "generations_run": sum(o["generations"] for o in outcomes)}, results, outcomes


def reference_sample(requests, results, outcomes, fraction=0.05, seed=5, budget=4e8):
    """Bit-exactness of served results against the reference CPU engine on
    a seeded sample of the requests: requests are visited in a seeded random
    order and taken while the sample is below `fraction` of all requests,
    skipping any whose reference run would evaluate more than `budget`
    program-cases (population x completed generations x train cases, about a
    minute of one core); each sampled request's published result (early exit
    at generation g) must equal `gagp run --engine cpu --generations g` of
    the reference, run one per core in parallel."""
    import random
    from concurrent.futures import ThreadPoolExecutor
    if not os.path.exists(REF_GAGP):
        return None
    want = max(1, int(round(fraction * len(requests))))
    order = list(range(len(requests)))
    random.Random(seed).shuffle(order)
    picked, skipped = [], 0
    for i in order:
        task, population, _, rseed, _ = requests[i]
        cases = len(json.load(open(os.path.join(os.path.dirname(task), "train.cases.json")))["cases"])
        cost = population * max(1, outcomes[i]["generations"]) * cases
        if cost > budget:
            skipped += 1
            continue
        picked.append((i, cost))
        if len(picked) == want:
            break

    def run(item):
        i, _ = item
        task, population, _, rseed, _ = requests[i]
        with tempfile.TemporaryDirectory() as out:
            subprocess.run([REF_GAGP, "run", task, "--engine", "cpu", "--population", str(population), "--generations",
                            str(outcomes[i]["generations"]), "--seed", str(rseed), "--out", out, "--quiet"], check=True)
            ref = json.load(open(os.path.join(out, "result.json")))
        return json.dumps(strip_timing(ref), sort_keys=True) == results[i]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(len(picked), 8) or 1) as pool:
        same = list(pool.map(run, picked))
    return {"sampled": len(picked), "of": len(requests), "identical": "%d/%d" % (sum(same), len(same)),
            "requests": [{"index": i, "population": requests[i][1], "generations": outcomes[i]["generations"],
                          "program_cases": cost} for i, cost in picked],
            "skipped_over_budget": skipped, "budget_program_cases": budget, "seconds": round(time.perf_counter() - t0, 1)}


def serving(n, rate, stress_rate, check_reference):
    """The same request stream served with continuous batching and by a
    server that admits one search at a time (max_live 1): completed
    searches/s over the makespan, p50/p99 latency from arrival to published
    result, and whether every result is identical in both. The stream is
    Poisson at `rate` requests/s, again at `stress_rate` (above the
    one-at-a-time capacity, so throughput shows the difference), and, for
    capacity, the same requests all arriving at t=0. Half the requests are
    list reductions. With `check_reference`, a seeded 5% sample of the
    results is compared with the reference CPU engine."""
    from gagp_amd import workloads

    def compare(requests):
        batched, rb, ob = serve_stream(requests, 0)
        serial, rs, _ = serve_stream(requests, 1)
        return {"continuous_batching": batched, "one_at_a_time": serial,
                "throughput_ratio": round(batched["searches_per_s"] / serial["searches_per_s"], 2),
                "identical_results": "%d/%d" % (sum(a == b for a, b in zip(rb, rs)), n)}, rb, ob

    with tempfile.TemporaryDirectory() as tmp:
        requests = workloads.served_requests(tmp, n=n, seed=5, rate=rate, lists=True)
        at_rate, results, outcomes = compare(requests)
        stress = [(task, population, generations, seed, arrival * rate / stress_rate)
                  for task, population, generations, seed, arrival in requests]
        at_stress, _, _ = compare(stress)
        burst = [(task, population, generations, seed, 0.0) for task, population, generations, seed, _ in requests]
        all_at_once, _, _ = compare(burst)
        sample = reference_sample(requests, results, outcomes) if check_reference else None
    out = {"requests": n, "rate_per_s": rate, "early_exit": True,
           "populations": "log-uniform 256-16384", "train_cases": "log-uniform 16-2048",
           "generation_caps": "uniform 32-256"}
    out.update(at_rate)
    out["at_%g_per_s" % stress_rate] = at_stress
    out["all_at_once"] = all_at_once
    out["reference_sample"] = sample
    return out


def strip_timing(r):
    """The result.json members that must equal the reference engine's (its
    engine name, task path and timings are not compared)."""
    out = {k: r[k] for k in ("search", "best", "test", "population", "generations", "seed", "train_cases", "format") if k in r}
    out["history"] = [{k: v for k, v in h.items() if not k.endswith("_ms")} for h in r["history"]]
    return out


def calibrate_copy(mib):
    """One streaming kernel of known size for the HBM counters: reads and
    writes `mib` MiB (int32 negation, one read and one write per element)."""
    import torch
    a = torch.ones(mib << 18, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.neg(a, out=b)
    torch.cuda.synchronize()
    print("copy calibration: %d MiB read, %d MiB written" % (mib, mib), file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--cpu-generations", type=int, default=GENERATIONS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0, help="run a streaming kernel of this size first (counter scale)")
    ap.add_argument("--batch-searches", type=int, default=64, help="searches of the batching comparison (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=60, help="requests of the serving comparison (0: skip)")
    ap.add_argument("--no-loop-heavy", dest="loop_heavy", action="store_false", help="skip the loop-heavy search")
    ap.add_argument("--no-graph", action="store_true", help="stream launches instead of graph replays everywhere")
    ap.add_argument("--no-whole", dest="whole", action="store_false", help="skip the whole-search timings")
    ap.add_argument("--no-alone", dest="alone", action="store_false",
                    help="batching: skip the one-at-a-time leg (for tools that trace every dispatch)")
    args = ap.parse_args()
    global GRAPH
    GRAPH = not args.no_graph
    if args.gpus != 1:
        print("bench.py measures one GPU: each search runs on a single device and the engine has no multi-GPU "
              "path; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.warmup + args.steps > GENERATIONS:
        sys.exit("warmup + steps must not exceed the %d generations of the search" % GENERATIONS)
    if args.copy_mib:
        calibrate_copy(args.copy_mib)
    b = gagp_amd.Batch()
    b.add(TASK, None, POPULATION, GENERATIONS, 0)
    b.prepare()
    b.step(args.warmup, use_graph=GRAPH)
    before = b.counters(0)
    t0 = time.perf_counter()
    b.step(args.steps, use_graph=GRAPH)  # graph replays, synchronized before returning
    seconds = time.perf_counter() - t0
    after = b.counters(0)
    value = (after["program_cases"] - before["program_cases"]) / seconds
    psteps = min(args.steps, 16)
    ms, counts = profile(args.warmup, psteps)
    roof, kernels, vm = roofline(ms, counts, psteps)
    peak_lane_ops = valu_peak()
    # Window-independent figures: whole searches end to end.
    whole = None if not args.whole else {
        "quintic": end_to_end("quintic symbolic regression (the headline search), all 128 generations", TASK, POPULATION,
                              GENERATIONS, 0, False),
        "c1_polynomial_example": end_to_end(
            "the reference's examples/polynomial (y = x^3 - 2x + 1, constants 1..3; the reference publishes 0.4 s per "
            "run on an RTX 3090, README.md:53)", os.path.join(ROOT, "tests", "golden", "tasks", "polynomial", "task.json"),
            1024, 64, 0, not args.no_cpu),
        "c1": end_to_end("y = x^3 - 2x + 1 with constants {1, 2}, 21 train / 40 test cases",
                         os.path.join(ROOT, "tests", "golden", "tasks", "poly_c1", "task.json"), 1024, 64, 0, not args.no_cpu),
    }
    loop = {form: loop_heavy(form, 0 if args.no_cpu else 2) for form in LOOP_FORMS} if args.loop_heavy else None
    batch = batching(args.batch_searches, 64, args.alone) if args.batch_searches else None
    serve = serving(args.serve_requests, 20.0, 60.0, not args.no_cpu) if args.serve_requests else None
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_generations)
    line = {
        "metric": "program-case evals/sec", "value": round(value, 1), "unit": "program-case evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(seconds * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded quintic task generated in-repo)",
        "config": {"workload": "quintic symbolic regression, population 4096, 128 generations, 256 train cases, "
                               "default resident steady-state search; one step = one generation",
                   "population": POPULATION, "generations": GENERATIONS, "train_cases": TRAIN_CASES},
        "roofline": roof, "roofline_valu": valu_roofline(roof["kernel"], roof["avg_launch_ms"], peak_lane_ops),
        "valu_peak_lane_ops_per_s": round(peak_lane_ops, 1), "cpu_baseline": cpu, "kernels": kernels, "vm": vm, "whole_searches": whole, "loop_heavy": loop,
        "batching": batch,
        "serving": serve,
    }
    if cpu:
        line["vs_cpu_reference"] = round(value / cpu["value"], 2)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
