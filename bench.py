"""Benchmark of the device GAGP engine on one MI355X.

Workload: the quintic symbolic-regression search (seeded quintic with
coefficients in [-5,5]; 256 train / 512 test cases, x in [-50,50];
add/sub/mul + constants 0..9; 64 nodes; population 4096, 128 generations;
absolute error; the default resident steady-state search). A step is one
generation (one steady-state round: selection, crossover, mutation,
lowering, VM scoring of every non-reused child on every case, placement).

Prints one JSON line: program-case evaluations per second over K timed steps
(after W warmup steps) with the population resident in HBM, the roofline
block of the dominant kernel (the VM), and the reference CPU engine's rate on
a bounded sample of the same search timed on this host.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import gagp_amd  # noqa: E402

TASK = os.path.join(ROOT, "tests", "golden", "tasks", "quintic", "task.json")
POPULATION, GENERATIONS = 4096, 128
HBM_PEAK_GBS = 8000.0
REF_GAGP = os.path.join(ROOT, "oracle", "_ref", "gagp")


def program_case_rate(batch, index, steps):
    before = batch.counters(index)
    t0 = time.perf_counter()
    batch.step(steps)
    t1 = time.perf_counter()
    after = batch.counters(index)
    return before, after, t1 - t0


def roofline(steps_warm, steps):
    """Dominant kernel (k_eval) with HIP events around every launch."""
    b = gagp_amd.Batch()
    b.add(TASK, None, POPULATION, GENERATIONS, 0)
    b.set_profile(True)
    b.prepare()
    b.step(steps_warm, use_graph=False)
    t_before = b.kernel_times()
    c0 = b.counters(0)
    b.step(steps, use_graph=False)
    c1 = b.counters(0)
    t_after = b.kernel_times()
    eval_ms = (t_after["k_eval"] - t_before.get("k_eval", 0.0)) / steps
    per_kernel = {k: round((t_after[k] - t_before.get(k, 0.0)) / steps, 4) for k in t_after if not k.startswith("k_init")
                  and k not in ("k_eval_init", "k_lower_members", "k_seed")}
    # Algorithmic HBM bytes of one k_eval launch: each evaluated program's
    # code (8 B per instruction) and constants read once, the case set
    # (int64 + tag + flag per input and answer) read once, results written.
    programs = (c1["programs"] - c0["programs"]) / steps
    code_words = (c1["code_words"] - c0["code_words"]) / steps
    cases = 256
    case_bytes = cases * (10 + 10)
    algorithmic = code_words * 8 + programs * 16 * 8 + case_bytes + programs * 12
    vm_instr = (c1["vm_instructions"] - c0["vm_instructions"]) / steps
    achieved = algorithmic / (eval_ms * 1e-3) / 1e9
    return {
        "bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
        "kernel": "k_eval", "avg_launch_ms": round(eval_ms, 4), "bytes_per_launch": int(algorithmic),
        "programs_per_launch": programs, "vm_instructions_per_launch": vm_instr,
        "vm_instructions_per_s": round(vm_instr / (eval_ms * 1e-3), 1),
    }, per_kernel


def cpu_baseline(generations):
    """The reference CPU engine (built from its sources) on the same search,
    first `generations` generations, timed on this host (one thread)."""
    if not os.path.exists(REF_GAGP):
        return None
    with tempfile.TemporaryDirectory() as out:
        t0 = time.perf_counter()
        subprocess.run([REF_GAGP, "run", TASK, "--engine", "cpu", "--population", str(POPULATION), "--generations",
                        str(generations), "--seed", "0", "--out", out, "--quiet"], check=True)
        wall = time.perf_counter() - t0
        ref = json.load(open(os.path.join(out, "result.json")))
    # The same search on the device evaluates exactly the same children.
    b = gagp_amd.Batch()
    b.add(TASK, None, POPULATION, generations, 0)
    b.run()
    mine = b.result(0)
    same = mine["best"] == ref["best"] and [h["best_fitness"] for h in mine["history"]] == \
        [h["best_fitness"] for h in ref["history"]]
    cases = b.evaluated_program_cases(0)
    search_s = ref["time_ms"]["search"] / 1e3
    return {"value": round(cases / search_s, 1), "unit": "program-case evals/s", "cores": 1, "kind": "reference",
            "sample": "reference CPU engine (gagp run --engine cpu), quintic P%d seed 0, first %d generations: "
                      "%d program-cases in %.2f s search time (%.2f s wall); results identical to the device: %s"
                      % (POPULATION, generations, cases, search_s, wall, same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--cpu-generations", type=int, default=GENERATIONS)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs each search on a single device; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.warmup + args.steps > GENERATIONS:
        sys.exit("warmup + steps must not exceed the %d generations of the search" % GENERATIONS)
    b = gagp_amd.Batch()
    b.add(TASK, None, POPULATION, GENERATIONS, 0)
    b.prepare()
    b.step(args.warmup)
    before, after, seconds = program_case_rate(b, 0, args.steps)
    value = (after["program_cases"] - before["program_cases"]) / seconds
    roof, per_kernel = roofline(args.warmup, min(args.steps, 16))
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_generations)
    line = {
        "metric": "program-case evals/sec", "value": round(value, 1), "unit": "program-case evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(seconds * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": "quintic symbolic regression, P4096, 128 generations, 256 train cases, "
                               "default steady-state search (one step = one generation)",
                   "population": POPULATION, "generations": GENERATIONS, "train_cases": 256},
        "roofline": roof, "cpu_baseline": cpu, "kernel_ms_per_step": per_kernel,
    }
    if cpu:
        line["vs_cpu_reference"] = round(value / cpu["value"], 2)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
