"""Writes the golden fixtures of the parity tests.

Runs in the build container only: it executes the reference's own CPU engine
(`gagp run --engine cpu`, built from the reference sources by oracle/ref.mk
into oracle/_ref/gagp) on the committed tasks and keeps each result.json
under tests/golden/runs/<name>/. It also copies the reference's bytecode
fixture files (cpp/tests/fixtures/runtime/*.json: programs, cases and
expected values or error names) into tests/golden/bytecode/.

    python tests/golden/make_golden.py            (from the repository root)
"""
import json
import os
import shutil
import subprocess
import sys

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
}


def run_reference(task, population, generations, seed, out, save_all, search=None):
    cmd = [REF, "run", task, "--engine", "cpu", "--population", str(population), "--generations", str(generations),
           "--seed", str(seed), "--out", out, "--quiet"]
    if search:
        cmd += ["--search", search]
    if save_all:
        cmd.append("--save-all")
    subprocess.run(cmd, check=True)
    if save_all:
        # history.json and population.json are large; keep the population only.
        os.remove(os.path.join(out, "history.json"))


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    workloads.poly_c1(os.path.join(TASKS, "poly_c1"))
    workloads.quintic(os.path.join(TASKS, "quintic"))
    mixed = workloads.mixed_searches(os.path.join(TASKS, "mixed"), n=8, seed=4, p_range=(64, 512), c_range=(16, 256))
    runs = dict(RUNS)
    for i, (task, population) in enumerate(mixed):
        runs["mixed_%d" % i] = (os.path.relpath(os.path.dirname(task), TASKS), population, 8, 11 + i, False)
    manifest = {}
    for name, spec in runs.items():
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
    with open(os.path.join(GOLDEN, "runs", "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    if os.path.isdir(REFERENCE_FIXTURES):
        dst = os.path.join(GOLDEN, "bytecode")
        os.makedirs(dst, exist_ok=True)
        for name in sorted(os.listdir(REFERENCE_FIXTURES)):
            if name.endswith(".json"):
                shutil.copy(os.path.join(REFERENCE_FIXTURES, name), os.path.join(dst, name))


if __name__ == "__main__":
    main()
