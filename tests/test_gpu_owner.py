"""The owner seam on the device (gagp_owner_*): the reference's generation
loops, restated here from run_phase_population_steady
(cpp/src/evolution/evolve.cpp:331-441) and run_phase_population (:253-322),
drive the device population one round or generation at a time -- host-side
ranks, selection values, std::mt19937_64 seeds and victims, device-side
batches, commits and reproduction -- and must reproduce the reference CPU
engine's golden runs: every generation's best fitness, mean fitness, best
wrong count and best program size, and the final best program, byte for
byte."""
import ctypes
import functools
import json
import os

import numpy as np
import pytest

import gagp_amd
from _util import GOLDEN, oracle

pytestmark = pytest.mark.gpu
MANIFEST = json.load(open(os.path.join(GOLDEN, "runs", "manifest.json")))
RUNS = ["polynomial_p256_g20_s123", "polynomial_steady_b128_rs3_p512_g16_s4", "quintic_p512_g16_s3",
        "list_count_gt_rev_p300_g8_s6", "stmt_count_gt_p300_g8_s6", "fit_polynomial_exact_p256_g12_s4",
        "fit_list_sum_even_disc_p512_g10_s5", "init_let_sum_p256_g8_s3"]


class Words:
    """std::mt19937_64(seed) outputs, in order (the oracle's generator)."""

    def __init__(self, seed, n):
        self.out = (ctypes.c_uint64 * n)()
        oracle().orc_mt19937_64(seed, self.out, n)
        self.k = 0

    def __call__(self):
        x = self.out[self.k]
        self.k += 1
        return x


def drive(name):
    """run_phase_population_steady over a gagp_amd.Owner; returns the
    history rows and the final best (fitness, wrong, program)."""
    m = MANIFEST[name]
    gold = json.load(open(os.path.join(GOLDEN, "runs", name, "result.json")))
    search = gold["search"]
    P, G, seed = m["population"], m["generations"], m["seed"]
    owner = gagp_amd.Owner(os.path.join(GOLDEN, m["task"]), os.path.join(GOLDEN, m["search"]) if "search" in m else None,
                           P, seed, os.path.join(GOLDEN, m["init"]) if "init" in m else None)
    B = owner.B
    assert owner.P == P and B == search["schedule"]["batch"]
    pressure = min(max(search["selection"]["size"], 1), P)
    p_mut = search["mutation"]["probability"]
    rs = search["schedule"]["replacement_size"]
    sites = search["ranking"]["tie_break"] == "grammar_sites"
    rounds, per_epoch = P * G // B, P // B
    rng = Words(seed, rounds * (2 * B * rs + 1) + 4096)

    fitness, wrong, counts = owner.evaluate()
    assert all(c[0] > 0 for c in counts)
    has_wrong = owner.has_wrong
    nodes = owner.node_counts(sites) if has_wrong else [0] * P
    owner.begin_steady(B)

    def better(a, b):
        if fitness[a] != fitness[b]:
            return fitness[a] > fitness[b]
        if has_wrong:
            if wrong[a] != wrong[b]:
                return wrong[a] < wrong[b]
            if nodes[a] != nodes[b]:
                return nodes[a] < nodes[b]
        return a < b

    key = functools.cmp_to_key(lambda a, b: -1 if better(a, b) else (1 if better(b, a) else 0))
    order = sorted(range(P), key=key)
    selection = list(fitness)

    def ranks():
        if has_wrong:
            for r, member in enumerate(order):
                selection[member] = float(P - r)

    ranks()
    best = order[0]
    history = []

    def record():
        b = order[0]
        total = np.longdouble(0)
        for f in fitness:
            total += np.longdouble(f)
        program = owner.export_member(b)
        history.append({"best_fitness": fitness[b], "mean_fitness": float(total / np.longdouble(P)),
                        "best_wrong": wrong[b] if has_wrong else -1, "best_nodes": len(program["nodes"])})

    def victim():
        while True:
            worst = rng() % P
            for _ in range(1, rs):
                c = rng() % P
                if better(worst, c):
                    worst = c
            if worst != best:
                return worst

    for rnd in range(rounds):
        if rnd % per_epoch == 0:
            record()
        reuse = owner.make_batch(selection, pressure, p_mut, rng())
        scores, scored_wrong, scored_counts = owner.evaluate_batch()
        child_nodes = owner.batch_nodes(sites) if has_wrong else [0] * B
        placed = [-1] * P
        child_fitness = [fitness[s] if s >= 0 else scores[b] for b, s in enumerate(reuse)]
        child_wrong = [(wrong[s] if s >= 0 else scored_wrong[b]) if has_wrong else 0 for b, s in enumerate(reuse)]
        child_count = [(nodes[s] if s >= 0 else child_nodes[b]) if has_wrong else 0 for b, s in enumerate(reuse)]
        assert all(scored_counts[b][0] > 0 for b, s in enumerate(reuse) if s < 0)
        for b in range(B):
            v = victim()
            fitness[v] = child_fitness[b]
            if has_wrong:
                wrong[v] = child_wrong[b]
            nodes[v] = child_count[b]
            placed[v] = b
            if not has_wrong:
                selection[v] = fitness[v]
            if better(v, best):
                best = v
        order.sort(key=key)
        assert order[0] == best
        ranks()
        owner.commit_batch([(mbr, placed[mbr]) for mbr in range(P) if placed[mbr] >= 0])
    if len(history) < G:
        record()
    b = order[0]
    return history, {"fitness": fitness[b], "wrong": wrong[b] if has_wrong else -1, "program": owner.export_member(b)}


GENERATIONAL = ["polynomial_gen_e1_p256_g12_s3", "poly_c1_gen_e3r_p512_g24_s9", "list_count_gt_rev_gen_e1_p256_g8_s2",
                "stmt_count_gt_gen_e1_p256_g8_s2", "fit_poly_c1_disc_gen_e3r_p256_g10_s6", "init_poly_c1_gen_e3r_p512_g12_s4",
                "polynomial_gen_e40_p256_g10_s7"]


def mean(fitness):
    total = np.longdouble(0)
    for f in fitness:
        total += np.longdouble(f)
    return float(total / np.longdouble(len(fitness)))


def drive_generational(name):
    """run_phase_population over a gagp_amd.Owner."""
    m = MANIFEST[name]
    gold = json.load(open(os.path.join(GOLDEN, "runs", name, "result.json")))
    search = gold["search"]
    P, G, seed = m["population"], m["generations"], m["seed"]
    owner = gagp_amd.Owner(os.path.join(GOLDEN, m["task"]), os.path.join(GOLDEN, m["search"]) if "search" in m else None,
                           P, seed, os.path.join(GOLDEN, m["init"]) if "init" in m else None)
    pressure = min(max(search["selection"]["size"], 1), P)
    p_mut, elites = search["mutation"]["probability"], search["elitism"]
    sites = search["ranking"]["tie_break"] == "grammar_sites"
    rng = Words(seed, G + 16)
    state = {}

    def evaluate():
        fitness, wrong, counts = owner.evaluate()
        has_wrong = owner.has_wrong
        nodes = owner.node_counts(sites) if has_wrong else []

        def better(a, b):
            if fitness[a] != fitness[b]:
                return fitness[a] > fitness[b]
            if has_wrong:
                if wrong[a] != wrong[b]:
                    return wrong[a] < wrong[b]
                if nodes[a] != nodes[b]:
                    return nodes[a] < nodes[b]
            return a < b

        ranking = sorted(range(P), key=functools.cmp_to_key(lambda a, b: -1 if better(a, b) else (1 if better(b, a) else 0)))
        selection = list(fitness)
        if has_wrong:
            for r, member in enumerate(ranking):
                selection[member] = float(P - r)
        state.update(fitness=fitness, wrong=wrong if has_wrong else None, ranking=ranking, selection=selection,
                     counts=counts)

    history = []
    for gen in range(G):
        evaluate()
        b = state["ranking"][0]
        history.append({"best_fitness": state["fitness"][b], "mean_fitness": mean(state["fitness"]),
                        "best_wrong": state["wrong"][b] if state["wrong"] is not None else -1,
                        "best_nodes": len(owner.export_member(b)["nodes"])})
        owner.reproduce(state["selection"], pressure, p_mut, rng(), elites)
    evaluate()
    # reused members (the elites at least) were not scored again
    assert sum(1 for c in state["counts"] if c[0] == 0) >= elites
    b = state["ranking"][0]
    return history, {"fitness": state["fitness"][b], "wrong": state["wrong"][b] if state["wrong"] is not None else -1,
                     "program": owner.export_member(b)}


@pytest.mark.parametrize("name", GENERATIONAL)
def test_owner_generational_loop_reproduces_reference_run(name):
    print("owner generational", name, flush=True)
    history, best = drive_generational(name)
    gold = json.load(open(os.path.join(GOLDEN, "runs", name, "result.json")))
    want = [{k: row[k] for k in ("best_fitness", "mean_fitness", "best_wrong", "best_nodes")} for row in gold["history"]]
    assert history == want
    assert best["fitness"] == gold["best"]["fitness"] and best["wrong"] == gold["best"]["wrong"]
    assert best["program"] == gold["best"]["program"]


@pytest.mark.parametrize("name", RUNS)
def test_owner_loop_reproduces_reference_run(name):
    print("owner", name, flush=True)
    history, best = drive(name)
    gold = json.load(open(os.path.join(GOLDEN, "runs", name, "result.json")))
    want = [{k: row[k] for k in ("best_fitness", "mean_fitness", "best_wrong", "best_nodes")} for row in gold["history"]]
    assert history == want
    assert best["fitness"] == gold["best"]["fitness"] and best["wrong"] == gold["best"]["wrong"]
    assert best["program"] == gold["best"]["program"]


def test_owner_rejects_out_of_order_calls():
    m = MANIFEST["polynomial_p100_g10_s5"]
    owner = gagp_amd.Owner(os.path.join(GOLDEN, m["task"]), None, m["population"], m["seed"])
    with pytest.raises(gagp_amd.GagpError):
        owner.evaluate_batch()  # no batch made
    with pytest.raises(gagp_amd.GagpError):
        owner.begin_steady(owner.B + 1)  # not the configuration's batch
    owner.make_batch([0.0] * owner.P, 2, 0.3, 1)
    with pytest.raises(gagp_amd.GagpError):
        owner.commit_batch([(0, 0)])  # not evaluated
    owner.evaluate_batch()
    with pytest.raises(gagp_amd.GagpError):
        owner.commit_batch([(owner.P, 0)])  # member outside the population
    owner.commit_batch([(0, 0)])
    with pytest.raises(gagp_amd.GagpError):
        owner.reproduce([0.0] * owner.P, 2, 0.3, 1, 1)  # a steady-state search
