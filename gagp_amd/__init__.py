"""MI355X-native GAGP search engine.

The engine is C++17 + HIP (gfx950) behind the C ABI declared in
include/gagp_hip.h and built into ``gagp_amd/libgagp_hip.so`` (and the
``gagp_amd/bin/gagp`` command line). This module is a thin ctypes binding of
that ABI for tests and benchmarks; it never computes anything itself and
raises when the native library is missing.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgagp_hip.so")
CLI_PATH = os.path.join(HERE, "bin", "gagp")

GAGP_OK, GAGP_ERR_INVALID, GAGP_ERR_UNSUPPORTED, GAGP_ERR_DEVICE = 0, 1, 2, 3
TAG_INT, TAG_FLOAT, TAG_BOOL, TAG_INT_LIST, TAG_INVALID = 0, 1, 2, 5, 9


class Value(ctypes.Structure):
    _fields_ = [("i", ctypes.c_int64), ("b", ctypes.c_uint8), ("tag", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 6)]


class Instr(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint8), ("has_a", ctypes.c_uint8), ("has_b", ctypes.c_uint8), ("pad", ctypes.c_uint8),
                ("a", ctypes.c_int32), ("b", ctypes.c_int32)]


class Program(ctypes.Structure):
    _fields_ = [("code", ctypes.POINTER(Instr)), ("code_len", ctypes.c_int32), ("fuel", ctypes.POINTER(ctypes.c_uint32)),
                ("consts", ctypes.POINTER(Value)), ("n_consts", ctypes.c_int32), ("n_locals", ctypes.c_int32)]


class Policy(ctypes.Structure):
    _fields_ = [("metric", ctypes.c_int32), ("proximity", ctypes.c_int32), ("summary_count", ctypes.c_int32),
                ("summary_weight", ctypes.c_double), ("summary_cap", ctypes.c_double)]


class GagpError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("gagp status %d: %s" % (status, message))
        self.status = status


_lib = None


def build(jobs: int = 8) -> None:
    """Compiles the HIP engine for gfx950 in place (make -C gagp_amd)."""
    subprocess.run(["make", "-C", HERE, "-j%d" % jobs], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("native engine not built: %s is missing (run gagp_amd.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.gagp_last_error.restype = ctypes.c_char_p
    L.gagp_device_count.restype = ctypes.c_int
    L.gagp_session_create.argtypes = [ctypes.POINTER(Value), ctypes.POINTER(ctypes.c_uint8), ctypes.c_int32, ctypes.c_int32,
                                      ctypes.POINTER(Value), ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_double, Policy, ctypes.POINTER(ctypes.c_void_p)]
    L.gagp_session_eval.argtypes = [ctypes.c_void_p, ctypes.POINTER(Program), ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    L.gagp_session_eval_static.argtypes = L.gagp_session_eval.argtypes
    L.gagp_session_eval_cases.argtypes = [ctypes.c_void_p, ctypes.POINTER(Program), ctypes.POINTER(Value),
                                          ctypes.POINTER(ctypes.c_int32)]
    L.gagp_session_last_kernel_ms.argtypes = [ctypes.c_void_p]
    L.gagp_session_last_kernel_ms.restype = ctypes.c_float
    L.gagp_session_destroy.argtypes = [ctypes.c_void_p]
    L.gagp_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    L.gagp_batch_add.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                 ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32)]
    L.gagp_batch_add_init.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_uint64, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32)]
    L.gagp_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.gagp_batch_planned_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]
    L.gagp_batch_result_json.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_size_t)]
    L.gagp_batch_prepare.argtypes = [ctypes.c_void_p]
    L.gagp_batch_step.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    L.gagp_batch_finish.argtypes = [ctypes.c_void_p]
    L.gagp_batch_max_rounds.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
    L.gagp_server_create.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    L.gagp_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_int32)]
    L.gagp_server_run.argtypes = [ctypes.c_void_p]
    L.gagp_server_set_max_live.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.gagp_server_outcome.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.gagp_server_result_json.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]
    L.gagp_server_destroy.argtypes = [ctypes.c_void_p]
    L.gagp_server_destroy.restype = None
    L.gagp_batch_counters.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]
    L.gagp_batch_evaluated_program_cases.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]
    L.gagp_batch_set_profile.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.gagp_batch_kernel_times_json.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.gagp_batch_set_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.gagp_valu_probe.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]
    L.gagp_batch_tuning_json.argtypes = L.gagp_batch_kernel_times_json.argtypes
    L.gagp_batch_destroy.argtypes = [ctypes.c_void_p]
    vp, i32, u64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_double
    P = ctypes.POINTER
    L.gagp_owner_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32, i32, u64, ctypes.c_char_p, P(vp)]
    for name in ("population_size", "batch_size", "has_wrong"):
        getattr(L, "gagp_owner_" + name).argtypes = [vp, P(i32)]
    L.gagp_owner_evaluate.argtypes = [vp, P(dbl), P(ctypes.c_uint32), P(u64)]
    L.gagp_owner_evaluate_batch.argtypes = L.gagp_owner_evaluate.argtypes
    L.gagp_owner_node_counts.argtypes = [vp, i32, P(ctypes.c_uint32)]
    L.gagp_owner_batch_nodes.argtypes = L.gagp_owner_node_counts.argtypes
    L.gagp_owner_begin_steady.argtypes = [vp, i32]
    L.gagp_owner_make_batch.argtypes = [vp, P(dbl), i32, dbl, u64, P(i32)]
    L.gagp_owner_reproduce.argtypes = [vp, P(dbl), i32, dbl, u64, i32]
    L.gagp_owner_commit_batch.argtypes = [vp, P(i32), i32]
    L.gagp_owner_export_member.argtypes = [vp, i32, ctypes.c_char_p, ctypes.c_size_t, P(ctypes.c_size_t)]
    L.gagp_owner_destroy.argtypes = [vp]
    L.gagp_owner_destroy.restype = None
    _lib = L
    return L


def _check(status):
    if status != GAGP_OK:
        raise GagpError(status, lib().gagp_last_error().decode())


def _text(fn, *args):
    needed = ctypes.c_size_t(0)
    fn(*args, None, 0, ctypes.byref(needed))
    buf = ctypes.create_string_buffer(needed.value)
    _check(fn(*args, buf, needed.value, ctypes.byref(needed)))
    return buf.value.decode()


def make_value(v) -> Value:
    x = Value()
    if isinstance(v, bool):
        x.tag, x.b = TAG_BOOL, int(v)
    elif isinstance(v, int):
        x.tag, x.i = TAG_INT, v
    elif isinstance(v, float):
        x.tag, x.i = TAG_FLOAT, ctypes.c_int64.from_buffer_copy(ctypes.c_double(v)).value
    elif isinstance(v, tuple) and v[0] == "list":
        x.tag, x.i = TAG_INT_LIST, v[1] * 65536 + v[2]
    elif v is None:
        x.tag = TAG_INVALID
    else:
        raise TypeError(v)
    return x


class Session:
    """FitnessSessionGpu-style evaluation of bytecode programs on a fixed case set."""

    def __init__(self, inputs, answers, fuel=1000, lanes=512, penalty=1000.0, metric=1, proximity=0, arena=(),
                 input_set=None):
        L = lib()
        n_cases = len(answers)
        n_inputs = len(inputs[0]) if n_cases and inputs else 0
        flat = (Value * max(1, n_cases * n_inputs))(*[make_value(v) for row in inputs for v in row])
        ans = (Value * n_cases)(*[make_value(v) for v in answers])
        s = None
        if input_set is not None:
            s = (ctypes.c_uint8 * (n_cases * n_inputs))(*[int(b) for row in input_set for b in row])
        ar = (ctypes.c_int64 * max(1, len(arena)))(*arena)
        self.h = ctypes.c_void_p()
        _check(L.gagp_session_create(flat, s, n_cases, n_inputs, ans, ar, len(arena), fuel, lanes, penalty,
                                     Policy(metric, proximity, 0, 0.1, 100.0), ctypes.byref(self.h)))
        self.n_cases = n_cases

    @staticmethod
    def _program(p, keep):
        code = (Instr * len(p["code"]))()
        for k, ins in enumerate(p["code"]):
            op, a, b = ins
            code[k].op = op
            code[k].has_a = a is not None
            code[k].has_b = b is not None
            code[k].a = a or 0
            code[k].b = b or 0
        consts = (Value * max(1, len(p["consts"])))(*[make_value(v) for v in p["consts"]])
        fuel = None
        if p.get("fuel") is not None:
            fuel = (ctypes.c_uint32 * len(p["fuel"]))(*p["fuel"])
        keep.extend([code, consts, fuel])
        return Program(code, len(p["code"]), fuel, consts, len(p["consts"]), p["n_locals"])

    def eval(self, programs, allocate=True):
        """programs: dicts {"code": [(op, a|None, b|None)], "consts": [...], "n_locals": n, "fuel": [...]|None}.
        Returns (fitness list, wrong list, counts list of (cases, errors, timeouts)).
        allocate=False calls the allocation-free entry point (gagp_session_eval_static)."""
        keep = []
        arr = (Program * len(programs))(*[self._program(p, keep) for p in programs])
        n = len(programs)
        fit = (ctypes.c_double * n)()
        wrong = (ctypes.c_uint32 * n)()
        counts = (ctypes.c_uint32 * (3 * n))()
        entry = lib().gagp_session_eval if allocate else lib().gagp_session_eval_static
        _check(entry(self.h, arr, n, fit, wrong, counts))
        return list(fit), list(wrong), [tuple(counts[3 * i:3 * i + 3]) for i in range(n)]

    def eval_cases(self, program):
        keep = []
        p = self._program(program, keep)
        values = (Value * self.n_cases)()
        errors = (ctypes.c_int32 * self.n_cases)()
        _check(lib().gagp_session_eval_cases(self.h, ctypes.byref(p), values, errors))
        return [(values[i].tag, values[i].i, values[i].b) for i in range(self.n_cases)], list(errors)

    def last_kernel_ms(self):
        return lib().gagp_session_last_kernel_ms(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().gagp_session_destroy(self.h)
            self.h = None


def valu_probe(blocks, iterations):
    """Device milliseconds of the int32 VALU probe (8 * iterations * 256 *
    blocks lane operations)."""
    ms = ctypes.c_float()
    _check(lib().gagp_valu_probe(blocks, iterations, ctypes.byref(ms)))
    return ms.value


class Batch:
    """Searches stepped together on the device (`gagp run --engine gpu`)."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        _check(lib().gagp_batch_create(ctypes.byref(self.h)))

    def add(self, task, search=None, population=0, generations=0, seed=None, init=None):
        """Admits one search; `init` is a population document (`gagp run
        --init`) whose first programs are the initial population."""
        idx = ctypes.c_int32(-1)
        _check(lib().gagp_batch_add_init(self.h, task.encode(), search.encode() if search else None, population,
                                         generations, int(seed is not None), seed or 0,
                                         init.encode() if init else None, ctypes.byref(idx)))
        return idx.value

    def planned_bytes(self, index):
        out = ctypes.c_uint64()
        _check(lib().gagp_batch_planned_bytes(self.h, index, ctypes.byref(out)))
        return out.value

    def set_profile(self, on=True):
        _check(lib().gagp_batch_set_profile(self.h, int(on)))

    def run(self, use_graph=True):
        _check(lib().gagp_batch_run(self.h, int(use_graph)))

    def prepare(self):
        _check(lib().gagp_batch_prepare(self.h))

    def step(self, rounds, use_graph=True):
        _check(lib().gagp_batch_step(self.h, rounds, int(use_graph)))

    def max_rounds(self):
        out = ctypes.c_int32()
        _check(lib().gagp_batch_max_rounds(self.h, ctypes.byref(out)))
        return out.value

    def finish(self):
        _check(lib().gagp_batch_finish(self.h))

    def counters(self, index):
        out = (ctypes.c_uint64 * 14)()
        _check(lib().gagp_batch_counters(self.h, index, out))
        names = ("programs", "program_cases", "errors", "timeouts", "vm_instructions", "code_words", "pool_words",
                 "children", "child_nodes", "child_consts", "parent_nodes", "parent_consts", "child_code_words",
                 "child_pool_words")
        return dict(zip(names, out))

    def result(self, index):
        return json.loads(_text(lib().gagp_batch_result_json, self.h, index))

    def evaluated_program_cases(self, index):
        out = ctypes.c_uint64()
        _check(lib().gagp_batch_evaluated_program_cases(self.h, index, ctypes.byref(out)))
        return out.value

    def kernel_times(self):
        return json.loads(_text(lib().gagp_batch_kernel_times_json, self.h))

    def set_autotune(self, mode=1):
        """0 off, 1 when a round evaluates at least 2^22 program-cases (the default), 2 always."""
        _check(lib().gagp_batch_set_autotune(self.h, int(mode)))

    def tuning(self):
        return json.loads(_text(lib().gagp_batch_tuning_json, self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().gagp_batch_destroy(self.h)
            self.h = None


class Owner:
    """The owner seam (gagp_owner_*): one steady-state search on the device
    whose generation loop the caller runs, as the reference's
    run_phase_population_steady drives its GpuPhaseOwner."""

    def __init__(self, task, search=None, population=0, seed=None, init=None):
        self.h = ctypes.c_void_p()
        _check(lib().gagp_owner_create(task.encode(), search.encode() if search else None, population or 0,
                                       int(seed is not None), ctypes.c_uint64(seed or 0), init.encode() if init else None,
                                       ctypes.byref(self.h)))
        self.P = self._int("population_size")
        self.B = self._int("batch_size")
        self.has_wrong = bool(self._int("has_wrong"))

    def __del__(self):
        if getattr(self, "h", None):
            lib().gagp_owner_destroy(self.h)
            self.h = None

    def _int(self, name):
        out = ctypes.c_int32()
        _check(getattr(lib(), "gagp_owner_" + name)(self.h, ctypes.byref(out)))
        return out.value

    def _scores(self, fn, n):
        fit, wrong, counts = (ctypes.c_double * n)(), (ctypes.c_uint32 * n)(), (ctypes.c_uint64 * (5 * n))()
        _check(fn(self.h, fit, wrong, counts))
        return list(fit), list(wrong), [list(counts[5 * i:5 * i + 5]) for i in range(n)]

    def evaluate(self):
        """(fitness[P], wrong[P], counts[P][5]) of the initial population."""
        return self._scores(lib().gagp_owner_evaluate, self.P)

    def node_counts(self, grammar_sites=False):
        out = (ctypes.c_uint32 * self.P)()
        _check(lib().gagp_owner_node_counts(self.h, int(grammar_sites), out))
        return list(out)

    def begin_steady(self, batch):
        _check(lib().gagp_owner_begin_steady(self.h, batch))

    def make_batch(self, selection, pressure, p_mut, seed):
        sel = (ctypes.c_double * self.P)(*selection)
        reuse = (ctypes.c_int32 * self.B)()
        _check(lib().gagp_owner_make_batch(self.h, sel, pressure, float(p_mut), ctypes.c_uint64(seed), reuse))
        return list(reuse)

    def reproduce(self, selection, pressure, p_mut, seed, elites):
        sel = (ctypes.c_double * self.P)(*selection)
        _check(lib().gagp_owner_reproduce(self.h, sel, pressure, float(p_mut), ctypes.c_uint64(seed), elites))

    def evaluate_batch(self):
        return self._scores(lib().gagp_owner_evaluate_batch, self.B)

    def batch_nodes(self, grammar_sites=False):
        out = (ctypes.c_uint32 * self.B)()
        _check(lib().gagp_owner_batch_nodes(self.h, int(grammar_sites), out))
        return list(out)

    def commit_batch(self, placements):
        flat = [x for pair in placements for x in pair]
        arr = (ctypes.c_int32 * max(1, len(flat)))(*flat)
        _check(lib().gagp_owner_commit_batch(self.h, arr, len(placements)))

    def export_member(self, member):
        return json.loads(_text(lib().gagp_owner_export_member, self.h, member))


class Server:
    """Request loop with continuous batching (gagp_server_*): submit requests
    with arrival offsets, run() serves them in real time."""

    def __init__(self, budget_bytes=64 << 30, poll_rounds=2, max_live=0):
        self.h = ctypes.c_void_p()
        _check(lib().gagp_server_create(budget_bytes, poll_rounds, ctypes.byref(self.h)))
        if max_live:
            _check(lib().gagp_server_set_max_live(self.h, max_live))

    def __del__(self):
        if getattr(self, "h", None):
            lib().gagp_server_destroy(self.h)
            self.h = None

    def submit(self, task, search=None, population=0, generations=0, seed=None, early_exit=False, arrival=0.0):
        t = ctypes.c_int32()
        _check(lib().gagp_server_submit(self.h, task.encode(), search.encode() if search else None, population or 0,
                                        generations or 0, int(seed is not None), ctypes.c_uint64(seed or 0),
                                        int(early_exit), float(arrival), ctypes.byref(t)))
        return t.value

    def run(self):
        _check(lib().gagp_server_run(self.h))

    def outcome(self, ticket):
        times = (ctypes.c_double * 3)()
        gens, status = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().gagp_server_outcome(self.h, ticket, times, ctypes.byref(gens), ctypes.byref(status)))
        return {"arrival": times[0], "admit": times[1], "done": times[2], "generations": gens.value,
                "status": status.value}

    def result(self, ticket):
        return json.loads(_text(lib().gagp_server_result_json, self.h, ticket))
