"""Shared helpers of the test suite: the oracle binding and result comparison."""
import ctypes
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")

E_NAMES = {-1: None, 0: "NameError", 1: "TypeError", 2: "ZeroDivisionError", 3: "ValueError", 4: "Timeout"}
OPCODES = {"PUSH_CONST": 0, "LOAD": 1, "STORE": 2, "NEG": 3, "NOT": 4, "ADD": 5, "SUB": 6, "MUL": 7, "DIV": 8, "MOD": 9,
           "LT": 10, "LE": 11, "GT": 12, "GE": 13, "EQ": 14, "NE": 15, "JMP": 16, "JMP_IF_FALSE": 17, "JMP_IF_TRUE": 18,
           "CALL_BUILTIN": 19, "RETURN": 20, "CHECK_LIST": 21, "CHECK_INT": 22}


class OVal(ctypes.Structure):
    _fields_ = [("i", ctypes.c_int64), ("b", ctypes.c_uint8), ("tag", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 6)]


class OIns(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint8), ("has_a", ctypes.c_uint8), ("has_b", ctypes.c_uint8), ("pad", ctypes.c_uint8),
                ("a", ctypes.c_int32), ("b", ctypes.c_int32)]


class OProg(ctypes.Structure):
    _fields_ = [("code", ctypes.POINTER(OIns)), ("code_len", ctypes.c_int32), ("fuel", ctypes.POINTER(ctypes.c_uint32)),
                ("consts", ctypes.POINTER(OVal)), ("n_consts", ctypes.c_int32), ("n_locals", ctypes.c_int32)]


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_LIB)
        L.orc_run.argtypes = [ctypes.POINTER(OProg), ctypes.POINTER(OVal), ctypes.POINTER(ctypes.c_uint8), ctypes.c_int,
                              ctypes.POINTER(ctypes.c_int64), ctypes.c_int, ctypes.POINTER(OVal), ctypes.POINTER(ctypes.c_int)]
        L.orc_eval_fitness.argtypes = [ctypes.POINTER(OProg), ctypes.c_int, ctypes.POINTER(OVal), ctypes.POINTER(ctypes.c_uint8),
                                       ctypes.c_int, ctypes.c_int, ctypes.POINTER(OVal), ctypes.POINTER(ctypes.c_int64),
                                       ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.POINTER(ctypes.c_uint32)]
        L.orc_canon.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_canon.restype = ctypes.c_double
        L.orc_splitmix_next.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
        L.orc_splitmix_bounded.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
        L.orc_mt19937_64.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
        _oracle = L
    return _oracle


def oval(v):
    """Python value -> OVal: int, bool, float, ('list', offset, length), None (invalid)."""
    x = OVal()
    if isinstance(v, bool):
        x.tag, x.b = 2, int(v)
    elif isinstance(v, int):
        x.tag, x.i = 0, v
    elif isinstance(v, float):
        x.tag = 1
        x.i = ctypes.c_int64.from_buffer_copy(ctypes.c_double(v)).value
    elif isinstance(v, tuple) and v[0] == "list":
        x.tag, x.i = 5, v[1] * 65536 + v[2]
    elif v is None:
        x.tag = 9
    else:
        raise TypeError(v)
    return x


def oprog(p, keep):
    """p: {"code": [(op, a|None, b|None)], "consts": [...], "n_locals": n, "fuel": list|None}."""
    code = (OIns * len(p["code"]))()
    for k, (op, a, b) in enumerate(p["code"]):
        code[k].op, code[k].has_a, code[k].has_b = op, a is not None, b is not None
        code[k].a, code[k].b = a or 0, b or 0
    consts = (OVal * max(1, len(p["consts"])))(*[oval(v) for v in p["consts"]])
    fuel = (ctypes.c_uint32 * len(p["fuel"]))(*p["fuel"]) if p.get("fuel") is not None else None
    keep.extend([code, consts, fuel])
    return OProg(code, len(p["code"]), fuel, consts, len(p["consts"]), p["n_locals"])


def oracle_run(p, inputs, arena=(), fuel=1000):
    keep = []
    prog = oprog(p, keep)
    vals = (OVal * max(1, len(inputs)))(*[oval(v) for v in inputs])
    sets = (ctypes.c_uint8 * max(1, len(inputs)))(*[0 if v is None else 1 for v in inputs])
    ar = (ctypes.c_int64 * max(1, len(arena)))(*arena)
    out = OVal()
    left = ctypes.c_int()
    err = oracle().orc_run(ctypes.byref(prog), vals, sets, len(inputs), ar, fuel, ctypes.byref(out), ctypes.byref(left))
    return err, out, left.value


def oracle_fitness(programs, inputs, answers, arena=(), fuel=1000, penalty=1000.0, lanes=512, metric=1, proximity=0):
    keep = []
    progs = (OProg * len(programs))(*[oprog(p, keep) for p in programs])
    n_cases, n_inputs = len(answers), (len(inputs[0]) if inputs else 0)
    flat = (OVal * max(1, n_cases * n_inputs))(*[oval(v) for row in inputs for v in row])
    sets = (ctypes.c_uint8 * max(1, n_cases * n_inputs))(*[0 if v is None else 1 for row in inputs for v in row])
    ans = (OVal * n_cases)(*[oval(v) for v in answers])
    ar = (ctypes.c_int64 * max(1, len(arena)))(*arena)
    n = len(programs)
    fit = (ctypes.c_double * n)()
    wrong = (ctypes.c_uint32 * n)()
    counts = (ctypes.c_uint32 * (3 * n))()
    oracle().orc_eval_fitness(progs, n, flat, sets, n_cases, n_inputs, ans, ar, fuel, penalty, lanes, metric, proximity,
                              fit, wrong, counts)
    return list(fit), list(wrong), [tuple(counts[3 * i:3 * i + 3]) for i in range(n)]


def load_fixture_scenarios(names=("scalar.json", "control_flow.json", "builtins.json", "typed_values.json")):
    """Scenarios of the reference's bytecode fixtures restricted to the value
    model of the oracle and the device VM (Int, Float, Bool scalars and
    IntList inputs; no payload-constructing builtins)."""
    supported_types = {"int", "float", "bool", "int_list"}
    out, skipped = [], []
    for name in names:
        doc = json.load(open(os.path.join(GOLDEN, "bytecode", name)))
        fuel = doc["fuel"]
        for s in doc["scenarios"]:
            prog = s["program"]
            ok = all(c["type"] in {"int", "float", "bool"} for c in prog["consts"])
            ok = ok and all(ins["op"] in OPCODES for ins in prog["code"])
            builtins = {ins.get("a") for ins in prog["code"] if ins["op"] == "CALL_BUILTIN"}
            ok = ok and builtins <= {0, 1, 2, 3, 4, 7, 12, 13, 14, 28, 29, 30, 31}
            code = [(OPCODES[ins["op"]], ins.get("a"), ins.get("b")) for ins in prog["code"]]
            consts = [scalar(c) for c in prog["consts"]] if ok else []
            cases = []
            for c in s["cases"]:
                types = [i["value"]["type"] for i in c["inputs"]]
                exp_type = c["expected"]["type"] if "expected" in c else None
                if not ok or not set(types) <= supported_types or (exp_type and exp_type not in {"int", "float", "bool"}):
                    skipped.append((name, s["intent"]))
                    continue
                cases.append(c)
            if ok and cases:
                out.append({"file": name, "intent": s["intent"], "fuel": s.get("fuel", fuel),
                            "program": {"code": code, "consts": consts, "n_locals": prog["n_locals"],
                                        "fuel": prog.get("instruction_fuel")},
                            "cases": cases})
    return out, skipped


def scalar(tv):
    t, v = tv["type"], tv["value"]
    if t == "int":
        return int(v)
    if t == "bool":
        return bool(v)
    if t == "float":
        return float(v)
    raise TypeError(t)


def case_inputs(case, n_locals):
    """Fixture case -> (locals list, arena)."""
    vals = [None] * n_locals
    arena = []
    for i in case["inputs"]:
        tv = i["value"]
        if i["idx"] >= n_locals:
            continue
        if tv["type"] == "int_list":
            off = len(arena)
            arena.extend(int(x) for x in tv["value"])
            vals[i["idx"]] = ("list", off, len(tv["value"]))
        else:
            vals[i["idx"]] = scalar(tv)
    return vals, arena


def strip_timing(result):
    """result.json members that must equal the reference engine's."""
    out = {k: result[k] for k in ("search", "best", "test", "population", "generations", "seed", "train_cases", "format")
           if k in result}
    out["history"] = [{k: v for k, v in h.items() if k not in ("elapsed_ms", "eval_ms", "repro_ms")}
                      for h in result["history"]]
    return out
