"""The C-ABI library and command line load and export every declared entry
point (no device calls), and the host-side documents behave."""
import ctypes
import json
import os
import re
import subprocess

import gagp_amd
from _util import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "gagp_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gagp_[a-z_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("gagp_session_create", "gagp_session_eval", "gagp_batch_create", "gagp_batch_add", "gagp_batch_run",
              "gagp_batch_result_json", "gagp_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gagp_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_are_reported_without_a_device():
    L = gagp_amd.lib()
    # Invalid arguments fail on the host with a message, before any device call.
    st = L.gagp_batch_add(None, b"x", None, 0, 0, 0, 0, None)
    assert st == gagp_amd.GAGP_ERR_INVALID and b"invalid" in L.gagp_last_error()


def test_batch_rejects_unsupported_grammar_on_the_host():
    b = gagp_amd.Batch()
    try:
        b.add(os.path.join(GOLDEN, "tasks", "fuel-profile", "task.json"))
    except gagp_amd.GagpError as e:
        assert e.status == gagp_amd.GAGP_ERR_UNSUPPORTED
    else:
        raise AssertionError("fuel-profile has a source fuel profile (fuel_events), which the engine declines")


def test_cli_usage_and_exit_codes():
    r = subprocess.run([gagp_amd.CLI_PATH, "run"], capture_output=True, text=True)
    assert r.returncode == 2 and "missing TASK" in r.stderr
    r = subprocess.run([gagp_amd.CLI_PATH, "run", "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--engine" in r.stdout


def test_workload_generators_are_deterministic(tmp_path):
    from gagp_amd import workloads
    workloads.poly_c1(str(tmp_path / "c1"))
    workloads.quintic(str(tmp_path / "c2"))
    for name, sub in (("poly_c1", "c1"), ("quintic", "c2")):
        for f in ("task.json", "grammar.json", "train.cases.json", "test.cases.json"):
            assert json.load(open(tmp_path / sub / f)) == json.load(open(os.path.join(GOLDEN, "tasks", name, f)))


def test_server_rejects_requests_before_and_after_run_on_the_host():
    s = gagp_amd.Server(poll_rounds=1)
    L = gagp_amd.lib()
    # No outcome before the server has run; bad handles fail with a message.
    times = (ctypes.c_double * 3)()
    st = L.gagp_server_outcome(s.h, 0, times, None, None)
    assert st == gagp_amd.GAGP_ERR_INVALID and b"no such request" in L.gagp_last_error()
    assert L.gagp_server_submit(None, b"x", None, 0, 0, 0, 0, 0, 0.0, None) == gagp_amd.GAGP_ERR_INVALID
    assert L.gagp_server_run(None) == gagp_amd.GAGP_ERR_INVALID


def test_planned_footprint_grows_with_population_and_cases_on_the_host():
    # The footprint is computed from the documents alone (no device calls).
    b = gagp_amd.Batch()
    task = os.path.join(GOLDEN, "tasks", "quintic", "task.json")
    small = b.planned_bytes(b.add(task, None, 512, 8, 0))
    large = b.planned_bytes(b.add(task, None, 4096, 8, 0))
    longer = b.planned_bytes(b.add(task, None, 512, 128, 0))
    assert 1 << 20 <= small < large and small < longer
    # Rows of 60 nodes, 24 B per node for members and 72 B per child row, the
    # bytecode rows and the hash table: at least 100 B per node at P = 4096.
    assert large >= 4096 * 60 * 100


def test_server_fails_bad_requests_alone_on_the_host():
    # Every request fails before any device call: an unreadable task, a
    # grammar the engine declines, and a search whose footprint exceeds the
    # whole budget. run() itself succeeds and each ticket carries its status.
    s = gagp_amd.Server(budget_bytes=1 << 20, poll_rounds=1)
    bad = s.submit(os.path.join(GOLDEN, "tasks", "no-such-task.json"))
    declined = s.submit(os.path.join(GOLDEN, "tasks", "fuel-profile", "task.json"))
    big = s.submit(os.path.join(GOLDEN, "tasks", "quintic", "task.json"), None, 4096, 8, 0)
    s.run()
    assert s.outcome(bad)["status"] == gagp_amd.GAGP_ERR_INVALID
    assert s.outcome(declined)["status"] == gagp_amd.GAGP_ERR_UNSUPPORTED
    assert s.outcome(big)["status"] == gagp_amd.GAGP_ERR_INVALID
    with __import__("pytest").raises(gagp_amd.GagpError, match="above the serving budget"):
        s.result(big)
