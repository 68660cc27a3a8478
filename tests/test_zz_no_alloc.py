"""The allocation-free evaluation entry point (gagp_session_eval_static)
against the allocating one (gagp_session_eval) through the C ABI: identical
fitness, wrong counts and per-program counts on two calls in a row, the
second with fewer programs and cases, and a launch above the fixed
workspace rejected before it reaches the device."""
import random

import pytest

import gagp_amd
from _util import oracle_fitness
from test_gpu_vm import random_program

pytestmark = pytest.mark.gpu


def both(session, programs):
    allocating = session.eval(programs, allocate=True)
    fixed = session.eval(programs, allocate=False)
    return allocating, fixed


def test_allocation_free_eval_matches_allocating_eval():
    rng = random.Random(31)
    # first call: 300 programs on 777 cases, discounted proximity (fractional
    # scores, so the canonical order of the reduction decides every total)
    programs = [random_program(rng, rng.randrange(1, 6)) for _ in range(300)]
    inputs = [[rng.randrange(-50, 51), rng.randrange(-9, 10)] for _ in range(777)]
    answers = [rng.randrange(-200, 200) for _ in range(777)]
    print("no-alloc: 300 programs x 777 cases, discounted", flush=True)
    s = gagp_amd.Session(inputs, answers, penalty=1000.0, lanes=512, proximity=1)
    allocating, fixed = both(s, programs)
    assert fixed == allocating
    assert fixed[0] == oracle_fitness(programs, inputs, answers, penalty=1000.0, lanes=512, proximity=1)[0]
    # second call, smaller: 40 programs on 65 cases of another session (the
    # workspace keeps the first call's larger contents)
    small = [random_program(rng, rng.randrange(1, 5)) for _ in range(40)]
    inputs2 = inputs[:65]
    answers2 = answers[:65]
    print("no-alloc: 40 programs x 65 cases, absolute", flush=True)
    s2 = gagp_amd.Session(inputs2, answers2, penalty=1000.0, lanes=512)
    allocating, fixed = both(s2, small)
    assert fixed == allocating
    assert fixed[0] == oracle_fitness(small, inputs2, answers2, penalty=1000.0, lanes=512)[0]
    # and the first session again, after the smaller launch
    print("no-alloc: first session again", flush=True)
    again = s.eval(programs[:7], allocate=False)
    assert again == s.eval(programs[:7], allocate=True)


def test_allocation_free_eval_rejects_launches_above_the_workspace():
    # 530 programs of 4002 instructions need about 17 MB: above the 16 MiB
    # workspace, refused on the host with an invalid-argument status.
    program = {"code": [(0, 0, None)] + [(3, None, None)] * 4000 + [(20, None, None)], "consts": [3], "n_locals": 1}
    s = gagp_amd.Session([[1]], [0], penalty=1000.0)
    print("no-alloc: oversized launch", flush=True)
    with pytest.raises(gagp_amd.GagpError) as e:
        s.eval([program] * 530, allocate=False)
    assert e.value.status == gagp_amd.GAGP_ERR_INVALID
    assert "workspace" in str(e.value)
    # the session still evaluates within the bound
    assert s.eval([program], allocate=False) == s.eval([program], allocate=True)
