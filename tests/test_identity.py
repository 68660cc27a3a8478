"""Grammar identity on the host (CPU): SHA-256 against the FIPS 180-2
example digests, and the grammar content hash against the grammar_hash the
reference engine wrote into the population.json of every --save-all golden
run (cpp/src/cli/run.cpp:364-365)."""
import hashlib
import json
import os
import subprocess

import pytest

from _util import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "gagp_amd", "_build")
MANIFEST = json.load(open(os.path.join(GOLDEN, "runs", "manifest.json")))
SAVED = sorted(n for n in MANIFEST if MANIFEST[n]["save_all"])


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    obj = os.path.join(BUILD, "identity.o")
    assert os.path.exists(obj), "build the engine first (__graft_entry__.build())"
    exe = str(tmp_path_factory.mktemp("identity") / "identity_driver")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "gagp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "identity_driver.cpp"), obj, "-o", exe], check=True)
    return exe


def sha(driver, data):
    return subprocess.run([driver, "sha"], input=data, capture_output=True, check=True).stdout.decode().strip()


@pytest.mark.parametrize("message,digest", [
    (b"abc", "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"),
    (b"", "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"),
    (b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq",
     "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"),
])
def test_sha256_known_answers(driver, message, digest):
    assert sha(driver, message) == digest


def test_sha256_padding_boundaries(driver):
    # Message lengths around the one- and two-block padding cases, checked
    # against Python's hashlib.
    for n in (55, 56, 57, 63, 64, 65, 119, 120, 127, 128, 1000):
        data = bytes((7 * i + 3) & 255 for i in range(n))
        assert sha(driver, data) == hashlib.sha256(data).hexdigest(), n


@pytest.mark.parametrize("name", SAVED)
def test_grammar_hash_matches_reference(driver, name):
    m = MANIFEST[name]
    task_path = os.path.join(GOLDEN, m["task"])
    task = json.load(open(task_path))
    grammar = os.path.join(os.path.dirname(task_path), task["grammar"])
    want = json.load(open(os.path.join(GOLDEN, "runs", name, "population.json")))["grammar_hash"]
    got = subprocess.run([driver, "grammar", grammar], capture_output=True, check=True, text=True).stdout.strip()
    assert got == want
