"""Facts the wavefront operators rely on, checked on the CPU.

- Crossover's tournaments draw on all lanes at once (phase_wave.hpp
  tournaments_w): draw k of a SplitMix64 stream must equal
  splitmix_mix(state + k * increment), and bounded() must take exactly one
  draw whenever that draw is not below (2^64 - n) mod n. Checked against the
  engine's own SplitMix (core.hpp) compiled for the host.
- The tile walk (phase_wave.hpp walk_tile_w) replaces the node-by-node frame
  walk (WaveWalk) for rows of at most 64 nodes: with q_i the open child
  slots before node i, node j < i is an ancestor of i exactly when
  min_{j<k<=i} q_k >= q_j. Checked exhaustively on short arity sequences
  (valid and malformed) and on random long ones against a direct simulation
  of the frame walk: the same depth and incoming site depth at every node,
  and the same failures (no open slot, a bad arity, children at depth 64)."""
import itertools
import os
import random
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAX_ARITY = 8  # grammar.hpp kMaxArity

SPLITMIX_CHECK = r"""
#include <cstdio>
#include <initializer_list>
#include "core.hpp"
using namespace gagp_amd;
int main() {
  const std::uint64_t seeds[] = {0ull, 1ull, 0x617089e168ull, 0xfffffffffffffff0ull, 0x9e3779b97f4a7c15ull};
  int bad = 0;
  for (std::uint64_t seed : seeds) {
    SplitMix r(seed);
    for (int k = 0; k < 128; ++k) {
      const std::uint64_t direct = splitmix_mix(seed + std::uint64_t(k) * 0x9e3779b97f4a7c15ULL);
      if (r.next() != direct) ++bad;
    }
    // bounded(n): one draw when it is not below the rejection threshold
    for (std::uint64_t n : {1ull, 2ull, 3ull, 836ull, 4096ull, 15670ull, 1000003ull}) {
      SplitMix a(seed), b(seed);
      const std::uint64_t threshold = (0ULL - n) % n;
      const std::uint64_t v = splitmix_mix(seed);
      if (v >= threshold) {
        if (a.bounded(n) != v % n) ++bad;
        b.next();
        if (a.next() != b.next()) ++bad;
      }
      if ((n & (n - 1)) == 0 && threshold != 0) ++bad;  // powers of two never reject
    }
  }
  std::printf("%d\n", bad);
  return bad != 0;
}
"""


def test_splitmix_draws_are_counter_based(tmp_path):
    src = tmp_path / "splitmix_check.cpp"
    src.write_text(SPLITMIX_CHECK)
    exe = tmp_path / "splitmix_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "gagp_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout


def frame_walk(arity, flags, n):
    """The serial walk (WaveWalk) over nodes [0, n): per node the depth and
    incoming site depth before visiting it (also at node n when it exists),
    and whether the walk fails; `open` tells whether slots remain after n."""
    pend, inc = [1], [0]
    depth, incoming = [], []
    ok = True
    for i in range(min(n + 1, len(arity))):
        while pend and pend[-1] == 0:
            pend.pop()
            inc.pop()
        if i == n:
            if not pend:
                ok = False
            depth.append(len(pend))
            incoming.append(inc[-1] if inc else 0)
            break
        if not pend:
            ok = False
            break
        depth.append(len(pend))
        incoming.append(inc[-1])
        a = arity[i]
        if a < 0 or a > MAX_ARITY:
            ok = False
            break
        pend[-1] -= 1
        if a:
            if len(pend) == 64:
                ok = False
                break
            value = inc[-1] + flags[i]
            pend.append(a)
            inc.append(value)
    while pend and pend[-1] == 0:
        pend.pop()
    return ok, depth, incoming, bool(pend)


def tile_walk(arity, flags, n):
    """walk_tile_w's arithmetic: prefix sums of open slots and the ancestor
    criterion, for lanes 0..n (n <= 64)."""
    lanes = min(n + 1, len(arity), 64)
    q = []
    acc = 1
    for i in range(lanes):
        q.append(acc)
        if i < n:
            acc += arity[i] - 1
    depth, inc = [], []
    for i in range(lanes):
        d, s, m = 1, 0, q[i]
        for j in range(i - 1, -1, -1):
            m = min(m, q[j + 1])
            if m >= q[j]:
                d += 1
                s += flags[j]
        depth.append(d)
        inc.append(s)
    ok = all(not (q[i] < 1 or arity[i] < 0 or arity[i] > MAX_ARITY or (arity[i] > 0 and depth[i] >= 64)) for i in range(n))
    return ok, depth, inc, acc


def agree(arity, flags, n):
    ok_s, depth_s, inc_s, open_s = frame_walk(arity, flags, n)
    ok_t, depth_t, inc_t, open_t = tile_walk(arity, flags, n)
    if n < len(arity):
        # a walk to node n (node_depth_w, site_depth_at_w): node n itself
        # needs an open slot, the device's `open >= 1`
        ok_t = ok_t and open_t >= 1
    assert ok_t == ok_s, (arity, n)
    if ok_s:
        k = len(depth_s)
        assert depth_t[:k] == depth_s and inc_t[:k] == inc_s, (arity, flags, n)
        if n == len(arity):  # the splice's closing check: no slot left open
            assert (open_t >= 1) == open_s, (arity, n)


def test_tile_walk_matches_frame_walk_exhaustively():
    for length in range(1, 8):
        for arity in itertools.product([-1, 0, 1, 2, 3, 9], repeat=length):
            flags = [(i * 7 + sum(arity)) % 2 for i in range(length)]
            for n in range(length + 1):
                agree(list(arity), flags, n)


@pytest.mark.parametrize("seed", range(8))
def test_tile_walk_matches_frame_walk_on_long_rows(seed):
    rnd = random.Random(seed)
    for _ in range(300):
        length = rnd.randint(1, 64)
        # mostly well-formed prefix trees, sometimes chains to depth 64 and bad arities
        arity = [rnd.choice([0, 0, 1, 2, 2, 3]) for _ in range(length)]
        if rnd.random() < 0.2:
            arity = [1] * (length - 1) + [rnd.choice([0, 1])]
        if rnd.random() < 0.1:
            arity[rnd.randrange(length)] = rnd.choice([-1, MAX_ARITY + 1])
        flags = [rnd.randint(0, 1) for _ in range(length)]
        for n in {length, rnd.randint(0, length - 1), 0}:
            agree(arity, flags, n)
