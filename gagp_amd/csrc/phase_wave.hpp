// Wavefront forms of the per-child operators: one wavefront (64 lanes) works
// on one child. Every argument and result is uniform across the wavefront.
// Row passes (copies, site counts, subtree ends, constant ranges,
// comparisons, hashes) run lane-parallel over 64-node tiles. The prefix-order
// tree walks (depths, site depths, the splice's bounds) take all the nodes of
// a row of at most 64 at once from open-slot prefix sums (walk_tile_w);
// longer rows keep a pending-children stack one entry per lane (entry k in
// lane k, read with readlane and written with a lane select), so a walk step
// is a few VALU/SALU instructions with no memory traffic. Derivation and
// lowering run the shared bodies of phase.hpp uniformly with their stacks in
// lane registers; a derivation chooses each production with one lane per
// production.
//
// Every function returns exactly what its one-thread counterpart in phase.hpp
// returns and consumes the same random draws in the same order; those
// counterparts restate the reference bodies cited there
// (cpp/src/evolution/repro/gpu/phase_operators.cuh:46-265,
// cpp/src/evolution/repro/gpu/device/phase_variation.cuh:81-293).
//
// Tiles are aligned to multiples of 64, so node i of a row is always touched
// by lane i % 64: a lane reads back what it wrote itself. Values written by
// one lane and read by all (uniform reads of a written row) are followed by a
// workgroup fence first.
#pragma once
#include "phase.hpp"

namespace gagp_amd {

__device__ inline int wlane() { return int(threadIdx.x & 63u); }
// This wavefront's index in its block, marked uniform for the compiler (it
// cannot see that threadIdx.x / 64 is the same on every lane), so values
// derived from it stay in scalar registers.
__device__ inline int wave_index() { return __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)); }
__device__ inline std::uint64_t lanes_below() { return (1ull << wlane()) - 1ull; }
__device__ inline int wave_first(std::uint64_t mask) { return __ffsll((long long)mask) - 1; }
__device__ inline int rdlane(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
// Lane `lane` (uniform) of `old` replaced by v: one compare and one select.
__device__ inline int wrlane(int v, int lane, int old) { return wlane() == lane ? v : old; }
// Orders this wavefront's earlier global stores before its later loads.
__device__ inline void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}
__device__ inline int wave_incl_scan(int x) {
  const int lane = wlane();
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x;
}
__device__ inline std::uint64_t wave_sum_u64(std::uint64_t x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Copies the phase tables into this block's LDS (all kThreads threads take
// part; `lds` holds phase_tables_bytes(t) bytes, 16-byte aligned) and returns
// a view of the copy. The operators then read expression, production,
// domain and minimum-cost entries at LDS latency instead of one dependent
// global round trip per lookup.
template <int kThreads>
__device__ inline PhaseView stage_phase_tables(const PhaseTables& t, unsigned char* lds) {
  std::size_t off = 0;
  const auto copy = [&](GAS const void* src, std::size_t bytes) {
    unsigned* dst = reinterpret_cast<unsigned*>(lds + off);
    GAS const unsigned* from = (GAS const unsigned*)src;
    for (std::size_t k = threadIdx.x; k < bytes / 4; k += kThreads) dst[k] = from[k];
    off += table_bytes(bytes);
    return static_cast<const void*>(dst);
  };
  PhaseView v;
  v.expr = static_cast<const PExpr*>(copy(t.expr, sizeof(PExpr) * std::size_t(t.n_expr)));
  v.prod = static_cast<const PProd*>(copy(t.prod, sizeof(PProd) * std::size_t(t.n_prod)));
  v.nt = static_cast<const PNt*>(copy(t.nt, sizeof(PNt) * std::size_t(t.n_nt)));
  v.dom = static_cast<const PDomain*>(copy(t.dom, sizeof(PDomain) * std::size_t(t.n_dom)));
  v.values = static_cast<const Val*>(copy(t.values, sizeof(Val) * std::size_t(t.n_values)));
  v.minimum = static_cast<const unsigned*>(
      copy(t.minimum, sizeof(unsigned) * std::size_t(t.n_expr) * std::size_t(t.depth_count) * std::size_t(t.logical_count)));
  v.n_expr = t.n_expr;
  v.n_prod = t.n_prod;
  v.n_nt = t.n_nt;
  v.n_dom = t.n_dom;
  v.n_values = t.n_values;
  v.depth_count = t.depth_count;
  v.logical_count = t.logical_count;
  v.site_depth = t.site_depth;
  v.root = t.root;
  __syncthreads();
  return v;
}

// ---- lane-register stacks for the shared derivation and lowering bodies ----
// Pending derivation tasks: entry k in lane k % 64 of group k / 64. Named
// registers rather than arrays, so no access can fall back to memory.
struct WaveTasks {
  static_assert(kPending <= 192, "three groups of 64 lanes");
  // Task k is lane k mod 64 of group k / 64, packed in two registers:
  // expression | scope count << 17, and depth | logical depth << 16 (the
  // derivation keeps expressions below 2^17, scope counts and depths small
  // and non-negative). Groups are chosen by value selects only: a branch per
  // group lets the optimizer address the registers as an array in scratch
  // memory, whose loads then wait for the derivation's row stores.
  unsigned a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
  __device__ GenTask get(int k) const {
    const int ln = k & 63, grp = k >> 6;
    const unsigned ra0 = unsigned(rdlane(int(a0), ln)), ra1 = unsigned(rdlane(int(a1), ln)), ra2 = unsigned(rdlane(int(a2), ln));
    const unsigned rb0 = unsigned(rdlane(int(b0), ln)), rb1 = unsigned(rdlane(int(b1), ln)), rb2 = unsigned(rdlane(int(b2), ln));
    const unsigned a = grp == 0 ? ra0 : grp == 1 ? ra1 : ra2, b = grp == 0 ? rb0 : grp == 1 ? rb1 : rb2;
    return GenTask{int(a & 0x1ffffu), int(b & 0xffffu), int(a >> 17), int(b >> 16)};
  }
  __device__ void set(int k, const GenTask& x) {
    const unsigned a = unsigned(x.e) | unsigned(x.sc) << 17, b = unsigned(x.d) | unsigned(x.l) << 16;
    const int lane = wlane();
    const bool h0 = lane == k, h1 = lane + 64 == k, h2 = lane + 128 == k;
    a0 = h0 ? a : a0; b0 = h0 ? b : b0;
    a1 = h1 ? a : a1; b1 = h1 ? b : b1;
    a2 = h2 ? a : a2; b2 = h2 ? b : b2;
  }
};
__device__ inline double rdlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = unsigned(rdlane(int(b), lane)), hi = unsigned(rdlane(int(b >> 32), lane));
  return __longlong_as_double((long long)(std::uint64_t(hi) << 32 | lo));
}
// choose_production for a wavefront's derivation: lane i reads production i
// and its minimum cost in one step, the largest eligible weight is a
// max-reduction (exact in any order), each lane divides its own weight, and
// the sum and the draw's scan add the eligible weights in production order
// from registers, so every rounding is the in-order loop's.
__device__ inline int choose_production(const WaveTasks&, const PhaseView& g, const PNt& nt, int d, int l, unsigned available,
                                        SplitMix& random) {
  if (nt.production_count > 64) return choose_production_in_order(g, nt, d, l, available, random);
  const int lane = wlane();
  const bool in = lane < nt.production_count;
  const PProd p = in ? g.prod[nt.production_offset + lane] : PProd{-1, 0.0};
  const bool eligible = in && g_minimum(g, p.expression, d, l) <= available;
  double maximum = 1;
  if (!nt.normalized) {
    double m = eligible && p.weight > 0 ? p.weight : 0;
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(m, off, 64);
      m = o > m ? o : m;
    }
    maximum = rdlane_d(m, 0);
  }
  if (!(maximum > 0)) return kChooseCapacity;
  const double w = nt.normalized ? p.weight : p.weight / maximum;
  const std::uint64_t mask = __ballot(eligible);
  double total = 0;
  for (std::uint64_t m = mask; m; m &= m - 1) total += rdlane_d(w, wave_first(m));
  double choice = random.unit() * total;
  int chosen = kChooseNone;
  for (std::uint64_t m = mask; m; m &= m - 1) {
    const int i = wave_first(m);
    const int e = rdlane(p.expression, i);
    if (e < 0 || e >= g.n_expr) return kChooseInvalid;
    chosen = e;
    choice -= rdlane_d(w, i);
    if (choice < 0) break;
  }
  return chosen;
}

// Open nodes of the definite-assignment check: frame k in lane k.
struct WaveDefFrames {
  int origin = 0, seen = 0, saved = 0, branch = 0;
  __device__ DefFrame get(int k) const {
    return DefFrame{rdlane(origin, k), rdlane(seen, k), unsigned(rdlane(saved, k)), unsigned(rdlane(branch, k))};
  }
  __device__ void set(int k, const DefFrame& f) {
    origin = wrlane(f.origin, k, origin);
    seen = wrlane(f.seen, k, seen);
    saved = wrlane(int(f.saved), k, saved);
    branch = wrlane(int(f.branch), k, branch);
  }
};
// Open operations of the lowering walk: frame k in lane k.
struct WaveFrames {
  using Def = WaveDefFrames;
  int origin = 0, packed = 0, first = 0, second = 0, base = 0;  // packed: seen | args << 8
  __device__ Frame get(int k) const {
    const int p = rdlane(packed, k);
    return Frame{rdlane(origin, k), p & 0xff, rdlane(first, k), rdlane(second, k), rdlane(base, k),
                 {std::uint8_t(p >> 8), std::uint8_t(p >> 16), std::uint8_t(p >> 24)}};
  }
  __device__ void set(int k, const Frame& f) {
    origin = wrlane(f.origin, k, origin);
    packed = wrlane((f.seen & 0xff) | int(f.args[0]) << 8 | int(f.args[1]) << 16 | int(unsigned(f.args[2]) << 24), k, packed);
    first = wrlane(f.first_jump, k, first);
    second = wrlane(f.second_jump, k, second);
    base = wrlane(f.base_stack, k, base);
  }
};

// Row of one child with its counts kept uniform in registers.
__device__ inline void copy_row_w(const CRow& from, Row& to) {
  for (int i = wlane(); i < from.n; i += 64) { to.origin[i] = from.origin[i]; to.site[i] = from.site[i]; }
  for (int i = wlane(); i < from.nc; i += 64) to.cval[i] = from.cval[i];
  to.n = from.n;
  to.nc = from.nc;
}

__device__ inline int count_constants_w(const PhaseView& g, const CRow& t, int from, int to) {
  int count = 0;
  for (int tile = from & ~63; tile < to; tile += 64) {
    const int i = tile + wlane();
    count += __popcll(__ballot(i >= from && i < to && is_constant(g, t.origin[i])));
  }
  return count;
}

// choose_site (phase_operators.cuh:55-83): counted = one bounded draw over
// the eligible count; otherwise one reservoir draw per eligible site.
__device__ inline int choose_site_w(const PhaseView& g, const CRow& t, SplitMix& random, bool counted, int nt, int scope) {
  if (t.n <= 0) return -1;
  if (counted) {
    unsigned count = 0;
    for (int tile = 0; tile < t.n; tile += 64) {
      const int i = tile + wlane();
      count += unsigned(__popcll(__ballot(i < t.n && site_matches(g, t, i, nt, scope))));
    }
    if (!count) return -1;
    unsigned rank = unsigned(random.bounded(count));
    for (int tile = 0; tile < t.n; tile += 64) {
      const int i = tile + wlane();
      const bool eligible = i < t.n && site_matches(g, t, i, nt, scope);
      const std::uint64_t m = __ballot(eligible);
      const unsigned c = unsigned(__popcll(m));
      if (rank < c) return tile + wave_first(__ballot(eligible && unsigned(__popcll(m & lanes_below())) == rank));
      rank -= c;
    }
    return -1;
  }
  int selected = -1;
  unsigned count = 0;
  for (int tile = 0; tile < t.n; tile += 64) {
    const int i = tile + wlane();
    std::uint64_t m = __ballot(i < t.n && site_matches(g, t, i, nt, scope));
    while (m) {
      const int j = wave_first(m);
      m &= m - 1;
      if (random.bounded(++count) == 0) selected = tile + j;
    }
  }
  return selected;
}

// subtree_end: the first node after `start` where the pending-slot count
// 1 + sum(arity - 1) reaches zero, by a wavefront scan per tile.
__device__ inline int subtree_end_w(const PhaseView& g, const CRow& t, int start) {
  if (start < 0 || start >= t.n || t.n > kPhaseCapacity) return -1;
  int pending = 1;
  for (int tile = start & ~63; tile < t.n; tile += 64) {
    const int i = tile + wlane();
    const bool live = i >= start && i < t.n;
    const int a = live ? arity_of(g, t.origin[i]) : 1;
    const bool bad = live && (a < 0 || a > kMaxArity);
    const int scan = wave_incl_scan(live ? a - 1 : 0);
    const std::uint64_t events = __ballot(bad || (live && pending + scan == 0));
    if (events) {
      const int j = wave_first(events);
      return rdlane(int(bad), j) ? -1 : tile + j + 1;
    }
    pending += rdlane(scan, 63);
  }
  return -1;
}

// One step of a prefix-order walk: pops finished frames, then accounts for a
// node of arity `a` (pushed when it has children). `pend` holds the
// remaining-children count of frame k in lane k, `inc` its incoming site
// depth. Returns false on a malformed row.
struct WaveWalk {
  int pend, inc, depth;
  __device__ WaveWalk() : pend(wrlane(1, 0, 0)), inc(0), depth(1) {}
  __device__ bool pop_finished() {
    while (depth && rdlane(pend, depth - 1) == 0) --depth;
    return depth != 0;
  }
  // After pop_finished(): the node's incoming site depth before its own site.
  __device__ int incoming() const { return rdlane(inc, depth - 1); }
  __device__ bool visit(int a, int site_value) {
    pend = wrlane(rdlane(pend, depth - 1) - 1, depth - 1, pend);
    if (a) {
      if (depth == 64) return false;
      pend = wrlane(a, depth, pend);
      inc = wrlane(site_value, depth, inc);
      ++depth;
    }
    return true;
  }
};

// node_depth (phase_variation.cuh:93-110).
// The same walk over the nodes [0, n) of one tile (n <= 64) at once: lane i
// holds node i's arity `a` and site flag `f`. q_i = 1 + sum_{k<i} (a_k - 1)
// is the number of open child slots before node i, and node j < i is an
// ancestor of i exactly when min_{j<k<=i} q_k >= q_j (i lies in j's subtree,
// which ends where the slot count first drops below q_j). Each lane i <= n
// gets its depth (1 + ancestors: WaveWalk's depth before visiting i) and its
// incoming site depth (the site flags of its ancestors); `ok` is false where
// the walk would fail on one of the nodes [0, n): no open slot, an arity
// outside [0, kMaxArity], or children below depth 64. `open` is q_n.
struct TileWalk { int depth, inc, open; bool ok; };
__device__ inline TileWalk walk_tile_w(int a, int f, int n) {
  const int lane = wlane();
  const int delta = lane < n ? a - 1 : 0;
  const int q = 1 + wave_incl_scan(delta) - delta;
  int depth = 1, inc = 0, m = q;
  for (int j = (n < 63 ? n : 63) - 1; j >= 0; --j) {  // lanes 0..n, ancestors j < lane
    const int qj = rdlane(q, j), qn = rdlane(q, j + 1), fj = rdlane(f, j);
    if (j < lane) {
      m = qn < m ? qn : m;
      if (m >= qj) { ++depth; inc += fj; }
    }
  }
  TileWalk w;
  w.depth = depth;
  w.inc = inc;
  w.open = rdlane(q, n < 64 ? n : 63) + (n < 64 ? 0 : rdlane(delta, 63));
  w.ok = __ballot(lane < n && (q < 1 || a < 0 || a > kMaxArity || (a > 0 && depth >= 64))) == 0;
  return w;
}

__device__ inline int node_depth_w(const PhaseView& g, const CRow& t, int position) {
  if (position < 0 || position >= t.n || t.n > kPhaseCapacity) return -1;
  if (position < 64) {
    const int i = wlane();
    const TileWalk w = walk_tile_w(i <= position ? arity_of(g, t.origin[i]) : 0, 0, position);
    return w.ok && w.open >= 1 ? rdlane(w.depth, position) : -1;
  }
  WaveWalk w;
  for (int tile = 0; tile <= position; tile += 64) {
    const int i = tile + wlane();
    const int av = i <= position ? arity_of(g, t.origin[i]) : 0;
    const int last = position - tile < 63 ? position - tile : 63;
    for (int j = 0; j <= last; ++j) {
      if (!w.pop_finished()) return -1;
      if (tile + j == position) return w.depth;
      const int a = rdlane(av, j);
      if (a < 0 || a > kMaxArity) return -1;
      if (!w.visit(a, 0)) return -1;
    }
  }
  return -1;
}

// site_depth_of with a position: the incoming grammar-invocation depth there.
__device__ inline int site_depth_at_w(const PhaseView& g, const CRow& t, int position) {
  if (t.n <= 0 || t.n > kPhaseCapacity || position < 0 || position >= t.n) return -1;
  if (position < 64) {
    const int i = wlane();
    const bool in = i <= position;
    const TileWalk w = walk_tile_w(in ? arity_of(g, t.origin[i]) : 0, in && t.site[i] >= 0 ? 1 : 0, position);
    return w.ok && w.open >= 1 ? rdlane(w.inc, position) : -1;
  }
  WaveWalk w;
  for (int tile = 0; tile <= position; tile += 64) {
    const int i = tile + wlane();
    const int av = i <= position ? arity_of(g, t.origin[i]) : 0;
    const int sv = i <= position && t.site[i] >= 0 ? 1 : 0;
    const int last = position - tile < 63 ? position - tile : 63;
    for (int j = 0; j <= last; ++j) {
      if (!w.pop_finished()) return -1;
      const int in = w.incoming();
      if (tile + j == position) return in;
      const int a = rdlane(av, j);
      if (a < 0 || a > kMaxArity) return -1;
      if (!w.visit(a, in + rdlane(sv, j))) return -1;
    }
  }
  return -1;
}

// splice (phase_variation.cuh:147-207) with both subtree ends given: copies
// the three ranges tile by tile (constants by their rank within each range)
// and walks each tile for the depth and grammar-invocation depth bounds.
__device__ inline GenResult splice_w(const PhaseView& g, const CRow& base, int destination, int end, const CRow& donor,
                                     int source, int donor_end, Row& out, int out_capacity, int max_depth, int site_limit) {
  GenResult r{ST_INVALID, 0, 0, 0};
  if (out_capacity <= 0 || out_capacity > kPhaseCapacity || max_depth <= 0 || max_depth > 64 || base.n <= 0 ||
      base.n > kPhaseCapacity || donor.n <= 0 || donor.n > kPhaseCapacity || destination < 0 || destination >= base.n ||
      source < 0 || source >= donor.n)
    return r;
  const int a = base.site[destination], b = donor.site[source];
  if (a < 0 || b < 0 || site_nt(g, a) != site_nt(g, b) || site_scope(g, a) != site_scope(g, b)) return r;
  if (end < 0 || donor_end < 0) return r;
  const int length = base.n - (end - destination) + donor_end - source;
  if (length > out_capacity) { r.status = ST_CAPACITY; return r; }
  const int donor_first = count_constants_w(g, donor, 0, source), tail_first = count_constants_w(g, base, 0, end);
  const int in_end = destination + donor_end - source, shift = (end - destination) - (donor_end - source);
  int head_run = 0, in_run = 0, tail_run = 0;
  WaveWalk w;
  int maximum = 1, site_max = 0;
  // One tile: the walk runs on all lanes at once (walk_tile_w).
  const bool one_tile = length <= 64;
  bool deep = false, site_deep = false;
  int open = 0;
  for (int tile = 0; tile < length; tile += 64) {
    const int i = tile + wlane();
    const bool live = i < length;
    const int seg = i < destination ? 0 : i < in_end ? 1 : 2;
    const CRow& tree = seg == 1 ? donor : base;
    const int index = seg == 1 ? source + i - destination : seg == 0 ? i : i + shift;
    const int o = live ? tree.origin[index] : 0;
    const int st = live ? tree.site[index] : -1;
    const int ar = live ? arity_of(g, o) : 0;
    const bool constant = live && is_constant(g, o);
    const std::uint64_t m0 = __ballot(constant && seg == 0), m1 = __ballot(constant && seg == 1), m2 = __ballot(constant && seg == 2);
    bool bad = false;
    if (constant) {
      const std::uint64_t mine = seg == 0 ? m0 : seg == 1 ? m1 : m2;
      const int k = __popcll(mine & lanes_below()) + (seg == 0 ? head_run : seg == 1 ? in_run : tail_run);
      const int src = seg == 0 ? k : seg == 1 ? donor_first + k : tail_first + k;
      const int dst = head_run + in_run + tail_run + __popcll((m0 | m1 | m2) & lanes_below());
      if (src >= tree.nc) bad = true;
      else out.cval[dst] = tree.cval[src];
    }
    if (live) { out.origin[i] = o; out.site[i] = st; }
    if (__ballot(bad)) return r;
    head_run += __popcll(m0);
    in_run += __popcll(m1);
    tail_run += __popcll(m2);
    if (one_tile) {
      const int f = st >= 0 ? 1 : 0;
      const TileWalk tw = walk_tile_w(ar, f, length);
      if (!tw.ok) return r;
      open = tw.open;
      deep = __ballot(live && tw.depth > max_depth) != 0;
      site_deep = site_limit && __ballot(live && tw.inc + f > site_limit) != 0;
      break;
    }
    const int lanes = length - tile < 64 ? length - tile : 64;
    for (int j = 0; j < lanes; ++j) {
      const int aj = rdlane(ar, j);
      if (!w.pop_finished() || aj < 0 || aj > kMaxArity) return r;
      maximum = w.depth > maximum ? w.depth : maximum;
      // The destination keeps its own (site) entry; its flag equals the donor's.
      const int value = w.incoming() + (rdlane(st, j) >= 0 ? 1 : 0);
      site_max = value > site_max ? value : site_max;
      if (!w.visit(aj, value)) return r;
    }
  }
  out.site[destination] = a;
  if (one_tile ? open >= 1 : w.pop_finished()) return r;
  if (one_tile ? deep : maximum > max_depth) { r.status = ST_CAPACITY; return r; }
  if (one_tile ? site_deep : site_limit && site_max > site_limit) { r.status = ST_CAPACITY; return r; }
  r.nodes = length;
  r.constants = head_run + in_run + tail_run;
  r.status = ST_OK;
  out.n = length;
  out.nc = r.constants;
  return r;
}

// Origins of a row read 64 at a time, one per lane, and handed out uniformly.
struct WaveOrigins {
  const CRow& t;
  int tile = -64, tile_origin = 0;
  __device__ int operator()(int i) {
    if (i - tile >= 64 || i < tile) {
      tile = i & ~63;
      const int k = tile + wlane();
      tile_origin = k < t.n ? t.origin[k] : 0;
    }
    return rdlane(tile_origin, i - tile);
  }
};
// locals_defined by the whole wavefront (phase_operators.cuh:148-154).
__device__ inline bool locals_defined_w(const PhaseView& g, const CRow& t) {
  wave_fence();
  WaveDefFrames frames;
  WaveOrigins origin_at{t};
  return locals_defined_with(g, t.n, origin_at, frames);
}

// The parent and donor tournaments of crossover_child on all lanes at once.
// SplitMix64 is counter-based: draw k of the stream is the mix of the state
// plus (k + 1) increments, so lane k computes draw k directly and loads its
// candidate's keys; the selection then replays tournament()'s comparisons in
// draw order from registers. When a draw would be rejected by bounded()
// (below (2^64 - P) mod P), or the draws exceed the lanes, the in-order
// tournaments run from the unchanged state instead.
__device__ inline void tournaments_w(const Keys& k, int population, int pressure, SplitMix& random, int& parent, int& donor) {
  const int draws = 2 * pressure;
  const std::uint64_t n = std::uint64_t(population), threshold = (0ULL - n) % n;
  const int lane = wlane();
  const bool mine = lane < draws;
  const std::uint64_t v = splitmix_mix(random.s + std::uint64_t(lane) * 0x9e3779b97f4a7c15ULL);
  if (draws > 64 || __ballot(mine && v < threshold) != 0) {
    parent = tournament(k, population, pressure, random);
    donor = tournament(k, population, pressure, random);
    return;
  }
  const int c = mine ? int(v % n) : 0;
  const double f = mine ? k.fitness[c] : 0.0;
  const int w = mine && k.has_wrong ? int(k.wrong[c]) : 0, m = mine && k.has_wrong ? int(k.nodes[c]) : 0;
  // better() / selected_over() of draw i over the running best, in registers.
  const auto select = [&](int first) {
    int best = rdlane(c, first), bw = rdlane(w, first), bn = rdlane(m, first);
    double bf = rdlane_d(f, first);
    for (int i = first + 1; i < first + pressure; ++i) {
      const int ci = rdlane(c, i), wi = rdlane(w, i), ni = rdlane(m, i);
      const double fi = rdlane_d(f, i);
      bool over;
      if (fi != bf) over = fi > bf;
      else if (!k.has_wrong) over = false;
      else if (wi != bw) over = unsigned(wi) < unsigned(bw);
      else if (ni != bn) over = unsigned(ni) < unsigned(bn);
      else over = ci < best;
      if (over) { best = ci; bf = fi; bw = wi; bn = ni; }
    }
    return best;
  };
  parent = select(0);
  donor = select(pressure);
  random.s += std::uint64_t(draws) * 0x9e3779b97f4a7c15ULL;
}

// crossover_child (phase_operators.cuh:156-199).
template <class MemberRow>
__device__ inline void crossover_w(const PhaseView& g, const SlotInfo& s, int child, MemberRow in, Row& out, const Keys& keys,
                                   int population, int pressure, std::uint64_t seed, const OpPolicy& pol, OpStats& st) {
  SplitMix random(seed ^ (0x617089e168ULL + std::uint64_t(child)));
  st = OpStats{-1, 0, 0, 0, 0, 0, 0, -1, -1};
  int parent, donor;
  tournaments_w(keys, population, pressure, random, parent, donor);
  st.parent = parent;
  st.donor = donor;
  const CRow a = in(parent);
  if (a.n <= 0) { st.error = 1; return; }
  const int total = a.n;
  copy_row_w(a, out);
  if (pol.crossover_probability < 1 && random.unit() >= pol.crossover_probability) return;
  st.crossover_attempted = 1;
  if (pol.uniform_phase) (void)random.bounded(1);
  const CRow b = in(donor);
  for (int attempt = 0; attempt < pol.crossover_attempts; ++attempt) {
    const int dest = choose_site_w(g, a, random, pol.counted_sites != 0, -1, -1);
    if (dest < 0) break;
    const int contract = a.site[dest];
    const int src = b.n > 0 ? choose_site_w(g, b, random, pol.counted_sites != 0, site_nt(g, contract), site_scope(g, contract)) : -1;
    if (src < 0) { ++st.crossover_rejected; continue; }
    const int ae = subtree_end_w(g, a, dest), be = subtree_end_w(g, b, src);
    if (ae < 0 || be < 0) { st.error = 3; return; }
    if (total - (ae - dest) + (be - src) > s.max_variable_nodes) { ++st.crossover_rejected; continue; }
    const GenResult r = splice_w(g, a, dest, ae, b, src, be, out, s.row_capacity, s.max_depth, s.logical_depth);
    if (r.status == ST_OK && (!s.statement || locals_defined_w(g, as_const(out)))) { st.crossover = 1; return; }
    copy_row_w(a, out);  // failed speculation restores the phase
    ++st.crossover_rejected;
  }
}

// mutate_child (phase_operators.cuh:202-245): `x` is mutated in place through
// the private rows `tmp` (generated subtree) and `spliced`.
__device__ inline void mutate_w(const PhaseView& g, const SlotInfo& s, int child, Row& x, Row& tmp, Row& spliced, double probability,
                                std::uint64_t seed, const OpPolicy& pol, OpStats& st) {
  if (st.error) return;
  if (x.n <= 0) { st.error = 4; return; }
  const int total = x.n;
  SplitMix random(seed ^ (0x124ab512bf01ULL + std::uint64_t(child)));
  if (random.unit() >= probability) return;
  if (pol.uniform_phase) (void)random.bounded(1);
  wave_fence();  // x was written by the crossover lanes; its site entries are read uniformly below
  const CRow a = as_const(x);
  const int dest = choose_site_w(g, a, random, pol.counted_sites != 0, -1, -1);
  if (dest < 0) return;
  const int end = subtree_end_w(g, a, dest), depth = node_depth_w(g, a, dest);
  if (end < 0 || depth < 0) { st.error = 6; return; }
  const int site = a.site[dest];
  const int incoming = s.logical_depth ? site_depth_at_w(g, a, dest) : 0;
  if (incoming < 0 || incoming > s.logical_depth) { st.error = 6; return; }
  const int logical = s.logical_depth - incoming;
  const int replaced = end - dest;
  const int whole = s.max_variable_nodes - (total - replaced), phase = kPhaseCapacity - (a.n - replaced);
  const int available = whole < phase ? whole : phase, uncapped = available < kPhaseCapacity ? available : kPhaseCapacity;
  const bool capped = pol.mutation_max_nodes > 0 && pol.mutation_max_nodes < uncapped;
  const int bounded = capped ? pol.mutation_max_nodes : uncapped;
  if (bounded > s.row_capacity) { st.error = 10; return; }  // node budget exceeds the physical row
  WaveTasks pending;
  // Every lane runs the derivation and stores the same values, so each lane
  // reads back its own stores below.
  const GenResult gen = generate_phase_with(g, site_entry(site), site_parent_scope(site), s.max_depth - depth + 1, bounded,
                                            random.next(), tmp.origin, tmp.cval, tmp.site, logical, pending);
  if (capped && gen.status == ST_CAPACITY) return;
  if (gen.status != ST_OK) { st.error = 7; return; }
  const CRow sub{tmp.origin, tmp.site, tmp.cval, gen.nodes, gen.constants};
  const int sub_end = subtree_end_w(g, sub, 0);
  const GenResult r = splice_w(g, a, dest, end, sub, 0, sub_end, spliced, s.row_capacity, s.max_depth, s.logical_depth);
  if (r.status != ST_OK) { st.error = 8; return; }
  // An undefined local read rejects this mutation; the crossed child stands.
  if (s.statement && !locals_defined_w(g, as_const(spliced))) return;
  copy_row_w(as_const(spliced), x);
  st.mutation = 1;
}

// compare_child (phase_operators.cuh:247-265).
__device__ inline void compare_w(const CRow& parent_row, const CRow& child_row, int population, OpStats& st) {
  if (st.error) return;
  if (st.parent < 0 || st.parent >= population) { st.error = 9; return; }
  bool equal = parent_row.n == child_row.n && parent_row.nc == child_row.nc, reusable = true;
  if (equal) {
    bool diff = false;
    for (int i = wlane(); i < parent_row.n; i += 64) diff |= parent_row.origin[i] != child_row.origin[i];
    bool raw = false;
    for (int i = wlane(); i < parent_row.nc; i += 64) {
      const Val x = parent_row.cval[i], y = child_row.cval[i];
      diff |= !(x.tag == y.tag && (x.tag == TAG_BOOL ? x.b == y.b : x.i == y.i));
      raw |= !(x.i == y.i && x.b == y.b);
    }
    equal = __ballot(diff) == 0;
    reusable = __ballot(raw) == 0;
  }
  st.changed = !equal;
  st.reusable_parent = equal && reusable ? st.parent : -1;
}

__device__ inline std::uint64_t row_hash_w(const CRow& r) {
  std::uint64_t h = 0;
  for (int i = wlane(); i < r.n; i += 64) h += row_hash_node(i, r.origin[i]);
  for (int i = wlane(); i < r.nc; i += 64) h += row_hash_constant(i, r.cval[i]);
  return wave_sum_u64(h) + row_hash_counts(r.n, r.nc);
}

// Lowering of one row by the whole wavefront (uniform), origins read from a
// per-tile register by lane. `ws`: this wavefront's lowering record (in LDS
// for the kWalks instantiation; every lane stores the same values).
template <bool kWalks = false>
__device__ inline LowerResult lower_phase_w(const PhaseView& g, const CRow& t, const std::uint8_t* local_types, int local_count,
                                            int expected_type, Ins* code, int code_capacity, Val* pool, int pool_capacity,
                                            WalkStore<kWalks>& ws) {
  WaveFrames frames;
  WaveOrigins origin_at{t};
  return lower_phase_with<WaveFrames, WaveOrigins&, kWalks>(g, t, origin_at, local_types, local_count, expected_type, code,
                                                            code_capacity, pool, pool_capacity, frames, ws);
}

}  // namespace gagp_amd
