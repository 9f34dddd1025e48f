"""Pins the oracle (the CPU restatement) before anything is checked against it.

Each check is chosen so a plausible slip (a wrong shift in the RNG, an
off-by-one in the Lemire threshold, a different tie rule, a missed quirk)
fails it: golden vectors from the reference's own host code, libstdc++'s own
distribution, the reference's unit-test vectors, closed forms and brute force.
"""
import hashlib
import itertools
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_lib as O

GRAPHS = ["k4", "c100", "star6", "triangle", "weighted_general", "pattern_dups", "er2k", "chunglu4k",
          "rmat_s10", "delaunay1k"]


def sha(b):
    return hashlib.sha256(b).hexdigest()


# ---------------------------------------------------------------- RNG
def test_mt19937_64_standard_value():
    # [rand.predef]: the 10000th invocation of a default-constructed mt19937_64
    assert O.lib().orc_mt64_nth(5489, 10000) == 9981545732273789042


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_uniform_matches_libstdcxx():
    src = r"""
#include <random>
#include <cstdio>
#include <cstdint>
int main(int c, char** v) {
  unsigned long long seed = strtoull(v[1], 0, 10), a = strtoull(v[2], 0, 10), b = strtoull(v[3], 0, 10);
  int k = atoi(v[4]);
  std::mt19937_64 g(seed);
  for (int i = 0; i < k; i++) { std::uniform_int_distribution<uint64_t> d(a, b); printf("%llu\n", (unsigned long long)d(g)); }
}"""
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "u.cpp")
        open(p, "w").write(src.replace("strtoull", "std::strtoull").replace("atoi", "std::atoi")
                           .replace("#include <cstdint>", "#include <cstdint>\n#include <cstdlib>"))
        exe = os.path.join(td, "u")
        subprocess.run(["g++", "-O1", "-o", exe, p], check=True)
        # small ranges (rejection path taken often), 99, and a huge range near 2^63
        for seed, a, b in [(42, 0, 99), (7, 0, 2), (1, 5, 5), (3, 0, (1 << 63) + 12345), (9, 10, 1000003)]:
            want = [int(x) for x in subprocess.run([exe, str(seed), str(a), str(b), "2000"], check=True,
                                                   capture_output=True, text=True).stdout.split()]
            got = np.zeros(2000, np.uint64)
            O.lib().orc_uniform_draws(seed, a, b, 2000, got)
            assert got.tolist() == want


def test_uniform_is_uniform():
    got = np.zeros(60000, np.uint64)
    O.lib().orc_uniform_draws(123, 0, 2, 60000, got)
    counts = np.bincount(got.astype(np.int64), minlength=3)
    assert counts.sum() == 60000 and np.all(np.abs(counts - 20000) < 600)


# ---------------------------------------------------------------- loader and streams
@pytest.mark.parametrize("name", GRAPHS)
def test_loader_matches_reference(name, golden, graph_path):
    g = O.load_mtx(graph_path(name))
    want = golden["graphs"][name]
    assert g.n == want["n"] and g.nnz == want["nnz"]
    body = g.row_ptr.tobytes() + g.col.tobytes() + g.deg.tobytes()
    assert sha(body) == want["csr_sha256"]


@pytest.mark.parametrize("name", GRAPHS)
def test_init_stream_matches_reference(name, golden, graph_path):
    g = O.load_mtx(graph_path(name))
    u, v = O.init_stream(g)
    want = golden["graphs"][name]["init"]
    assert len(u) == want["count"]
    pairs = np.stack([u, v], axis=1).astype(np.uint32)
    assert sha(pairs.tobytes()) == want["sha256"]


@pytest.mark.parametrize("name", GRAPHS)
def test_bench_streams_match_reference(name, golden, graph_path):
    g = O.load_mtx(graph_path(name))
    for key, want in golden["graphs"][name]["bench"].items():
        seed, size, pct = (int(x) for x in key.split("/"))
        u, v, t = O.bench_stream(g, size, pct, seed)
        rec = O.records(u, v, t)
        assert want["count"] == size
        got = sha(rec.tobytes())
        if got != want["sha256"]:
            head = np.stack([u, v, t], axis=1)[:len(want["head"])].tolist()
            assert head == want["head"], key
            assert got == want["sha256"], key


# ---------------------------------------------------------------- color selection (reference unit vectors)
def _engine(n, edges, **cfg):
    g = O.Csr.from_edges(n, edges)
    return g, O.Engine(g, O.default_config(**cfg))


def test_find_color_bitmap_vectors():
    # A/test/unit/test_color_select.cu:62-134,252-273
    g, e = _engine(2, [(0, 1)])
    for bm, want in [(0, 1), (0x1, 2), (0b10111, 4), (0b11, 3)]:
        b = e.bitmaps()
        b[0] = bm
        e.set_bitmaps(b)
        assert e.find_color(0) == want


def test_find_color_bitmap_full_falls_back_to_row_scan():
    # The reference expects 65 for a full bitmap (A/test/unit/test_color_select.cu:137-160) but the
    # implementation scans the row colours (A/include/color_select.cuh:20-52): the code defines parity.
    g, e = _engine(4, [(0, 1), (0, 2), (0, 3)])
    w = e.words()
    w[1], w[2], w[3] = 1, 2, 4
    e.set_words(w)
    b = e.bitmaps()
    b[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    e.set_bitmaps(b)
    assert e.find_color(0) == 3
    w[1], w[2], w[3] = 0, 0, 0
    e.set_words(w)
    assert e.find_color(0) == 1


def test_find_color_bitmap_full_saturated_returns_256():
    n = 257
    g, e = _engine(n, [(0, i) for i in range(1, n)])
    w = e.words()
    w[1:256] = np.arange(1, 256)       # colours 1..255 all present
    w[256] = 300                       # colours >= 256 are not marked
    e.set_words(w)
    b = e.bitmaps()
    b[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    e.set_bitmaps(b)
    # vertex 0 has degree 256 > 64: force bitmap mode with a huge threshold
    g2 = O.Csr.from_edges(n, [(0, i) for i in range(1, n)])
    e2 = O.Engine(g2, O.default_config(histogram_threshold=1 << 30))
    e2.set_words(w)
    e2.set_bitmaps(b)
    assert e2.find_color(0) == 256


def test_find_color_histogram_vectors():
    # A/test/unit/test_color_select.cu:163-249 plus the all-nonzero minimum rule
    g, e = _engine(2, [(0, 1)], histogram_threshold=0)
    h = e.hist()
    assert h.shape == (2, 1024)
    cases = [({}, 1), ({0: 3, 1: 2, 2: 1, 3: 5, 4: 4}, 6), ({0: 2, 1: 1, 2: 0, 3: 3}, 3)]
    for bins, want in cases:
        h[:] = 0
        for k, c in bins.items():
            h[0, k] = c
        e.set_hist(h)
        assert e.find_color(0) == want
    h[0, :] = 5
    h[0, 700] = 2
    h[0, 900] = 2
    e.set_hist(h)
    assert e.find_color(0) == 701        # minimum count, lowest index on ties


# ---------------------------------------------------------------- engine: closed forms (reference integration tests)
def _run_all_adds(g, cfg=None):
    e = O.Engine(g, cfg or O.default_config())
    u, v = g.edges()
    e.submit(u, v)
    st = e.finish()
    return e, st


def _proper(g_or_edges, colors, n=None):
    if isinstance(g_or_edges, O.Csr):
        u, v = g_or_edges.edges()
        n = g_or_edges.n
    else:
        u, v = g_or_edges
    bad, unc = O.validate_edges(n, u, v, colors)
    return bad == 0 and unc == 0


def test_k4_needs_four_colors():
    g = O.Csr.from_edges(4, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
    e, st = _run_all_adds(g)
    c = e.colors()
    assert _proper(g, c) and sorted(c.tolist()) == [1, 2, 3, 4]


def test_c100_at_most_three():
    g = O.Csr.from_edges(100, [(i, (i + 1) % 100) for i in range(100)])
    e, st = _run_all_adds(g)
    c = e.colors()
    assert _proper(g, c) and c.max() <= 3


def test_triangle_then_delete():
    g = O.Csr.from_edges(3, [(0, 1), (1, 2), (0, 2)])
    e, _ = _run_all_adds(g)
    e.submit(np.array([0]), np.array([2]), np.array([1]))
    c = e.colors()
    assert c[0] != c[1] and c[1] != c[2] and np.all(c > 0)


def test_star_mixed_stream():
    g = O.Csr.from_edges(6, [(0, i) for i in range(1, 6)])
    e, _ = _run_all_adds(g)
    e.submit(np.array([0, 0, 1]), np.array([4, 5, 2]), np.array([1, 1, 0]))
    e.finish()
    c = e.colors()
    assert _proper((np.array([0, 0, 0, 1]), np.array([1, 2, 3, 2])), c, n=6)


def test_add_conflict_lower_priority_recolors_version_plus_one():
    # A/test/unit/test_local_update.cu:196-269: equal colours -> the lower-pi endpoint
    # recolours to a different non-zero colour with version +1; the winner is untouched.
    for d0, d1, loser in [(10, 5, 1), (10, 10, 0)]:
        # vertex 0 gets d0 pendant neighbours, vertex 1 gets d1
        edges, nxt = [], 2
        for _ in range(d0):
            edges.append((0, nxt)); nxt += 1
        for _ in range(d1):
            edges.append((1, nxt)); nxt += 1
        g = O.Csr.from_edges(nxt, edges)
        e = O.Engine(g)
        w = e.words()
        w[0] = (0 << 16) | 3
        w[1] = (0 << 16) | 3
        e.set_words(w)
        e.apply(np.array([0]), np.array([1]))
        w2 = e.words()
        win = 1 - loser
        assert w2[win] == w[win]
        assert (w2[loser] & 0xFFFF) not in (0, 3) and (w2[loser] >> 16) == 1


def test_uncolored_pair_is_not_a_conflict_and_gets_colored_version_one():
    g = O.Csr.from_edges(2, [])
    e = O.Engine(g)
    e.apply(np.array([0]), np.array([1]))
    w = e.words()
    assert (w[0] & 0xFFFF) == 1 and (w[1] & 0xFFFF) == 2 and (w[0] >> 16) == 1 and (w[1] >> 16) == 1


def test_out_of_range_update_is_skipped():
    g = O.Csr.from_edges(3, [(0, 1)])
    e = O.Engine(g)
    before = e.words().copy()
    e.apply(np.array([0, 7]), np.array([9, 1]))
    assert np.array_equal(e.words(), before)


def test_insert_takes_first_hole_even_with_later_copy():
    # Q8: a hole before an existing copy yields a duplicate; delete removes the first match only.
    g = O.Csr.from_edges(4, [(0, 1), (0, 2), (0, 3)])
    e = O.Engine(g)
    e.apply(np.array([0]), np.array([1]), np.array([1]))     # hole at slot 0 of row 0
    e.apply(np.array([0]), np.array([3]), np.array([0]))     # 3 already at slot 2: goes into the hole
    rp, col = e.csr()
    row0 = col[rp[0]:rp[1]].tolist()
    assert row0[:3] == [3, 2, 3]
    e.apply(np.array([0]), np.array([3]), np.array([1]))
    rp, col = e.csr()
    assert col[rp[0]:rp[1]].tolist()[:3] == [0xFFFFFFFF, 2, 3]


def test_overflow_then_drain_and_drop():
    # rows of capacity 1 (slack 0, min 1): the second neighbour overflows; capacity-1 buffer drops
    g = O.Csr.from_edges(4, [])
    e = O.Engine(g, O.default_config(slack_pct=0, min_row_capacity=1, overflow_capacity=1))
    e.apply(np.array([0, 0, 0]), np.array([1, 2, 3]))
    assert e.overflow_count() == 2          # 0-2 and 0-3 overflowed on row 0; rows 2 and 3 had room
    st = e.finish()
    # one entry fits the 1-entry buffer (the other is lost), and row 0 is still full after compaction
    assert st["overflow_count"] == 2 and st["overflow_drained"] == 1 and st["overflow_failed"] == 1


def test_repair_makes_random_streams_proper():
    rng = np.random.default_rng(5)
    for trial in range(6):
        n = int(rng.integers(20, 300))
        m = int(rng.integers(n, 6 * n))
        g = O.Csr.from_edges(n, rng.integers(0, n, size=(m, 2)))
        e, st = _run_all_adds(g)
        assert _proper(g, e.colors())
        u, v, t = O.bench_stream(g, 2000, int(rng.integers(20, 81)), trial)
        e.submit(u, v, t)
        st = e.finish()
        lu, lv = O.live_edges_after(g, u, v, t)
        bad, unc = O.validate_edges(n, lu, lv, e.colors())
        assert st["t2_remaining"] == 0 and bad == 0 and unc == 0


def test_repair_every_epoch_publishes_proper_snapshots():
    rng = np.random.default_rng(11)
    n = 400
    g = O.Csr.from_edges(n, rng.integers(0, n, size=(3000, 2)))
    cfg = O.default_config(epoch_size=256, batch_size=256, flags=O.FLAG_REPAIR_EVERY_EPOCH)
    e = O.Engine(g, cfg)
    u, v = O.init_stream(g)
    e.submit(u, v)
    e.finish()
    su, sv, st = O.bench_stream(g, 256 * 6, 50, 42)
    for k in range(6):
        sl = slice(256 * k, 256 * (k + 1))
        e.apply(su[sl], sv[sl], st[sl])
        s = e.epoch()
        lu, lv = O.live_edges_after(g, su[:sl.stop], sv[:sl.stop], st[:sl.stop])
        bad, unc = O.validate_edges(n, lu, lv, e.snapshot())
        assert s["snapshot_published"] == 1 and bad == 0 and unc == 0


def _chromatic(n, edges):
    adj = [set() for _ in range(n)]
    for a, b in edges:
        adj[a].add(b); adj[b].add(a)
    for k in range(1, n + 1):
        for assign in itertools.product(range(k), repeat=n):
            if all(assign[a] != assign[b] for a, b in edges):
                return k
    return n


def test_max_color_at_least_chromatic_number_small_graphs():
    rng = np.random.default_rng(3)
    for _ in range(8):
        n = int(rng.integers(4, 8))
        raw = rng.integers(0, n, size=(int(rng.integers(n, 3 * n)), 2))
        g = O.Csr.from_edges(n, raw)
        u, v = g.edges()
        if len(u) == 0:
            continue
        e, _ = _run_all_adds(g)
        c = e.colors()
        assert _proper(g, c)
        assert c.max() >= _chromatic(n, list(zip(u.tolist(), v.tolist())))


# ---------------------------------------------------------------- tier-2 and the static greedy
def test_tier2_recolors_higher_id_endpoint_first_fit():
    # Path 0-1-2 all colour 1 (and bitmaps/histograms irrelevant to tier-2).  Pass 1 collects the
    # higher endpoint of each conflicting edge: {1, 2}; greedy in ascending order: 1 -> 2 (neighbours
    # 0, 2 have colour 1), then 2 sees neighbour 1 with colour 2 -> first free is 1 == its colour: no change.
    g = O.Csr.from_edges(3, [(0, 1), (1, 2)])
    e = O.Engine(g)
    e.set_words(np.array([1, 1, 1], np.uint32))
    st = e.tier2()
    assert e.colors().tolist() == [1, 2, 1] and st["t2_fixed"] == 1
    assert (e.words()[1] >> 16) == 1


def test_tier2_colors_uncolored_vertices_with_edges():
    g = O.Csr.from_edges(4, [(0, 1), (2, 3)])
    e = O.Engine(g)
    e.set_words(np.array([0, 1, 2, 0], np.uint32))
    e.tier2()
    assert e.colors().tolist() == [2, 1, 2, 1]


def _first_fit(g):
    color = [0] * g.n
    for v in range(g.n):
        nb = {color[w] for w in g.col[g.row_ptr[v]:g.row_ptr[v + 1]].tolist()}
        c = 1
        while c in nb:
            c += 1
        color[v] = c
    return max(color) if color else 0


def test_static_greedy_is_first_fit(graph_path):
    for name in ["k4", "c100", "er2k", "rmat_s10", "delaunay1k"]:
        g = O.load_mtx(graph_path(name))
        assert O.static_greedy(g) == _first_fit(g)


# ---------------------------------------------------------------- the reference's ablations (Makefile:58-78)
def test_ablation_force_bitmap_ignores_the_histogram():
    # A/include/color_select.cuh:98-100: a hub selects from its bitmap.  The
    # histogram's first zero is colour 2, the bitmap's first clear bit colour 4.
    g, e = _engine(2, [(0, 1)], histogram_threshold=0)
    h = e.hist()
    h[:] = 0
    h[0, 0] = 1
    e.set_hist(h)
    b = e.bitmaps()
    b[0] = 0b111
    e.set_bitmaps(b)
    assert e.find_color(0) == 2
    e.set_flags(O.FLAG_ABL_FORCE_BITMAP)
    assert e.find_color(0) == 4
    e.set_flags(O.FLAG_ABL_FORCE_HISTOGRAM)            # the reference's default dispatch
    assert e.find_color(0) == 2


def test_ablation_no_overflow_drops_unbuffered():
    # A/src/kernels/local_update.cu:28-38: a full row drops the edge; nothing is buffered or counted
    g = O.Csr.from_edges(4, [])
    e = O.Engine(g, O.default_config(slack_pct=0, min_row_capacity=1, overflow_capacity=8,
                                     flags=O.FLAG_ABL_NO_OVERFLOW))
    e.apply(np.array([0, 0, 0]), np.array([1, 2, 3]))
    assert e.overflow_count() == 0
    rp, col = e.csr()
    assert col[rp[0]:rp[1]].tolist() == [1]                 # row 0 kept its first neighbour only
    assert col[rp[2]:rp[3]].tolist() == [0] and col[rp[3]:rp[4]].tolist() == [0]
    st = e.finish()
    assert st["overflow_count"] == 0 and st["overflow_drained"] == 0


def test_ablation_live_degree_priority_picks_the_other_loser():
    # A/src/kernels/conflict_repair.cu:46-58: a vertex's own priority comes from
    # its live row, its neighbours' from the loaded degree.  Vertex 0 (deg0 3,
    # live 1 after two deletes) and vertex 1 (deg0 2) share colour 1: frozen
    # priorities make 1 the loser, the live degree makes 0 the loser.
    edges = [(0, 1), (0, 2), (0, 3), (1, 4)]
    for flags, loser in ((0, 1), (O.FLAG_ABL_LIVE_DEGREE, 0)):
        g, e = _engine(5, edges, flags=flags)
        e.apply(np.array([0, 0]), np.array([2, 3]), np.array([1, 1]))    # deletes: row 0 keeps vertex 1
        w = np.array([1, 1, 2, 3, 2], np.uint32)                        # version 0, colours
        e.set_words(w)
        e.tier1()
        after = e.words()
        assert (after[loser] & 0xFFFF) != 1 and (after[loser] >> 16) == 1
        assert after[1 - loser] == w[1 - loser]


# ---------------------------------------------------------------- device quirks, hand-computed states
# Each case sets a small state by hand (words, bitmaps, histograms, rows), runs
# one oracle step and compares every affected array with values worked out by
# hand from the reference lines cited.  tests/test_oracle_mutants.py plants a
# slip in each of these oracle functions and requires one of these to fail.
def _w(color, version):
    return (version << 16) | color


def test_delete_reduces_with_the_neighbours_colour_and_propagates():
    # A/src/kernels/local_update.cu:126-141, recolor.cu:50-82,145-163, color_select.cuh:120-157,188-209.
    # Histogram rows for every vertex (threshold 0).  Path 1 - 0 - 2; colours 0:3 (v5), 1:1 (v2), 2:2 (v7).
    # DEL(0,1): freed_u = colour(1) = 1, freed_v = colour(0) = 3 (both read before anything changes).
    #   hist[0] bin 1 -> 0, hist[1] bin 3 -> 0.
    #   reduce(0, 1): 1 < 3 and the first zero bin of hist[0] is colour 1 < 3 -> recolour 0 to 1 (v6);
    #     propagate 3 -> 1 along row 0 (now only vertex 2): bitmap[2] |= bit(1), hist[2] bin 3 -1, bin 1 +1.
    #   reduce(1, 3): 3 >= colour(1) = 1 -> nothing.
    g, e = _engine(3, [(0, 1), (0, 2)], histogram_threshold=0)
    e.set_words(np.array([_w(3, 5), _w(1, 2), _w(2, 7)], np.uint32))
    h = e.hist()
    h[:] = 0
    h[0, 0] = 1          # vertex 0 sees colour 1 (vertex 1)
    h[0, 1] = 1          # and colour 2 (vertex 2)
    h[1, 2] = 1          # vertex 1 sees colour 3
    h[2, 2] = 1          # vertex 2 sees colour 3
    e.set_hist(h)
    e.apply(np.array([0]), np.array([1]), np.array([1]))
    assert e.words().tolist() == [_w(1, 6), _w(1, 2), _w(2, 7)]
    want = np.zeros((3, 1024), np.uint32)
    want[0, 1] = 1
    want[2, 0] = 1
    assert np.array_equal(e.hist(), want)
    assert e.bitmaps().tolist() == [0, 0, 0b1]


def test_delete_leaves_bitmap_bits_and_reduces_after_a_rebuild():
    # Bitmap rows (threshold 64).  Star 0 - {1, 2, 3}; colours 0:4 (v1), 1:1, 2:3, 3:2.
    # A delete never clears a bitmap bit (color_select.cuh:188-209), so right after DEL(0,1)
    # bitmap[0] = {1,2,3} still holds colour 1 and the first clear bit is 4 = colour(0): no reduction.
    # The rebuild (epoch_maintenance.cu:113-128) makes bitmap[0] = {2,3} (rows 0: {2,3}, 1: {},
    # 2: {4}, 3: {4}).  DEL(0,2): freed_u = 3 < 4 and the first clear bit is colour 1 -> 0 becomes 1 (v2);
    # propagate 4 -> 1 to the one neighbour left, 3: bitmap[3] = {4} | {1} (the old bit stays set).
    # freed_v = 4 >= colour(2) = 3: vertex 2 keeps its colour.
    g, e = _engine(4, [(0, 1), (0, 2), (0, 3)])
    words = np.array([_w(4, 1), _w(1, 1), _w(3, 1), _w(2, 1)], np.uint32)
    e.set_words(words)
    e.set_bitmaps(np.array([0b0111, 0b1000, 0b1000, 0b1000], np.uint64))
    e.apply(np.array([0]), np.array([1]), np.array([1]))
    assert e.words().tolist() == words.tolist()
    assert e.bitmaps().tolist() == [0b0111, 0b1000, 0b1000, 0b1000]
    e.rebuild_bitmaps()
    assert e.bitmaps().tolist() == [0b0110, 0, 0b1000, 0b1000]
    e.apply(np.array([0]), np.array([2]), np.array([1]))
    assert e.words().tolist() == [_w(1, 2), _w(1, 1), _w(3, 1), _w(2, 1)]
    assert e.bitmaps().tolist() == [0b0110, 0, 0b1000, 0b1001]


def test_delete_of_a_conflicting_edge_does_not_reduce():
    # recolor.cu:155: reduction needs freed < current.  Edge 0-1 with both endpoints colour 3 and empty
    # bitmaps (first clear bit: colour 1): freed == current on both sides, so neither endpoint moves.
    g, e = _engine(2, [(0, 1)])
    words = np.array([_w(3, 4), _w(3, 9)], np.uint32)
    e.set_words(words)
    e.apply(np.array([0]), np.array([1]), np.array([1]))
    assert e.words().tolist() == words.tolist()


def test_add_conflict_recolour_propagates_to_the_losers_row():
    # local_update.cu:110-122 then recolor.cu:87-111 and :50-82.  Vertex 1 (deg0 1) loses to
    # vertex 0 (deg0 2) on colour 2.  Vertex 1's row: {0, 3} after the insert; vertex 3 is a hub
    # (threshold 1, deg0 2) with histogram bins {2: 1}.  bitmap[1] = {1, 2} -> vertex 1 takes colour 3 (v5).
    # Propagation 2 -> 3: bitmap[0] |= bit(3) (keeps bit 2), hist[0] bin 2 -1 and bin 3 +1;
    # vertex 3: bitmap |= bit(3), hist bin 2 -1 -> 0 and bin 3 +1.
    # Hubs (deg0 > 1): vertex 0 -> histogram row 0, vertex 2 -> row 1, vertex 3 -> row 2.
    g, e = _engine(5, [(0, 2), (0, 4), (1, 3), (3, 2)], histogram_threshold=1)
    e.set_words(np.array([_w(2, 1), _w(2, 4), _w(1, 1), _w(5, 1), _w(7, 1)], np.uint32))
    e.set_bitmaps(np.array([0, 0b0011, 0, 0, 0], np.uint64))
    h = e.hist()
    h[:] = 0
    h[0, 1] = 1          # vertex 0's histogram counts colour 2
    h[2, 1] = 1          # vertex 3's histogram counts colour 2 (vertex 1)
    e.set_hist(h)
    # bitmap/histogram bookkeeping of the insert itself (color_select.cuh:160-183): vertex 0 gets
    # +colour(1) = 2 in hist bin 2 (now 2) and bit 2; vertex 1 gets bit 2 (already set).
    e.apply(np.array([0]), np.array([1]), np.array([0]))
    w = e.words()
    assert w.tolist() == [_w(2, 1), _w(3, 5), _w(1, 1), _w(5, 1), _w(7, 1)]
    want = np.zeros((3, 1024), np.uint32)
    want[0, 1] = 1       # 1 + 1 (insert) - 1 (propagated 2 -> 3)
    want[0, 2] = 1
    want[2, 2] = 1       # vertex 3: 2 -> 3
    assert np.array_equal(e.hist(), want)
    assert e.bitmaps().tolist() == [0b0110, 0b0011, 0, 0b0100, 0]


def test_compaction_is_a_stable_left_pack():
    # epoch_maintenance.cu:20-43: survivors keep their order, the tail becomes INVALID.
    # Row 0 = [1 2 3 4 5] in a 20-slot row; deleting 2 and 4 leaves [1 _ 3 _ 5].
    g, e = _engine(6, [(0, k) for k in range(1, 6)])
    e.apply(np.array([0, 0]), np.array([2, 4]), np.array([1, 1]))
    e.compact()
    rp, col = e.csr()
    inv = 0xFFFFFFFF
    assert rp[1] - rp[0] == 20
    assert col[rp[0]:rp[1]].tolist() == [1, 3, 5] + [inv] * 17
    assert col[rp[2]:rp[3]].tolist() == [inv] * 8 and col[rp[1]:rp[2]].tolist() == [0] + [inv] * 7
    # A later insert takes the first hole (slot 3); a re-insert of 3 finds it at slot 1 first.
    e.apply(np.array([0, 0]), np.array([4, 3]), np.array([0, 0]))
    rp, col = e.csr()
    assert col[rp[0]:rp[0] + 5].tolist() == [1, 3, 5, 4, inv]


def test_tier1_marks_its_own_colour_before_selecting():
    # conflict_repair.cu:100-112: the loser ORs its own colour into its bitmap, then selects.
    # Vertex 0 (pi below vertex 1's: equal degree, lower id) shares colour 2 with vertex 1;
    # its bitmap {1} lacks colour 2, so without the own bit it would pick 2 again.
    g, e = _engine(2, [(0, 1)])
    e.set_words(np.array([_w(2, 3), _w(2, 9)], np.uint32))
    e.set_bitmaps(np.array([0b001, 0], np.uint64))
    assert e.tier1_sweep() == 1
    assert e.words().tolist() == [_w(3, 4), _w(2, 9)]
    assert e.bitmaps().tolist() == [0b011, 0]


def test_tier1_stops_on_the_fourth_equal_conflict_count():
    # streamgc.cu:1378-1389: a count equal to the previous one three times in a row stops the loop
    # (counts c, c, c, c -> 4 sweeps).  Vertex 0 (histogram row, drifted: every bin 1 except its own
    # colour 5) loses to vertex 1 on colour 5 in every sweep and re-selects colour 5, one conflict
    # per sweep; histograms are not rebuilt (epoch_maintenance.cu:119-122), so the drift persists.
    g, e = _engine(2, [(0, 1)], histogram_threshold=0)
    e.set_words(np.array([_w(5, 0), _w(5, 0)], np.uint32))
    h = e.hist()
    h[:] = 1
    h[0, 4] = 0
    e.set_hist(h)
    st = e.tier1()
    assert st["t1_iterations"] == 4 and st["t1_conflicts_last"] == 1
    assert e.words().tolist() == [_w(5, 4), _w(5, 0)]
    # counts 2, 1, 1, 1, 1: a changed count resets the plateau -> 5 sweeps.  Vertices 2-3 hold a
    # conflict that the first sweep fixes (vertex 2's exact histogram: colour 7 only -> colour 1).
    g, e = _engine(4, [(0, 1), (2, 3)], histogram_threshold=0)
    e.set_words(np.array([_w(5, 0), _w(5, 0), _w(7, 0), _w(7, 0)], np.uint32))
    h = e.hist()
    h[:] = 0
    h[0, :] = 1
    h[0, 4] = 0
    h[2, 6] = 1
    e.set_hist(h)
    st = e.tier1()
    assert st["t1_iterations"] == 5 and st["t1_conflicts_last"] == 1
    assert e.words().tolist() == [_w(5, 5), _w(5, 0), _w(1, 1), _w(7, 0)]
    # with the plateau out of reach the loop ends at max_repair_iters
    g, e = _engine(2, [(0, 1)], histogram_threshold=0, plateau_iters=100, max_repair_iters=20)
    e.set_words(np.array([_w(5, 0), _w(5, 0)], np.uint32))
    h = e.hist()
    h[:] = 1
    h[0, 4] = 0
    e.set_hist(h)
    st = e.tier1()
    assert st["t1_iterations"] == 20 and e.words().tolist() == [_w(5, 20), _w(5, 0)]


def test_drain_retries_overflow_in_buffer_order():
    # local_update.cu:27-38 buffers (0,3), (0,4), (0,5) in that order behind the full row [1 2];
    # DEL(0,1) frees one slot; the reference-order epoch compacts row 0 to [2 _] and drains
    # (epoch_maintenance.cu:66-87): the first entry takes the slot, the other two fail.  The
    # reference drains with a thread per entry; one wave in entry order is the deterministic replay.
    g, e = _engine(6, [(0, 1), (0, 2)], slack_pct=0, min_row_capacity=2, overflow_capacity=16)
    e.apply(np.array([0, 0, 0, 0]), np.array([3, 4, 5, 1]), np.array([0, 0, 0, 1]))
    assert e.overflow_count() == 3
    st = e.epoch()
    assert (st["overflow_count"], st["overflow_drained"], st["overflow_failed"]) == (3, 3, 2)
    rp, col = e.csr()
    assert col[rp[0]:rp[1]].tolist() == [2, 3]
    assert e.overflow_count() == 0


def test_rebuild_leaves_hub_bitmaps_and_histograms_alone():
    # epoch_maintenance.cu:119-123: histogram vertices keep their bitmap; the reference never
    # recounts histograms.  Star 0 - {1, 2, 3}, threshold 2: vertex 0 is the only hub.
    g, e = _engine(4, [(0, 1), (0, 2), (0, 3)], histogram_threshold=2)
    e.set_words(np.array([_w(4, 1), _w(1, 1), _w(2, 1), _w(3, 1)], np.uint32))
    e.set_bitmaps(np.array([0xDEAD, 0, 0, 0], np.uint64))
    h = e.hist()
    h[:] = 0
    h[0, 9] = 3          # drifted counts stay as they are
    e.set_hist(h)
    e.rebuild_bitmaps()
    assert e.bitmaps().tolist() == [0xDEAD, 0b1000, 0b1000, 0b1000]
    assert e.hist()[0, 9] == 3 and e.hist().sum() == 3
    # the exact-histogram mode (no reference counterpart) recounts the hub's bins, bitmap still untouched
    e.set_flags(O.FLAG_EXACT_HISTOGRAM)
    e.rebuild_bitmaps()
    want = np.zeros((1, 1024), np.uint32)
    want[0, 0:3] = 1
    assert np.array_equal(e.hist(), want)
    assert e.bitmaps().tolist() == [0xDEAD, 0b1000, 0b1000, 0b1000]


def test_tier1_colours_uncoloured_vertices_with_edges_and_counts_them():
    # conflict_repair.cu:35-43: an uncoloured vertex with a live slot takes a colour, version +1,
    # and counts as a conflict so the loop runs again; an isolated uncoloured vertex stays as it is.
    g, e = _engine(3, [(0, 1)])
    e.set_words(np.array([_w(0, 7), _w(1, 0), _w(0, 3)], np.uint32))
    e.set_bitmaps(np.array([0b1, 0, 0], np.uint64))
    assert e.tier1_sweep() == 1
    assert e.words().tolist() == [_w(2, 8), _w(1, 0), _w(0, 3)]
    # through the loop: sweep 1 counts it, sweep 2 finds nothing -> 2 iterations
    g, e = _engine(3, [(0, 1)])
    e.set_words(np.array([_w(0, 7), _w(1, 0), _w(0, 3)], np.uint32))
    st = e.tier1()
    assert st["t1_iterations"] == 2 and st["t1_conflicts_last"] == 0
    assert e.words().tolist() == [_w(2, 8), _w(1, 0), _w(0, 3)]
