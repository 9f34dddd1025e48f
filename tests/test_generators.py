"""Pins the synthetic inputs (Chung-Lu, R-MAT, Delaunay, core numbers and the
streams built on them) by what the mathematics fixes, independently of the
generators' own code:

- core numbers against networkx.core_number;
- Chung-Lu: exactly m distinct edges, the rank-degree slope -1/(gamma-1) of the
  expected-degree sequence w_i ~ (i + i0)^(-1/(gamma-1)), the maximum degree
  and E[d^2]/E[d] against the weights, and a permutation that keeps the degree
  multiset;
- R-MAT: the distinct-edge count and sum of squared degrees against their
  exact expectations under the quadrant probabilities;
- Delaunay: Euler's count m = 3n - 3 - h with h from an independent convex hull;
- determinism across OpenMP thread counts, and seeded golden hashes
  (tests/golden/make_generator_hashes.py).
"""
import hashlib
import importlib.util
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import streamgc_amd as S

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _degrees(g):
    return np.diff(g.row_ptr.astype(np.int64))


def _well_formed(g):
    deg = _degrees(g)
    u = np.repeat(np.arange(g.n, dtype=np.int64), deg)
    v = g.col_idx.astype(np.int64)
    assert (u != v).all()
    fwd = set((u[u < v] << 32 | v[u < v]).tolist())
    back = set((v[u > v] << 32 | u[u > v]).tolist())
    assert fwd == back and len(fwd) == len(v) // 2          # symmetric, no duplicates
    for x in range(0, g.n, max(1, g.n // 64)):                # rows sorted
        row = g.col_idx[g.row_ptr[x]:g.row_ptr[x + 1]]
        assert (np.diff(row.astype(np.int64)) > 0).all()


def _nx(g):
    import networkx as nx
    h = nx.Graph()
    h.add_nodes_from(range(g.n))
    deg = _degrees(g)
    u = np.repeat(np.arange(g.n, dtype=np.int64), deg)
    v = g.col_idx.astype(np.int64)
    h.add_edges_from(zip(u[u < v].tolist(), v[u < v].tolist()))
    return h


@pytest.mark.parametrize("make", [
    lambda: S.gen_rmat(10, 16.0, seed=1),
    lambda: S.gen_rmat(11, 40.0, seed=2),
    lambda: S.gen_chung_lu(4096, 32768, 2.3, 600.0, seed=3),
    lambda: S.gen_chung_lu(3000, 6000, 2.1, 300.0, seed=4),
    lambda: S.gen_delaunay(700, seed=5),
])
def test_core_numbers_match_networkx(make):
    import networkx as nx
    g = make()
    want = nx.core_number(_nx(g))
    got = S.core_numbers(g)
    assert got.tolist() == [want[i] for i in range(g.n)]


def _chung_lu_weights(n, m, gamma, max_degree):
    """The expected-degree sequence the generator is specified by, with i0 solved
    from max/mean weight = max_degree / (2m/n) (bisection on a log scale)."""
    alpha = 1.0 / (gamma - 1.0)
    want = max_degree / (2.0 * m / n)
    lo, hi = 1e-3, float(n)
    for _ in range(80):
        mid = np.sqrt(lo * hi)
        w = (np.arange(n) + mid) ** -alpha
        if mid ** -alpha / w.mean() > want:
            lo = mid
        else:
            hi = mid
    i0 = np.sqrt(lo * hi)
    return i0, (np.arange(n) + i0) ** -alpha


@pytest.mark.parametrize("gamma,max_degree", [(2.3, 900.0), (2.1, 1500.0), (2.8, 400.0)])
def test_chung_lu_degree_sequence(gamma, max_degree):
    n, m = 1 << 14, 1 << 17
    g = S.gen_chung_lu(n, m, gamma, max_degree, seed=5, permute=False)
    _well_formed(g)
    deg = _degrees(g)
    assert deg.sum() == 2 * m
    i0, w = _chung_lu_weights(n, m, gamma, max_degree)
    # rank-degree slope of the unpermuted ids: -1/(gamma-1)
    idx = np.arange(100, n // 4)
    slope = np.polyfit(np.log(idx + i0), np.log(np.maximum(deg[idx], 0.5)), 1)[0]
    assert abs(slope + 1.0 / (gamma - 1.0)) < 0.04, slope
    # the top vertex's degree is its expected 2m w_0 / sum(w) of raw draws, less the repeated pairs
    # dropped (the heavier the tail, the more of them: about 10% at gamma 2.3, 20% at 2.1)
    expect0 = 2 * m * w[0] / w.sum()
    assert 0.7 * expect0 < deg[0] <= 1.05 * expect0
    # E[d^2]/E[d] from the weights; the repeated hub pairs dropped take mass from the top of the
    # sequence (8% at gamma 2.3, 21% at 2.1), so the realised value sits a little below the weights'
    pred = 2 * m * (w ** 2).sum() / w.sum() ** 2 + 1
    got = (deg.astype(np.float64) ** 2).mean() / deg.mean()
    assert 0.7 * pred < got <= 1.02 * pred, (got, pred)
    # the permutation keeps the degree multiset and removes the id-degree trend
    gp = S.gen_chung_lu(n, m, gamma, max_degree, seed=5, permute=True)
    dp = _degrees(gp)
    assert np.array_equal(np.sort(dp), np.sort(deg))
    assert abs(np.corrcoef(np.arange(n), dp)[0, 1]) < 0.05


@pytest.mark.parametrize("scale,ef,seed", [(9, 8.0, 1), (10, 16.0, 3), (10, 40.0, 4)])
def test_rmat_distinct_edges_and_degrees(scale, ef, seed):
    a, b, c = 0.57, 0.19, 0.19
    d = 1 - a - b - c
    g = S.gen_rmat(scale, ef, a, b, c, seed=seed)
    _well_formed(g)
    N = 1 << scale
    ids = np.arange(N)
    quad = np.array([[a, b], [c, d]])
    P = np.ones((N, N))                      # P[s, d]: one draw lands on (s, d)
    for lvl in range(scale):
        bit = (ids >> lvl) & 1
        P *= quad[bit[:, None], bit[None, :]]
    draws = round(ef * N)
    Q = P + P.T                              # the unordered pair {s, d}
    hit = 1 - (1 - Q) ** draws
    np.fill_diagonal(hit, 0)                 # self-loops dropped
    iu = np.triu_indices(N, 1)
    exp_m = hit[iu].sum()
    sd = np.sqrt((hit[iu] * (1 - hit[iu])).sum())
    m = g.num_edges // 2
    assert abs(m - exp_m) < 4 * sd + 1, (m, exp_m, sd)
    deg = _degrees(g)
    exp_deg = hit.sum(axis=1)
    exp_d2 = (exp_deg ** 2 + (hit * (1 - hit)).sum(axis=1)).sum()
    assert abs((deg.astype(np.float64) ** 2).sum() / exp_d2 - 1) < 0.06
    assert abs(deg[0] / exp_deg[0] - 1) < 0.1     # vertex 0 is the R-MAT hub (quadrant a)


@pytest.mark.parametrize("points,seed", [(1024, 42), (4096, 7)])
def test_delaunay_euler_count(points, seed):
    from scipy.spatial import ConvexHull
    g = S.gen_delaunay(points, seed=seed)
    _well_formed(g)
    rng = np.random.default_rng(seed)
    pts = rng.random((points, 2))            # the same points the generator triangulates
    h = len(ConvexHull(pts).vertices)
    assert g.num_edges // 2 == 3 * points - 3 - h
    assert _degrees(g).min() >= 2


def test_delaunay_degrees_stay_below_the_histogram_threshold():
    g = S.gen_delaunay(1 << 14, seed=42)
    assert _degrees(g).max() < 64 and abs(_degrees(g).mean() - 6) < 0.05


def _load_specs():
    spec = importlib.util.spec_from_file_location("mgh", os.path.join(HERE, "golden", "make_generator_hashes.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_generator_golden_hashes():
    mod = _load_specs()
    with open(os.path.join(HERE, "golden", "generators.json")) as f:
        want = json.load(f)
    assert sorted(want) == sorted(mod.SPECS)
    for name, (kind, p) in mod.SPECS.items():
        assert mod.compute(kind, p) == want[name], name


def test_generators_are_deterministic_across_thread_counts():
    code = ("import sys, hashlib, numpy as np; sys.path.insert(0, %r); import streamgc_amd as S\n"
            "g = S.gen_chung_lu(1 << 15, 1 << 18, 2.3, 2000.0, seed=9)\n"
            "r = S.gen_rmat(13, 24.0, seed=9)\n"
            "s = S.dense_growth_stream(r, 30000, seed=9)\n"
            "h = hashlib.sha256()\n"
            "for a in (g.row_ptr, g.col_idx, r.row_ptr, r.col_idx, S.core_numbers(r), s): h.update(np.ascontiguousarray(a).tobytes())\n"
            "print(h.hexdigest())\n") % ROOT
    out = []
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        out.append(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                                  check=True).stdout.strip())
    assert len(set(out)) == 1 and len(out[0]) == 64
