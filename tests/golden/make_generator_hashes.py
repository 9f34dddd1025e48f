"""Writes tests/golden/generators.json: seeded SHA-256 hashes of the synthetic
inputs (Chung-Lu, R-MAT, Delaunay graphs as CSR; core numbers; the dense-
growth, uniform and hold-out streams).  tests/test_generators.py recomputes
them, so a change to any generator's draws or order shows up as a mismatch.

Usage: python tests/golden/make_generator_hashes.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

SPECS = {
    "chung_lu_4096_32768_g2.3_seed42": ("chung_lu", dict(n=4096, m=32768, gamma=2.3, max_degree=600.0, seed=42)),
    "chung_lu_4096_32768_g2.3_seed42_unpermuted": ("chung_lu", dict(n=4096, m=32768, gamma=2.3, max_degree=600.0,
                                                                    seed=42, permute=False)),
    "rmat_s10_ef16_seed42": ("rmat", dict(scale=10, edge_factor=16.0, seed=42)),
    "rmat_s12_ef48_seed7": ("rmat", dict(scale=12, edge_factor=48.0, seed=7)),
    "delaunay_1024_seed42": ("delaunay", dict(num_points=1024, seed=42)),
    "core_rmat_s12_ef48_seed7": ("core", dict(scale=12, edge_factor=48.0, seed=7)),
    "dense_growth_rmat_s12_ef48_seed7_20000": ("dense_growth", dict(scale=12, edge_factor=48.0, seed=7, size=20000)),
    "uniform_rmat_s10_ef16_seed42_5000": ("uniform", dict(scale=10, edge_factor=16.0, seed=42, size=5000)),
    "holdout_chung_lu_4096_10pct_seed42": ("holdout", dict(n=4096, m=32768, pct=10, seed=42)),
}


def _h(*arrays):
    s = hashlib.sha256()
    for a in arrays:
        s.update(np.ascontiguousarray(a).tobytes())
    return s.hexdigest()


def compute(kind, p):
    import streamgc_amd as S
    if kind == "chung_lu":
        g = S.gen_chung_lu(p["n"], p["m"], p["gamma"], p["max_degree"], seed=p["seed"],
                           permute=p.get("permute", True))
        return _h(g.row_ptr, g.col_idx)
    if kind == "rmat":
        g = S.gen_rmat(p["scale"], p["edge_factor"], seed=p["seed"])
        return _h(g.row_ptr, g.col_idx)
    if kind == "delaunay":
        g = S.gen_delaunay(p["num_points"], seed=p["seed"])
        return _h(g.row_ptr, g.col_idx)
    if kind == "core":
        g = S.gen_rmat(p["scale"], p["edge_factor"], seed=p["seed"])
        return _h(S.core_numbers(g))
    if kind == "dense_growth":
        g = S.gen_rmat(p["scale"], p["edge_factor"], seed=p["seed"])
        return _h(S.dense_growth_stream(g, p["size"], seed=p["seed"]))
    if kind == "uniform":
        g = S.gen_rmat(p["scale"], p["edge_factor"], seed=p["seed"])
        return _h(S.uniform_stream(g, p["size"], seed=p["seed"]))
    if kind == "holdout":
        g = S.gen_chung_lu(p["n"], p["m"], 2.3, 600.0, seed=p["seed"])
        base, st = S.holdout_stream(g, pct=p["pct"], seed=p["seed"])
        return _h(base.row_ptr, base.col_idx, st)
    raise ValueError(kind)


def main():
    out = {name: compute(kind, p) for name, (kind, p) in SPECS.items()}
    with open(os.path.join(HERE, "generators.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
