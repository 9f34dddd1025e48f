"""Oracle colour counts at full configuration scale (committed as scale_reference.json).

The oracle replays the same seeded workload the benchmark runs -- the init
stream in the reference's epoch order, then the churn stream with a repair and
publish every epoch -- and records the maximum colour (the reference's colour
count, A/src/streamgc.cu:549-553) after the init repair, after every full
epoch, and after the final repair of the stream's tail (final_max_color).
The GPU engine's parallel results are compared against these numbers.

Usage: python tests/golden/make_scale_reference.py c1 [c2 c3 ...]   (CPU; c2 and c3 take minutes)
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
import streamgc_amd as S  # noqa: E402

OUT = os.path.join(HERE, "scale_reference.json")
E = 32768


def graph(name):
    if name == "c1":
        return S.gen_delaunay(1 << 17, seed=42), 100000
    if name == "c2":   # 1,000,000 updates: 30 full epochs and a 16,960-update tail
        return S.gen_chung_lu(3997962, 34681189, 2.3, 15000.0, seed=42, permute=True), 1000000
    if name == "c3":
        return S.gen_rmat(20, 48.0, seed=42), 1000000
    raise SystemExit(f"unknown workload {name}")


def stream(name, g, size):
    if name == "c3":   # dense growth inside the 90th-percentile k-core, 2 inserts per delete
        return S.dense_growth_stream(g, size, core_pct=90, ins_per_del=2, seed=42)
    return S.generate_bench_stream_edges(g, size, 50, 42)


def run(name, exact):
    g, size = graph(name)
    og = O.Csr(g.n, g.row_ptr, g.col_idx, g.degree)
    flags = O.FLAG_EXACT_HISTOGRAM if exact else 0
    eng = O.Engine(og, O.default_config(flags=flags), threads=os.cpu_count() or 1)
    init = S.generate_init_stream(g)
    t = time.time()
    eng.submit(init["u"], init["v"], init["type"], max_epochs=1)
    fin = eng.finish()
    t_init = time.time() - t
    eng.set_flags(flags | O.FLAG_REPAIR_EVERY_EPOCH)
    bench = stream(name, g, size)
    t = time.time()
    eps = eng.submit(bench["u"], bench["v"], bench["type"])
    last = eng.finish()
    t_bench = time.time() - t
    return {"vertices": g.n, "edges": g.num_edges // 2, "stream": size, "epoch_size": E,
            "exact_histogram": exact, "init_max_color": fin["max_color"],
            "epoch_max_color": [e["max_color"] for e in eps], "final_max_color": last["max_color"],
            "static_greedy": S.static_greedy_coloring(g), "oracle_seconds": round(t_init + t_bench, 1)}


def main():
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in sys.argv[1:]:
        for exact in (False, True):
            key = f"{name}/{'exact' if exact else 'reference'}"
            data[key] = run(name, exact)
            print(key, data[key]["init_max_color"], data[key]["final_max_color"], data[key]["oracle_seconds"],
                  flush=True)
            json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
