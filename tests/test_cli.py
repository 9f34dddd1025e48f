"""streaming_gc, the reference's command line (A/src/streamgc.cu:653-1581).

CPU: the usage error.  GPU: a deterministic run on a golden graph writes
the reference's three result files, prints the lines the reference's
benchmark scripts parse (A/test/benchmark/bench_throughput.sh), and its
colouring is the oracle's for the same init stream, churn stream and
repairs.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import streamgc_amd as S

CLI = os.path.join(os.path.dirname(S.__file__), "streaming_gc")


def _cli():
    if not os.path.exists(CLI):
        pytest.fail("streaming_gc is not built (python -m streamgc_amd.build)")
    return CLI


def test_usage_without_graph():
    r = subprocess.run([_cli(), "--stream", "10"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert r.stderr.startswith("Usage: ") and "--graph <path.mtx>" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["delaunay1k", "chunglu4k"])
def test_cli_run_matches_the_oracle(name, graph_path, tmp_path):
    stream, pct, seed, epoch = 3000, 50, 7, 96
    r = subprocess.run([_cli(), "--graph", graph_path(name), "--stream", str(stream), "--insert-pct", str(pct),
                        "--seed", str(seed), "--epoch-size", str(epoch), "--deterministic", "--validate",
                        "--hw-tag", "test_hw", "--output-dir", str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    for key in ("Throughput:", "effective:", "Init time:", "Bench time:", "repair:", "Colors used:", "ratio:",
                "Coloring valid: 0 violations", "Done."):
        assert key in out, key

    g = S.load_graph_csr(graph_path(name))
    og = O.Csr(g.n, g.row_ptr, g.col_idx, g.degree)
    orc = O.Engine(og, O.default_config(epoch_size=epoch))
    init = S.generate_init_stream(g)
    orc.submit(init["u"], init["v"], init["type"])
    orc.finish()
    bench = S.generate_bench_stream_edges(g, stream, pct, seed)
    orc.submit(bench["u"], bench["v"], bench["type"])
    orc.finish()
    want = orc.colors()

    path = tmp_path / "out" / f"streamgc_1gpu_stream{stream}_{pct}pct.coloring"
    got = S.read_coloring(path)
    assert np.array_equal(got, want)
    txt = open(str(path) + ".txt").read().splitlines()
    assert txt[:3] == [f"# vertices: {g.n}", f"# colors: {want.max()}", "# algorithm: streamgc"]
    meta = json.load(open(str(path) + ".meta"))
    greedy = O.static_greedy(og)
    assert meta["graph"] == name and meta["hw_tag"] == "test_hw"
    assert (meta["num_vertices"], meta["num_edges"]) == (g.n, g.num_edges)
    assert meta["num_colors_used"] == int(want.max())
    assert meta["static_greedy_colors"] == greedy
    assert meta["color_ratio"] == round(int(want.max()) / greedy, 2)
    assert (meta["stream_size"], meta["insert_pct"], meta["seed"], meta["num_gpus"]) == (stream, pct, seed, 1)
    assert f"Colors used: {want.max()}" in out
    assert f"Static greedy: {greedy} colors" in out
    m = re.search(r"Throughput: ([\d.]+) M updates/sec \(effective: ([\d.]+)", out)
    assert float(m.group(1)) >= float(m.group(2)) > 0
