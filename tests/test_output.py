"""The reference's result files, written by the product (no device needed).

Formats: A/src/streamgc.cu:595-642 (.coloring raw u16, .coloring.txt) and
A/src/streamgc.cu:1488-1573 (.coloring.meta JSON, keys in this order).
"""
import json

import numpy as np

import streamgc_amd as S

META_KEYS = ["algorithm", "graph", "num_vertices", "num_edges", "num_colors_used", "static_greedy_colors",
             "color_ratio", "init_time_ms", "bench_time_ms", "bench_stream_ms", "repair_time_ms",
             "throughput_mups", "effective_throughput_mups", "classify_ms", "local_kernel_ms",
             "boundary_kernel_ms", "gpu_sync_ms", "ghost_sync_ms", "epoch_maintenance_ms", "memcpy_ms",
             "query_latency_us", "query_count", "partition_time_ms", "partition_info", "stream_size",
             "insert_pct", "seed", "num_gpus", "hw_tag"]


def test_coloring_files_round_trip(tmp_path):
    colors = np.array([3, 1, 2, 0, 65535, 7], np.uint16)
    p = tmp_path / "streamgc_1gpu_stream1000_50pct.coloring"
    S.write_coloring(p, colors)
    assert p.stat().st_size == 2 * len(colors)
    assert np.array_equal(S.read_coloring(p), colors)
    lines = open(str(p) + ".txt").read().splitlines()
    assert lines[:3] == ["# vertices: 6", "# colors: 65535", "# algorithm: streamgc"]
    assert lines[3:] == [f"{i} {c}" for i, c in enumerate(colors.tolist())]


def test_empty_coloring(tmp_path):
    p = tmp_path / "e.coloring"
    S.write_coloring(p, np.zeros(0, np.uint16))
    assert p.stat().st_size == 0
    assert open(str(p) + ".txt").read().splitlines() == ["# vertices: 0", "# colors: 0", "# algorithm: streamgc"]


def test_meta_has_the_reference_keys_in_order(tmp_path):
    p = tmp_path / "x.coloring.meta"
    S.write_coloring_meta(p, graph='we"ird', hw_tag="MI355X", num_vertices=100, num_edges=400, num_colors_used=9,
                          static_greedy_colors=6, insert_pct=50, stream_size=100000, seed=42, init_time_ms=12.5,
                          bench_time_ms=20.0, bench_stream_ms=16.0, repair_time_ms=4.0, query_count=3)
    text = open(p).read()
    meta = json.loads(text)
    assert list(meta) == META_KEYS
    assert meta["graph"] == 'we"ird' and meta["hw_tag"] == "MI355X"
    assert meta["color_ratio"] == 1.5                      # 9 / 6, two decimals
    assert meta["throughput_mups"] == 6.25                 # 100000 / 16 ms / 1000
    assert meta["effective_throughput_mups"] == 5.0        # 100000 / 20 ms / 1000
    assert meta["num_gpus"] == 1 and meta["partition_info"] == [{"gpu": 0, "owned": 100, "ghost": 0, "edges": 400}]
