"""Mutation check of the oracle's own tests.

Each mutant below is a plausible slip in oracle/oracle.cpp (a dropped call, a
swapped argument, a reversed loop, an off-by-one in a stop rule).  The test
writes a copy of the oracle with that one slip into a temporary directory,
builds it with g++, and runs tests/test_oracle.py against the copy (through
SGC_ORACLE_LIB).  The oracle's tests must fail on every mutant: a mutant that
survives marks a behaviour the oracle restates without a check.

CPU only; the copies never leave the temporary directory.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "oracle.cpp")

# name -> (text in oracle.cpp, replacement).  Each text must occur exactly once.
MUTANTS = {
    # A/src/kernels/local_update.cu:139-141 and recolor.cu:145-163
    "delete_without_reduction": (
        "            try_reduce(u, fu);\n            try_reduce(v, fv);\n", ""),
    "reduction_freed_colours_swapped": (
        "            try_reduce(u, fu);\n            try_reduce(v, fv);\n",
        "            try_reduce(u, fv);\n            try_reduce(v, fu);\n"),
    "reduction_on_equal_freed_colour": (
        "        if (freed >= cur) return;", "        if (freed > cur) return;"),
    # A/src/kernels/recolor.cu:50-82,101-111
    "recolour_without_propagation": (
        "        word[v] = pack(newc, (uint16_t)(version_of(word[v]) + 1));\n        propagate(v, oldc, newc);\n",
        "        word[v] = pack(newc, (uint16_t)(version_of(word[v]) + 1));\n"),
    # A/include/color_select.cuh:120-157 (bits are only ever set)
    "propagation_clears_the_old_bit": (
        "        if (newc != 0 && newc <= 64) bitmap[v] |= 1ULL << (newc - 1);\n",
        "        if (oldc != 0 && oldc <= 64) bitmap[v] &= ~(1ULL << (oldc - 1));\n"
        "        if (newc != 0 && newc <= 64) bitmap[v] |= 1ULL << (newc - 1);\n"),
    "propagation_histogram_signs_swapped": (
        "            if (oldc != 0 && oldc <= HIST_BINS) h[oldc - 1] -= 1;\n"
        "            if (newc != 0 && newc <= HIST_BINS) h[newc - 1] += 1;\n",
        "            if (oldc != 0 && oldc <= HIST_BINS) h[oldc - 1] += 1;\n"
        "            if (newc != 0 && newc <= HIST_BINS) h[newc - 1] -= 1;\n"),
    # A/include/color_select.cuh:188-209
    "delete_clears_the_bitmap_bit": (
        "        if (c == 0) return;\n        if (hd_map[v] != INVALID && c <= HIST_BINS) hist[(size_t)hd_map[v] * HIST_BINS + c - 1] -= 1;\n",
        "        if (c == 0) return;\n        if (c <= 64) bitmap[v] &= ~(1ULL << (c - 1));\n"
        "        if (hd_map[v] != INVALID && c <= HIST_BINS) hist[(size_t)hd_map[v] * HIST_BINS + c - 1] -= 1;\n"),
    # A/src/kernels/local_update.cu:95-101 (Q1: v's histogram counts u's colour twice)
    "first_colouring_without_the_early_count": (
        "                on_new_neighbor(v, color(u));\n                word[v] = pack(fas(v), 1);\n",
        "                word[v] = pack(fas(v), 1);\n"),
    # A/src/kernels/epoch_maintenance.cu:20-43
    "compaction_packs_in_reverse": (
        "            for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)\n"
        "                if (col[j] != INVALID) col[w++] = col[j];\n",
        "            std::vector<uint32_t> keep;\n"
        "            for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)\n"
        "                if (col[j] != INVALID) keep.push_back(col[j]);\n"
        "            for (size_t k = keep.size(); k-- > 0;) col[w++] = keep[k];\n"),
    # A/src/kernels/epoch_maintenance.cu:66-87
    "drain_newest_first": (
        "        for (uint32_t i = 0; i < m; i++) {\n            uint32_t v = overflow[i].first, nb = overflow[i].second;\n",
        "        for (uint32_t k = 0; k < m; k++) {\n            uint32_t i = m - 1 - k;\n"
        "            uint32_t v = overflow[i].first, nb = overflow[i].second;\n"),
    # A/src/kernels/epoch_maintenance.cu:119-123
    "hub_bitmaps_rebuilt": (
        "                if (!exact) continue;\n",
        "                if (!exact) {\n"
        "                    uint64_t hb = 0;\n"
        "                    for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {\n"
        "                        uint32_t nb = col[j];\n"
        "                        if (nb == INVALID || nb >= n) continue;\n"
        "                        uint16_t c = color(nb);\n"
        "                        if (c != 0 && c <= 64) hb |= 1ULL << (c - 1);\n"
        "                    }\n"
        "                    bitmap[v] = hb;\n"
        "                    continue;\n"
        "                }\n"),
    # A/src/kernels/conflict_repair.cu:35-43,100-112
    "tier1_without_own_colour_bit": (
        "                if (mine <= 64) bitmap[v] |= 1ULL << (mine - 1);\n", ""),
    "tier1_uncoloured_not_counted": (
        "                    word[v] = pack(fas(v), (uint16_t)(version_of(word[v]) + 1));\n                    conflicts++;\n",
        "                    word[v] = pack(fas(v), (uint16_t)(version_of(word[v]) + 1));\n"),
    # A/src/streamgc.cu:1378-1389
    "plateau_stop_one_sweep_later": (
        "                if (++plateau >= cfg.plateau_iters) return;", "                if (++plateau > cfg.plateau_iters) return;"),
    # A/src/streamgc.cu:1197-1206 (Tier-2 fixes the higher-id endpoint)
    "tier2_fixes_the_lower_id_endpoint": (
        "                    if (color(u) == color(v) && color(u) != 0) to_fix.push_back(v);",
        "                    if (color(u) == color(v) && color(u) != 0) to_fix.push_back(u);"),
    # A/include/color_select.cuh:71-88 (lowest index among equal minima)
    "histogram_minimum_tie_takes_the_last": (
        "            if (h[c] < mc) { mc = h[c]; mi = c; }", "            if (h[c] <= mc) { mc = h[c]; mi = c; }"),
    # A/src/kernels/local_update.cu:22-26 (Q8: first hole, or an existing copy)
    "insert_skips_holes_before_a_copy": (
        "            if (col[j] == INVALID) { col[j] = nb; return; }\n            if (col[j] == nb) return;\n        }\n"
        "        if (cfg.flags & FLAG_ABL_NO_OVERFLOW) return;",
        "            if (col[j] == nb) return;\n        }\n"
        "        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)\n"
        "            if (col[j] == INVALID) { col[j] = nb; return; }\n"
        "        if (cfg.flags & FLAG_ABL_NO_OVERFLOW) return;"),
}


def _build_and_test(name, src, tmp):
    old, new = MUTANTS[name]
    assert src.count(old) == 1, f"{name}: the mutated text no longer occurs exactly once in oracle.cpp"
    d = os.path.join(tmp, name)
    os.makedirs(d)
    cpp = os.path.join(d, "oracle.cpp")
    with open(cpp, "w") as f:
        f.write(src.replace(old, new))
    so = os.path.join(d, "liboracle.so")
    b = subprocess.run(["g++", "-O1", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-w", "-o", so, cpp],
                       capture_output=True, text=True)
    if b.returncode != 0:
        return name, None, b.stderr[-2000:]
    env = dict(os.environ, SGC_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle.py"), "-q", "-x",
                        "-p", "no:cacheprovider"], capture_output=True, text=True, env=env, cwd=ROOT)
    return name, r.returncode, r.stdout[-600:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_every_oracle_mutant_fails_the_oracle_tests():
    with open(ORACLE) as f:
        src = f.read()
    tmp = tempfile.mkdtemp(prefix="sgc_mut_")
    try:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            results = list(ex.map(lambda n: _build_and_test(n, src, tmp), sorted(MUTANTS)))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    broken = [(n, out) for n, rc, out in results if rc is None]
    assert not broken, f"mutants that do not build: {broken}"
    survivors = [n for n, rc, _ in results if rc == 0]
    assert not survivors, f"oracle mutants that pass tests/test_oracle.py: {survivors}"
    # pytest exits 1 on failed tests; any other code (a crash, a collection error) is not a kill
    odd = [(n, rc, out) for n, rc, out in results if rc not in (0, 1)]
    assert not odd, f"mutant runs that ended abnormally: {odd}"
