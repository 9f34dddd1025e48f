"""T_rc microbenchmark (A/benchmark/microbench/t_rc_atomic.cu:18-62) through the C ABI.

Each wavefront's lane 0 bumps its word's version and toggles its colour's low
bit by dependent CAS.  A word no other wavefront touches ends, after R runs
of I iterations, at version R*I (16-bit) with colour (R*I) & 1; a word all
wavefronts share has had one toggle per successful CAS, so its colour bit is
its version's parity.
"""
import pytest

import streamgc_amd as S

pytestmark = pytest.mark.gpu


def _word(total):
    return ((total & 0xFFFF) << 16) | (total & 1)


def test_single_wavefront_latency_and_final_word():
    r = S.microbench_trc(waves=1, iters=20001, runs=3)
    print(f"T_rc {r['ns']:.1f} ns", flush=True)
    assert 0 < r["ns"] < 1e5
    assert r["final"] == (_word(3 * 20001), _word(3 * 20001))


def test_wavefronts_on_their_own_words_all_complete():
    r = S.microbench_trc(waves=4096, iters=2000, runs=3)
    print(f"4096 words: {r['ns']:.1f} ns per iteration of one wavefront", flush=True)
    assert r["final"] == (_word(6000), _word(6000))


def test_shared_word_keeps_version_and_colour_in_step():
    r = S.microbench_trc(waves=256, shared=True, iters=1000, runs=3)
    w = r["final"][0]
    assert r["final"][0] == r["final"][1]
    assert (w >> 16) > 0 and (w & 0xFFFF) == ((w >> 16) & 1)


def test_bad_arguments_are_refused():
    with pytest.raises(S.SgcError):
        S.microbench_trc(waves=0)
    with pytest.raises(S.SgcError):
        S.microbench_trc(runs=0)
