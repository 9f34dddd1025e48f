"""The host's reading of the captured rounds' control block (CPU only).

streamgc_amd/csrc/sgc_round_control.hpp decides, after a run of replays,
which rounds ran, which round the host finishes, and where the next replays
start.  A control block that cannot be right (a stop round behind the rounds
just replayed, or past them; a next round that does not match) must end the
run with an error: before the check, a stale stop round sent the host loop
back to rounds it had run and it replayed them without end.  The header is
plain C++, so the test compiles it with g++ and drives it directly, including
a host loop over a simulated device.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include "sgc_round_control.hpp"
using namespace sgc;

static int fails = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL line %d: %s\n", __LINE__, #c); fails++; } } while (0)

// Host loop of batch_run over a simulated device: replays go in chunks of
// `chunk`; `stops` lists rounds that stop for the host; `stale` makes the
// device report stop round 0 every time (the fault the check guards against).
static int host_loop(unsigned rounds, unsigned chunk, const unsigned* stops, unsigned nstops, bool stale,
                     unsigned* iters, unsigned* host_rounds) {
    unsigned r = 0, next_stop = 0;
    *iters = *host_rounds = 0;
    while (r < rounds) {
        if (++*iters > 4 * rounds + 4) return -2;          // would never end
        const unsigned end = r + chunk < rounds ? r + chunk : rounds;
        unsigned ctl_round = end, stop = 0, stop_round = 0;
        while (next_stop < nstops && stops[next_stop] < r) next_stop++;
        if (next_stop < nstops && stops[next_stop] < end) {
            stop = 1;
            stop_round = ctl_round = stops[next_stop];
        }
        if (stale && stop) stop_round = ctl_round = 0;
        RoundReading rd;
        if (!read_rounds(r, end, rounds, ctl_round, stop, stop_round, &rd)) return -1;
        if (rd.stopped) { ++*host_rounds; next_stop++; }
        r = rd.next;
    }
    return 0;
}

int main() {
    RoundReading rd;
    // every replay ran
    CHECK(read_rounds(0, 4, 10, 4, 0, 0, &rd) && !rd.stopped && rd.done == 4 && rd.next == 4);
    // a stop at round 3 of [2, 6): rounds 2 ran, the host finishes 3, replays resume at 4
    CHECK(read_rounds(2, 6, 10, 3, 1, 3, &rd) && rd.stopped && rd.done == 3 && rd.host == 3 && rd.next == 4);
    CHECK(read_rounds(2, 6, 10, 2, 1, 2, &rd) && rd.stopped && rd.done == 2 && rd.next == 3);
    CHECK(read_rounds(9, 10, 10, 9, 1, 9, &rd) && rd.next == 10);
    // stale or impossible readings
    CHECK(!read_rounds(5, 8, 10, 2, 1, 2, &rd));      // stop round behind the replayed rounds
    CHECK(!read_rounds(5, 8, 10, 8, 1, 8, &rd));      // stop round past them
    CHECK(!read_rounds(0, 4, 10, 3, 0, 0, &rd));      // no stop, but the device did not reach `end`
    CHECK(!read_rounds(0, 4, 10, 5, 0, 0, &rd));
    CHECK(!read_rounds(2, 6, 10, 4, 1, 3, &rd));      // stopped, next round moved past the stop
    CHECK(!read_rounds(2, 6, 10, 3, 7, 3, &rd));      // not a stop flag the device writes
    CHECK(!read_rounds(4, 4, 10, 4, 0, 0, &rd));      // empty replay range
    CHECK(!read_rounds(8, 12, 10, 12, 0, 0, &rd));    // past the plan
    // host loops: end, every stop finished once, at most `rounds` iterations
    unsigned it, hr;
    const unsigned stops[] = {0, 3, 4, 9, 15, 31};
    CHECK(host_loop(32, 1, stops, 6, false, &it, &hr) == 0 && hr == 6 && it == 32);
    CHECK(host_loop(32, 8, stops, 6, false, &it, &hr) == 0 && hr == 6 && it <= 32);
    CHECK(host_loop(32, 32, stops, 6, false, &it, &hr) == 0 && hr == 6);
    CHECK(host_loop(32, 5, nullptr, 0, false, &it, &hr) == 0 && hr == 0 && it == 7);
    // a device that reports a stale stop round: an error, not an endless replay
    CHECK(host_loop(32, 8, stops + 1, 5, true, &it, &hr) == -1);
    std::printf("%s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_round_control_reading():
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "rc.cpp")
        with open(src, "w") as f:
            f.write(DRIVER)
        exe = os.path.join(td, "rc")
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "streamgc_amd", "csrc"),
                        "-o", exe, src], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout
