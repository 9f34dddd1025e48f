// sgc_round_control.hpp -- the host's reading of a captured-round control
// block after a run of replays (plain C++: the engine includes it, and a CPU
// test compiles it on its own).
//
// The host replays rounds [r, end) and then reads the control block back.
// Either every replay ran (no stop, the device's next round is `end`), or the
// device stopped at some round s with r <= s < end: that replay ended early
// for the host to finish, and the replays after it did nothing, so the
// device's next round is still s.  Anything else -- a stop round outside
// [r, end), a next round that does not match -- is a control block the host
// cannot act on: the run ends with an error instead of replaying rounds
// again (each accepted reading moves r forward, so the host loop runs at
// most `rounds` times).
#pragma once

#include <cstdint>

namespace sgc {

struct RoundReading {
    uint32_t done;       // rounds fully run by the replays: [r, done)
    uint32_t host;       // the round the host finishes next (valid when stopped)
    uint32_t next;       // the first round of the next replays
    bool stopped;
};

// Returns false when the control block cannot be trusted.
inline bool read_rounds(uint32_t r, uint32_t end, uint32_t rounds, uint32_t ctl_round, uint32_t ctl_stop,
                        uint32_t ctl_stop_round, RoundReading* out) {
    if (r >= end || end > rounds) return false;
    if (ctl_stop == 0) {
        if (ctl_round != end) return false;
        *out = RoundReading{end, 0, end, false};
        return true;
    }
    if (ctl_stop != 1 || ctl_stop_round < r || ctl_stop_round >= end || ctl_round != ctl_stop_round) return false;
    *out = RoundReading{ctl_stop_round, ctl_stop_round, ctl_stop_round + 1, true};
    return true;
}

}  // namespace sgc
