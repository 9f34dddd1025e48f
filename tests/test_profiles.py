"""The profile summaries under profiles/ quote the commands that made them.
Each quoted command must run committed files only: bench.py with options it
has, rocprofv3 with the program straight after `--`, and no working-file
paths.  (CPU only: the commands are parsed, not run.)"""
import os
import re
import shlex

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(ROOT, "profiles")
MDS = sorted(f for f in os.listdir(PROFILES) if f.endswith(".md"))


def quoted_commands(path):
    """Lines of the ``` blocks of a summary, with trailing comments removed."""
    text = open(path).read()
    out = []
    for block in re.findall(r"```\n(.*?)```", text, re.S):
        for line in block.splitlines():
            line = line.split("  #")[0].strip()
            if line:
                out.append(line)
    return out


@pytest.mark.parametrize("name", MDS)
def test_summary_commands_run_committed_files(name):
    cmds = quoted_commands(os.path.join(PROFILES, name))
    assert cmds, f"{name} quotes no command"
    parser = bench.arg_parser()
    for cmd in cmds:
        assert "_work" not in cmd and "/root/reference" not in cmd, (name, cmd)
        tok = shlex.split(cmd)
        if tok[0] == "rocprofv3":
            at = tok.index("--")
            assert tok[at + 1] in ("python3", "python"), (name, cmd)   # the program itself, no launcher
            prog = tok[at + 1:]
        else:
            prog = tok
        if "bench.py" not in prog:
            continue
        args = prog[prog.index("bench.py") + 1:]
        try:
            parser.parse_args(args)
        except SystemExit:
            pytest.fail(f"{name}: bench.py does not take {args}")


def test_every_summary_is_checked():
    assert len(MDS) >= 9, MDS
