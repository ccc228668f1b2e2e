"""bench-dp-chaining's anchors file and the tal-bench-dp-chaining command
line without a GPU (-parse-only): the file round trip, the counts, and the
errors on malformed input."""
import os
import subprocess

import numpy as np
import pytest

import synth as S
import tal_chain as T

needs_cli = pytest.mark.skipif(not os.path.exists(T.CLI_PATH), reason="tal-bench-dp-chaining not built")


def test_anchor_file_round_trip(tmp_path):
    a, off = S.chain_generate(7, 0, 40, seed=3)
    a["r"][0] = (1 << 64) - 1  # full 64-bit r
    a["q"][1] = -5
    p = tmp_path / "anchors.txt"
    T.write_anchor_file(p, a, off)
    b, boff = T.read_anchor_file(p)
    assert np.array_equal(boff, off)
    assert all(np.array_equal(a[f], b[f]) for f in ("r", "q", "l"))


@needs_cli
def test_cli_parse_only_counts(tmp_path):
    a, off = S.chain_generate(9, 0, 50, seed=4)
    p = tmp_path / "anchors.txt"
    T.write_anchor_file(p, a, off)
    r = subprocess.run([T.CLI_PATH, str(p), "-parse-only"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert f"9 sets, {len(a)} anchors" in r.stderr


@needs_cli
@pytest.mark.parametrize("text,msg", [("3\n1 2 3\n4 5 6\n", "truncated"), ("-1\n", "bad set size")])
def test_cli_rejects_malformed_files(tmp_path, text, msg):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    r = subprocess.run([T.CLI_PATH, str(p), "-parse-only"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and msg in r.stderr


@needs_cli
def test_cli_without_arguments():
    r = subprocess.run([T.CLI_PATH], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "Anchors_file_name" in r.stderr
