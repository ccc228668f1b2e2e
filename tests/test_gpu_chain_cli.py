"""tal-bench-dp-chaining on the GPU: bench-dp-chaining's anchors file in, its
DEBUG output ("f p" per anchor, set after set) out, equal to the oracle's f
and p for every anchor."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import synth as S
import tal_chain as T

pytestmark = pytest.mark.gpu


def test_cli_debug_output_matches_oracle(tmp_path):
    print("[gpu] tal-bench-dp-chaining 150 sets", flush=True)
    assert os.path.exists(T.CLI_PATH), "tal-bench-dp-chaining not built"
    a, off = S.chain_generate(150, 0, 1500, seed=41)
    p = tmp_path / "anchors.txt"
    T.write_anchor_file(p, a, off)
    r = subprocess.run([T.CLI_PATH, str(p), "-debug"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([ln.split() for ln in r.stdout.splitlines()], dtype=np.int64)
    f, pp, _ = O.chain_dp_sets(a, off, nthreads=8)
    assert got.shape == (len(a), 2)
    assert np.array_equal(got[:, 0], f.astype(np.int64))
    assert np.array_equal(got[:, 1], pp.astype(np.int64))
    assert "M anchors/s" in r.stderr
