"""tal-bench-sw (bench-smith-waterman's command line on the engine, through
the PairWiseSW C++ shim) against the oracle: a pair file of two lines per
pair, one score printed per pair."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import tal_sw as T

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(T.LIB_PATH), "tal-bench-sw")


def _pairs_file(path, rng, n, hi):
    A = np.frombuffer(b"ACGT", np.uint8)
    s1, s2 = [], []
    with open(path, "wb") as f:
        for _ in range(n):
            a = A[rng.integers(0, 4, rng.integers(1, hi + 1))]
            b = np.resize(a, int(rng.integers(1, hi + 1))).copy()
            m = rng.random(len(b)) < 0.2
            b[m] = A[rng.integers(0, 4, int(m.sum()))]
            s1.append(a)
            s2.append(b)
            f.write(a.tobytes() + b"\n" + b.tobytes() + b"\n")
    return s1, s2


@pytest.mark.parametrize("precision,hi", [(32, 300), (16, 600), (8, 124)])
@pytest.mark.parametrize("weights", [(1, -1, 1, 1), (2, -3, -5, -2)])
def test_cli_matches_oracle(tmp_path, precision, hi, weights):
    print(f"[gpu] tal-bench-sw precision={precision} weights={weights}", flush=True)
    rng = np.random.default_rng(precision * 10 + weights[0])
    path = tmp_path / "pairs.txt"
    s1, s2 = _pairs_file(path, rng, 1500, hi)
    args = [CLI, "-pairs", str(path), "-match", str(weights[0]), "-mismatch", str(weights[1]),
            "-gapo", str(weights[2]), "-gape", str(weights[3]), "-precision", str(precision)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = np.array([int(x) for x in r.stdout.splitlines()[:len(s1)]])
    pairs, buf = O.sw_make_batch(s1, s2, precision)
    want = O.sw_scores(precision, pairs, buf, weights, nthreads=8)["score"]
    assert np.array_equal(got, want)
    assert "SW GCUPS" in r.stdout
