"""tal-bench-bsw (bench-bsw's command line on the engine) against the oracle.

The CLI reads a text pair file in 20,000-pair chunks into fixed slots and
scores each chunk with getScores16 on one object, as bench-bsw does
(benchmarks/bench-bsw.cpp:232-263); its -maxi lines must equal the oracle's
outputs for the same chunks on one oracle object, which carries the stale F
scratch from chunk to chunk like the reference object."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import synth as S
import tal_bsw as T

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(T.LIB_PATH), "tal-bench-bsw")


def _say(msg):
    print(msg, flush=True)


def _maxi(text):
    rows = []
    for line in text.splitlines():
        a, rest = line.split(" (")
        tq, rest = rest.split(") ")
        tle, qle = tq.split()
        gscore, max_off, gtle = rest.split()
        rows.append((int(a), int(tle), int(qle), int(gscore), int(max_off), int(gtle)))
    return np.array(rows, np.int64).reshape(-1, 6)


def _want(pairs):
    return np.stack([pairs[f].astype(np.int64)
                     for f in ("score", "tle", "qle", "gscore", "max_off", "gtle")], 1)


def _sample(n):
    # bench-bsw's slots hold 254 reference and 126 query bases
    pairs, ref, qer, _ = S.generate(2, 3 * n)
    keep = np.where((pairs["len1"] <= 254) & (pairs["len2"] <= 126))[0][:n]
    assert len(keep) == n
    return np.ascontiguousarray(pairs[keep]), ref, qer


def test_cli_matches_oracle_across_chunks(tmp_path):
    n = 40100                       # three chunks: 20000, 20000, 100
    pairs, ref, qer = _sample(n)
    path = tmp_path / "pairs.txt"
    T.write_pairfile(path, pairs, ref, qer)
    slotted = [c for c in T.read_pairfile(path)]
    want = []
    orc = O.Oracle()
    for cp, cr, cq in slotted:
        want.append(_want(orc.scores16(cp.copy(), cr, cq, 100)))
    want = np.concatenate(want)
    _say(f"[gpu] tal-bench-bsw -maxi on {n} pairs")
    r = subprocess.run([CLI, "-pairs", str(path), "-maxi"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert f"Total pairs processed: {n}" in r.stdout
    got = _maxi(r.stderr)
    assert got.shape == want.shape
    bad = np.where((got != want).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} pairs differ; first {bad[:5]}: {got[bad[:3]]} vs {want[bad[:3]]}"


def test_cli_auto_matches_library(tmp_path):
    n = 3001
    pairs, ref, qer = _sample(n)
    pairs["h0"] = np.where(np.arange(n) % 3 == 0, 0, pairs["h0"])
    path = tmp_path / "pairs.txt"
    T.write_pairfile(path, pairs, ref, qer)
    (cp, cr, cq), = list(T.read_pairfile(path))
    eng = T.BandedPairWiseSW()
    want = _want(eng.get_scores_auto(cp.copy(), cr, cq, 100))
    eng.close()
    orc = O.Oracle()
    is8 = T.is_8bit(cp)
    p8 = np.ascontiguousarray(cp[is8])
    p16 = np.ascontiguousarray(cp[~is8])
    assert 0 < len(p8) < n
    _say(f"[gpu] tal-bench-bsw -auto on {n} pairs ({len(p8)} 8-bit)")
    r = subprocess.run([CLI, "-pairs", str(path), "-maxi", "-auto"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = _maxi(r.stderr)
    assert np.array_equal(got, want)
    # the 16-bit pairs of the mix score as a 16-bit call on a fresh object
    # (no 16-bit pair ran before them on either side)
    assert np.array_equal(got[~is8], _want(orc.scores16(p16, cr, cq, 100)))
