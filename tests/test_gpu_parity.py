"""GPU parity: the HIP engine against the oracle, bit-exact on every output.

Inputs are seeded samples of the synthetic workloads plus hand-built edge
cases; every test calls the engine through the C ABI (libtal_bsw.so).
"""
import numpy as np
import pytest

import oracle as O
import synth as S
import tal_bsw as T
from helpers import assert_same_outputs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def engine():
    e = T.BandedPairWiseSW()
    yield e
    e.close()


def _oracle16(pairs, ref, qer, w, orc=None):
    o = orc or O.Oracle()
    return o.scores16(pairs.copy(), ref, qer, w)


@pytest.mark.parametrize("cfg,n", [(1, 4001), (2, 4003), (5, 4005)])
def test_workload_16bit(engine, cfg, n):
    _say(f"[gpu] workload {cfg} 16-bit n={n}")
    pairs, ref, qer, w = S.generate(cfg, n)
    want = _oracle16(pairs, ref, qer, 100)
    got = engine.get_scores16(pairs.copy(), ref, qer, 100)
    assert_same_outputs(got, want, f"workload {cfg}")


def test_workload3_per_band(engine):
    pairs, ref, qer, w = S.generate(3, 1200)
    for wv in (50, 100, 200, 500):
        sel = np.where(w == wv)[0]
        _say(f"[gpu] workload 3 w={wv} n={len(sel)}")
        p = pairs[sel].copy()
        want = _oracle16(p, ref, qer, wv)
        got = engine.get_scores16(p.copy(), ref, qer, wv)
        assert_same_outputs(got, want, f"workload 3 w={wv}")


def test_8bit_and_auto(engine):
    _say("[gpu] workload 2 8-bit + dispatcher")
    pairs, ref, qer, _ = S.generate(2, 6000)
    el = (pairs["len1"] <= 127) & (pairs["len2"] <= 127)
    p8 = pairs[el].copy()
    want8 = O.Oracle().scores8(p8.copy(), ref, qer, 100)
    got8 = engine.get_scores8(p8.copy(), ref, qer, 100)
    assert_same_outputs(got8, want8, "8-bit")
    # dispatcher: 8-bit pairs in order through getScores8, the rest through getScores16
    is8 = ((pairs["len1"] <= 127) & (pairs["len2"] <= 127) & (pairs["h0"] >= 0)
           & (pairs["h0"] + np.minimum(pairs["len1"], pairs["len2"]) <= 127))
    assert engine.count_8bit(pairs) == int(is8.sum())
    o = O.Oracle()
    want = pairs.copy()
    a = o.scores8(pairs[is8].copy(), ref, qer, 100)
    b = o.scores16(pairs[~is8].copy(), ref, qer, 100)
    for f in T.OUT_FIELDS:
        want[f][is8] = a[f]
        want[f][~is8] = b[f]
    got = engine.get_scores_auto(pairs.copy(), ref, qer, 100)
    assert_same_outputs(got, want, "auto")
