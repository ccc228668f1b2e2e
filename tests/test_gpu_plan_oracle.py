"""GPU tests of replayable plans against the oracle: both halves of the
dispatcher (8-bit and 16-bit) of two consecutive plan launches on one handle
equal two consecutive calls on one oracle object, on samples of workloads 2,
3 (each band class) and 5, and on a batch built to drive the 16-bit
stale-column rerun path (tests/stale_batch.py), with the rerun rounds inside
the graph and with none (every rerun then runs after the replay)."""
import numpy as np
import pytest

import oracle as O
import stale_batch
import synth as S
import tal_bsw as T
from helpers import DeviceArrays, assert_same_outputs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def oracle_auto(o, pairs, ref, qer, w):
    """One dispatcher call on oracle object o: getScores8 on the pairs the
    8/16-bit rule picks (input order), getScores16 on the rest."""
    is8 = T.is_8bit(pairs)
    out = pairs.copy()
    a = o.scores8(pairs[is8].copy(), ref, qer, w)
    b = o.scores16(pairs[~is8].copy(), ref, qer, w)
    for f in T.OUT_FIELDS:
        out[f][is8] = a[f]
        out[f][~is8] = b[f]
    return out


def _plan_twice(pairs, ref, qer, w, rounds=None, label=""):
    """Two launches of one plan on a fresh handle, each checked against the
    same call sequence on a fresh oracle object; returns the handle's stats
    of both launches."""
    d = DeviceArrays(pairs, ref, qer)
    eng = T.BandedPairWiseSW()
    o = O.Oracle()
    stats = []
    try:
        if rounds is not None:
            eng.set_plan_rounds(rounds)
        plan = eng.plan(d.ptr(0), d.ptr(1), d.ptr(2), len(pairs), w)
        try:
            for call in range(2):
                plan.launch()
                stats.append(eng.stats())
                got = d.download(0, pairs.copy())
                want = oracle_auto(o, pairs, ref, qer, w)
                assert_same_outputs(got, want, f"{label} launch {call}")
        finally:
            plan.close()
    finally:
        eng.close()
        d.close()
        o.close()
    return stats


@pytest.mark.parametrize("cfg,n", [(2, 4000), (5, 4000)])
def test_plan_workload_sample(cfg, n):
    _say(f"[gpu] plan vs oracle: workload {cfg}, {n} pairs")
    pairs, ref, qer, _ = S.generate(cfg, n, first=10_000)
    st = _plan_twice(pairs, ref, qer, 100, label=f"workload {cfg}")
    if cfg == 2:
        assert st[0]["pairs8"] > 0 and st[0]["pairs16"] > 0


def test_plan_workload3_band_classes():
    pairs, ref, qer, w = S.generate(3, 2400, first=3000)
    for wv in (50, 100, 200, 500):
        sel = np.where(w == wv)[0]
        _say(f"[gpu] plan vs oracle: workload 3 w={wv}, {len(sel)} pairs")
        _plan_twice(pairs[sel].copy(), ref, qer, wv, label=f"workload 3 w={wv}")


def _stale_inputs():
    """The stale-column batch, then 8-bit pairs of workload 2 after it (so the
    16-bit groups are the batch's own)."""
    pairs, ref, qer = stale_batch.make(32, 29)
    w2, r2, q2, _ = S.generate(2, 3000)
    w2 = w2[T.is_8bit(w2)].copy()
    w2["idr"] += len(ref)
    w2["idq"] += len(qer)
    return np.concatenate([pairs, w2]), np.concatenate([ref, r2]), np.concatenate([qer, q2])


def test_stale_batch_has_teeth():
    """The oracle reads nonzero stale values in this batch, and ignoring them
    (F[ncol] cleared per group) changes some outputs, so a rerun path that did
    nothing would fail the tests below."""
    pairs, ref, qer = _stale_inputs()
    o = O.Oracle()
    a = oracle_auto(o, pairs, ref, qer, 100)
    assert o.stats()["stale_nonzero"] > 0
    z = O.Oracle()
    z.set_stale_mode(1)
    b = oracle_auto(z, pairs, ref, qer, 100)
    assert any((a[f] != b[f]).any() for f in T.OUT_FIELDS)


def test_stale_batch_eager():
    _say("[gpu] stale-column batch, eager 16-bit calls")
    pairs, ref, qer = _stale_inputs()
    sel = ~T.is_8bit(pairs)
    p16 = pairs[sel].copy()
    eng = T.BandedPairWiseSW()
    o = O.Oracle()
    for call in range(2):
        got = eng.get_scores16(p16.copy(), ref, qer, 100)
        st = eng.stats()
        assert st["reruns16"] > 0, f"call {call}: the rerun path did not run"
        assert_same_outputs(got, o.scores16(p16.copy(), ref, qer, 100), f"eager call {call}")
    eng.close()


def test_stale_batch_settled_in_launch():
    """In input order each unit's two groups share a wave: the long group,
    whose run does not read F[ncol], is final when it ends, so the group that
    takes its true mask from it is settled and run again inside the main
    launch (not by the resolution after it); outputs as the oracle's."""
    _say("[gpu] stale-column batch, groups settled inside the launch")
    pairs, ref, qer = stale_batch.make(32, 29)
    eng = T.BandedPairWiseSW()
    eng.set_tuning(8, 0)
    o = O.Oracle()
    got = eng.get_scores16(pairs.copy(), ref, qer, 100)
    st = eng.stats()
    eng.close()
    assert st["settled16"] > 0, "no group was settled inside the launch"
    assert st["settled16"] <= st["reruns16"]
    assert_same_outputs(got, o.scores16(pairs.copy(), ref, qer, 100), "settled in the launch")


@pytest.mark.parametrize("rounds", [2, 0])
def test_stale_batch_plan(rounds):
    _say(f"[gpu] stale-column batch, plan with {rounds} rerun rounds in the graph")
    pairs, ref, qer = _stale_inputs()
    st = _plan_twice(pairs, ref, qer, 100, rounds=rounds, label=f"rounds {rounds}")
    for s in st:
        assert s["reruns16"] > 0 and s["pairs8"] > 0
