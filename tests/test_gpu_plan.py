"""GPU tests of replayable batch plans (hipGraph capture of the dispatcher
pipeline) and of the autotuner: a plan's launches give the same outputs, call
after call, as the eager dispatcher on another handle, and every autotuner
configuration gives the same outputs."""
import numpy as np
import pytest

import oracle as O
import synth as S
import tal_bsw as T
from helpers import DeviceArrays, assert_same_outputs

pytestmark = pytest.mark.gpu


def _batch(n=24000, first=0):
    return S.generate(2, n, first)[:3]


def test_plan_replays_match_eager_sequence():
    print("[gpu] plan: 3 replays vs eager dispatcher", flush=True)
    pairs, ref, qer = _batch()
    da = DeviceArrays(pairs, ref, qer)
    db = DeviceArrays(pairs, ref, qer)
    a, b = T.BandedPairWiseSW(), T.BandedPairWiseSW()
    try:
        plan = a.plan(da.ptr(0), da.ptr(1), da.ptr(2), len(pairs), 100)
        for call in range(3):
            plan.launch()
            st = a.stats()
            b.get_scores_device(0, db.ptr(0), db.ptr(1), db.ptr(2), len(pairs), 100)
            got = da.download(0, pairs.copy())
            want = db.download(0, pairs.copy())
            assert_same_outputs(got, want, f"replay {call}")
            assert st["pairs8"] + st["pairs16"] == len(pairs)
        # the first 3000 pairs against the oracle (dispatcher split on a fresh object)
        sub = pairs[:3000].copy()
        is8 = T.is_8bit(sub)
        o = O.Oracle()
        w8 = o.scores8(sub[is8].copy(), ref, qer, 100)
        fresh = T.BandedPairWiseSW()
        got = fresh.get_scores_auto(sub.copy(), ref, qer, 100)
        fresh.close()
        for f in T.OUT_FIELDS:
            assert (got[f][is8] == w8[f]).all()
        plan.close()
    finally:
        a.close()
        b.close()
        da.close()
        db.close()


def test_plan_scores_changed_inputs():
    print("[gpu] plan: inputs changed between replays", flush=True)
    pairs, ref, qer = _batch(9000, 777)
    da = DeviceArrays(pairs, ref, qer)
    eng = T.BandedPairWiseSW()
    try:
        plan = eng.plan(da.ptr(0), da.ptr(1), da.ptr(2), len(pairs), 100)
        plan.launch()
        changed = pairs.copy()
        is8 = T.is_8bit(changed)
        sel = np.where(~is8)[0][::3]
        changed["h0"][sel] = (changed["h0"][sel] * 7 + 11) % 400 + 130  # stays 16-bit
        assert (T.is_8bit(changed) == is8).all()
        da.upload(0, changed)
        plan.launch()
        got = da.download(0, changed.copy())
        # the same two calls eagerly on a fresh handle
        ref_eng = T.BandedPairWiseSW()
        ref_eng.get_scores_auto(pairs.copy(), ref, qer, 100)
        want = ref_eng.get_scores_auto(changed.copy(), ref, qer, 100)
        ref_eng.close()
        assert_same_outputs(got, want, "changed inputs")
        plan.close()
    finally:
        eng.close()
        da.close()


def test_plan_rejects_a_changed_split():
    print("[gpu] plan: changed 8/16 split", flush=True)
    pairs, ref, qer = _batch(5000, 123)
    da = DeviceArrays(pairs, ref, qer)
    eng = T.BandedPairWiseSW()
    try:
        plan = eng.plan(da.ptr(0), da.ptr(1), da.ptr(2), len(pairs), 100)
        changed = pairs.copy()
        k = int(np.where(~T.is_8bit(changed))[0][0])
        changed["len1"][k] = 40
        changed["len2"][k] = min(40, int(changed["len2"][k]))
        changed["h0"][k] = 5
        assert T.is_8bit(changed).sum() == T.is_8bit(pairs).sum() + 1
        da.upload(0, changed)
        with pytest.raises(T.TalError) as ei:
            plan.launch()
        assert ei.value.code == T.TAL_ERANGE
        plan.close()
        # the handle still works, and its state was not touched by the rejected replay
        got = eng.get_scores_auto(pairs.copy(), ref, qer, 100)
        fresh = T.BandedPairWiseSW()
        want = fresh.get_scores_auto(pairs.copy(), ref, qer, 100)
        fresh.close()
        assert_same_outputs(got, want, "after rejected replay")
    finally:
        eng.close()
        da.close()


def test_autotune_configurations_agree():
    print("[gpu] autotune", flush=True)
    pairs, ref, qer = _batch(16000, 50000)
    da = DeviceArrays(pairs, ref, qer)
    eng = T.BandedPairWiseSW()
    try:
        res = eng.autotune(da.ptr(0), da.ptr(1), da.ptr(2), len(pairs), 100, sample=8000)
        assert len(res["trials"]) == 16 and res["ms"] > 0
        assert (res["waves_per_cu"], res["order"]) in {(t["waves_per_cu"], t["order"])
                                                        for t in res["trials"]}
        # the batch itself is untouched by the tuning runs
        back = da.download(0, pairs.copy())
        assert (back.view(np.uint8) == pairs.view(np.uint8)).all()
        got = eng.get_scores_auto(pairs.copy(), ref, qer, 100)
        for wv, order in ((4, 0), (16, 1)):
            other = T.BandedPairWiseSW()
            other.set_tuning(wv, order)
            want = other.get_scores_auto(pairs.copy(), ref, qer, 100)
            other.close()
            assert_same_outputs(got, want, f"tuning {wv}/{order}")
    finally:
        eng.close()
        da.close()
