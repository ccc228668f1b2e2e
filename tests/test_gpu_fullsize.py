"""Full-size parity on the GPU: every pair of workloads 1, 2, 3 (each band
class) and 5, through the graph-replayed plan (the dispatcher's semantics)
and through eager getScores16 calls, matched against the oracle's SHA-256
digests of the six output arrays (tests/golden/bsw_digests.json, written by
tests/golden/make_digests.py from oracle/bsw_oracle.c on the same seeded
inputs).  Where the oracle saw a nonzero stale F[ncol] decide a tail, the
engine's rerun counter must show that the rerun path ran."""
import hashlib
import json
import os

import numpy as np
import pytest

import synth as S
import tal_bsw as T
from helpers import DeviceArrays

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MANIFEST = os.path.join(HERE, "golden", "bsw_digests.json")
IN_FIELDS = ["idr", "idq", "len1", "len2", "h0"]


def _cases():
    if not os.path.exists(MANIFEST):  # written by tests/golden/make_digests.py
        return []
    with open(MANIFEST) as f:
        return json.load(f)["cases"]


def _say(msg):
    print(msg, flush=True)


def _inputs(c):
    pairs, ref, qer, w = S.generate(c["workload"], c["workload_pairs"])
    if c["band_class"] is not None:
        pairs = pairs[np.where(w == c["band_class"])[0]].copy()
    h = hashlib.sha256()
    for f in IN_FIELDS:
        h.update(np.ascontiguousarray(pairs[f], "<i4").tobytes())
    h.update(ref.tobytes())
    h.update(qer.tobytes())
    assert h.hexdigest() == c["inputs_sha256"], "the generator's output changed"
    return pairs, ref, qer


def _digest(pairs):
    m = np.stack([np.ascontiguousarray(pairs[f], "<i4") for f in T.OUT_FIELDS])
    return hashlib.sha256(m.tobytes()).hexdigest()


@pytest.mark.parametrize("case", [c["name"] for c in _cases()])
def test_fullsize_plan(case):
    c = next(x for x in _cases() if x["name"] == case)
    _say(f"[gpu] full size, plan: {case} ({c['pairs']} pairs, w={c['w']})")
    pairs, ref, qer = _inputs(c)
    d = DeviceArrays(pairs, ref, qer)
    eng = T.BandedPairWiseSW()
    try:
        plan = eng.plan(d.ptr(0), d.ptr(1), d.ptr(2), len(pairs), c["w"])
        plan.launch()
        st = eng.stats()
        got = d.download(0, pairs.copy())
        plan.close()
    finally:
        eng.close()
        d.close()
    assert st["pairs8"] == c["pairs_8bit"]
    assert _digest(got) == c["digest_auto"], f"{case}: plan outputs differ from the oracle's"
    if c["stale_nonzero_groups"]["auto"] > 0:
        assert st["reruns16"] > 0, f"{case}: the stale-column rerun path did not run"


@pytest.mark.parametrize("case", [c["name"] for c in _cases()])
def test_fullsize_eager16(case):
    c = next(x for x in _cases() if x["name"] == case)
    _say(f"[gpu] full size, eager getScores16: {case} ({c['pairs']} pairs, w={c['w']})")
    pairs, ref, qer = _inputs(c)
    d = DeviceArrays(pairs, ref, qer)
    eng = T.BandedPairWiseSW()
    try:
        eng.get_scores_device(16, d.ptr(0), d.ptr(1), d.ptr(2), len(pairs), c["w"])
        st = eng.stats()
        got = d.download(0, pairs.copy())
    finally:
        eng.close()
        d.close()
    assert _digest(got) == c["digest_all16"], f"{case}: outputs differ from the oracle's"
    if c["stale_nonzero_groups"]["all16"] > 0:
        assert st["reruns16"] > 0, f"{case}: the stale-column rerun path did not run"
