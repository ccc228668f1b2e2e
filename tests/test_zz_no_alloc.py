"""GPU tests of tal_bsw_get_scores_auto_device_noalloc, the dispatcher whose
call scratch comes from the library's static device array: on each input, two
calls in a row (the second on a smaller batch) through the no-allocation entry
on one handle give the same outputs and path counters as the same two calls
through tal_bsw_get_scores_auto_device on another handle, and the first pair
of calls also equals the oracle.  Covered: a mixed 8/16-bit sample of the
2M-pair uneven batch, long queries (up to 2000 columns, band 500) whose state
slabs make the static array cap the grid, and a batch that drives the 16-bit
stale-column rerun path; a batch too large for the array is refused before
any work."""

import numpy as np
import pytest

import oracle as O
import stale_batch
import synth as S
import tal_bsw as T
from helpers import DeviceArrays, assert_same_outputs

pytestmark = pytest.mark.gpu

# the dispatcher's own scratch per pair: flags, two indices, one pair copy
PER_PAIR = 2 + 4 + 4 + T.SEQPAIR_DTYPE.itemsize
STATIC_BYTES = 64 << 20


def _say(msg):
    print(msg, flush=True)


def _oracle_auto(pairs, ref, qer, w):
    o = O.Oracle()
    try:
        is8 = T.is_8bit(pairs)
        out = pairs.copy()
        a = o.scores8(pairs[is8].copy(), ref, qer, w)
        b = o.scores16(pairs[~is8].copy(), ref, qer, w)
        for f in T.OUT_FIELDS:
            out[f][is8] = a[f]
            out[f][~is8] = b[f]
        return out
    finally:
        o.close()


def _two_handles(calls, check_oracle=True):
    """Runs `calls` ((pairs, ref, qer, w) each) in order on two fresh handles,
    one through the no-allocation entry and one through the existing device
    dispatcher, and compares every call's outputs and counters."""
    a = T.BandedPairWiseSW()
    b = T.BandedPairWiseSW()
    stats = []
    try:
        for k, (pairs, ref, qer, w) in enumerate(calls):
            _say(f"[gpu] no-alloc call {k}: {len(pairs)} pairs, w={w}")
            d1 = DeviceArrays(pairs, ref, qer)
            d2 = DeviceArrays(pairs, ref, qer)
            try:
                a.get_scores_device_noalloc(d1.ptr(0), d1.ptr(1), d1.ptr(2), len(pairs), w)
                sa = a.stats()
                b.get_scores_device(0, d2.ptr(0), d2.ptr(1), d2.ptr(2), len(pairs), w)
                sb = b.stats()
                got = d1.download(0, pairs.copy())
                want = d2.download(0, pairs.copy())
            finally:
                d1.close()
                d2.close()
            assert_same_outputs(got, want, f"call {k}")
            for f in ("pairs8", "pairs16", "groups8", "groups16", "reruns16"):
                assert sa[f] == sb[f], f"call {k}: {f} {sa[f]} != {sb[f]}"
            if check_oracle and k == 0:
                assert_same_outputs(got, _oracle_auto(pairs, ref, qer, w), f"call {k} vs oracle")
            stats.append(sa)
    finally:
        a.close()
        b.close()
    return stats


def test_mixed_batch_two_calls():
    pairs, ref, qer, _ = S.generate(2, 26000)
    st = _two_handles([(pairs[:20000].copy(), ref, qer, 100),
                       (pairs[20000:].copy(), ref, qer, 100)])
    assert st[0]["pairs8"] > 0 and st[0]["pairs16"] > 0


def test_long_queries_two_calls():
    pairs, ref, qer, w = S.generate(3, 2400, first=3000)
    sel = pairs[np.where(w == 500)[0]].copy()
    assert sel["len2"].max() > 1000
    _two_handles([(sel, ref, qer, 500), (sel[: len(sel) // 3].copy(), ref, qer, 500)])


def test_stale_batch_two_calls():
    pairs, ref, qer = stale_batch.make(32, 29)
    st = _two_handles([(pairs, ref, qer, 100), (pairs[: len(pairs) // 2].copy(), ref, qer, 100)])
    assert st[0]["reruns16"] > 0, "the rerun path did not run"


def test_too_large_is_refused_before_work():
    n = STATIC_BYTES // PER_PAIR + 1
    _say(f"[gpu] no-alloc call over the static scratch: {n} pairs")
    pairs = np.zeros(n, dtype=T.SEQPAIR_DTYPE)
    pairs["score"] = -7
    ref = np.zeros(16, dtype=np.uint8)
    d = DeviceArrays(pairs, ref, ref)
    eng = T.BandedPairWiseSW()
    try:
        with pytest.raises(T.TalError) as ei:
            eng.get_scores_device_noalloc(d.ptr(0), d.ptr(1), d.ptr(2), n, 100)
        assert ei.value.code == T.TAL_ERANGE
        assert "static scratch" in str(ei.value)
        back = d.download(0, np.empty_like(pairs))
        assert (back["score"] == -7).all(), "outputs were written"
    finally:
        eng.close()
        d.close()
