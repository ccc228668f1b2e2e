"""Allocation failure on the GPU, forced by fault injection: with
TAL_BSW_FAIL_ALLOC=k set when a handle is created, the handle's k-th scratch
request fails as if hipMalloc had run out of memory.  For every k that falls
inside one dispatcher call, the call returns TAL_ENOMEM with a message naming
the allocation and leaves the caller's pairs untouched; the same handle then
scores the batch exactly like the oracle on a fresh object (the failed call
left no state behind)."""
import os

import numpy as np
import pytest

import oracle as O
import synth as S
import tal_bsw as T
from helpers import assert_same_outputs

pytestmark = pytest.mark.gpu


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


def _handle_failing_at(k):
    os.environ["TAL_BSW_FAIL_ALLOC"] = str(k)
    try:
        return T.BandedPairWiseSW()
    finally:
        del os.environ["TAL_BSW_FAIL_ALLOC"]


def test_every_scratch_request_can_fail_cleanly():
    pairs, ref, qer, _ = S.generate(2, 3000)
    assert T.is_8bit(pairs).any() and (~T.is_8bit(pairs)).any()
    want = _oracle_auto(pairs, ref, qer, 100)
    failed = 0
    for k in range(1, 200):
        print(f"[gpu] allocation failure injected at scratch request {k}", flush=True)
        eng = _handle_failing_at(k)
        try:
            work = pairs.copy()
            try:
                eng.get_scores_auto(work, ref, qer, 100)
                done = True
            except T.TalError as e:
                assert e.code == T.TAL_ENOMEM, str(e)
                assert "alloc" in str(e), str(e)
                assert (work == pairs).all(), "a failed call wrote outputs"
                failed += 1
                done = False
            if done:  # k lies beyond the call's requests
                assert_same_outputs(work, want, f"request {k} beyond the call")
                break
            again = pairs.copy()
            eng.get_scores_auto(again, ref, qer, 100)
            assert_same_outputs(again, want, f"after a failure at request {k}")
        finally:
            eng.close()
    else:
        pytest.fail("the call never got past the injected failures")
    # the host staging (3), the dispatcher (6) and both paths' buffers
    assert failed >= 15, f"only {failed} scratch requests were exercised"
