"""GPU parity on edge cases: every output field of every pair against the
oracle, on seeded random batches built to reach the rarely taken branches
(band narrowing with small w, empty and one-base sequences, ambiguous and
arbitrary byte codes, scores that wrap int16/int8, asymmetric gaps, scores
outside the fast path's range, and multi-call sequences on one handle)."""
import zlib

import numpy as np
import pytest

import oracle as O
import tal_bsw as T
from helpers import assert_same_outputs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _batch(rng, n, len_lo, len_hi, h0_lo, h0_hi, alphabet=4, related=0.7, zero_len=0.0):
    refs, qs = [], []
    for _ in range(n):
        l1 = int(rng.integers(len_lo, len_hi + 1))
        l2 = int(rng.integers(len_lo, len_hi + 1))
        if rng.random() < zero_len:
            l1 = 0 if rng.random() < 0.5 else l1
            l2 = 0 if rng.random() < 0.5 else l2
        r = rng.integers(0, alphabet, l1).astype(np.uint8)
        if rng.random() < related and l1 > 0:
            q = np.resize(r, l2).copy()
            m = rng.random(l2) < 0.05
            q[m] = rng.integers(0, alphabet, int(m.sum()))
        else:
            q = rng.integers(0, alphabet, l2).astype(np.uint8)
        refs.append(r)
        qs.append(q)
    h0 = [int(x) for x in rng.integers(h0_lo, h0_hi + 1, n)]
    return O.make_pairs(refs, qs, h0)


CASES16 = [
    # name, params, n, len range, h0 range, alphabet, w
    ("small-w", {}, 997, (1, 200), (0, 80), 4, 7),
    ("w30-mixed", {}, 1021, (0, 260), (0, 150), 5, 30),
    ("wide", {}, 640, (1, 120), (0, 60), 4, 500),
    ("zero-lengths", {}, 333, (0, 40), (0, 40), 4, 100),
    ("N-and-bytes", {}, 520, (1, 150), (0, 90), 7, 100),
    ("high-bytes", {}, 300, (1, 90), (0, 90), 256, 50),
    ("h0-wrap", {}, 256, (1, 120), (32600, 40000), 4, 100),
    ("h0-negative", {}, 256, (1, 120), (-300, 20), 4, 100),
    ("asym-gaps", dict(o_del=5, e_del=2, o_ins=7, e_ins=1), 800, (1, 200), (0, 100), 5, 60),
    ("zdrop-5", dict(zdrop=5), 640, (1, 200), (0, 100), 4, 100),
    ("big-scores", dict(w_match=9, w_mismatch=11), 400, (1, 150), (0, 100), 4, 100),
    ("w0", {}, 200, (0, 60), (0, 60), 4, 0),
]


@pytest.mark.parametrize("name,params,n,lens,h0s,alpha,w", CASES16, ids=[c[0] for c in CASES16])
def test_edges_16bit(name, params, n, lens, h0s, alpha, w):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    pairs, ref, qer = _batch(rng, n, *lens, *h0s, alphabet=alpha, zero_len=0.1)
    _say(f"[gpu] edges16 {name} n={n} w={w}")
    want = O.Oracle(**params).scores16(pairs.copy(), ref, qer, w)
    eng = T.BandedPairWiseSW(**params)
    got = eng.get_scores16(pairs.copy(), ref, qer, w)
    eng.close()
    assert_same_outputs(got, want, name)


CASES8 = [
    ("small-w", {}, 1001, (1, 127), (0, 60), 4, 9),
    ("w-over-127", {}, 700, (1, 127), (0, 60), 4, 200),
    ("zero-lengths", {}, 300, (0, 30), (0, 30), 4, 100),
    ("bytes", {}, 400, (1, 127), (0, 120), 256, 100),
    ("h0-wrap", {}, 300, (1, 127), (100, 300), 4, 100),
    ("asym", dict(o_del=5, e_del=2, o_ins=7, e_ins=1, zdrop=40), 640, (1, 127), (0, 100), 5, 40),
]


@pytest.mark.parametrize("name,params,n,lens,h0s,alpha,w", CASES8, ids=[c[0] for c in CASES8])
def test_edges_8bit(name, params, n, lens, h0s, alpha, w):
    rng = np.random.default_rng(zlib.crc32((name + "8").encode()))
    pairs, ref, qer = _batch(rng, n, *lens, *h0s, alphabet=alpha, zero_len=0.1)
    _say(f"[gpu] edges8 {name} n={n} w={w}")
    want = O.Oracle(**params).scores8(pairs.copy(), ref, qer, w)
    eng = T.BandedPairWiseSW(**params)
    got = eng.get_scores8(pairs.copy(), ref, qer, w)
    eng.close()
    assert_same_outputs(got, want, name)


def test_call_sequence_keeps_scratch():
    """Several calls on one handle match the same calls on one oracle object
    (the 16-bit F scratch carries from call to call)."""
    rng = np.random.default_rng(7)
    o = O.Oracle()
    eng = T.BandedPairWiseSW()
    for k, (lo, hi) in enumerate([(100, 260), (20, 90), (150, 151), (30, 200), (1, 255)]):
        pairs, ref, qer = _batch(rng, 301 + 37 * k, lo, hi, 0, 120, alphabet=4)
        _say(f"[gpu] call {k}")
        want = o.scores16(pairs.copy(), ref, qer, 100)
        got = eng.get_scores16(pairs.copy(), ref, qer, 100)
        assert_same_outputs(got, want, f"call {k}")
    eng.close()


def test_small_n():
    rng = np.random.default_rng(11)
    eng = T.BandedPairWiseSW()
    for n in (1, 2, 31, 33, 63, 65):
        pairs, ref, qer = _batch(rng, n, 1, 200, 0, 100)
        _say(f"[gpu] n={n}")
        assert_same_outputs(eng.get_scores16(pairs.copy(), ref, qer, 100),
                            O.Oracle().scores16(pairs.copy(), ref, qer, 100), f"16 n={n}")
        p8, r8, q8 = _batch(rng, n, 1, 127, 0, 60)
        assert_same_outputs(eng.get_scores8(p8.copy(), r8, q8, 100),
                            O.Oracle().scores8(p8.copy(), r8, q8, 100), f"8 n={n}")
    empty = np.zeros(0, T.SEQPAIR_DTYPE)
    eng.get_scores16(empty, np.zeros(1, np.uint8), np.zeros(1, np.uint8), 100)
    eng.close()


def test_errors_are_reported():
    eng = T.BandedPairWiseSW()
    p8, r8, q8 = O.make_pairs([np.zeros(200, np.uint8)], [np.zeros(20, np.uint8)], [5])
    with pytest.raises(T.TalError) as e:
        eng.get_scores8(p8, r8, q8, 100)
    assert e.value.code == T.TAL_ERANGE
    bad = p8.copy()
    bad["idr"] = 10**6
    with pytest.raises(T.TalError) as e:
        eng.get_scores16(bad, r8, q8, 100)
    assert e.value.code == T.TAL_EINVAL
    # the handle still works after an error
    ok = eng.get_scores16(p8.copy(), r8, q8, 100)
    assert_same_outputs(ok, O.Oracle().scores16(p8.copy(), r8, q8, 100), "after error")
    eng.close()


def test_device_api_in_place():
    import ctypes
    rng = np.random.default_rng(5)
    pairs, ref, qer = _batch(rng, 777, 1, 250, 0, 120)
    want = O.Oracle().scores16(pairs.copy(), ref, qer, 100)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    bufs = []
    for a in (pairs, ref, qer):
        d = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(d), a.nbytes) == 0
        assert hip.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0
        bufs.append(d)
    eng = T.BandedPairWiseSW()
    eng.get_scores_device(16, bufs[0].value, bufs[1].value, bufs[2].value, len(pairs), 100)
    got = pairs.copy()
    assert hip.hipMemcpy(got.ctypes.data, bufs[0], got.nbytes, 2) == 0
    eng.close()
    for d in bufs:
        hip.hipFree(d)
    assert_same_outputs(got, want, "device API")
