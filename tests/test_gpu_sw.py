"""GPU parity of the full inter-task SW engine (libtal_sw.so) against the
oracle, bit-exact, through the C ABI."""
import numpy as np
import pytest

import oracle as O
import synth as S
import tal_sw as T
from helpers import DeviceArrays

pytestmark = pytest.mark.gpu

PARAMS = [(1, -1, 1, 1), (2, -4, -6, -1), (1, -3, -5, -2), (100, -50, 30, 20), (3, 5, -2, 4)]


def _say(msg):
    print(msg, flush=True)


def _rand_batch(rng, n, bits, lo, hi, alpha=b"ACGT"):
    A = np.frombuffer(alpha, np.uint8)
    s1, s2 = [], []
    for _ in range(n):
        a = A[rng.integers(0, len(A), rng.integers(lo, hi + 1))]
        L = int(rng.integers(lo, hi + 1))
        b = np.resize(a, L).copy() if len(a) else A[rng.integers(0, len(A), L)]
        m = rng.random(L) < 0.15
        b[m] = A[rng.integers(0, len(A), int(m.sum()))]
        s1.append(a)
        s2.append(b)
    return T.make_batch(s1, s2, bits)


def _check(eng, bits, pairs, buf, label):
    want = O.sw_scores(bits, pairs.copy(), buf, eng.params, nthreads=8)["score"]
    got = eng.get_scores(bits, pairs.copy(), buf)["score"]
    bad = np.where(got != want)[0]
    assert len(bad) == 0, f"{label}: {len(bad)} of {len(pairs)} differ; first {bad[:5]}: " \
                          f"{got[bad[:5]]} vs {want[bad[:5]]}"


@pytest.mark.parametrize("params", PARAMS)
@pytest.mark.parametrize("bits,lo,hi,n", [(8, 0, 124, 1001), (8, 60, 127, 517),
                                          (16, 0, 300, 1003), (32, 0, 300, 1005)])
def test_random_batches(bits, lo, hi, n, params):
    _say(f"[gpu] sw bits={bits} len=[{lo},{hi}] n={n} params={params}")
    eng = T.PairWiseSW(*params)
    rng = np.random.default_rng(bits * 7 + lo + n)
    pairs, buf = _rand_batch(rng, n, bits, lo, hi)
    _check(eng, bits, pairs, buf, f"bits={bits} {params}")
    eng.close()


@pytest.mark.parametrize("bits", [16, 32])
def test_long_sequences(bits):
    _say(f"[gpu] sw long bits={bits}")
    rng = np.random.default_rng(bits)
    for params in [(1, -1, 1, 1), (100, -50, 30, 20), (2, -4, -6, -1)]:
        eng = T.PairWiseSW(*params)
        pairs, buf = _rand_batch(rng, 70, bits, 700, 1024)
        _check(eng, bits, pairs, buf, f"long {params}")
        eng.close()


def test_arbitrary_bytes_and_padding_codes():
    _say("[gpu] sw arbitrary bytes")
    rng = np.random.default_rng(77)
    for bits in (8, 16, 32):
        eng = T.PairWiseSW(2, -4, -6, -1)
        pairs, buf = _rand_batch(rng, 300, bits, 0, 100, alpha=bytes(range(256)))
        _check(eng, bits, pairs, buf, f"bytes bits={bits}")
        eng.close()


def test_mixed_bounds_in_one_wave():
    """Groups of very different bounds next to each other, so that the four
    units of a wave differ in rows and columns (the masked sweep)."""
    _say("[gpu] sw mixed bounds")
    rng = np.random.default_rng(5)
    A = np.frombuffer(b"ACGT", np.uint8)
    for bits in (8, 16, 32):
        g = T.GROUP[bits]
        hi = 124 if bits == 8 else 600
        s1, s2 = [], []
        for grp in range(24):
            L1 = int(rng.integers(1, hi + 1))
            L2 = int(rng.integers(1, hi + 1))
            for _ in range(g):
                s1.append(A[rng.integers(0, 4, rng.integers(0, L1 + 1))])
                s2.append(A[rng.integers(0, 4, rng.integers(0, L2 + 1))])
        pairs, buf = T.make_batch(s1, s2, bits)
        for params in [(1, -1, 1, 1), (2, -4, -6, -1)]:
            eng = T.PairWiseSW(*params)
            _check(eng, bits, pairs, buf, f"mixed bits={bits} {params}")
            eng.close()


def test_workload_samples():
    for name in ("sw16", "sw8", "sw32"):
        bits, _, lo, hi = S.SW_WORKLOADS[name]
        pairs, buf = S.sw_generate(bits, 3000 if bits != 8 else 20000, lo, hi, seed=3)
        _say(f"[gpu] sw workload {name} sample n={len(pairs)}")
        for params in [(1, -1, 1, 1), (2, -4, -6, -1)]:
            eng = T.PairWiseSW(*params)
            _check(eng, bits, pairs, buf, f"{name} {params}")
            eng.close()


def test_edges():
    _say("[gpu] sw edges")
    eng = T.PairWiseSW(2, -4, -6, -1)
    A = np.frombuffer(b"ACGT", np.uint8)
    for bits in (8, 16, 32):
        # empty batch
        eng.get_scores(bits, np.zeros(0, T.PAIR_DTYPE), np.zeros(4, np.uint8))
        # zero lengths, single pair, a group of only empty sequences
        cases = [([A[:0]], [A[:0]]), ([A[:3]], [A[:3]]), ([A[:0]] * 40, [A[:2]] * 40),
                 ([A[:4]] * 5, [A[:0]] * 5)]
        for s1, s2 in cases:
            pairs, buf = T.make_batch(s1, s2, bits)
            _check(eng, bits, pairs, buf, f"edge bits={bits}")
        # the longest sequences the class takes
        L = T.MAX_LEN[bits]
        rng = np.random.default_rng(bits)
        s = [A[rng.integers(0, 4, L)] for _ in range(3)]
        pairs, buf = T.make_batch(s, [x[::-1].copy() for x in s], bits)
        _check(eng, bits, pairs, buf, f"max len bits={bits}")
    # ids out of order and repeated: sequences are located by id
    s1 = [A[np.arange(k % 4, k % 4 + 20) % 4] for k in range(8)]
    pairs, buf = T.make_batch(s1, [x[::-1].copy() for x in s1], 16)
    pairs["id"] = [3, 1, 1, 7, 0, 2, 6, 5]
    pairs["len1"] = 20
    pairs["len2"] = 20
    _check(eng, 16, pairs, buf, "ids")
    eng.close()


def test_errors_leave_handle_usable():
    _say("[gpu] sw errors")
    eng = T.PairWiseSW(1, -3, -5, -2)
    A = np.frombuffer(b"ACGT", np.uint8)
    pairs, buf = T.make_batch([A[:10]] * 4, [A[:10]] * 4, 16)
    bad = pairs.copy()
    bad["len1"][2] = 1025
    with pytest.raises(T.TalError) as e:
        eng.get_scores(16, bad, buf)
    assert e.value.code == T.TAL_ERANGE
    bad = pairs.copy()
    bad["id"][1] = 4            # sequences past the buffer
    with pytest.raises(T.TalError) as e:
        eng.get_scores(16, bad, buf)
    assert e.value.code == T.TAL_EINVAL
    p8, b8 = T.make_batch([A[:10]] * 4, [A[:10]] * 4, 8)
    p8["len2"][0] = 128
    with pytest.raises(T.TalError) as e:
        eng.get_scores(8, p8, b8)
    assert e.value.code == T.TAL_ERANGE
    with pytest.raises(T.TalError) as e:
        eng.get_scores(12, pairs.copy(), buf)
    assert e.value.code == T.TAL_EINVAL
    _check(eng, 16, pairs, buf, "after errors")
    eng.close()


def test_device_path_and_stats():
    _say("[gpu] sw device path")
    rng = np.random.default_rng(8)
    pairs, buf = _rand_batch(rng, 2000, 16, 50, 400)
    eng = T.PairWiseSW(2, -4, -6, -1)
    want = O.sw_scores(16, pairs.copy(), buf, eng.params, nthreads=8)["score"]
    d = DeviceArrays(pairs, buf)
    try:
        eng.get_scores_device(16, d.ptr(0), d.ptr(1), buf.nbytes, len(pairs))
        out = d.download(0, pairs.copy())
    finally:
        d.close()
    assert np.array_equal(out["score"], want)
    st = eng.stats()
    assert st["ms_kernel"] > 0 and st["units"] == (len(pairs) + 31) // 32
    assert st["max_rows"] <= 400 and st["max_cols"] <= 400
    eng.close()


def test_regrouping_engages_only_when_exact():
    """Penalty weights on DNA: pairs are regrouped by length (stats say so)
    and the scores still equal the oracle's group semantics; each condition
    that could let padding matter keeps the reference's groups."""
    _say("[gpu] sw regrouping conditions")
    rng = np.random.default_rng(31)
    A = np.frombuffer(b"ACGT", np.uint8)
    pairs, buf = _rand_batch(rng, 700, 16, 0, 400)
    eng = T.PairWiseSW(2, -4, -6, -1)
    _check(eng, 16, pairs, buf, "regroup")
    assert eng.stats()["regrouped"] == 1
    # positive gap weights: padding matters
    eng2 = T.PairWiseSW(1, -1, 1, 1)
    _check(eng2, 16, pairs, buf, "positive gaps")
    assert eng2.stats()["regrouped"] == 0
    # a seq2 holding the seq1 padding code 'B'
    p2, b2 = pairs.copy(), buf.copy()
    k = int(np.argmax(p2["len2"]))
    b2[(2 * int(p2["id"][k]) + 1) * 1024 + 3] = ord("B")
    _check(eng, 16, p2, b2, "B in seq2")
    assert eng.stats()["regrouped"] == 0
    # match * min(len1, len2) past the int16 lane
    eng3 = T.PairWiseSW(40, -4, -6, -1)
    s = [A[rng.integers(0, 4, 900)] for _ in range(40)]
    p3, b3 = T.make_batch(s, [x.copy() for x in s], 16)
    _check(eng3, 16, p3, b3, "int16 bound")
    assert eng3.stats()["regrouped"] == 0
    for e in (eng, eng2, eng3):
        e.close()


def test_regrouping_keeps_zero_bound_groups_8bit():
    _say("[gpu] sw regrouping with 8-bit zero-bound groups")
    rng = np.random.default_rng(32)
    A = np.frombuffer(b"ACGT", np.uint8)
    s1, s2 = [], []
    for grp in range(20):
        longest = 126 if grp % 3 == 0 else 100
        for k in range(64):
            L = longest if k == 5 else int(rng.integers(0, 100))
            a = A[rng.integers(0, 4, L)]
            s1.append(a)
            s2.append(np.resize(a, int(rng.integers(0, 100))).copy())
    pairs, buf = T.make_batch(s1, s2, 8)
    eng = T.PairWiseSW(1, -3, -5, -2)
    _check(eng, 8, pairs, buf, "8-bit zero groups")
    assert eng.stats()["regrouped"] == 1
    eng.close()


@pytest.mark.parametrize("split", [1, 2, 4])
def test_every_unit_split(split, monkeypatch):
    """The DP kernel's unit split (one unit per quarter, or split over two or
    four quarters) gives the same scores, in both the reference's groups and
    regrouped order, for ragged and uniform bounds."""
    _say(f"[gpu] sw split={split}")
    monkeypatch.setenv("TAL_SW_SPLIT", str(split))
    rng = np.random.default_rng(40 + split)
    for bits, lo, hi, n in [(16, 0, 300, 901), (8, 0, 124, 1200), (32, 0, 300, 450),
                            (16, 700, 1024, 90)]:
        pairs, buf = _rand_batch(rng, n, bits, lo, hi)
        for params in [(1, -1, 1, 1), (2, -4, -6, -1)]:
            eng = T.PairWiseSW(*params)
            _check(eng, bits, pairs, buf, f"split={split} bits={bits} {params}")
            assert eng.stats()["split"] == split
            eng.close()
