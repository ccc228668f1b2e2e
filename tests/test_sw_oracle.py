"""Pins of the full inter-task SW oracle (oracle/sw_oracle.c), CPU only.

The oracle restates PairWiseSW_{8,16,32}::getScores of the reference's
AVX-512BW build.  It is checked against:
  * the reference itself, compiled from its sources into oracle/_ref by
    `make -C oracle ref` (skipped where that library is absent or the CPU
    lacks AVX-512BW);
  * the reference's scalar routine smithWatermanOnePair, restated
    independently, wherever the group padding cannot change a score;
  * a brute-force local alignment over explicit gap lengths on tiny inputs;
  * closed forms (identical sequences) and the int8 group-bound wrap of the
    8-bit class;
and each planted mistake is shown to fail at least one of these checks.
"""
import itertools

import numpy as np
import pytest

import oracle as O

PARAMS = [(1, -1, 1, 1), (2, -4, -6, -1), (1, -3, -5, -2), (100, -50, 30, 20), (3, 5, -2, 4),
          (5, -4, -10, -1)]


def _rand_batch(rng, n, bits, lo, hi, alpha=b"ACGT", related=True):
    A = np.frombuffer(alpha, np.uint8)
    s1, s2 = [], []
    for _ in range(n):
        a = A[rng.integers(0, len(A), rng.integers(lo, hi + 1))]
        L = int(rng.integers(lo, hi + 1))
        if related and len(a):
            b = np.resize(a, L).copy()
            m = rng.random(L) < 0.15
            b[m] = A[rng.integers(0, len(A), int(m.sum()))]
        else:
            b = A[rng.integers(0, len(A), L)]
        s1.append(a)
        s2.append(b)
    return O.sw_make_batch(s1, s2, bits)


ref = pytest.mark.skipif(not O.ref_sw_available(), reason="oracle/_ref/libpwsw_ref.so not built")


@ref
@pytest.mark.parametrize("bits,lo,hi", [(8, 0, 124), (8, 60, 127), (16, 0, 300), (32, 0, 300),
                                        (16, 900, 1020)])
@pytest.mark.parametrize("params", PARAMS)
def test_oracle_equals_reference_build(bits, lo, hi, params):
    rng = np.random.default_rng(bits * 1000 + lo + hash(params) % 997)
    n = 150 if hi > 500 else 301
    pairs, buf = _rand_batch(rng, n, bits, lo, hi)
    want = O.ref_sw_scores(bits, pairs.copy(), buf, params)
    if want is None:
        pytest.skip("CPU lacks AVX-512BW")
    got = O.sw_scores(bits, pairs.copy(), buf, params)
    assert np.array_equal(got["score"], want["score"])


@ref
def test_oracle_equals_reference_arbitrary_bytes():
    # every byte value, padding codes 'B' and 'D' inside the sequences
    rng = np.random.default_rng(5)
    alpha = bytes(range(256))
    for bits in (8, 16, 32):
        pairs, buf = _rand_batch(rng, 200, bits, 0, 100, alpha=alpha, related=False)
        for params in PARAMS[:3]:
            want = O.ref_sw_scores(bits, pairs.copy(), buf, params)
            if want is None:
                pytest.skip("CPU lacks AVX-512BW")
            assert np.array_equal(O.sw_scores(bits, pairs.copy(), buf, params)["score"],
                                  want["score"])


@pytest.mark.parametrize("bits", [8, 16, 32])
def test_scalar_agrees_where_padding_cannot_matter(bits):
    # penalties negative and mismatch below match: a padded cell never beats
    # the real cells, so the group score equals each pair's own DP
    rng = np.random.default_rng(11 + bits)
    hi = 120 if bits == 8 else 400
    pairs, buf = _rand_batch(rng, 200, bits, 0, hi)
    for params in [(2, -4, -6, -1), (1, -3, -5, -2), (5, -4, -10, -1)]:
        if bits == 8 and params[0] == 5:
            continue  # scores would pass the int8 range
        got = O.sw_scores(bits, pairs.copy(), buf, params)
        want = O.sw_scalar(pairs.copy(), buf, params, slot=O.SW_SLOT[bits])
        # the 8-bit class scores whole groups of len >= 125 as 0
        keep = np.ones(len(pairs), bool)
        if bits == 8:
            g = np.arange(len(pairs)) // 64
            for k in np.unique(g):
                sel = g == k
                if max(pairs["len1"][sel].max(), pairs["len2"][sel].max()) > 124:
                    keep[sel] = False
            # and the int8 lanes wrap once a score passes 127
            keep &= want["score"] <= 120
        assert np.array_equal(got["score"][keep], want["score"][keep])


def test_scalar_agrees_on_unpadded_groups_any_weights():
    # lengths equal and a multiple of 4 across the group: no padding at all
    rng = np.random.default_rng(3)
    for bits, L in ((16, 64), (32, 96)):
        g = O.SW_GROUP[bits]
        s1 = [np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, L)] for _ in range(g * 3)]
        s2 = [np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, L)] for _ in range(g * 3)]
        pairs, buf = O.sw_make_batch(s1, s2, bits)
        for params in [(1, -1, 1, 1), (3, 5, -2, 4), (2, -4, -6, -1)]:
            got = O.sw_scores(bits, pairs.copy(), buf, params)
            want = O.sw_scalar(pairs.copy(), buf, params, slot=O.SW_SLOT[bits])
            assert np.array_equal(got["score"], want["score"]), params


def _brute(a, b, match, mismatch, gopen, gext):
    """max over local alignments by explicit gap lengths:
    H(i,j) = max(0, H(i-1,j-1)+s, max_k H(i-k,j)+open+(k-1)ext,
                 max_k H(i,j-k)+open+(k-1)ext), with H = 0 on the borders."""
    n, m = len(a), len(b)
    H = np.zeros((n + 1, m + 1), np.int64)
    best = 0
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            v = max(0, H[i - 1, j - 1] + (match if a[i - 1] == b[j - 1] else mismatch))
            for k in range(1, i + 1):
                v = max(v, H[i - k, j] + gopen + (k - 1) * gext)
            for k in range(1, j + 1):
                v = max(v, H[i, j - k] + gopen + (k - 1) * gext)
            H[i, j] = v
            best = max(best, v)
    return best


def test_brute_force_tiny():
    rng = np.random.default_rng(9)
    seqs1, seqs2 = [], []
    for _ in range(64):
        seqs1.append(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, rng.integers(1, 7))])
        seqs2.append(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, rng.integers(1, 7))])
    for params in [(2, -3, -4, -1), (1, -1, -2, -2), (3, -2, -3, 0)]:
        pairs, buf = O.sw_make_batch(seqs1, seqs2, 32)
        # one pair per 16-lane group at a time, so padding cannot matter
        got = O.sw_scores(32, pairs.copy(), buf, params)["score"]
        want = [_brute(seqs1[k], seqs2[k], *params) for k in range(len(seqs1))]
        assert list(got) == want, params


def test_identical_sequences_closed_form():
    rng = np.random.default_rng(2)
    for bits, L in ((8, 100), (16, 700), (32, 1000)):
        s = [np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, L)] for _ in range(3)]
        pairs, buf = O.sw_make_batch(s, s, bits)
        got = O.sw_scores(bits, pairs, buf, (1, -3, -5, -2))
        assert list(got["score"]) == [L] * 3


def test_int8_group_bound_wraps():
    # _8.cpp:293,308: int8 maxLen rounded up to 128 wraps to -128, loops empty
    rng = np.random.default_rng(4)
    A = np.frombuffer(b"ACGT", np.uint8)
    for L, zero in ((124, False), (125, True), (127, True)):
        s1 = [A[rng.integers(0, 4, 40)] for _ in range(64)]
        s1[7] = A[rng.integers(0, 4, L)]
        pairs, buf = O.sw_make_batch(s1, [x.copy() for x in s1], 8)
        got = O.sw_scores(8, pairs, buf, (1, -3, -5, -2))["score"]
        assert (got == 0).all() == zero, L
        if not zero:
            assert got[0] == 40


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes_are_caught(mutation):
    """Every planted mistake fails at least one of the pins above."""
    rng = np.random.default_rng(21)
    caught = []
    try:
        O.sw_set_mutation(mutation)
        pairs, buf = _rand_batch(rng, 96, 16, 0, 60)
        if O.ref_sw_available():
            for params in PARAMS:
                want = O.ref_sw_scores(16, pairs.copy(), buf, params)
                if want is not None:
                    got = O.sw_scores(16, pairs.copy(), buf, params)
                    caught.append(not np.array_equal(got["score"], want["score"]))
        for params in [(2, -4, -6, -1), (1, -3, -5, -2)]:
            got = O.sw_scores(16, pairs.copy(), buf, params)
            want = O.sw_scalar(pairs.copy(), buf, params, slot=1024)
            caught.append(not np.array_equal(got["score"], want["score"]))
        s = [np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 30)] for _ in range(4)]
        p2, b2 = O.sw_make_batch(s, s, 16)
        caught.append(list(O.sw_scores(16, p2, b2, (1, -3, -5, -2))["score"]) != [30] * 4)
    finally:
        O.sw_set_mutation(0)
    assert any(caught)
