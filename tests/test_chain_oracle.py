"""Pins of the DP chaining oracle (oracle/chain_oracle.c), CPU only.

The oracle restates dp_chain::mm_dp_vectorized (the AVX-512 build of
src/chaining/parallel_chaining_32_bit.h).  It is checked against the
reference itself, compiled from its header into oracle/_ref by
`make -C oracle ref` (skipped where absent or the CPU lacks AVX-512BW/CD),
against closed forms, and each planted mistake is shown to fail a check."""
import numpy as np
import pytest

import oracle as O
import synth as S

ref = pytest.mark.skipif(not O.ref_chain_available(), reason="oracle/_ref/libchain_ref.so not built")


def _ref_or_skip(*args, **kw):
    out = O.ref_chain_dp_sets(*args, **kw)
    if out is None:
        pytest.skip("CPU lacks AVX-512BW/CD")
    return out


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@ref
@pytest.mark.parametrize("seed,nsets,lo,hi", [(1, 80, 1, 40), (2, 40, 1, 3000), (3, 12, 3000, 6000)])
def test_oracle_equals_reference_build(seed, nsets, lo, hi):
    a, off = S.chain_generate(nsets, lo, hi, seed=seed)
    assert _same(O.chain_dp_sets(a, off, nthreads=8), _ref_or_skip(a, off))


@ref
@pytest.mark.parametrize("kw", [dict(bw=20, max_iter=50, dr=300, dq=200),
                                dict(bw=5000, max_iter=8, dr=5000, dq=5000),
                                dict(bw=0, max_iter=5000, dr=100000, dq=100000),
                                dict(bw=500, max_iter=3, dr=50, dq=5000),
                                dict(bw=500, max_iter=700, dr=-1, dq=5000),
                                dict(bw=500, max_iter=5000, dr=5000, dq=-3)])
def test_oracle_equals_reference_other_parameters(kw):
    a, off = S.chain_generate(30, 1, 2500, seed=7)
    assert _same(O.chain_dp_sets(a, off, nthreads=8, **kw), _ref_or_skip(a, off, **kw))


@ref
def test_oracle_equals_reference_ties_and_duplicates():
    # repeated anchors and equal diagonals give equal scores: p must be the
    # highest tying j, on both the scalar (<= 6 candidates) and vector paths
    rng = np.random.default_rng(5)
    sets = []
    for m in (2, 5, 7, 8, 40, 300):
        r = np.sort(rng.integers(0, 2000, m)).astype(np.uint64)
        r[1::3] = r[0::3][:len(r[1::3])]
        q = (r.astype(np.int64) // 2 + rng.integers(0, 3, m)).astype(np.int32)
        sets.append((r, q, np.full(m, 15, np.int32)))
    n = sum(len(s[0]) for s in sets)
    a = np.zeros(n, O.ANCHOR_DTYPE)
    off = np.zeros(len(sets) + 1, np.int64)
    k = 0
    for i, (r, q, ln) in enumerate(sets):
        a["r"][k:k + len(r)], a["q"][k:k + len(r)], a["l"][k:k + len(r)] = r, q, ln
        k += len(r)
        off[i + 1] = k
    a["r"] = np.maximum.accumulate(a["r"])
    assert _same(O.chain_dp_sets(a, off), _ref_or_skip(a, off))


def _rounding_set(m=200, seed=4):
    """All seed lengths 100, so avg_qspan = 100: the vector path's float gap
    (dd * (float)(0.01 * 100) = dd * 1.0f) and the scalar path's double gap
    ((dd * .01) * 100, truncated) differ for some dd (29, 58, 116, ...: 29
    and 28 for dd = 29) -- a check that each anchor takes the reference's
    path.  The q drifts make such dd between nearby anchors."""
    rng = np.random.default_rng(seed)
    a = np.zeros(m, O.ANCHOR_DTYPE)
    a["r"] = 5000 + 40 * np.arange(m)
    a["q"] = 40 * np.arange(m) + rng.choice([0, 29, 58, 116, 5], m)
    a["l"] = 100
    return a, np.array([0, m], np.int64)


def _overflow_set(m=2000, seed=6):
    """Half the seeds have lengths near 2^31 (and q far away, so that they
    chain with nothing), so avg_qspan is about 1e9 and dd * 0.01 * avg_qspan
    passes 2^31 from dd of about 200: those gap costs leave int32 range, where
    the reference's x86 conversions give INT32_MIN (the GPU's own conversion
    saturates, so the engine handles this case by hand).  Such a cost only
    shows in a score when f[j] + oc is negative, oc = ddr < 0, where the score
    wraps to a large value; so some neighbouring r of the short seeds are
    swapped (r out of order is within what the reference accepts).  Short r
    steps give both the scalar (<= 6 candidates) and the vector path."""
    rng = np.random.default_rng(seed)
    a = np.zeros(m, O.ANCHOR_DTYPE)
    steps = rng.choice([1, 2, 30, 400, 4000], m)
    r = 10_000 + np.cumsum(steps)
    q = (r - 10_000 + rng.integers(-400, 401, m)).clip(0)
    for k in rng.choice(m - 1, m // 5, replace=False):
        r[k], r[k + 1] = r[k + 1], r[k]
    big = rng.random(m) < 0.5
    ln = rng.integers(15, 29, m)
    ln[big] = (1 << 31) - 1 - rng.integers(0, 1000, int(big.sum()))
    q[big] = 50_000_000 + rng.integers(0, 1000, int(big.sum()))
    a["r"], a["q"], a["l"] = r, q, ln
    return a, np.array([0, m], np.int64)


@ref
def test_oracle_equals_reference_gap_overflow():
    a, off = _overflow_set()
    assert _same(O.chain_dp_sets(a, off), _ref_or_skip(a, off))


def test_overflow_set_reaches_the_limit():
    # some candidate of the overflow set has a gap cost at or past 2^31 on
    # each path, so the set pins the out-of-range conversions
    a, _ = _overflow_set()
    avg = np.float32(np.float32(a["l"].astype(np.uint64).sum()) / np.float32(len(a)))
    dd = np.arange(0, 501)
    f32 = dd.astype(np.float32) * np.float32(0.01 * np.float64(avg))
    assert (f32 >= 2.0 ** 31).any() and (f32 < 2.0 ** 31).any()


@ref
def test_oracle_equals_reference_gap_rounding():
    a, off = _rounding_set()
    assert _same(O.chain_dp_sets(a, off), _ref_or_skip(a, off))


def test_gap_paths_differ_on_the_rounding_set():
    # the two gap formulas give different integers here, so the set above
    # pins which one each anchor uses
    dd = np.arange(1, 40)
    f32 = np.floor(dd.astype(np.float32) * np.float32(0.01 * np.float32(100.0))).astype(int)
    f64 = ((dd.astype(np.float64) * .01) * np.float64(np.float32(100.0))).astype(int)
    assert (f32 != f64).any()


def test_closed_forms():
    # one anchor: f = l, p = -1, v = l
    a = np.zeros(1, O.ANCHOR_DTYPE)
    a["r"], a["q"], a["l"] = 100, 50, 17
    f, p, v = O.chain_dp_sets(a, [0, 1])
    assert (int(f[0]), int(p[0]), int(v[0])) == (17, -1, 17)
    # a perfect diagonal of anchors 10 apart with l = 15: no gap (dd = 0),
    # overlap min(10, 10, 15) = 10, so f[i] = 15 + 10 i and p[i] = i - 1
    m = 50
    a = np.zeros(m, O.ANCHOR_DTYPE)
    a["r"] = 1000 + 10 * np.arange(m)
    a["q"] = 10 * np.arange(m)
    a["l"] = 15
    f, p, v = O.chain_dp_sets(a, [0, m])
    assert list(f) == [15 + 10 * i for i in range(m)]
    assert list(p) == [-1] + list(range(m - 1))
    assert list(v) == list(f)


@pytest.mark.parametrize("mutation", [1, 2, 3, 4])
def test_planted_mistakes_are_caught(mutation):
    a, off = S.chain_generate(20, 1, 1500, seed=9)
    caught = []
    try:
        O.chain_set_mutation(mutation)
        got = O.chain_dp_sets(a, off)
        if O.ref_chain_available():
            want = O.ref_chain_dp_sets(a, off)
            if want is not None:
                caught.append(not _same(got, want))
        m = 50
        d = np.zeros(m, O.ANCHOR_DTYPE)
        d["r"] = 1000 + 10 * np.arange(m)
        d["q"] = 10 * np.arange(m)
        d["l"] = 15
        f, p, _ = O.chain_dp_sets(d, [0, m])
        caught.append(list(f) != [15 + 10 * i for i in range(m)] or
                      list(p) != [-1] + list(range(m - 1)))
        if O.ref_chain_available():
            for ra, roff in (_rounding_set(), _overflow_set()):
                want = O.ref_chain_dp_sets(ra, roff)
                if want is not None:
                    caught.append(not _same(O.chain_dp_sets(ra, roff), want))
    finally:
        O.chain_set_mutation(0)
    assert any(caught)
