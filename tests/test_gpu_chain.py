"""GPU parity of the DP chaining engine (libtal_chain.so) against the oracle
(bit-exact f, p, v), through the C ABI."""
import numpy as np
import pytest

import oracle as O
import synth as S
import tal_chain as T
from helpers import DeviceArrays

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _check(eng, a, off, label, **kw):
    params = dict(bw=eng.params.bw, max_iter=eng.params.max_iter)
    want = O.chain_dp_sets(a, off, nthreads=8, **params, **kw)
    got = eng.dp(a, off, **kw)
    for name, x, y in zip("fpv", got, want):
        bad = np.where(x != y)[0]
        assert len(bad) == 0, f"{label}: {name} differs at {len(bad)} anchors, first {bad[:5]}"


@pytest.mark.parametrize("nsets,lo,hi", [(300, 1, 60), (200, 1, 3000), (20, 4000, 9000)])
def test_synthetic_sets(nsets, lo, hi):
    _say(f"[gpu] chain sets={nsets} size=[{lo},{hi}]")
    a, off = S.chain_generate(nsets, lo, hi, seed=nsets + hi)
    eng = T.DpChain()
    _check(eng, a, off, "default")
    st = eng.stats()
    assert st["sets"] == nsets and st["predecessors"] > 0
    eng.close()


@pytest.mark.parametrize("nsets,size,ring", [(300, 256, 512), (300, 257, 1024),
                                              (2000, 40, 512), (30, 3000, 1024)])
def test_block_shape_follows_set_size(nsets, size, ring, monkeypatch):
    """Batches of at most 256 anchors a set on average run one wave a set
    with a 512-slot ring, longer ones two waves a set with 1024 slots; both
    bit-exact."""
    _say(f"[gpu] chain shape sets={nsets} size={size}")
    monkeypatch.delenv("TAL_CHAIN_RING", raising=False)
    monkeypatch.delenv("TAL_CHAIN_WPB", raising=False)
    a, off = S.chain_generate(nsets, size, size, seed=size)
    assert len(a) == nsets * size
    eng = T.DpChain()
    _check(eng, a, off, f"sets={nsets} size={size}")
    assert eng.stats()["lds_ring"] == ring
    eng.close()


@pytest.mark.parametrize("kw,ctor", [(dict(dr=300, dq=200), dict(bw=20, max_iter=50)),
                                     (dict(dr=5000, dq=5000), dict(bw=5000, max_iter=8)),
                                     (dict(dr=100000, dq=100000), dict(bw=0, max_iter=5000)),
                                     (dict(dr=50, dq=5000), dict(bw=500, max_iter=3)),
                                     (dict(dr=5000, dq=5000), dict(bw=500, max_iter=100)),
                                     (dict(dr=-1, dq=5000), dict(bw=500, max_iter=700)),
                                     (dict(dr=5000, dq=-3), dict(bw=500, max_iter=5000)),
                                     (dict(dr=2147483000, dq=2147483000),
                                      dict(bw=1000, max_iter=900))])
def test_parameters_and_ring_wrap(kw, ctor):
    """Small max_iter keeps the predecessor windows inside the ring, whose
    slots are reused under 3000-anchor sets; dr = -1 (every r within reach)
    makes windows of max_iter; bw + min(dq, dr) >= 2^31 turns the gap table
    off."""
    _say(f"[gpu] chain params {kw} {ctor}")
    a, off = S.chain_generate(60, 1, 3000, seed=17)
    eng = T.DpChain(**ctor)
    _check(eng, a, off, f"{kw} {ctor}", **kw)
    eng.close()


@pytest.mark.parametrize("ring,wpb,table", [(128, 1, 1), (128, 4, 0), (512, 2, 1), (4096, 1, 1),
                                            (1024, 4, 0), (1024, 1, 0), (256, 2, 1),
                                            (2048, 4, 1)])
def test_ring_sizes_and_far_predecessors(ring, wpb, table, monkeypatch):
    """A 128-slot ring (64 in reach) sends most windows partly to HBM; 4096
    none.  Blocks of 1, 2 and 4 waves sharing a set; gap costs from the
    table or computed."""
    _say(f"[gpu] chain ring={ring} waves/block={wpb} table={table}")
    monkeypatch.setenv("TAL_CHAIN_RING", str(ring))
    monkeypatch.setenv("TAL_CHAIN_WPB", str(wpb))
    monkeypatch.setenv("TAL_CHAIN_TABLE", str(table))
    a, off = S.chain_generate(80, 1, 2500, seed=23)
    eng = T.DpChain()
    _check(eng, a, off, f"ring={ring} wpb={wpb} table={table}")
    st = eng.stats()
    assert st["lds_ring"] == ring
    if ring == 128:
        assert st["far_anchors"] > 0
    if ring == 4096:
        assert st["far_anchors"] == 0
    eng.close()


def test_rounding_ties_and_edges():
    _say("[gpu] chain rounding, ties, edges")
    rng = np.random.default_rng(3)
    eng = T.DpChain()
    # the gap-rounding set: float and double gap costs differ at dd = 29, 58 ...
    m = 300
    a = np.zeros(m, T.ANCHOR_DTYPE)
    a["r"] = 5000 + 40 * np.arange(m)
    a["q"] = 40 * np.arange(m) + rng.choice([0, 29, 58, 116, 5], m)
    a["l"] = 100
    _check(eng, a, np.array([0, m]), "rounding")
    # duplicates (ties), empty sets, single anchors, zero-length seeds
    parts, offs = [], [0]
    for size in (0, 1, 2, 0, 7, 6, 65, 64, 130, 1):
        r = np.sort(rng.integers(0, 3000, size)).astype(np.uint64)
        if size > 3:
            r[1::3] = r[0::3][:len(r[1::3])]
        part = np.zeros(size, T.ANCHOR_DTYPE)
        part["r"] = r
        part["q"] = (r.astype(np.int64) // 2 + rng.integers(0, 3, size)).astype(np.int32)
        part["l"] = rng.choice([0, 15, 15, 20], size)
        parts.append(part)
        offs.append(offs[-1] + size)
    _check(eng, np.concatenate(parts), np.array(offs), "edges")
    # r beyond 32 bits (the window uses 64-bit r, the scores its low 32 bits)
    a2, off2 = S.chain_generate(10, 100, 500, seed=8)
    a2["r"] += np.uint64(1 << 33)
    _check(eng, a2, off2, "64-bit r")
    # sets whose r span 2^31 or more, across the 2^32 boundary (the window
    # test needs all 64 bits)
    a3, off3 = S.chain_generate(12, 200, 900, seed=9)
    for k in range(12):
        b, e = off3[k], off3[k + 1]
        h = (b + e) // 2
        a3["r"][b:h] += np.uint64((1 << 32) - 3000)
        a3["r"][h:e] += np.uint64((1 << 32) + (1 << 31))
    _check(eng, a3, off3, "wide sets")
    eng.close()


@pytest.mark.parametrize("swaps", [0.02, 0.3])
def test_unsorted_sets(swaps):
    """r out of order within a set (the reference accepts it; its st scan
    then stops at the first r within reach, not at a sorted bound): the
    in-order st search instead of the binary search."""
    _say(f"[gpu] chain unsorted swaps={swaps}")
    rng = np.random.default_rng(31)
    a, off = S.chain_generate(120, 1, 2500, seed=29)
    for k in range(0, 120, 2):
        b, e = int(off[k]), int(off[k + 1])
        if e - b < 2:
            continue
        idx = rng.choice(e - b - 1, max(1, int((e - b) * swaps)), replace=False) + b
        r = a["r"].copy()
        r[idx], r[idx + 1] = a["r"][idx + 1], a["r"][idx]
        a["r"] = r
    eng = T.DpChain()
    _check(eng, a, off, f"unsorted {swaps}")
    eng.close()


def test_gap_costs_out_of_int32_range():
    """Gap costs at or past 2^31 convert to INT32_MIN as the reference's x86
    conversions do (the set of tests/test_chain_oracle.py, pinned against the
    reference build there)."""
    _say("[gpu] chain gap overflow")
    from test_chain_oracle import _overflow_set
    eng = T.DpChain()
    for seed in (6, 7):
        a, off = _overflow_set(2000, seed)
        _check(eng, a, off, f"overflow seed={seed}")
    eng.close()


def test_errors_and_device_path():
    _say("[gpu] chain errors and device path")
    a, off = S.chain_generate(50, 10, 800, seed=5)
    eng = T.DpChain()
    bad = a.copy()
    bad["l"][7] = -1
    with pytest.raises(T.TalError) as e:
        eng.dp(bad, off)
    assert e.value.code == T.TAL_ERANGE
    boff = off.copy()
    boff[3], boff[4] = boff[4], boff[3]
    with pytest.raises(T.TalError) as e:
        eng.dp(a, boff)
    assert e.value.code == T.TAL_EINVAL
    want = O.chain_dp_sets(a, off, nthreads=8)
    n = len(a)
    d = DeviceArrays(a, off, np.zeros(n, np.uint32), np.zeros(n, np.int32), np.zeros(n, np.int32))
    try:
        eng.dp_device(d.ptr(0), d.ptr(1), len(off) - 1, d.ptr(2), d.ptr(3), d.ptr(4))
        got = [d.download(2, np.zeros(n, np.uint32)), d.download(3, np.zeros(n, np.int32)),
               d.download(4, np.zeros(n, np.int32))]
    finally:
        d.close()
    for x, y in zip(got, want):
        assert np.array_equal(x, y)
    eng.close()
