"""The vectorised restatement (oracle/bsw_simd.cpp, the CPU baseline) equals
the lane-by-lane oracle (oracle/bsw_oracle.c) on every output of every pair:
samples of the workloads in both widths, every edge batch of the GPU edge
tests, and a sequence of calls on one object (the F scratch carries over).
CPU only."""
import zlib

import numpy as np
import pytest

import oracle as O
import synth as S
import tal_bsw as T
from helpers import assert_same_outputs
from test_gpu_edges import CASES8, CASES16, _batch

KEYS = ("o_del", "e_del", "o_ins", "e_ins", "zdrop", "end_bonus", "w_match", "w_mismatch")


def _params(over):
    d = dict(o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100, end_bonus=5, w_match=1, w_mismatch=4)
    d.update(over)
    return tuple(d[k] for k in KEYS), d


def test_simd_build_is_selected():
    assert O.simd_width() in (256, 512)


@pytest.mark.parametrize("cfg,n", [(1, 3000), (2, 6000), (5, 3000)])
def test_simd_workloads(cfg, n):
    pairs, ref, qer, _ = S.generate(cfg, n, first=777)
    is8 = T.is_8bit(pairs)
    for bits, sel in ((16, np.ones(n, bool)), (8, is8)):
        p = pairs[sel].copy()
        want = O.run_mt(p.copy(), ref, qer, 100, bits=bits, chunk=1000)
        got = O.simd_run_mt(p.copy(), ref, qer, 100, bits=bits, chunk=1000)
        assert_same_outputs(got, want, f"workload {cfg} {bits}-bit")


def test_simd_workload3_bands():
    pairs, ref, qer, w = S.generate(3, 1600)
    for wv in (50, 100, 200, 500):
        p = pairs[w == wv].copy()
        want = O.run_mt(p.copy(), ref, qer, wv, bits=16)
        got = O.simd_run_mt(p.copy(), ref, qer, wv, bits=16)
        assert_same_outputs(got, want, f"workload 3 w={wv}")


@pytest.mark.parametrize("name,params,n,lens,h0s,alpha,w", CASES16, ids=[c[0] for c in CASES16])
def test_simd_edges16(name, params, n, lens, h0s, alpha, w):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    pairs, ref, qer = _batch(rng, n, *lens, *h0s, alphabet=alpha, zero_len=0.1)
    tup, d = _params(params)
    want = O.Oracle(**d).scores16(pairs.copy(), ref, qer, w)
    got = O.simd_run_mt(pairs.copy(), ref, qer, w, bits=16, chunk=n, params=tup)
    assert_same_outputs(got, want, name)


@pytest.mark.parametrize("name,params,n,lens,h0s,alpha,w", CASES8, ids=[c[0] for c in CASES8])
def test_simd_edges8(name, params, n, lens, h0s, alpha, w):
    rng = np.random.default_rng(zlib.crc32((name + "8").encode()))
    pairs, ref, qer = _batch(rng, n, *lens, *h0s, alphabet=alpha, zero_len=0.1)
    tup, d = _params(params)
    want = O.Oracle(**d).scores8(pairs.copy(), ref, qer, w)
    got = O.simd_run_mt(pairs.copy(), ref, qer, w, bits=8, chunk=n, params=tup)
    assert_same_outputs(got, want, name)


def test_simd_chunks_share_one_object():
    """Chunks on one thread run on one object, so a later chunk sees the F
    scratch an earlier one left, as on the reference object."""
    pairs, ref, qer, _ = S.generate(5, 5000)
    want = O.run_mt(pairs.copy(), ref, qer, 100, bits=16, chunk=700, nthreads=1)
    got = O.simd_run_mt(pairs.copy(), ref, qer, 100, bits=16, chunk=700, nthreads=1)
    assert_same_outputs(got, want, "chunks")
