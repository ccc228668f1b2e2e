"""The row-pair model (oracle/rowpair_model.c) on samples of the benchmark
workloads: every block of row i+1 it decides while row i is still sweeping
must carry exactly the column masks and the (trimmed) inputs that the oracle's
own group loop gives that block when row i+1 starts.  The model observes the
oracle's 16-bit loop (bandedSWA.cpp:2962-3356) and changes no result."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import synth as S

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")
LIB = os.path.join(ORACLE, "_build", "librowpair.so")


@pytest.fixture(scope="module")
def model():
    subprocess.check_call(["make", "-s", "-C", ORACLE, LIB])
    L = ctypes.CDLL(LIB)
    L.bsw_oracle_create.restype = ctypes.c_void_p
    return L


def _run(L, wl, n, cls=None):
    pairs, ref, qer, w = S.generate(wl, n)
    if cls is not None:
        pairs = pairs[np.where(w == cls)[0]].copy()
        w = cls
    w = int(np.asarray(w).ravel()[0])
    h = L.bsw_oracle_create(6, 1, 6, 1, 100, 5, 1, 4)
    try:
        L.rp_reset()
        L.bsw_oracle_get_scores16(ctypes.c_void_p(h), ctypes.c_void_p(pairs.ctypes.data),
                                  ctypes.c_void_p(ref.ctypes.data),
                                  ctypes.c_void_p(qer.ctypes.data), len(pairs), int(w))
    finally:
        L.bsw_oracle_destroy(ctypes.c_void_p(h))
    out = (ctypes.c_uint64 * 4)()
    L.rp_stats(out)
    return list(out)


@pytest.mark.parametrize("wl,n,cls", [(2, 3000, None), (1, 2000, None), (5, 3000, None),
                                      (3, 3000, 50), (3, 3000, 200), (3, 3000, 500)])
def test_early_blocks_match_the_oracle(model, wl, n, cls):
    blocks, early, mismatch, rows = _run(model, wl, n, cls)
    print(f"workload {wl} class {cls}: {rows} rows, {blocks} blocks, {early} early "
          f"({early / max(blocks, 1):.3f})")
    assert rows > 0 and blocks > 0
    assert mismatch == 0
    assert early > blocks // 4, "the rules should decide a good share of the blocks early"


def test_planted_mistake_is_caught(model):
    """With the oracle's range reset (:3267) dropped, row i+1's real start and
    scans differ from what the rules assume, and early blocks mismatch: the
    comparison has teeth."""
    L = model
    pairs, ref, qer, w = S.generate(2, 3000)
    w = int(np.asarray(w).ravel()[0])
    h = L.bsw_oracle_create(6, 1, 6, 1, 100, 5, 1, 4)
    try:
        L.bsw_oracle_set_mutation(ctypes.c_void_p(h), 7)
        L.rp_reset()
        L.bsw_oracle_get_scores16(ctypes.c_void_p(h), ctypes.c_void_p(pairs.ctypes.data),
                                  ctypes.c_void_p(ref.ctypes.data),
                                  ctypes.c_void_p(qer.ctypes.data), len(pairs), int(w))
    finally:
        L.bsw_oracle_destroy(ctypes.c_void_p(h))
    out = (ctypes.c_uint64 * 4)()
    L.rp_stats(out)
    assert out[2] > 0
