"""GPU tests of the serving loop (tal_bsw_server_*): every request's outputs
are the same alone and in the mix, bit-exact, whatever batch it ran in.

"Alone" is the request on a fresh engine handle through the 8/16-bit
dispatcher; the oracle (a fresh oracle object per request, 8-bit pairs then
16-bit pairs) checks both.
"""
import threading

import numpy as np
import pytest

import oracle as O
import synth as S
import tal_bsw as T
from helpers import assert_same_outputs

pytestmark = pytest.mark.gpu

# (workload, pairs, first pair): a mix of sizes and of the three generators
SPECS = [(1, 1500, 0), (2, 3000, 5000), (5, 900, 100), (2, 64, 0), (1, 33, 7),
         (5, 4000, 2000), (2, 7000, 10000), (1, 2500, 3000), (2, 1, 9), (5, 257, 50),
         (2, 5000, 40000), (1, 700, 900)]


def _requests(specs=SPECS):
    out = []
    for cfg, n, first in specs:
        pairs, ref, qer, _ = S.generate(cfg, n, first)
        out.append((pairs, ref, qer))
    return out


def _alone(pairs, ref, qer, w=100):
    e = T.BandedPairWiseSW()
    try:
        return e.get_scores_auto(pairs.copy(), ref, qer, w)
    finally:
        e.close()


def _oracle(pairs, ref, qer, w=100):
    is8 = T.is_8bit(pairs)
    want = pairs.copy()
    o = O.Oracle()
    a = o.scores8(pairs[is8].copy(), ref, qer, w)
    b = o.scores16(pairs[~is8].copy(), ref, qer, w)
    o.close()
    for f in T.OUT_FIELDS:
        want[f][is8] = a[f]
        want[f][~is8] = b[f]
    return want


def test_each_request_same_alone_and_in_mix():
    print("[gpu] serving: 12 requests, 2 workers, 32 MiB budget", flush=True)
    reqs = _requests()
    srv = T.Server(hbm_budget=32 << 20, workers=2)
    try:
        mine = [r[0].copy() for r in reqs]
        tickets = [srv.submit(p, r[1], r[2], 100) for p, r in zip(mine, reqs)]
        infos = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert all(i["status"] == T.TAL_OK for i in infos), infos
    assert max(i["batch_requests"] for i in infos) > 1, "no batch mixed requests"
    assert st["requests"] == len(reqs) and st["pairs"] == sum(len(r[0]) for r in reqs)
    for k, (r, got) in enumerate(zip(reqs, mine)):
        alone = _alone(*r)
        assert_same_outputs(got, alone, f"request {k} mix vs alone")
        if k % 3 == 0:
            assert_same_outputs(got, _oracle(*r), f"request {k} mix vs oracle")


def test_concurrent_submitters_and_admission_waits():
    """Four threads submit at once into a budget smaller than the offered
    requests, so admission has to wait for finished batches."""
    print("[gpu] serving: 4 submitter threads, 6 MiB budget", flush=True)
    reqs = _requests(SPECS * 2)
    srv = T.Server(hbm_budget=6 << 20, workers=3)
    mine = [r[0].copy() for r in reqs]
    results = [None] * len(reqs)

    def client(ks):
        for k in ks:
            t = srv.submit(mine[k], reqs[k][1], reqs[k][2], 100)
            results[k] = srv.wait(t)

    try:
        ths = [threading.Thread(target=client, args=(range(j, len(reqs), 4),)) for j in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join(timeout=120)
        st = srv.stats()
    finally:
        srv.close()
    assert all(r is not None and r["status"] == T.TAL_OK for r in results), results
    assert st["peak_bytes"] <= 6 << 20
    cache = {}
    for k, r in enumerate(reqs):
        key = k % len(SPECS)
        if key not in cache:
            cache[key] = _alone(*r)
        assert_same_outputs(mine[k], cache[key], f"request {k}")


def test_bad_requests_fail_alone():
    """A request over the 16-bit limit and one larger than the budget fail
    with their own status; the requests around them are served."""
    print("[gpu] serving: failing requests", flush=True)
    good = _requests([(1, 500, 0), (2, 800, 0)])
    bad_pairs, bad_ref, bad_qer, _ = S.generate(1, 10)
    bad_pairs = bad_pairs.copy()
    bad_pairs["len2"][3] = 40000  # beyond MAX_SEQ_LEN16 - 1
    big = S.generate(2, 30000)
    srv = T.Server(hbm_budget=2 << 20, workers=1)
    try:
        p0 = good[0][0].copy()
        p1 = good[1][0].copy()
        t0 = srv.submit(p0, good[0][1], good[0][2], 100)
        tb = srv.submit(bad_pairs, bad_ref, np.zeros(50000, np.uint8), 100)
        tbig = srv.submit(big[0].copy(), big[1], big[2], 100)
        t1 = srv.submit(p1, good[1][1], good[1][2], 100)
        i0, ib, ibig, i1 = (srv.wait(t) for t in (t0, tb, tbig, t1))
    finally:
        srv.close()
    assert ib["status"] == T.TAL_ERANGE and ib["batch"] == -1
    assert ibig["status"] == T.TAL_ERANGE
    assert i0["status"] == T.TAL_OK and i1["status"] == T.TAL_OK
    assert_same_outputs(p0, _alone(*good[0]), "good request 0")
    assert_same_outputs(p1, _alone(*good[1]), "good request 1")
