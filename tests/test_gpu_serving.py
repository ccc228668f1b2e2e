"""GPU tests of the serving loop (tal_bsw_server_*), host-batched and
device-resident: every request's outputs are the same alone and in the mix,
bit-exact, whatever batch or kernel launch it ran in.

"Alone" is the request on a fresh engine handle through the 8/16-bit
dispatcher; the oracle (a fresh oracle object per request, 8-bit pairs then
16-bit pairs) checks both.
"""
import threading
import time

import numpy as np
import pytest

import oracle as O
import stale_batch
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


MODES = [pytest.param(False, id="host-batched"), pytest.param(True, id="resident")]


@pytest.mark.parametrize("resident", MODES)
def test_each_request_same_alone_and_in_mix(resident):
    print(f"[gpu] serving (resident={resident}): 12 requests, 32 MiB budget", flush=True)
    reqs = _requests()
    srv = T.Server(hbm_budget=32 << 20, workers=2, resident=resident)
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


@pytest.mark.parametrize("resident", MODES)
def test_concurrent_submitters_and_admission_waits(resident):
    """Four threads submit at once into a budget smaller than the offered
    requests, so admission has to wait for finished requests and reuses the
    device memory they held."""
    print(f"[gpu] serving (resident={resident}): 4 submitter threads, 6 MiB budget", flush=True)
    reqs = _requests(SPECS * 2)
    srv = T.Server(hbm_budget=6 << 20, workers=3, resident=resident)
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


@pytest.mark.parametrize("resident", MODES)
def test_bad_requests_fail_alone(resident):
    """A request over the 16-bit limit and one larger than the budget fail
    with their own status; the requests around them are served."""
    print(f"[gpu] serving (resident={resident}): failing requests", flush=True)
    good = _requests([(1, 500, 0), (2, 800, 0)])
    bad_pairs, bad_ref, bad_qer, _ = S.generate(1, 10)
    bad_pairs = bad_pairs.copy()
    bad_pairs["len2"][3] = 40000  # beyond MAX_SEQ_LEN16 - 1
    big = S.generate(2, 30000)
    srv = T.Server(hbm_budget=2 << 20, workers=1, resident=resident)
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


def test_resident_stale_column_reruns():
    """A request whose 16-bit groups read nonzero stale F[ncol] values: its
    resolution task finds them, the reruns run in the resident kernel, and the
    outputs equal a fresh object's (oracle), next to ordinary requests."""
    print("[gpu] serving (resident): stale-column batch", flush=True)
    pairs, ref, qer = stale_batch.make(32, 29)
    w2, r2, q2, _ = S.generate(2, 3000)
    w2 = w2[T.is_8bit(w2)].copy()
    w2["idr"] += len(ref)
    w2["idq"] += len(qer)
    sp, sr, sq = (np.concatenate([pairs, w2]), np.concatenate([ref, r2]),
                  np.concatenate([qer, q2]))
    others = _requests([(1, 900, 0), (2, 2000, 100)])
    srv = T.Server(hbm_budget=64 << 20, resident=True)
    try:
        mine = sp.copy()
        t = srv.submit(mine, sr, sq, 100)
        oth = [r[0].copy() for r in others]
        ot = [srv.submit(p, r[1], r[2], 100) for p, r in zip(oth, others)]
        info = srv.wait(t)
        oinfo = [srv.wait(x) for x in ot]
        st = srv.stats()
    finally:
        srv.close()
    assert info["status"] == T.TAL_OK and all(i["status"] == T.TAL_OK for i in oinfo)
    assert st["reruns"] > 0, "no stale-column rerun ran"
    assert_same_outputs(mine, _oracle(sp, sr, sq), "stale-column request vs oracle")
    for k, (r, got) in enumerate(zip(others, oth)):
        assert_same_outputs(got, _alone(*r), f"request {k} next to it")


def test_resident_request_joins_running_kernel():
    """Requests submitted while a long request runs are taken by the waves it
    leaves idle: they finish first, in the same kernel launch."""
    print("[gpu] serving (resident): short requests overtake a long one", flush=True)
    pairs, ref, qer, w = S.generate(3, 20000)
    sel = np.where(w == 500)[0][:2000]
    long_req = (pairs[sel].copy(), ref, qer)
    shorts = _requests([(1, 400, k * 400) for k in range(4)])
    srv = T.Server(hbm_budget=256 << 20, resident=True, max_query_len=2048)
    try:
        lp = long_req[0].copy()
        tl = srv.submit(lp, long_req[1], long_req[2], 500)
        time.sleep(0.005)
        sp = [r[0].copy() for r in shorts]
        ts = [srv.submit(p, r[1], r[2], 100) for p, r in zip(sp, shorts)]
        done_at = {}
        while len(done_at) < 1 + len(ts):
            for t in [tl] + ts:
                if t not in done_at:
                    info = srv.wait(t, 0)
                    if info is not None:
                        done_at[t] = (time.perf_counter(), info)
            time.sleep(0.0005)
        st = srv.stats()
    finally:
        srv.close()
    assert all(v[1]["status"] == T.TAL_OK for v in done_at.values())
    assert all(done_at[t][0] < done_at[tl][0] for t in ts), "a short request waited for the long one"
    assert max(done_at[t][1]["batch_requests"] for t in ts) > 1
    assert st["batches"] == 1, f"{st['batches']} kernel launches"
    assert_same_outputs(lp, _alone(*long_req, w=500), "long request")
    for k, (r, got) in enumerate(zip(shorts, sp)):
        assert_same_outputs(got, _alone(*r), f"short request {k}")


def test_resident_kernel_leaves_when_idle_and_returns():
    """After a stretch without work the kernel leaves; the next request
    launches it again and is served the same."""
    print("[gpu] serving (resident): idle exit and relaunch", flush=True)
    reqs = _requests([(1, 300, 0), (2, 500, 0)])
    srv = T.Server(hbm_budget=16 << 20, resident=True)
    try:
        got = []
        for r in reqs:
            p = r[0].copy()
            info = srv.wait(srv.submit(p, r[1], r[2], 100))
            assert info["status"] == T.TAL_OK
            got.append(p)
            time.sleep(0.2)
        st = srv.stats()
    finally:
        srv.close()
    assert st["batches"] == 2, f"{st['batches']} kernel launches"
    for k, (r, p) in enumerate(zip(reqs, got)):
        assert_same_outputs(p, _alone(*r), f"request {k}")
