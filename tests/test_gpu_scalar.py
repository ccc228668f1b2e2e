"""GPU parity of the per-pair (scalar) path, tal_bsw_scalar_scores
(scalarBandedSWAWrapper, bandedSWA.cpp:242-260, and scalarBandedSWA,
:116-237), against the oracle's scalar restatement, on samples of the
workloads, a non-default scoring matrix, edge shapes, bench-bsw's non-AVX2
branch through the C++ shim (tal-bench-bsw -scalar), and its errors."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import synth as S
import tal_bsw as T
from helpers import assert_same_outputs

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(T.LIB_PATH), "tal-bench-bsw")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def engine():
    e = T.BandedPairWiseSW()
    yield e
    e.close()


@pytest.mark.parametrize("cfg,n", [(1, 3000), (2, 5000), (5, 3000)])
def test_scalar_workloads(engine, cfg, n):
    _say(f"[gpu] scalar path: workload {cfg}, {n} pairs")
    pairs, ref, qer, _ = S.generate(cfg, n, first=500)
    want = O.Oracle().scalar(pairs.copy(), ref, qer, 100)
    got = engine.scalar_scores(pairs.copy(), ref, qer, 100)
    assert_same_outputs(got, want, f"scalar workload {cfg}")


def test_scalar_workload3_bands(engine):
    pairs, ref, qer, w = S.generate(3, 1600)
    for wv in (50, 100, 200, 500):
        p = pairs[w == wv].copy()
        _say(f"[gpu] scalar path: workload 3 w={wv}, {len(p)} pairs")
        want = O.Oracle().scalar(p.copy(), ref, qer, wv)
        got = engine.scalar_scores(p.copy(), ref, qer, wv)
        assert_same_outputs(got, want, f"scalar workload 3 w={wv}")


def test_scalar_custom_matrix():
    """The matrix is the caller's: asymmetric scores, an N row and column
    of -2, and a larger maximum (which also widens the band clip)."""
    _say("[gpu] scalar path: custom matrix and gaps")
    mat = np.array([[2, -3, -1, -3, -2],
                    [-3, 3, -3, -1, -2],
                    [-1, -3, 2, -3, -2],
                    [-3, -1, -3, 4, -2],
                    [-2, -2, -2, -2, -2]], np.int8)
    kw = dict(o_del=5, e_del=2, o_ins=7, e_ins=1, zdrop=40, end_bonus=3, w_match=2, w_mismatch=3)
    pairs, ref, qer, _ = S.generate(5, 4000, first=9000)
    eng = T.BandedPairWiseSW(**kw)
    got = eng.scalar_scores(pairs.copy(), ref, qer, 60, mat=mat)
    eng.close()
    want = O.Oracle(**kw).scalar_mat(mat, pairs.copy(), ref, qer, 60)
    assert_same_outputs(got, want, "custom matrix")


def test_scalar_edges(engine):
    _say("[gpu] scalar path: edge shapes")
    rng = np.random.default_rng(77)
    refs, qrys, h0s = [], [], []
    shapes = [(0, 0), (0, 10), (10, 0), (1, 1), (1, 40), (40, 1), (300, 300), (2000, 1500),
              (127, 128), (128, 127)]
    for l1, l2 in shapes:
        for h0 in (0, 1, 7, 60, 30000):
            q = rng.integers(0, 5, l2).astype(np.uint8)
            r = q[:l1].copy() if l1 <= l2 else np.concatenate([q, rng.integers(0, 5, l1 - l2)])
            r = r.astype(np.uint8)
            sub = rng.random(len(r)) < 0.1
            r[sub] = rng.integers(0, 5, int(sub.sum()))
            refs.append(r)
            qrys.append(q)
            h0s.append(h0)
    pairs, ref, qer = O.make_pairs(refs, qrys, h0s)
    for w in (0, 3, 100, 1000):
        want = O.Oracle().scalar(pairs.copy(), ref, qer, w)
        got = engine.scalar_scores(pairs.copy(), ref, qer, w)
        assert_same_outputs(got, want, f"edges w={w}")
    # z-drop off (zdrop <= 0 skips the test, :210)
    e0 = T.BandedPairWiseSW(zdrop=0)
    want = O.Oracle(zdrop=0).scalar(pairs.copy(), ref, qer, 100)
    got = e0.scalar_scores(pairs.copy(), ref, qer, 100)
    e0.close()
    assert_same_outputs(got, want, "zdrop 0")


def test_scalar_errors(engine):
    _say("[gpu] scalar path: errors")
    pairs, ref, qer, _ = S.generate(1, 64)
    bad = qer.copy()
    bad[int(pairs["idq"][5]) + 3] = 5
    with pytest.raises(T.TalError) as e:
        engine.scalar_scores(pairs.copy(), ref, bad, 100)
    assert e.value.code == T.TAL_ERANGE
    badr = ref.copy()
    badr[int(pairs["idr"][7])] = 9
    with pytest.raises(T.TalError) as e:
        engine.scalar_scores(pairs.copy(), badr, qer, 100)
    assert e.value.code == T.TAL_ERANGE
    # the handle still works
    want = O.Oracle().scalar(pairs.copy(), ref, qer, 100)
    assert_same_outputs(engine.scalar_scores(pairs.copy(), ref, qer, 100), want, "after errors")


def test_cli_scalar_branch(tmp_path):
    """bench-bsw's branch for builds without AVX2 (bench-bsw.cpp:256-258),
    compiled against the C++ shim: scalarBandedSWAWrapper per chunk."""
    pairs, ref, qer, _ = S.generate(2, 3000)
    keep = np.where((pairs["len1"] <= 254) & (pairs["len2"] <= 126))[0][:2500]
    pairs = np.ascontiguousarray(pairs[keep])
    path = tmp_path / "pairs.txt"
    T.write_pairfile(path, pairs, ref, qer)
    (cp, cr, cq), = list(T.read_pairfile(path))
    want = O.Oracle().scalar(cp.copy(), cr, cq, 100)
    _say("[gpu] tal-bench-bsw -scalar")
    r = subprocess.run([CLI, "-pairs", str(path), "-maxi", "-scalar"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    rows = []
    for line in r.stderr.splitlines():
        a, rest = line.split(" (")
        tq, rest = rest.split(") ")
        tle, qle = tq.split()
        gscore, max_off, gtle = rest.split()
        rows.append((int(a), int(tle), int(qle), int(gscore), int(max_off), int(gtle)))
    got = np.array(rows)
    exp = np.stack([want[f] for f in ("score", "tle", "qle", "gscore", "max_off", "gtle")], 1)
    assert np.array_equal(got, exp)
