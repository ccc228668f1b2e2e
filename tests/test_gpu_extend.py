"""GPU parity of the extension pipeline (libtal_extend.so) through the C ABI:
the producer's pairs and bytes equal the model's (tests/extend_model.py),
and every extension score equals the BSW oracle's on the same pairs in the
same order, the right pairs' h0 coming from the left scores."""
import numpy as np
import pytest

import extend_model as M
import oracle as O
import synth as S
import tal_bsw as TB
import tal_chain as TC
import tal_extend as TE
from helpers import DeviceArrays

pytestmark = pytest.mark.gpu

IN_FIELDS = ["idr", "idq", "id", "len1", "len2", "h0", "seqid", "regid"]


def _say(msg):
    print(msg, flush=True)


def _check(eng, d, label):
    got = eng.run(d["genome"], d["reads"], d["read_off"], d["anchors"], d["anchor_off"])
    want, wl, wr = M.pipeline(d["genome"], d["reads"], d["read_off"], d["anchors"],
                              d["anchor_off"])
    for fld in TE.READ_DTYPE.names:
        bad = np.where(got[fld] != want[fld])[0]
        assert len(bad) == 0, f"{label}: read field {fld} differs at {bad[:5]}"
    gl, gr, gref, gqer = eng.pairs()
    assert len(gl) == len(wl) and len(gr) == len(wr), label
    for fld in IN_FIELDS + TB.OUT_FIELDS:
        assert np.array_equal(gl[fld], wl[fld]), f"{label}: left {fld}"
        assert np.array_equal(gr[fld], wr[fld]), f"{label}: right {fld}"
    st = eng.stats()
    assert st["left_pairs"] == len(wl) and st["right_pairs"] == len(wr)
    return got, gref, gqer


def test_pipeline_on_simulated_reads():
    _say("[gpu] extend pipeline 3000 reads")
    d = S.extend_generate(3000, 150, 2_000_000, seed=21)
    eng = TE.Extender()
    got, ref, qer = _check(eng, d, "sim")
    assert (got["score"] >= got["left_score"]).mean() > 0.99
    eng.close()


def test_pipeline_edges():
    _say("[gpu] extend edges")
    rng = np.random.default_rng(4)
    L = 5000
    genome = rng.integers(0, 4, L, dtype=np.uint8)
    reads, anchors, aoff, roff = [], [], [0], [0]

    def add(read, anc):
        reads.append(np.asarray(read, np.uint8))
        roff.append(roff[-1] + len(read))
        a = np.zeros(len(anc), TC.ANCHOR_DTYPE)
        for i, (r, q, ln) in enumerate(anc):
            a[i] = (r, q, ln)
        anchors.append(a)
        aoff.append(aoff[-1] + len(anc))

    add(genome[100:250], [])                                # no anchors: no chain
    add(genome[300:450], [(300, 0, 150)])                   # the whole read: no sides
    add(genome[0:150], [(20, 20, 40), (60, 60, 40)])        # left side clipped by the genome start
    add(genome[L - 150:L], [(L - 100, 50, 30)])             # right side clipped by the genome end
    add(genome[1000:1120], [(1000, 0, 19), (1010, 10, 19), (1050, 50, 19), (1100, 100, 20)])
    noisy = genome[2000:2150].copy()
    noisy[70] = (noisy[70] + 1) % 4
    add(noisy, [(2000, 0, 19), (2020, 20, 19), (2040, 40, 19), (2080, 80, 19), (2100, 100, 19),
                (4000, 5, 19)])
    # a long read whose chain has 990 anchors: a walk over many register windows
    add(genome[2500:4500], [(2500 + q, q, 19) for q in range(0, 1980, 2)])
    # two equal seeds on diagonals too far apart to chain: f ties, the first wins
    add(genome[600:750], [(600, 0, 30), (3600, 100, 30)])
    # a chain whose back-pointers skip runs of off-diagonal anchors (window refills)
    chain = [(1500 + q, q, 19) for q in (0, 40, 80, 120)]
    junk = [(1500 + q + 2 + k, int(rng.integers(0, 131)), 19) for q in (0, 40, 80)
            for k in range(0, 36, 3)]
    add(genome[1500:1650], sorted(chain + junk))
    d = dict(genome=genome, reads=np.concatenate(reads), read_off=np.array(roff, np.int64),
             anchors=np.concatenate(anchors), anchor_off=np.array(aoff, np.int64))
    eng = TE.Extender()
    got, _, _ = _check(eng, d, "edges")
    assert got["chain_anchors"][6] == 990
    assert got["chain_score"][7] == 30 and got["seed_r"][7] == 600
    assert got["chain_anchors"][8] >= 4
    assert got["chain_anchors"][0] == 0 and got["left_pair"][0] == -1
    assert got["left_pair"][1] == -1 and got["right_pair"][1] == -1
    assert got["score"][1] == 150
    eng.close()


def test_pipeline_errors():
    _say("[gpu] extend errors")
    rng = np.random.default_rng(5)
    genome = rng.integers(0, 4, 1000, dtype=np.uint8)
    eng = TE.Extender()
    a = np.zeros(1, TC.ANCHOR_DTYPE)
    a[0] = (10, 140, 20)  # q + l past the 150-base read
    with pytest.raises(TE.TalError) as e:
        eng.run(genome, genome[:150], np.array([0, 150]), a, np.array([0, 1]))
    assert e.value.code == TE.TAL_EINVAL
    long_read = rng.integers(0, 4, 40000, dtype=np.uint8)
    a[0] = (10, 0, 19)
    with pytest.raises(TE.TalError) as e:
        eng.run(np.concatenate([genome, long_read]), long_read, np.array([0, 40000]), a,
                np.array([0, 1]))
    assert e.value.code == TE.TAL_ERANGE
    # the handle still works afterwards
    d = S.extend_generate(50, 150, 100_000, seed=8)
    _check(eng, d, "after errors")
    eng.close()


def _device_inputs(d):
    arrays = [np.ascontiguousarray(d["genome"], np.uint8),
              np.ascontiguousarray(d["reads"], np.uint8),
              np.ascontiguousarray(d["read_off"], np.int64),
              np.ascontiguousarray(d["anchors"], TC.ANCHOR_DTYPE),
              np.ascontiguousarray(d["anchor_off"], np.int64)]
    return arrays, DeviceArrays(*arrays)


def test_device_entry_points():
    """tal_extend_run_device on device buffers gives what tal_extend_run
    gives on host buffers; tal_extend_produce_device on the chaining
    oracle's f and p gives the model's pairs, bytes and seed records."""
    _say("[gpu] extend device entry points 500 reads")
    d = S.extend_generate(500, 150, 500_000, seed=31)
    n = len(d["read_off"]) - 1
    host = TE.Extender()
    want = host.run(d["genome"], d["reads"], d["read_off"], d["anchors"], d["anchor_off"])
    host.close()
    arrays, dev = _device_inputs(d)
    out = np.zeros(n, TE.READ_DTYPE)
    dout = DeviceArrays(out)
    eng = TE.Extender()
    eng.run_device(dev.ptr(0), len(arrays[0]), dev.ptr(1), dev.ptr(2), n, dev.ptr(3), dev.ptr(4),
                   dout.ptr(0))
    got = dout.download(0, out.copy())
    for fld in TE.READ_DTYPE.names:
        assert np.array_equal(got[fld], want[fld]), f"run_device: {fld}"
    # the producer alone, on the oracle's chaining results
    f, p, _ = O.chain_dp_sets(d["anchors"], d["anchor_off"], nthreads=8)
    fp = DeviceArrays(np.ascontiguousarray(f, np.uint32), np.ascontiguousarray(p, np.int32))
    eng.produce_device(dev.ptr(0), len(arrays[0]), dev.ptr(1), dev.ptr(2), n, dev.ptr(3),
                       dev.ptr(4), fp.ptr(0), fp.ptr(1), dout.ptr(0))
    got = dout.download(0, out.copy())
    mout, ml, mr, mref, mqer = M.produce(d["genome"], d["reads"], d["read_off"], d["anchors"],
                                         d["anchor_off"], f, p)
    for fld in ("seed_r", "seed_q", "seed_len", "chain_score", "chain_anchors", "left_pair",
                "right_pair"):
        assert np.array_equal(got[fld], mout[fld]), f"produce_device: {fld}"
    gl, gr, gref, gqer = eng.pairs()
    for fld in IN_FIELDS:
        assert np.array_equal(gl[fld], ml[fld]), f"produce_device: left {fld}"
        assert np.array_equal(gr[fld], mr[fld]), f"produce_device: right {fld}"
    assert np.array_equal(gref, mref) and np.array_equal(gqer, mqer)
    for h in (fp, dout, dev):
        h.close()
    eng.close()


def test_invalid_offsets_are_reported():
    """Negative or decreasing anchor or read offsets give TAL_EINVAL from the
    producer (nothing outside the offsets' ranges is read), and the handle
    still works afterwards."""
    _say("[gpu] extend invalid offsets")
    rng = np.random.default_rng(9)
    genome = rng.integers(0, 4, 2000, dtype=np.uint8)
    reads = np.concatenate([genome[100:250], genome[500:650]])
    a = np.zeros(8, TC.ANCHOR_DTYPE)
    for i in range(8):
        a[i] = (100 + 10 * i, 10 * i, 19)
    f = np.full(8, 19, np.uint32)
    p = np.full(8, -1, np.int32)
    out = np.zeros(2, TE.READ_DTYPE)
    eng = TE.Extender()
    cases = [("decreasing anchor offsets", [0, 150, 300], [0, 5, 3]),
             ("negative anchor start", [0, 150, 300], [-4, 2, 8]),
             ("decreasing read offsets", [0, 150, 100], [0, 8, 8])]
    for label, roff, aoff in cases:
        dev = DeviceArrays(genome, reads, np.array(roff, np.int64), a, np.array(aoff, np.int64),
                           f, p, out)
        with pytest.raises(TE.TalError) as e:
            eng.produce_device(dev.ptr(0), len(genome), dev.ptr(1), dev.ptr(2), 2, dev.ptr(3),
                               dev.ptr(4), dev.ptr(5), dev.ptr(6), dev.ptr(7))
        assert e.value.code == TE.TAL_EINVAL, label
        assert "offsets" in str(e.value), label
        dev.close()
    with pytest.raises(TE.TalError) as e:
        eng.run(genome, reads, np.array([0, 150, 100]), a, np.array([0, 8, 8]))
    assert e.value.code == TE.TAL_EINVAL
    # a negative genome length is refused before any kernel runs
    dev = DeviceArrays(genome, reads, np.array([0, 150, 300], np.int64), a,
                       np.array([0, 4, 8], np.int64), f, p, out)
    with pytest.raises(TE.TalError) as e:
        eng.produce_device(dev.ptr(0), -1, dev.ptr(1), dev.ptr(2), 2, dev.ptr(3),
                           dev.ptr(4), dev.ptr(5), dev.ptr(6), dev.ptr(7))
    assert e.value.code == TE.TAL_EINVAL and "genome_len" in str(e.value)
    dev.close()
    # an anchor whose reference position is near 2^64 (r + l wraps) is the
    # best chain of read 0 and must be reported, not read through
    a2 = a.copy()
    a2[1]["r"] = np.uint64(2**64 - 5)
    f2 = f.copy()
    f2[1] = 60
    dev = DeviceArrays(genome, reads, np.array([0, 150, 300], np.int64), a2,
                       np.array([0, 4, 8], np.int64), f2, p, out)
    with pytest.raises(TE.TalError) as e:
        eng.produce_device(dev.ptr(0), len(genome), dev.ptr(1), dev.ptr(2), 2, dev.ptr(3),
                           dev.ptr(4), dev.ptr(5), dev.ptr(6), dev.ptr(7))
    assert e.value.code == TE.TAL_EINVAL and "outside" in str(e.value)
    dev.close()
    d = S.extend_generate(40, 150, 100_000, seed=12)
    _check(eng, d, "after invalid offsets")
    eng.close()
