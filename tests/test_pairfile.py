"""Pair files without a GPU: the text reader/writer (bench-bsw's pair-file
format, benchmarks/bench-bsw.cpp:117-169) and the packed mmap format.

The text reader is checked against an independent parse of the file in
Python (three lines per pair: h0, reference, query; base code = character
minus '0'), not against the writer alone, so a reader and writer that agree
on a wrong layout still fail."""
import os
import subprocess

import numpy as np
import pytest

import tal_bsw as T

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(T.LIB_PATH), "tal-bench-bsw")


def _random_pairs(n, seed, max1=254, max2=126):
    rng = np.random.default_rng(seed)
    len1 = rng.integers(0, max1 + 1, n)
    len2 = rng.integers(0, max2 + 1, n)
    len1[:3] = [0, max1, 1]
    len2[:3] = [max2, 0, 1]
    pairs = np.zeros(n, T.SEQPAIR_DTYPE)
    pairs["len1"], pairs["len2"] = len1, len2
    pairs["h0"] = rng.integers(-20, 200, n)
    pairs["idr"][1:] = np.cumsum(len1)[:-1]
    pairs["idq"][1:] = np.cumsum(len2)[:-1]
    ref = rng.integers(0, 5, max(1, int(len1.sum())), dtype=np.uint8)
    qer = rng.integers(0, 5, max(1, int(len2.sum())), dtype=np.uint8)
    return pairs, ref, qer


def _parse_text(path):
    """Independent parse: list of (h0, ref codes, query codes)."""
    lines = open(path).read().split("\n")
    assert lines[-1] == ""
    lines = lines[:-1]
    assert len(lines) % 3 == 0
    out = []
    for k in range(0, len(lines), 3):
        out.append((int(lines[k]), np.frombuffer(lines[k + 1].encode(), np.uint8) - 48,
                    np.frombuffer(lines[k + 2].encode(), np.uint8) - 48))
    return out


def _seq(buf, off, n):
    return buf[off:off + n]


def test_text_write_matches_independent_parse(tmp_path):
    pairs, ref, qer = _random_pairs(257, 1)
    path = tmp_path / "p.txt"
    T.write_pairfile(path, pairs, ref, qer)
    parsed = _parse_text(path)
    assert len(parsed) == len(pairs)
    for k, (h0, r, q) in enumerate(parsed):
        p = pairs[k]
        assert h0 == p["h0"]
        assert np.array_equal(r, _seq(ref, p["idr"], p["len1"]))
        assert np.array_equal(q, _seq(qer, p["idq"], p["len2"]))


def test_text_read_chunks_and_slots(tmp_path):
    pairs, ref, qer = _random_pairs(257, 2)
    path = tmp_path / "p.txt"
    T.write_pairfile(path, pairs, ref, qer)
    parsed = _parse_text(path)
    k = 0
    sizes = []
    for cp, cr, cq in T.read_pairfile(path, max_pairs=100):
        sizes.append(len(cp))
        for j in range(len(cp)):
            h0, r, q = parsed[k]
            p = cp[j]
            # bench-bsw's loadPairs layout: slot offsets, id = index in the
            # chunk, every other field -1
            assert (p["idr"], p["idq"], p["id"]) == (j * T.REF_SLOT, j * T.QER_SLOT, j)
            assert (p["len1"], p["len2"], p["h0"]) == (len(r), len(q), h0)
            for f in ["seqid", "regid"] + T.OUT_FIELDS:
                assert p[f] == -1
            assert np.array_equal(_seq(cr, p["idr"], p["len1"]), r)
            assert np.array_equal(_seq(cq, p["idq"], p["len2"]), q)
            k += 1
    assert sizes == [100, 100, 57]
    assert k == len(parsed)


def test_text_edge_files(tmp_path):
    # CRLF line ends, a trailing h0 line with no reference (a clean end in
    # bench-bsw) and an empty file
    path = tmp_path / "crlf.txt"
    path.write_bytes(b"3\r\n0123\r\n012\r\n7\r\n")
    chunks = list(T.read_pairfile(path))
    assert len(chunks) == 1 and len(chunks[0][0]) == 1
    p = chunks[0][0][0]
    assert (p["h0"], p["len1"], p["len2"]) == (3, 4, 3)
    assert list(chunks[0][1][:4]) == [0, 1, 2, 3]
    empty = tmp_path / "empty.txt"
    empty.write_bytes(b"")
    assert list(T.read_pairfile(empty)) == []


def test_text_errors(tmp_path):
    odd = tmp_path / "odd.txt"
    odd.write_text("5\n0123\n")          # a reference without its query
    with pytest.raises(T.TalError) as e:
        list(T.read_pairfile(odd))
    assert e.value.code == T.TAL_EINVAL
    # a slot of S bytes holds S-2 bases
    fits = tmp_path / "fits.txt"
    fits.write_text("1\n" + "0" * 254 + "\n" + "1" * 126 + "\n")
    (cp, _, _), = list(T.read_pairfile(fits))
    assert (cp[0]["len1"], cp[0]["len2"]) == (254, 126)
    for body in ("0" * 255 + "\n" + "1" * 10, "0" * 10 + "\n" + "1" * 127):
        long = tmp_path / "long.txt"
        long.write_text("1\n" + body + "\n")
        with pytest.raises(T.TalError) as e:
            list(T.read_pairfile(long))
        assert e.value.code == T.TAL_ERANGE
    with pytest.raises(T.TalError) as e:
        list(T.read_pairfile(tmp_path / "missing.txt"))
    assert e.value.code == T.TAL_EINVAL


def test_pack_round_trip(tmp_path):
    pairs, ref, qer = _random_pairs(300, 3, max1=1000, max2=600)
    path = tmp_path / "p.pack"
    T.write_pack(path, pairs, ref, qer)
    assert os.path.getsize(path) == 64 + pairs.nbytes + ref.nbytes + qer.nbytes
    with T.MappedPack(path) as m:
        assert np.array_equal(m.pairs, pairs)
        assert np.array_equal(m.ref, ref)
        assert np.array_equal(m.qer, qer)
        m.pairs["score"] = 7                 # private mapping: the file keeps its bytes
    with T.MappedPack(path) as m:
        assert np.array_equal(m.pairs, pairs)
    T.write_pack(tmp_path / "e.pack", pairs[:0], ref[:0], qer[:0])
    with T.MappedPack(tmp_path / "e.pack") as m:
        assert len(m.pairs) == 0 and len(m.ref) == 0


def test_pack_rejects_bad_files(tmp_path):
    pairs, ref, qer = _random_pairs(20, 4)
    good = tmp_path / "g.pack"
    T.write_pack(good, pairs, ref, qer)
    raw = good.read_bytes()
    cases = {"magic": b"X" + raw[1:], "short": raw[:-1], "header": raw[:40],
             "count": raw[:8] + (10**12).to_bytes(8, "little") + raw[16:]}
    for name, data in cases.items():
        bad = tmp_path / f"{name}.pack"
        bad.write_bytes(data)
        with pytest.raises(T.TalError) as e:
            T.MappedPack(bad)
        assert e.value.code == T.TAL_EINVAL, name


def test_cli_usage_without_gpu():
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stdout
    r = subprocess.run([CLI, "-pairs", "x", "-qer_slot", "2"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2
