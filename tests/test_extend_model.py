"""The extension-pair producer's model (tests/extend_model.py) on the CPU:
its pairs point at the right bytes, h0 follows the rules, and the library
exports what include/tal_extend.h declares."""
import ctypes
import os
import re

import numpy as np
import pytest

import extend_model as M
import oracle as O
import synth as S
import tal_extend as TE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _data(nreads=60, seed=5):
    return S.extend_generate(nreads, 150, 200_000, seed=seed)


def test_generator_anchors_are_kmer_matches():
    d = _data()
    g, rd, a, off = d["genome"], d["reads"].reshape(-1, 150), d["anchors"], d["anchor_off"]
    good = 0
    for k in range(len(off) - 1):
        for x in a[off[k]:off[k + 1]]:
            r, q = int(x["r"]), int(x["q"])
            good += np.array_equal(g[r:r + 19], rd[k, q:q + 19])
    assert good >= len(a) - 6 * 60  # all but the spurious ones (6 a read) match


def test_model_pairs_point_at_the_sides():
    d = _data()
    f, p, _ = O.chain_dp_sets(d["anchors"], d["anchor_off"])
    out, left, right, ref, qer = M.produce(d["genome"], d["reads"], d["read_off"], d["anchors"],
                                           d["anchor_off"], f, p)
    reads = d["reads"].reshape(-1, 150)
    for k in range(len(out)):
        o = out[k]
        qb, ln, rb = int(o["seed_q"]), int(o["seed_len"]), int(o["seed_r"])
        # the seed is an exact match
        assert np.array_equal(d["genome"][rb:rb + ln], reads[k, qb:qb + ln])
        if o["left_pair"] >= 0:
            sp = left[o["left_pair"]]
            assert sp["h0"] == ln and sp["len2"] == qb and sp["seqid"] == k
            assert np.array_equal(qer[sp["idq"]:sp["idq"] + sp["len2"]], reads[k, :qb][::-1])
            assert np.array_equal(ref[sp["idr"]:sp["idr"] + sp["len1"]],
                                  d["genome"][rb - sp["len1"]:rb][::-1])
        if o["right_pair"] >= 0:
            sp = right[o["right_pair"]]
            assert sp["len2"] == 150 - qb - ln
            assert np.array_equal(qer[sp["idq"]:sp["idq"] + sp["len2"]], reads[k, qb + ln:])
    # most reads chain to their origin: the seed's diagonal is near it
    near = np.abs(out["seed_r"].astype(np.int64) - out["seed_q"] - d["origin"]) < 10
    assert near.mean() > 0.9


def test_model_pipeline_right_h0_is_left_score():
    d = _data(30, seed=6)
    out, left, right = M.pipeline(d["genome"], d["reads"], d["read_off"], d["anchors"],
                                  d["anchor_off"])
    for k in range(len(out)):
        rp, lp = out["right_pair"][k], out["left_pair"][k]
        if rp >= 0 and lp >= 0:
            assert right["h0"][rp] == left["score"][lp]
        # an extension never scores below its h0
        if lp >= 0:
            assert left["score"][lp] >= left["h0"][lp]


def test_extend_header_declares_binding_exports():
    text = open(os.path.join(ROOT, "include", "tal_extend.h")).read()
    decl = set(re.findall(r"^\w[\w\s\*]*?\b(tal_extend_\w+)\s*\(", text, re.M))
    assert decl == set(TE.EXPORTS)
    if os.path.exists(TE.LIB_PATH):
        lib = ctypes.CDLL(TE.LIB_PATH)
        for name in TE.EXPORTS:
            assert hasattr(lib, name), name


def test_read_record_layout():
    assert TE.READ_DTYPE.itemsize == 40


def test_extend_header_layouts_match_the_binding(tmp_path):
    """tal_extend.h compiles as C11 with warnings as errors, and its structs
    have the sizes and field offsets that the Python binding assumes."""
    from helpers import check_c_layouts
    check_c_layouts(tmp_path, ROOT, "tal_extend.h", dtypes=[("tal_extend_read", TE.READ_DTYPE)],
                    structs=[("tal_extend_params", TE.Params), ("tal_extend_stats", TE.Stats)])
