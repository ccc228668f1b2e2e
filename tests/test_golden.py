"""Golden fixtures (tests/golden/bsw_golden.{json,npz}).

CPU: the seeded generator still produces the recorded inputs, and the oracle
still produces the recorded outputs.  GPU: the engine produces them, through
the C ABI, each case on a fresh handle.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as G  # noqa: E402

with open(os.path.join(HERE, "golden", "bsw_golden.json")) as _f:
    MANIFEST = json.load(_f)
CASES = {c["name"]: c for c in MANIFEST["cases"]}


def _fixture(name):
    with np.load(os.path.join(HERE, "golden", "bsw_golden.npz")) as z:
        return z[name]


def _inputs(c):
    pairs, ref, qer = G.case_inputs(c["workload"], c["prefix"], c["select"])
    assert len(pairs) == c["pairs"]
    return pairs, ref, qer


def test_fixture_files_consistent():
    for name, c in CASES.items():
        out = _fixture(name)
        assert out.shape == (c["pairs"], 6) and out.dtype == np.dtype("<i4")
        assert hashlib.sha256(out.tobytes()).hexdigest() == c["outputs_sha256"], name


@pytest.mark.parametrize("name", list(CASES))
def test_generator_reproduces_inputs(name):
    c = CASES[name]
    pairs, ref, qer = _inputs(c)
    assert G.input_digest(pairs, ref, qer) == c["inputs_sha256"]


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_outputs(name):
    c = CASES[name]
    pairs, ref, qer = _inputs(c)
    got = G.oracle_outputs(pairs, ref, qer, c["bits"], c["w"])
    np.testing.assert_array_equal(got, _fixture(name))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_engine_matches_fixture(name):
    import tal_bsw as T
    c = CASES[name]
    print(f"[gpu] golden {name} n={c['pairs']}", flush=True)
    pairs, ref, qer = _inputs(c)
    eng = T.BandedPairWiseSW(**MANIFEST["params"])
    try:
        fn = eng.get_scores16 if c["bits"] == 16 else eng.get_scores8
        got = G.output_matrix(fn(pairs.copy(), ref, qer, c["w"]))
    finally:
        eng.close()
    want = _fixture(name)
    bad = np.where((got != want).any(axis=1))[0]
    assert bad.size == 0, f"{name}: {bad.size} pairs differ, first {bad[:5].tolist()}"
