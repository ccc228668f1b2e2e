"""The host side of the GPU library under AddressSanitizer and UBSan:
_asan/tal-bench-bsw links _asan/libtal_bsw.so, whose host code (C ABI,
validation, dispatcher, plans, file I/O) is sanitized while its device code
is the ordinary build.  The sanitized command line scores a pair file in its
16-bit, dispatcher and scalar modes with the same outputs as the ordinary
library, refuses a malformed file with an error, and no sanitizer fires.
Leak checking is off: the HIP runtime keeps allocations until exit."""
import os
import subprocess

import numpy as np
import pytest

import synth as S
import tal_bsw as T

pytestmark = pytest.mark.gpu

ASAN_CLI = os.path.join(os.path.dirname(T.LIB_PATH), "_asan", "tal-bench-bsw")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _maxi(text):
    rows = []
    for line in text.splitlines():
        a, rest = line.split(" (")
        tq, rest = rest.split(") ")
        tle, qle = tq.split()
        gscore, max_off, gtle = rest.split()
        rows.append((int(a), int(tle), int(qle), int(gscore), int(max_off), int(gtle)))
    return np.array(rows, np.int64).reshape(-1, 6)


def _want(pairs):
    return np.stack([pairs[f].astype(np.int64)
                     for f in ("score", "tle", "qle", "gscore", "max_off", "gtle")], 1)


def _clean(out):
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]


def _run(args):
    print(f"[gpu] sanitized tal-bench-bsw {' '.join(args)}", flush=True)
    return subprocess.run([ASAN_CLI, *args], capture_output=True, text=True, timeout=300,
                          env=ENV)


@pytest.mark.parametrize("mode", ["16", "auto", "scalar"])
def test_sanitized_cli_scores_like_the_library(tmp_path, mode):
    assert os.path.exists(ASAN_CLI), "build() makes the sanitized command line"
    pairs, ref, qer, _ = S.generate(2, 9000)
    keep = np.where((pairs["len1"] <= 254) & (pairs["len2"] <= 126))[0][:2500]
    pairs = np.ascontiguousarray(pairs[keep])
    path = tmp_path / "pairs.txt"
    T.write_pairfile(path, pairs, ref, qer)
    (cp, cr, cq), = list(T.read_pairfile(path))
    eng = T.BandedPairWiseSW()
    try:
        if mode == "16":
            want = eng.get_scores16(cp.copy(), cr, cq, 100)
        elif mode == "auto":
            want = eng.get_scores_auto(cp.copy(), cr, cq, 100)
        else:
            want = eng.scalar_scores(cp.copy(), cr, cq, 100)
    finally:
        eng.close()
    flags = {"16": [], "auto": ["-auto"], "scalar": ["-scalar"]}[mode]
    r = _run(["-pairs", str(path), "-maxi", *flags])
    _clean(r.stdout + r.stderr)
    assert r.returncode == 0, r.stderr[-3000:]
    assert np.array_equal(_maxi(r.stderr), _want(want))


def test_sanitized_cli_refuses_a_truncated_file(tmp_path):
    assert os.path.exists(ASAN_CLI), "build() makes the sanitized command line"
    path = tmp_path / "bad.txt"
    path.write_text("12\n0123012301\n3210\n7\n012301\n")  # a reference with no query
    r = _run(["-pairs", str(path), "-maxi"])
    _clean(r.stdout + r.stderr)
    assert r.returncode != 0
