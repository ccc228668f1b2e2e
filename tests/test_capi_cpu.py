"""C ABI checks that need no GPU: the library loads, exports every entry
point include/tal_bsw.h declares, the SeqPair layout is the reference's, and
argument errors are reported without touching a device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import tal_bsw as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tal_bsw.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tal_bsw_\w+)\s*\(", text)))


def test_header_declares_binding_exports():
    assert _declared() == sorted(T.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(T.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_sw_header_declares_binding_exports():
    import tal_sw
    text = open(os.path.join(ROOT, "include", "tal_sw.h")).read()
    declared = sorted(set(re.findall(r"\b(tal_sw_\w+)\s*\(", text)))
    assert declared == sorted(tal_sw.EXPORTS)
    lib = ctypes.CDLL(tal_sw.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_sw_pair_layout(tmp_path):
    src = tmp_path / "swlayout.c"
    src.write_text("""
#include <stddef.h>
#include "tal_sw.h"
_Static_assert(sizeof(tal_sw_pair) == 12, "size");
_Static_assert(offsetof(tal_sw_pair, len1) == 4, "len1");
_Static_assert(offsetof(tal_sw_pair, len2) == 6, "len2");
_Static_assert(offsetof(tal_sw_pair, score) == 8, "score");
int main(void) { return 0; }
""")
    subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                           "-o", str(tmp_path / "swlayout.o")])
    import tal_sw
    assert tal_sw.PAIR_DTYPE.itemsize == 12


def test_seqpair_layout(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stddef.h>
#include "tal_bsw.h"
_Static_assert(sizeof(tal_seqpair) == 56, "size");
_Static_assert(offsetof(tal_seqpair, h0) == 20, "h0");
_Static_assert(offsetof(tal_seqpair, score) == 32, "score");
_Static_assert(offsetof(tal_seqpair, max_off) == 52, "max_off");
_Static_assert(sizeof(tal_bsw_params) == 32, "params");
int main(void) { return 0; }
""")
    subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                           "-o", str(tmp_path / "layout.o")])
    assert T.SEQPAIR_DTYPE.itemsize == 56


def test_cpp_shim_compiles(tmp_path):
    src = tmp_path / "shim.cpp"
    src.write_text('#include "bandedSWA_tal.hpp"\nint main(){return 0;}\n')
    subprocess.check_call(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                           str(src)])


def test_bsw_shim_compiles_reference_call_pattern(tmp_path):
    """bench-bsw's calls (benchmarks/bench-bsw.cpp:213-218, 253, 256-258),
    both branches, and scalarBandedSWA's signature (bandedSWA.h:124-129)
    compile against the shim unchanged."""
    src = tmp_path / "bsw_caller.cpp"
    src.write_text("""
#include "bandedSWA_tal.hpp"
void bwa_fill_scmat(int a, int b, int ambig, int8_t mat[25]);
void run(SeqPair *seqPairArray, uint8_t *seqBufRef, uint8_t *seqBufQer, int numPairs) {
    int8_t mat[25];
    bwa_fill_scmat(1, 4, -1, mat);
    int zdrop = 100, w = 100, end_bonus = 5, w_open = 6, w_extend = 1;
    BandedPairWiseSW *bsw = new BandedPairWiseSW(w_open, w_extend, w_open, w_extend,
                                                 zdrop, end_bonus, mat, 1, 4, 1);
    bsw->getScores16(seqPairArray, seqBufRef, seqBufQer, numPairs, 1, w);
    bsw->getScores8(seqPairArray, seqBufRef, seqBufQer, numPairs, 1, w);
    bsw->scalarBandedSWAWrapper(seqPairArray, seqBufRef, seqBufQer, numPairs, 1, w);
    SeqPair *p = seqPairArray;
    int qle, tle, gtle, gscore, max_off;
    int score = bsw->scalarBandedSWA(p->len2, seqBufQer + p->idq, p->len1, seqBufRef + p->idr,
                                     w, p->h0, &qle, &tle, &gtle, &gscore, &max_off);
    (void)score;
    delete bsw;
}
""")
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Wshadow", "-Werror", "-fsyntax-only",
                           "-I", os.path.join(ROOT, "include"), str(src)])


def test_pairwisesw_shim_compiles_reference_call_pattern(tmp_path):
    # the calls bench-smith-waterman.cpp:203-223 makes, unchanged
    src = tmp_path / "pwsw.cpp"
    src.write_text("""
#include "PairWiseSW_tal.hpp"
int run(SeqPair *seqPairArray, uint8_t *seqBuf, int64_t numPairs, int32_t numThreads) {
    CycleBreakup ticks;
    PairWiseSW_32 *pwsw32 = new PairWiseSW_32(1, -1, 1, 1);
    pwsw32->getScores(seqPairArray, seqBuf, numPairs, numThreads);
    ticks = pwsw32->getTicks();
    PairWiseSW_16 *pwsw16 = new PairWiseSW_16(1, -1, 1, 1);
    pwsw16->getScores(seqPairArray, seqBuf, numPairs, numThreads);
    PairWiseSW_8 *pwsw8 = new PairWiseSW_8(1, -1, 1, 1);
    pwsw8->getScores(seqPairArray, seqBuf, numPairs, numThreads);
    delete pwsw32; delete pwsw16; delete pwsw8;
    return (int)(ticks.swTicks + MAX_SEQ_LEN_8 + MAX_SEQ_LEN_16 + MAX_SEQ_LEN_32);
}
static_assert(sizeof(SeqPair) == 12, "SeqPair of PairWiseSW.h:44-49");
int main() { return 0; }
""")
    subprocess.check_call(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                           str(src)])


def test_sw_arguments_without_gpu():
    import tal_sw
    L = tal_sw.load()
    assert [L.tal_sw_slot(b) for b in (8, 16, 32, 12)] == [128, 1024, 1024, 0]
    assert L.tal_sw_get_scores(None, 16, None, None, 0, 0) == tal_sw.TAL_EINVAL
    assert L.tal_sw_last_error(None) == b"null handle"
    cli = os.path.join(os.path.dirname(tal_sw.LIB_PATH), "tal-bench-sw")
    r = subprocess.run([cli], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "pairFileName not specified" in r.stdout


def test_chain_header_and_shim():
    import tal_chain
    text = open(os.path.join(ROOT, "include", "tal_chain.h")).read()
    declared = sorted(set(re.findall(r"\b(tal_chain_\w+)\s*\(", text)))
    assert declared == sorted(tal_chain.EXPORTS)
    lib = ctypes.CDLL(tal_chain.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert tal_chain.ANCHOR_DTYPE.itemsize == 16
    L = tal_chain.load()
    assert L.tal_chain_dp(None, None, None, 0, 0, 0, None, None, None) == tal_chain.TAL_EINVAL


def test_chain_shim_compiles_reference_call_pattern(tmp_path):
    # the calls bench-dp-chaining.cpp:69-101 makes, unchanged
    src = tmp_path / "chain.cpp"
    src.write_text("""
#include <vector>
#include "parallel_chaining_tal.hpp"
int run(std::vector<anchor_t> &anc) {
    dp_chain obj(5000,5000,500,25,5000,1.000000,0,1);
    int32_t n = anc.size();
    num_bits_t *f = (num_bits_t*)malloc(sizeof(num_bits_t)*(n+16)); f = f + 16;
    int32_t *p = (int32_t*)malloc(sizeof(num_bits_t)*(n+16)); p = p + 16;
    int32_t *v = (int32_t*)malloc(sizeof(num_bits_t)*(n+16)); v = v + 16;
    num_bits_t *anchor_r, *anchor_q, *anchor_l;
    create_SoA_Anchors_32_bit(&anc[0], anc.size(), anchor_r, anchor_q, anchor_l);
    obj.mm_dp_vectorized(anc.size(), &anc[0], anchor_r, anchor_q, anchor_l, f, p, v, 5000, 5000, NULL, NULL);
    return (int)f[0] + p[0] + v[0];
}
int main() { return 0; }
""")
    subprocess.check_call(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                           str(src)])


def test_create_rejects_bad_params():
    L = T.load()
    h = ctypes.c_void_p()
    assert L.tal_bsw_create(None, 0, ctypes.byref(h)) == T.TAL_EINVAL
    p = T.Params(6, 0, 6, 1, 100, 5, 1, 4)  # e_del = 0 would divide by zero
    assert L.tal_bsw_create(ctypes.byref(p), 0, ctypes.byref(h)) == T.TAL_EINVAL
    p = T.Params(6, 1, 6, 1, 100, 5, 300, 4)  # w_match must fit int8
    assert L.tal_bsw_create(ctypes.byref(p), 0, ctypes.byref(h)) == T.TAL_EINVAL
    assert not h.value


def test_null_handle_calls():
    L = T.load()
    assert L.tal_bsw_count_8bit(None, None, 0) == -1
    assert L.tal_bsw_get_scores16(None, None, None, 0, None, 0, 0, 100) == T.TAL_EINVAL
    assert L.tal_bsw_last_error(None) == b"null handle"


def test_header_layouts_match_the_bindings(tmp_path):
    """Every struct that a Python binding mirrors has, in its C header, the
    binding's size and field offsets (the headers compile as C11 with
    warnings as errors)."""
    import tal_chain
    import tal_sw
    from helpers import check_c_layouts
    check_c_layouts(tmp_path, ROOT, "tal_bsw.h", dtypes=[("tal_seqpair", T.SEQPAIR_DTYPE)],
                    structs=[("tal_bsw_params", T.Params), ("tal_bsw_stats", T.Stats),
                             ("tal_bsw_tuning", T.Tuning),
                             ("tal_bsw_server_config", T.ServerConfig),
                             ("tal_bsw_request_info", T.RequestInfo),
                             ("tal_bsw_server_stats", T.ServerStats), ("tal_bsw_pack", T.Pack)])
    check_c_layouts(tmp_path, ROOT, "tal_sw.h", dtypes=[("tal_sw_pair", tal_sw.PAIR_DTYPE)],
                    structs=[("tal_sw_stats", tal_sw.Stats)])
    check_c_layouts(tmp_path, ROOT, "tal_chain.h",
                    dtypes=[("tal_chain_anchor", tal_chain.ANCHOR_DTYPE)],
                    structs=[("tal_chain_params", tal_chain.Params),
                             ("tal_chain_stats", tal_chain.Stats)])


@pytest.mark.parametrize("modname", ["tal_bsw", "tal_sw", "tal_chain", "tal_extend"])
def test_missing_library_fails_loudly(modname, monkeypatch, tmp_path):
    """Without its built library a binding raises; there is no fallback."""
    import importlib
    mod = importlib.import_module(modname)
    monkeypatch.setattr(mod, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(mod, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        mod.load()
