"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.  The oracle is a
CPU restatement of TAL's BandedPairWiseSW (src/alignment/BSW/bandedSWA.cpp);
see bsw_oracle.h for what each entry point restates.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# SeqPair, bandedSWA.h:90-99 (14 x int32 = 56 bytes)
SEQPAIR_FIELDS = ["idr", "idq", "id", "len1", "len2", "h0", "seqid", "regid",
                  "score", "tle", "gtle", "qle", "gscore", "max_off"]
SEQPAIR_DTYPE = np.dtype([(f, "<i4") for f in SEQPAIR_FIELDS])
OUT_FIELDS = ["score", "tle", "gtle", "qle", "gscore", "max_off"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i8 = ctypes.c_int32, ctypes.c_int8
        L.bsw_oracle_create.restype = P
        L.bsw_oracle_create.argtypes = [ctypes.c_int] * 6 + [i8, i8]
        L.bsw_oracle_destroy.argtypes = [P]
        for name in ("bsw_oracle_get_scores16", "bsw_oracle_get_scores8",
                     "bsw_oracle_scalar"):
            getattr(L, name).argtypes = [P, P, P, P, i32, i32]
        L.bsw_oracle_scalar_mat.argtypes = [P, P, P, P, P, i32, i32]
        L.bsw_oracle_band_dp.restype = ctypes.c_int
        L.bsw_oracle_band_dp.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P,
                                         i32, ctypes.c_int] + [P] * 5
        L.bsw_oracle_set_stale_mode.argtypes = [P, ctypes.c_int]
        L.bsw_oracle_set_mutation.argtypes = [P, ctypes.c_int]
        L.bsw_oracle_stats.argtypes = [P, P]
        L.bsw_oracle_f16_nonzero_mask.restype = ctypes.c_uint32
        L.bsw_oracle_f16_nonzero_mask.argtypes = [P, ctypes.c_int]
        L.bsw_oracle_band_cells.restype = ctypes.c_uint64
        L.bsw_oracle_band_cells.argtypes = [P, ctypes.c_int, ctypes.c_int, i32]
        L.bsw_simd_run_mt.restype = ctypes.c_int
        L.bsw_simd_run_mt.argtypes = [ctypes.c_int] * 6 + [i8, i8, ctypes.c_int,
                                      P, P, P, i32, i32, i32, ctypes.c_int]
        L.bsw_simd_width.restype = ctypes.c_int
        L.bsw_oracle_run_mt.restype = ctypes.c_int
        L.bsw_oracle_run_mt.argtypes = [ctypes.c_int] * 6 + [i8, i8, ctypes.c_int,
                                        P, P, P, i32, i32, i32, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Oracle:
    """One restated BandedPairWiseSW object (its scratch persists across calls,
    as the reference object's does)."""

    def __init__(self, o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100,
                 end_bonus=5, w_match=1, w_mismatch=4):
        self.params = (o_del, e_del, o_ins, e_ins, zdrop, end_bonus, w_match, w_mismatch)
        self._h = lib().bsw_oracle_create(*self.params)
        if not self._h:
            raise MemoryError("bsw_oracle_create failed")

    def close(self):
        if self._h:
            lib().bsw_oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _run(self, fn, pairs, ref, qer, w):
        assert pairs.dtype == SEQPAIR_DTYPE and pairs.flags.c_contiguous
        ref = np.ascontiguousarray(ref, dtype=np.uint8)
        qer = np.ascontiguousarray(qer, dtype=np.uint8)
        fn(self._h, _ptr(pairs), _ptr(ref), _ptr(qer), len(pairs), int(w))
        return pairs

    def scores16(self, pairs, ref, qer, w):
        return self._run(lib().bsw_oracle_get_scores16, pairs, ref, qer, w)

    def scores8(self, pairs, ref, qer, w):
        return self._run(lib().bsw_oracle_get_scores8, pairs, ref, qer, w)

    def scalar(self, pairs, ref, qer, w):
        return self._run(lib().bsw_oracle_scalar, pairs, ref, qer, w)

    def scalar_mat(self, mat, pairs, ref, qer, w):
        """The scalar routine with a caller-given 5x5 matrix (mat[r*5+q])."""
        m = np.ascontiguousarray(mat, dtype=np.int8).reshape(25)
        assert pairs.dtype == SEQPAIR_DTYPE and pairs.flags.c_contiguous
        ref = np.ascontiguousarray(ref, dtype=np.uint8)
        qer = np.ascontiguousarray(qer, dtype=np.uint8)
        lib().bsw_oracle_scalar_mat(self._h, _ptr(m), _ptr(pairs), _ptr(ref), _ptr(qer),
                                    len(pairs), int(w))
        return pairs

    def band_dp(self, query, target, w, h0):
        q = np.ascontiguousarray(query, dtype=np.uint8)
        t = np.ascontiguousarray(target, dtype=np.uint8)
        outs = [ctypes.c_int() for _ in range(5)]
        s = lib().bsw_oracle_band_dp(self._h, len(q), _ptr(q), len(t), _ptr(t),
                                     int(w), int(h0), *[ctypes.byref(x) for x in outs])
        qle, tle, gtle, gscore, max_off = (x.value for x in outs)
        return dict(score=s, qle=qle, tle=tle, gtle=gtle, gscore=gscore, max_off=max_off)

    def set_stale_mode(self, mode):
        lib().bsw_oracle_set_stale_mode(self._h, int(mode))

    def set_mutation(self, m):
        """Plants a known mistake (tests of the pins only)."""
        lib().bsw_oracle_set_mutation(self._h, int(m))

    def stats(self):
        out = np.zeros(3, dtype=np.uint64)
        lib().bsw_oracle_stats(self._h, _ptr(out))
        return dict(groups=int(out[0]), stale_decisive=int(out[1]), stale_nonzero=int(out[2]))

    def f16_nonzero_mask(self, col):
        return int(lib().bsw_oracle_f16_nonzero_mask(self._h, int(col)))

    def band_cells(self, len1, len2, w):
        return int(lib().bsw_oracle_band_cells(self._h, int(len1), int(len2), int(w)))


def run_mt(pairs, ref, qer, w, bits=16, nthreads=1, chunk=20000, params=None):
    """Each thread owns one oracle object and takes contiguous chunks."""
    p = params or (6, 1, 6, 1, 100, 5, 1, 4)
    ref = np.ascontiguousarray(ref, dtype=np.uint8)
    qer = np.ascontiguousarray(qer, dtype=np.uint8)
    rc = lib().bsw_oracle_run_mt(*p, int(bits), _ptr(pairs), _ptr(ref), _ptr(qer),
                                 len(pairs), int(w), int(chunk), int(nthreads))
    if rc != 0:
        raise MemoryError("bsw_oracle_run_mt failed")
    return pairs


def simd_run_mt(pairs, ref, qer, w, bits=16, nthreads=1, chunk=20000, params=None):
    """The vectorised restatement (bsw_simd.cpp, the CPU baseline): each
    thread owns one object and takes contiguous chunks, as run_mt."""
    p = params or (6, 1, 6, 1, 100, 5, 1, 4)
    ref = np.ascontiguousarray(ref, dtype=np.uint8)
    qer = np.ascontiguousarray(qer, dtype=np.uint8)
    rc = lib().bsw_simd_run_mt(*p, int(bits), _ptr(pairs), _ptr(ref), _ptr(qer),
                               len(pairs), int(w), int(chunk), int(nthreads))
    if rc != 0:
        raise MemoryError("bsw_simd_run_mt failed")
    return pairs


def simd_width():
    """512 if the AVX-512BW build of the vectorised restatement runs here,
    else 256 (AVX2)."""
    return int(lib().bsw_simd_width())


def make_pairs(refs, queries, h0s):
    """Pack python lists of code arrays into (pairs, ref_buf, qer_buf)."""
    n = len(refs)
    pairs = np.zeros(n, dtype=SEQPAIR_DTYPE)
    rlen = np.array([len(r) for r in refs], dtype=np.int64)
    qlen = np.array([len(q) for q in queries], dtype=np.int64)
    roff = np.concatenate([[0], np.cumsum(rlen)[:-1]]) if n else np.zeros(0, np.int64)
    qoff = np.concatenate([[0], np.cumsum(qlen)[:-1]]) if n else np.zeros(0, np.int64)
    ref = np.zeros(max(1, int(rlen.sum())), dtype=np.uint8)
    qer = np.zeros(max(1, int(qlen.sum())), dtype=np.uint8)
    for k in range(n):
        ref[roff[k]:roff[k] + rlen[k]] = refs[k]
        qer[qoff[k]:qoff[k] + qlen[k]] = queries[k]
    pairs["idr"], pairs["idq"], pairs["id"] = roff, qoff, np.arange(n)
    pairs["len1"], pairs["len2"], pairs["h0"] = rlen, qlen, h0s
    pairs["seqid"] = pairs["regid"] = -1
    for f in OUT_FIELDS:
        pairs[f] = -1
    return pairs, ref, qer


# ---------------------------------------------------------------------------
# Full inter-task SW (sw_oracle.c): PairWiseSW_{8,16,32}::getScores.

# SeqPair of PairWiseSW.h:44-49 (12 bytes)
SW_PAIR_DTYPE = np.dtype([("id", "<i4"), ("len1", "<i2"), ("len2", "<i2"), ("score", "<i4")])
SW_GROUP = {8: 64, 16: 32, 32: 16}        # SIMD_WIDTH of the AVX-512BW build
SW_SLOT = {8: 128, 16: 1024, 32: 1024}    # MAX_SEQ_LEN_8/16/32
REF_SW_PATH = os.path.join(HERE, "_ref", "libpwsw_ref.so")


def _sw_lib():
    L = lib()
    if not getattr(L, "_sw_ready", False):
        P, i32 = ctypes.c_void_p, ctypes.c_int32
        L.sw_oracle_get_scores.restype = ctypes.c_int
        L.sw_oracle_get_scores.argtypes = [ctypes.c_int] + [i32] * 4 + [P, P, i32, ctypes.c_int]
        L.sw_oracle_scalar.argtypes = [i32] * 4 + [P, P, i32, i32]
        L.sw_oracle_set_mutation.argtypes = [ctypes.c_int]
        L._sw_ready = True
    return L


def sw_scores(bits, pairs, seq_buf, params=(1, -1, 1, 1), nthreads=1):
    """Oracle getScores of the bits-wide class; scores written into pairs."""
    assert pairs.dtype == SW_PAIR_DTYPE and pairs.flags.c_contiguous
    seq_buf = np.ascontiguousarray(seq_buf, dtype=np.uint8)
    rc = _sw_lib().sw_oracle_get_scores(int(bits), *[int(x) for x in params], _ptr(pairs),
                                        _ptr(seq_buf), len(pairs), int(nthreads))
    if rc != 0:
        raise ValueError(f"bad width {bits}")
    return pairs


def sw_scalar(pairs, seq_buf, params=(1, -1, 1, 1), slot=1024):
    """smithWatermanOnePair on each pair's own lengths."""
    seq_buf = np.ascontiguousarray(seq_buf, dtype=np.uint8)
    _sw_lib().sw_oracle_scalar(*[int(x) for x in params], _ptr(pairs), _ptr(seq_buf),
                               len(pairs), int(slot))
    return pairs


def sw_set_mutation(m):
    """Plants a known mistake in sw_scores (tests of the pins only)."""
    _sw_lib().sw_oracle_set_mutation(int(m))


_ref_sw = None


def ref_sw_available():
    return os.path.exists(REF_SW_PATH)


def ref_sw_scores(bits, pairs, seq_buf, params=(1, -1, 1, 1), nthreads=1):
    """The reference's own getScores (oracle/_ref, built from its sources by
    `make -C oracle ref`).  Returns None when the CPU lacks AVX-512BW."""
    global _ref_sw
    if _ref_sw is None:
        _ref_sw = ctypes.CDLL(REF_SW_PATH)
        _ref_sw.ref_sw_get_scores.restype = ctypes.c_int
        _ref_sw.ref_sw_get_scores.argtypes = [ctypes.c_int] + [ctypes.c_int32] * 4 + [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int]
    n = len(pairs)
    g = SW_GROUP[bits]
    # getScores writes the padding entries of the last group
    buf = np.zeros(((n + g - 1) // g) * g, SW_PAIR_DTYPE)
    buf[:n] = pairs
    seq_buf = np.ascontiguousarray(seq_buf, dtype=np.uint8)  # read only by getScores
    rc = _ref_sw.ref_sw_get_scores(int(bits), *[int(x) for x in params], _ptr(buf),
                                   _ptr(seq_buf), n, int(nthreads))
    if rc == -2:
        return None
    if rc != 0:
        raise ValueError(f"bad width {bits}")
    pairs["score"] = buf["score"][:n]
    return pairs


def sw_make_batch(seq1s, seq2s, bits):
    """(pairs, seq_buf) in the reference's layout: pair k's sequences at
    seq_buf + 2*k*S and + (2k+1)*S (S = MAX_SEQ_LEN of the class)."""
    S = SW_SLOT[bits]
    n = len(seq1s)
    pairs = np.zeros(n, SW_PAIR_DTYPE)
    buf = np.zeros(max(1, 2 * n * S), np.uint8)
    for k in range(n):
        a = np.asarray(seq1s[k], np.uint8)
        b = np.asarray(seq2s[k], np.uint8)
        buf[2 * k * S:2 * k * S + len(a)] = a
        buf[(2 * k + 1) * S:(2 * k + 1) * S + len(b)] = b
    pairs["id"] = np.arange(n)
    pairs["len1"] = [len(a) for a in seq1s]
    pairs["len2"] = [len(b) for b in seq2s]
    return pairs, buf


# ---------------------------------------------------------------------------
# DP chaining (chain_oracle.c): dp_chain::mm_dp_vectorized.

# anchor_t of parallel_chaining_32_bit.h:56-65 (16 bytes)
ANCHOR_DTYPE = np.dtype([("r", "<u8"), ("q", "<i4"), ("l", "<i4")])
REF_CHAIN_PATH = os.path.join(HERE, "_ref", "libchain_ref.so")
# bench-dp-chaining's object and call (bench-dp-chaining.cpp:69,101)
CHAIN_DEFAULTS = dict(bw=500, max_iter=5000, dr=5000, dq=5000)


def _chain_lib():
    L = lib()
    if not getattr(L, "_chain_ready", False):
        P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.chain_oracle_dp.argtypes = [P, i64, i32, i32, i32, i32, P, P, P]
        L.chain_oracle_dp_sets.argtypes = [P, P, i32, i32, i32, i32, i32, P, P, P, ctypes.c_int]
        L.chain_oracle_set_mutation.argtypes = [ctypes.c_int]
        L._chain_ready = True
    return L


def chain_dp_sets(anchors, offsets, bw=500, max_iter=5000, dr=5000, dq=5000, nthreads=1):
    """Oracle chaining of every set; returns (f uint32, p int32, v int32)."""
    anchors = np.ascontiguousarray(anchors, ANCHOR_DTYPE)
    offsets = np.ascontiguousarray(offsets, np.int64)
    n = len(anchors)
    f = np.zeros(max(n, 1), np.uint32)
    p = np.zeros(max(n, 1), np.int32)
    v = np.zeros(max(n, 1), np.int32)
    _chain_lib().chain_oracle_dp_sets(_ptr(anchors), _ptr(offsets), len(offsets) - 1, bw,
                                      max_iter, dr, dq, _ptr(f), _ptr(p), _ptr(v), nthreads)
    return f[:n], p[:n], v[:n]


def chain_set_mutation(m):
    _chain_lib().chain_oracle_set_mutation(int(m))


_ref_chain = None


def ref_chain_available():
    return os.path.exists(REF_CHAIN_PATH)


def ref_chain_dp_sets(anchors, offsets, bw=500, max_iter=5000, dr=5000, dq=5000):
    """The reference's own mm_dp_vectorized (oracle/_ref, built from its
    header by `make -C oracle ref`), set by set.  None when the CPU lacks
    AVX-512BW/CD."""
    global _ref_chain
    if _ref_chain is None:
        _ref_chain = ctypes.CDLL(REF_CHAIN_PATH)
        P, i32 = ctypes.c_void_p, ctypes.c_int32
        _ref_chain.ref_chain_dp.restype = ctypes.c_int
        _ref_chain.ref_chain_dp.argtypes = [P, P, P, ctypes.c_int64, i32, i32, i32, i32, P, P, P]
    n = len(anchors)
    r = np.ascontiguousarray(anchors["r"], np.uint64)
    q = np.ascontiguousarray(anchors["q"], np.int32)
    ln = np.ascontiguousarray(anchors["l"], np.int32)
    f = np.zeros(max(n, 1), np.uint32)
    p = np.zeros(max(n, 1), np.int32)
    v = np.zeros(max(n, 1), np.int32)
    for k in range(len(offsets) - 1):
        a, b = int(offsets[k]), int(offsets[k + 1])
        if b == a:
            continue
        rc = _ref_chain.ref_chain_dp(_ptr(r[a:b]), _ptr(q[a:b]), _ptr(ln[a:b]), b - a, bw,
                                     max_iter, dr, dq, _ptr(f[a:b]), _ptr(p[a:b]), _ptr(v[a:b]))
        if rc == -2:
            return None
    return f[:n], p[:n], v[:n]
