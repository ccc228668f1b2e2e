"""Python binding of libtal_sw.so (the C ABI in include/tal_sw.h): the full
inter-task Smith-Waterman engine.

Mirrors the reference's PairWiseSW_{8,16,32} interface
(src/alignment/Smith-Waterman/PairWiseSW_16.h:44-72): construct with the four
weights, then call get_scores(bits, pairs, seq_buf) on a SeqPair array and the
sequence buffer; scores are written in place.  The GPU library must be built
(build.py); there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtal_sw.so")

# SeqPair of PairWiseSW.h:44-49 (12 bytes)
PAIR_DTYPE = np.dtype([("id", "<i4"), ("len1", "<i2"), ("len2", "<i2"), ("score", "<i4")])
GROUP = {8: 64, 16: 32, 32: 16}        # pairs per group (SIMD_WIDTH of the AVX-512BW build)
SLOT = {8: 128, 16: 1024, 32: 1024}    # MAX_SEQ_LEN_8/16/32
MAX_LEN = {8: 127, 16: 1024, 32: 1024}

TAL_OK, TAL_EINVAL, TAL_ENOMEM, TAL_EDEVICE, TAL_ERANGE = 0, 1, 2, 3, 4
_STATUS = {1: "TAL_EINVAL", 2: "TAL_ENOMEM", 3: "TAL_EDEVICE", 4: "TAL_ERANGE", 5: "TAL_EAGAIN"}

# every C-ABI entry point of include/tal_sw.h
EXPORTS = ["tal_sw_slot", "tal_sw_create", "tal_sw_destroy", "tal_sw_last_error",
           "tal_sw_get_scores", "tal_sw_get_scores_device", "tal_sw_get_stats"]


class Stats(ctypes.Structure):
    _fields_ = [("ms_kernel", ctypes.c_float), ("ms_total", ctypes.c_float),
                ("units", ctypes.c_int32), ("max_rows", ctypes.c_int32),
                ("max_cols", ctypes.c_int32), ("regrouped", ctypes.c_int32),
                ("split", ctypes.c_int32)]


class TalError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Loads libtal_sw.so; raises if it is not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, i32, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_size_t
        L.tal_sw_slot.restype = i32
        L.tal_sw_slot.argtypes = [ctypes.c_int]
        L.tal_sw_create.restype = ctypes.c_int
        L.tal_sw_create.argtypes = [i32, i32, i32, i32, ctypes.c_int, ctypes.POINTER(P)]
        L.tal_sw_destroy.argtypes = [P]
        L.tal_sw_last_error.restype = ctypes.c_char_p
        L.tal_sw_last_error.argtypes = [P]
        L.tal_sw_get_scores.restype = ctypes.c_int
        L.tal_sw_get_scores.argtypes = [P, ctypes.c_int, P, P, sz, i32]
        L.tal_sw_get_scores_device.restype = ctypes.c_int
        L.tal_sw_get_scores_device.argtypes = [P, ctypes.c_int, P, P, sz, i32, P]
        L.tal_sw_get_stats.restype = ctypes.c_int
        L.tal_sw_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class PairWiseSW:
    """One engine handle on one GPU (the reference's PairWiseSW_{8,16,32}
    objects in one: the class is chosen per call by `bits`)."""

    def __init__(self, w_match=1, w_mismatch=-1, w_open=1, w_extend=1, device=0):
        L = load()
        h = ctypes.c_void_p()
        rc = L.tal_sw_create(int(w_match), int(w_mismatch), int(w_open), int(w_extend),
                             int(device), ctypes.byref(h))
        if rc != TAL_OK:
            raise TalError(rc, "tal_sw_create failed")
        self._h = h
        self.params = (w_match, w_mismatch, w_open, w_extend)

    def close(self):
        if getattr(self, "_h", None):
            load().tal_sw_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != TAL_OK:
            raise TalError(rc, load().tal_sw_last_error(self._h).decode())

    def get_scores(self, bits, pairs, seq_buf):
        """getScores of the bits-wide class on host buffers; scores in place."""
        assert pairs.dtype == PAIR_DTYPE and pairs.flags.c_contiguous
        seq_buf = np.ascontiguousarray(seq_buf, dtype=np.uint8)
        self._check(load().tal_sw_get_scores(self._h, int(bits), _ptr(pairs), _ptr(seq_buf),
                                             seq_buf.nbytes, len(pairs)))
        return pairs

    def get_scores_device(self, bits, d_pairs, d_seq, seq_len, n, stream=0):
        self._check(load().tal_sw_get_scores_device(self._h, int(bits), ctypes.c_void_p(d_pairs),
                                                    ctypes.c_void_p(d_seq), int(seq_len), int(n),
                                                    ctypes.c_void_p(stream)))

    def stats(self):
        st = Stats()
        self._check(load().tal_sw_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in Stats._fields_}


def make_batch(seq1s, seq2s, bits):
    """(pairs, seq_buf) in the reference's layout: pair k's sequences at
    seq_buf + 2*k*S and seq_buf + (2k+1)*S, S = SLOT[bits]."""
    S = SLOT[bits]
    n = len(seq1s)
    pairs = np.zeros(n, PAIR_DTYPE)
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


def cells(pairs):
    """Work unit of the GCUPS figures: len1 * len2 per pair, the convention
    of bench-smith-waterman.cpp:150."""
    return pairs["len1"].astype(np.int64) * pairs["len2"].astype(np.int64)
