"""Python binding of libtal_extend.so (the C ABI in include/tal_extend.h):
the extension-pair producer and the end-to-end extension pipeline (anchors
-> DP chaining -> best chain -> left/right extension pairs -> banded SW of
both sides).  The GPU library must be built (build.py); there is no CPU
fallback."""
import ctypes
import os

import numpy as np

import tal_bsw
import tal_chain

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtal_extend.so")

READ_DTYPE = np.dtype([("seed_r", "<u8"), ("seed_q", "<i4"), ("seed_len", "<i4"),
                       ("chain_score", "<i4"), ("chain_anchors", "<i4"), ("left_pair", "<i4"),
                       ("right_pair", "<i4"), ("left_score", "<i4"), ("score", "<i4")])

TAL_OK, TAL_EINVAL, TAL_ENOMEM, TAL_EDEVICE, TAL_ERANGE = 0, 1, 2, 3, 4

# every C-ABI entry point of include/tal_extend.h
EXPORTS = ["tal_extend_create", "tal_extend_destroy", "tal_extend_last_error",
           "tal_extend_run_device", "tal_extend_run", "tal_extend_produce_device",
           "tal_extend_get_pairs", "tal_extend_get_stats"]


class Params(ctypes.Structure):
    _fields_ = [("pad", ctypes.c_int32), ("w", ctypes.c_int32), ("dr", ctypes.c_int32),
                ("dq", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("ms_chain", ctypes.c_float), ("ms_produce", ctypes.c_float),
                ("ms_left", ctypes.c_float), ("ms_right", ctypes.c_float),
                ("ms_total", ctypes.c_float), ("reads", ctypes.c_int32),
                ("left_pairs", ctypes.c_int32), ("right_pairs", ctypes.c_int32),
                ("ref_bytes", ctypes.c_int64), ("qer_bytes", ctypes.c_int64)]


class TalError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{code}: {msg}")
        self.code = code


_lib = None


def load():
    """Loads libtal_extend.so; raises if it is not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.tal_extend_create.restype = ctypes.c_int
        L.tal_extend_create.argtypes = [ctypes.POINTER(Params), P, P, ctypes.c_int,
                                        ctypes.POINTER(P)]
        L.tal_extend_destroy.argtypes = [P]
        L.tal_extend_last_error.restype = ctypes.c_char_p
        L.tal_extend_last_error.argtypes = [P]
        L.tal_extend_run_device.restype = ctypes.c_int
        L.tal_extend_run_device.argtypes = [P, P, i64, P, P, i32, P, P, P, P]
        L.tal_extend_run.restype = ctypes.c_int
        L.tal_extend_run.argtypes = [P, P, i64, P, P, i32, P, P, P]
        L.tal_extend_produce_device.restype = ctypes.c_int
        L.tal_extend_produce_device.argtypes = [P, P, i64, P, P, i32, P, P, P, P, P, P]
        L.tal_extend_get_pairs.restype = ctypes.c_int
        L.tal_extend_get_pairs.argtypes = [P, P, P, P, P]
        L.tal_extend_get_stats.restype = ctypes.c_int
        L.tal_extend_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Extender:
    """One pipeline handle.  Chaining parameters default to bench-dp-chaining's
    and extension ones to bench-bsw's."""

    def __init__(self, pad=100, w=100, dr=5000, dq=5000, chain=None, bsw=None, device=0):
        L = load()
        self.params = Params(pad, w, dr, dq)
        c = chain or dict(max_dist_x=5000, max_dist_y=5000, bw=500, max_skip=25, max_iter=5000,
                          gap_scale=1.0, is_cdna=0, n_segs=1)
        b = bsw or dict(o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100, end_bonus=5, w_match=1,
                        w_mismatch=4)
        self.chain_params = tal_chain.Params(**c)
        self.bsw_params = tal_bsw.Params(**b)
        h = ctypes.c_void_p()
        rc = L.tal_extend_create(ctypes.byref(self.params), ctypes.byref(self.chain_params),
                                 ctypes.byref(self.bsw_params), int(device), ctypes.byref(h))
        if rc != TAL_OK:
            raise TalError(rc, "tal_extend_create failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().tal_extend_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != TAL_OK:
            raise TalError(rc, load().tal_extend_last_error(self._h).decode())

    def run(self, genome, reads, read_off, anchors, anchor_off):
        """The whole pipeline on host arrays; returns READ_DTYPE[nreads]."""
        genome = np.ascontiguousarray(genome, np.uint8)
        reads = np.ascontiguousarray(reads, np.uint8)
        read_off = np.ascontiguousarray(read_off, np.int64)
        anchors = np.ascontiguousarray(anchors, tal_chain.ANCHOR_DTYPE)
        anchor_off = np.ascontiguousarray(anchor_off, np.int64)
        nreads = len(read_off) - 1
        out = np.zeros(max(nreads, 1), READ_DTYPE)
        self._check(load().tal_extend_run(self._h, _ptr(genome), len(genome), _ptr(reads),
                                          _ptr(read_off), nreads, _ptr(anchors),
                                          _ptr(anchor_off), _ptr(out)))
        return out[:nreads]

    def run_device(self, d_genome, genome_len, d_reads, d_read_off, nreads, d_anchors,
                   d_anchor_off, d_out, stream=0):
        c = ctypes.c_void_p
        self._check(load().tal_extend_run_device(self._h, c(d_genome), int(genome_len),
                                                 c(d_reads), c(d_read_off), int(nreads),
                                                 c(d_anchors), c(d_anchor_off), c(d_out),
                                                 c(stream)))

    def produce_device(self, d_genome, genome_len, d_reads, d_read_off, nreads, d_anchors,
                       d_anchor_off, d_f, d_p, d_out, stream=0):
        """The producer alone on chaining results already in device memory
        (f, p); the pairs and buffers then come from pairs()."""
        c = ctypes.c_void_p
        self._check(load().tal_extend_produce_device(self._h, c(d_genome), int(genome_len),
                                                     c(d_reads), c(d_read_off), int(nreads),
                                                     c(d_anchors), c(d_anchor_off), c(d_f),
                                                     c(d_p), c(d_out), c(stream)))

    def pairs(self):
        """(left pairs, right pairs, ref bytes, query bytes) of the last call."""
        st = self.stats()
        left = np.zeros(max(st["left_pairs"], 1), tal_bsw.SEQPAIR_DTYPE)
        right = np.zeros(max(st["right_pairs"], 1), tal_bsw.SEQPAIR_DTYPE)
        ref = np.zeros(max(st["ref_bytes"], 1), np.uint8)
        qer = np.zeros(max(st["qer_bytes"], 1), np.uint8)
        self._check(load().tal_extend_get_pairs(self._h, _ptr(left), _ptr(right), _ptr(ref),
                                                _ptr(qer)))
        return (left[:st["left_pairs"]], right[:st["right_pairs"]], ref[:st["ref_bytes"]],
                qer[:st["qer_bytes"]])

    def stats(self):
        st = Stats()
        self._check(load().tal_extend_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in Stats._fields_}
