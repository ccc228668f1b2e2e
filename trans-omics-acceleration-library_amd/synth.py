"""Deterministic synthetic extension-pair workloads (ctypes over
libtal_synth.so, built from csrc/bsw_synth.c).  See csrc/bsw_synth.h for the
workload definitions."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtal_synth.so")

SEQPAIR_FIELDS = ["idr", "idq", "id", "len1", "len2", "h0", "seqid", "regid",
                  "score", "tle", "gtle", "qle", "gscore", "max_off"]
SEQPAIR_DTYPE = np.dtype([(f, "<i4") for f in SEQPAIR_FIELDS])

SEEDS = {1: 0xB5A0000000000001, 2: 0xB5A0000000000002, 3: 0xB5A0000000000003,
         4: 0xB5A0000000000004, 5: 0xB5A0000000000005}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.tal_synth_shape.restype = ctypes.c_int
        L.tal_synth_shape.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int64,
                                      ctypes.c_int32, P, P, P, P]
        L.tal_synth_fill.restype = ctypes.c_int
        L.tal_synth_fill.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int64,
                                     ctypes.c_int32, P, P, P, P, P, ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def shape(cfg, n, first=0, seed=None):
    seed = SEEDS[cfg] if seed is None else seed
    len1 = np.empty(n, np.int32)
    len2 = np.empty(n, np.int32)
    h0 = np.empty(n, np.int32)
    w = np.empty(n, np.int32)
    if lib().tal_synth_shape(cfg, seed, first, n, _p(len1), _p(len2), _p(h0), _p(w)):
        raise ValueError(f"unknown workload {cfg}")
    return len1, len2, h0, w


def generate(cfg, n, first=0, seed=None, nthreads=8):
    """Returns (pairs, ref, qer, w) for pairs [first, first+n) of workload cfg."""
    seed = SEEDS[cfg] if seed is None else seed
    len1, len2, h0, w = shape(cfg, n, first, seed)
    roff = np.zeros(n, np.int64)
    qoff = np.zeros(n, np.int64)
    if n:
        roff[1:] = np.cumsum(len1.astype(np.int64))[:-1]
        qoff[1:] = np.cumsum(len2.astype(np.int64))[:-1]
    ref = np.zeros(max(1, int(len1.sum())), np.uint8)
    qer = np.zeros(max(1, int(len2.sum())), np.uint8)
    if int(len1.sum()) >= 2**31 or int(len2.sum()) >= 2**31:
        raise ValueError("sequence buffers exceed the int32 offsets of SeqPair")
    pairs = np.zeros(n, SEQPAIR_DTYPE)
    if lib().tal_synth_fill(cfg, seed, first, n, _p(roff), _p(qoff), _p(pairs),
                            _p(ref), _p(qer), int(nthreads)):
        raise ValueError(f"unknown workload {cfg}")
    return pairs, ref, qer, w
