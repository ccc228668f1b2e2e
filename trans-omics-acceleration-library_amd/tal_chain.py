"""Python binding of libtal_chain.so (the C ABI in include/tal_chain.h): the
DP chaining engine, a drop-in for dp_chain::mm_dp_vectorized
(src/chaining/parallel_chaining_32_bit.h:216-412) over many anchor sets.
The GPU library must be built (build.py); there is no CPU fallback."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtal_chain.so")

# anchor_t of parallel_chaining_32_bit.h:56-65 (16 bytes)
ANCHOR_DTYPE = np.dtype([("r", "<u8"), ("q", "<i4"), ("l", "<i4")])

TAL_OK, TAL_EINVAL, TAL_ENOMEM, TAL_EDEVICE, TAL_ERANGE = 0, 1, 2, 3, 4
_STATUS = {1: "TAL_EINVAL", 2: "TAL_ENOMEM", 3: "TAL_EDEVICE", 4: "TAL_ERANGE", 5: "TAL_EAGAIN"}

# every C-ABI entry point of include/tal_chain.h
EXPORTS = ["tal_chain_create", "tal_chain_destroy", "tal_chain_last_error", "tal_chain_dp",
           "tal_chain_dp_device", "tal_chain_get_stats"]


class Params(ctypes.Structure):
    _fields_ = [("max_dist_x", ctypes.c_int32), ("max_dist_y", ctypes.c_int32),
                ("bw", ctypes.c_int32), ("max_skip", ctypes.c_int32),
                ("max_iter", ctypes.c_int32), ("gap_scale", ctypes.c_float),
                ("is_cdna", ctypes.c_int32), ("n_segs", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("ms_kernel", ctypes.c_float), ("ms_total", ctypes.c_float),
                ("sets", ctypes.c_int32), ("lds_ring", ctypes.c_int32),
                ("predecessors", ctypes.c_int64), ("far_anchors", ctypes.c_int64)]


CLI_PATH = os.path.join(HERE, "tal-bench-dp-chaining")


def write_anchor_file(path, anchors, offsets):
    """bench-dp-chaining's anchors file (bench-dp-chaining.cpp:74-82): per set
    its anchor count, then one "r q l" line per anchor."""
    anchors = np.asarray(anchors, ANCHOR_DTYPE)
    with open(path, "w") as fh:
        for k in range(len(offsets) - 1):
            b, e = int(offsets[k]), int(offsets[k + 1])
            fh.write(f"{e - b}\n")
            fh.writelines(f"{int(x['r'])} {int(x['q'])} {int(x['l'])}\n" for x in anchors[b:e])


def read_anchor_file(path):
    """(anchors, offsets) of a bench-dp-chaining anchors file."""
    toks = open(path).read().split()
    parts, offs, i = [], [0], 0
    while i < len(toks):
        m = int(toks[i])
        i += 1
        if m < 0 or i + 3 * m > len(toks):
            raise ValueError(f"{path}: bad or truncated set {len(offs) - 1}")
        a = np.zeros(m, ANCHOR_DTYPE)
        vals = np.array(toks[i:i + 3 * m], dtype=object).reshape(m, 3) if m else None
        if m:
            a["r"] = [int(x) for x in vals[:, 0]]
            a["q"] = [int(x) for x in vals[:, 1]]
            a["l"] = [int(x) for x in vals[:, 2]]
        parts.append(a)
        offs.append(offs[-1] + m)
        i += 3 * m
    anchors = np.concatenate(parts) if parts else np.zeros(0, ANCHOR_DTYPE)
    return anchors, np.array(offs, np.int64)


class TalError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Loads libtal_chain.so; raises if it is not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, i32 = ctypes.c_void_p, ctypes.c_int32
        L.tal_chain_create.restype = ctypes.c_int
        L.tal_chain_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(P)]
        L.tal_chain_destroy.argtypes = [P]
        L.tal_chain_last_error.restype = ctypes.c_char_p
        L.tal_chain_last_error.argtypes = [P]
        L.tal_chain_dp.restype = ctypes.c_int
        L.tal_chain_dp.argtypes = [P, P, P, i32, i32, i32, P, P, P]
        L.tal_chain_dp_device.restype = ctypes.c_int
        L.tal_chain_dp_device.argtypes = [P, P, P, i32, i32, i32, P, P, P, P]
        L.tal_chain_get_stats.restype = ctypes.c_int
        L.tal_chain_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class DpChain:
    """One engine handle (the reference's dp_chain object); the defaults are
    bench-dp-chaining's (bench-dp-chaining.cpp:69)."""

    def __init__(self, max_dist_x=5000, max_dist_y=5000, bw=500, max_skip=25, max_iter=5000,
                 gap_scale=1.0, is_cdna=0, n_segs=1, device=0):
        L = load()
        self.params = Params(max_dist_x, max_dist_y, bw, max_skip, max_iter, gap_scale, is_cdna,
                             n_segs)
        h = ctypes.c_void_p()
        rc = L.tal_chain_create(ctypes.byref(self.params), int(device), ctypes.byref(h))
        if rc != TAL_OK:
            raise TalError(rc, "tal_chain_create failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().tal_chain_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != TAL_OK:
            raise TalError(rc, load().tal_chain_last_error(self._h).decode())

    def dp(self, anchors, offsets, dr=5000, dq=5000):
        """mm_dp_vectorized over every set; returns (f uint32, p int32, v int32)."""
        anchors = np.ascontiguousarray(anchors, ANCHOR_DTYPE)
        offsets = np.ascontiguousarray(offsets, np.int64)
        n = len(anchors)
        f = np.zeros(max(n, 1), np.uint32)
        p = np.zeros(max(n, 1), np.int32)
        v = np.zeros(max(n, 1), np.int32)
        self._check(load().tal_chain_dp(self._h, _ptr(anchors), _ptr(offsets), len(offsets) - 1,
                                        int(dr), int(dq), _ptr(f), _ptr(p), _ptr(v)))
        return f[:n], p[:n], v[:n]

    def dp_device(self, d_anchors, d_offsets, nsets, d_f, d_p, d_v, dr=5000, dq=5000, stream=0):
        c = ctypes.c_void_p
        self._check(load().tal_chain_dp_device(self._h, c(d_anchors), c(d_offsets), int(nsets),
                                               int(dr), int(dq), c(d_f), c(d_p), c(d_v),
                                               c(stream)))

    def stats(self):
        st = Stats()
        self._check(load().tal_chain_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in Stats._fields_}
