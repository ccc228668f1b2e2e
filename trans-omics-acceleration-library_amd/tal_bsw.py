"""Python binding of libtal_bsw.so (the C ABI in include/tal_bsw.h).

Mirrors the reference's BandedPairWiseSW interface
(src/alignment/BSW/bandedSWA.h:113-336): construct with the scoring
parameters, then call get_scores16 / get_scores8 on a SeqPair array and the
packed reference/query byte buffers; outputs are written in place.  The GPU
library must be built (build.py); there is no CPU fallback.
"""
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtal_bsw.so")

SEQPAIR_FIELDS = ["idr", "idq", "id", "len1", "len2", "h0", "seqid", "regid",
                  "score", "tle", "gtle", "qle", "gscore", "max_off"]
SEQPAIR_DTYPE = np.dtype([(f, "<i4") for f in SEQPAIR_FIELDS])
OUT_FIELDS = ["score", "tle", "gtle", "qle", "gscore", "max_off"]

TAL_OK, TAL_EINVAL, TAL_ENOMEM, TAL_EDEVICE, TAL_ERANGE, TAL_EAGAIN = 0, 1, 2, 3, 4, 5
_STATUS = {1: "TAL_EINVAL", 2: "TAL_ENOMEM", 3: "TAL_EDEVICE", 4: "TAL_ERANGE", 5: "TAL_EAGAIN"}

# every C-ABI entry point of include/tal_bsw.h
EXPORTS = ["tal_bsw_create", "tal_bsw_destroy", "tal_bsw_last_error",
           "tal_bsw_get_scores16", "tal_bsw_get_scores8", "tal_bsw_get_scores_auto",
           "tal_bsw_get_scores16_device", "tal_bsw_get_scores8_device",
           "tal_bsw_get_scores_auto_device", "tal_bsw_get_scores_auto_device_noalloc",
           "tal_bsw_scalar_scores",
           "tal_bsw_scalar_scores_device", "tal_bsw_count_8bit", "tal_bsw_get_stats",
           "tal_bsw_plan_create", "tal_bsw_plan_launch", "tal_bsw_plan_destroy",
           "tal_bsw_autotune", "tal_bsw_set_tuning", "tal_bsw_set_plan_rounds",
           "tal_bsw_server_create", "tal_bsw_server_submit", "tal_bsw_server_wait",
           "tal_bsw_server_get_stats", "tal_bsw_server_destroy",
           "tal_bsw_pairfile_open", "tal_bsw_pairfile_read", "tal_bsw_pairfile_close",
           "tal_bsw_pairfile_write", "tal_bsw_pack_write", "tal_bsw_pack_map",
           "tal_bsw_pack_unmap"]


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("o_del", "e_del", "o_ins", "e_ins", "zdrop", "end_bonus", "w_match", "w_mismatch")]


class Stats(ctypes.Structure):
    _fields_ = [("ms_main16", ctypes.c_float), ("ms_main8", ctypes.c_float),
                ("ms_rerun16", ctypes.c_float)] + [
        (n, ctypes.c_int32) for n in ("groups16", "groups8", "reruns16", "lds_state16",
                                      "pairs8", "pairs16")] + [("ms_replay", ctypes.c_float),
                                                               ("settled16", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int32), ("order", ctypes.c_int32),
                ("ms", ctypes.c_float), ("trials", ctypes.c_int32),
                ("trial_waves", ctypes.c_int32 * 16), ("trial_order", ctypes.c_int32 * 16),
                ("trial_ms", ctypes.c_float * 16)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("workers", ctypes.c_int32),
                ("max_batch_pairs", ctypes.c_int32), ("resident", ctypes.c_int32),
                ("max_query_len", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("batch_requests", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_int64), ("pairs", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("peak_bytes", ctypes.c_int64),
                ("busy_ms", ctypes.c_double), ("reruns", ctypes.c_int64)]


class Pack(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_void_p), ("n", ctypes.c_int64),
                ("ref", ctypes.c_void_p), ("ref_len", ctypes.c_size_t),
                ("qer", ctypes.c_void_p), ("qer_len", ctypes.c_size_t),
                ("mapping", ctypes.c_void_p), ("mapping_len", ctypes.c_size_t)]


class TalError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Loads libtal_bsw.so; raises if it is not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, i32, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_size_t
        L.tal_bsw_create.restype = ctypes.c_int
        L.tal_bsw_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(P)]
        L.tal_bsw_destroy.argtypes = [P]
        L.tal_bsw_last_error.restype = ctypes.c_char_p
        L.tal_bsw_last_error.argtypes = [P]
        for name in ("tal_bsw_get_scores16", "tal_bsw_get_scores8", "tal_bsw_get_scores_auto"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [P, P, P, sz, P, sz, i32, i32]
        for name in ("tal_bsw_get_scores16_device", "tal_bsw_get_scores8_device",
                     "tal_bsw_get_scores_auto_device", "tal_bsw_get_scores_auto_device_noalloc"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [P, P, P, P, i32, i32, P]
        L.tal_bsw_count_8bit.restype = i32
        L.tal_bsw_count_8bit.argtypes = [P, P, i32]
        L.tal_bsw_get_stats.restype = ctypes.c_int
        L.tal_bsw_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        L.tal_bsw_plan_create.restype = ctypes.c_int
        L.tal_bsw_plan_create.argtypes = [P, P, P, P, i32, i32, ctypes.POINTER(P)]
        L.tal_bsw_plan_launch.restype = ctypes.c_int
        L.tal_bsw_plan_launch.argtypes = [P, P]
        L.tal_bsw_plan_destroy.argtypes = [P]
        L.tal_bsw_autotune.restype = ctypes.c_int
        L.tal_bsw_autotune.argtypes = [P, P, P, P, i32, i32, i32, ctypes.POINTER(Tuning)]
        L.tal_bsw_set_tuning.restype = ctypes.c_int
        L.tal_bsw_set_tuning.argtypes = [P, i32, i32]
        L.tal_bsw_set_plan_rounds.restype = ctypes.c_int
        L.tal_bsw_scalar_scores.restype = ctypes.c_int
        L.tal_bsw_scalar_scores.argtypes = [P, P, P, P, sz, P, sz, i32, i32]
        L.tal_bsw_scalar_scores_device.restype = ctypes.c_int
        L.tal_bsw_scalar_scores_device.argtypes = [P, P, P, P, P, i32, i32, P]
        L.tal_bsw_set_plan_rounds.argtypes = [P, i32]
        L.tal_bsw_server_create.restype = ctypes.c_int
        L.tal_bsw_server_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int,
                                            ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]
        L.tal_bsw_server_submit.restype = ctypes.c_int
        L.tal_bsw_server_submit.argtypes = [P, P, P, sz, P, sz, i32, i32,
                                            ctypes.POINTER(ctypes.c_uint64)]
        L.tal_bsw_server_wait.restype = ctypes.c_int
        L.tal_bsw_server_wait.argtypes = [P, ctypes.c_uint64, i32, ctypes.POINTER(RequestInfo)]
        L.tal_bsw_server_get_stats.restype = ctypes.c_int
        L.tal_bsw_server_get_stats.argtypes = [P, ctypes.POINTER(ServerStats)]
        L.tal_bsw_server_destroy.argtypes = [P]
        L.tal_bsw_pairfile_open.restype = ctypes.c_int
        L.tal_bsw_pairfile_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
        L.tal_bsw_pairfile_read.restype = ctypes.c_int
        L.tal_bsw_pairfile_read.argtypes = [P, P, P, sz, P, sz, i32, ctypes.POINTER(i32)]
        L.tal_bsw_pairfile_close.argtypes = [P]
        L.tal_bsw_pairfile_write.restype = ctypes.c_int
        L.tal_bsw_pairfile_write.argtypes = [ctypes.c_char_p, P, P, P, i32]
        L.tal_bsw_pack_write.restype = ctypes.c_int
        L.tal_bsw_pack_write.argtypes = [ctypes.c_char_p, P, ctypes.c_int64, P, sz, P, sz]
        L.tal_bsw_pack_map.restype = ctypes.c_int
        L.tal_bsw_pack_map.argtypes = [ctypes.c_char_p, ctypes.POINTER(Pack)]
        L.tal_bsw_pack_unmap.argtypes = [ctypes.POINTER(Pack)]
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class BandedPairWiseSW:
    """One engine handle on one GPU (the reference object's counterpart)."""

    def __init__(self, o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100, end_bonus=5,
                 w_match=1, w_mismatch=4, device=0):
        L = load()
        self.params = Params(o_del, e_del, o_ins, e_ins, zdrop, end_bonus, w_match, w_mismatch)
        h = ctypes.c_void_p()
        rc = L.tal_bsw_create(ctypes.byref(self.params), int(device), ctypes.byref(h))
        if rc != TAL_OK:
            raise TalError(rc, "tal_bsw_create failed")
        self._h = h
        self._plans = weakref.WeakSet()

    def close(self):
        if getattr(self, "_h", None):
            # a plan uses its handle's buffers and stream: release plans first
            for p in list(self._plans):
                p.close()
            load().tal_bsw_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != TAL_OK:
            raise TalError(rc, load().tal_bsw_last_error(self._h).decode())

    def _host(self, fn, pairs, ref, qer, w):
        assert pairs.dtype == SEQPAIR_DTYPE and pairs.flags.c_contiguous
        ref = np.ascontiguousarray(ref, dtype=np.uint8)
        qer = np.ascontiguousarray(qer, dtype=np.uint8)
        self._check(fn(self._h, _ptr(pairs), _ptr(ref), ref.nbytes, _ptr(qer), qer.nbytes,
                       len(pairs), int(w)))
        return pairs

    def get_scores16(self, pairs, ref, qer, w):
        return self._host(load().tal_bsw_get_scores16, pairs, ref, qer, w)

    def get_scores8(self, pairs, ref, qer, w):
        return self._host(load().tal_bsw_get_scores8, pairs, ref, qer, w)

    def get_scores_auto(self, pairs, ref, qer, w):
        return self._host(load().tal_bsw_get_scores_auto, pairs, ref, qer, w)

    def scalar_scores(self, pairs, ref, qer, w, mat=None):
        """scalarBandedSWAWrapper: every pair on its own, scored with the 5x5
        matrix mat (default: bwa_fill_scmat of the handle's match/mismatch,
        ambiguity -1)."""
        m = fill_scmat(self.params.w_match, self.params.w_mismatch) if mat is None else mat
        m = np.ascontiguousarray(m, dtype=np.int8).reshape(25)
        assert pairs.dtype == SEQPAIR_DTYPE and pairs.flags.c_contiguous
        ref = np.ascontiguousarray(ref, dtype=np.uint8)
        qer = np.ascontiguousarray(qer, dtype=np.uint8)
        self._check(load().tal_bsw_scalar_scores(self._h, _ptr(m), _ptr(pairs), _ptr(ref),
                                                 ref.nbytes, _ptr(qer), qer.nbytes, len(pairs),
                                                 int(w)))
        return pairs

    def scalar_scores_device(self, mat, d_pairs, d_ref, d_qer, n, w, stream=0):
        m = np.ascontiguousarray(mat, dtype=np.int8).reshape(25)
        self._check(load().tal_bsw_scalar_scores_device(
            self._h, _ptr(m), ctypes.c_void_p(d_pairs), ctypes.c_void_p(d_ref),
            ctypes.c_void_p(d_qer), int(n), int(w), ctypes.c_void_p(stream or None)))

    def get_scores_device(self, bits, d_pairs, d_ref, d_qer, n, w, stream=0):
        """Device pointers (ints); bits in {16, 8, 0 (auto)}."""
        fn = {16: load().tal_bsw_get_scores16_device, 8: load().tal_bsw_get_scores8_device,
              0: load().tal_bsw_get_scores_auto_device}[bits]
        self._check(fn(self._h, ctypes.c_void_p(d_pairs), ctypes.c_void_p(d_ref),
                       ctypes.c_void_p(d_qer), int(n), int(w), ctypes.c_void_p(stream or None)))

    def get_scores_device_noalloc(self, d_pairs, d_ref, d_qer, n, w, stream=0):
        """The dispatcher on device pointers, with all call scratch taken from
        the library's static device array (no device allocation)."""
        self._check(load().tal_bsw_get_scores_auto_device_noalloc(
            self._h, ctypes.c_void_p(d_pairs), ctypes.c_void_p(d_ref), ctypes.c_void_p(d_qer),
            int(n), int(w), ctypes.c_void_p(stream or None)))

    def stats(self):
        st = Stats()
        self._check(load().tal_bsw_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in Stats._fields_}

    def count_8bit(self, pairs):
        return int(load().tal_bsw_count_8bit(self._h, _ptr(pairs), len(pairs)))

    def plan(self, d_pairs, d_ref, d_qer, n, w):
        """A replayable, graph-captured plan of a device batch (Plan)."""
        p = Plan(self, d_pairs, d_ref, d_qer, n, w)
        self._plans.add(p)
        return p

    def autotune(self, d_pairs, d_ref, d_qer, n, w, sample=0):
        """Times the launch configurations on the batch (device pointers) and
        applies the fastest; returns the trials."""
        t = Tuning()
        self._check(load().tal_bsw_autotune(self._h, ctypes.c_void_p(d_pairs),
                                            ctypes.c_void_p(d_ref), ctypes.c_void_p(d_qer),
                                            int(n), int(w), int(sample), ctypes.byref(t)))
        k = t.trials
        return {"waves_per_cu": t.waves_per_cu, "order": t.order, "ms": t.ms,
                "trials": [{"waves_per_cu": t.trial_waves[i], "order": t.trial_order[i],
                            "ms": t.trial_ms[i]} for i in range(k)]}

    def set_tuning(self, waves_per_cu, order):
        self._check(load().tal_bsw_set_tuning(self._h, int(waves_per_cu), int(order)))

    def set_plan_rounds(self, rounds):
        """Stale-column rerun rounds captured in the graph of later plans."""
        self._check(load().tal_bsw_set_plan_rounds(self._h, int(rounds)))


class Plan:
    """tal_bsw_plan_*: the dispatcher pipeline of one device batch captured in
    a hipGraph; launch() scores the batch's current contents."""

    def __init__(self, eng, d_pairs, d_ref, d_qer, n, w):
        self._eng = eng
        p = ctypes.c_void_p()
        eng._check(load().tal_bsw_plan_create(eng._h, ctypes.c_void_p(d_pairs),
                                              ctypes.c_void_p(d_ref), ctypes.c_void_p(d_qer),
                                              int(n), int(w), ctypes.byref(p)))
        self._p = p

    def launch(self, stream=0):
        self._eng._check(load().tal_bsw_plan_launch(self._p, ctypes.c_void_p(stream or None)))

    def close(self):
        if getattr(self, "_p", None):
            load().tal_bsw_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()


class Server:
    """Streamed serving with continuous batching (tal_bsw_server_* of
    include/tal_bsw.h).  Each request is scored as one dispatcher call on a
    fresh engine object, whatever it is batched with.  resident=True runs the
    device-resident loop (one persistent kernel pulling tasks) instead of
    host-launched batches."""

    def __init__(self, hbm_budget=0, workers=0, max_batch_pairs=0, device=0, o_del=6, e_del=1,
                 o_ins=6, e_ins=1, zdrop=100, end_bonus=5, w_match=1, w_mismatch=4,
                 resident=False, max_query_len=0):
        L = load()
        self.params = Params(o_del, e_del, o_ins, e_ins, zdrop, end_bonus, w_match, w_mismatch)
        cfg = ServerConfig(int(hbm_budget), int(workers), int(max_batch_pairs),
                           1 if resident else 0, int(max_query_len))
        h = ctypes.c_void_p()
        rc = L.tal_bsw_server_create(ctypes.byref(self.params), int(device), ctypes.byref(cfg),
                                     ctypes.byref(h))
        if rc != TAL_OK:
            raise TalError(rc, "tal_bsw_server_create failed")
        self._h = h
        self._live = {}

    def submit(self, pairs, ref, qer, w):
        """Queues a request; the arrays are kept alive until it is waited for."""
        assert pairs.dtype == SEQPAIR_DTYPE and pairs.flags.c_contiguous
        ref = np.ascontiguousarray(ref, dtype=np.uint8)
        qer = np.ascontiguousarray(qer, dtype=np.uint8)
        t = ctypes.c_uint64()
        rc = load().tal_bsw_server_submit(self._h, _ptr(pairs), _ptr(ref), ref.nbytes, _ptr(qer),
                                          qer.nbytes, len(pairs), int(w), ctypes.byref(t))
        if rc != TAL_OK:
            raise TalError(rc, "submit failed")
        self._live[t.value] = (pairs, ref, qer)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """Returns the request's record (status, batch, batch_requests,
        queue_ms, latency_ms), or None if it is not finished in time."""
        info = RequestInfo()
        rc = load().tal_bsw_server_wait(self._h, ctypes.c_uint64(ticket), int(timeout_ms),
                                        ctypes.byref(info))
        if rc == TAL_EAGAIN:
            return None
        self._live.pop(ticket, None)
        return {f: getattr(info, f) for f, _ in RequestInfo._fields_ if f != "pad"}

    def stats(self):
        st = ServerStats()
        rc = load().tal_bsw_server_get_stats(self._h, ctypes.byref(st))
        if rc != TAL_OK:
            raise TalError(rc, "stats failed")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            load().tal_bsw_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __del__(self):
        self.close()


def fill_scmat(match, mismatch, ambig=-1):
    """The 5x5 matrix bench-bsw builds (bwa_fill_scmat, bench-bsw.cpp:56-64):
    match on the diagonal of codes 0-3, -mismatch off it, ambig in row and
    column 4."""
    m = np.full((5, 5), ambig, np.int8)
    for i in range(4):
        for j in range(4):
            m[i, j] = match if i == j else -mismatch
    return m.reshape(25)


def _smax(w_match, w_mismatch):
    # max over the object's scores: match, -mismatch penalty, ambiguity -1, and 0
    return max(0, int(np.int8(w_match)), int(np.int8(-int(np.int8(w_mismatch)))), -1)


def pair_band(len2, w, o_del=6, e_del=1, o_ins=6, e_ins=1, end_bonus=5, w_match=1,
              w_mismatch=4, **_):
    """Per-pair band of the 16-bit wrapper (bandedSWA.cpp:2882-2904)."""
    sm = _smax(w_match, w_mismatch)
    ql = (np.asarray(len2, np.int64) * sm) & 0xFFFF
    b = np.full(ql.shape, w, np.int64)
    for o, e in ((o_ins, e_ins), (o_del, e_del)):
        t = (ql + end_bonus - o) & 0xFFFF
        bb = np.maximum(np.trunc(t / e).astype(np.int64) + 1, 1)
        b = np.minimum(b, bb) & 0xFFFF
    return b


def band_cells(len1, len2, w, **params):
    """Cells inside each pair's band, sum_i max(0, min(len2, i+b+1) - max(0, i-b))
    (the work unit of the GCUPS figures)."""
    len1 = np.asarray(len1, np.int64)
    len2 = np.asarray(len2, np.int64)
    if len1.size == 0:
        return np.zeros(0, np.int64)
    keys, inv = np.unique((np.maximum(len1, 0) << 20) | np.maximum(len2, 0), return_inverse=True)
    l1, l2 = keys >> 20, keys & 0xFFFFF
    b = pair_band(l2, w, **params)
    out = np.zeros(keys.shape, np.int64)
    for i in range(int(l1.max())):
        span = np.minimum(l2, i + b + 1) - np.maximum(0, i - b)
        out += np.where(i < l1, np.maximum(span, 0), 0)
    return out[inv.reshape(-1)]


def is_8bit(pairs, w_match=1, w_mismatch=4, **_):
    """Dispatcher rule of tal_bsw_get_scores_auto (include/tal_bsw.h)."""
    sm = _smax(w_match, w_mismatch)
    l1, l2, h0 = (pairs[f].astype(np.int64) for f in ("len1", "len2", "h0"))
    return ((l1 >= 0) & (l2 >= 0) & (l1 <= 127) & (l2 <= 127) & (h0 >= 0)
            & (h0 + sm * np.minimum(l1, l2) <= 127))


REF_SLOT, QER_SLOT, CHUNK = 256, 128, 20000   # bandedSWA.h:47-51


def read_pairfile(path, max_pairs=CHUNK, ref_slot=REF_SLOT, qer_slot=QER_SLOT):
    """Yields (pairs, ref, qer) chunks of a bench-bsw text pair file
    (benchmarks/bench-bsw.cpp:117-169 layout: fixed slots per pair)."""
    L = load()
    f = ctypes.c_void_p()
    rc = L.tal_bsw_pairfile_open(os.fsencode(path), ctypes.byref(f))
    if rc != TAL_OK:
        raise TalError(rc, f"cannot open {path}")
    try:
        while True:
            pairs = np.zeros(max_pairs, SEQPAIR_DTYPE)
            ref = np.zeros(max_pairs * ref_slot, np.uint8)
            qer = np.zeros(max_pairs * qer_slot, np.uint8)
            n = ctypes.c_int32()
            rc = L.tal_bsw_pairfile_read(f, _ptr(pairs), _ptr(ref), ref_slot, _ptr(qer), qer_slot,
                                         max_pairs, ctypes.byref(n))
            if rc != TAL_OK:
                raise TalError(rc, f"reading {path}")
            if n.value == 0:
                return
            yield pairs[:n.value], ref[:n.value * ref_slot], qer[:n.value * qer_slot]
    finally:
        L.tal_bsw_pairfile_close(f)


def write_pairfile(path, pairs, ref, qer):
    """Writes pairs as a bench-bsw text pair file (h0, reference, query lines)."""
    pairs = np.ascontiguousarray(pairs, SEQPAIR_DTYPE)
    ref = np.ascontiguousarray(ref, np.uint8)
    qer = np.ascontiguousarray(qer, np.uint8)
    rc = load().tal_bsw_pairfile_write(os.fsencode(path), _ptr(pairs), _ptr(ref), _ptr(qer),
                                       len(pairs))
    if rc != TAL_OK:
        raise TalError(rc, f"writing {path}")


def write_pack(path, pairs, ref, qer):
    """Writes a packed pair file (header, SeqPair records, reference bytes, query bytes)."""
    pairs = np.ascontiguousarray(pairs, SEQPAIR_DTYPE)
    ref = np.ascontiguousarray(ref, np.uint8)
    qer = np.ascontiguousarray(qer, np.uint8)
    rc = load().tal_bsw_pack_write(os.fsencode(path), _ptr(pairs), len(pairs), _ptr(ref),
                                   ref.nbytes, _ptr(qer), qer.nbytes)
    if rc != TAL_OK:
        raise TalError(rc, f"writing {path}")


class MappedPack:
    """A packed pair file mapped copy-on-write; pairs/ref/qer are numpy views of
    the mapping, valid until close()."""

    def __init__(self, path):
        self._p = Pack()
        rc = load().tal_bsw_pack_map(os.fsencode(path), ctypes.byref(self._p))
        if rc != TAL_OK:
            raise TalError(rc, f"cannot map {path}")
        p = self._p

        def view(addr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(addr)
            return np.frombuffer(buf, dtype)

        self.pairs = view(p.pairs, p.n, SEQPAIR_DTYPE)
        self.ref = view(p.ref, p.ref_len, np.uint8)
        self.qer = view(p.qer, p.qer_len, np.uint8)

    def close(self):
        if self._p.mapping:
            self.pairs = self.ref = self.qer = None
            load().tal_bsw_pack_unmap(ctypes.byref(self._p))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
