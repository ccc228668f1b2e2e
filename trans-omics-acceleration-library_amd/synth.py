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


# ---------------------------------------------------------------------------
# Full inter-task SW batches (PairWiseSW_{8,16,32} layout, bench-smith-
# waterman's input): pair k's seq1 at seq_buf + 2kS and seq2 at (2k+1)S, as
# ASCII bases.  seq2 is seq1 read from a random offset with 10% substitutions,
# so that the alignments are local and the scores non-trivial.

SW_PAIR_DTYPE = np.dtype([("id", "<i4"), ("len1", "<i2"), ("len2", "<i2"), ("score", "<i4")])
SW_SLOT = {8: 128, 16: 1024, 32: 1024}

# name -> (class bits, pairs, min length, max length)
SW_WORKLOADS = {
    "sw16": (16, 200_000, 200, 1000),
    "sw8": (8, 1_000_000, 30, 124),
    "sw32": (32, 50_000, 200, 1000),
}


def sw_generate(bits, n, lo, hi, seed=7, sub=0.10):
    """(pairs, seq_buf) for n pairs with len1, len2 ~ U[lo, hi]."""
    S = SW_SLOT[bits]
    rng = np.random.default_rng(seed)
    bases = np.frombuffer(b"ACGT", np.uint8)
    len1 = rng.integers(lo, hi + 1, n)
    len2 = rng.integers(lo, hi + 1, n)
    buf = np.zeros((max(n, 1), 2, S), np.uint8)
    src = bases[rng.integers(0, 4, (max(n, 1), S), dtype=np.uint8)]
    buf[:, 0, :] = src
    shift = rng.integers(0, max(1, S // 8), max(n, 1))
    idx = (np.arange(S)[None, :] + shift[:, None]) % S
    q = np.take_along_axis(src, idx, axis=1)
    mut = rng.random(q.shape) < sub
    q[mut] = bases[rng.integers(0, 4, int(mut.sum()), dtype=np.uint8)]
    buf[:, 1, :] = q
    pos = np.arange(S)[None, :]
    buf[:, 0, :][pos >= len1[:, None]] = 0
    buf[:, 1, :][pos >= len2[:, None]] = 0
    pairs = np.zeros(n, SW_PAIR_DTYPE)
    pairs["id"] = np.arange(n)
    pairs["len1"] = len1
    pairs["len2"] = len2
    return pairs, buf.reshape(-1)


# ---------------------------------------------------------------------------
# DP chaining inputs (bench-dp-chaining's anchor sets): per set, anchors of a
# few true chains (near one diagonal, r and q advancing together with small
# drift) and random anchors, sorted by r; l is the seed length.

ANCHOR_DTYPE = np.dtype([("r", "<u8"), ("q", "<i4"), ("l", "<i4")])

# name -> (sets, min anchors, max anchors)
CHAIN_WORKLOADS = {
    "chain": (4000, 200, 4000),
}


def chain_generate(nsets, lo, hi, seed=11, noise=0.3, kmer=(15, 28)):
    """(anchors, offsets): nsets sets of U[lo, hi] anchors each."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, nsets)
    offsets = np.zeros(nsets + 1, np.int64)
    offsets[1:] = np.cumsum(sizes)
    n = int(offsets[-1])
    anchors = np.zeros(n, ANCHOR_DTYPE)
    for k in range(nsets):
        m = int(sizes[k])
        base = int(rng.integers(0, 3_000_000_000))
        nchains = int(rng.integers(1, 4))
        is_noise = rng.random(m) < noise
        chain_id = rng.integers(0, nchains, m)
        starts_r = base + rng.integers(0, 200_000, nchains)
        starts_q = rng.integers(0, 5_000, nchains)
        step = rng.integers(1, 40, m)
        pos = np.zeros(m, np.int64)
        for c in range(nchains):
            sel = (chain_id == c) & ~is_noise
            pos[sel] = np.cumsum(step[sel])
        drift = rng.integers(-4, 5, m)
        r = starts_r[chain_id] + pos
        q = starts_q[chain_id] + pos + drift
        r[is_noise] = base + rng.integers(0, 220_000, int(is_noise.sum()))
        q[is_noise] = rng.integers(0, 60_000, int(is_noise.sum()))
        order = np.argsort(r, kind="stable")
        a, b = int(offsets[k]), int(offsets[k + 1])
        anchors["r"][a:b] = r[order]
        anchors["q"][a:b] = np.maximum(q[order], 0)
        anchors["l"][a:b] = rng.integers(kmer[0], kmer[1] + 1, m)
    return anchors, offsets


# Reads for the extension pipeline (extension-pair producer): reads sampled
# from a random genome with substitutions and short indels; each read's
# anchors are its k-mers that match the genome along the read's true
# alignment (exact runs of k), plus a few random spurious ones, sorted by r.
EXTEND_WORKLOADS = {
    # reads, read length, genome length
    "extend": (200_000, 150, 50_000_000),
}


def extend_generate(nreads, read_len, genome_len, seed=13, sub=0.01, indel=0.001, k=19,
                    spurious=6):
    """dict(genome uint8 codes 0-3, reads uint8 codes, read_off int64[nreads+1],
    anchors ANCHOR_DTYPE, anchor_off int64[nreads+1], origin int64[nreads])."""
    rng = np.random.default_rng(seed)
    genome = rng.integers(0, 4, genome_len, dtype=np.uint8)
    span = read_len + read_len // 4
    origin = rng.integers(0, genome_len - span, nreads)
    # edit process per read base: 0 match/sub, 1 insertion, 2 deletion (skip
    # a genome base, then take the next)
    ops = rng.choice(3, (nreads, read_len), p=[1 - 2 * indel, indel, indel]).astype(np.int8)
    is_ins = ops == 1
    gstep = np.where(is_ins, 0, np.where(ops == 2, 2, 1))
    gpos = origin[:, None] + np.cumsum(gstep, axis=1) - 1
    gpos = np.where(is_ins, -1, gpos)
    gidx = np.clip(gpos, 0, genome_len - 1)
    reads = np.where(is_ins, rng.integers(0, 4, (nreads, read_len)), genome[gidx]).astype(np.uint8)
    subs = (rng.random((nreads, read_len)) < sub) & ~is_ins
    reads[subs] = (reads[subs] + rng.integers(1, 4, int(subs.sum()))).astype(np.uint8) % 4
    # a read base matches when it is not inserted, not substituted; a k-mer
    # at i is an anchor when bases i..i+k-1 match on consecutive genome bases
    ok = ~is_ins & ~subs
    consec = np.ones((nreads, read_len), bool)
    consec[:, 1:] = gpos[:, 1:] == gpos[:, :-1] + 1
    good = ok & consec
    good[:, 0] = ok[:, 0]
    # run length of good bases ending at each position (a base starts a new
    # run when it is not consecutive with the previous one)
    run = np.zeros((nreads, read_len), np.int32)
    run[:, 0] = ok[:, 0]
    for i in range(1, read_len):
        run[:, i] = np.where(ok[:, i], np.where(consec[:, i], run[:, i - 1] + 1, 1), 0)
    end_ok = run[:, k - 1:] >= k  # k-mer ending at i + k - 1 starting at i
    rr, ii = np.nonzero(end_ok)
    a_r = gpos[rr, ii].astype(np.uint64)
    a_q = ii.astype(np.int32)
    # spurious anchors: random positions near and far
    sr = np.repeat(np.arange(nreads), spurious)
    s_r = rng.integers(0, genome_len - k, nreads * spurious).astype(np.uint64)
    s_q = rng.integers(0, read_len - k, nreads * spurious).astype(np.int32)
    read_id = np.concatenate([rr, sr])
    r_all = np.concatenate([a_r, s_r])
    q_all = np.concatenate([a_q, s_q])
    order = np.lexsort((q_all, r_all, read_id))
    anchors = np.zeros(len(order), ANCHOR_DTYPE)
    anchors["r"] = r_all[order]
    anchors["q"] = q_all[order]
    anchors["l"] = k
    counts = np.bincount(read_id, minlength=nreads)
    anchor_off = np.zeros(nreads + 1, np.int64)
    anchor_off[1:] = np.cumsum(counts)
    read_off = np.arange(nreads + 1, dtype=np.int64) * read_len
    return dict(genome=genome, reads=reads.reshape(-1), read_off=read_off, anchors=anchors,
                anchor_off=anchor_off, origin=origin.astype(np.int64))
