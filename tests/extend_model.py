"""Plain-Python model of the extension-pair producer's rules
(include/tal_extend.h, steps 1-4), the checker of the GPU producer, and of
the pipeline's score flow on top of the BSW oracle.  Test infrastructure."""
import numpy as np

import oracle as O
import tal_bsw as TB
import tal_extend as TE


def produce(genome, reads, read_off, anchors, anchor_off, f, p, pad=100, match=1):
    """(reads READ_DTYPE with seed fields and pair indexes, left pairs, right
    pairs, ref bytes, query bytes) as the producer builds them."""
    n = len(read_off) - 1
    L = len(genome)
    out = np.zeros(n, TE.READ_DTYPE)
    out["left_pair"] = out["right_pair"] = -1
    left, right, ref, qer = [], [], [], []
    roff = qoff = 0
    for k in range(n):
        a0, a1 = int(anchor_off[k]), int(anchor_off[k + 1])
        read = reads[int(read_off[k]):int(read_off[k + 1])]
        m = len(read)
        if a1 == a0:
            continue
        fs = [int(x) for x in f[a0:a1]]
        e = fs.index(max(fs))  # the first maximum
        out["chain_score"][k] = fs[e]
        segs, steps, j = [], 0, e
        cur = None
        while j >= 0 and steps < a1 - a0:
            x = anchors[a0 + j]
            r, q, ln = int(x["r"]), int(x["q"]), int(x["l"])
            d = r - q
            if cur is not None and d == cur[0] and q + ln >= cur[1]:
                cur = [d, min(cur[1], q), max(cur[2], q + ln)]
            else:
                if cur is not None:
                    segs.append(cur)
                cur = [d, q, q + ln]
            steps += 1
            pj = int(p[a0 + j])
            j = pj if pj < j else -1
        segs.append(cur)
        best = None
        for d, lo, hi in segs:  # longest, then smaller q
            if best is None or hi - lo > best[2] - best[1] or (hi - lo == best[2] - best[1] and lo < best[1]):
                best = (d, lo, hi)
        d, qb, qe = best
        rb = d + qb
        out["chain_anchors"][k] = steps
        out["seed_r"][k], out["seed_q"][k], out["seed_len"][k] = rb, qb, qe - qb
        h0 = match * (qe - qb)
        out["left_score"][k] = out["score"][k] = h0
        lq, lt = qb, min(qb + pad, rb)
        if lq == 0 or lt == 0:
            lq = lt = 0
        re = rb + (qe - qb)
        rq = m - qe
        rt = min(rq + pad, L - re)
        if rq == 0 or rt == 0:
            rq = rt = 0
        if lq:
            sp = np.zeros(1, TB.SEQPAIR_DTYPE)
            sp["idr"], sp["idq"], sp["id"] = roff, qoff, len(left)
            sp["len1"], sp["len2"], sp["h0"], sp["seqid"] = lt, lq, h0, k
            out["left_pair"][k] = len(left)
            left.append(sp)
            ref.append(genome[rb - lt:rb][::-1])
            qer.append(read[:qb][::-1])
        if rq:
            sp = np.zeros(1, TB.SEQPAIR_DTYPE)
            sp["idr"], sp["idq"], sp["id"] = roff + lt, qoff + lq, len(right)
            sp["len1"], sp["len2"], sp["h0"], sp["seqid"] = rt, rq, h0, k
            out["right_pair"][k] = len(right)
            right.append(sp)
            ref.append(genome[re:re + rt])
            qer.append(read[qe:])
        roff += lt + rt
        qoff += lq + rq
    cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt)
    return (out, cat(left, TB.SEQPAIR_DTYPE), cat(right, TB.SEQPAIR_DTYPE), cat(ref, np.uint8),
            cat(qer, np.uint8))


def pipeline(genome, reads, read_off, anchors, anchor_off, pad=100, w=100, match=1):
    """The whole pipeline's read records: the chaining oracle, the producer
    model, the left pass, h0 of the right pairs, the right pass (both passes
    16-bit, on one BSW oracle object)."""
    f, p, _ = O.chain_dp_sets(anchors, anchor_off, nthreads=8)
    out, left, right, ref, qer = produce(genome, reads, read_off, anchors, anchor_off, f, p, pad,
                                         match)
    orc = O.Oracle()
    refb = ref if len(ref) else np.zeros(1, np.uint8)
    qerb = qer if len(qer) else np.zeros(1, np.uint8)
    if len(left):
        left = orc.scores16(left.copy(), refb, qerb, w)
    for k in range(len(out)):
        lp = out["left_pair"][k]
        h0 = int(left["score"][lp]) if lp >= 0 else match * int(out["seed_len"][k])
        out["left_score"][k] = out["score"][k] = h0
        rp = out["right_pair"][k]
        if rp >= 0:
            right["h0"][rp] = h0
    if len(right):
        right = orc.scores16(right.copy(), refb, qerb, w)
    for k in range(len(out)):
        rp = out["right_pair"][k]
        if rp >= 0:
            out["score"][k] = right["score"][rp]
    orc.close()
    return out, left, right
