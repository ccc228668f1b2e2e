"""Batches that make the 16-bit path read a stale, nonzero F[ncol].

In the reference's AVX-512 16-bit kernel the head/tail scans run over the
narrowed range (bandedSWA.cpp:3267), so a row whose narrowed end reaches the
group's last column ncol scans column F[ncol] before the group has stored it;
the value there is whatever an earlier group (or call) on the same object
left (the F scratch is only cleared for columns below ncol, :3034-3035).

Each unit here is two 32-pair groups:
  * A: 32 identical pairs of 200 bases (h0 50), whose final F is nonzero
    around column 150;
  * B: queries of 30-150 bases (the first exactly 150, so ncol = 150), each
    against a reference of 128-260 bases (never 8-bit) that is, in thirds,
    the query with 2-50% substitutions plus flank, the query with a 3-39 base
    block removed (an insertion in the query) and 0-15% substitutions, or
    unrelated; h0 5-149.
Row 47 of B ends at column 148 (= 47 + w + 1 with w = 100), its narrowing
reaches ncol, and lanes whose scores died out early then take their tail
from the stale value.  The oracle's stale counters show which units hit it.
"""
import numpy as np

import oracle as O


def make(units, seed=1):
    rng = np.random.default_rng(seed)
    refs, qrys, h0s = [], [], []
    for _ in range(units):
        for _ in range(32):
            s = rng.integers(0, 4, 200).astype(np.uint8)
            refs.append(s)
            qrys.append(s.copy())
            h0s.append(50)
        for lane in range(32):
            l2 = 150 if lane == 0 else int(rng.integers(30, 151))
            l1 = int(rng.integers(128, 261))
            q = rng.integers(0, 4, l2).astype(np.uint8)
            kind = rng.integers(0, 3)
            if kind == 0:
                r = np.concatenate([q, rng.integers(0, 4, 300)]).astype(np.uint8)[:l1]
                sub = rng.random(len(r)) < rng.choice([0.02, 0.2, 0.5])
            elif kind == 1:
                a = int(rng.integers(5, max(6, l2 // 2)))
                k = int(rng.integers(3, 40))
                r = np.concatenate([q[:a], q[a + k:], rng.integers(0, 4, 300)]).astype(np.uint8)[:l1]
                sub = rng.random(len(r)) < rng.choice([0.0, 0.05, 0.15])
            else:
                r = rng.integers(0, 4, l1).astype(np.uint8)
                sub = np.zeros(l1, bool)
            r[sub] = rng.integers(0, 4, int(sub.sum()))
            refs.append(r)
            qrys.append(q)
            h0s.append(int(rng.integers(5, 150)))
    return O.make_pairs(refs, qrys, h0s)
