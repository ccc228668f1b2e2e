#!/usr/bin/env python3
"""Writes bsw_digests.json: the oracle's outputs on every pair of the full
synthetic workloads 1, 2, 3 (one case per band class) and 5, as SHA-256
digests, together with how the vector restatement relates to the scalar one
on those pairs.

For each case (inputs regenerate from the seeded generator, synth.py):
  * "auto"  : the dispatcher's semantics on a fresh object -- getScores8 on
              the pairs the 8/16-bit rule sends to the 8-bit path (in input
              order), getScores16 on the rest;
  * "all16" : getScores16 on every pair, on a fresh object (bench-bsw's call);
  digest = SHA-256 of the six output arrays (score, tle, gtle, qle, gscore,
  max_off), each little-endian int32 in input order, concatenated.
  * vec_ne_scalar: pairs whose outputs differ between the vector path and the
    scalar routine (bandedSWA.cpp:116-237), per variant.
  * toggles (on "all16"): pairs whose outputs change when one vector-only rule
    of the 16-bit path is switched off --
      m7    : the range reset beg = nbeg; end = nend before the head/tail scans
              (bandedSWA.cpp:3267);
      stale : the stale F[ncol] carried between groups (zeroed per group);
      m8    : the F of column tail, one past a lane's band, stored
              (:3150-3155; stored as 0 instead);
    and, for the pairs where vector != scalar, how many each set of toggles
    changes, with the first such pairs listed.
  * stale_decisive: 16-bit groups whose tail scan read a stale F[ncol] that
    decided some lane's tail; stale_nonzero: those where that stale value was
    nonzero in a deciding lane (the groups the engine must rerun, since it
    first assumes zero).

The reference's BSW cannot be compiled here (it needs bwa-mem2's macro.h), so
these digests are the oracle's; tests/test_gpu_fullsize.py checks the engine
against them on the GPU.  Takes about 20 minutes on 8 cores.

Run from the repository root after building:  python tests/golden/make_digests.py
"""
import concurrent.futures as cf
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "trans-omics-acceleration-library_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
import synth as S  # noqa: E402

OUT_FIELDS = ["score", "tle", "gtle", "qle", "gscore", "max_off"]
IN_FIELDS = ["idr", "idq", "len1", "len2", "h0"]
PARAMS = dict(o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100, end_bonus=5, w_match=1, w_mismatch=4)
SCALAR_CHUNK = 250_000
LISTED = 200

# (name, workload, pairs of the workload, band class or None = every pair at w = 100)
CASES = [
    ("workload1", 1, 100_000, None),
    ("workload2", 2, 2_000_000, None),
    ("workload3_w50", 3, 500_000, 50),
    ("workload3_w100", 3, 500_000, 100),
    ("workload3_w200", 3, 500_000, 200),
    ("workload3_w500", 3, 500_000, 500),
    ("workload5", 5, 1_000_000, None),
]
TOGGLES = {"m7": ("mutation", 7), "stale": ("stale_mode", 1), "m8": ("mutation", 8)}


def case_inputs(cfg, n, wclass):
    pairs, ref, qer, w = S.generate(cfg, n, nthreads=1)
    if wclass is None:
        return pairs, ref, qer, 100
    return pairs[np.where(w == wclass)[0]].copy(), ref, qer, wclass


def is_8bit(pairs):
    """The dispatcher's rule (include/tal_bsw.h): len1, len2 <= 127 and
    h0 + match * min(len1, len2) <= 127, with match = 1 here."""
    return ((pairs["len1"] <= 127) & (pairs["len2"] <= 127) & (pairs["h0"] >= 0)
            & (pairs["h0"] + np.minimum(pairs["len1"], pairs["len2"]) <= 127))


def outputs(p):
    return np.stack([np.ascontiguousarray(p[f], "<i4") for f in OUT_FIELDS])


def digest(m):
    return hashlib.sha256(np.ascontiguousarray(m, "<i4").tobytes()).hexdigest()


def input_digest(pairs, ref, qer):
    h = hashlib.sha256()
    for f in IN_FIELDS:
        h.update(np.ascontiguousarray(pairs[f], "<i4").tobytes())
    h.update(ref.tobytes())
    h.update(qer.tobytes())
    return h.hexdigest()


def task(kind, case, arg=None):
    """One oracle run of a case; returns (kind, case, arg, result)."""
    name, cfg, n, wclass = case
    pairs, ref, qer, w = case_inputs(cfg, n, wclass)
    o = O.Oracle(**PARAMS)
    if kind == "auto":
        sel = is_8bit(pairs)
        out = pairs.copy()
        a = o.scores8(pairs[sel].copy(), ref, qer, w)
        b = o.scores16(pairs[~sel].copy(), ref, qer, w)
        st = o.stats()
        for f in OUT_FIELDS:
            out[f][sel] = a[f]
            out[f][~sel] = b[f]
        res = dict(out=outputs(out), n8=int(sel.sum()), stale=st,
                   inputs_sha256=input_digest(pairs, ref, qer))
    elif kind == "all16":
        out = o.scores16(pairs.copy(), ref, qer, w)
        res = dict(out=outputs(out), stale=o.stats())
    elif kind in TOGGLES:
        how, v = TOGGLES[kind]
        (o.set_mutation if how == "mutation" else o.set_stale_mode)(v)
        res = dict(out=outputs(o.scores16(pairs.copy(), ref, qer, w)))
    elif kind == "scalar":
        lo = arg
        hi = min(len(pairs), lo + SCALAR_CHUNK)
        res = dict(out=outputs(o.scalar(pairs[lo:hi].copy(), ref, qer, w)))
    else:
        raise ValueError(kind)
    o.close()
    return kind, name, arg, res


def main():
    t0 = time.time()
    npairs = {c[0]: len(case_inputs(c[1], c[2], c[3])[0]) for c in CASES}
    jobs = []
    for c in CASES:
        for kind in ("auto", "all16", *TOGGLES):
            jobs.append((kind, c, None))
        for lo in range(0, npairs[c[0]], SCALAR_CHUNK):
            jobs.append(("scalar", c, lo))
    # longest first: workload 3 at w = 500, then workload 2
    weight = {"workload3_w500": 0, "workload2": 1, "workload3_w200": 2}
    jobs.sort(key=lambda j: (weight.get(j[1][0], 9), j[0] == "scalar"))
    results = {}
    workers = int(os.environ.get("DIGEST_WORKERS", "7"))
    with cf.ProcessPoolExecutor(max_workers=workers) as ex:
        futs = [ex.submit(task, *j) for j in jobs]
        for k, f in enumerate(cf.as_completed(futs)):
            kind, name, arg, res = f.result()
            results[(name, kind, arg)] = res
            print(f"[{time.time() - t0:6.0f}s] {k + 1}/{len(jobs)} {name} {kind} {arg or ''}",
                  flush=True)
    manifest = {"params": PARAMS, "out_fields": OUT_FIELDS,
                "digest": "sha256 of the six output arrays (int32 LE, input order), concatenated",
                "note": "inputs regenerate from the seeded synthetic workloads (synth.py); "
                        "each variant runs on a fresh object", "cases": []}
    for name, cfg, n, wclass in CASES:
        auto = results[(name, "auto", None)]
        all16 = results[(name, "all16", None)]
        m = npairs[name]
        scal = np.concatenate([results[(name, "scalar", lo)]["out"]
                               for lo in range(0, m, SCALAR_CHUNK)], axis=1)
        ne_auto = (auto["out"] != scal).any(axis=0)
        ne_all16 = (all16["out"] != scal).any(axis=0)
        changed = {t: (results[(name, t, None)]["out"] != all16["out"]).any(axis=0)
                   for t in TOGGLES}
        mask = np.zeros(m, np.int32)
        for b, t in enumerate(TOGGLES):
            mask |= changed[t].astype(np.int32) << b
        idx = np.where(ne_all16)[0]
        by_set = {}
        for b in range(1 << len(TOGGLES)):
            label = "+".join(t for k, t in enumerate(TOGGLES) if (b >> k) & 1) or "none"
            cnt = int((mask[idx] == b).sum())
            if cnt:
                by_set[label] = cnt
        manifest["cases"].append({
            "name": name, "workload": cfg, "workload_pairs": n, "band_class": wclass,
            "w": wclass or 100, "pairs": m, "inputs_sha256": auto["inputs_sha256"],
            "pairs_8bit": auto["n8"],
            "digest_auto": digest(auto["out"]), "digest_all16": digest(all16["out"]),
            "digest_scalar": digest(scal),
            "stale_decisive_groups": {"auto": auto["stale"]["stale_decisive"],
                                      "all16": all16["stale"]["stale_decisive"]},
            "stale_nonzero_groups": {"auto": auto["stale"]["stale_nonzero"],
                                     "all16": all16["stale"]["stale_nonzero"]},
            "vec_ne_scalar": {"auto": int(ne_auto.sum()), "all16": int(ne_all16.sum())},
            "toggle_changes_all16": {t: int(changed[t].sum()) for t in TOGGLES},
            "vec_ne_scalar_all16_by_toggles": by_set,
            "vec_ne_scalar_all16_first": [[int(i), int(mask[i])] for i in idx[:LISTED]],
        })
        print(name, json.dumps(manifest["cases"][-1]["vec_ne_scalar"]),
              json.dumps(manifest["cases"][-1]["toggle_changes_all16"]), flush=True)
    with open(os.path.join(HERE, "bsw_digests.json"), "w") as f:
        json.dump(manifest, f, indent=1)
        f.write("\n")
    print(f"done in {time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
