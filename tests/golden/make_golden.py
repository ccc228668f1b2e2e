#!/usr/bin/env python3
"""Writes the golden fixtures of the BSW engine: bsw_golden.json (manifest)
and bsw_golden.npz (expected outputs).

Each case is a prefix of one of the seeded synthetic workloads (the
generator is deterministic per pair, so the inputs are regenerated rather
than stored; the manifest keeps a SHA-256 of the input bytes so that a change
of the generator is caught) together with the oracle's six outputs per pair.
The reference has no tests or vectors for this path, and its own build needs a
header it does not ship, so the expected outputs come from the oracle; the
fixtures freeze them, so that a later change to the oracle or the engine
that alters any output is caught.

Run from the repository root after building:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

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

# (name, workload, prefix length, bits, w, selection)
#   selection "all": every pair of the prefix; "w=<v>": pairs whose workload-3
#   band class is v; "8bit": pairs with len1, len2 <= 127.
CASES = [
    ("workload1_16", 1, 2048, 16, 100, "all"),
    ("workload2_16", 2, 2048, 16, 100, "all"),
    ("workload2_8", 2, 4096, 8, 100, "8bit"),
    ("workload3_w50", 3, 1600, 16, 50, "w=50"),
    ("workload3_w100", 3, 1600, 16, 100, "w=100"),
    ("workload3_w200", 3, 1600, 16, 200, "w=200"),
    ("workload3_w500", 3, 1600, 16, 500, "w=500"),
    ("workload5_16", 5, 2048, 16, 100, "all"),
]


def case_inputs(cfg, m, sel):
    """Pairs of the case (outputs zeroed), ref, qer."""
    pairs, ref, qer, w = S.generate(cfg, m)
    if sel == "8bit":
        idx = np.where((pairs["len1"] <= 127) & (pairs["len2"] <= 127))[0]
    elif sel.startswith("w="):
        idx = np.where(w == int(sel[2:]))[0]
    else:
        idx = np.arange(m)
    return pairs[idx].copy(), ref, qer


def input_digest(pairs, ref, qer):
    h = hashlib.sha256()
    for f in IN_FIELDS:
        h.update(np.ascontiguousarray(pairs[f], "<i4").tobytes())
    h.update(ref.tobytes())
    h.update(qer.tobytes())
    return h.hexdigest()


def output_matrix(pairs):
    return np.stack([pairs[f].astype("<i4") for f in OUT_FIELDS], axis=1)


def oracle_outputs(pairs, ref, qer, bits, w):
    o = O.Oracle(**PARAMS)
    out = (o.scores16 if bits == 16 else o.scores8)(pairs.copy(), ref, qer, w)
    o.close()
    return output_matrix(out)


def main():
    manifest = {"params": PARAMS, "out_fields": OUT_FIELDS, "in_fields": IN_FIELDS,
                "note": "inputs regenerate from the seeded synthetic workloads "
                        "(trans-omics-acceleration-library_amd/synth.py); each case is a "
                        "fresh object", "cases": []}
    arrays = {}
    for name, cfg, m, bits, w, sel in CASES:
        pairs, ref, qer = case_inputs(cfg, m, sel)
        outs = oracle_outputs(pairs, ref, qer, bits, w)
        arrays[name] = outs
        manifest["cases"].append({
            "name": name, "workload": cfg, "prefix": m, "bits": bits, "w": w, "select": sel,
            "pairs": int(len(pairs)), "inputs_sha256": input_digest(pairs, ref, qer),
            "outputs_sha256": hashlib.sha256(outs.tobytes()).hexdigest()})
        print(f"{name}: {len(pairs)} pairs", flush=True)
    np.savez_compressed(os.path.join(HERE, "bsw_golden.npz"), **arrays)
    with open(os.path.join(HERE, "bsw_golden.json"), "w") as f:
        json.dump(manifest, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
