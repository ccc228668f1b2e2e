"""Shared helpers of the test suite."""
import numpy as np

OUT_FIELDS = ["score", "tle", "gtle", "qle", "gscore", "max_off"]


def diff_mask(a, b):
    d = np.zeros(len(a), bool)
    for f in OUT_FIELDS:
        d |= a[f] != b[f]
    return d


def assert_same_outputs(got, want, label=""):
    d = diff_mask(got, want)
    if d.any():
        k = np.where(d)[0][:5]
        rows = [(int(i), [int(got[f][i]) for f in OUT_FIELDS], [int(want[f][i]) for f in OUT_FIELDS])
                for i in k]
        raise AssertionError(f"{label}: {int(d.sum())} of {len(d)} pairs differ; first: {rows}")
