"""The synthetic workload generator is deterministic and has the stated
shapes."""
import numpy as np

import synth as S


def test_prefix_and_sample_regenerate_identically():
    a = S.generate(2, 300)
    b = S.generate(2, 100, first=200)
    np.testing.assert_array_equal(a[0]["len1"][200:], b[0]["len1"])
    for k in range(100):
        pa, pb = a[0][200 + k], b[0][k]
        np.testing.assert_array_equal(a[1][pa["idr"]:pa["idr"] + pa["len1"]],
                                      b[1][pb["idr"]:pb["idr"] + pb["len1"]])
        np.testing.assert_array_equal(a[2][pa["idq"]:pa["idq"] + pa["len2"]],
                                      b[2][pb["idq"]:pb["idq"] + pb["len2"]])


def test_shapes():
    l1, l2, h0, w = S.shape(1, 5000)
    assert (l2 == 151).all() and l1.min() >= 151 and l1.max() <= 250
    assert h0.min() >= 10 and h0.max() <= 60
    l1, l2, h0, w = S.shape(2, 5000)
    assert l2.min() >= 30 and l2.max() <= 250 and ((l1 - l2) >= 0).all() and ((l1 - l2) <= 100).all()
    l1, l2, h0, w = S.shape(3, 5000)
    assert l2.min() >= 100 and l2.max() <= 2000 and set(np.unique(w)) == {50, 100, 200, 500}
    assert 330 < np.median(l2) < 470
    l1, l2, h0, w = S.shape(5, 5000)
    assert h0.min() >= 10 and h0.max() <= 30000


def test_codes():
    p, r, q, _ = S.generate(5, 2000)
    assert r.max() <= 4 and q.max() <= 4
    assert (r == 4).any()
    p, r, q, _ = S.generate(1, 2000)
    assert r.max() <= 3 and q.max() <= 3
