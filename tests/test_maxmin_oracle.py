"""The CPU restatement of maxmin (oracle/maxmin_ref.py) against what fixes
it independently: the extreme values are numpy's, and among equal
extremes the location is the smallest (k, j, i)."""
import numpy as np

import maxmin_ref as R


def lexi_first(a, v):
    ks, js, is_ = np.nonzero(a == v)
    order = np.lexsort((is_, js, ks))
    q = order[0]
    return int(is_[q]) + 1, int(js[q]) + 1, int(ks[q]) + 1


def test_maxmin_ref_random_with_ties():
    rng = np.random.default_rng(2)
    for trial in range(20):
        nk, nj, ni = rng.integers(1, 6, size=3)
        a = rng.integers(-3, 4, size=(nk, nj, ni)).astype(np.float32)  # many ties
        f = np.zeros((nk + 1, nj + 1, ni + 1), np.float32)
        f[1:, 1:, 1:] = a
        (vmax, lmax), (vmin, lmin) = R.maxmin(f, ni, nj, nk)
        assert vmax == a.max() and vmin == a.min()
        assert lmax == lexi_first(a, a.max()) and lmin == lexi_first(a, a.min()), trial
