"""Device maxmin (cm1g_maxmin) against the CPU restatement of src/maxmin.F
on fields with tied extremes, a NaN, and each staggering."""
import numpy as np
import pytest

import maxmin_ref as R
from common import case, make_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("stag", [None, "u", "v", "w"])
def test_maxmin_matches_reference_order(require_gpu, stag):
    cfgd, cols, fields = case(ni=24, nj=20, nk=12)
    g = make_gpu(cfgd, cols, fields)
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    izz, jzz, kzz = ni + (stag == "u"), nj + (stag == "v"), nk + (stag == "w")
    rng = np.random.default_rng(7)
    box = np.full((g.pz, g.py, ni + 7), 1.0e6, np.float32)  # halo values must not count
    inner = rng.integers(-50, 51, size=(kzz, jzz, izz)).astype(np.float32) * 0.25
    inner[3, 4, 5] = 99.0
    inner[2, 9, 7] = 99.0      # the tie at smaller k wins
    inner[5, 1, 1] = -77.0
    inner[5, 0, 3] = -77.0     # same k: smaller j wins
    inner[0, 0, 0] = np.nan
    box[1:kzz + 1, 3:jzz + 3, 3:izz + 3] = inner
    g.put("dum3", box)
    (vmax, lmax), (vmin, lmin) = g.maxmin("dum3", stag)
    f = np.zeros((kzz + 1, jzz + 1, izz + 1), np.float32)
    f[1:, 1:, 1:] = inner
    (rmax, rlmax), (rmin, rlmin) = R.maxmin(f, izz, jzz, kzz)
    assert (vmax, lmax) == (rmax, rlmax) == (99.0, (8, 10, 3))
    assert (vmin, lmin) == (rmin, rlmin) == (-77.0, (4, 1, 6))


def test_maxmin_of_a_stepped_field(require_gpu):
    cfgd, cols, fields = case(ni=40, nj=36, nk=30, amp=3.0)
    g = make_gpu(cfgd, cols, fields)
    g.step(3)
    for name, stag in (("wa", "w"), ("ua", "u"), ("tha", None)):
        a = g.interior(name, stag)
        f = np.zeros((a.shape[0] + 1, a.shape[1] + 1, a.shape[2] + 1), np.float32)
        f[1:, 1:, 1:] = a
        want = R.maxmin(f, a.shape[2], a.shape[1], a.shape[0])
        got = g.maxmin(name, stag)
        assert got == want, (name, got, want)


def test_maxmin_keeps_the_sign_of_a_zero_extreme(require_gpu):
    """maxmin reports the value at the point its scan met first: with every
    value zero and -0.0 at (1,1,1), both extremes are -0.0 there; with -0.0
    elsewhere, +0.0 at (1,1,1)."""
    cfgd, cols, fields = case(ni=24, nj=20, nk=12)
    g = make_gpu(cfgd, cols, fields)
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    for pos in ((0, 0, 0), (3, 4, 5)):
        box = np.zeros((g.pz, g.py, ni + 7), np.float32)
        inner = np.zeros((nk, nj, ni), np.float32)
        inner[pos] = -0.0
        box[1:nk + 1, 3:nj + 3, 3:ni + 3] = inner
        g.put("dum3", box)
        (vmax, lmax), (vmin, lmin) = g.maxmin("dum3")
        f = np.zeros((nk + 1, nj + 1, ni + 1), np.float32)
        f[1:, 1:, 1:] = inner
        (rmax, rlmax), (rmin, rlmin) = R.maxmin(f, ni, nj, nk)
        assert lmax == rlmax == (1, 1, 1) and lmin == rlmin == (1, 1, 1)
        first_negative = pos == (0, 0, 0)
        assert np.signbit(vmax) == np.signbit(rmax) == first_negative
        assert np.signbit(vmin) == np.signbit(rmin) == first_negative
