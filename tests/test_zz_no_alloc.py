"""cm1g_maxmin_noalloc (no counterpart in the reference): maxmin of a device
field for callers that cannot allocate, with its partial results in the
library's file-scope device array.  Checked through the C ABI against
cm1g_maxmin on the same fields, bit for bit in value and position, on two
calls in a row (the second on a smaller grid, so the array still holds the
first call's larger set of partials), and on a field with ties and a NaN."""
import ctypes as C

import numpy as np
import pytest

from cm1_amd import _native as N
from common import case, make_gpu

pytestmark = pytest.mark.gpu

STAG = {None: 0, "u": 1, "v": 2, "w": 3}


def noalloc(g, name, stag):
    L = N.lib()
    g.sync()
    p = C.c_void_p()
    N.check(L.cm1g_field_ptr(g.h, N.FID[name], C.byref(p)), "cm1g_field_ptr")
    vmax, vmin = C.c_float(), C.c_float()
    lmax, lmin = (C.c_int32 * 3)(), (C.c_int32 * 3)()
    N.check(L.cm1g_maxmin_noalloc(p, g.ni, g.nj, g.nk, STAG[stag], None, C.byref(vmax), lmax,
                                  C.byref(vmin), lmin), "cm1g_maxmin_noalloc")
    return (vmax.value, tuple(lmax)), (vmin.value, tuple(lmin))


def same(a, b):
    (va, la), (wa, ma) = a
    (vb, lb), (wb, mb) = b
    return (np.float32(va).view(np.uint32) == np.float32(vb).view(np.uint32) and la == lb
            and np.float32(wa).view(np.uint32) == np.float32(wb).view(np.uint32) and ma == mb)


def test_noalloc_equals_maxmin_two_calls(require_gpu):
    big = make_gpu(*case(ni=200, nj=150, nk=40, amp=3.0, seed=11))
    small = make_gpu(*case(ni=24, nj=20, nk=12, amp=2.0, seed=12))
    big.step(2)
    small.step(1)
    for g in (big, small):
        for name, stag in (("wa", "w"), ("ua", "u"), ("va", "v"), ("tha", None)):
            want = g.maxmin(name, stag)
            got = noalloc(g, name, stag)
            assert same(got, want), (g.ni, name, got, want)


def test_noalloc_ties_and_nan(require_gpu):
    g = make_gpu(*case(ni=24, nj=20, nk=12))
    ni, nj, nk = g.ni, g.nj, g.nk
    box = np.full((g.pz, g.py, ni + 7), 1.0e6, np.float32)
    inner = np.zeros((nk, nj, ni), np.float32)
    inner[3, 4, 5] = inner[2, 9, 7] = 99.0
    inner[5, 1, 1] = inner[5, 0, 3] = -77.0
    inner[0, 0, 0] = np.nan
    box[1:nk + 1, 3:nj + 3, 3:ni + 3] = inner
    g.put("dum3", box)
    got = noalloc(g, "dum3", None)
    assert got == ((99.0, (8, 10, 3)), (-77.0, (4, 1, 6)))
    assert same(got, g.maxmin("dum3", None))


def test_noalloc_rejects_sizes_beyond_the_limit(require_gpu):
    L = N.lib()
    vmax, vmin = C.c_float(), C.c_float()
    lmax, lmin = (C.c_int32 * 3)(), (C.c_int32 * 3)()
    g = make_gpu(*case(ni=24, nj=20, nk=12))
    g.sync()
    p = C.c_void_p()
    N.check(L.cm1g_field_ptr(g.h, N.FID["tha"], C.byref(p)), "cm1g_field_ptr")
    rc = L.cm1g_maxmin_noalloc(p, 16385, 20, 12, 0, None, C.byref(vmax), lmax, C.byref(vmin), lmin)
    assert rc == 1 and b"CM1G_NOALLOC_MAXN" in L.cm1g_last_error()
