"""A caller that writes the 3-D arrays through cm1g_field_ptr between steps.

After solve3 the first RK stage may read ua, va, wa, ppi (and, after a
step that exchanged the a-array and 3-D-array buffers, every a-array) in
place of u3d, v3d, w3d, pp3d, because they are equal.  A host holding the
device address of u3d or pp3d may break that equality; the library must
then step exactly as the reference does, which uses u3d and pp3d as they
are (src/solve2.F, src/sound.F: no copies before the first stage).  Here
u3d and pp3d are perturbed through their device addresses after two steps,
the oracle's u3d and pp3d get the same perturbation, and one more step
must agree bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from cm1_amd import _native as N
from common import PROG, case, interior, make_gpu, make_oracle

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemcpy.restype = C.c_int
    lib.hipDeviceSynchronize.restype = C.c_int
    return lib


def perturb_through_pointer(g, name, delta):
    """Add delta (an array over the field's interior box) to the device field
    through the address cm1g_field_ptr hands out; returns that address."""
    p = C.c_void_p()
    N.check(g.L.cm1g_field_ptr(g.h, N.FID[name], C.byref(p)), "cm1g_field_ptr")
    lib = hip()
    n = g.px * g.py * g.pz
    buf = np.empty(n, np.float32)
    assert lib.hipDeviceSynchronize() == 0
    assert lib.hipMemcpy(buf.ctypes.data, p.value, n * 4, D2H) == 0
    box = buf.reshape(g.pz, g.py, g.px)
    nk, nj, ni = delta.shape
    # element (i,j,k) of the layout sits at (i + CM1G_IOFF) + (j+2)*px + k*px*py
    box[1:nk + 1, 3:nj + 3, 32:ni + 32] += delta
    assert lib.hipMemcpy(p.value, buf.ctypes.data, n * 4, H2D) == 0
    assert lib.hipDeviceSynchronize() == 0
    return p.value


@pytest.mark.parametrize("bc", [2, 1], ids=["open", "periodic"])
def test_writes_through_field_ptr_between_steps(require_gpu, bc):
    cfgd, cols, fields = case(ni=40, nj=36, nk=30, amp=3.0, bc=bc)
    g = make_gpu(cfgd, cols, fields)
    o = make_oracle(cfgd, cols, fields)
    g.step(2)
    o.step(2)
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    rng = np.random.default_rng(11)
    du = (rng.standard_normal((nk, nj, ni + 1)) * 0.25).astype(np.float32)
    dp = (rng.standard_normal((nk, nj, ni)) * 1.0e-4).astype(np.float32)
    # the same perturbation on both sides, in float32 arithmetic
    ou = interior(o.f("u3d"), nk, nj, ni, "u")
    ou[...] = ou + du
    op = interior(o.f("pp3d"), nk, nj, ni)
    op[...] = op + dp
    perturb_through_pointer(g, "u3d", du)
    perturb_through_pointer(g, "pp3d", dp)
    assert np.array_equal(g.interior("u3d", "u").view(np.uint32),
                          interior(o.f("u3d"), nk, nj, ni, "u").view(np.uint32))
    g.step(1)
    o.step(1)
    for name, stag in PROG:
        a = g.interior(name, stag)
        b = interior(o.f(name), nk, nj, ni, stag)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    # the perturbation reached the state: the step differs from an
    # unperturbed one
    h = make_gpu(cfgd, cols, fields)
    h.step(3)
    assert not np.array_equal(h.interior("ua", "u"), g.interior("ua", "u"))
    g.close()
    h.close()
