"""Device-state checkpoint (cm1g_state_save / cm1g_state_load): a run that is
saved between steps and resumed in a fresh model continues bit for bit, and a
state saved from one configuration is refused by another."""
import numpy as np
import pytest

from cm1_amd import _native as N
from common import PROG, case, make_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kw", [dict(bc=2, irdamp=1, zd=10000.0), dict(bc=1, npt=2)],
                         ids=["open-rdamp", "periodic-tracers"])
def test_resume_is_bitwise(require_gpu, kw):
    cfgd, cols, fields = case(ni=40, nj=36, nk=30, amp=3.0, **kw)
    a = make_gpu(cfgd, cols, fields)
    a.step(3)
    buf = a.save_state()
    a.step(4)
    b = make_gpu(cfgd, cols, fields)
    b.step(1)                      # a different state, overwritten by the load
    b.load_state(buf)
    b.step(4)
    names = [n for n, _ in PROG] + [f"pta{t}" for t in range(cfgd["npt"])]
    for n in names:
        x, y = a.get(n), b.get(n)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), n
    sa, sb = a.scalars(), b.scalars()
    assert sa == sb, (sa, sb)
    assert a.budgets() == b.budgets()


def test_les_resume_is_bitwise(require_gpu):
    """The dry LES of the reference fixture, saved after 310 steps (TKE
    active) and resumed in a fresh model with the same turbulence set-up:
    the next 20 steps equal the uninterrupted run's bit for bit, the TKE and
    the eddy viscosities included."""
    import os
    from common import load_turb_fixture
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    cfgd, cols, fields, _, _, _ = load_turb_fixture(os.path.join(gold, "ref_les_tke_impl1_16x12x24_400.npz"))
    a = make_gpu(cfgd, cols, fields)
    a.step(310)
    buf = a.save_state()
    a.step(20)
    b = make_gpu(cfgd, cols, fields)
    b.step(1)
    b.load_state(buf)
    b.step(20)
    assert float(a.get("tkea").max()) > 1.0e-3
    names = [n for n, _ in PROG if not n.startswith("qa")] + [f"pta{t}" for t in range(cfgd["npt"])] + \
        ["tkea", "kmh", "khv"]
    for n in names:
        x, y = a.get(n), b.get(n)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), n
    assert a.scalars() == b.scalars()


def test_load_refuses_other_configuration(require_gpu):
    cfgd, cols, fields = case(ni=40, nj=36, nk=30)
    a = make_gpu(cfgd, cols, fields)
    buf = a.save_state()
    cfgd2, cols2, fields2 = case(ni=48, nj=36, nk=30)
    b = make_gpu(cfgd2, cols2, fields2)
    with pytest.raises(N.CM1GError):
        b.load_state(np.concatenate([buf, np.zeros(1 << 20, np.uint8)]))
