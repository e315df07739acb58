"""Ensembles (cm1g_ensemble_*): members of different grid sizes, soundings
and bubbles advanced together, one member-indexed launch per kernel of the
step, replayed as one HIP graph.  Every member must equal the same model
stepped on its own, bit for bit, over the whole padded box, whatever the
column-tile height."""
import numpy as np
import pytest

from cm1_amd import model
from common import PROG, case, make_gpu

pytestmark = pytest.mark.gpu

SHAPES = [dict(ni=24, nj=20, nk=16, umax=10.0, amp=1.5, seed=3),
          dict(ni=40, nj=36, nk=30, umax=25.0, amp=3.0, seed=4),
          dict(ni=70, nj=18, nk=22, umax=18.0, amp=2.0, seed=5, bfx=0.3)]
NAMES = [n for n, _ in PROG] + ["u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "rr", "rf"]


def members(common):
    out = []
    for s in SHAPES:
        kw = dict(s)
        kw.update(common)
        out.append(case(**kw))
    return out


def check_equal(ens_models, solo_models, names):
    for a, b in zip(ens_models, solo_models):
        for n in names:
            x, y = a.get(n), b.get(n)
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (a.ni, n)
        assert a.scalars() == b.scalars()


@pytest.mark.parametrize("common,ty", [(dict(bc=2), 4), (dict(bc=2), 1),
                                       (dict(bc=1, npt=2, irdamp=1), 2),
                                       (dict(bc=1, psolver=2, nsound=8, advwenos=1, advwenov=1), 4)],
                         ids=["open-ty4", "open-ty1", "periodic-tracers-ty2", "sounde-weno-ty4"])
def test_ensemble_equals_members_alone(require_gpu, common, ty):
    cases = members(common)
    ens_models = [make_gpu(*c) for c in cases]
    solo = [make_gpu(*c) for c in cases]
    e = model.Ensemble(ens_models)
    e.set_tiles(ty)
    names = NAMES + [f"pta{t}" for t in range(common.get("npt", 0))]
    for step in range(3):
        e.step(2)
        for s in solo:
            s.step(2)
        e.sync()
        check_equal(ens_models, solo, names)
    e.close()


@pytest.mark.parametrize("order", [1, 2], ids=["deepest-first", "largest-first"])
def test_member_order_changes_no_result(require_gpu, order):
    """The order of the members' workgroups in the merged launches is a
    schedule choice only: every member still equals itself stepped alone."""
    cases = members(dict(bc=2))
    ens_models = [make_gpu(*c) for c in cases]
    solo = [make_gpu(*c) for c in cases]
    e = model.Ensemble(ens_models)
    e.set_order(order)
    e.step(3)
    for s in solo:
        s.step(3)
    e.sync()
    check_equal(ens_models, solo, NAMES)
    e.close()


def test_schedule_tuner(require_gpu):
    """The autotuner over tile height and member order takes ordinary steps
    (10 * (steps + 1) of them) and keeps one of the nine schedules."""
    cases = members(dict(bc=2))
    ens_models = [make_gpu(*c) for c in cases]
    solo = [make_gpu(*c) for c in cases]
    e = model.Ensemble(ens_models)
    ty, od, ms = e.tune_schedule(1)
    assert ty in (1, 2, 4) and od in (0, 1, 2) and len(ms) == 9 and all(v > 0.0 for v in ms.values())
    for s in solo:
        s.step(20)
    e.sync()
    check_equal(ens_models, solo, NAMES)
    e.close()


def test_ensemble_tune(require_gpu):
    cases = members(dict(bc=2))
    ens_models = [make_gpu(*c) for c in cases]
    solo = [make_gpu(*c) for c in cases]
    e = model.Ensemble(ens_models)
    best, ms = e.tune(2)
    assert best in (1, 2, 4) and all(v > 0.0 for v in ms.values()), (best, ms)
    # the tuner's 4 * (2 + 1) steps are ordinary steps of the run
    for s in solo:
        s.step(12)
    e.sync()
    check_equal(ens_models, solo, NAMES)
    e.close()


def test_mismatched_members_refused(require_gpu):
    a = make_gpu(*case(ni=24, nj=20, nk=16))
    b = make_gpu(*case(ni=24, nj=20, nk=16, imoist=0))
    before = a.get("ua")
    e = model.Ensemble([a, b])
    from cm1_amd import _native as N
    with pytest.raises(N.CM1GError):
        e.step(1)
    e.sync()
    # nothing ran: the members are where they were and still step alone
    assert np.array_equal(a.get("ua"), before)
    a.step(1)
    e.close()
