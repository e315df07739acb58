"""Large steps replayed as captured HIP graphs (cm1g_set_graphs, the default)
against the same steps launched kernel by kernel.

The scalar updates exchange buffers with a partner array, so consecutive
steps start from different buffer arrangements and replay different graphs;
the runs below cover several of them, with host work and synchronisations
between the replays (every `get` waits for the stream), and compare every
prognostic array over the whole padded box after every step, bit for bit.
"""
import numpy as np
import pytest

from common import PROG, case, make_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kw", [dict(bc=2, npt=2, irdamp=1, zd=10000.0),
                                dict(bc=1, psolver=2, nsound=8, advwenos=1, advwenov=1),
                                dict(bc=2, imoist=0)],
                         ids=["open-moist-tracers", "periodic-sounde-weno", "open-dry"])
def test_graph_steps_equal_eager(require_gpu, kw):
    cfgd, cols, fields = case(ni=40, nj=36, nk=30, amp=3.0, **kw)
    eager = make_gpu(cfgd, cols, fields)
    eager.set_graphs(False)
    graph = make_gpu(cfgd, cols, fields)
    names = [n for n, _ in PROG] + ["u3d", "v3d", "w3d", "th3d", "pp3d", "rho", "prs", "ppx"]
    names += [f"pta{t}" for t in range(cfgd["npt"])]
    if not cfgd["imoist"]:
        names = [n for n in names if not n.startswith("qa")]
    for step in range(7):
        ce = eager.step(1)
        cg = graph.step(1)
        assert ce == cg, (step, ce, cg)
        for n in names:
            a, b = eager.get(n), graph.get(n)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (step, n)
    assert eager.scalars() == graph.scalars()


def test_graphs_off_and_on_again(require_gpu):
    """Switching graphs off drops the captured steps; switching back on
    captures afresh from the current arrangement."""
    cfgd, cols, fields = case(ni=24, nj=20, nk=16, amp=3.0)
    a = make_gpu(cfgd, cols, fields)
    b = make_gpu(cfgd, cols, fields)
    b.set_graphs(False)
    a.step(2)
    a.set_graphs(False)
    a.step(1)
    a.set_graphs(True)
    a.step(2)
    b.step(5)
    for n, _ in PROG:
        assert np.array_equal(a.get(n).view(np.uint32), b.get(n).view(np.uint32)), n
