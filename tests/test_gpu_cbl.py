"""The dry convective-boundary-layer configuration (doubly periodic,
dx=dy=dz=20 m, explicit acoustic solver, WENO5 scalar advection on every
stage, a mixed layer under an 8 K/km inversion, six Gaussian tracer blobs)
on a reduced grid: the GPU against the oracle, bit for bit, and the dry
mass kept by the pi' adjustment."""
import numpy as np
import pytest

from cm1_amd import init as hinit
from common import interior, make_gpu, make_oracle

pytestmark = pytest.mark.gpu


def cbl_case(ni=64, nj=48, nk=40):
    cols, fields, dzn = hinit.cbl_state(ni, nj, nk, 20.0, 20.0, 20.0, seed=3)
    cfgd = dict(ni=ni, nj=nj, nk=nk, wbc=1, ebc=1, sbc=1, nbc=1, imoist=0, ptype=0,
                psolver=2, advwenos=1, advwenov=0, apmasscon=1, irbc=4, nsound=10, npt=6,
                pdtra=1, dx=20.0, dy=20.0, dz=dzn, dt=0.25, alph=0.6, kdiv=0.1, csmax=350.0,
                irdamp=0, rdalpha=3.3333333333e-3, zd=0.9 * nk * 20.0)
    return cfgd, cols, fields


def test_cbl_matches_oracle(require_gpu):
    cfgd, cols, fields = cbl_case()
    g = make_gpu(cfgd, cols, fields)
    o = make_oracle(cfgd, cols, fields)
    tm0, _ = g.budgets()
    g.step(4)
    o.step(4)
    nk, nj, ni = cfgd["nk"], cfgd["nj"], cfgd["ni"]
    names = [("ua", "u"), ("va", "v"), ("wa", "w"), ("tha", None), ("ppi", None), ("rho", None),
             ("prs", None)] + [(f"pta{t}", None) for t in range(6)]
    for name, stag in names:
        a = g.interior(name, stag)
        b = interior(o.f(name), nk, nj, ni, stag)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    tm1, _ = g.budgets()
    assert abs(tm1 - tm0) <= 1.0e-6 * tm0, (tm0, tm1)
    for t in range(6):
        assert g.interior(f"pta{t}").min() >= 0.0


@pytest.mark.parametrize("doimpl", [1, 0])
def test_cbl_les_matches_oracle(require_gpu, doimpl):
    """The same reduced CBL with the workload's subgrid turbulence (TKE
    scheme, semi-slip surface, 0.24 K m/s surface heat flux; implicit
    vertical diffusion as in the workload, and explicit) and an initial TKE
    of up to 0.6 m^2/s^2 below 600 m, so that the TKE's length scale,
    production, transport and advection all act from the first step: the
    GPU equals the oracle bit for bit after 12 steps, the TKE, the eddy
    viscosities and the six tracers (diffused with theta in one pass)
    included."""
    from cm1_amd.workloads import CBL_TURB
    cfgd, cols, fields = cbl_case()
    cfgd.update(CBL_TURB)
    cfgd["doimpl"] = doimpl
    nk, nj, ni = cfgd["nk"], cfgd["nj"], cfgd["ni"]
    tke = hinit.blank(ni, nj, nk)
    zf = cols["zf"]
    rng = np.random.default_rng(5)
    for k in range(2, nk + 1):
        if zf[k] < 600.0:
            tke[k, 3:nj + 3, 3:ni + 3] = (0.3 * (1.0 + rng.uniform(-1.0, 1.0, (nj, ni)))
                                          * (1.0 - zf[k] / 600.0)).astype(np.float32)
    fields = dict(fields, tkea=tke)
    g = make_gpu(cfgd, cols, fields)
    o = make_oracle(cfgd, cols, fields)
    g.step(12)
    o.step(12)
    names = [("ua", "u"), ("va", "v"), ("wa", "w"), ("tha", None), ("ppi", None), ("rho", None),
             ("prs", None), ("tkea", "w"), ("kmh", "w"), ("khv", "w"), ("lenscl", "w")] + \
        [(f"pta{t}", None) for t in range(6)]
    for name, stag in names:
        a = g.interior(name, stag)
        b = interior(o.f(name), nk, nj, ni, stag)
        nb = int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32)))
        assert nb == 0, (name, nb)
    assert float(g.interior("tkea", "w").max()) > 0.05
    g.close()
