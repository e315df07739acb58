"""Dry LES with the TKE subgrid model on the GPU against the reference.

Golden vectors (tests/golden/make_ref_fixture.py turb): the reference
executable runs the switches of its dry convective boundary-layer case
(run/config_files/les_ConvBoundLayer: TKE scheme, semi-slip surface with the
log-law friction velocity, 0.24 K m/s specified surface heat flux, mixed-
layer sounding, +-0.25 K noise) on a 16x12x24 doubly periodic grid,
dx = dz = 20 m, dt = 0.25 s, sounde, WENO5 scalars, two passive tracers, with
explicit and with implicit vertical diffusion, over 400 steps.  The
subgrid TKE stays zero for about 300 steps and then grows.

From the reference's t=0 state the device path (cm1g_set_turbulence,
cm1g_step with its turbulence kernels, captured as a graph) must equal the
reference bit for bit at every output time (u, v, w, theta', pi', TKE),
and at the end rho, prs and the tracers; the eddy viscosities and
diffusivities, length scale, N^2, deformation and dissipation computed by
cm1g_sfc_and_turb from the final state must equal the ones the reference
wrote.
"""
import glob
import os

import numpy as np
import pytest

from common import TURB_OUT, interior, load_turb_fixture, make_gpu

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
LES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_les_*.npz")))
DIAG = ("kmh", "kmv", "khh", "khv", "lenscl", "nm", "defv", "defh", "dissten")


def compare(g, z, n, keys, ni, nj, nk):
    bad = []
    for key in keys:
        name, stag = TURB_OUT[key]
        ref = z[f"o{n}_{key}"]
        got = interior(g.get(name), nk, nj, ni, stag)
        nb = int(np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32)))
        if nb:
            bad.append((key, nb, float(np.abs(got.astype(np.float64) - ref).max())))
    return bad


@pytest.mark.parametrize("path", LES, ids=[os.path.basename(p) for p in LES])
def test_les_equals_reference(require_gpu, path):
    cfgd, cols, fields, nsteps, nout, z = load_turb_fixture(path)
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    g = make_gpu(cfgd, cols, fields)
    per = nsteps // nout
    for n in range(1, nout + 1):
        g.step(per)
        bad = compare(g, z, n, ("u", "v", "w", "th", "pi", "tke"), ni, nj, nk)
        print(f"{os.path.basename(path)}: step {n * per}: {bad}", flush=True)
        assert not bad, (n, bad)
    bad = compare(g, z, nout, ("rho", "prs"), ni, nj, nk)
    for m in range(cfgd["npt"]):
        got = interior(g.get(f"pta{m}"), nk, nj, ni)
        if not np.array_equal(got, z[f"o{nout}_pt{m + 1}"]):
            bad.append((f"pt{m + 1}", int(np.count_nonzero(got != z[f"o{nout}_pt{m + 1}"]))))
    assert not bad, bad
    assert float(z[f"o{nout}_tke"].max()) > 0.1
    g.sfc_and_turb()
    bad = compare(g, z, nout, DIAG, ni, nj, nk)
    print(f"diagnostics: {bad}", flush=True)
    assert not bad, bad
    g.close()


def test_les_graphs_equal_launches(require_gpu):
    """The captured step graphs of the turbulent step replay exactly: 20
    steps from the explicit fixture's state after 330 steps with graphs and
    kernel by kernel, every prognostic field equal bit for bit."""
    path = [p for p in LES if "impl0" in p][0]
    cfgd, cols, fields, nsteps, nout, z = load_turb_fixture(path)
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    a = make_gpu(cfgd, cols, fields)
    b = make_gpu(cfgd, cols, fields)
    b.set_graphs(False)
    a.step(330)
    b.step(330)
    for _ in range(20):
        a.step(1)
        b.step(1)
    for name, stag in (("ua", "u"), ("va", "v"), ("wa", "w"), ("tha", None), ("ppi", None),
                       ("tkea", "w"), ("pta0", None)):
        x = interior(a.get(name), nk, nj, ni, stag)
        y = interior(b.get(name), nk, nj, ni, stag)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
    assert float(interior(a.get("tkea"), nk, nj, ni, "w").max()) > 0.0
    a.close()
    b.close()


def test_les_ensemble_equals_members_alone(require_gpu):
    """Two LES members (the fixture's t=0 state, and the same with theta
    perturbed below 200 m; both with an initial TKE so the scheme acts from
    the first step) advanced together by the batched ensemble equal each
    member stepped alone, bit for bit, the TKE and eddy viscosities
    included."""
    from cm1_amd import model
    path = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "ref_les_tke_impl1_*.npz")))[0]
    cfgd, cols, fields, _, _, _ = load_turb_fixture(path)
    nk, nj, ni = cfgd["nk"], cfgd["nj"], cfgd["ni"]
    rng = np.random.default_rng(11)
    tke = fields["tkea"].copy()
    tke[2:nk // 2, 3:nj + 3, 3:ni + 3] = rng.uniform(0.05, 0.5, (nk // 2 - 2, nj, ni)).astype(np.float32)
    f1 = dict(fields, tkea=tke)
    tha = fields["tha"].copy()
    tha[1:11, 3:nj + 3, 3:ni + 3] += rng.uniform(-0.2, 0.2, (10, nj, ni)).astype(np.float32)
    f2 = dict(f1, tha=tha)
    cases = [(cfgd, cols, f1), (cfgd, cols, f2)]
    ens_models = [make_gpu(*c) for c in cases]
    solo = [make_gpu(*c) for c in cases]
    e = model.Ensemble(ens_models)
    names = [("ua", "u"), ("va", "v"), ("wa", "w"), ("tha", None), ("ppi", None), ("rho", None),
             ("tkea", "w"), ("kmh", "w"), ("khv", "w")] + [(f"pta{t}", None) for t in range(cfgd["npt"])]
    for _ in range(3):
        e.step(2)
        for s in solo:
            s.step(2)
        e.sync()
        for a, b in zip(ens_models, solo):
            for n, stag in names:
                x, y = a.interior(n, stag), b.interior(n, stag)
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), n
    assert float(solo[1].interior("tkea", "w").max()) > 0.01
    e.close()
