"""The CPU oracle against the reference CM1 build (FP32).

* Golden vectors (tests/golden/*.npz, written by the reference executable
  via tests/golden/make_ref_fixture.py): from the reference's own t=0
  state, the oracle must reproduce its state after 10 steps bit for bit.
* Live (when oracle/_ref/cm1_sp.exe is built): the 64x64x40 Kessler
  supercell of the reference, bit for bit, with the implicit (sound) and
  explicit (sounde) acoustic solvers and 5th-order or WENO5 advection.
"""
import glob
import os

import numpy as np
import pytest

from common import PROG, interior, load_fixture, make_oracle, nsound_for

HERE = os.path.dirname(os.path.abspath(__file__))
# the supercell fixtures (the LES ones, ref_les_*, have their own tests)
FIXTURES = sorted(p for p in glob.glob(os.path.join(HERE, "golden", "ref_*.npz"))
                  if not os.path.basename(p).startswith("ref_les_"))
STAG = dict(PROG)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_matches_reference_fixture(path):
    cfgd, cols, fields, nsteps, target, z = load_fixture(path)
    cfgd["nsound"] = nsound_for(cfgd, cols["zf"])
    if "zh" in z:
        # the vertical grid restated in single precision equals the
        # reference's own (its written scalar-level heights)
        assert np.array_equal(cols["zh"][1:cfgd["nk"] + 1], z["zh"])
    o = make_oracle(cfgd, cols, fields)
    # base state derived by the oracle equals the reference's prs0
    assert np.array_equal(o.kf("prs0")[1:cfgd["nk"] + 1], z["prs0"])
    # t=0 density from getset equals the reference's
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    assert np.array_equal(interior(o.f("rho"), nk, nj, ni), z["t0_rho"])
    o.step(nsteps)
    for name, ref in target.items():
        got = interior(o.f(name), nk, nj, ni, STAG.get(name))
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), (name, int(np.count_nonzero(got != ref)))


REF_EXE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "cm1_sp.exe")


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="reference executable not built")
@pytest.mark.parametrize("bc,nsteps,opts", [
    (2, 60, {}),
    (1, 20, {}),
    (1, 20, dict(psolver=2, advwenos=1, advwenov=0)),
    (2, 30, dict(psolver=3, advwenos=2, advwenov=2)),
], ids=["open-psolver3", "periodic-psolver3", "periodic-sounde-weno", "open-weno-last-stage"])
def test_oracle_matches_reference_live(bc, nsteps, opts, tmp_path):
    """64x64x40 Kessler supercell run by the reference, then by the oracle
    from the reference's t=0 state: bit for bit after nsteps."""
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_ref_fixture
    fx = make_ref_fixture.make("live", 64, 64, 40, nsteps, 6.0, bc, workdir=str(tmp_path / "ref"),
                               opts=opts, threads=4, outdir=str(tmp_path))
    test_oracle_matches_reference_fixture(fx)


LES_FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_les_*.npz")))


@pytest.mark.parametrize("path", LES_FIXTURES, ids=[os.path.basename(p) for p in LES_FIXTURES])
def test_oracle_les_matches_reference_fixture(path):
    """Dry LES with the TKE subgrid model, the semi-slip surface with the
    log-law friction velocity and the 0.24 K m/s surface heat flux
    (src/turb.F, src/turbtend.F, src/sfcphys.F), explicit and implicit
    vertical diffusion: from the reference's t=0 state the oracle equals
    the reference bit for bit at every output time over 400 steps (u, v,
    w, theta', pi', TKE; at the end rho, prs, the tracers), and its eddy
    viscosities, diffusivities, length scale, N^2, deformation and
    dissipation computed from the final state equal those the reference
    wrote.  The TKE is zero for the first 300 steps and active after."""
    from common import TURB_OUT, load_turb_fixture
    cfgd, cols, fields, nsteps, nout, z = load_turb_fixture(path)
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    o = make_oracle(cfgd, cols, fields)
    assert np.array_equal(o.kf("prs0")[1:nk + 1], z["prs0"])
    per = nsteps // nout
    for n in range(1, nout + 1):
        o.step(per)
        for key, (name, stag) in TURB_OUT.items():
            ref = z.get(f"o{n}_{key}") if f"o{n}_{key}" in z else None
            if ref is None or key in ("kmh", "kmv", "khh", "khv", "lenscl", "nm", "defv", "defh", "dissten"):
                continue
            got = interior(o.f(name), nk, nj, ni, stag)
            assert np.array_equal(got, ref), (n, key, int(np.count_nonzero(got != ref)))
        for m in range(cfgd["npt"]):
            if f"o{n}_pt{m + 1}" in z:
                assert np.array_equal(interior(o.f(f"pta{m}"), nk, nj, ni), z[f"o{n}_pt{m + 1}"])
    assert float(z[f"o{nout}_tke"].max()) > 0.1
    # the diagnostics the reference wrote come from the output time's own
    # turbulence set-up
    o.L.cm1o_set_libm(0)
    o.L.cm1o_sfc_and_turb(o.h)
    for key in ("kmh", "kmv", "khh", "khv", "lenscl", "nm", "defv", "defh", "dissten"):
        name, stag = TURB_OUT[key]
        got = interior(o.f(name), nk, nj, ni, stag)
        assert np.array_equal(got, z[f"o{nout}_{key}"]), (key, int(np.count_nonzero(got != z[f"o{nout}_{key}"])))
