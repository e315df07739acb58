"""The CPU oracle against the reference CM1 build (FP32).

* Golden vectors (tests/golden/*.npz, written by the reference executable
  via tests/golden/make_ref_fixture.py): from the reference's own t=0
  state, the oracle must reproduce its state after 10 steps bit for bit.
* Live (when oracle/_ref/cm1_sp.exe is built): the 64x64x40 Kessler
  supercell of the reference for 60 steps, bit for bit.
"""
import glob
import os

import numpy as np
import pytest

from common import PROG, interior, load_fixture, make_oracle, nsound_for

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_*.npz")))
STAG = dict(PROG)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_matches_reference_fixture(path):
    cfgd, cols, fields, nsteps, target, z = load_fixture(path)
    cfgd["nsound"] = nsound_for(cfgd)
    o = make_oracle(cfgd, cols, fields)
    # base state derived by the oracle equals the reference's prs0
    assert np.array_equal(o.kf("prs0")[1:cfgd["nk"] + 1], z["prs0"])
    # t=0 density from getset equals the reference's
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    assert np.array_equal(interior(o.f("rho"), nk, nj, ni), z["t0_rho"])
    o.step(nsteps)
    for name, ref in target.items():
        got = interior(o.f(name), nk, nj, ni, STAG[name])
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), (name, int(np.count_nonzero(got != ref)))


REF_EXE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "cm1_sp.exe")


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="reference executable not built")
@pytest.mark.parametrize("bc", [2, 1])
def test_oracle_matches_reference_live(bc, tmp_path):
    import refrun
    from common import load_fixture as _unused  # noqa: F401
    import tempfile
    nsteps = 60 if bc == 2 else 20
    ni, nj, nk, dt = 64, 64, 40, 6.0
    wd = str(tmp_path / "ref")
    refrun.run_reference(wd, dict(nx=ni, ny=nj, nz=nk, dtl=dt, timax=nsteps * dt,
                                  tapfrq=nsteps * dt, wbc=bc, ebc=bc, sbc=bc, nbc=bc),
                         threads=4)
    S = refrun.read_grads(wd, "cm1out_s.ctl")
    U = refrun.read_grads(wd, "cm1out_u.ctl")
    V = refrun.read_grads(wd, "cm1out_v.ctl")
    W = refrun.read_grads(wd, "cm1out_w.ctl")
    fx = os.path.join(tempfile.mkdtemp(), "live.npz")
    out = dict(meta=np.array([ni, nj, nk, nsteps, bc], np.int32), dt=np.array([dt], np.float32))
    for t, tag in ((0, "t0"), (1, "tN")):
        out.update({f"{tag}_u": U["u"][t], f"{tag}_v": V["v"][t], f"{tag}_w": W["w"][t],
                    f"{tag}_th": S["thpert"][t], f"{tag}_pi": S["pipert"][t],
                    f"{tag}_qv": S["qv"][t], f"{tag}_qc": S["qc"][t], f"{tag}_qr": S["qr"][t],
                    f"{tag}_rho": S["rho"][t], f"{tag}_prs": S["prs"][t]})
    for n in ("pi0", "th0", "qv0", "prs0"):
        out[n] = S[n][0, :, 0, 0]
    np.savez(fx, **out)
    test_oracle_matches_reference_fixture(fx)
