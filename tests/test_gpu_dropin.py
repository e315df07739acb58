"""The CM1 model program itself with its dynamical-core modules replaced by
the GPU drop-ins (cm1_amd/fortran/dropin/: adv_module, sound_module,
sounde_module), built by cm1_amd/fortran/build_dropin.py from the
reference's own sources with solve2.F, mp_driver.F and cm1.F unchanged.

The program runs the namelist a golden fixture was made with, and its
GrADS output (u, v, w, theta', pi', qv, qc, qr, rho, prs after N steps)
must equal the reference executable's output stored in the fixture, bit
for bit."""
import os

import numpy as np
import pytest

import refrun

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cm1_amd", "fortran", "_cm1", "cm1_gpu.exe")
GOLDEN = os.path.join(ROOT, "tests", "golden")

FIXTURES = ["ref_config0_64x64x40_10", "ref_periodic_20x24x18_10", "ref_open_weno_24x20x18_10",
            "ref_open_sounde_weno2_24x20x18_10", "ref_open_rdamp_24x20x18_10",
            "ref_stretch_dt3_24x20x96_10"]


def overrides_of(z):
    """The namelist values the fixture's reference run used
    (tests/golden/make_ref_fixture.py)."""
    ni, nj, nk, nsteps, bc = [int(v) for v in z["meta"]]
    dt = float(z["dt"][0])
    opts = [int(v) for v in z["opts"]] if "opts" in z else []
    psolver, advwenos, advwenov, irdamp, npt = (opts + [3, 0, 0, 0, 0][len(opts):])[:5]
    rdalpha, zd = [float(v) for v in z["rdamp"]] if "rdamp" in z else (3.3333333333e-3, 15000.0)
    g = [float(v) for v in z["grid"]] if "grid" in z else [1000.0, 1000.0, 500.0, 0.0]
    ov = dict(nx=ni, ny=nj, nz=nk, dtl=dt, timax=nsteps * dt, tapfrq=nsteps * dt, wbc=bc, ebc=bc,
              sbc=bc, nbc=bc, rdalpha=rdalpha, zd=zd, dx=g[0], dy=g[1], dz=g[2], psolver=psolver,
              advwenos=advwenos, advwenov=advwenov, irdamp=irdamp)
    if int(g[3]) == 1:
        ov.update(stretch_z=1, dz_bot=g[4], dz_top=g[5], str_bot=g[6], str_top=g[7], ztop=g[8])
    if npt > 0:
        ov.update(iptra=1, npt=npt)
    return ov


@pytest.mark.parametrize("name", FIXTURES)
def test_dropin_program_equals_reference(require_gpu, name, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("drop-in program not built (cm1_amd/fortran/build_dropin.py needs the reference sources)")
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    work = str(tmp_path / "run")
    print(f"drop-in run {name}", flush=True)
    out, wall = refrun.run_reference(work, overrides_of(z), threads=1, exe=EXE, timeout=600)
    S = refrun.read_grads(work, "cm1out_s.ctl")
    U = refrun.read_grads(work, "cm1out_u.ctl")
    V = refrun.read_grads(work, "cm1out_v.ctl")
    W = refrun.read_grads(work, "cm1out_w.ctl")
    got = {"u": U["u"], "v": V["v"], "w": W["w"], "th": S["thpert"], "pi": S["pipert"], "qv": S["qv"],
           "qc": S["qc"], "qr": S["qr"], "rho": S["rho"], "prs": S["prs"]}
    bad = {}
    for key, arr in got.items():
        for t, tag in ((0, "t0"), (1, "tN")):
            want = z[f"{tag}_{key}"]
            n = int(np.count_nonzero(arr[t].view(np.uint32) != want.view(np.uint32)))
            if n:
                bad[f"{tag}_{key}"] = n
    print(f"{name}: {wall:.1f} s; differing values {bad}")
    assert not bad, bad


def test_dropin_restart_round_trip(require_gpu, tmp_path):
    """The reference's restart files (src/restart_write.F, restart_format 1)
    in the drop-in program: a continuous 10-step run writes restarts at 30 s
    and 60 s equal, in every field record (the s, u, v, w files), to the
    reference executable's; a run restarted (irst = 1) from the reference
    executable's 30 s restart writes a 60 s restart equal to them too."""
    if not os.path.exists(EXE):
        pytest.skip("drop-in program not built (cm1_amd/fortran/build_dropin.py needs the reference sources)")
    if not os.path.exists(refrun.REF_EXE):
        pytest.skip("reference executable not built")
    import shutil
    import subprocess
    ov = dict(nx=24, ny=20, nz=18, dtl=6.0, timax=60.0, tapfrq=60.0, rstfrq=30.0, wbc=2, ebc=2, sbc=2,
              nbc=2)
    ref, gpu, rst = str(tmp_path / "ref"), str(tmp_path / "gpu"), str(tmp_path / "rst")
    refrun.run_reference(ref, ov, threads=1)
    refrun.run_reference(gpu, ov, threads=1, exe=EXE)

    def same(a, b, n, kind):
        fa = os.path.join(a, f"cm1rst_{n:06d}_{kind}.dat")
        fb = os.path.join(b, f"cm1rst_{n:06d}_{kind}.dat")
        with open(fa, "rb") as x, open(fb, "rb") as y:
            return x.read() == y.read()

    for n in (1, 2):
        for kind in "suvw":
            assert same(ref, gpu, n, kind), (n, kind)
    # restart the drop-in program from the reference's 30 s files
    os.makedirs(rst)
    for kind in "xsuvw":
        shutil.copy(os.path.join(ref, f"cm1rst_000001_{kind}.dat"), rst)
    with open(os.path.join(rst, "namelist.input"), "w") as f:
        f.write(refrun.namelist_text(dict(ov, irst=1, rstnum=1)))
    r = subprocess.run([EXE], cwd=rst, capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, OMP_NUM_THREADS="1", OMP_STACKSIZE="512M"))
    assert r.returncode == 0 and "Program terminated normally" in r.stdout + r.stderr, (r.stdout + r.stderr)[-2000:]
    for kind in "suvw":
        assert same(ref, rst, 2, kind), kind


LES = ["impl1", "impl0"]


@pytest.mark.parametrize("impl", LES)
def test_dropin_program_les_equals_reference(require_gpu, impl, tmp_path):
    """The drop-in program with turb_module and turbtend_module replaced as
    well (sfc_and_turb, turbs, turbt, turbu, turbv, turbw on the GPU, advw's
    TKE call through cm1g_advw_scalar; the first sfc_and_turb after the
    restart keeps the restart's surface values, getsfc = .false.) on the dry
    LES namelist of the fixture, restarted from
    the reference executable's restart files after 300 steps (the TKE is
    active from there on): after the last 100 steps its GrADS output (u, v,
    w, theta', pi', TKE, rho, prs, the tracers, km, kh) equals the reference
    executable's bit for bit."""
    if not os.path.exists(EXE):
        pytest.skip("drop-in program not built (cm1_amd/fortran/build_dropin.py needs the reference sources)")
    import sys
    sys.path.insert(0, GOLDEN)
    import make_ref_fixture as mrf
    z = np.load(os.path.join(GOLDEN, f"ref_les_tke_{impl}_16x12x24_400.npz"))
    rst = np.load(os.path.join(GOLDEN, f"les_rst_{impl}_16x12x24_300.npz"))
    ni, nj, nk, nsteps, bc, nout = [int(v) for v in z["meta"]]
    dt = float(z["dt"][0])
    psolver, advwenos, _, _, npt, doimpl = [int(v) for v in z["opts"]]
    dx, _, dz = [float(v) for v in z["grid"]]
    ov = dict(nx=ni, ny=nj, nz=nk, dtl=dt, timax=nsteps * dt, tapfrq=nsteps * dt / nout,
              wbc=1, ebc=1, sbc=1, nbc=1, dx=dx, dy=dx, dz=dz, psolver=psolver,
              advwenos=advwenos, advwenov=0, irdamp=0, doimpl=doimpl, iptra=1, npt=npt,
              irst=1, rstnum=1, **mrf.TURB_NAMELIST)
    ov.update({"output_" + k: 1 for k in mrf.TURB_OUT})
    files = {k[:-4] + ".dat": rst[k].tobytes() for k in rst.files if k.startswith("cm1rst")}
    files["LANDUSE.TBL"] = np.load(os.path.join(GOLDEN, "les_inputs.npz"))["landuse_tbl"].tobytes()
    work = str(tmp_path / "run")
    print(f"drop-in LES run {impl}, from step {int(rst['meta'][0])}", flush=True)
    out, wall = refrun.run_reference(work, ov, threads=1, exe=EXE, timeout=600, files=files)
    S = refrun.read_grads(work, "cm1out_s.ctl")
    U = refrun.read_grads(work, "cm1out_u.ctl")
    V = refrun.read_grads(work, "cm1out_v.ctl")
    W = refrun.read_grads(work, "cm1out_w.ctl")
    got = {"u": U["u"], "v": V["v"], "w": W["w"], "th": S["thpert"], "pi": S["pipert"], "tke": W["tke"],
           "rho": S["rho"], "prs": S["prs"], "kmh": W["kmh"], "khv": W["khv"]}
    for m in range(1, npt + 1):
        got[f"pt{m}"] = S[f"pt{m}"]
    bad = {}
    for key, arr in got.items():
        want = z[f"o{nout}_{key}"]
        c = int(np.count_nonzero(arr[nout].view(np.uint32) != want.view(np.uint32)))
        if c:
            bad[key] = c
    print(f"drop-in LES {impl}: {wall:.1f} s; differing values {bad}", flush=True)
    assert not bad, bad
    assert float(z[f"o{nout}_tke"].max()) > 0.1
