"""Generate golden vectors from the reference CM1 executable.

Runs ``oracle/_ref/cm1_sp.exe`` (built from the reference's own Fortran
sources by ``oracle/build_ref.py``) on a small Kessler supercell case and
stores its GrADS output at t=0 and after N large steps as float32 arrays:

* ``t0_*``: the initial state (u, v, w, theta', pi', qv, qc, qr) and the
  base state (pi0, th0, qv0, prs0) the reference wrote at t=0;
* ``tN_*``: the same prognostics plus rho and prs after N steps.

* ``opts``: the dynamical-core switches the run used (psolver, advwenos,
  advwenov, irdamp, npt), and ``rdamp``: (rdalpha, zd) of the Rayleigh
  damper;
* with npt > 0 (iptra = 1): ``t0_ptN``/``tN_ptN``, the passive tracers the
  reference initialises (src/init3d.F:328-349) and advects;
* ``grid``: dx, dy, the namelist dz, stretch_z and (dz_bot, dz_top, str_bot,
  str_top, ztop), and ``zh``: the scalar-level heights the reference wrote
  (m), which pin the vertical-grid set-up.

``python tests/golden/make_ref_fixture.py stretched`` writes the fixtures of
the stretched 96-level grid of the 1024x1024x96 supercell (dx = 250 m) on a
small horizontal domain.

Usage: python tests/golden/make_ref_fixture.py [name ni nj nk nsteps dt wbc
[psolver advwenos advwenov [irdamp zd [npt]]]]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import refrun  # noqa: E402


OPT_KEYS = ("psolver", "advwenos", "advwenov", "irdamp", "npt")
OPT_DEFAULTS = dict(psolver=3, advwenos=0, advwenov=0, irdamp=0, npt=0)
RDALPHA = 3.3333333333e-3
STRETCH_KEYS = ("dz_bot", "dz_top", "str_bot", "str_top", "ztop")
# the vertical grid of the 1024x1024x96 supercell workload (stretch_z = 1)
CONFIG1_STRETCH = dict(dz_bot=50.0, dz_top=500.0, str_bot=500.0, str_top=17000.0, ztop=30000.0)


def make(name, ni, nj, nk, nsteps, dt, bc, workdir="/tmp/cm1_fixture", opts=None, threads=1,
         outdir=HERE, zd=15000.0, dx=1000.0, dy=1000.0, dz=500.0, stretch=None):
    """Run the reference and write <outdir>/<name>.npz.  opts: namelist
    switches of the dynamical core (psolver, advwenos, advwenov, irdamp);
    zd: the damping height when irdamp = 1; dx, dy, dz: the namelist grid
    spacings; stretch: dict(dz_bot, dz_top, str_bot, str_top, ztop) for
    stretch_z = 1 (param6)."""
    o = dict(OPT_DEFAULTS)
    o.update(opts or {})
    npt = o.pop("npt")
    ov = dict(nx=ni, ny=nj, nz=nk, dtl=dt, timax=nsteps * dt, tapfrq=nsteps * dt,
              wbc=bc, ebc=bc, sbc=bc, nbc=bc, rdalpha=RDALPHA, zd=zd, dx=dx, dy=dy, dz=dz, **o)
    if stretch is not None:
        ov.update(stretch_z=1, **{k: float(stretch[k]) for k in STRETCH_KEYS})
    if npt > 0:
        ov.update(iptra=1, npt=npt)
    o["npt"] = npt
    refrun.run_reference(workdir, ov, threads=threads)
    S = refrun.read_grads(workdir, "cm1out_s.ctl")
    U = refrun.read_grads(workdir, "cm1out_u.ctl")
    V = refrun.read_grads(workdir, "cm1out_v.ctl")
    W = refrun.read_grads(workdir, "cm1out_w.ctl")
    out = dict(meta=np.array([ni, nj, nk, nsteps, bc], np.int32), dt=np.array([dt], np.float32),
               opts=np.array([o[k] for k in OPT_KEYS], np.int32),
               rdamp=np.array([RDALPHA, zd], np.float32),
               grid=np.array([dx, dy, dz, 0 if stretch is None else 1] +
                             [0.0 if stretch is None else stretch[k] for k in STRETCH_KEYS],
                             np.float32))
    for t, tag in ((0, "t0"), (1, "tN")):
        out[f"{tag}_u"] = U["u"][t]
        out[f"{tag}_v"] = V["v"][t]
        out[f"{tag}_w"] = W["w"][t]
        out[f"{tag}_th"] = S["thpert"][t]
        out[f"{tag}_pi"] = S["pipert"][t]
        out[f"{tag}_qv"] = S["qv"][t]
        out[f"{tag}_qc"] = S["qc"][t]
        out[f"{tag}_qr"] = S["qr"][t]
        out[f"{tag}_rho"] = S["rho"][t]
        out[f"{tag}_prs"] = S["prs"][t]
        for n in range(1, npt + 1):
            out[f"{tag}_pt{n}"] = S[f"pt{n}"][t]
    for n in ("pi0", "th0", "qv0", "prs0"):
        out[n] = S[n][0, :, 0, 0]
    # the reference's scalar-level heights (written in km) pin the grid set-up
    out["zh"] = S["zhval"][0, :, 0, 0]
    path = os.path.join(outdir, f"{name}.npz")
    np.savez_compressed(path, **{k: np.ascontiguousarray(v) for k, v in out.items()})
    return path


# the reference's dry convective boundary layer switches (run/config_files/
# les_ConvBoundLayer/namelist.input) as far as the device path carries
# them: TKE subgrid model, semi-slip surface with the neutral log-law
# friction velocity over a specified roughness, specified surface heat
# flux, the mixed-layer sounding with uniform 10 m/s wind and +-0.25 K
# random theta noise
TURB_NAMELIST = dict(cm1setup=1, sgsmodel=1, tconfig=1, bcturbs=1, horizturb=0, ramp_sgs=0,
                     bbc=3, tbc=1, isfcflx=1, sfcmodel=1, set_flx=1, cnst_shflx=0.24,
                     set_znt=1, cnst_znt=0.1, lu0=3, xland0=1.0, imoist=0, ptype=0,
                     isnd=14, iwnd=6, iinit=0, irandp=1)
TURB_OUT = ("tke", "km", "kh", "nm", "def", "dissten")


def make_turb(name, ni, nj, nk, nsteps, nout, dt, doimpl, dx=20.0, dz=20.0, npt=2, psolver=2,
              advwenos=1, workdir="/tmp/cm1_fixture_turb", outdir=HERE, threads=4):
    """Dry LES with the TKE scheme: the reference from t=0 over nsteps large
    steps with output every nsteps/nout steps.  Stores the t=0 state and,
    at every output time, u, v, w, theta', pi', tke (``o{n}_*``); at the
    last time also rho, prs, the tracers and the turbulence diagnostics
    (km, kh, lenscl, nm, defv, defh, dissten)."""
    ov = dict(nx=ni, ny=nj, nz=nk, dtl=dt, timax=nsteps * dt, tapfrq=nsteps * dt / nout,
              wbc=1, ebc=1, sbc=1, nbc=1, dx=dx, dy=dx, dz=dz, psolver=psolver,
              advwenos=advwenos, advwenov=0, irdamp=0, doimpl=doimpl, iptra=1 if npt else 0,
              npt=max(npt, 1), **TURB_NAMELIST)
    ov.update({"output_" + k: 1 for k in TURB_OUT})
    refrun.run_reference(workdir, ov, threads=threads, tables=("LANDUSE.TBL",))
    S = refrun.read_grads(workdir, "cm1out_s.ctl")
    U = refrun.read_grads(workdir, "cm1out_u.ctl")
    V = refrun.read_grads(workdir, "cm1out_v.ctl")
    W = refrun.read_grads(workdir, "cm1out_w.ctl")
    out = dict(meta=np.array([ni, nj, nk, nsteps, 1, nout], np.int32), dt=np.array([dt], np.float32),
               opts=np.array([psolver, advwenos, 0, 0, npt, doimpl], np.int32),
               grid=np.array([dx, dx, dz], np.float32),
               surface=np.array([TURB_NAMELIST["cnst_shflx"], TURB_NAMELIST["cnst_znt"]], np.float32))
    for t in range(nout + 1):
        tag = "t0" if t == 0 else f"o{t}"
        out[f"{tag}_u"] = U["u"][t]
        out[f"{tag}_v"] = V["v"][t]
        out[f"{tag}_w"] = W["w"][t]
        out[f"{tag}_th"] = S["thpert"][t]
        out[f"{tag}_pi"] = S["pipert"][t]
        out[f"{tag}_tke"] = W["tke"][t]
    for n in range(1, npt + 1):
        out[f"t0_pt{n}"] = S[f"pt{n}"][0]
        out[f"o{nout}_pt{n}"] = S[f"pt{n}"][nout]
    for k in ("rho", "prs"):
        out[f"o{nout}_{k}"] = S[k][nout]
    for k in ("kmh", "kmv", "khh", "khv", "lenscl", "nm", "defv", "defh", "dissten"):
        out[f"o{nout}_{k}"] = W[k][nout]
    for n in ("pi0", "th0", "qv0", "prs0"):
        out[n] = S[n][0, :, 0, 0]
    out["zh"] = S["zhval"][0, :, 0, 0]
    path = os.path.join(outdir, f"{name}.npz")
    np.savez_compressed(path, **{k: np.ascontiguousarray(v) for k, v in out.items()})
    return path


def make_turb_restart(name, ni, nj, nk, nsteps, nout, dt, doimpl, nrst, dx=20.0, dz=20.0, npt=2, psolver=2,
                      advwenos=1, workdir="/tmp/cm1_fixture_turb_rst", outdir=HERE, threads=4):
    """The run of make_turb with restart files written after nrst steps:
    stores every cm1rst_000001_*.dat as bytes."""
    import glob as _glob
    ov = dict(nx=ni, ny=nj, nz=nk, dtl=dt, timax=nsteps * dt, tapfrq=nsteps * dt / nout,
              rstfrq=nrst * dt, wbc=1, ebc=1, sbc=1, nbc=1, dx=dx, dy=dx, dz=dz, psolver=psolver,
              advwenos=advwenos, advwenov=0, irdamp=0, doimpl=doimpl, iptra=1 if npt else 0,
              npt=max(npt, 1), **TURB_NAMELIST)
    ov.update({"output_" + k: 1 for k in TURB_OUT})
    refrun.run_reference(workdir, ov, threads=threads, tables=("LANDUSE.TBL",))
    out = {os.path.basename(f).replace(".", "_"): np.frombuffer(open(f, "rb").read(), dtype=np.uint8)
           for f in sorted(_glob.glob(os.path.join(workdir, "cm1rst_000001_*.dat")))}
    out["meta"] = np.array([nrst], np.int32)
    path = os.path.join(outdir, f"{name}.npz")
    np.savez_compressed(path, **out)
    return path


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "turb":
        # 400 steps of 0.25 s: the subgrid TKE appears after about 300
        for doimpl in (1, 0):
            print(make_turb(f"ref_les_tke_impl{doimpl}_16x12x24_400", 16, 12, 24, 400, 4, 0.25, doimpl))
        # the reference's restart files after 300 of the 400 steps (the state
        # a run of the program restarted from them continues exactly; the
        # drop-in program restarts there instead of stepping from t=0)
        for doimpl in (1, 0):
            print(make_turb_restart(f"les_rst_impl{doimpl}_16x12x24_300", 16, 12, 24, 400, 4, 0.25, doimpl, 300))
        # the land-use table the surface model reads at start-up (input data
        # of the reference's les_ConvBoundLayer case), for runs of the
        # program where the reference's directory is absent
        with open(os.path.join(refrun.REF_RUN, "LANDUSE.TBL"), "rb") as f:
            tbl = np.frombuffer(f.read(), dtype=np.uint8)
        np.savez_compressed(os.path.join(HERE, "les_inputs.npz"), landuse_tbl=tbl)
    elif len(sys.argv) > 1 and sys.argv[1] == "stretched":
        # dt = 1.5 s: the reference's dtsmall picks nsound = 4; dt = 3 s: nsound = 8
        for dt, tag in ((1.5, "dt1.5"), (3.0, "dt3")):
            print(make(f"ref_stretch_{tag}_24x20x96_10", 24, 20, 96, 10, dt, 2,
                       dx=250.0, dy=250.0, dz=500.0, stretch=CONFIG1_STRETCH))
    elif len(sys.argv) > 1:
        a = sys.argv[1:]
        opts = dict(zip(OPT_KEYS[:4], (int(v) for v in a[7:11])))
        zd = float(a[11]) if len(a) > 11 else 15000.0
        if len(a) > 12:
            opts["npt"] = int(a[12])
        print(make(a[0], int(a[1]), int(a[2]), int(a[3]), int(a[4]), float(a[5]), int(a[6]),
                   opts=opts, zd=zd))
    else:
        print(make("ref_open_24x20x18_10", 24, 20, 18, 10, 6.0, 2))
        print(make("ref_periodic_20x24x18_10", 20, 24, 18, 10, 6.0, 1))
