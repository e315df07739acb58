"""Golden vectors of the reference's GrADS output (src/writeout.F) and
statistics (src/statpack.F).

Runs ``oracle/_ref/cm1_sp.exe`` with the namelist of a state fixture (so
from the same t=0 state) and writes output at t = 0 and at the end, with
every output switch the device writer supports turned on, and statistics at
t = 0, halfway and at the end with every statistics switch the device
statpack supports turned on:

* ``grads_config0_64x64x40_300``: the 64x64x40 supercell of
  ``ref_config0_64x64x40_60.npz`` over 300 large steps (30 minutes, long
  enough for rain to reach the ground);
* ``grads_stretch_24x20x96_10``: the stretched 96-level grid of
  ``ref_stretch_dt3_24x20x96_10.npz`` over 10 steps (its descriptors list
  the levels);
* ``grads_les_16x12x24_400``: the dry LES of ``ref_les_tke_impl1_16x12x24_400.npz``
  (TKE scheme, implicit vertical diffusion, semi-slip surface, 0.24 K m/s
  heat flux) over 400 steps, the TKE active over the last ~100, with the
  subgrid turbulence's output and statistics switches (output: tke, km,
  kh, dissten, nm, def; statistics: tke, km, kh) on besides the dry ones.

Each fixture stores:

* ``{s,u,v,w}_digest``: the SHA-256 of every record of ``cm1out_{s,u,v,w}.dat``
  (one record per level of each variable, both output times, in file order);
* ``{s,u,v,w,metadata}_ctl``: the descriptor files, byte for byte;
* ``metadata_dat``: ``cm1out_metadata.dat`` (mtime, nstep, dt, nwrite);
* ``rain``, ``prate``: the two 2-D records of the last output time as values, so a
  failing comparison can say by how much;
* ``stats_dat``, ``stats_ctl``: ``cm1out_stats.dat`` (three records) and its
  descriptor.

Usage: python tests/golden/make_grads_fixture.py [les]
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)
import refrun  # noqa: E402
from make_ref_fixture import CONFIG1_STRETCH, RDALPHA, STRETCH_KEYS, TURB_NAMELIST  # noqa: E402

# output switches (param9) the device writer supports, all on
SWITCHES = ("rain", "zh", "basestate", "th", "thpert", "prs", "prspert", "pi", "pipert",
            "rho", "rhopert", "qv", "qvpert", "q", "buoyancy", "u", "upert", "uinterp", "v",
            "vpert", "vinterp", "w", "winterp")
# statistics switches (param10) the device statpack supports, all on
STATS = ("w", "wlevs", "u", "v", "pipert", "prspert", "thpert", "q", "sfcprs", "wsp", "cfl", "tmass",
         "tmois", "qmass")


def digests(path, reclen):
    data = open(path, "rb").read()
    assert len(data) % reclen == 0, (path, len(data), reclen)
    return np.array([hashlib.sha256(data[o:o + reclen]).hexdigest().encode()
                     for o in range(0, len(data), reclen)], dtype="S64")


def text(path):
    return np.frombuffer(open(path, "rb").read(), dtype=np.uint8)


def make(name, ni, nj, nk, nsteps, dt, dx, stretch=None, workdir="/tmp/cm1_grads_fixture",
         outdir=HERE):
    """The namelist make_ref_fixture.make writes for the same arguments, with
    output at t = 0 and t = nsteps*dt and the writer's switches on."""
    ov = dict(nx=ni, ny=nj, nz=nk, dtl=dt, timax=nsteps * dt, tapfrq=nsteps * dt,
              statfrq=nsteps * dt / 2,
              wbc=2, ebc=2, sbc=2, nbc=2, rdalpha=RDALPHA, zd=15000.0,
              dx=dx, dy=dx, dz=500.0, psolver=3, advwenos=0, advwenov=0, irdamp=0)
    if stretch is not None:
        ov.update(stretch_z=1, **{k: float(stretch[k]) for k in STRETCH_KEYS})
    ov.update({"output_" + k: 1 for k in SWITCHES})
    ov.update({"stat_" + k: 1 for k in STATS})
    refrun.run_reference(workdir, ov, threads=8)
    out = dict(meta=np.array([ni, nj, nk, nsteps, 2], np.int32), dt=np.array([dt], np.float32),
               switches=np.array([s.encode() for s in SWITCHES], dtype="S16"),
               stats=np.array([s.encode() for s in STATS], dtype="S16"))
    recl = dict(s=ni * nj * 4, u=(ni + 1) * nj * 4, v=ni * (nj + 1) * 4, w=ni * nj * 4)
    for f, r in recl.items():
        out[f"{f}_digest"] = digests(os.path.join(workdir, f"cm1out_{f}.dat"), r)
        out[f"{f}_ctl"] = text(os.path.join(workdir, f"cm1out_{f}.ctl"))
    out["metadata_ctl"] = text(os.path.join(workdir, "cm1out_metadata.ctl"))
    out["metadata_dat"] = text(os.path.join(workdir, "cm1out_metadata.dat"))
    out["stats_ctl"] = text(os.path.join(workdir, "cm1out_stats.ctl"))
    out["stats_dat"] = text(os.path.join(workdir, "cm1out_stats.dat"))
    S = refrun.read_grads(workdir, "cm1out_s.ctl")
    out["rain"] = S["rain"][1, 0]
    out["prate"] = S["prate"][1, 0]
    path = os.path.join(outdir, f"{name}.npz")
    np.savez_compressed(path, **out)
    return path


# the LES case: the dry switches of the device writer and the turbulence's
LES_SWITCHES = ("zh", "basestate", "th", "thpert", "prs", "prspert", "pi", "pipert", "rho",
                "rhopert", "buoyancy", "u", "upert", "uinterp", "v", "vpert", "vinterp", "w",
                "winterp", "tke", "km", "kh", "dissten", "nm", "def")
# (stat_wsp and stat_cfl are left out: with a surface model and a subgrid
# model the reference adds wsp10/hpbl and the diffusion numbers ksh/ksv)
LES_STATS = ("w", "u", "v", "pipert", "prspert", "thpert", "sfcprs", "tmass", "tke", "km", "kh")


def make_les(name, state, workdir="/tmp/cm1_grads_fixture_les", outdir=HERE):
    """The namelist make_ref_fixture.make_turb wrote for the LES state
    fixture ``state``, with output at t = 0 and at its last step and
    LES_SWITCHES on (every other output switch off)."""
    z = np.load(os.path.join(HERE, state + ".npz"))
    ni, nj, nk, nsteps, _, _ = [int(v) for v in z["meta"]]
    dt = float(z["dt"][0])
    psolver, advwenos, _, _, npt, doimpl = [int(v) for v in z["opts"]]
    dx, _, dz = [float(v) for v in z["grid"]]
    ov = dict(nx=ni, ny=nj, nz=nk, dtl=dt, timax=nsteps * dt, tapfrq=nsteps * dt,
              statfrq=nsteps * dt / 2,
              wbc=1, ebc=1, sbc=1, nbc=1, dx=dx, dy=dx, dz=dz, psolver=psolver,
              advwenos=advwenos, advwenov=0, irdamp=0, doimpl=doimpl, iptra=1 if npt else 0,
              npt=max(npt, 1), **TURB_NAMELIST)
    ov.update({"output_" + k: 0 for k in ("rain", "qv", "q")})
    ov.update({"output_" + k: 1 for k in LES_SWITCHES})
    ov.update({"stat_" + k: 1 for k in LES_STATS})
    refrun.run_reference(workdir, ov, threads=4, tables=("LANDUSE.TBL",))
    S = refrun.read_grads(workdir, "cm1out_s.ctl")
    # the same t = 0 state as the state fixture
    assert np.array_equal(S["thpert"][0], z["t0_th"]), "t = 0 state differs from " + state
    out = dict(meta=np.array([ni, nj, nk, nsteps, 2], np.int32), dt=np.array([dt], np.float32),
               switches=np.array([s.encode() for s in LES_SWITCHES], dtype="S16"),
               stats=np.array([s.encode() for s in LES_STATS], dtype="S16"))
    recl = dict(s=ni * nj * 4, u=(ni + 1) * nj * 4, v=ni * (nj + 1) * 4, w=ni * nj * 4)
    for f, r in recl.items():
        out[f"{f}_digest"] = digests(os.path.join(workdir, f"cm1out_{f}.dat"), r)
        out[f"{f}_ctl"] = text(os.path.join(workdir, f"cm1out_{f}.ctl"))
    out["metadata_ctl"] = text(os.path.join(workdir, "cm1out_metadata.ctl"))
    out["metadata_dat"] = text(os.path.join(workdir, "cm1out_metadata.dat"))
    out["stats_ctl"] = text(os.path.join(workdir, "cm1out_stats.ctl"))
    out["stats_dat"] = text(os.path.join(workdir, "cm1out_stats.dat"))
    W = refrun.read_grads(workdir, "cm1out_w.ctl")
    for k in ("tke", "kmh", "dissten"):
        out[k] = W[k][1]
    path = os.path.join(outdir, f"{name}.npz")
    np.savez_compressed(path, **out)
    return path


if __name__ == "__main__":
    if sys.argv[1:] == ["les"]:
        print(make_les("grads_les_16x12x24_400", "ref_les_tke_impl1_16x12x24_400"))
        sys.exit(0)
    print(make("grads_config0_64x64x40_300", 64, 64, 40, 300, 6.0, 1000.0))
    print(make("grads_stretch_24x20x96_10", 24, 20, 96, 10, 3.0, 250.0, stretch=CONFIG1_STRETCH))
