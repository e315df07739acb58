"""Device GrADS writer (cm1g_write_grads) and statpack (cm1g_statpack)
against the reference's own output.

Golden vectors (tests/golden/make_grads_fixture.py): the reference
executable, started from the t=0 state of a state fixture with every output
switch the writer supports on, writes output at t=0 and after N large steps;
the fixture keeps the SHA-256 of every record of cm1out_{s,u,v,w}.dat, the
descriptor files and cm1out_metadata.dat.  The GPU model starts from the
same state, writes output time 1, steps N times and writes output time 2:

* every record (each level of each variable, both times, including the 2-D
  rain and prate after rain has reached the ground, and the derived
  th0+th', prs-prs0, rho-rho0, qv-qv0, buoyancy and the averages to scalar
  points) has the reference's digest, i.e. equals it bit for bit;
* the descriptors (.ctl) and the metadata file equal the reference's byte
  for byte, for the unstretched 64x64x40 grid (300 steps) and for the
  stretched 96-level grid (10 steps), whose descriptors list the levels;
* the three statistics records (t=0, halfway, end: 55 values each, the
  maxima/minima of w, u, v, pi', p', th', q, prate, surface pressure, wind
  speed, w at five heights, the levels of the w and wind-speed extremes,
  the Courant number, and the dry-air, moisture and per-species mass
  totals) equal the reference's cm1out_stats.dat bit for bit, and the stats
  descriptor byte for byte.
"""
import hashlib
import os

import numpy as np
import pytest

from common import load_fixture, load_turb_fixture, make_gpu, nsound_for

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [("grads_config0_64x64x40_300", "ref_config0_64x64x40_60", 0),
         ("grads_stretch_24x20x96_10", "ref_stretch_dt3_24x20x96_10", 1)]


def ctl_records(text):
    """(name, level) of each record of one output time, from a descriptor."""
    lines = text.splitlines()
    i0 = next(i for i, l in enumerate(lines) if l.startswith("vars"))
    recs = []
    for l in lines[i0 + 1:i0 + 1 + int(lines[i0].split()[1])]:
        name, nl = l.split()[:2]
        recs += [(name, k) for k in range(1, max(1, int(nl)) + 1)]
    return recs


@pytest.mark.parametrize("case,state,stretch", CASES, ids=[c[0] for c in CASES])
def test_grads_output_equals_the_reference(require_gpu, tmp_path, case, state, stretch):
    g = np.load(os.path.join(GOLD, case + ".npz"))
    ni, nj, nk, nsteps, ntimes = [int(v) for v in g["meta"]]
    dt = float(g["dt"][0])
    switches = [s.decode() for s in g["switches"]]
    cfgd, cols, fields, _, _, _ = load_fixture(os.path.join(GOLD, state + ".npz"))
    assert (cfgd["ni"], cfgd["nj"], cfgd["nk"], cfgd["dt"]) == (ni, nj, nk, dt)
    cfgd["nsound"] = nsound_for(cfgd, cols["zf"])
    m = make_gpu(cfgd, cols, fields)
    out = str(tmp_path)
    stats = [s.decode() for s in g["stats"]]
    m.write_grads(out, 1, switches, iorigin=2, stretch_z=stretch, mtime=0.0, nstep=0)
    rec = [m.statpack(out, 1, stats, mtime=0.0)]
    m.step(nsteps // 2)
    rec.append(m.statpack(out, 2, stats, mtime=(nsteps // 2) * dt))
    m.step(nsteps - nsteps // 2)
    rec.append(m.statpack(out, 3, stats, mtime=nsteps * dt))
    m.write_grads(out, 2, switches, iorigin=2, stretch_z=stretch, mtime=nsteps * dt, nstep=nsteps)
    print(f"{case}: {nsteps} steps, output written to {out}", flush=True)
    compare_files(g, out, ni, nj)
    # the 2-D records of the last time as values (rain has reached the
    # ground in the 300-step case)
    s = np.fromfile(os.path.join(out, "cm1out_s.dat"), dtype="<f4")
    per_t = s.size // ntimes
    rain = s[per_t:per_t + ni * nj].reshape(nj, ni)
    prate = s[per_t + ni * nj:per_t + 2 * ni * nj].reshape(nj, ni)
    assert np.array_equal(rain, g["rain"]) and np.array_equal(prate, g["prate"])
    if nsteps >= 300:
        assert g["rain"].max() > 0.0 and g["prate"].max() > 0.0
    compare_stats(g, out, rec)
    m.close()


def test_les_grads_output_equals_the_reference(require_gpu, tmp_path):
    """The dry LES (TKE scheme, implicit vertical diffusion, semi-slip
    surface, 0.24 K m/s heat flux) from the reference's t=0 state: output at
    t=0 and after 400 steps, the TKE active, with the subgrid turbulence's
    switches on (tke, kmh, kmv, khh, khv, lenscl, dissten, nm, defv, defh,
    epst) besides the dry ones, and statistics at t=0, 200 and 400 steps
    with the turbulence's tkemax/min, kmh/kmv and khh/khv extremes.  As the
    reference writes after its end-of-step sfc_and_turb, cm1g_sfc_and_turb
    runs before each write.  Every record, the statistics, the descriptors
    and the metadata equal the reference's."""
    g = np.load(os.path.join(GOLD, "grads_les_16x12x24_400.npz"))
    ni, nj, nk, nsteps, ntimes = [int(v) for v in g["meta"]]
    dt = float(g["dt"][0])
    switches = [s.decode() for s in g["switches"]]
    cfgd, cols, fields, n_ref, _, _ = load_turb_fixture(os.path.join(GOLD, "ref_les_tke_impl1_16x12x24_400.npz"))
    assert (cfgd["ni"], cfgd["nj"], cfgd["nk"], cfgd["dt"], n_ref) == (ni, nj, nk, dt, nsteps)
    stats = [s.decode() for s in g["stats"]]
    m = make_gpu(cfgd, cols, fields)
    out = str(tmp_path)
    m.sfc_and_turb()
    m.write_grads(out, 1, switches, iorigin=2, stretch_z=0, mtime=0.0, nstep=0)
    rec = [m.statpack(out, 1, stats, mtime=0.0)]
    m.step(nsteps // 2)
    m.sfc_and_turb()
    rec.append(m.statpack(out, 2, stats, mtime=(nsteps // 2) * dt))
    m.step(nsteps - nsteps // 2)
    m.sfc_and_turb()
    rec.append(m.statpack(out, 3, stats, mtime=nsteps * dt))
    m.write_grads(out, 2, switches, iorigin=2, stretch_z=0, mtime=nsteps * dt, nstep=nsteps)
    print(f"LES: {nsteps} steps, output written to {out}", flush=True)
    compare_files(g, out, ni, nj)
    compare_stats(g, out, rec)
    w = np.fromfile(os.path.join(out, "cm1out_w.dat"), dtype="<f4").reshape(ntimes, -1, nk + 1, nj, ni)
    assert float(w[1, 1].max()) > 0.05 and np.array_equal(w[1, 1], g["tke"])
    m.close()


def compare_stats(g, out, rec):
    """The three statistics records (values) and cm1out_stats.dat/.ctl
    against the reference's."""
    ref_stats = np.frombuffer(g["stats_dat"].tobytes(), dtype="<f4").reshape(3, -1)
    names = ctl_records(g["stats_ctl"].tobytes().decode())
    got = np.array(rec)
    assert got.shape == ref_stats.shape, (got.shape, ref_stats.shape)
    bad = [(t, names[n][0], float(got[t, n]), float(ref_stats[t, n]))
           for t in range(3) for n in range(got.shape[1])
           if got[t, n].view(np.uint32) != ref_stats[t, n].view(np.uint32)]
    print(f"statistics: {got.size} values, {len(bad)} differ {bad[:8]}", flush=True)
    assert not bad, bad[:20]
    assert open(os.path.join(out, "cm1out_stats.dat"), "rb").read() == g["stats_dat"].tobytes()
    assert open(os.path.join(out, "cm1out_stats.ctl"), "rb").read() == g["stats_ctl"].tobytes()


def compare_files(g, out, ni, nj):
    """cm1out_{s,u,v,w}.ctl and the metadata files byte for byte; every
    record of cm1out_{s,u,v,w}.dat against the reference's digest."""
    recl = dict(s=ni * nj * 4, u=(ni + 1) * nj * 4, v=ni * (nj + 1) * 4, w=ni * nj * 4)
    for f, r in recl.items():
        ref_ctl = g[f"{f}_ctl"].tobytes()
        got_ctl = open(os.path.join(out, f"cm1out_{f}.ctl"), "rb").read()
        assert got_ctl == ref_ctl, (f, got_ctl.decode(), ref_ctl.decode())
        data = open(os.path.join(out, f"cm1out_{f}.dat"), "rb").read()
        ref = [d.decode() for d in g[f"{f}_digest"]]
        assert len(data) == r * len(ref), (f, len(data), r * len(ref))
        names = ctl_records(ref_ctl.decode())
        bad = []
        for n, want in enumerate(ref):
            if hashlib.sha256(data[n * r:(n + 1) * r]).hexdigest() != want:
                bad.append((n // len(names) + 1,) + names[n % len(names)])
        print(f"cm1out_{f}.dat: {len(ref)} records, {len(bad)} differ", flush=True)
        assert not bad, (f, bad[:20])
    assert open(os.path.join(out, "cm1out_metadata.ctl"), "rb").read() == g["metadata_ctl"].tobytes()
    assert open(os.path.join(out, "cm1out_metadata.dat"), "rb").read() == g["metadata_dat"].tobytes()
