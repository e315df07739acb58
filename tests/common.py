"""Shared helpers of the parity tests: one initial state loaded into both the
HIP model and the CPU oracle, and field comparisons."""
import ctypes as C

import numpy as np

import pyoracle
from cm1_amd import init as hinit

# a small stretch_z = 1 grid for nk = 30: 100 m levels to 500 m, analytic
# stretching to 500 m levels at 5 km, 10 km top (nk1, nk2, nk3 = 5, 15, 10);
# the namelist dz stays 500 m
SMALL_STRETCH = dict(dz_bot=100.0, dz_top=500.0, str_bot=500.0, str_top=5000.0, ztop=10000.0)

# prognostic fields compared after stepping: (gpu/oracle name, staggering)
PROG = [("ua", "u"), ("va", "v"), ("wa", "w"), ("tha", None), ("ppi", None),
        ("qa0", None), ("qa1", None), ("qa2", None), ("rho", None), ("prs", None)]


def case(ni=40, nj=36, nk=30, dx=1000.0, dy=1000.0, dz=500.0, dt=6.0, bc=2, imoist=1,
         ptype=1, nsound=6, amp=3.0, umax=25.0, stretch=None, noise=0.1, seed=7, npt=0,
         bfx=0.45, bfy=0.55, brh=None, psolver=3, advwenos=0, advwenov=0, irdamp=0, zd=None):
    cfgd = dict(ni=ni, nj=nj, nk=nk, wbc=bc, ebc=bc, sbc=bc, nbc=bc, imoist=imoist,
                ptype=ptype, psolver=psolver, advwenos=advwenos, advwenov=advwenov,
                apmasscon=1, irbc=4,
                nsound=nsound, npt=npt, pdtra=1, dx=dx, dy=dy, dz=dz, dt=dt, alph=0.6,
                kdiv=0.1, csmax=350.0, irdamp=irdamp, rdalpha=3.3333333333e-3,
                zd=(0.7 * nk * dz if zd is None else zd))
    bub = [dict(x=bfx * ni * dx, y=bfy * nj * dy, z=1400.0, amp=amp,
                rh=(8.0 * dx if brh is None else brh), rv=1400.0)]
    cols, fields, dzn = hinit.initial_state(ni, nj, nk, dx, dy, dz, stretch=stretch,
                                            imoist=imoist, umax=umax, bubbles=bub,
                                            noise_amp=noise, noise_seed=seed, npt=npt)
    if stretch is not None:
        cfgd["dz"] = dzn
    return cfgd, cols, fields


def make_oracle(cfgd, cols, fields, libm=0, fp64=False):
    """libm=0: exp/pow rounded as the reference (host C library); libm=1:
    as the HIP kernels (double precision, rounded once).  fp64: the
    double-precision build, started from the same (single-precision) state."""
    cfg = pyoracle.make_config(**cfgd)
    o = pyoracle.Oracle(cfg, libm=libm, fp64=fp64)
    for n in ("zf", "zh", "mh", "mf"):
        o.kf(n)[:cfgd["nk"] + 3] = cols[n]
    o.L.cm1o_set_extrap(o.h, None)
    for n in ("pi0", "th0", "qv0", "u0", "v0"):
        if n in cols:
            o.kf(n)[:cfgd["nk"] + 3] = cols[n]
    o.base_derive()
    for name, arr in fields.items():
        o.f(name)[:, :, :] = arr
    o.getset(True)
    return o


def make_gpu(cfgd, cols, fields):
    from cm1_amd import model
    d = dict(cfgd)
    turb = {k: d.pop(k) for k in TURB_KEYS if k in d}
    cfg = model.make_config(**d)
    return model.setup(cfg, cols, fields, turb=turb if turb.get("cm1setup") else None)


def interior(a, nk, nj, ni, stag=None):
    ei, ej, ek = (1 if stag == "u" else 0), (1 if stag == "v" else 0), (1 if stag == "w" else 0)
    return a[1:nk + 1 + ek, 3:nj + 3 + ej, 3:ni + 3 + ei]


def rel_l2(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    den = np.sqrt((b * b).sum())
    num = np.sqrt(((a - b) ** 2).sum())
    if den == 0.0:
        return 0.0 if num == 0.0 else np.inf
    return num / den


def load_fixture(path):
    """Reference t=0 state -> (cfgd, cols, fields), plus the t=N outputs."""
    import os
    z = np.load(path)
    ni, nj, nk, nsteps, bc = [int(v) for v in z["meta"]]
    dt = float(z["dt"][0])
    opts = [int(v) for v in z["opts"]] if "opts" in z else [3, 0, 0]
    psolver, advwenos, advwenov = opts[:3]
    irdamp = opts[3] if len(opts) > 3 else 0
    npt = opts[4] if len(opts) > 4 else 0
    rdalpha, zd = [float(v) for v in z["rdamp"]] if "rdamp" in z else (3.3333333333e-3, 15000.0)
    dx, dy, dz, stretch = 1000.0, 1000.0, 500.0, None
    if "grid" in z:
        g = [float(v) for v in z["grid"]]
        dx, dy, dz = g[:3]
        if int(g[3]) == 1:
            stretch = dict(zip(("dz_bot", "dz_top", "str_bot", "str_top", "ztop"), g[4:9]))
    cfgd = dict(ni=ni, nj=nj, nk=nk, wbc=bc, ebc=bc, sbc=bc, nbc=bc, imoist=1, ptype=1,
                psolver=psolver, advwenos=advwenos, advwenov=advwenov, apmasscon=1, irbc=4,
                nsound=0, npt=npt, pdtra=1, dx=dx, dy=dy, dz=dz, dt=dt, alph=0.6,
                kdiv=0.1, csmax=350.0, irdamp=irdamp, rdalpha=rdalpha, zd=zd)
    cols, _ = hinit.vertical_grid(nk, dz, stretch)
    for n in ("pi0", "th0", "qv0"):
        c = np.zeros(nk + 3, np.float32)
        c[1:nk + 1] = z[n]
        cols[n] = c
    # base-state winds: the reference starts from ua = u0, va = v0
    # (horizontally uniform columns)
    for n, key in (("u0", "t0_u"), ("v0", "t0_v")):
        c = np.zeros(nk + 3, np.float32)
        c[1:nk + 1] = z[key][:, 0, 0]
        cols[n] = c
    f = {n: hinit.blank(ni, nj, nk) for n in ("ua", "va", "wa", "tha", "ppi", "qa0", "qa1", "qa2")}
    f["ua"][1:nk + 1, 3:nj + 3, 3:ni + 4] = z["t0_u"]
    f["va"][1:nk + 1, 3:nj + 4, 3:ni + 3] = z["t0_v"]
    f["wa"][1:nk + 2, 3:nj + 3, 3:ni + 3] = z["t0_w"]
    f["tha"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z["t0_th"]
    f["ppi"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z["t0_pi"]
    f["qa0"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z["t0_qv"]
    f["qa1"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z["t0_qc"]
    f["qa2"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z["t0_qr"]
    target = {"ua": z["tN_u"], "va": z["tN_v"], "wa": z["tN_w"], "tha": z["tN_th"],
              "ppi": z["tN_pi"], "qa0": z["tN_qv"], "qa1": z["tN_qc"], "qa2": z["tN_qr"],
              "rho": z["tN_rho"], "prs": z["tN_prs"]}
    # passive tracers (iptra = 1): pta(n) is the reference's pt(n+1)
    for n in range(npt):
        f[f"pta{n}"] = hinit.blank(ni, nj, nk)
        f[f"pta{n}"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z[f"t0_pt{n + 1}"]
        target[f"pta{n}"] = z[f"tN_pt{n + 1}"]
    return cfgd, cols, f, nsteps, target, z


def nsound_for(cfgd, zf=None):
    """dtsmall (src/misclibs.F:818-915) for psolver=2 or 3 (3-D), in Python;
    zf: the grid's zf column, whose smallest level thickness psolver 2 uses
    (src/param.F:8522-8541); None for an unstretched grid of spacing dz."""
    if cfgd.get("psolver", 3) == 2:
        mdz = cfgd["dz"] if zf is None else hinit.min_dz(zf, cfgd["nk"])
        mn = min(cfgd["dx"], cfgd["dy"], mdz)
        dtsm = np.float32(0.50) * np.float32(mn) / np.float32(360.0)
    else:
        dtsm = np.float32(0.60) * np.float32(min(cfgd["dx"], cfgd["dy"])) / np.float32(360.0)
    ns = max(int(round(cfgd["dt"] / float(dtsm))), 4)
    if ns % 2:
        ns += 1
    if cfgd["dt"] / float(np.float32(ns)) > float(dtsm):
        ns += 2
    return ns


# LES subgrid-turbulence switches of the device path (the keys the oracle
# and cm1g_set_turbulence take), for the reference's dry CBL case
TURB_KEYS = ("cm1setup", "doimpl", "bbc", "isfcflx", "cnst_shflx", "cnst_znt")


def load_turb_fixture(path):
    """Reference t=0 state of an LES fixture (make_ref_fixture.make_turb) ->
    (cfgd, cols, fields, nsteps, nout, z); cfgd carries the turbulence
    switches (TURB_KEYS) besides the dynamical-core configuration."""
    z = np.load(path)
    ni, nj, nk, nsteps, bc, nout = [int(v) for v in z["meta"]]
    dt = float(z["dt"][0])
    psolver, advwenos, _, _, npt, doimpl = [int(v) for v in z["opts"]]
    dx, dy, dz = [float(v) for v in z["grid"]]
    shflx, znt = [float(v) for v in z["surface"]]
    cfgd = dict(ni=ni, nj=nj, nk=nk, wbc=bc, ebc=bc, sbc=bc, nbc=bc, imoist=0, ptype=0,
                psolver=psolver, advwenos=advwenos, advwenov=0, apmasscon=1, irbc=4,
                nsound=0, npt=npt, pdtra=1, dx=dx, dy=dy, dz=dz, dt=dt, alph=0.6,
                kdiv=0.1, csmax=350.0, irdamp=0, rdalpha=3.3333333333e-3, zd=15000.0,
                cm1setup=1, doimpl=doimpl, bbc=3, isfcflx=1, cnst_shflx=shflx, cnst_znt=znt)
    cols, _ = hinit.vertical_grid(nk, dz, None)
    for n in ("pi0", "th0", "qv0"):
        c = np.zeros(nk + 3, np.float32)
        c[1:nk + 1] = z[n]
        cols[n] = c
    for n, key in (("u0", "t0_u"), ("v0", "t0_v")):
        c = np.zeros(nk + 3, np.float32)
        c[1:nk + 1] = z[key][:, 0, 0]
        cols[n] = c
    f = {n: hinit.blank(ni, nj, nk) for n in ("ua", "va", "wa", "tha", "ppi", "tkea")}
    f["ua"][1:nk + 1, 3:nj + 3, 3:ni + 4] = z["t0_u"]
    f["va"][1:nk + 1, 3:nj + 4, 3:ni + 3] = z["t0_v"]
    f["wa"][1:nk + 2, 3:nj + 3, 3:ni + 3] = z["t0_w"]
    f["tha"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z["t0_th"]
    f["ppi"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z["t0_pi"]
    f["tkea"][1:nk + 2, 3:nj + 3, 3:ni + 3] = z["t0_tke"]
    for n in range(npt):
        f[f"pta{n}"] = hinit.blank(ni, nj, nk)
        f[f"pta{n}"][1:nk + 1, 3:nj + 3, 3:ni + 3] = z[f"t0_pt{n + 1}"]
    cfgd["nsound"] = nsound_for(cfgd)
    return cfgd, cols, f, nsteps, nout, z


# fixture key -> (model field, staggering) of the turbulence outputs
TURB_OUT = {"u": ("ua", "u"), "v": ("va", "v"), "w": ("wa", "w"), "th": ("tha", None),
            "pi": ("ppi", None), "tke": ("tkea", "w"), "rho": ("rho", None), "prs": ("prs", None),
            "kmh": ("kmh", "w"), "kmv": ("kmv", "w"), "khh": ("khh", "w"), "khv": ("khv", "w"),
            "lenscl": ("lenscl", "w"), "nm": ("nm", "w"), "defv": ("defv", "w"),
            "defh": ("defh", "w"), "dissten": ("dissten", "w")}
