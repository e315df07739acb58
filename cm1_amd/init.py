"""Host-side set-up of grids, base states and initial perturbations.

This is initialisation code (it runs once, on the host), restating the
parts of the reference's preprocessors the supported configurations use:

* vertical grid: unstretched, or the analytic stretching ``stretch_z=1``
  (``src/param.F:7790-7960``), plus the metric terms zh, mh, mf
  (``src/param.F:8140-8165``);
* Weisman-Klemp analytic sounding ``isnd=5`` (``src/base.F:363-455``) with
  the hydrostatic iteration and ``rslf`` (``src/cm1libs.F:16-62``);
* wind profile ``iwnd=4``, ``u0 = umax*tanh(z/3 km)`` (``src/base.F:2396-2408``);
  the reference hard-codes ``umax=35``; the value is a parameter here;
* warm bubbles of the ``iinit=1`` form (``src/init3d.F:450-487``) with
  configurable amplitude, centres and radii, and seeded uniform theta noise.

Arrays are returned in the library's logical layout: 3-D fields as
``(nk+3, nj+7, ni+7)`` for k=0..nk+2, j=-2..nj+4, i=-2..ni+4; columns as
``(nk+3,)`` for k=0..nk+2.
"""
import math

import numpy as np

F32 = np.float32

# constants (testcase=0 values of set_constants, src/constants.F:69-170)
G, RD, RV, CP = F32(9.81), F32(287.04), F32(461.5), F32(1005.7)
P00 = F32(1.0e5)
EPS = RD / RV
REPS = RV / RD


def _nint(x):
    """Fortran NINT of a REAL: nearest integer, halves away from zero."""
    x = float(x)
    return int(math.floor(x + 0.5)) if x >= 0.0 else -int(math.floor(-x + 0.5))


def vertical_grid(nk, dz, stretch=None):
    """zf, zh, mh, mf columns (k=0..nk+2) and the namelist dz.

    Every operation is single precision, in the reference's order
    (``src/param.F:7790-7810`` for the unstretched levels,
    ``src/param.F:7819-7950`` for ``stretch_z=1``,
    ``src/param.F:8140-8170`` for zh, mh and mf).  ``dz`` is the namelist
    value: the reference keeps it as the metric scale (``rdz = 1/dz``,
    ``mh = dz/(zf(k+1)-zf(k))``) also on a stretched grid.

    ``stretch`` = dict(dz_bot, dz_top, str_bot, str_top, ztop) selects
    stretch_z=1; settings the reference stops on raise ValueError."""
    dz = F32(dz)
    zf = np.zeros(nk + 3, F32)
    for k in range(0, nk + 3):
        zf[k] = dz * F32(k - 1)
    zh = np.zeros(nk + 3, F32)
    mh = np.ones(nk + 3, F32)
    mf = np.ones(nk + 3, F32)
    if stretch is None:
        for k in range(0, nk + 2):
            zh[k] = F32(0.5) * (zf[k] + zf[k + 1])
        return dict(zf=zf, zh=zh, mh=mh, mf=mf), float(dz)
    dzb, dzt = F32(stretch["dz_bot"]), F32(stretch["dz_top"])
    sb, st, zt = F32(stretch["str_bot"]), F32(stretch["str_top"]), F32(stretch["ztop"])
    nominal = F32(0.5) * (dzb + dzt)
    nk1 = _nint(sb / dzb)
    nk3 = _nint((zt - st) / dzt)
    nk2 = nk - (nk1 + nk3)
    # the reference's consistency checks (src/param.F:7827-7911)
    if abs(F32(nk1) * dzb - sb) > F32(0.01):
        raise ValueError("stretched grid: depth of bottom layer does not divide by dz_bot")
    if abs(F32(nk2) * nominal - (st - sb)) >= F32(0.01):
        raise ValueError("stretched grid: depth of middle layer does not divide by nominal_dz")
    if abs(F32(nk3) * dzt - (zt - st)) > F32(0.01):
        raise ValueError("stretched grid: depth of top layer does not divide by dz_top")
    test_len = dzb * F32(nk1) + F32(0.5) * (dzb + dzt) * F32(nk2) + dzt * F32(nk3)
    if nk1 < 0 or nk2 < 0 or nk3 < 0 or abs(zt - test_len) > F32(1.0e-6) * zt:
        raise ValueError(f"stretched grid: nk1,nk2,nk3 = {nk1},{nk2},{nk3} do not give ztop")
    nominal = (st - sb) / F32(nk2)
    c2 = (nominal - dzb) / (nominal * nominal * F32(nk2 - 1))
    c1 = (dzb / nominal) - c2 * nominal
    for k in range(1, nk1 + 2):
        zf[k] = F32(k - 1) * dzb
    for k in range(nk1 + 1, nk1 + nk2 + 2):
        m = F32(k - 1 - nk1)
        zf[k] = zf[nk1 + 1] + (c1 + c2 * m * nominal) * m * nominal
    for k in range(nk1 + nk2 + 2, nk1 + nk2 + nk3 + 2):
        zf[k] = zf[k - 1] + dzt
    zf[0] = -zf[2]
    zf[nk + 2] = zf[nk + 1] + (zf[nk + 1] - zf[nk])
    for k in range(0, nk + 2):
        zh[k] = F32(0.5) * (zf[k + 1] + zf[k])
        mh[k] = dz / (zf[k + 1] - zf[k])
    zh[0] = -zh[1]
    zh[nk + 1] = zh[nk] + F32(2.0) * (zf[nk + 1] - zh[nk])
    for k in range(1, nk + 2):
        mf[k] = dz / (zh[k] - zh[k - 1])
    mf[0] = mf[1]
    mf[nk + 2] = mf[nk + 1]
    return dict(zf=zf, zh=zh, mh=mh, mf=mf), float(dz)


def min_dz(zf, nk):
    """Smallest level thickness, min over k=1..nk of zf(k+1)-zf(k) in single
    precision (``src/param.F:8522-8541``); dtsmall uses it for psolver 2."""
    zf = np.asarray(zf, F32)
    return float(np.min(zf[2:nk + 2] - zf[1:nk + 1]))


def rslf(p, t):
    esl = 611.2 * math.exp(17.67 * (t - 273.15) / (t - 29.65))
    esl = min(esl, p * 0.5)
    return float(EPS) * esl / (p - esl)


def wk_sounding(zh, nk, imoist=1, qv_pbl=0.014, qv_scale=1.0):
    """Weisman-Klemp base state (src/base.F:363-455): th0, pi0, qv0 columns."""
    z_trop, th_trop, t_trop, th_sfc, prs_sfc = 12000.0, 343.0, 213.0, 300.0, 100000.0
    g, cp, rd = float(G), float(CP), float(RD)
    reps = float(REPS)
    pi_sfc = (prs_sfc / 1.0e5) ** (rd / cp)
    qv_sfc = rslf(prs_sfc, th_sfc * pi_sfc)
    thv_sfc = th_sfc * (1.0 + qv_sfc * reps) / (1.0 + qv_sfc)
    th0 = np.zeros(nk + 3)
    rh0 = np.zeros(nk + 3)
    for k in range(1, nk + 1):
        z = float(zh[k])
        if z < z_trop:
            th0[k] = th_sfc + (th_trop - th_sfc) * ((z / z_trop) ** 1.25)
            rh0[k] = (1.0 - 0.75 * ((z / z_trop) ** 1.25)) if imoist else 0.0
        else:
            th0[k] = th_trop * math.exp((g / (t_trop * cp)) * (z - z_trop))
            rh0[k] = 0.25 if imoist else 0.0
    th0[0], th0[nk + 1] = th0[1], th0[nk]
    qv0 = np.zeros(nk + 3)
    pi0 = np.zeros(nk + 3)
    for _ in range(20):
        thv0 = th0 * (1.0 + reps * qv0) / (1.0 + qv0)
        pi0[1] = pi_sfc - g * float(zh[1]) / (cp * 0.5 * (thv_sfc + thv0[1]))
        for k in range(2, nk + 1):
            pi0[k] = pi0[k - 1] - g * float(zh[k] - zh[k - 1]) / (cp * 0.5 * (thv0[k] + thv0[k - 1]))
        for k in range(1, nk + 1):
            prs = 1.0e5 * pi0[k] ** (cp / rd)
            q = rh0[k] * rslf(prs, th0[k] * pi0[k]) * qv_scale
            qv0[k] = min(q, qv_pbl) if imoist else 0.0
    return dict(th0=th0.astype(F32), pi0=pi0.astype(F32), qv0=qv0.astype(F32))


def wind_iwnd4(zh, nk, umax=35.0):
    """u0(k) = umax*tanh(z/3 km) at u points (src/base.F:2396-2408)."""
    u0 = np.zeros(nk + 3, F32)
    for k in range(1, nk + 1):
        u0[k] = F32(umax * math.tanh(float(zh[k]) / 3000.0))
    return u0


def blank(ni, nj, nk):
    return np.zeros((nk + 3, nj + 7, ni + 7), F32)


def initial_state(ni, nj, nk, dx, dy, dz, *, stretch=None, imoist=1, umax=35.0,
                  bubbles=None, noise_amp=0.0, noise_seed=0, noise_zmax=None,
                  qv_scale=1.0, npt=0, tracers=None):
    """Columns and t=0 fields for a WK supercell-type case.

    bubbles: list of dict(x, y, z, amp, rh, rv) in metres/K (iinit=1 form);
    default one 1 K bubble at the domain centre, 10 km / 1.4 km radii.
    """
    cols, dzn = vertical_grid(nk, dz, stretch)
    snd = wk_sounding(cols["zh"], nk, imoist=imoist, qv_scale=qv_scale)
    cols.update(snd)
    cols["u0"] = wind_iwnd4(cols["zh"], nk, umax)
    f = {n: blank(ni, nj, nk) for n in ("ua", "va", "wa", "tha", "ppi", "qa0", "qa1", "qa2")}
    for k in range(1, nk + 1):
        f["ua"][k, :, :] = cols["u0"][k]
        if imoist:
            f["qa0"][k, :, :] = cols["qv0"][k]
    if bubbles is None:
        bubbles = [dict(x=0.5 * ni * dx, y=0.5 * nj * dy, z=1400.0, amp=1.0, rh=10000.0, rv=1400.0)]
    xh = (np.arange(1, ni + 1) - 0.5) * dx
    yh = (np.arange(1, nj + 1) - 0.5) * dy
    zh = cols["zh"][1:nk + 1].astype(np.float64)
    th = np.zeros((nk, nj, ni))
    for b in bubbles:
        beta = np.sqrt(((xh[None, None, :] - b["x"]) / b["rh"]) ** 2 +
                       ((yh[None, :, None] - b["y"]) / b["rh"]) ** 2 +
                       ((zh[:, None, None] - b["z"]) / b["rv"]) ** 2)
        th += np.where(beta < 1.0, b["amp"] * np.cos(0.5 * np.pi * beta) ** 2, 0.0)
    if noise_amp > 0.0:
        rng = np.random.default_rng(noise_seed)
        nz = rng.uniform(-noise_amp, noise_amp, size=th.shape)
        if noise_zmax is not None:
            nz *= (zh < noise_zmax)[:, None, None]
        th += nz
    f["tha"][1:nk + 1, 3:nj + 3, 3:ni + 3] = th.astype(F32)
    for n in range(npt):
        a = blank(ni, nj, nk)
        if tracers is not None:
            a[1:nk + 1, 3:nj + 3, 3:ni + 3] = tracers[n]
        f[f"pta{n}"] = a
    return cols, f, dzn


def cbl_sounding(zh, nk, th_ml=300.0, z_inv=1000.0, gamma=0.008, prs_sfc=100000.0):
    """Dry convective-boundary-layer base state: theta = th_ml up to z_inv,
    then th_ml + gamma (z - z_inv); pi0 hydrostatic as in wk_sounding."""
    g, cp, rd = float(G), float(CP), float(RD)
    pi_sfc = (prs_sfc / 1.0e5) ** (rd / cp)
    th0 = np.zeros(nk + 3)
    for k in range(1, nk + 1):
        z = float(zh[k])
        th0[k] = th_ml + (gamma * (z - z_inv) if z > z_inv else 0.0)
    th0[0], th0[nk + 1] = th0[1], th0[nk]
    pi0 = np.zeros(nk + 3)
    pi0[1] = pi_sfc - g * float(zh[1]) / (cp * 0.5 * (th_ml + th0[1]))
    for k in range(2, nk + 1):
        pi0[k] = pi0[k - 1] - g * float(zh[k] - zh[k - 1]) / (cp * 0.5 * (th0[k] + th0[k - 1]))
    return dict(th0=th0.astype(F32), pi0=pi0.astype(F32), qv0=np.zeros(nk + 3, F32))


def cbl_state(ni, nj, nk, dx, dy, dz, *, u=10.0, npt=6, blob_sigma=200.0, noise_amp=0.1,
              noise_zmax=500.0, seed=0):
    """Columns and t=0 fields of the dry CBL case: the cbl_sounding base state,
    uniform u, +-noise_amp K theta noise below noise_zmax, and npt passive
    tracers initialised as seeded Gaussian blobs (amplitude 1, blob_sigma)."""
    cols, dzn = vertical_grid(nk, dz, None)
    cols.update(cbl_sounding(cols["zh"], nk))
    cols["u0"] = np.zeros(nk + 3, F32)
    cols["u0"][1:nk + 1] = u
    f = {n: blank(ni, nj, nk) for n in ("ua", "va", "wa", "tha", "ppi")}
    f["ua"][1:nk + 1] = u
    rng = np.random.default_rng(seed)
    zh = cols["zh"][1:nk + 1].astype(np.float64)
    nzk = int(np.count_nonzero(zh < noise_zmax))
    f["tha"][1:nzk + 1, 3:nj + 3, 3:ni + 3] = rng.uniform(-noise_amp, noise_amp,
                                                          size=(nzk, nj, ni)).astype(F32)
    xh = ((np.arange(1, ni + 1) - 0.5) * dx).astype(np.float32)
    yh = ((np.arange(1, nj + 1) - 0.5) * dy).astype(np.float32)
    for n in range(npt):
        cx, cy = rng.uniform(0.2, 0.8) * ni * dx, rng.uniform(0.2, 0.8) * nj * dy
        cz = rng.uniform(0.2, 0.6) * float(zh[-1])
        gx = np.exp(-0.5 * ((xh - cx) / blob_sigma) ** 2).astype(F32)
        gy = np.exp(-0.5 * ((yh - cy) / blob_sigma) ** 2).astype(F32)
        gz = np.exp(-0.5 * ((zh - cz) / blob_sigma) ** 2).astype(F32)
        a = blank(ni, nj, nk)
        a[1:nk + 1, 3:nj + 3, 3:ni + 3] = gz[:, None, None] * (gy[:, None] * gx[None, :])[None, :, :]
        f[f"pta{n}"] = a
    return cols, f, dzn
