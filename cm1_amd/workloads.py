"""Synthetic initial states of the benchmark workloads, shared by bench.py and
the tests (a test runs the same construction at a smaller horizontal size).

Each function returns ``(cfgd, cols, fields)``: the namelist-style config
keywords of :func:`cm1_amd.model.make_config`, the vertical columns and the
t=0 fields, in the layout of :mod:`cm1_amd.init`.
"""
import numpy as np

from . import init as hinit

# the stretched vertical grid of the 1024x1024x96 supercell: 50 m levels up
# to 500 m, analytic stretching to 500 m levels at 17 km, 30 km top
# (stretch_z = 1, src/param.F:7819-7950); the namelist dz (the metric scale)
# is 500 m
CONFIG1_STRETCH = dict(dz_bot=50.0, dz_top=500.0, str_bot=500.0, str_top=17000.0, ztop=30000.0)
CONFIG1_DZ = 500.0


def config1(ni=1024, nj=1024, nk=96, seed=2024, nsound=8, dt=1.5):
    """Supercell on the stretched grid: dx=dy=250 m, dt=1.5 s, nsound=8,
    Kessler, 5th-order advection, open lateral boundaries, Weisman-Klemp
    sounding with u = 30 tanh(z/3 km), eight seeded +-2 K bubbles (10 km /
    1.4 km radii, centred at 1.4 km) at random horizontal positions."""
    dx = dy = 250.0
    rng = np.random.default_rng(seed)
    bubbles = []
    for _ in range(8):
        bubbles.append(dict(x=float(rng.uniform(0.1, 0.9) * ni * dx),
                            y=float(rng.uniform(0.1, 0.9) * nj * dy), z=1400.0,
                            amp=float(rng.choice([-2.0, 2.0])), rh=10000.0, rv=1400.0))
    cols, fields, dzn = hinit.initial_state(ni, nj, nk, dx, dy, CONFIG1_DZ, stretch=CONFIG1_STRETCH,
                                            umax=30.0, bubbles=bubbles)
    cfgd = dict(ni=ni, nj=nj, nk=nk, wbc=2, ebc=2, sbc=2, nbc=2, imoist=1, ptype=1,
                psolver=3, advwenos=0, advwenov=0, apmasscon=1, irbc=4, nsound=nsound, npt=0,
                pdtra=1, dx=dx, dy=dy, dz=dzn, dt=dt, alph=0.6, kdiv=0.1, csmax=350.0)
    return cfgd, cols, fields


def ensemble_shapes(n=48, seed=2025):
    """The members' draws: (ni, nj, nk, qv scale, umax, bubble amplitude)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ni, nj = (int(rng.choice([96, 128, 192, 256])) for _ in range(2))
        nk = int(rng.choice([40, 56, 64]))
        qvs, umax, amp = float(rng.uniform(0.9, 1.1)), float(rng.uniform(10, 30)), float(rng.uniform(1, 3))
        out.append((ni, nj, nk, qvs, umax, amp))
    return out


def ensemble_member(ni, nj, nk, qvs, umax, amp):
    """One ensemble member: dx=dy=1 km, a 20 km model top, dt=6 s, nsound=6,
    Weisman-Klemp sounding with qv scaled by qvs, u = umax tanh(z/3 km), a
    centred bubble of amplitude amp."""
    dx = 1000.0
    dz = 20000.0 / nk
    bub = [dict(x=0.5 * ni * dx, y=0.5 * nj * dx, z=1400.0, amp=amp, rh=10000.0, rv=1400.0)]
    cols, fields, dzn = hinit.initial_state(ni, nj, nk, dx, dx, dz, umax=umax, bubbles=bub,
                                            qv_scale=qvs)
    cfgd = dict(ni=ni, nj=nj, nk=nk, wbc=2, ebc=2, sbc=2, nbc=2, imoist=1, ptype=1,
                psolver=3, advwenos=0, advwenov=0, apmasscon=1, irbc=4, nsound=6, npt=0,
                pdtra=1, dx=dx, dy=dx, dz=dzn, dt=6.0, alph=0.6, kdiv=0.1, csmax=350.0)
    return cfgd, cols, fields


def ensemble_small_shapes(n=96, seed=2027):
    """Small members, where one member alone fills a fraction of the chip:
    (nx, ny) from {24, 32, 48, 64}, nz from {24, 32, 40}, the same
    perturbations as the 48-member draw."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ni, nj = (int(rng.choice([24, 32, 48, 64])) for _ in range(2))
        nk = int(rng.choice([24, 32, 40]))
        qvs, umax, amp = float(rng.uniform(0.9, 1.1)), float(rng.uniform(10, 30)), float(rng.uniform(1, 3))
        out.append((ni, nj, nk, qvs, umax, amp))
    return out


def ensemble(n=48, seed=2025):
    """Batched supercell ensemble: n members with (nx, ny) drawn from
    {96,128,192,256} and nz from {40,56,64}, qv scaled by U(0.9,1.1), shear
    u = U(10,30) tanh(z/3 km), a centred U(1,3) K bubble."""
    return [ensemble_member(*s) for s in ensemble_shapes(n, seed)]


def ensemble_small(n=96, seed=2027):
    """Batched ensemble of small members (see ensemble_small_shapes)."""
    return [ensemble_member(*s) for s in ensemble_small_shapes(n, seed)]


def spinup(ni=128, nj=128, nk=40, seed=7):
    """Moist supercell case for the surrogate: 128x128x40, dx=dy=1 km,
    dz=500 m, dt=6 s, nsound=6, Kessler, four seeded +3 K bubbles."""
    dx = 1000.0
    rng = np.random.default_rng(seed)
    bubbles = [dict(x=float(rng.uniform(0.2, 0.8) * ni * dx), y=float(rng.uniform(0.2, 0.8) * nj * dx),
                    z=1400.0, amp=3.0, rh=10000.0, rv=1400.0) for _ in range(4)]
    cols, fields, dzn = hinit.initial_state(ni, nj, nk, dx, dx, 500.0, umax=25.0, bubbles=bubbles)
    cfgd = dict(ni=ni, nj=nj, nk=nk, wbc=2, ebc=2, sbc=2, nbc=2, imoist=1, ptype=1,
                psolver=3, advwenos=0, advwenov=0, apmasscon=1, irbc=4, nsound=6, npt=0,
                pdtra=1, dx=dx, dy=dx, dz=dzn, dt=6.0, alph=0.6, kdiv=0.1, csmax=350.0)
    return cfgd, cols, fields


# the subgrid turbulence of the CBL workload: the TKE scheme with implicit
# vertical diffusion, the semi-slip surface (log-law friction velocity over
# a 0.1 m roughness length) and the 0.24 K m/s surface heat flux, the
# switches of the reference's run/config_files/les_ConvBoundLayer
CBL_TURB = dict(cm1setup=1, doimpl=1, bbc=3, isfcflx=1, cnst_shflx=0.24, cnst_znt=0.1)


def cbl(ni=1536, nj=1536, nk=128, seed=2026):
    """Dry convective boundary layer: 1536x1536x128, dx=dy=dz=20 m,
    doubly periodic, dt=0.25 s, explicit acoustic solver (psolver=2), WENO5
    scalar advection on every stage, 300 K mixed layer to 1 km then 8 K/km,
    u=10 m/s, +-0.1 K noise below 500 m, six Gaussian tracer blobs
    (sigma 200 m); its subgrid turbulence and surface flux are CBL_TURB."""
    cols, fields, dzn = hinit.cbl_state(ni, nj, nk, 20.0, 20.0, 20.0, seed=seed)
    cfgd = dict(ni=ni, nj=nj, nk=nk, wbc=1, ebc=1, sbc=1, nbc=1, imoist=0, ptype=0,
                psolver=2, advwenos=1, advwenov=0, apmasscon=1, irbc=4, nsound=0, npt=6,
                pdtra=1, dx=20.0, dy=20.0, dz=dzn, dt=0.25, alph=0.6, kdiv=0.1, csmax=350.0)
    return cfgd, cols, fields
