"""Run the reference CM1 executable (``oracle/_ref/cm1_sp.exe``) and read its output.

Test infrastructure only (parity fixtures and the CPU baseline timing).

* :func:`namelist_text` writes a complete ``namelist.input`` in the
  reference's format (groups and keys as read by ``src/param.F:185-280``),
  from our own table of values for the dynamical-core configurations:
  5th-order advection, psolver 3, Kessler (ptype 1), no turbulence,
  radiation, Coriolis or Rayleigh damping, GrADS output.
* :func:`run_reference` runs the executable in a scratch directory.
* :func:`read_grads` reads the GrADS direct-access output described by the
  ``.ctl`` files the reference writes (``src/writeout.F:9818-9905``): one
  record per horizontal slab, little-endian REAL*4.
"""
import os
import re
import shutil
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_EXE = os.path.join(HERE, "_ref", "cm1_sp.exe")
# the reference's run directory (its data tables), read by fixture scripts only
REF_RUN = "/root/reference/run"


def _fmt(v):
    if isinstance(v, bool):
        return ".true." if v else ".false."
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return repr(v)
    return str(v)


# Group order as read by src/param.F:185-280.  Values are ours.
def default_groups():
    g = {}
    g["param0"] = dict(nx=64, ny=64, nz=40, ppnode=128, timeformat=2, timestats=2,
                       terrain_flag=False, procfiles=False, outunits=1)
    g["param1"] = dict(dx=1000.0, dy=1000.0, dz=500.0, dtl=6.0, timax=60.0,
                       run_time=-999.9, tapfrq=60.0, rstfrq=-3600.0, statfrq=1.0e9,
                       prclfrq=1.0e9)
    g["param2"] = dict(cm1setup=0, testcase=0, adapt_dt=0, irst=0, rstnum=1, iconly=0,
                       hadvordrs=5, vadvordrs=5, hadvordrv=5, vadvordrv=5,
                       advwenos=0, advwenov=0, weno_order=5, apmasscon=1, idiff=0,
                       mdiff=0, difforder=6, imoist=1, ipbl=0, sgsmodel=1, tconfig=1,
                       bcturbs=1, horizturb=0, doimpl=1, irdamp=0, hrdamp=0,
                       psolver=3, ptype=1, ihail=1, iautoc=1, cuparam=0, icor=0,
                       betaplane=0, lspgrad=0, eqtset=2, idiss=0, efall=0, rterm=0,
                       wbc=2, ebc=2, sbc=2, nbc=2, bbc=1, tbc=1, irbc=4, roflux=0,
                       nudgeobc=0, isnd=5, iwnd=4, itern=0, iinit=1, irandp=0,
                       ibalance=0, iorigin=2, axisymm=0, imove=0, iptra=0, npt=1,
                       pdtra=1, iprcl=0, nparcels=1)
    g["param3"] = dict(kdiff2=75.0, kdiff6=0.040, fcor=0.00005, kdiv=0.10, alph=0.60,
                       rdalpha=3.3333333333e-3, zd=15000.0, xhd=100000.0,
                       alphobc=60.0, umove=0.0, vmove=0.0, v_t=7.0, l_h=100.0,
                       lhref1=100.0, lhref2=1000.0, l_inf=75.0, ndcnst=250.0,
                       nt_c=250.0, csound=300.0, cstar=30.0)
    g["param11"] = dict(radopt=0, dtrad=300.0, ctrlat=36.68, ctrlon=-98.35, year=2009,
                        month=5, day=15, hour=21, minute=38, second=0)
    g["param12"] = dict(isfcflx=0, sfcmodel=0, oceanmodel=0, initsfc=1, tsk0=299.28,
                        tmn0=297.28, xland0=2.0, lu0=16, season=1, cecd=3, pertflx=0,
                        cnstce=0.001, cnstcd=0.001, isftcflx=0, iz0tlnd=0,
                        oml_hml0=50.0, oml_gamma=0.14, set_flx=0, cnst_shflx=0.24,
                        cnst_lhflx=5.2e-5, set_znt=0, cnst_znt=0.16, set_ust=0,
                        cnst_ust=0.25, ramp_sgs=1, ramp_time=1800.0, t2p_avg=1)
    g["param4"] = dict(stretch_x=0, dx_inner=1000.0, dx_outer=7000.0,
                       nos_x_len=40000.0, tot_x_len=120000.0)
    g["param5"] = dict(stretch_y=0, dy_inner=1000.0, dy_outer=7000.0,
                       nos_y_len=40000.0, tot_y_len=120000.0)
    g["param6"] = dict(stretch_z=0, ztop=18000.0, str_bot=0.0, str_top=2000.0,
                       dz_bot=125.0, dz_top=500.0)
    g["param7"] = dict(bc_temp=1, ptc_top=250.0, ptc_bot=300.0, viscosity=25.0,
                       pr_num=0.72)
    g["param8"] = {f"var{i}": 0.0 for i in range(1, 21)}
    out = dict(output_format=1, output_filetype=1, output_interp=0)
    for k in ("rain", "sws", "svs", "sps", "srs", "sgs", "sus", "shs", "coldpool",
              "sfcflx", "sfcparams", "sfcdiags", "psfc", "zs"):
        out["output_" + k] = 0
    out["output_rain"] = 1
    for k in ("zh", "basestate", "th", "thpert", "prs", "prspert", "pi", "pipert",
              "rho", "rhopert", "tke", "km", "kh", "qv", "qvpert", "q", "dbz",
              "buoyancy", "u", "upert", "uinterp", "v", "vpert", "vinterp", "w",
              "winterp", "vort", "pv", "uh", "pblten", "dissten", "fallvel", "nm",
              "def", "radten", "cape", "cin", "lcl", "lfc", "pwat", "lwp",
              "thbudget", "qvbudget", "ubudget", "vbudget", "wbudget", "pdcomp"):
        out["output_" + k] = 0
    for k in ("zh", "basestate", "thpert", "prs", "pipert", "rho", "qv", "q", "u",
              "v", "w"):
        out["output_" + k] = 1
    g["param9"] = out
    g["param16"] = dict(restart_format=1, restart_filetype=2, restart_reset_frqtim=True,
                        restart_file_theta=False, restart_file_dbz=False,
                        restart_file_th0=False, restart_file_prs0=False,
                        restart_file_pi0=False, restart_file_rho0=False,
                        restart_file_qv0=False, restart_file_u0=False,
                        restart_file_v0=False, restart_file_zs=False,
                        restart_file_zh=False, restart_file_zf=False,
                        restart_file_diags=False, restart_use_theta=False)
    g["param10"] = {("stat_" + k): 0 for k in (
        "w", "wlevs", "u", "v", "rmw", "pipert", "prspert", "thpert", "q", "tke", "km",
        "kh", "div", "rh", "rhi", "the", "cloud", "sfcprs", "wsp", "cfl", "vort",
        "tmass", "tmois", "qmass", "tenerg", "mo", "tmf", "pcn", "qsrc")}
    g["param13"] = {("prcl_" + k): 0 for k in (
        "th", "t", "prs", "ptra", "q", "nc", "km", "kh", "tke", "dbz", "b", "vpg",
        "vort", "rho", "qsat", "sfc")}
    g["param14"] = dict(dodomaindiag=False, diagfrq=60.0)
    g["param15"] = dict(doazimavg=False, azimavgfrq=3600.0, rlen=300000.0,
                        do_adapt_move=False, adapt_move_frq=3600.0)
    g["param17"] = dict(les_subdomain_shape=1, les_subdomain_xlen=200000.0,
                        les_subdomain_ylen=200000.0, les_subdomain_dlen=200000.0,
                        les_subdomain_trnslen=5000.0)
    g["param18"] = dict(do_recycle_w=False, do_recycle_s=False, do_recycle_e=False,
                        do_recycle_n=False, recycle_width_dx=6.0, recycle_depth_m=1500.0,
                        recycle_cap_loc_m=4000.0, recycle_inj_loc_m=0.0)
    g["param19"] = dict(do_lsnudge=False, do_lsnudge_u=False, do_lsnudge_v=False,
                        do_lsnudge_th=False, do_lsnudge_qv=False, lsnudge_tau=1800.0,
                        lsnudge_start=3600.0, lsnudge_end=7200.0, lsnudge_ramp_time=600.0)
    g["param20"] = dict(do_ib=False, ib_init=4, top_cd=0.4, side_cd=0.4)
    g["param21"] = dict(hurr_vg=40.0, hurr_rad=40000.0, hurr_vgpl=-0.70, hurr_rotate=0.0)
    g["nssl2mom_params"] = dict(alphah=0, alphahl=0.5, ccn=0.6e9, cnor=8.e6, cnoh=4.e4)
    return g


def namelist_text(overrides=None):
    """Full namelist.input text; ``overrides`` maps key -> value (any group)."""
    g = default_groups()
    for key, val in (overrides or {}).items():
        for grp in g.values():
            if key in grp:
                grp[key] = val
                break
        else:
            raise KeyError(f"unknown namelist key {key}")
    lines = []
    for name, grp in g.items():
        lines.append(f" &{name}")
        for k, v in grp.items():
            lines.append(f" {k} = {_fmt(v)},")
        lines.append(" /")
        lines.append("")
    return "\n".join(lines) + "\n"


def run_reference(workdir, overrides=None, threads=None, exe=REF_EXE, timeout=3600,
                  extra_env=None, tables=(), files=None):
    """Run the reference in ``workdir``; returns (stdout text, wall seconds).
    ``tables``: data files of the reference's run directory the case reads
    (e.g. LANDUSE.TBL for a surface model), copied into ``workdir``;
    ``files``: {name: bytes} written into ``workdir`` (the same tables taken
    from a fixture where the reference's directory is absent)."""
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    if os.path.exists(workdir):
        shutil.rmtree(workdir)
    os.makedirs(workdir)
    for t in tables:
        shutil.copy(os.path.join(REF_RUN, t), os.path.join(workdir, t))
    for name, data in (files or {}).items():
        with open(os.path.join(workdir, name), "wb") as f:
            f.write(data)
    with open(os.path.join(workdir, "namelist.input"), "w") as f:
        f.write(namelist_text(overrides))
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    env.setdefault("OMP_STACKSIZE", "512M")
    if extra_env:
        env.update(extra_env)
    t0 = time.time()
    r = subprocess.run([exe], cwd=workdir, env=env, capture_output=True, text=True,
                       timeout=timeout)
    wall = time.time() - t0
    out = r.stdout + r.stderr
    with open(os.path.join(workdir, "cm1.print.out"), "w") as f:
        f.write(out)
    if r.returncode != 0 or "Program terminated normally" not in out:
        raise RuntimeError(f"reference run failed (rc={r.returncode}):\n{out[-3000:]}")
    return out, wall


def step_times(out):
    """Per-step wall times from the reference's 'timing for time step' lines."""
    ts = []
    for line in out.splitlines():
        if "timing for time step" in line:
            m = re.findall(r"[-+]?\d*\.\d+(?:[eE][-+]?\d+)?", line)
            if m:
                ts.append(float(m[-1]))
    return ts


def _parse_ctl(path):
    with open(path) as f:
        txt = f.read()
    m = re.search(r"xdef\s+(\d+)", txt, re.I)
    nx = int(m.group(1))
    ny = int(re.search(r"ydef\s+(\d+)", txt, re.I).group(1))
    nz = int(re.search(r"zdef\s+(\d+)", txt, re.I).group(1))
    nt = int(re.search(r"tdef\s+(\d+)", txt, re.I).group(1))
    dset = re.search(r"dset\s+\^?(\S+)", txt, re.I).group(1)
    lines = txt.splitlines()
    i0 = next(i for i, l in enumerate(lines) if l.lower().startswith("vars"))
    nv = int(lines[i0].split()[1])
    vars_ = []
    for l in lines[i0 + 1:i0 + 1 + nv]:
        p = l.split()
        vars_.append((p[0], max(1, int(p[1]))))
    return dict(nx=nx, ny=ny, nz=nz, nt=nt, dset=dset, vars=vars_)


def read_grads(workdir, ctl):
    """Read one GrADS dataset; returns {var: array[nt, nlev, ny, nx]}."""
    d = _parse_ctl(os.path.join(workdir, ctl))
    data = np.fromfile(os.path.join(workdir, d["dset"]), dtype="<f4")
    nx, ny = d["nx"], d["ny"]
    per_t = sum(nl for _, nl in d["vars"]) * nx * ny
    nt = data.size // per_t
    data = data[:nt * per_t].reshape(nt, per_t)
    out, off = {}, 0
    for name, nl in d["vars"]:
        n = nl * nx * ny
        out[name] = data[:, off:off + n].reshape(nt, nl, ny, nx)
        off += n
    return out
