#!/usr/bin/env python3
"""Throughput of the CM1 dynamical-core large step on one MI355X.

Metric: grid-cell updates per second per full RK3 large step (FP32),
``nx*ny*nz / t_step``, on the 1024x1024x96 supercell workload:
dx=dy=250 m, stretched dz 50-500 m (stretch_z=1), dt=1.5 s, 8 acoustic
substeps per large step, Kessler microphysics, 5th-order advection, open
lateral boundaries, eight seeded +-2 K bubbles; every array resident in HBM.
Synthetic initial state (Weisman-Klemp sounding, u = 30 tanh(z/3 km)).

One JSON line is printed with the roofline of the dominant kernel and a
CPU baseline: the reference CM1 OpenMP build (oracle/_ref/cm1_sp.exe) timed
on a bounded sample of the same case on this host's cores.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE counters (calibrated on a plain copy) and recorded
# in this summary
PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_supercell_1024x1024x96.md")

from cm1_amd import workloads  # noqa: E402
from cm1_amd.workloads import CONFIG1_STRETCH as STRETCH  # noqa: E402

workload_config1 = workloads.config1
workload_ensemble = workloads.ensemble
workload_spinup = workloads.spinup
workload_cbl = workloads.cbl


def bench_ensemble(args):
    """Ensemble throughput: the members stepped one after another (each step
    one captured graph) against the batched ensemble step (one member-indexed
    launch per kernel, one graph per step) at the fixed tile height and at
    the autotuned one.  Prints one JSON line."""
    import torch
    from cm1_amd import model
    import cm1_amd._native as N
    import ctypes as C
    L = N.lib()
    work = workloads.ensemble_small(args.ensemble) if args.small else workload_ensemble(args.ensemble)
    members = []
    for cfgd, cols, fields in work:
        members.append(model.setup(model.make_config(**cfgd), cols, fields))
    del work
    cells = sum(m.ni * m.nj * m.nk for m in members)
    print(f"ensemble: {len(members)} members, {cells} cells", file=sys.stderr, flush=True)

    def seq_step():
        for m in members:
            N.check(L.cm1g_step(m.h, None), "cm1g_step")

    for _ in range(args.warmup):
        seq_step()
    for m in members:
        m.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        seq_step()
    for m in members:
        m.sync()
    ms_seq = (time.perf_counter() - t0) * 1000.0 / args.steps
    print(f"ensemble: one by one {ms_seq:.2f} ms/step", file=sys.stderr, flush=True)

    e = model.Ensemble(members)

    def ens_timed(nsteps):
        e.sync()
        t = time.perf_counter()
        e.step(nsteps)
        e.sync()
        return (time.perf_counter() - t) * 1000.0 / nsteps

    e.set_tiles(4)
    e.step(max(1, args.warmup))
    ms_fixed = ens_timed(args.steps)
    print(f"ensemble: batched, 4-row tiles {ms_fixed:.2f} ms/step", file=sys.stderr, flush=True)
    best, order, tune_ms = e.tune_schedule(max(2, args.steps // 4))
    ms_tuned = ens_timed(args.steps)
    print(f"ensemble: batched, tuned ({best}-row tiles, member order {order}) {ms_tuned:.2f} ms/step",
          file=sys.stderr, flush=True)
    e.close()
    shapes = sorted({(m.ni, m.nj, m.nk) for m in members})
    line = {
        "metric": "ensemble member cell-updates/s per RK3 step (FP32)",
        "value": cells / (ms_tuned / 1000.0),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_tuned,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"supercell ensemble, {len(members)} members, " +
                                ("(nx,ny) in {24,32,48,64}, nz in {24,32,40}" if args.small else
                                 "(nx,ny) in {96,128,192,256}, nz in {40,56,64}") +
                                ", dx=1km dt=6s nsound=6 Kessler"),
                   "members": len(members), "cells": cells, "shapes": len(shapes),
                   "ms_per_step_one_by_one": round(ms_seq, 3),
                   "ms_per_step_batched_fixed_tiles": round(ms_fixed, 3),
                   "ms_per_step_batched_tuned": round(ms_tuned, 3),
                   "tuned_tile_rows": best,
                   "tuned_member_order": ["as given", "deepest columns first", "largest first"][order],
                   "tuner_ms_per_step": {f"order{o}_ty{h}": round(v, 3) for (o, h), v in tune_ms.items()},
                   "gain_batched_tuned_vs_one_by_one": round(ms_seq / ms_tuned, 3)},
    }
    print(json.dumps(line))
    return 0


def bench_surrogate(args):
    """Learned Kessler surrogate (one JSON line).

    (a) The 1024x1024x96 supercell itself: a plain spin-up of --spinup steps,
    (state, change-over-kessler) pairs harvested on 6 steps spread over its
    second half, the MLP trained on the GPU, then 60 further steps from the
    end of the spin-up with plain kessler, with the network in its place
    and with a zero network (no autoconversion, accretion or evaporation):
    the domain rain accumulation and step time of each.
    (b) A 128x128x40 dx=1 km case (dt=6 s, as the ensemble's members): the
    same harvest, training and rain comparison, then a 48-member ensemble
    batched with that network against its members stepped one by one with
    plain Kessler."""
    import torch
    from cm1_amd import model
    from cm1_amd import surrogate as S
    import cm1_amd._native as N
    L = N.lib()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    def train(pairs):
        norm = S.fit_norm(pairs)
        x, t = S.normalise(pairs, norm)
        net = S.Net(seed=1)
        net.set_norm(norm)
        t0 = time.perf_counter()
        losses = net.train(x, t, steps=args.train_steps, batch=65536, lr=1.0e-3, log_every=2000)
        return net, norm, losses, time.perf_counter() - t0

    def compare(m, state, net, norm, nsteps=60):
        zero = S.Net(seed=2)
        zero.set_params(np.zeros(zero.nparams, np.float32))
        zn = norm.copy()
        zn[16:20] = 0.0
        zero.set_norm(zn)
        out = {}
        for name, n in (("plain", None), ("surrogate", net), ("no_network", zero)):
            m.load_state(state)
            S.use(m, n)
            r0 = m.scalars()["train"]
            c0 = S.counts(m)["fallback_cells"]
            m.step(1)
            m.sync()
            t1 = time.perf_counter()
            m.step(nsteps - 1)
            m.sync()
            ms = (time.perf_counter() - t1) * 1000.0 / (nsteps - 1)
            out[name] = dict(rain=m.scalars()["train"] - r0, ms=ms,
                             fallback=S.counts(m)["fallback_cells"] - c0)
        S.use(m, None)
        zero.close()
        p = out["plain"]["rain"]
        for k in out:
            out[k]["rel_error"] = abs(out[k]["rain"] - p) / p if p > 0 else None
        return out

    # (a) the 1024x1024x96 supercell
    cfg1, cols1, fields1 = workload_config1()
    m = model.setup(model.make_config(**cfg1), cols1, fields1)
    del fields1
    n1 = args.spinup
    at = {n1 // 2 + q * (n1 - n1 // 2) // 6 for q in range(6)}
    t0 = time.perf_counter()
    pairs = S.harvest_spread(m, n1, at, 16_000_000, 50)
    state = m.save_state()
    log(f"surrogate (a): {n1} plain steps in {time.perf_counter() - t0:.1f} s, {pairs.shape[0]} pairs")
    net_a, norm_a, losses_a, tt_a = train(pairs)
    del pairs
    log(f"surrogate (a): trained in {tt_a:.1f} s, losses {losses_a}")
    res_a = compare(m, state, net_a, norm_a)
    log(f"surrogate (a): {res_a}")
    cells1 = cfg1["ni"] * cfg1["nj"] * cfg1["nk"]
    ncell1 = cells1 * 60
    m.close()
    del state

    # (b) the 128x128x40 case and the ensemble
    cfgd, cols, fields = workload_spinup()
    mb = model.setup(model.make_config(**cfgd), cols, fields)
    nb = 300
    pairs = S.harvest_spread(mb, nb, set(range(nb // 2, nb)), 12_000_000, 20)
    state = mb.save_state()
    net_b, norm_b, losses_b, tt_b = train(pairs)
    del pairs
    res_b = compare(mb, state, net_b, norm_b)
    log(f"surrogate (b): {res_b}")
    mb.close()
    work = workload_ensemble(48)
    members = [model.setup(model.make_config(**c_), co, fi) for c_, co, fi in work]
    del work
    cells_e = sum(mm.ni * mm.nj * mm.nk for mm in members)
    for mm in members:
        N.check(L.cm1g_step(mm.h, None), "cm1g_step")
    for mm in members:
        mm.sync()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        for mm in members:
            N.check(L.cm1g_step(mm.h, None), "cm1g_step")
    for mm in members:
        mm.sync()
    ms_seq = (time.perf_counter() - t1) * 1000.0 / args.steps
    for mm in members:
        S.use(mm, net_b)
    e = model.Ensemble(members)
    e.step(1)
    e.sync()
    t1 = time.perf_counter()
    e.step(args.steps)
    e.sync()
    ms_ens = (time.perf_counter() - t1) * 1000.0 / args.steps
    e.close()
    log(f"surrogate (b): ensemble {ms_seq:.2f} ms one by one (plain), {ms_ens:.2f} ms batched with the surrogate")
    ms_sur = res_a["surrogate"]["ms"]
    line = {
        "metric": "cell-updates/s per RK3 step (FP32), Kessler surrogate in place of kessler",
        "value": cells1 / (ms_sur / 1000.0),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": 59,
        "warmup": 1,
        "ms_per_step": ms_sur,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (network: bf16 operands, f32 accumulation)",
        "data": "synthetic",
        "config": {"workload": f"supercell 1024x1024x96 (as the main bench) after {n1} plain steps, 60 steps "
                               "with the surrogate; plus a 128x128x40 case and a 48-member ensemble",
                   "train_steps": args.train_steps, "batch": 65536,
                   "a_1024": {"train_seconds": round(tt_a, 2), "final_loss": losses_a[-1][1],
                              "runs": res_a, "fallback_fraction": res_a["surrogate"]["fallback"] / ncell1},
                   "b_128": {"train_seconds": round(tt_b, 2), "final_loss": losses_b[-1][1], "runs": res_b},
                   "ensemble_cells": cells_e,
                   "ensemble_ms_per_step_one_by_one_plain": round(ms_seq, 3),
                   "ensemble_ms_per_step_batched_surrogate": round(ms_ens, 3),
                   "ensemble_gain": round(ms_seq / ms_ens, 3)},
    }
    print(json.dumps(line))
    return 0


def bench_cbl(args):
    """Throughput of the dry CBL workload (one JSON line)."""
    import torch
    from cm1_amd import model
    t0 = time.perf_counter()
    cfgd, cols, fields = workload_cbl()
    cfg = model.make_config(**cfgd)
    m = model.setup(cfg, cols, fields, turb=workloads.CBL_TURB)
    del fields
    print(f"cbl: set-up {time.perf_counter() - t0:.1f} s, nsound {cfg.nsound}", file=sys.stderr, flush=True)
    tm0, _ = m.budgets()
    m.step(max(1, args.warmup))
    m.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    cfl = m.step(args.steps)
    m.sync()
    ms = (time.perf_counter() - t1) * 1000.0 / args.steps
    tm1, _ = m.budgets()
    cells = cfgd["ni"] * cfgd["nj"] * cfgd["nk"]
    line = {
        "metric": "cell-updates/s per RK3 step (FP32)",
        "value": cells / (ms / 1000.0),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(1, args.warmup),
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "dry CBL 1536x1536x128 dx=dz=20m dt=0.25s periodic sounde WENO5 6 tracers, "
                               "TKE subgrid turbulence (implicit vertical diffusion), semi-slip surface, "
                               "0.24 K m/s surface heat flux",
                   "nsound": cfg.nsound, "cflmax": cfl,
                   "dry_mass_rel_change": (tm1 - tm0) / tm0},
    }
    print(json.dumps(line))
    return 0


def host_cpu():
    """The host's CPU model, sockets and physical cores (/proc/cpuinfo), the
    CPUs this process may run on, and the CPU quota of its cgroup."""
    model, pairs = "unknown CPU", set()
    phys = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name") and model == "unknown CPU":
                    model = ln.split(":", 1)[1].strip()
                elif ln.startswith("physical id"):
                    phys = ln.split(":", 1)[1].strip()
                elif ln.startswith("core id"):
                    pairs.add((phys, ln.split(":", 1)[1].strip()))
    except OSError:
        pass
    sockets = len({p for p, _ in pairs}) or None
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    return dict(cpu=model, sockets=sockets, physical_cores=len(pairs) or None, logical_cpus=os.cpu_count(),
                affinity_cpus=affinity, cgroup_cpu_quota=quota)


def cpu_baseline(threads=None):
    """Reference CM1 (FP32 OpenMP build) on a 256x256x96 sample of the case,
    plus a one-thread run of the 64x64x40 supercell."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    host = host_cpu()
    if threads is None:
        # every CPU this job may use: the cgroup quota where one is set (the
        # host's other cores belong to other jobs), else the affinity set
        threads = int(host["cgroup_cpu_quota"] or host["affinity_cpus"])
        threads = max(1, min(threads, host["affinity_cpus"]))
    base = dict(unit="cell-updates/s", cores=threads, threads=threads, kind="reference", **host)
    try:
        import refrun
    except Exception as e:  # pragma: no cover
        return dict(base, value=None, sample=f"unavailable: {e}")
    if not os.path.exists(refrun.REF_EXE):
        return dict(base, value=None, sample="oracle/_ref/cm1_sp.exe not built")
    env = {"OMP_PROC_BIND": "spread", "OMP_PLACES": "cores"}

    def timed(ov, nthreads, tag):
        wd = os.path.join("/tmp", f"cm1_cpu_baseline_{tag}_{os.getpid()}")
        out, _ = refrun.run_reference(wd, ov, threads=nthreads, timeout=900, extra_env=env)
        ts = refrun.step_times(out)
        if len(ts) > 1:
            ts = ts[1:]
        return float(np.median(ts)) if ts else float("nan")

    # the 1024x1024x96 case on a 256x256x96 horizontal sample: same dx,
    # stretched grid, Kessler, open boundaries; dt = 3 s makes the
    # reference's own dtsmall choose nsound = 8 (15 acoustic substeps per
    # step, as the GPU runs at dt = 1.5 s with nsound fixed at 8), so a step
    # does the same work
    ni, nj, nk, nsteps, dt = 256, 256, 96, 6, 3.0
    ov = dict(nx=ni, ny=nj, nz=nk, dx=250.0, dy=250.0, dz=500.0, dtl=dt, timax=nsteps * dt,
              tapfrq=1.0e9, stretch_z=1, ztop=STRETCH["ztop"], str_bot=STRETCH["str_bot"],
              str_top=STRETCH["str_top"], dz_bot=STRETCH["dz_bot"], dz_top=STRETCH["dz_top"])
    try:
        t = timed(ov, threads, "sample")
    except Exception as e:
        return dict(base, value=None, sample=f"run failed: {str(e)[:200]}")
    res = dict(base, value=ni * nj * nk / t,
               sample=(f"reference CM1 FP32 OpenMP build ({threads} threads, OMP_PROC_BIND=spread, "
                       f"OMP_PLACES=cores), {ni}x{nj}x{nk} horizontal sample of the 1024x1024x96 case "
                       f"(dx=250 m, stretched dz, Kessler, open BCs, dt={dt:g} s so that its dtsmall picks "
                       f"nsound=8: 15 acoustic substeps per step as on the GPU), {nsteps} steps, median "
                       f"per-step time {t:.3f} s excluding step 1"))
    # the 64x64x40 supercell (dx = 1 km, dz = 500 m, dt = 6 s) on one thread
    try:
        t1 = timed(dict(nx=64, ny=64, nz=40, timax=8 * 6.0, tapfrq=1.0e9), 1, "config0")
        res["config0_1thread"] = dict(value=64 * 64 * 40 / t1, unit="cell-updates/s",
                                      ms_per_step=round(t1 * 1000.0, 2), steps=8)
    except Exception as e:
        res["config0_1thread"] = dict(value=None, error=str(e)[:200])
    return res


def measured_traffic(kernel):
    """Bytes per launch of `kernel` from the committed counter summary."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith(f"traffic {kernel} bytes_per_launch="):
                    return float(line.split("=", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ni", type=int, default=1024)
    ap.add_argument("--nj", type=int, default=1024)
    ap.add_argument("--nk", type=int, default=96)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="threads of the CPU baseline (default: the CPUs this job may use)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy (reference kernel for HBM counters)")
    ap.add_argument("--small", action="store_true",
                    help="with --ensemble: members from the small-grid draw (24-64 columns, 24-40 levels)")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="benchmark an ensemble of this many members instead (batched steps)")
    ap.add_argument("--cbl", action="store_true",
                    help="benchmark the dry convective-boundary-layer workload (1536x1536x128) instead")
    ap.add_argument("--surrogate", action="store_true",
                    help="train and benchmark the learned Kessler surrogate instead")
    ap.add_argument("--spinup", type=int, default=1000,
                    help="plain 1024x1024x96 steps before the surrogate comparison")
    ap.add_argument("--train-steps", type=int, default=10000, help="Adam steps for the surrogate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU horizontal decomposition is not implemented", file=sys.stderr)
        return 2

    if args.ensemble > 0:
        return bench_ensemble(args)
    if args.surrogate:
        return bench_surrogate(args)
    if args.cbl:
        return bench_cbl(args)

    import torch
    from cm1_amd import model
    import cm1_amd._native as N
    import ctypes as C
    from cm1_amd import bytes_model

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cfgd, cols, fields = workload_config1(args.ni, args.nj, args.nk)
    cfg = model.make_config(**cfgd)
    m = model.setup(cfg, cols, fields)
    del fields
    for _ in range(args.warmup):
        m.step(1)
    m.sync()
    L = N.lib()

    def timed(nsteps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c = 0.0
        for _ in range(nsteps):
            c = m.step(1)
        m.sync()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / nsteps, c

    # the production path: each step replayed as one captured HIP graph
    ms, cfl = timed(args.steps)
    cells = cfgd["ni"] * cfgd["nj"] * cfgd["nk"]
    value = cells / (ms / 1000.0)
    # the same steps launched kernel by kernel, with HIP events around every
    # launch of the acoustic substep kernel on the model's stream
    kt_ok = L.cm1g_ktimer_start(m.h, N.KT_SOUND_SUBSTEP) == 0
    ms_eager, _ = timed(args.steps)

    roof = None
    if kt_ok:
        tot = C.c_double()
        nl = C.c_int64()
        nbytes = C.c_double()
        name = C.create_string_buffer(64)
        rc = L.cm1g_ktimer_stop(m.h, C.byref(tot), C.byref(nl), C.byref(nbytes), name, 64)
        if rc == 0 and nl.value > 0 and tot.value > 0:
            avg_ms = tot.value / nl.value
            gbs = nbytes.value / (avg_ms * 1e-3) / 1e9
            roof = dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(gbs / HBM_PEAK_GBS, 4), traffic=measured_traffic(name.value.decode()),
                        kernel=name.value.decode(), avg_ms=round(avg_ms, 4),
                        launches=nl.value, bytes_per_launch=nbytes.value,
                        fields_per_cell=bytes_model.SOUND_SUBSTEP,
                        achieved_fused_distinct=round(gbs * bytes_model.SOUND_SUBSTEP_FUSED /
                                                      bytes_model.SOUND_SUBSTEP, 1))
    # whole-step algorithmic rate (frozen byte table)
    step_bytes = cells * bytes_model.bytes_per_cell_step(cfgd["nsound"])
    step_gbs = step_bytes / (ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_threads)

    line = {
        "metric": "cell-updates/s per RK3 step (FP32)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"supercell {cfgd['ni']}x{cfgd['nj']}x{cfgd['nk']} dx=250m stretched-dz "
                               f"dt=1.5s nsound=8 Kessler open-BC 8 bubbles",
                   "nx": cfgd["ni"], "ny": cfgd["nj"], "nz": cfgd["nk"], "nsound": cfgd["nsound"],
                   "cflmax": cfl,
                   "ms_per_step_kernel_by_kernel": round(ms_eager, 3),
                   "step_algorithmic_GBps": round(step_gbs, 1),
                   "step_bytes_per_cell": bytes_model.bytes_per_cell_step(cfgd["nsound"])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
