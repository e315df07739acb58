#!/usr/bin/env python3
"""Build the CM1 model program with its dynamical-core modules replaced by
the GPU drop-ins of ``cm1_amd/fortran/dropin/``.

The reference's own Fortran sources (default ``/root/reference/src``) are
compiled where they lie, except the files that define a module the drop-in
directory defines (``adv_module`` in ``adv.F``, ``sound_module`` in
``sound.F``, ...): those are left out of the build entirely, and every
caller (``solve2.F``, ``mp_driver.F``, ``pdcomp.F``, ``cm1.F``) is compiled
unchanged against the drop-in modules of the same name.  The program links
``cm1_amd/libcm1gpu.so``.  Nothing from the reference is copied into the
repository: the preprocessed sources, objects and the executable live in the
git-ignored ``cm1_amd/fortran/_cm1/``.

Recipe: every reference ``.F`` goes through ``cpp -C -P -traditional
-DOPENMP`` as the reference's own preprocessing step does; flang ``-O2
-fopenmp -ffp-contract=off`` (the reference's CPU build has no FMA); module
dependencies are found from ``module``/``use`` lines and files compile level
by level.  One compiler work-around: ``src/radlib3d.F`` names a dummy
argument ``INT``, which flang cannot re-read from the ``.mod`` it writes, so
the preprocessed copy renames that dummy (``INT`` tokens not followed by
``(``); radiation is never called by the configurations the drop-in runs.

Usage: python cm1_amd/fortran/build_dropin.py [--src DIR] [-j N]
"""
import argparse
import concurrent.futures as cf
import glob
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_cm1")
EXE = os.path.join(OUT, "cm1_gpu.exe")
FFLAGS = ["-O2", "-fopenmp", "-ffp-contract=off"]


def _modules_and_uses(text):
    mods, uses = set(), set()
    for line in text.splitlines():
        s = line.strip().lower()
        m = re.match(r"module\s+(\w+)\s*$", s)
        if m and m.group(1) != "procedure":
            mods.add(m.group(1))
        m = re.match(r"use\s*(?:,\s*\w+\s*::)?\s*(\w+)", s)
        if m:
            uses.add(m.group(1))
    return mods, uses


def _write_if_changed(path, text):
    if os.path.exists(path):
        with open(path) as f:
            if f.read() == text:
                return False
    with open(path, "w") as f:
        f.write(text)
    return True


def build(src="/root/reference/src", jobs=8, verbose=False):
    if not os.path.isdir(src):
        print(f"reference sources not found at {src}; drop-in program not built", file=sys.stderr)
        return None
    os.makedirs(OUT, exist_ok=True)
    # our sources: the C-ABI binding module and the drop-in modules
    ours = [os.path.join(HERE, "cm1gpu.f90")] + sorted(glob.glob(os.path.join(HERE, "dropin", "*.f90")))
    replaced = set()
    units = {}  # name -> (path compiled, provides, uses, changed)
    for p in ours:
        with open(p) as f:
            text = f.read()
        mods, uses = _modules_and_uses(text)
        replaced |= mods
        units[os.path.basename(p)] = [p, mods, uses, True]
    replaced.discard("cm1gpu")
    left_out = []
    for f in sorted(os.listdir(src)):
        if not f.endswith(".F"):
            continue
        cmd = ["cpp", "-C", "-P", "-traditional", "-Wno-invalid-pp-token", "-ffreestanding",
               "-DOPENMP", os.path.join(src, f)]
        text = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
        if f == "radlib3d.F":
            text = re.sub(r"\bINT\b(?!\s*\()", "INTSTRD", text)
        mods, uses = _modules_and_uses(text)
        if mods & replaced:
            left_out.append(f)
            continue
        dst = os.path.join(OUT, f[:-2] + ".f90")
        changed = _write_if_changed(dst, text)
        units[f] = [dst, mods, uses, changed]
    provides = {}
    for name, (_, mods, _, _) in units.items():
        for m in mods:
            provides[m] = name
    deps = {n: {provides[u] for u in u_[2] if u in provides and provides[u] != n} for n, u_ in units.items()}
    order, done = [], set()
    while len(done) < len(units):
        lv = sorted(n for n in units if n not in done and deps[n] <= done)
        if not lv:
            raise SystemExit("cyclic module dependencies")
        order.append(lv)
        done |= set(lv)
    rebuilt = set()

    def obj_of(n):
        return os.path.join(OUT, os.path.splitext(n)[0] + ".o")

    def compile_one(n):
        path = units[n][0]
        obj = obj_of(n)
        tmp = obj + ".part"
        cmd = ["flang"] + FFLAGS + ["-J", OUT, "-c", path, "-o", tmp]
        r = subprocess.run(cmd, cwd=OUT, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{n}: {r.stderr[-3000:]}")
        os.replace(tmp, obj)
        return n

    for lv in order:
        todo = []
        for n in lv:
            obj = obj_of(n)
            stale = (not os.path.exists(obj) or os.path.getsize(obj) == 0
                     or os.path.getmtime(obj) < os.path.getmtime(units[n][0]))
            if units[n][3] and n.endswith(".F") or stale or (deps[n] & rebuilt):
                todo.append(n)
        with cf.ThreadPoolExecutor(jobs) as ex:
            for n in ex.map(compile_one, todo):
                rebuilt.add(n)
    lib = os.path.join(PKG, "libcm1gpu.so")
    if rebuilt or not os.path.exists(EXE) or os.path.getmtime(EXE) < os.path.getmtime(lib):
        objs = [obj_of(n) for n in sorted(units)]
        cmd = ["flang"] + FFLAGS + objs + ["-o", EXE, "-L" + PKG, "-lcm1gpu",
                                          "-Wl,-rpath,$ORIGIN/../.."]
        subprocess.run(cmd, check=True)
    if verbose:
        print(f"left out (replaced by drop-ins): {left_out}")
    return EXE


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--src", default="/root/reference/src")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build(a.src, a.j, verbose=True))
