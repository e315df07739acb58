#!/usr/bin/env python3
"""Build the reference CM1 (cm1r21.1) CPU executable from its Fortran sources.

Test infrastructure only: this compiles the reference's own ``src/*.F`` where
they lie (default ``/root/reference/src``) into ``oracle/_ref/`` with flang,
so that the parity tests can run the real reference on the same inputs.
Nothing from the reference is copied into the repository; the preprocessed
``.f90`` files and objects live under ``oracle/_ref/`` (git-ignored).

Recipe (our own; the reference Makefile is not used):
  * every ``src/*.F`` is run through ``cpp -C -P -traditional`` with
    ``-DOPENMP`` (and ``-DDP`` for the FP64 variant), as the reference's
    own preprocessing step does;
  * flang 22 ``-O2 -fopenmp -ffp-contract=off`` (no FMA, the reference's
    default for its CPU builds), FP64 variant adds
    ``-fdefault-real-8 -fdefault-double-8``;
  * module dependencies are found by scanning ``module``/``use`` lines and
    files are compiled level by level in parallel.

One compiler work-around: ``src/radlib3d.F`` names a dummy argument ``INT``
(``radlib3d.F:27,237,308,491``).  flang writes ``int(n,kind=8)`` array
bounds into the generated ``.mod`` and then cannot re-read them because
``int`` resolves to that dummy.  The dummy is renamed ``INTSTRD`` in the
preprocessed copy (only ``INT`` tokens not followed by ``(``, so the
intrinsic call in ``search`` at ``radlib3d.F:457`` is untouched).
Radiation is never called by the configurations we run (``radopt=0``).
"""
import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def preprocess(src, dst, defs):
    cmd = ["cpp", "-C", "-P", "-traditional", "-Wno-invalid-pp-token",
           "-ffreestanding"] + defs + [src]
    out = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
    if os.path.basename(src) == "radlib3d.F":
        out = re.sub(r"\bINT\b(?!\s*\()", "INTSTRD", out)
    if os.path.exists(dst):
        with open(dst) as f:
            if f.read() == out:
                return False
    with open(dst, "w") as f:
        f.write(out)
    return True


def scan(path):
    mods, uses = set(), set()
    with open(path, errors="replace") as f:
        for line in f:
            s = line.strip().lower()
            m = re.match(r"module\s+(\w+)\s*$", s)
            if m and m.group(1) != "procedure":
                mods.add(m.group(1))
            m = re.match(r"use\s*(?:,\s*\w+\s*::)?\s*(\w+)", s)
            if m:
                uses.add(m.group(1))
    return mods, uses


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--src", default="/root/reference/src")
    ap.add_argument("--out", default=os.path.join(HERE, "_ref"))
    ap.add_argument("--dp", action="store_true", help="FP64 (-DDP) build")
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args()
    if not os.path.isdir(args.src):
        print(f"reference sources not found at {args.src}; skipping", file=sys.stderr)
        return 0
    tag = "dp" if args.dp else "sp"
    bdir = os.path.join(args.out, "build_" + tag)
    os.makedirs(bdir, exist_ok=True)
    defs = ["-DOPENMP"] + (["-DDP"] if args.dp else [])
    fflags = ["-O2", "-fopenmp", "-ffp-contract=off"]
    if args.dp:
        fflags += ["-fdefault-real-8", "-fdefault-double-8"]
    srcs = sorted(f for f in os.listdir(args.src) if f.endswith(".F"))
    changed = {}
    for f in srcs:
        changed[f] = preprocess(os.path.join(args.src, f),
                                os.path.join(bdir, f[:-2] + ".f90"), defs)
    provides, needs = {}, {}
    for f in srcs:
        mods, uses = scan(os.path.join(bdir, f[:-2] + ".f90"))
        for m in mods:
            provides[m] = f
        needs[f] = uses
    deps = {f: {provides[u] for u in needs[f] if u in provides and provides[u] != f}
            for f in srcs}
    # topological levels
    level, done = [], set()
    while len(done) < len(srcs):
        lv = [f for f in srcs if f not in done and deps[f] <= done]
        if not lv:
            raise SystemExit("cyclic module dependencies")
        level.append(lv)
        done |= set(lv)
    # a file must be rebuilt when it or anything it depends on was rebuilt
    rebuilt = set()

    def compile_one(f):
        # the object is written under a temporary name and renamed, so a
        # compile that is interrupted never leaves a stub that a later
        # incremental build would take as up to date
        obj = os.path.join(bdir, f[:-2] + ".o")
        tmp = obj + ".part"
        cmd = ["flang"] + fflags + ["-c", f[:-2] + ".f90", "-o", tmp]
        r = subprocess.run(cmd, cwd=bdir, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{f}: {r.stderr[-3000:]}")
        os.replace(tmp, obj)
        return f

    for lv in level:
        todo = [f for f in lv
                if changed[f] or (deps[f] & rebuilt)
                or not os.path.exists(os.path.join(bdir, f[:-2] + ".o"))
                or os.path.getsize(os.path.join(bdir, f[:-2] + ".o")) == 0]
        with cf.ThreadPoolExecutor(args.j) as ex:
            for f in ex.map(compile_one, todo):
                rebuilt.add(f)
    exe = os.path.join(args.out, "cm1_" + tag + ".exe")
    if rebuilt or not os.path.exists(exe):
        objs = [os.path.join(bdir, f[:-2] + ".o") for f in srcs]
        subprocess.run(["flang"] + fflags + objs + ["-o", exe], check=True)
    print(exe)
    return 0


if __name__ == "__main__":
    sys.exit(main())
