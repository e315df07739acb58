"""Build the cm1gpu HIP library (``cm1_amd/libcm1gpu.so``) for gfx950.

FP32 semantics follow the reference CPU build: no FMA contraction
(``-ffp-contract=off``), IEEE division and square root, denormals kept.
Packed FP32 instructions are disabled (``-packed-fp32-ops``).
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcm1gpu.so")

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
               "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-result", "-Wno-unused-value"]


def build(verbose=False, jobs=8):
    srcs = sorted(glob.glob(os.path.join(SRC, "*.hip")))
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(SRC, "_obj_" + os.path.basename(s) + ".o")
        cmd = ["hipcc"] + [f for f in HIPCC_FLAGS if f != "-shared"] + ["-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd))
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                          text=True)))
        objs.append(o)
    err = []
    for s, p in procs:
        out = p.communicate()[0]
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
        if p.returncode != 0:
            err.append(f"{s}:\n" + "\n".join(lines))
        elif verbose and lines:
            print("\n".join(lines))
    if err:
        raise RuntimeError("hipcc failed:\n" + "\n".join(err))
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    for o in objs:
        os.remove(o)
    build_fortran_host()
    return LIB


FDIR = os.path.join(HERE, "fortran")
FHOST = os.path.join(FDIR, "cm1g_host")


def build_fortran_host():
    """The ISO_C_BINDING module (fortran/cm1gpu.f90) and the Fortran host
    program (fortran/cm1g_host.f90) with flang, linked against the HIP
    library by rpath.  Skipped when no Fortran compiler is installed."""
    fc = None
    for cand in ("flang", "amdflang", "/opt/rocm/llvm/bin/flang"):
        try:
            subprocess.run([cand, "--version"], capture_output=True, check=True)
            fc = cand
            break
        except (OSError, subprocess.CalledProcessError):
            continue
    if fc is None:
        return None
    mdir = os.path.join(FDIR, "_mod")
    os.makedirs(mdir, exist_ok=True)
    cmd = [fc, "-O2", "-J", mdir, os.path.join(FDIR, "cm1gpu.f90"), os.path.join(FDIR, "cm1g_host.f90"),
           "-o", FHOST, "-L" + HERE, "-lcm1gpu", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
    return FHOST


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
