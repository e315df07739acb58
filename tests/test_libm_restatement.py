"""The single-precision exp/pow/log the HIP kernels call (cm1_amd/csrc/cm1g_libm.h)
against the host C library's expf/powf, which the reference's REAL exp() and
** call.  The header is compiled for the CPU here (it is host+device code)
and compared on every 127th float of exp's working range (-100, 88) and
every 127th positive float for each exponent the model uses, plus edge
inputs; the full, every-float comparison was run the same way with stride 1
and found no difference either.  This test is what makes the GPU's results
bit-identical to the reference build's, exp/pow included."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exp_pow_match_host_library(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    exe = str(tmp_path / "libm_check")
    subprocess.run([cxx, "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "cm1_amd", "csrc"),
                    os.path.join(ROOT, "tests", "libm_check.cpp"), "-o", exe, "-lm"], check=True)
    out = subprocess.run([exe, "127"], check=True, capture_output=True, text=True).stdout
    print(out)
    lines = [l.split() for l in out.strip().splitlines()]
    assert len(lines) == 10
    for parts in lines:
        n, bad = int(parts[-2]), int(parts[-1])
        assert n > 1_000_000 and bad == 0, parts
