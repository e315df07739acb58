"""The C-ABI library loads and exports every entry point include/cm1gpu.h
declares (no compute call: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cm1gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cm1g_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("cm1g_create", "cm1g_step", "cm1g_sound", "cm1g_advs", "cm1g_kessler_branch",
              "cm1g_bc", "cm1g_calccflquick", "cm1g_getdivx", "cm1g_budgets"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from cm1_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libcm1gpu.so not built (run __graft_entry__.build())")
    lib = C.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from cm1_amd import _native
    missing = [n for n in declared_functions() if n not in _native.SIGNATURES]
    assert not missing, missing


def test_config_struct_matches_header():
    """Field order of cm1g_config in the header and in the ctypes mirror."""
    from cm1_amd import _native
    text = open(HEADER).read()
    body = re.search(r"typedef struct cm1g_config\s*\{(.*?)\}\s*cm1g_config;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            first, *rest = [p.strip() for p in decl.split(",")]
            names.append(first.split()[-1])
            names.extend(rest)
    assert names == [f[0] for f in _native.Config._fields_]


def test_fortran_module_binds_every_declared_symbol():
    """cm1_amd/fortran/cm1gpu.f90 gives every C entry point a Fortran
    interface (bind(C, name=...)), so a Fortran host can call each one."""
    text = open(os.path.join(ROOT, "cm1_amd", "fortran", "cm1gpu.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(cm1g_\w+)"\)', text))
    missing = [n for n in declared_functions() if n not in bound]
    assert not missing, missing


def test_fortran_module_compiles(tmp_path):
    """The binding module compiles with the image's Fortran compiler."""
    import shutil
    import subprocess
    fc = shutil.which("flang") or shutil.which("flang-new")
    if fc is None:
        pytest.skip("no Fortran compiler")
    src = os.path.join(ROOT, "cm1_amd", "fortran", "cm1gpu.f90")
    r = subprocess.run([fc, "-c", src, "-J", str(tmp_path), "-o", str(tmp_path / "m.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_output_switches_match_header():
    """CM1G_OUT_* bits of the header, the ctypes mirror and the Fortran module
    (the subgrid turbulence's switches included)."""
    from cm1_amd import _native
    text = open(HEADER).read()
    hdr = {n.lower(): 1 << int(b) for n, b in re.findall(r"CM1G_OUT_(\w+) = 1u << (\d+)", text)}
    assert hdr == {**_native.OUT, **_native.OUT_TURB}
    assert not set(_native.OUT) & set(_native.OUT_TURB)
    f90 = open(os.path.join(ROOT, "cm1_amd", "fortran", "cm1gpu.f90")).read()
    fort = {n.lower(): int(v) for n, v in re.findall(r"CM1G_OUT_(\w+?)\s*= (\d+)", f90) if n != "ALL"}
    assert fort == hdr


def test_statistics_switches_match_header():
    """CM1G_STAT_* bits of the header, the ctypes mirror and the Fortran module."""
    from cm1_amd import _native
    text = open(HEADER).read()
    hdr = {n.lower(): 1 << int(b) for n, b in re.findall(r"CM1G_STAT_(\w+) = 1u << (\d+)", text)}
    assert hdr == {**_native.STAT, **_native.STAT_TURB}
    assert not set(_native.STAT) & set(_native.STAT_TURB)
    f90 = open(os.path.join(ROOT, "cm1_amd", "fortran", "cm1gpu.f90")).read()
    fort = {n.lower(): int(v) for n, v in re.findall(r"CM1G_STAT_(\w+?)\s*= (\d+)", f90)
            if n not in ("ALL", "MAXN")}
    assert fort == hdr
    assert int(re.search(r"#define CM1G_STAT_MAXN (\d+)", text).group(1)) == _native.STAT_MAXN
