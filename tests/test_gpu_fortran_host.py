"""The Fortran host binding (cm1_amd/fortran/cm1gpu.f90, ISO_C_BINDING) end
to end: a flang-built host program loads a state in the reference's array
shapes through cm1g_put_array, steps it with cm1g_step and reads it back
with cm1g_get_array.  The result must equal, bit for bit, the same case
driven through the Python host (same library, same calls)."""
import os
import subprocess

import numpy as np
import pytest

from common import case, make_gpu

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "cm1_amd", "fortran", "cm1g_host")

IN_FIELDS = [("ua", "u"), ("va", "v"), ("wa", "w"), ("tha", None), ("ppi", None),
             ("qa0", None), ("qa1", None), ("qa2", None)]
OUT_FIELDS = IN_FIELDS + [("rho", None), ("prs", None)]


def ref_shape(a, ni, nj, nk, stag):
    """Logical device box (k=0..nk+2, j=-2..nj+4, i=-2..ni+4) -> the
    reference array (kb:ke(+1), jb:je(+1), ib:ie(+1))."""
    ei, ej, ek = (stag == "u"), (stag == "v"), (stag == "w")
    return a[:nk + 2 + ek, :nj + 6 + ej, :ni + 6 + ei]


@pytest.mark.skipif(not os.path.exists(HOST), reason="Fortran host not built (no flang)")
def test_fortran_host_matches_python_host(require_gpu, tmp_path):
    from cm1_amd import _native as N
    nsteps = 3
    cfgd, cols, fields = case(ni=40, nj=36, nk=30, amp=3.0, bc=2)
    cfg = N.Config(**{k: cfgd[k] for k, _ in N.Config._fields_})
    ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        f.write(np.int32(nsteps).tobytes())
        f.write(bytes(cfg))
        for n in ("zf", "zh", "mh", "mf", "pi0", "th0", "qv0", "u0", "v0"):
            c = np.zeros(nk + 3, np.float32)
            c[:] = np.asarray(cols.get(n, np.zeros(nk + 3)), np.float32)[:nk + 3]
            f.write(c.tobytes())
        for n, stag in IN_FIELDS:
            f.write(np.ascontiguousarray(ref_shape(fields[n], ni, nj, nk, stag)).tobytes())
    print("fortran host: 40x36x30 open, 3 steps", flush=True)
    r = subprocess.run([HOST, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    g = make_gpu(cfgd, cols, fields)
    cfl = g.step(nsteps)
    tm, tw = g.budgets()
    raw = open(fout, "rb").read()
    head = np.frombuffer(raw[:24], np.float64)
    assert head[0] == np.float64(np.float32(cfl))
    assert head[1] == tm and head[2] == tw
    off = 24
    for n, stag in OUT_FIELDS:
        want = np.ascontiguousarray(ref_shape(g.get(n), ni, nj, nk, stag))
        got = np.frombuffer(raw[off:off + want.nbytes], np.float32).reshape(want.shape)
        off += want.nbytes
        # interior plus the halos the step fills
        assert np.array_equal(got[1:nk + 1], want[1:nk + 1]), n
    assert off == len(raw)
