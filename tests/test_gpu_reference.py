"""HIP path against the reference CM1's own output (golden vectors written
by the reference executable, tests/golden/make_ref_fixture.py), from the
reference's t=0 state over 10 (and 60) large steps.

The HIP kernels follow the reference's operation order without
contraction and evaluate exp/pow with the host C library's method
(cm1g_libm.h), so every prognostic field (u, v, w, theta', pi', qv, qc, qr,
rho, prs, and the passive tracers where present) must equal the reference
build's output bit for bit; the relative L2 distance (0) is printed beside
the 1e-5 tolerance a single-precision port is held to.
"""
import glob
import os

import numpy as np
import pytest

from common import PROG, load_fixture, make_gpu, nsound_for, rel_l2

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
# the supercell fixtures (the LES ones, ref_les_*, have their own tests)
FIXTURES = sorted(p for p in glob.glob(os.path.join(HERE, "golden", "ref_*.npz"))
                  if not os.path.basename(p).startswith("ref_les_"))
STAG = dict(PROG)
TOL = 1.0e-5


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_matches_reference_fixture(require_gpu, path):
    cfgd, cols, fields, nsteps, target, z = load_fixture(path)
    cfgd["nsound"] = nsound_for(cfgd, cols["zf"])
    g = make_gpu(cfgd, cols, fields)
    g.step(nsteps)
    for name, ref in target.items():
        got = g.interior(name, STAG.get(name))
        nbad = int(np.count_nonzero(got != ref))
        err = rel_l2(got, ref)
        print(name, "gpu vs reference rel L2", err, "differing values", nbad)
        assert err <= TOL, (name, err)
        assert nbad == 0, (name, nbad)
