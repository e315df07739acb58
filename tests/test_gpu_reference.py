"""HIP path against the reference CM1's own output (golden vectors written
by the reference executable, tests/golden/make_ref_fixture.py): from the
reference's t=0 state, 10 large steps on the GPU must match the
reference's state to per-field relative L2 <= 1e-5,
or 3x the case's FP32 noise floor where that is larger (see
tests/common.py:ulp_floor; only libm pow/exp results differ, by one ulp in
~0.06% of calls).
"""
import glob
import os

import numpy as np
import pytest

from common import PROG, load_fixture, make_gpu, nsound_for, parity_bound, rel_l2, ulp_floor

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_*.npz")))
STAG = dict(PROG)
TOL = 1.0e-5


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_matches_reference_fixture(require_gpu, path):
    cfgd, cols, fields, nsteps, target, z = load_fixture(path)
    cfgd["nsound"] = nsound_for(cfgd)
    g = make_gpu(cfgd, cols, fields)
    g.step(nsteps)
    floor = ulp_floor(cfgd, cols, fields, nsteps)
    for name, ref in target.items():
        got = g.interior(name, STAG[name])
        err = rel_l2(got, ref)
        print(name, err, int(np.count_nonzero(got != ref)), "floor", floor[name])
        assert err <= parity_bound(floor, name), (name, err, floor[name])
