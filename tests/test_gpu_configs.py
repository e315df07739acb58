"""The batched-ensemble and learned-surrogate workloads in their own shapes.

- The 48-member supercell ensemble, drawn exactly as the ensemble
  benchmark draws it ((nx, ny) in {96,128,192,256}, nz in {40,56,64},
  qv scale U(0.9,1.1), shear U(10,30) m/s, bubble U(1,3) K): after two
  batched steps every member equals the same member stepped alone, bit for
  bit, and the smallest and the largest members equal the CPU oracle.
- The learned Kessler surrogate on the stretched 96-level grid of the
  1024x1024x96 supercell (dx = 250 m, dt = 1.5 s, nsound = 8, eight seeded
  bubbles) at a reduced horizontal size: spin-up until it rains, harvest
  over the next 400 steps, train on the GPU, then 60 steps with the network in place of kessler
  against 60 plain steps from the same state; domain rain accumulation
  within 2% of plain Kessler, no negative species, fallback accounted.
"""
import numpy as np
import pytest

from cm1_amd import model, workloads
from common import PROG, interior, make_gpu, make_oracle

pytestmark = pytest.mark.gpu

NAMES = [n for n, _ in PROG] + ["u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "rr", "rf"]


def test_config3_ensemble_48_members(require_gpu):
    import time
    t0 = time.perf_counter()

    def log(msg):
        print(f"[{time.perf_counter() - t0:7.1f} s] {msg}", flush=True)

    shapes = workloads.ensemble_shapes(48)
    work = [workloads.ensemble_member(*s) for s in shapes]
    log("48 initial states built")
    members = [make_gpu(*w) for w in work]
    log("48 members set up")
    e = model.Ensemble(members)
    nsteps = 2
    e.step(nsteps)
    e.sync()
    e.close()
    log(f"{nsteps} batched steps done")
    for q, w in enumerate(work):
        solo = make_gpu(*w)
        solo.set_graphs(False)
        solo.step(nsteps)
        for n in NAMES:
            a, b = members[q].get(n), solo.get(n)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (q, shapes[q], n)
        assert members[q].scalars() == solo.scalars(), q
        solo.close()
        if q % 8 == 7:
            log(f"members 0..{q} equal to themselves stepped alone")
    cells = [s[0] * s[1] * s[2] for s in shapes]
    for q in (int(np.argmin(cells)), int(np.argmax(cells))):
        cfgd = work[q][0]
        o = make_oracle(*work[q])
        o.step(nsteps)
        ni, nj, nk = cfgd["ni"], cfgd["nj"], cfgd["nk"]
        for name, stag in PROG:
            a = members[q].interior(name, stag)
            b = interior(o.f(name), nk, nj, ni, stag)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (q, shapes[q], name)
        log(f"member {q} {shapes[q][:3]} equals the oracle")
    for m in members:
        m.close()


def test_config4_surrogate_on_the_stretched_grid(require_gpu):
    import torch
    from cm1_amd import surrogate as S
    torch.cuda.init()
    cfgd, cols, fields = workloads.config1(ni=192, nj=192)
    assert cfgd["nk"] == 96 and cfgd["dt"] == 1.5 and cfgd["nsound"] == 8
    m = make_gpu(cfgd, cols, fields)
    # plain spin-up until rain reaches the ground (at most 4000 steps,
    # 100 minutes) and has kept falling for 200 more steps, then 400 steps
    # with pairs harvested on every fourth, then training (10,000 Adam
    # steps of 65,536 pairs)
    steps, first = 0, None
    while steps < 4000:
        r_before = m.scalars()["train"]
        m.step(100)
        steps += 100
        r_after = m.scalars()["train"]
        print(f"spin-up step {steps}: rain {r_after:.3e}", flush=True)
        if r_after > 0.0 and first is None:
            first = steps
        if first is not None and steps >= first + 200 and r_after > r_before:
            break
    assert first is not None and m.scalars()["train"] > 0.0, steps
    r0h = m.scalars()["train"]
    pairs = S.harvest_spread(m, 400, set(range(0, 400, 4)), capacity=12_000_000, keep_one_in=20)
    state = m.save_state()
    print(f"{pairs.shape[0]} pairs harvested over 400 steps (rain over them "
          f"{m.scalars()['train'] - r0h:.3e})", flush=True)
    norm = S.fit_norm(pairs)
    x, t = S.normalise(pairs, norm)
    net = S.Net(seed=1)
    net.set_norm(norm)
    losses = net.train(x, t, steps=10000, batch=65536, lr=1.0e-3, log_every=2000)
    print(f"spin-up {steps} steps, {pairs.shape[0]} pairs, losses {losses}", flush=True)
    del pairs, x, t

    nsteps = 60
    m.load_state(state)
    r0 = m.scalars()["train"]
    m.step(nsteps)
    rain_plain = m.scalars()["train"] - r0
    assert rain_plain > 0.0, "no rain reached the ground in the 60 compared steps"

    m.load_state(state)
    S.use(m, net)
    c0 = S.counts(m)
    m.step(nsteps)
    c1 = S.counts(m)
    rain_sur = m.scalars()["train"] - r0
    for q in ("qa0", "qa1", "qa2"):
        a = m.interior(q)
        assert np.all(np.isfinite(a)) and a.min() >= 0.0, q
    err = abs(rain_sur - rain_plain) / rain_plain
    fb = c1["fallback_cells"] - c0["fallback_cells"]
    ncell = m.ni * m.nj * m.nk * nsteps
    print(f"rain over {nsteps} steps: plain {rain_plain:.6e} surrogate {rain_sur:.6e} (rel error {err:.2e}); "
          f"fallback cells {fb} of {ncell}; counts {c1}")
    assert err <= 0.02, err
    assert c1["surrogate_steps"] - c0["surrogate_steps"] == nsteps
    assert c1["unusable"] == c0["unusable"]
    assert 0 <= fb <= ncell // 100, fb
    S.use(m, None)
    net.close()
    m.close()
