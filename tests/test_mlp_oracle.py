"""The surrogate MLP's CPU oracle (oracle/mlp_ref.py) against what fixes it
independently: finite differences of the float64 loss, hand-picked BF16
roundings, and SiLU's derivative."""
import numpy as np

import mlp_ref as R


def test_bf16_rounding_known_values():
    one = 1.0
    cases = {one: one, one + 2.0 ** -8: one,                      # tie -> even (down)
             one + 3 * 2.0 ** -8: one + 2.0 ** -6,                # tie -> even (up)
             one + 2.0 ** -8 + 2.0 ** -20: one + 2.0 ** -7,       # above the tie
             -2.5: -2.5, 3.0e-3: np.float32(3.0e-3)}
    for x, want in cases.items():
        got = R.bf16(np.float32(x))
        if x == 3.0e-3:
            # 3e-3 = 1.536 * 2^-9 = (196.608 / 128) * 2^-9: rounds to 197/128
            want = (197.0 / 128.0) * 2.0 ** -9
        assert got == want, (x, got, want)


def test_dsilu_closed_form():
    z = np.linspace(-6, 6, 41)
    h = 1.0e-6
    num = (R.silu(z + h) - R.silu(z - h)) / (2 * h)
    assert np.max(np.abs(num - R.dsilu(z))) < 1.0e-8


def test_gradient_finite_differences():
    rng = np.random.default_rng(4)
    p = rng.standard_normal(R.NPARAM) * 0.1
    x = rng.standard_normal((7, R.NI))
    t = rng.standard_normal((7, R.NO))
    loss, g = R.loss_grad(p, x, t, rnd=False)
    assert np.isclose(loss, np.mean((R.forward(p, x, rnd=False)[0] - t) ** 2))
    # central differences on parameters spread over every block
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in R.SIZES])
    idx = [o + k for o, e in zip(offs[:-1], offs[1:]) for k in rng.choice(e - o, min(5, e - o), replace=False)]
    h = 1.0e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (R.loss_grad(p + e, x, t, rnd=False)[0] - R.loss_grad(p - e, x, t, rnd=False)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1.0e-6 * (1.0 + abs(g[i])), (i, num, g[i])


def test_rounding_is_applied_to_products():
    """With BF16 operand rounding the forward differs from the exact one by
    about BF16 precision, not more."""
    rng = np.random.default_rng(5)
    p = rng.standard_normal(R.NPARAM) * 0.1
    x = rng.standard_normal((50, R.NI))
    y0 = R.forward(p, x, rnd=False)[0]
    y1 = R.forward(p, x, rnd=True)[0]
    rel = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
    assert 1.0e-5 < rel < 3.0e-2, rel
