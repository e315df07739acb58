"""CPU oracle of the surrogate MLP (test infrastructure only).

The network of cm1_amd/csrc/cm1g_mlp.hip restated in float64 numpy:
8 -> 128 -> 128 -> 128 -> 4, SiLU hidden layers, mean squared error on the
normalised outputs.  Every matrix product rounds both operands to BF16
(round to nearest, ties to even) exactly where the device kernels do and
accumulates in float64, so the device result (FP32 accumulation) differs
only by accumulation rounding.  Parameters use the flat layout of
include/cm1gpu.h (W1 b1 W2 b2 W3 b3 W4 b4, weights row-major [in][out]).

Pinned by test_mlp_oracle.py: finite differences of the float64 loss
(without BF16 rounding) against the analytic gradient, the BF16 rounding
against hand-picked values, and SiLU's derivative against its closed form.
"""
import numpy as np

NI, NH, NO = 8, 128, 4
SIZES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
         ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NH, NO)), ("b4", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16(a):
    """Round float32 values to BF16 (nearest, ties to even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _mm(a, b, rnd):
    # operands enter the matrix cores as BF16 rounded from the FP32 values
    if rnd:
        return bf16(np.asarray(a, np.float32)) @ bf16(np.asarray(b, np.float32))
    return a @ b


def forward(p, x, rnd=True):
    """Outputs and saved activations of the network on rows x [n, 8]."""
    w = unpack(p)
    z1 = _mm(x, w["W1"], rnd) + w["b1"]
    h1 = silu(z1)
    z2 = _mm(h1, w["W2"], rnd) + w["b2"]
    h2 = silu(z2)
    z3 = _mm(h2, w["W3"], rnd) + w["b3"]
    h3 = silu(z3)
    y = _mm(h3, w["W4"], rnd) + w["b4"]
    return y, (z1, h1, z2, h2, z3, h3)


def loss_grad(p, x, t, rnd=True):
    """Mean squared error over all outputs and its gradient (flat)."""
    w = unpack(p)
    y, (z1, h1, z2, h2, z3, h3) = forward(p, x, rnd)
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / d.size
    g = {}
    g["W4"] = _mm(h3.T, dy, rnd)
    g["b4"] = dy.sum(0)
    dz3 = _mm(dy, w["W4"].T, rnd) * dsilu(z3)
    g["W3"] = _mm(h2.T, dz3, rnd)
    g["b3"] = dz3.sum(0)
    dz2 = _mm(dz3, w["W3"].T, rnd) * dsilu(z2)
    g["W2"] = _mm(h1.T, dz2, rnd)
    g["b2"] = dz2.sum(0)
    dz1 = _mm(dz2, w["W2"].T, rnd) * dsilu(z1)
    g["W1"] = _mm(np.asarray(x).T, dz1, rnd)
    g["b1"] = dz1.sum(0)
    return loss, pack(g)
