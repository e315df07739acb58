"""Learned surrogate for kessler (autoconversion, accretion, rain
evaporation), trained on the GPU by the library's own MLP kernels
(cm1g_mlp_*); sedimentation and the saturation adjustment stay exact.

Workflow: harvest (state, change) pairs from plain-Kessler runs
(:func:`harvest`), fit the normalisation (:func:`fit_norm`), train
(:meth:`Net.train`), then hand the network to a model
(``Model`` + :func:`use`): its kessler becomes the network's, with the
projection (water conserved by construction, no negative species) and the
per-column fallback to plain kessler done in the kernels.  Device buffers are torch tensors (torch
provides memory and nothing else here).
"""
import ctypes as C

import numpy as np

from . import _native as N

NI, NO = 8, 4
INPUTS = ("T", "p", "rho", "qv", "qc", "qr", "dt", "z/ztop")
OUTPUTS = ("dtheta", "dqv", "dqc", "dqr")


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Net:
    """The 8 -> 128 -> 128 -> 128 -> 4 network (cm1g_mlp)."""

    def __init__(self, seed=0):
        self.L = N.lib()
        h = C.c_void_p()
        N.check(self.L.cm1g_mlp_create(C.c_uint64(seed), C.byref(h)), "cm1g_mlp_create")
        self.h = h
        n = C.c_int64()
        N.check(self.L.cm1g_mlp_nparams(C.byref(n)), "cm1g_mlp_nparams")
        self.nparams = n.value
        self.norm = None

    def close(self):
        if getattr(self, "h", None):
            self.L.cm1g_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        a = np.empty(self.nparams, np.float32)
        N.check(self.L.cm1g_mlp_get_params(self.h, a.ctypes.data), "cm1g_mlp_get_params")
        return a

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        N.check(self.L.cm1g_mlp_set_params(self.h, p.ctypes.data), "cm1g_mlp_set_params")

    def set_norm(self, norm):
        norm = np.ascontiguousarray(norm, np.float32)
        assert norm.size == 4 * NI + 2 * NO
        N.check(self.L.cm1g_mlp_set_norm(self.h, norm.ctypes.data), "cm1g_mlp_set_norm")
        self.norm = norm

    def forward(self, x):
        """Fused inference on normalised rows x (torch, cuda, float32 [n, 8])."""
        import torch
        x = x.contiguous()
        y = torch.empty((x.shape[0], NO), dtype=torch.float32, device=x.device)
        N.check(self.L.cm1g_mlp_forward(self.h, _ptr(x), x.shape[0], _ptr(y)), "cm1g_mlp_forward")
        N.check(self.L.cm1g_mlp_sync(self.h), "cm1g_mlp_sync")
        return y

    def loss_grad(self, x, t, want_grad=False):
        loss = C.c_float()
        g = np.empty(self.nparams, np.float32) if want_grad else None
        N.check(self.L.cm1g_mlp_loss_grad(self.h, _ptr(x), _ptr(t), x.shape[0], C.byref(loss),
                                          g.ctypes.data if want_grad else None), "cm1g_mlp_loss_grad")
        return (loss.value, g) if want_grad else loss.value

    def train(self, x, t, steps, batch=65536, lr=1.0e-3, seed=0, log_every=0, lr_end=0.05):
        """Adam on minibatches drawn (with replacement) from normalised x, t
        (torch cuda tensors), the rate decaying from lr to lr*lr_end along a
        cosine; returns the loss every `log_every` steps."""
        import torch
        gen = torch.Generator(device=x.device)
        gen.manual_seed(seed)
        n = x.shape[0]
        batch = min(batch, n)
        losses = []
        for s in range(steps):
            idx = torch.randint(0, n, (batch,), device=x.device, generator=gen)
            xb, tb = x[idx].contiguous(), t[idx].contiguous()
            want = log_every and (s % log_every == 0 or s == steps - 1)
            loss = C.c_float()
            N.check(self.L.cm1g_mlp_loss_grad(self.h, _ptr(xb), _ptr(tb), batch,
                                              C.byref(loss) if want else None, None), "cm1g_mlp_loss_grad")
            rate = lr * (lr_end + (1.0 - lr_end) * 0.5 * (1.0 + np.cos(np.pi * s / max(1, steps - 1))))
            N.check(self.L.cm1g_mlp_adam(self.h, C.c_float(rate)), "cm1g_mlp_adam")
            if want:
                losses.append((s, loss.value))
        N.check(self.L.cm1g_mlp_sync(self.h), "cm1g_mlp_sync")
        return losses


def harvest(models, steps, capacity, keep_one_in=20):
    """Run plain-Kessler models `steps` large steps each, collecting the
    (inputs, changes over kessler) pairs of their Kessler branches; returns
    a torch cuda tensor [n, 12] (at most `capacity` rows: later steps are
    dropped once it is full)."""
    import torch
    L = N.lib()
    buf = torch.empty((capacity, 12), dtype=torch.float32, device="cuda")
    used = 0
    for m in models:
        room = capacity - used
        if room <= 0:
            break
        part = buf[used:]
        N.check(L.cm1g_set_harvest(m.h, _ptr(part), room, keep_one_in), "cm1g_set_harvest")
        m.step(steps)
        cnt = (C.c_int64 * 6)()
        N.check(L.cm1g_surrogate_counts(m.h, cnt), "cm1g_surrogate_counts")
        N.check(L.cm1g_set_harvest(m.h, None, 0, 1), "cm1g_set_harvest")
        used += min(cnt[0], room)
    return buf[:used]


def harvest_spread(model, nsteps, at, capacity, keep_one_in=20):
    """Run a plain-Kessler model `nsteps` large steps, collecting the
    (inputs, changes over kessler) pairs of the steps whose index (0-based)
    is in `at`, so the training data spread over the run; returns a torch
    cuda tensor [n, 12] (at most `capacity` rows)."""
    import torch
    L = N.lib()
    buf = torch.empty((capacity, 12), dtype=torch.float32, device="cuda")
    used = 0
    for s in range(nsteps):
        if s in at and used < capacity:
            # the count restarts at every cm1g_set_harvest
            N.check(L.cm1g_set_harvest(model.h, _ptr(buf[used:]), capacity - used, keep_one_in),
                    "cm1g_set_harvest")
            model.step(1)
            got = counts(model)["harvested"]
            N.check(L.cm1g_set_harvest(model.h, None, 0, 1), "cm1g_set_harvest")
            used += min(got, capacity - used)
        else:
            model.step(1)
    return buf[:used]


def fit_norm(pairs, margin=0.1):
    """Normalisation of raw pairs [n, 12]: per-column mean and standard
    deviation of inputs and changes (a constant column, such as dt, gets a
    standard deviation of 1), and the range of normalised inputs seen,
    widened by `margin` of its width on each side (the network is trusted
    inside it)."""
    import torch
    a = pairs.double()
    mean = a.mean(0)
    std = a.std(0)
    std = torch.where(std > 0.0, std, torch.ones_like(std))
    x = (a[:, :NI] - mean[:NI]) / std[:NI]
    lo, hi = x.min(0).values, x.max(0).values
    w = torch.clamp(hi - lo, min=1.0e-3)
    lo, hi = lo - margin * w, hi + margin * w
    parts = [mean[:NI], std[:NI], mean[NI:], std[NI:], lo, hi]
    return np.concatenate([p.cpu().numpy() for p in parts]).astype(np.float32)


def normalise(pairs, norm):
    """Normalised inputs and targets (torch) of raw pairs [n, 12]."""
    import torch
    nm = torch.as_tensor(norm, device=pairs.device)
    x = (pairs[:, :NI] - nm[:NI]) / nm[NI:2 * NI]
    t = (pairs[:, NI:] - nm[2 * NI:2 * NI + NO]) / nm[2 * NI + NO:2 * NI + 2 * NO]
    return x.contiguous(), t.contiguous()


def use(model, net):
    """Make `net` the kessler of `model` (None: plain kessler)."""
    N.check(N.lib().cm1g_set_surrogate(model.h, net.h if net is not None else None), "cm1g_set_surrogate")


def counts(model):
    cnt = (C.c_int64 * 6)()
    N.check(N.lib().cm1g_surrogate_counts(model.h, cnt), "cm1g_surrogate_counts")
    return dict(harvested=cnt[0], surrogate_steps=cnt[1], fallback_cells=cnt[2], harvest_calls=cnt[3],
                out_of_range=cnt[4], unusable=cnt[5])
