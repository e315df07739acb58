"""Python host mirror of the CM1 driver around the cm1gpu C ABI.

``Model`` owns one device-resident model (``cm1g_create``); fields move
between host and device only through :meth:`Model.put` / :meth:`Model.get`
(the residency protocol a Fortran host uses around its non-ported code).
"""
import ctypes as C
import os

import numpy as np

from . import _native as N


def make_config(min_dz=0.0, **kw):
    """Config from namelist-style keywords; nsound <= 0 asks cm1g_dtsmall,
    which for psolver 2 needs the grid's smallest level thickness min_dz
    (0: an unstretched grid of spacing dz)."""
    d = dict(ni=64, nj=64, nk=40, wbc=2, ebc=2, sbc=2, nbc=2, imoist=1, ptype=1,
             psolver=3, advwenos=0, advwenov=0, apmasscon=1, irbc=4, nsound=0, npt=0,
             pdtra=1, dx=1000.0, dy=1000.0, dz=500.0, dt=6.0, alph=0.6, kdiv=0.1,
             csmax=350.0, irdamp=0, rdalpha=3.3333333333e-3, zd=15000.0)
    d.update(kw)
    cfg = N.Config(**d)
    if cfg.nsound <= 0:
        ns = C.c_int32()
        N.check(N.lib().cm1g_dtsmall(C.byref(cfg), float(min_dz), C.byref(ns)), "cm1g_dtsmall")
        cfg.nsound = ns.value
    return cfg


class Model:
    """One device-resident CM1 dynamical core."""

    def __init__(self, cfg):
        self.L = N.lib()
        self.cfg = cfg
        h = C.c_void_p()
        N.check(self.L.cm1g_create(C.byref(cfg), C.byref(h)), "cm1g_create")
        self.h = h
        px, py, pz = C.c_int64(), C.c_int64(), C.c_int64()
        N.check(self.L.cm1g_dims(self.h, C.byref(px), C.byref(py), C.byref(pz)), "cm1g_dims")
        self.px, self.py, self.pz = px.value, py.value, pz.value
        self.ni, self.nj, self.nk = cfg.ni, cfg.nj, cfg.nk

    def close(self):
        if getattr(self, "h", None):
            self.L.cm1g_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- columns
    def put_column(self, name, col):
        a = np.zeros(self.nk + 3, np.float32)
        col = np.asarray(col, np.float32)
        a[:min(len(col), self.nk + 3)] = col[:self.nk + 3]
        N.check(self.L.cm1g_put_column(self.h, N.KID[name], a.ctypes.data), "put_column " + name)

    def get_column(self, name):
        a = np.zeros(self.nk + 3, np.float32)
        N.check(self.L.cm1g_get_column(self.h, N.KID[name], a.ctypes.data), "get_column " + name)
        return a

    def set_columns(self, cols):
        for n in ("zf", "zh", "mh", "mf"):
            if n in cols:
                self.put_column(n, cols[n])
        N.check(self.L.cm1g_metrics(self.h), "cm1g_metrics")
        for n in ("pi0", "th0", "qv0", "u0", "v0"):
            if n in cols:
                self.put_column(n, cols[n])
        N.check(self.L.cm1g_base_derive(self.h), "cm1g_base_derive")

    def set_turbulence(self, cm1setup=1, doimpl=1, bbc=3, isfcflx=1, cnst_shflx=0.24, cnst_znt=0.1):
        """LES subgrid turbulence (cm1g_set_turbulence): the TKE scheme with
        the semi-slip surface and a specified surface heat flux; after
        set_columns, before getset."""
        t = N.TurbConfig(cm1setup=cm1setup, doimpl=doimpl, bbc=bbc, isfcflx=isfcflx,
                         cnst_shflx=cnst_shflx, cnst_znt=cnst_znt)
        N.check(self.L.cm1g_set_turbulence(self.h, C.byref(t)), "cm1g_set_turbulence")

    def sfc_and_turb(self, getsfc=True):
        N.check(self.L.cm1g_sfc_and_turb(self.h, 1 if getsfc else 0), "cm1g_sfc_and_turb")

    # ---- 3-D fields in the logical (nk+3, nj+7, ni+7) box
    def put(self, name, arr):
        a = np.ascontiguousarray(arr, np.float32)
        assert a.shape == (self.pz, self.py, self.ni + 7), a.shape
        N.check(self.L.cm1g_put_field(self.h, N.FID[name], a.ctypes.data), "put " + name)

    def get(self, name):
        a = np.empty((self.pz, self.py, self.ni + 7), np.float32)
        N.check(self.L.cm1g_get_field(self.h, N.FID[name], a.ctypes.data), "get " + name)
        return a

    def get_2d(self, name):
        n = self.px * self.py + (self.px + self.py) * self.pz
        a = np.empty(n, np.float32)
        N.check(self.L.cm1g_get_2d(self.h, N.D2[name], a.ctypes.data), "get_2d " + name)
        return a

    def interior(self, name, stag=None):
        a = self.get(name)
        ni, nj, nk = self.ni, self.nj, self.nk
        ei, ej, ek = (1 if stag == "u" else 0), (1 if stag == "v" else 0), (1 if stag == "w" else 0)
        return a[1:nk + 1 + ek, 3:nj + 3 + ej, 3:ni + 3 + ei]

    # ---- stepping
    def getset(self, fresh=True):
        N.check(self.L.cm1g_getset(self.h, 1 if fresh else 0), "cm1g_getset")

    def step(self, n=1):
        cfl = C.c_float()
        for _ in range(n):
            N.check(self.L.cm1g_step(self.h, C.byref(cfl)), "cm1g_step")
        return cfl.value

    def set_tiles(self, ty):
        """Rows of 64 columns per workgroup of the column kernels (1, 2, 4)."""
        N.check(self.L.cm1g_set_tiles(self.h, int(ty)), "cm1g_set_tiles")

    def set_graphs(self, on):
        """Replay each large step as a captured HIP graph (default) or launch
        kernel by kernel (cm1g_set_graphs)."""
        N.check(self.L.cm1g_set_graphs(self.h, 1 if on else 0), "cm1g_set_graphs")

    def sync(self):
        N.check(self.L.cm1g_sync(self.h), "cm1g_sync")

    def save_state(self):
        """Device-state checkpoint (cm1g_state_save) as a uint8 array."""
        n = C.c_int64()
        N.check(self.L.cm1g_state_bytes(self.h, C.byref(n)), "cm1g_state_bytes")
        buf = np.empty(n.value, np.uint8)
        N.check(self.L.cm1g_state_save(self.h, buf.ctypes.data, n.value), "cm1g_state_save")
        return buf

    def load_state(self, buf):
        buf = np.ascontiguousarray(buf, np.uint8)
        N.check(self.L.cm1g_state_load(self.h, buf.ctypes.data, buf.size), "cm1g_state_load")

    def maxmin(self, name, stag=None):
        """Device maxmin of a field's interior: (vmax, (i,j,k)), (vmin, (i,j,k))."""
        st = {None: 0, "u": 1, "v": 2, "w": 3}[stag]
        vmax, vmin = C.c_float(), C.c_float()
        lmax, lmin = (C.c_int32 * 3)(), (C.c_int32 * 3)()
        N.check(self.L.cm1g_maxmin(self.h, N.FID[name], st, C.byref(vmax), lmax, C.byref(vmin), lmin),
                "cm1g_maxmin")
        return (vmax.value, tuple(lmax)), (vmin.value, tuple(lmin))

    def budgets(self):
        a, b = C.c_double(), C.c_double()
        N.check(self.L.cm1g_budgets(self.h, C.byref(a), C.byref(b)), "cm1g_budgets")
        return a.value, b.value

    def write_grads(self, directory, nwrite, switches=None, iorigin=2, stretch_z=0, mtime=0.0,
                    nstep=0):
        """GrADS output time ``nwrite`` of the variables named in ``switches``
        (the reference's output_* names, default all but the subgrid
        turbulence's) into ``directory``."""
        names = N.OUT if switches is None else switches
        mask = 0
        for n in names:
            mask |= N.OUT[n] if n in N.OUT else N.OUT_TURB[n]
        N.check(self.L.cm1g_write_grads(self.h, os.fsencode(directory), int(nwrite), mask,
                                        int(iorigin), int(stretch_z), float(mtime), int(nstep)),
                "cm1g_write_grads")

    def statpack(self, directory, nrec, switches=None, mtime=0.0):
        """Statistics record ``nrec`` of the switches named in ``switches``
        (the reference's stat_* names, default all) appended to
        ``directory``/cm1out_stats.dat; returns the record's values."""
        names = N.STAT if switches is None else switches
        mask = 0
        for n in names:
            mask |= N.STAT[n] if n in N.STAT else N.STAT_TURB[n]
        r = (C.c_float * N.STAT_MAXN)()
        n = C.c_int32()
        N.check(self.L.cm1g_statpack(self.h, os.fsencode(directory), int(nrec), mask, float(mtime), r,
                                     C.byref(n)), "cm1g_statpack")
        return np.array(r[:n.value], np.float32)

    def scalars(self):
        a = (C.c_double * 4)()
        N.check(self.L.cm1g_scalars(self.h, a), "cm1g_scalars")
        return dict(mass1=a[0], train=a[1], cflmax=a[2], mass2=a[3])


def setup(cfg, cols, fields, fresh=True, turb=None):
    """Create a model, load columns and t=0 fields, run getset; turb: the
    keywords of :meth:`Model.set_turbulence` (None: no subgrid turbulence)."""
    m = Model(cfg)
    m.set_columns(cols)
    if turb:
        m.set_turbulence(**turb)
    for name, arr in fields.items():
        m.put(name, arr)
    m.getset(fresh)
    return m


class Ensemble:
    """Members (set-up :class:`Model` objects of any size) advanced together:
    one launch per kernel of the large step for all members, the merged step
    replayed as one HIP graph (cm1g_ensemble_*).  :meth:`sync` before using a
    member directly."""

    def __init__(self, members):
        self.L = N.lib()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = C.c_void_p()
        N.check(self.L.cm1g_ensemble_create(arr, len(self.members), C.byref(h)), "cm1g_ensemble_create")
        self.h = h

    def step(self, n=1):
        for _ in range(n):
            N.check(self.L.cm1g_ensemble_step(self.h), "cm1g_ensemble_step")

    def sync(self):
        N.check(self.L.cm1g_ensemble_sync(self.h), "cm1g_ensemble_sync")

    def set_tiles(self, ty):
        N.check(self.L.cm1g_ensemble_set_tiles(self.h, int(ty)), "cm1g_ensemble_set_tiles")

    def tune(self, steps):
        """Autotune the column-tile height; returns (best, {ty: ms/step})."""
        best = C.c_int()
        ms = (C.c_float * 3)()
        N.check(self.L.cm1g_ensemble_tune(self.h, int(steps), C.byref(best), ms), "cm1g_ensemble_tune")
        return best.value, {1: ms[0], 2: ms[1], 4: ms[2]}

    def set_order(self, mode):
        N.check(self.L.cm1g_ensemble_set_order(self.h, int(mode)), "cm1g_ensemble_set_order")

    def tune_schedule(self, steps):
        """Autotune tile height and member order; returns
        (best height, best order, {(order, height): ms/step})."""
        ty, od = C.c_int(), C.c_int()
        ms = (C.c_float * 9)()
        N.check(self.L.cm1g_ensemble_tune_schedule(self.h, int(steps), C.byref(ty), C.byref(od), ms),
                "cm1g_ensemble_tune_schedule")
        return ty.value, od.value, {(o, h): ms[3 * o + q] for o in range(3) for q, h in enumerate((1, 2, 4))}

    def close(self):
        if self.h:
            self.L.cm1g_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
