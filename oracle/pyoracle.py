"""ctypes view of the CPU oracle ``libcm1oracle.so`` (test infrastructure only).

Only tests, ``__graft_entry__.smoke()`` and the bench's ``cpu_baseline`` leg
use this module; the product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcm1oracle.so")
# the same restatement carried out in double precision (REAL -> REAL(8), the
# analogue of the reference's -DDP build)
LIB64_PATH = os.path.join(HERE, "libcm1oracle64.so")

FIELDS = ["ua", "u3d", "uten", "uten1", "va", "v3d", "vten", "vten1",
          "wa", "w3d", "wten", "wten1", "rru", "rrv", "rrw",
          "ppi", "pp3d", "ppten", "sten", "sadv", "ppx",
          "tha", "th3d", "thten", "thten1", "thterm",
          "qpten", "qtten", "qvten", "qcten",
          "rho", "rr", "rf", "prs", "divx",
          "dum1", "dum2", "dum3", "dum4", "dum5", "dum6", "dum7", "dum8", "dum9",
          "t11", "t12", "t13", "t22", "t23", "t33",
          "qa0", "qa1", "qa2", "q3d0", "q3d1", "q3d2", "qten0", "qten1", "qten2"]
FID = {n: i for i, n in enumerate(FIELDS)}
F_PTA0 = len(FIELDS)
for _n in range(8):
    FID[f"pta{_n}"] = F_PTA0 + _n
    FID[f"pt3d{_n}"] = F_PTA0 + 8 + _n
    FID[f"ptten{_n}"] = F_PTA0 + 16 + _n
# LES subgrid turbulence (after the tracers)
for _i, _n in enumerate(["tkea", "tke3d", "tketen", "kmh", "kmv", "khh", "khv", "nm", "defv",
                          "defh", "dissten", "lenscl"]):
    FID[_n] = F_PTA0 + 24 + _i
KFIELDS = ["mh", "rmh", "mf", "rmf", "c1", "c2", "zh", "zf", "pi0", "th0", "prs0",
           "qv0", "rho0", "rr0", "rf0", "rrf0", "thv0", "rth0", "qc0", "qi0", "u0", "v0",
           "tauh", "tauf"]
KID = {n: i for i, n in enumerate(KFIELDS)}
D2 = {"rain": 0, "prate": 1, "radbcw": 2, "radbce": 3, "radbcs": 4, "radbcn": 5,
      "ust": 6, "u1": 7, "v1": 8, "s1": 9, "thflux": 10, "znt": 11}


_INTS = ("ni", "nj", "nk", "wbc", "ebc", "sbc", "nbc", "imoist", "ptype", "psolver",
         "advwenos", "advwenov", "apmasscon", "irbc", "nsound", "npt", "pdtra")
_REALS = ("dx", "dy", "dz", "dt", "alph", "kdiv", "csmax")


_TAIL = (("irdamp", "i"), ("rdalpha", "r"), ("zd", "r"),
         ("cm1setup", "i"), ("doimpl", "i"), ("bbc", "i"), ("isfcflx", "i"),
         ("cnst_shflx", "r"), ("cnst_znt", "r"))


class Config(C.Structure):
    _fields_ = [(n, C.c_int32) for n in _INTS] + [(n, C.c_float) for n in _REALS] + \
        [(n, C.c_int32 if t == "i" else C.c_float) for n, t in _TAIL]


class Config64(C.Structure):
    _fields_ = [(n, C.c_int32) for n in _INTS] + [(n, C.c_double) for n in _REALS] + \
        [(n, C.c_int32 if t == "i" else C.c_double) for n, t in _TAIL]


def _load(fp64=False):
    path = LIB64_PATH if fp64 else LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (run make -C oracle)")
    lib = C.CDLL(path)
    real = C.c_double if fp64 else C.c_float
    vp, fp = C.c_void_p, C.POINTER(real)
    lib.cm1o_create.restype = vp
    lib.cm1o_create.argtypes = [C.POINTER(Config64 if fp64 else Config)]
    lib.cm1o_destroy.argtypes = [vp]
    for name in ("cm1o_field", "cm1o_kfield", "cm1o_2dfield", "cm1o_hfield"):
        getattr(lib, name).restype = fp
        getattr(lib, name).argtypes = [vp, C.c_int]
    lib.cm1o_set_extrap.argtypes = [vp, fp]
    lib.cm1o_dims.argtypes = [vp, C.POINTER(C.c_int64)] * 1 + [C.POINTER(C.c_int64)] * 2
    lib.cm1o_scalars.restype = C.POINTER(C.c_double)
    lib.cm1o_scalars.argtypes = [vp]
    lib.cm1o_base_derive.argtypes = [vp]
    lib.cm1o_getset.argtypes = [vp, C.c_int]
    lib.cm1o_step.argtypes = [vp]
    for name in ("cm1o_solve1", "cm1o_solve2", "cm1o_mp_driver", "cm1o_solve3", "cm1o_sfc_and_turb"):
        getattr(lib, name).argtypes = [vp]
    lib.cm1o_bc.argtypes = [vp, C.c_int, C.c_int]
    lib.cm1o_advs.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
    for name in ("cm1o_advu", "cm1o_advv", "cm1o_advw"):
        getattr(lib, name).argtypes = [vp, C.c_int]
    lib.cm1o_sound.argtypes = [vp, C.c_int, real]
    lib.cm1o_kessler_branch.argtypes = [vp]
    lib.cm1o_getdiv.argtypes = [vp, C.c_int]
    lib.cm1o_getdivx.argtypes = [vp]
    lib.cm1o_calccflquick.argtypes = [vp, real, C.c_int, C.c_int, C.c_int]
    lib.cm1o_pdefq.argtypes = [vp, real, C.c_int]
    lib.cm1o_set_libm.argtypes = [C.c_int]
    return lib


_LIB = {}


def lib(fp64=False):
    if fp64 not in _LIB:
        _LIB[fp64] = _load(fp64)
    return _LIB[fp64]


def make_config(**kw):
    d = dict(ni=64, nj=64, nk=40, wbc=2, ebc=2, sbc=2, nbc=2, imoist=1, ptype=1,
             psolver=3, advwenos=0, advwenov=0, apmasscon=1, irbc=4, nsound=4, npt=0,
             pdtra=1, dx=1000.0, dy=1000.0, dz=500.0, dt=6.0, alph=0.6, kdiv=0.1,
             csmax=350.0, irdamp=0, rdalpha=3.3333333333e-3, zd=15000.0)
    d.update(kw)
    return Config(**d)


class Oracle:
    """One oracle model; arrays are numpy views shaped (k, j, i) with
    index offsets i0=2, j0=2 (i=-2 at index 0), k0=0."""

    def __init__(self, cfg, libm=0, fp64=False):
        """libm: 0 evaluates exp/pow with the host C library (the
        reference's rounding), 1 in double precision rounded once (the HIP
        kernels' rounding).  fp64: run the double-precision build; cfg is
        then converted to a Config64 and the array views are float64."""
        self.L = lib(fp64)
        self.fp64 = fp64
        if fp64 and not isinstance(cfg, Config64):
            cfg = Config64(**{n: getattr(cfg, n) for n, _ in Config64._fields_})
        self.cfg = cfg
        self.libm = libm
        self.h = self.L.cm1o_create(C.byref(cfg))
        px, py, pz = C.c_int64(), C.c_int64(), C.c_int64()
        self.L.cm1o_dims(self.h, C.byref(px), C.byref(py), C.byref(pz))
        self.px, self.py, self.pz = px.value, py.value, pz.value

    def __del__(self):
        try:
            self.L.cm1o_destroy(self.h)
        except Exception:
            pass

    def f(self, name):
        p = self.L.cm1o_field(self.h, FID[name])
        return np.ctypeslib.as_array(p, shape=(self.pz, self.py, self.px))

    def kf(self, name):
        p = self.L.cm1o_kfield(self.h, KID[name])
        return np.ctypeslib.as_array(p, shape=(self.pz + 1,))

    def d2(self, name):
        p = self.L.cm1o_2dfield(self.h, D2[name])
        return np.ctypeslib.as_array(p, shape=(self.py, self.px))

    def scalars(self):
        p = self.L.cm1o_scalars(self.h)
        return dict(mass1=p[0], train=p[1], cflmax=p[2], mass2=p[3])

    def base_derive(self):
        self.L.cm1o_base_derive(self.h)

    def getset(self, fresh=True):
        self.L.cm1o_set_libm(self.libm)
        self.L.cm1o_getset(self.h, 1 if fresh else 0)

    def step(self, n=1):
        self.L.cm1o_set_libm(self.libm)
        for _ in range(n):
            self.L.cm1o_step(self.h)
