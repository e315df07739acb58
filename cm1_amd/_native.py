"""ctypes binding of the cm1gpu C ABI (``include/cm1gpu.h``).

The library is the only compute path: if ``libcm1gpu.so`` is missing or
fails to load, :func:`lib` raises; nothing falls back to a CPU
implementation.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CM1GPU_LIB", os.path.join(HERE, "libcm1gpu.so"))

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
_PTA0 = len(FIELDS)
for _n in range(8):
    FID[f"pta{_n}"] = _PTA0 + _n
    FID[f"pt3d{_n}"] = _PTA0 + 8 + _n
    FID[f"ptten{_n}"] = _PTA0 + 16 + _n
# LES subgrid turbulence (after the tracers; allocated by cm1g_set_turbulence)
TURB_FIELDS = ["tkea", "tke3d", "tketen", "kmh", "kmv", "khh", "khv", "nm", "defv", "defh",
               "dissten", "lenscl"]
for _i, _n in enumerate(TURB_FIELDS):
    FID[_n] = _PTA0 + 24 + _i
NFIELDS = _PTA0 + 24 + len(TURB_FIELDS)
KFIELDS = ["mh", "rmh", "mf", "rmf", "c1", "c2", "zh", "zf", "pi0", "th0", "prs0",
           "qv0", "rho0", "rr0", "rf0", "rrf0", "thv0", "rth0", "qc0", "qi0", "u0", "v0",
           "tauh", "tauf"]
KID = {n: i for i, n in enumerate(KFIELDS)}
D2 = {"rain": 0, "prate": 1, "radbcw": 2, "radbce": 3, "radbcs": 4, "radbcn": 5,
      "ust": 6, "u1": 7, "v1": 8, "s1": 9, "thflux": 10, "znt": 11}
KT_SOUND_SUBSTEP = 1
BC = {"s": 0, "p": 1, "u": 2, "v": 3, "w": 4, "w0": 5}

ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "device allocation failed",
          4: "unsupported", 5: "satadj did not converge", 6: "CFL >= 1.5",
          7: "output file could not be written"}
# GrADS output switches of cm1g_write_grads (the reference's output_* names)
OUT = {n: 1 << i for i, n in enumerate(
    ("rain", "zh", "th", "thpert", "prs", "prspert", "pi", "pipert", "rho", "rhopert", "qv",
     "qvpert", "q", "buoyancy", "uinterp", "vinterp", "winterp", "basestate", "u", "upert", "v",
     "vpert", "w"))}
# the subgrid turbulence's output switches (need the turbulence set); not
# part of the default "all"
OUT_TURB = {n: 1 << (23 + i) for i, n in enumerate(("tke", "km", "kh", "dissten", "nm", "def"))}
# statistics switches of cm1g_statpack (the reference's stat_* names)
STAT = {n: 1 << i for i, n in enumerate(
    ("w", "wlevs", "u", "v", "pipert", "prspert", "thpert", "q", "sfcprs", "wsp", "cfl", "tmass",
     "tmois", "qmass"))}
# the subgrid turbulence's statistics switches (need the turbulence set);
# not part of the default "all"
STAT_TURB = {n: 1 << (14 + i) for i, n in enumerate(("tke", "km", "kh"))}
STAT_MAXN = 128


class Config(C.Structure):
    """cm1g_config: the namelist subset the hot path reads."""
    _fields_ = [(n, C.c_int32) for n in (
        "ni", "nj", "nk", "wbc", "ebc", "sbc", "nbc", "imoist", "ptype", "psolver",
        "advwenos", "advwenov", "apmasscon", "irbc", "nsound", "npt", "pdtra")] + \
        [(n, C.c_float) for n in ("dx", "dy", "dz", "dt", "alph", "kdiv", "csmax")] + \
        [("irdamp", C.c_int32), ("rdalpha", C.c_float), ("zd", C.c_float)]


class TurbConfig(C.Structure):
    """cm1g_turb_config: the LES subgrid-turbulence switches."""
    _fields_ = [(n, C.c_int32) for n in ("cm1setup", "doimpl", "bbc", "isfcflx")] + \
        [(n, C.c_float) for n in ("cnst_shflx", "cnst_znt")]


class CM1GError(RuntimeError):
    pass


_LIB = None

# exported symbols and their signatures (checked by the CPU test suite)
SIGNATURES = {
    "cm1g_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "cm1g_destroy": (C.c_int, [C.c_void_p]),
    "cm1g_last_error": (C.c_char_p, []),
    "cm1g_dims": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "cm1g_put_column": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cm1g_get_column": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cm1g_metrics": (C.c_int, [C.c_void_p]),
    "cm1g_base_derive": (C.c_int, [C.c_void_p]),
    "cm1g_put_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cm1g_get_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cm1g_put_array": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "cm1g_get_array": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "cm1g_put_2d": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cm1g_get_2d": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cm1g_field_ptr": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "cm1g_sync": (C.c_int, [C.c_void_p]),
    "cm1g_getset": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_set_turbulence": (C.c_int, [C.c_void_p, C.POINTER(TurbConfig)]),
    "cm1g_sfc_and_turb": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_turbs": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cm1g_turbt": (C.c_int, [C.c_void_p]),
    "cm1g_turbu": (C.c_int, [C.c_void_p]),
    "cm1g_turbv": (C.c_int, [C.c_void_p]),
    "cm1g_turbw": (C.c_int, [C.c_void_p]),
    "cm1g_advw_scalar": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cm1g_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "cm1g_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_set_tiles": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_ensemble_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "cm1g_ensemble_destroy": (C.c_int, [C.c_void_p]),
    "cm1g_ensemble_step": (C.c_int, [C.c_void_p]),
    "cm1g_ensemble_sync": (C.c_int, [C.c_void_p]),
    "cm1g_ensemble_set_tiles": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_ensemble_tune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "cm1g_mlp_nparams": (C.c_int, [C.POINTER(C.c_int64)]),
    "cm1g_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "cm1g_mlp_destroy": (C.c_int, [C.c_void_p]),
    "cm1g_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cm1g_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cm1g_mlp_set_norm": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cm1g_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "cm1g_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.POINTER(C.c_float), C.c_void_p]),
    "cm1g_mlp_adam": (C.c_int, [C.c_void_p, C.c_float]),
    "cm1g_mlp_sync": (C.c_int, [C.c_void_p]),
    "cm1g_set_surrogate": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cm1g_set_harvest": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]),
    "cm1g_surrogate_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "cm1g_solve1": (C.c_int, [C.c_void_p]),
    "cm1g_solve2": (C.c_int, [C.c_void_p]),
    "cm1g_mp_driver": (C.c_int, [C.c_void_p]),
    "cm1g_solve3": (C.c_int, [C.c_void_p]),
    "cm1g_advs": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                            C.c_float, C.c_double]),
    "cm1g_advu": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_advv": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_advw": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_sound": (C.c_int, [C.c_void_p, C.c_int, C.c_float]),
    "cm1g_kessler_branch": (C.c_int, [C.c_void_p]),
    "cm1g_ktimer_start": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_ktimer_stop": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double), C.c_char_p, C.c_int]),
    "cm1g_pdefq": (C.c_int, [C.c_void_p, C.c_float, C.c_int]),
    "cm1g_bc": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "cm1g_calccflquick": (C.c_int, [C.c_void_p, C.c_float, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_float)]),
    "cm1g_getdiv": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_getdivx": (C.c_int, [C.c_void_p]),
    "cm1g_maxmin": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int32),
                              C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    "cm1g_maxmin_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                      C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                      C.POINTER(C.c_int32)]),
    "cm1g_put_2d_array": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cm1g_get_2d_array": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cm1g_k_fallout": (C.c_int, [C.c_void_p]),
    "cm1g_fallout": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "cm1g_geterain": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                C.POINTER(C.c_double)]),
    "cm1g_kessler": (C.c_int, [C.c_void_p]),
    "cm1g_satadj": (C.c_int, [C.c_void_p]),
    "cm1g_ensemble_set_order": (C.c_int, [C.c_void_p, C.c_int]),
    "cm1g_ensemble_tune_schedule": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                              C.POINTER(C.c_float)]),
    "cm1g_state_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "cm1g_state_save": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "cm1g_state_load": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "cm1g_dtsmall": (C.c_int, [C.POINTER(Config), C.c_float, C.POINTER(C.c_int32)]),
    "cm1g_budgets": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "cm1g_scalars": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "cm1g_write_grads": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_uint32, C.c_int, C.c_int,
                                   C.c_double, C.c_int]),
    "cm1g_statpack": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_uint32, C.c_double,
                                C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
}


def lib():
    """Load ``libcm1gpu.so`` (raises if it is missing)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise CM1GError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                            "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc, what=""):
    if rc != 0:
        msg = lib().cm1g_last_error().decode(errors="replace")
        raise CM1GError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")
