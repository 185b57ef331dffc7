"""ctypes driver for the reference TEMPO FP64 build (oracle/_ref/libtempo_ref.so).

Test infrastructure only: the library is compiled from the reference sources
by ``make -C oracle ref`` and is used to pin the C restatement and to produce
golden vectors.  Nothing in the product imports this module.

Fields are numpy float64 arrays of shape (nz, ncol) in C order, i.e. the
Fortran (ncol, nz) layout of tempo_run (src/module_mp_tempo_driver.F90:207-241)
with one column per contiguous index and levels strided by ncol.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libtempo_ref.so")          # -O0: parity oracle
REF_SO_O2 = os.path.join(HERE, "_ref", "libtempo_ref_o2.so")    # -O2: CPU timing baseline

# order of the state pointer array in ref_run (oracle/ref_bridge.F90)
STATE_FIELDS = ("t", "p", "w", "dz", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr",
                "nc", "nwfa", "nifa", "ng", "qb", "th", "pii")
DIAG2_FIELDS = ("rain_precip", "ice_precip", "snow_precip", "graupel_precip",
                "frozen_fraction", "frz_rain_precip", "max_hail_sfc", "max_hail_column")
DIAG3_FIELDS = ("refl10cm", "re_cloud", "re_ice", "re_snow", "mvd_r", "mvd_g", "nc_mixing_ratio")

DEFAULT_FLAGS = dict(aerosolaware=True, hailaware=True, semi_sedi=False, refl_from_melting=False,
                     turn_off_micro=False, cloud_condensation=True, refl10cm=True, re_cloud=True,
                     re_ice=True, re_snow=True, max_hail=True, mvd_r=False, mvd_g=False,
                     nc_mixing_ratio=False, ml_for_bl_nc=False, ml_for_nc=False, nc_opt=0)
FLAG_ORDER = ("aerosolaware", "hailaware", "semi_sedi", "refl_from_melting", "turn_off_micro",
              "cloud_condensation", "refl10cm", "re_cloud", "re_ice", "re_snow", "max_hail",
              "mvd_r", "mvd_g", "nc_mixing_ratio", "ml_for_bl_nc", "ml_for_nc", "nc_opt")

_lib = None
_so = REF_SO


def select(optimised):
    """Choose the -O2 build (timing) or the -O0 build (parity, the default)
    before the first call."""
    global _so
    if _lib is not None:
        raise RuntimeError("reference library already loaded")
    _so = REF_SO_O2 if optimised else REF_SO


def available(optimised=False):
    return os.path.exists(REF_SO_O2 if optimised else REF_SO)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_so):
            raise RuntimeError("reference build missing: run `make -C oracle ref`")
        _lib = ctypes.CDLL(_so)
        _lib.ref_init.restype = ctypes.c_int
        _lib.ref_init.argtypes = [ctypes.POINTER(ctypes.c_int)] + [ctypes.c_void_p] * 4
        _lib.ref_run.restype = ctypes.c_int
        _lib.ref_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
        for fn in ("ref_gamma_p", "ref_rslf", "ref_rsif"):
            getattr(_lib, fn).restype = ctypes.c_double
            getattr(_lib, fn).argtypes = [ctypes.c_double, ctypes.c_double]
        _lib.ref_snow_moments.restype = None
        _lib.ref_snow_moments.argtypes = [ctypes.c_double, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_double)]
        _lib.ref_get_array.restype = ctypes.c_int
        _lib.ref_get_array.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_keep = []


def init(tables, **flags):
    """Initialise the reference (tempo_init sequence) with in-memory tables.

    ``tables`` holds the flat arrays in reference file order: 'ccn' float32
    (8820), 'fz', 'qs', 'qg' float64.
    """
    f = dict(DEFAULT_FLAGS)
    f.update(flags)
    arr = (ctypes.c_int * 18)(*([int(f[k]) for k in FLAG_ORDER] + [0]))
    ccn = np.ascontiguousarray(tables["ccn"], dtype=np.float32)
    fz = np.ascontiguousarray(tables["fz"], dtype=np.float64)
    qs = np.ascontiguousarray(tables["qs"], dtype=np.float64)
    qg = np.ascontiguousarray(tables["qg"], dtype=np.float64)
    rc = lib().ref_init(arr, _ptr(ccn), _ptr(fz), _ptr(qs), _ptr(qg))
    if rc != 0:
        raise RuntimeError("ref_init failed")
    return f


def run(state, dt, itimestep, land=None):
    """One tempo_run step over all columns; state arrays are updated in place.

    Returns (diag2, diag3) dicts of the driver diagnostics.
    """
    nz, ncol = state["p"].shape
    ptrs = (ctypes.c_void_p * 19)()
    arrays = {}
    for i, name in enumerate(STATE_FIELDS):
        a = state.get(name)
        if a is None:
            ptrs[i] = None
            continue
        assert a.shape == (nz, ncol) and a.dtype == np.float64 and a.flags.c_contiguous, name
        arrays[name] = a
        ptrs[i] = a.ctypes.data
    d2 = {k: np.zeros(ncol) for k in DIAG2_FIELDS}
    d3 = {k: np.zeros((nz, ncol)) for k in DIAG3_FIELDS}
    p2 = (ctypes.c_void_p * 8)(*[d2[k].ctypes.data for k in DIAG2_FIELDS])
    p3 = (ctypes.c_void_p * 7)(*[d3[k].ctypes.data for k in DIAG3_FIELDS])
    landp = None
    if land is not None:
        land = np.ascontiguousarray(land, dtype=np.int32)
        landp = land.ctypes.data
    rc = lib().ref_run(ncol, nz, float(dt), int(itimestep), ptrs, landp, p2, p3)
    if rc != 0:
        raise RuntimeError("ref_run failed")
    return d2, d3


def gamma_p(a, x):
    return lib().ref_gamma_p(a, x)


def rslf(p, t):
    return lib().ref_rslf(p, t)


def rsif(p, t):
    return lib().ref_rsif(p, t)


def snow_moments(rs, tc):
    out = (ctypes.c_double * 10)()
    lib().ref_snow_moments(rs, tc, out)
    return np.array(out[:])


_ARRAY_SIZES = {1: 10000, 2: 10000, 3: 370000, 4: 3520, 5: 3520, 6: 3520, 7: 1100, 8: 220}


def get_array(which):
    n = _ARRAY_SIZES[which]
    out = np.zeros(n)
    rc = lib().ref_get_array(which, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    if rc != 0:
        raise RuntimeError("ref_get_array failed")
    return out
