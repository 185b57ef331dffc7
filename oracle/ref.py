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
REF_SO = os.path.join(HERE, "_ref", "libtempo_ref.so")          # FP64 -O0: parity oracle
REF_SO_O2 = os.path.join(HERE, "_ref", "libtempo_ref_o2.so")    # FP64 -O2: CPU timing baseline
REF_SO_SP = os.path.join(HERE, "_ref", "libtempo_ref_sp.so")    # -DSINGLE_PREC -O2: the reference FP32 path
VARIANTS = {"o0": REF_SO, "o2": REF_SO_O2, "sp": REF_SO_SP}

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

_libs = {}
_variant = "o0"


def select(variant):
    """Default build for the calls below: 'o0' (FP64 parity oracle, the
    default), 'o2' (FP64 timing baseline) or 'sp' (single precision).  Each
    build is a separate library with its own module state."""
    global _variant
    if variant not in VARIANTS:
        raise ValueError(variant)
    _variant = variant


def available(variant="o0"):
    return os.path.exists(VARIANTS[variant])


def lib(variant=None):
    v = variant or _variant
    if v not in _libs:
        so = VARIANTS[v]
        if not os.path.exists(so):
            raise RuntimeError("reference build missing: run `make -C oracle ref`")
        L = ctypes.CDLL(so)
        L.ref_init.restype = ctypes.c_int
        L.ref_init.argtypes = [ctypes.POINTER(ctypes.c_int)] + [ctypes.c_void_p] * 4
        L.ref_run.restype = ctypes.c_int
        L.ref_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                              ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
        for fn in ("ref_gamma_p", "ref_rslf", "ref_rsif"):
            getattr(L, fn).restype = ctypes.c_double
            getattr(L, fn).argtypes = [ctypes.c_double, ctypes.c_double]
        L.ref_snow_moments.restype = None
        L.ref_snow_moments.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.ref_surface_emissions.restype = None
        L.ref_surface_emissions.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 4
        L.ref_get_array.restype = ctypes.c_int
        L.ref_get_array.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        _libs[v] = L
    return _libs[v]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_keep = []


def init(tables, variant=None, **flags):
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
    rc = lib(variant).ref_init(arr, _ptr(ccn), _ptr(fz), _ptr(qs), _ptr(qg))
    if rc != 0:
        raise RuntimeError("ref_init failed")
    return f


def run(state, dt, itimestep, land=None, variant=None):
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
    # boundary-layer cloud inputs (qc_bl, qcfrac_bl), absent unless given
    bl = [state.get(n) for n in ("qc_bl", "qcfrac_bl")]
    for a in bl:
        assert a is None or (a.shape == (nz, ncol) and a.dtype == np.float64 and a.flags.c_contiguous)
    f_bl = lib(variant).ref_set_bl
    f_bl.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    f_bl.restype = None
    f_bl(*[None if a is None else a.ctypes.data for a in bl])
    rc = lib(variant).ref_run(ncol, nz, float(dt), int(itimestep), ptrs, landp, p2, p3)
    f_bl(None, None)
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


def surface_emissions(dt, nwfa, nwfa2d, kts, variant=None):
    """tempo_aerosol_surface_emissions of the reference build on nwfa
    (nj, nk, ni) C-order (== Fortran (ni, nk, nj)), in place; kts 1-based."""
    assert nwfa.dtype == np.float64 and nwfa.flags.c_contiguous and nwfa2d.flags.c_contiguous
    nj, nk, ni = nwfa.shape
    lib(variant).ref_surface_emissions(dt, _ptr(nwfa), _ptr(nwfa2d), ni, nk, nj, kts)


def check_and_update(which, arrays, dt, variant=None, has_nc=True, has_ncsave=False):
    """The reference's public cloud/ice/snow_check_and_update
    (src/module_mp_tempo_main.F90:891-1134) on n points; ``arrays`` maps the
    argument names to float64 arrays (l_* int32 0/1) updated in place."""
    L = lib(variant)
    n = len(arrays["rho"])
    P = ctypes.c_void_p
    ptr = lambda k: arrays[k].ctypes.data
    if which == "cloud":
        f = L.ref_cloud_check
        f.argtypes = [ctypes.c_int] + [P] * 11 + [ctypes.c_double, ctypes.c_int, ctypes.c_int]
        f.restype = None
        f(n, *[ptr(k) for k in ("rho", "l_qc", "qc1d", "nc1d", "ncsave", "rc", "nc", "qcten", "ncten", "ilamc",
                                "mvd_c")], float(dt), int(has_nc), int(has_ncsave))
    elif which == "ice":
        f = L.ref_ice_check
        f.argtypes = [ctypes.c_int] + [P] * 9 + [ctypes.c_double]
        f.restype = None
        f(n, *[ptr(k) for k in ("rho", "l_qi", "qi1d", "ni1d", "ri", "ni", "qiten", "niten", "ilami")], float(dt))
    else:
        f = L.ref_snow_check
        f.argtypes = [ctypes.c_int] + [P] * 5 + [ctypes.c_double]
        f.restype = None
        f(n, *[ptr(k) for k in ("rho", "l_qs", "qs1d", "rs", "qsten")], float(dt))
