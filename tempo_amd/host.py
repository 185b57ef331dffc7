"""Python host of the MI355X TEMPO column step, over the C ABI in
include/tempo_hip.h (libtempo_hip.so, loaded with ctypes).

Mirrors the reference driver API (src/module_mp_tempo_driver.F90):

    cfg = tempo_init(aerosolaware_flag=..., hailaware_flag=..., ...)
    diags = tempo_run(cfg, dt, itimestep, t=..., p=..., w=..., dz=..., qv=..., ...)

Fields are torch tensors on the GPU, shape (nz, ncol), C-contiguous -- the
Fortran (ims:ime, kms:kme) layout with one column per contiguous index --
all float32 (production path) or all float64 (FP64 twin).  Passing None for
an optional field means "absent", exactly like an absent optional argument
of tempo_run; state is updated in place.  torch supplies device memory and
the current stream only; every operation runs in the HIP library, and there
is no fallback when the library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtempo_hip.so")
CCN_PATH = os.path.join(HERE, "data", "ccn_activate.bin")

OK, E_ARG, E_UNSUPPORTED, E_TABLE, E_HIP, E_NOMEM = 0, -1, -2, -3, -4, -5
FP32, FP64 = 4, 8
TABLE_QG, TABLE_QS, TABLE_FZ = 1, 2, 3
TUNE_SED_WAVE, TUNE_BLOCK_A, TUNE_BLOCK_C, TUNE_ORDER, TUNE_FUSED = 1, 2, 3, 4, 5
TUNE_HOST_SLABS, TUNE_HOST_THREADS, TUNE_WAVES_A, TUNE_CCN_LDS = 6, 7, 8, 9
TABLE_SIZES = {TABLE_QG: 5 * 16_867_449, TABLE_QS: 12 * 455_877, TABLE_FZ: 4 * 3_388_275 + 2 * 9_157_500}


class TempoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tempo_hip error {code}: {msg}")
        self.code = code


class Cfg(ctypes.Structure):
    """tempo_hip_cfg: ty_tempo_cfgs (src/module_mp_tempo_cfgs.F90:14-38)."""
    _fields_ = [(n, ctypes.c_int) for n in (
        "aerosolaware_flag", "hailaware_flag", "ml_for_bl_nc_flag", "ml_for_nc_flag", "semi_sedi_flag",
        "refl10cm_from_melting_flag", "turn_off_micro_flag", "cloud_condensation_flag", "verbose",
        "check_tables", "nc_option")] + [("nc_value", ctypes.c_double)] + [(n, ctypes.c_int) for n in (
            "refl10cm_flag", "re_cloud_flag", "re_ice_flag", "re_snow_flag", "max_hail_diameter_flag",
            "rain_med_vol_diam_flag", "graupel_med_vol_diam_flag", "cloud_number_mixing_ratio_flag")]


FIELD_NAMES = ("t", "th", "pii", "p", "w", "dz", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr",
               "nc", "nwfa", "nifa", "ng", "qb")
# boundary-layer cloud inputs (read only; with the cloud-number network they
# enter the cloud effective radius)
BL_NAMES = ("qc_bl", "qcfrac_bl")


class Fields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in FIELD_NAMES] + [("land", ctypes.c_void_p)] + \
        [(n, ctypes.c_void_p) for n in BL_NAMES]


DIAG2_NAMES = ("rain_precip", "ice_liquid_equiv_precip", "snow_liquid_equiv_precip",
               "graupel_liquid_equiv_precip", "frozen_fraction", "frz_rain_precip",
               "max_hail_diameter_sfc", "max_hail_diameter_column")
DIAG3_NAMES = ("refl10cm", "re_cloud", "re_ice", "re_snow", "rain_med_vol_diam", "graupel_med_vol_diam",
               "cloud_number_mixing_ratio")


class Diags(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in DIAG2_NAMES + DIAG3_NAMES]


class Dims(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("nz", ctypes.c_int),
                ("ldk", ctypes.c_longlong), ("ldj", ctypes.c_longlong),
                ("dldk", ctypes.c_longlong), ("dldj", ctypes.c_longlong),
                ("d2ldj", ctypes.c_longlong), ("nlev", ctypes.c_void_p)]


_lib = None


def lib():
    """The HIP library; raises when it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with `python -m tempo_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.tempo_hip_cfg_default.argtypes = [ctypes.POINTER(Cfg)]
        L.tempo_hip_init.argtypes = [ctypes.POINTER(Cfg), ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.tempo_hip_reserve.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.tempo_hip_run.argtypes = [vp, ctypes.POINTER(Fields), ctypes.POINTER(Diags), ctypes.POINTER(Dims),
                                    ctypes.c_double, ctypes.c_int, ctypes.c_int, vp]
        L.tempo_hip_run_host.argtypes = [vp, ctypes.POINTER(Fields), ctypes.POINTER(Diags), ctypes.POINTER(Dims),
                                         ctypes.c_double, ctypes.c_int]
        L.tempo_hip_run_host_prec.argtypes = [vp, ctypes.POINTER(Fields), ctypes.POINTER(Diags), ctypes.POINTER(Dims),
                                              ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.tempo_hip_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(vp)]
        L.tempo_hip_host_free.argtypes = [vp]
        L.tempo_hip_branch_signature.argtypes = [vp, ctypes.POINTER(Fields), ctypes.POINTER(Diags),
                                                 ctypes.POINTER(Dims), ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                 vp, vp]
        L.tempo_hip_table_export.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
        L.tempo_hip_table_write.argtypes = [vp, ctypes.c_char_p]
        L.tempo_hip_table_build_ms.argtypes = [vp]
        L.tempo_hip_table_build_ms.restype = ctypes.c_double
        L.tempo_hip_substeps.argtypes = [vp, vp, ctypes.c_size_t]
        L.tempo_hip_profile.argtypes = [vp, ctypes.c_int]
        L.tempo_hip_tune.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.tempo_hip_tlm.argtypes = [vp, ctypes.POINTER(Fields), ctypes.POINTER(Fields), ctypes.POINTER(Diags),
                                    ctypes.POINTER(Diags), ctypes.POINTER(Dims), ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, vp]
        L.tempo_hip_adjoint.argtypes = [vp, ctypes.POINTER(Fields), ctypes.POINTER(Fields), ctypes.POINTER(Diags),
                                        ctypes.POINTER(Fields), ctypes.POINTER(Dims), ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, vp]
        L.tempo_hip_profile_read.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.tempo_hip_aerosol_surface_emissions.argtypes = [vp, ctypes.c_double, vp, vp, ctypes.POINTER(Dims),
                                                          ctypes.c_int, ctypes.c_int, vp]
        L.tempo_hip_aerosol_surface_emissions_host.argtypes = [vp, ctypes.c_double, vp, vp, ctypes.POINTER(Dims),
                                                               ctypes.c_int]
        L.tempo_hip_set_run_cfg.argtypes = [vp, ctypes.POINTER(Cfg)]
        L.tempo_hip_cloud_check_and_update.argtypes = [vp, ctypes.c_int] + [vp] * 11 + [ctypes.c_double, ctypes.c_int, vp]
        L.tempo_hip_ice_check_and_update.argtypes = [vp, ctypes.c_int] + [vp] * 9 + [ctypes.c_double, ctypes.c_int, vp]
        L.tempo_hip_snow_check_and_update.argtypes = [vp, ctypes.c_int] + [vp] * 5 + [ctypes.c_double, ctypes.c_int, vp]
        L.tempo_hip_cloud_check_and_update_host.argtypes = [vp, ctypes.c_int] + [vp] * 11 + [ctypes.c_double]
        L.tempo_hip_ice_check_and_update_host.argtypes = [vp, ctypes.c_int] + [vp] * 9 + [ctypes.c_double]
        L.tempo_hip_snow_check_and_update_host.argtypes = [vp, ctypes.c_int] + [vp] * 5 + [ctypes.c_double]
        L.tempo_hip_last_error.restype = ctypes.c_char_p
        L.tempo_hip_finalize.argtypes = [vp]
        L.tempo_hip_finalize.restype = None
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise TempoError(rc, lib().tempo_hip_last_error().decode())


def default_cfg(**overrides):
    c = Cfg()
    lib().tempo_hip_cfg_default(ctypes.byref(c))
    for k, v in overrides.items():
        if k == "single_moment_nc_opt":
            if v in ("land", "ocean"):
                c.nc_option = 0 if v == "land" else 1
            else:
                c.nc_option, c.nc_value = 2, float(v)
            continue
        if not hasattr(c, k):
            raise TypeError(f"unknown option {k}")
        setattr(c, k, int(v) if k != "nc_value" else float(v))
    return c


class TempoHip:
    """One initialised TEMPO context on the current HIP device (tempo_init)."""

    def __init__(self, table_dir=None, ccn_path=CCN_PATH, **cfg):
        self.cfg = default_cfg(**cfg)
        ctx = ctypes.c_void_p()
        td = table_dir.encode() if table_dir else None
        _check(lib().tempo_hip_init(ctypes.byref(self.cfg), td, ccn_path.encode() if ccn_path else None,
                                    ctypes.byref(ctx)))
        self._ctx = ctx

    def set_run_cfg(self, **changes):
        """Change run-time flags for later steps (a host editing its
        ty_tempo_cfgs between tempo_run calls); aerosolaware_flag and
        hailaware_flag stay as initialised."""
        c = Cfg()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self.cfg), ctypes.sizeof(Cfg))
        for k, v in changes.items():
            if not hasattr(c, k):
                raise TypeError(f"unknown option {k}")
            setattr(c, k, int(v) if k != "nc_value" else float(v))
        _check(lib().tempo_hip_set_run_cfg(self._ctx, ctypes.byref(c)))
        self.cfg = c

    def close(self):
        if self._ctx:
            lib().tempo_hip_finalize(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def table_build_ms(self):
        return lib().tempo_hip_table_build_ms(self._ctx)

    def reserve(self, ncol, nz):
        _check(lib().tempo_hip_reserve(self._ctx, int(ncol), int(nz)))

    def run(self, state, dt, itimestep, nlev=None, land=None, diags=True, stream=None, signature=None):
        """tempo_run over every column of ``state`` (dict of (nz, ncol) tensors).

        ``diags``: True allocates every diagnostic, False none, or a dict of
        preallocated tensors (name -> (ncol,) or (nz, ncol)) to write into.
        ``signature``: an int32 cuda tensor of shape (3, 2, nz, ncol) runs
        the same step through tempo_hip_branch_signature and receives the
        branch signatures of its points ([phase A, final update][comparisons,
        discrete outcomes]).
        Returns the dict of diagnostic tensors (2-D ones of shape (ncol,)).
        """
        import torch
        ref = state["p"]
        nz, ncol = ref.shape
        dtype = ref.dtype
        if dtype not in (torch.float32, torch.float64):
            raise TypeError("fields must be float32 or float64")
        f = Fields()
        keep = []
        for name in FIELD_NAMES + BL_NAMES:
            a = state.get(name)
            if a is None:
                continue
            if a.dtype != dtype or tuple(a.shape) != (nz, ncol) or not a.is_contiguous() or not a.is_cuda:
                raise ValueError(f"field {name}: need a contiguous {dtype} cuda tensor of shape {(nz, ncol)}")
            setattr(f, name, a.data_ptr())
        if land is not None:
            land = land.to(device=ref.device, dtype=torch.int32).contiguous()
            keep.append(land)
            f.land = land.data_ptr()
        d = Diags()
        out = {}
        if isinstance(diags, dict):
            for n, a in diags.items():
                want = (ncol,) if n in DIAG2_NAMES else (nz, ncol)
                if n not in DIAG2_NAMES + DIAG3_NAMES:
                    raise ValueError(f"unknown diagnostic {n}")
                if a.dtype != dtype or tuple(a.shape) != want or not a.is_contiguous() or not a.is_cuda:
                    raise ValueError(f"diagnostic {n}: need a contiguous {dtype} cuda tensor of shape {want}")
                setattr(d, n, a.data_ptr())
            out = diags
        elif diags:
            for n in DIAG2_NAMES:
                out[n] = torch.zeros(ncol, dtype=dtype, device=ref.device)
                setattr(d, n, out[n].data_ptr())
            for n in DIAG3_NAMES:
                out[n] = torch.zeros((nz, ncol), dtype=dtype, device=ref.device)
                setattr(d, n, out[n].data_ptr())
        dims = Dims(ncol, 1, nz, ncol, nz * ncol, ncol, nz * ncol, ncol, None)
        if nlev is not None:
            nlev = nlev.to(device=ref.device, dtype=torch.int32).contiguous()
            keep.append(nlev)
            dims.nlev = nlev.data_ptr()
        s = stream if stream is not None else torch.cuda.current_stream(ref.device).cuda_stream
        prec = FP32 if dtype == torch.float32 else FP64
        if signature is not None:
            if signature.dtype != torch.int32 or tuple(signature.shape) != (3, 2, nz, ncol) or \
                    not signature.is_contiguous() or not signature.is_cuda:
                raise ValueError(f"signature: need a contiguous int32 cuda tensor of shape {(3, 2, nz, ncol)}")
            _check(lib().tempo_hip_branch_signature(self._ctx, ctypes.byref(f), ctypes.byref(d), ctypes.byref(dims),
                                                    float(dt), int(itimestep), prec, signature.data_ptr(),
                                                    ctypes.c_void_p(s)))
            return out
        _check(lib().tempo_hip_run(self._ctx, ctypes.byref(f), ctypes.byref(d), ctypes.byref(dims), float(dt),
                                   int(itimestep), prec, ctypes.c_void_p(s)))
        return out

    # ---- shared argument marshalling for device tensors ----
    def _fields(self, state, nz, ncol, dtype, what="field"):
        f = Fields()
        for name in FIELD_NAMES + BL_NAMES:
            a = state.get(name) if state else None
            if a is None:
                continue
            if a.dtype != dtype or tuple(a.shape) != (nz, ncol) or not a.is_contiguous() or not a.is_cuda:
                raise ValueError(f"{what} {name}: need a contiguous {dtype} cuda tensor of shape {(nz, ncol)}")
            setattr(f, name, a.data_ptr())
        return f

    def _diags(self, diags, nz, ncol, dtype):
        d = Diags()
        for n, a in (diags or {}).items():
            if n not in DIAG2_NAMES + DIAG3_NAMES:
                raise ValueError(f"unknown diagnostic {n}")
            want = (ncol,) if n in DIAG2_NAMES else (nz, ncol)
            if a.dtype != dtype or tuple(a.shape) != want or not a.is_contiguous() or not a.is_cuda:
                raise ValueError(f"diagnostic {n}: need a contiguous {dtype} cuda tensor of shape {want}")
            setattr(d, n, a.data_ptr())
        return d

    def tlm(self, state, tangent, dt, itimestep, nlev=None, land=None, stream=None):
        """Tangent-linear step (tempo_hip_tlm): ``state`` is stepped in place and
        ``tangent`` (dict of tensors like the state; missing = zero) is
        overwritten with the output tangents.  Returns (diags, diag_tangents)."""
        import torch
        ref = state["p"]
        nz, ncol = ref.shape
        dtype = ref.dtype
        tangent = dict(tangent)
        for k, v in state.items():  # every output needs a tangent buffer
            if k not in tangent:
                tangent[k] = torch.zeros_like(v)
        f = self._fields(state, nz, ncol, dtype)
        ft = self._fields(tangent, nz, ncol, dtype, "tangent")
        out = {n: torch.zeros(ncol, dtype=dtype, device=ref.device) for n in DIAG2_NAMES}
        out.update({n: torch.zeros((nz, ncol), dtype=dtype, device=ref.device) for n in DIAG3_NAMES})
        tout = {n: torch.zeros_like(v) for n, v in out.items()}
        keep = []
        if land is not None:
            land = land.to(device=ref.device, dtype=torch.int32).contiguous()
            keep.append(land)
            f.land = land.data_ptr()
        dims = Dims(ncol, 1, nz, ncol, nz * ncol, ncol, nz * ncol, ncol, None)
        if nlev is not None:
            nlev = nlev.to(device=ref.device, dtype=torch.int32).contiguous()
            keep.append(nlev)
            dims.nlev = nlev.data_ptr()
        s = stream if stream is not None else torch.cuda.current_stream(ref.device).cuda_stream
        prec = FP32 if dtype == torch.float32 else FP64
        _check(lib().tempo_hip_tlm(self._ctx, ctypes.byref(f), ctypes.byref(ft),
                                   ctypes.byref(self._diags(out, nz, ncol, dtype)),
                                   ctypes.byref(self._diags(tout, nz, ncol, dtype)), ctypes.byref(dims), float(dt),
                                   int(itimestep), prec, ctypes.c_void_p(s)))
        return out, tout, tangent

    def adjoint(self, state, ybar, dbar, dt, itimestep, nlev=None, land=None, stream=None, wrt=None):
        """Adjoint step (tempo_hip_adjoint).  ``state``: the step's inputs (not
        modified); ``ybar``: cotangents of output fields; ``dbar``:
        cotangents of the precipitation diagnostics and refl10cm.  Returns
        the input cotangents of the fields named in ``wrt`` (default t, qv,
        qc, qr, qi, qs, qg, nc, nr; also ni, nwfa, nifa, ng, qb)."""
        import torch
        ref = state["p"]
        nz, ncol = ref.shape
        dtype = ref.dtype
        f = self._fields(state, nz, ncol, dtype)
        yb = self._fields(ybar, nz, ncol, dtype, "cotangent")
        db = self._diags(dbar, nz, ncol, dtype)
        names = wrt or ("t", "qv", "qc", "qr", "qi", "qs", "qg", "nc", "nr")
        xbar = {k: torch.zeros((nz, ncol), dtype=dtype, device=ref.device) for k in names if k in state}
        xb = self._fields(xbar, nz, ncol, dtype)
        keep = []
        if land is not None:
            land = land.to(device=ref.device, dtype=torch.int32).contiguous()
            keep.append(land)
            f.land = land.data_ptr()
        dims = Dims(ncol, 1, nz, ncol, nz * ncol, ncol, nz * ncol, ncol, None)
        if nlev is not None:
            nlev = nlev.to(device=ref.device, dtype=torch.int32).contiguous()
            keep.append(nlev)
            dims.nlev = nlev.data_ptr()
        s = stream if stream is not None else torch.cuda.current_stream(ref.device).cuda_stream
        prec = FP32 if dtype == torch.float32 else FP64
        _check(lib().tempo_hip_adjoint(self._ctx, ctypes.byref(f), ctypes.byref(yb), ctypes.byref(db),
                                       ctypes.byref(xb), ctypes.byref(dims), float(dt), int(itimestep), prec,
                                       ctypes.c_void_p(s)))
        return xbar

    def run_host(self, state, dt, itimestep, nlev=None, land=None, precision=FP64, diags=None):
        """tempo_run on host float64 numpy arrays (staged through device
        memory), stepped in ``precision`` (FP64 or FP32); ``diags``: the
        diagnostic arrays to fill (name -> float64 array of ncol or
        (nz, ncol)), else new ones."""
        nz, ncol = state["p"].shape
        f = Fields()
        for name in FIELD_NAMES + BL_NAMES:
            a = state.get(name)
            if a is not None:
                assert a.dtype == np.float64 and a.flags.c_contiguous and a.shape == (nz, ncol), name
                setattr(f, name, a.ctypes.data)
        if land is not None:
            land = np.ascontiguousarray(land, dtype=np.int32)
            f.land = land.ctypes.data
        if diags is None:
            out = {n: np.zeros(ncol) for n in DIAG2_NAMES}
            out.update({n: np.zeros((nz, ncol)) for n in DIAG3_NAMES})
        else:
            out = diags
            for n, a in out.items():
                assert a.dtype == np.float64 and a.flags.c_contiguous and a.shape == ((ncol,) if n in DIAG2_NAMES else (nz, ncol)), n
        d = Diags(*[out[n].ctypes.data if n in out else None for n in DIAG2_NAMES + DIAG3_NAMES])
        dims = Dims(ncol, 1, nz, ncol, nz * ncol, ncol, nz * ncol, ncol, None)
        if nlev is not None:
            nlev = np.ascontiguousarray(nlev, dtype=np.int32)
            dims.nlev = nlev.ctypes.data
        _check(lib().tempo_hip_run_host_prec(self._ctx, ctypes.byref(f), ctypes.byref(d), ctypes.byref(dims),
                                             float(dt), int(itimestep), int(precision)))
        return out

    # argument order of the public check routines (src/module_mp_tempo_main.F90:891-1134)
    CHECK_ARGS = {"cloud": ("rho", "l_qc", "qc1d", "nc1d", "ncsave", "rc", "nc", "qcten", "ncten", "ilamc", "mvd_c"),
                  "ice": ("rho", "l_qi", "qi1d", "ni1d", "ri", "ni", "qiten", "niten", "ilami"),
                  "snow": ("rho", "l_qs", "qs1d", "rs", "qsten")}

    def check_and_update(self, which, arrays, dt, stream=None):
        """cloud/ice/snow_check_and_update on n points: ``arrays`` maps the
        argument names of CHECK_ARGS[which] to 1-D device tensors (float32 or
        float64; l_* int32), updated in place; absent nc1d / ncsave = None.
        numpy float64 arrays instead run on the host-array entry points."""
        import torch
        names = self.CHECK_ARGS[which]
        n = len(arrays["rho"])
        on_host = isinstance(arrays["rho"], np.ndarray)
        ptrs = [None if arrays.get(k) is None else (arrays[k].ctypes.data if on_host else arrays[k].data_ptr())
                for k in names]
        fn = getattr(lib(), f"tempo_hip_{which}_check_and_update" + ("_host" if on_host else ""))
        if on_host:
            _check(fn(self._ctx, n, *ptrs, float(dt)))
        else:
            prec = FP32 if arrays["rho"].dtype == torch.float32 else FP64
            s = stream if stream is not None else torch.cuda.current_stream(arrays["rho"].device).cuda_stream
            _check(fn(self._ctx, n, *ptrs, float(dt), prec, ctypes.c_void_p(s)))

    def surface_emissions(self, nwfa, nwfa2d, dt, k=0, stream=None):
        """tempo_aerosol_surface_emissions (src/module_mp_tempo_driver.F90:516-529):
        nwfa[k, c] += nwfa2d[c] * dt in place; nwfa a (nz, ncol) device tensor,
        nwfa2d (ncol,), both float32 or both float64; k is 0-based (kts - kms)."""
        import torch
        nz, ncol = nwfa.shape
        assert nwfa.is_contiguous() and nwfa2d.is_contiguous() and nwfa2d.shape == (ncol,)
        assert nwfa.dtype == nwfa2d.dtype and nwfa.is_cuda and nwfa2d.is_cuda
        precision = FP32 if nwfa.dtype == torch.float32 else FP64
        dims = Dims(ncol, 1, nz, ncol, nz * ncol, ncol, nz * ncol, ncol, None)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(lib().tempo_hip_aerosol_surface_emissions(self._ctx, float(dt), nwfa.data_ptr(), nwfa2d.data_ptr(),
                                                         ctypes.byref(dims), int(k), precision, s))

    def surface_emissions_host(self, nwfa, nwfa2d, dt, k=0):
        """As surface_emissions on host float64 arrays (staged through device memory)."""
        nz, ncol = nwfa.shape
        assert nwfa.dtype == np.float64 and nwfa.flags.c_contiguous and nwfa2d.shape == (ncol,)
        nwfa2d = np.ascontiguousarray(nwfa2d, dtype=np.float64)
        dims = Dims(ncol, 1, nz, ncol, nz * ncol, ncol, nz * ncol, ncol, None)
        _check(lib().tempo_hip_aerosol_surface_emissions_host(self._ctx, float(dt), nwfa.ctypes.data,
                                                              nwfa2d.ctypes.data, ctypes.byref(dims), int(k)))

    def export_table(self, which):
        out = np.empty(TABLE_SIZES[which], dtype=np.float64)
        _check(lib().tempo_hip_table_export(self._ctx, which, out.ctypes.data, out.size))
        return out

    def write_tables(self, directory):
        os.makedirs(directory, exist_ok=True)
        _check(lib().tempo_hip_table_write(self._ctx, directory.encode()))

    def tune(self, precision, key, value):
        """Launch-shape setting for one precision (FP32/FP64); results are unchanged."""
        _check(lib().tempo_hip_tune(self._ctx, int(precision), int(key), int(value)))

    def profile(self, enable=True):
        """Record per-phase HIP events on subsequent runs."""
        _check(lib().tempo_hip_profile(self._ctx, int(enable)))

    def profile_read(self):
        """(ms per phase [prepass, level physics, sedimentation, final], runs)"""
        ms = np.zeros(4)
        runs = ctypes.c_int()
        _check(lib().tempo_hip_profile_read(self._ctx, ms.ctypes.data, 4, ctypes.byref(runs)))
        return ms, runs.value

    def substeps(self, ncol, decisions=False):
        """(10, ncol) substep counts and tops of the last run (rain, graupel,
        snow, ice, cloud); decisions: also the rows of the semi-Lagrangian
        rain and graupel decision hashes, (12, ncol)"""
        rows = 12 if decisions else 10
        out = np.empty(rows * ncol, dtype=np.int32)
        _check(lib().tempo_hip_substeps(self._ctx, out.ctypes.data, out.size))
        return out.reshape(rows, ncol)


class _PageLocked:
    """one tempo_hip_host_alloc block, released with its last array"""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib().tempo_hip_host_alloc(int(nbytes), ctypes.byref(p)))
        self.ptr = p.value

    def __del__(self):
        if self.ptr:
            lib().tempo_hip_host_free(self.ptr)
            self.ptr = None


def pinned_empty(shape, dtype=np.float64):
    """an uninitialised numpy array in page-locked host memory
    (tempo_hip_host_alloc): run_host moves such arrays without staging copies"""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    owner = _PageLocked(max(n, 1))
    buf = (ctypes.c_char * max(n, 1)).from_address(owner.ptr)
    buf._owner = owner  # the array keeps buf, buf keeps the block
    return np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)


# ---- reference-named wrappers ------------------------------------------------

def tempo_init(aerosolaware_flag=None, hailaware_flag=None, semi_sedi_flag=None, cloud_condensation_flag=None,
               refl10cm_from_melting_flag=None, ml_for_bl_nc_flag=None, ml_for_nc_flag=None,
               force_init_flag=None, table_dir=None, **cfg):
    """tempo_init (src/module_mp_tempo_driver.F90:42-182): returns the context."""
    flags = dict(aerosolaware_flag=aerosolaware_flag, hailaware_flag=hailaware_flag,
                 semi_sedi_flag=semi_sedi_flag, cloud_condensation_flag=cloud_condensation_flag,
                 refl10cm_from_melting_flag=refl10cm_from_melting_flag, ml_for_bl_nc_flag=ml_for_bl_nc_flag,
                 ml_for_nc_flag=ml_for_nc_flag)
    cfg.update({k: v for k, v in flags.items() if v is not None})
    return TempoHip(table_dir=table_dir, **cfg)


def tempo_run(ctx, dt, itimestep, t=None, th=None, pii=None, p=None, w=None, dz=None, qv=None, qc=None, qr=None,
              qi=None, qs=None, qg=None, ni=None, nr=None, nc=None, nwfa=None, nifa=None, ng=None, qb=None,
              land_input=None, nlev=None, qc_bl=None, qcfrac_bl=None):
    """tempo_run (src/module_mp_tempo_driver.F90:187-513) on (nz, ncol) device tensors."""
    state = dict(t=t, th=th, pii=pii, p=p, w=w, dz=dz, qv=qv, qc=qc, qr=qr, qi=qi, qs=qs, qg=qg, ni=ni, nr=nr,
                 nc=nc, nwfa=nwfa, nifa=nifa, ng=ng, qb=qb, qc_bl=qc_bl, qcfrac_bl=qcfrac_bl)
    return ctx.run({k: v for k, v in state.items() if v is not None}, dt, itimestep, nlev=nlev, land=land_input)


def tempo_aerosol_surface_emissions(ctx, dt, nwfa, nwfa2d, kts=0):
    """tempo_aerosol_surface_emissions (src/module_mp_tempo_driver.F90:516-529)
    on an (nz, ncol) device tensor nwfa and (ncol,) nwfa2d; kts 0-based."""
    ctx.surface_emissions(nwfa, nwfa2d, dt, kts)
