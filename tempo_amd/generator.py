"""Seeded, counter-based synthetic column generator for the TEMPO column step.

Every value is a pure function of (seed, column index, level, field id), so
any subset of columns can be regenerated bit-for-bit by any consumer (tests,
benchmark, golden-vector scripts) without generating the others.

Arrays are float64, shape (nz, ncol), C order: element (k, i) at k*ncol + i,
i.e. the Fortran (i, k) layout of the reference driver
(src/module_mp_tempo_driver.F90:207-241) with k = 0 at the surface.

Workloads (k = 1 is the surface; dz uniform unless stated):

  plain      1,024 x 60,  dz 250 m, dt 30 s, aerosol- and hail-aware off
  aerosol    100,000 x 90, dz 200 m, dt 20 s, aerosol-aware (nc, nwfa, nifa)
  ragged     131,072 columns of U{30..128} active levels, dt 20 s,
             hail-aware (ng, qb), 25 % convective columns; layers stretched
             from 25 m at the surface by 6 % per level up to 200 m (a
             convection-permitting grid: rain and hail falling through the
             thin low layers need 1-20 sedimentation substeps per step)
  adjoint    16,384 x 64, dz 250 m, dt 20 s, aerosol-aware
  multistep  262,144 x 90, dz 200 m, dt 20 s, aerosol- and hail-aware, 25 %
             convective columns as in the ragged workload, every level
             active
  cold       4,096 x 60, dz 250 m, dt 20 s, aerosol- and hail-aware; surface
             temperature U[258, 272.5] K per column and the first cloud layer
             resting on the surface, so that rain, snow, ice and graupel reach
             the ground below freezing (freezing rain)
"""
import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)

FIELD_IDS = {
    "layer1_base": 1, "layer1_thick": 2, "layer2_on": 3, "layer2_base": 4, "layer2_thick": 5,
    "rh": 6, "qc": 7, "qr": 8, "qi": 9, "qs": 10, "qg": 11, "w1": 12, "w2": 13,
    "nwfa1": 14, "nwfa2": 15, "nifa1": 16, "nifa2": 17, "rhog": 18, "conv": 19, "wmax": 20,
    "qr_sc": 21, "nlev": 22, "q1": 23, "q2": 24, "tsfc": 25,
}

WORKLOADS = {
    "plain": dict(ncol=1024, nz=60, dz=250.0, dt=30.0, aerosol=False, hail=False, seed=1001,
                  ragged=False, itimestep=2),
    "aerosol": dict(ncol=100_000, nz=90, dz=200.0, dt=20.0, aerosol=True, hail=False, seed=1002,
                    ragged=False, itimestep=2),
    "ragged": dict(ncol=131_072, nz=128, dz=200.0, dt=20.0, aerosol=False, hail=True, seed=1003,
                   ragged=True, itimestep=2, convective=0.25, dz_sfc=25.0, dz_growth=1.06),
    "adjoint": dict(ncol=16_384, nz=64, dz=250.0, dt=20.0, aerosol=True, hail=False, seed=1004,
                    ragged=False, itimestep=2),
    "multistep": dict(ncol=262_144, nz=90, dz=200.0, dt=20.0, aerosol=True, hail=True, seed=1005,
                      ragged=False, itimestep=2, convective=0.25),
    "cold": dict(ncol=4096, nz=60, dz=250.0, dt=20.0, aerosol=True, hail=True, seed=1006,
                 ragged=False, itimestep=2, t_sfc=(258.0, 272.5), surface_cloud=True),
}


def levels(cfg):
    """(layer depths dz, mid-level heights z) of a workload [m], float64 (nz,):
    uniform dz, or stretched from dz_sfc by dz_growth per level up to dz"""
    nz, dz = cfg["nz"], cfg["dz"]
    if "dz_sfc" in cfg:
        d = np.minimum(dz, cfg["dz_sfc"] * cfg["dz_growth"] ** np.arange(nz, dtype=np.float64))
    else:
        d = np.full(nz, dz)
    z = np.cumsum(d) - 0.5 * d
    return d, z


def _splitmix(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & MASK64
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & MASK64
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & MASK64
    return x ^ (x >> np.uint64(31))


def uniform(seed, col, level, field):
    """U[0,1) from the counter (seed, column, level, field): 53-bit mantissa."""
    col = np.asarray(col, dtype=np.uint64)
    level = np.asarray(level, dtype=np.uint64)
    key = (np.uint64(seed) << np.uint64(48)) ^ (np.uint64(FIELD_IDS[field]) << np.uint64(40)) \
        ^ (level << np.uint64(24)) ^ col
    with np.errstate(over="ignore"):
        h = _splitmix(_splitmix(key) ^ np.uint64(seed))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, col, level, field_a, field_b):
    """Standard normal via Box-Muller over two counter streams."""
    u1 = uniform(seed, col, level, field_a)
    u2 = uniform(seed, col, level, field_b)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def rslf(p, t):
    """Liquid saturation mixing ratio polynomial used to set relative humidity
    (same polynomial as src/module_mp_tempo_utils.F90:285-309)."""
    c = (.611583699e03, .444606896e02, .143177157e01, .264224321e-1, .299291081e-3,
         .203154182e-5, .702620698e-8, .379534310e-11, -.321582393e-13)
    x = np.maximum(-80.0, t - 273.16)
    esl = c[8]
    for ci in c[7::-1]:
        esl = ci + x * esl
    esl = np.minimum(esl, p * 0.15)
    return 0.622 * esl / (p - esl)


def generate(workload, ncol=None, col0=0, seed=None):
    """Generate inputs for ``ncol`` columns starting at global column ``col0``.

    Returns (state, meta): state maps field name -> (nz, ncol) float64 array
    (absent optional fields are omitted), meta holds dt, itimestep and, for
    ragged workloads, the per-column active level count 'nlev' (int32).
    """
    cfg = dict(WORKLOADS[workload])
    seed = cfg["seed"] if seed is None else seed
    n = cfg["ncol"] if ncol is None else ncol
    nz = cfg["nz"]
    cols = np.arange(col0, col0 + n, dtype=np.uint64)[None, :]
    lev = np.arange(nz, dtype=np.uint64)[:, None]
    col1 = cols[0]
    zero = np.zeros(1, dtype=np.uint64)

    dzl, z = levels(cfg)                                                  # layer depths, mid-points [m]
    zc = np.repeat(z[:, None], n, axis=1)
    g, rd = 9.8, 287.04
    zi = np.concatenate([[0.0], z])
    if "t_sfc" in cfg:
        # per-column surface temperature; hydrostatic pressure per column
        lo, hi = cfg["t_sfc"]
        tsfc = lo + (hi - lo) * uniform(seed, col1, zero, "tsfc")
        t = np.maximum(tsfc[None, :] - 6.5 * zc * 1e-3, 210.0)
        ti = np.maximum(tsfc[None, :] - 6.5 * zi[:, None] * 1e-3, 210.0)
        lnp = np.log(1.0e5) - np.cumsum(g * np.diff(zi)[:, None] / (rd * 0.5 * (ti[1:] + ti[:-1])), axis=0)
        p = np.exp(lnp)
    else:
        t = np.maximum(300.0 - 6.5 * zc * 1e-3, 210.0)
        # hydrostatic pressure from 1000 hPa at the surface with layer-mean temperature
        ti = np.maximum(300.0 - 6.5 * zi * 1e-3, 210.0)
        lnp = np.log(1.0e5) - np.cumsum(g * np.diff(zi) / (rd * 0.5 * (ti[1:] + ti[:-1])))
        p = np.repeat(np.exp(lnp)[:, None], n, axis=1)

    # one or two cloud layers per column: base U[0.5,8] km, thickness U[2,4] km
    # (the first from the surface up when the workload puts cloud there)
    def layer(bf, tf, surface=False):
        base = 0.0 if surface else 500.0 + 7500.0 * uniform(seed, col1, zero, bf)
        thick = 2000.0 + 2000.0 * uniform(seed, col1, zero, tf)
        return (zc >= base) & (zc <= (base + thick)[None, :])
    incloud = layer("layer1_base", "layer1_thick", cfg.get("surface_cloud", False))
    two = uniform(seed, col1, zero, "layer2_on") < 0.5
    incloud |= layer("layer2_base", "layer2_thick") & two[None, :]

    rh = np.where(incloud, 0.6 + 0.4 * uniform(seed, cols, lev, "rh"), 0.3)
    qv = rh * rslf(p, t)

    def lognormal(field, median, sigma=1.0, fb=None):
        return median * np.exp(sigma * normal(seed, cols, lev, field, fb or "q2"))
    hyd = {}
    for name in ("qc", "qr", "qi", "qs", "qg"):
        hyd[name] = np.where(incloud, lognormal(name, 1e-4, fb="q1"), 0.0)

    def number(q, rho_x, d):
        return q / (np.pi / 6.0 * rho_x * d ** 3)
    w = 0.5 * normal(seed, cols, lev, "w1", "w2")
    dzf = np.repeat(dzl[:, None], n, axis=1)
    state = dict(t=t, p=p, w=w, dz=dzf, qv=qv, qc=hyd["qc"], qr=hyd["qr"], qi=hyd["qi"],
                 qs=hyd["qs"], qg=hyd["qg"],
                 ni=number(hyd["qi"], 890.0, 50e-6), nr=number(hyd["qr"], 1000.0, 1e-3))

    conv_frac = cfg.get("convective", 0.0)
    if conv_frac > 0.0:
        conv = uniform(seed, col1, zero, "conv") < conv_frac
        wmax = 5.0 + 25.0 * uniform(seed, col1, zero, "wmax")
        zconv = 12000.0
        wconv = np.where(zc < zconv, wmax[None, :] * np.sin(np.pi * zc / zconv), 0.0)
        state["w"] = np.where(conv[None, :], wconv, state["w"])
        qr_sc = 1e-4 + (5e-3 - 1e-4) * uniform(seed, cols, lev, "qr_sc")
        sc = conv[None, :] & (t < 273.15) & (zc < zconv)
        state["qr"] = np.where(sc, qr_sc, state["qr"])
        state["nr"] = number(state["qr"], 1000.0, 1e-3)

    if cfg["aerosol"]:
        decay = np.exp(-zc / 2000.0)
        state["nc"] = number(hyd["qc"], 1000.0, 10e-6)
        state["nwfa"] = 3e8 * np.exp(normal(seed, cols, lev, "nwfa1", "nwfa2")) * decay
        state["nifa"] = 1e6 * np.exp(normal(seed, cols, lev, "nifa1", "nifa2")) * decay
    if cfg["hail"]:
        rhog = 300.0 + 600.0 * uniform(seed, cols, lev, "rhog")
        state["qb"] = 1000.0 * state["qg"] / rhog                       # L/kg
        state["ng"] = number(state["qg"], rhog, 2e-3)

    meta = dict(dt=cfg["dt"], itimestep=cfg["itimestep"], nz=nz, ncol=n, workload=workload,
                aerosol=cfg["aerosol"], hail=cfg["hail"])
    if cfg["ragged"]:
        nlev = 30 + np.floor(99.0 * uniform(seed, col1, zero, "nlev")).astype(np.int32)
        meta["nlev"] = nlev.astype(np.int32)
    state = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in state.items()}
    return state, meta


# ---- the same generator on a torch device -----------------------------------
# Bit-identical counters (int64 arithmetic wraps like uint64; logical shifts
# are masked), float64 transcendental math from the device's libm, so values
# agree with generate() to a few ulp.  Used to synthesise large batches
# directly in device memory.

def _s64(u):
    return u - (1 << 64) if u >= (1 << 63) else u


_C1, _C2, _C3 = _s64(0x9E3779B97F4A7C15), _s64(0xBF58476D1CE4E5B9), _s64(0x94D049BB133111EB)


def _shr(x, s):
    return (x >> s) & ((1 << (64 - s)) - 1)


def _splitmix_t(x):
    x = x + _C1
    x = (x ^ _shr(x, 30)) * _C2
    x = (x ^ _shr(x, 27)) * _C3
    return x ^ _shr(x, 31)


def uniform_t(seed, col, level, field):
    key = (seed << 48) ^ (FIELD_IDS[field] << 40) ^ (level << 24) ^ col
    h = _splitmix_t(_splitmix_t(key) ^ seed)
    return _shr(h, 11).double() * (1.0 / 9007199254740992.0)


def normal_t(seed, col, level, field_a, field_b):
    import torch
    u1 = uniform_t(seed, col, level, field_a)
    u2 = uniform_t(seed, col, level, field_b)
    return torch.sqrt(-2.0 * torch.log1p(-u1)) * torch.cos(2.0 * np.pi * u2)


def generate_torch(workload, ncol=None, col0=0, seed=None, device="cuda", dtype=None):
    """generate() evaluated with torch on ``device``; returns (state, meta) with
    (nz, ncol) tensors of ``dtype`` (default float64) and meta['nlev'] as an
    int32 tensor for ragged workloads."""
    import torch
    dtype = dtype or torch.float64
    cfg = dict(WORKLOADS[workload])
    seed = cfg["seed"] if seed is None else seed
    n = cfg["ncol"] if ncol is None else ncol
    nz = cfg["nz"]
    f64 = dict(dtype=torch.float64, device=device)
    cols = torch.arange(col0, col0 + n, dtype=torch.int64, device=device)[None, :]
    lev = torch.arange(nz, dtype=torch.int64, device=device)[:, None]
    col1 = cols[0]
    zero = torch.zeros(1, dtype=torch.int64, device=device)

    dzl, zl = levels(cfg)
    z = torch.tensor(zl, **f64)
    zc = z[:, None].expand(nz, n)
    g, rd = 9.8, 287.04
    zi = np.concatenate([[0.0], zl])
    if "t_sfc" in cfg:
        lo, hi = cfg["t_sfc"]
        tsfc = lo + (hi - lo) * uniform_t(seed, col1, zero, "tsfc")
        t = torch.clamp(tsfc[None, :] - 6.5 * zc * 1e-3, min=210.0)
        zit = torch.tensor(zi, **f64)
        ti = torch.clamp(tsfc[None, :] - 6.5 * zit[:, None] * 1e-3, min=210.0)
        dzi = torch.tensor(np.diff(zi), **f64)[:, None]
        lnp = np.log(1.0e5) - torch.cumsum(g * dzi / (rd * 0.5 * (ti[1:] + ti[:-1])), dim=0)
        p = torch.exp(lnp)
    else:
        t = torch.clamp(300.0 - 6.5 * zc * 1e-3, min=210.0)
        ti = np.maximum(300.0 - 6.5 * zi * 1e-3, 210.0)
        lnp = np.log(1.0e5) - np.cumsum(g * np.diff(zi) / (rd * 0.5 * (ti[1:] + ti[:-1])))
        p = torch.tensor(np.exp(lnp), **f64)[:, None].expand(nz, n)

    def layer(bf, tf, surface=False):
        base = 0.0 if surface else 500.0 + 7500.0 * uniform_t(seed, col1, zero, bf)
        thick = 2000.0 + 2000.0 * uniform_t(seed, col1, zero, tf)
        return (zc >= base) & (zc <= (base + thick)[None, :])
    incloud = layer("layer1_base", "layer1_thick", cfg.get("surface_cloud", False))
    two = uniform_t(seed, col1, zero, "layer2_on") < 0.5
    incloud = incloud | (layer("layer2_base", "layer2_thick") & two[None, :])

    def rslf_t(pp, tt):
        c = (.611583699e03, .444606896e02, .143177157e01, .264224321e-1, .299291081e-3,
             .203154182e-5, .702620698e-8, .379534310e-11, -.321582393e-13)
        x = torch.clamp(tt - 273.16, min=-80.0)
        esl = torch.full_like(x, c[8])
        for ci in c[7::-1]:
            esl = ci + x * esl
        esl = torch.minimum(esl, pp * 0.15)
        return 0.622 * esl / (pp - esl)

    rh = torch.where(incloud, 0.6 + 0.4 * uniform_t(seed, cols, lev, "rh"), torch.tensor(0.3, **f64))
    qv = rh * rslf_t(p, t)

    def lognormal(field, median, fb):
        return median * torch.exp(normal_t(seed, cols, lev, field, fb))
    zero_f = torch.tensor(0.0, **f64)
    hyd = {name: torch.where(incloud, lognormal(name, 1e-4, "q1"), zero_f)
           for name in ("qc", "qr", "qi", "qs", "qg")}

    def number(q, rho_x, d):
        return q / (np.pi / 6.0 * rho_x * d ** 3)
    w = 0.5 * normal_t(seed, cols, lev, "w1", "w2")
    state = dict(t=t, p=p, w=w, dz=torch.tensor(dzl, **f64)[:, None].expand(nz, n), qv=qv, qc=hyd["qc"], qr=hyd["qr"],
                 qi=hyd["qi"], qs=hyd["qs"], qg=hyd["qg"],
                 ni=number(hyd["qi"], 890.0, 50e-6), nr=number(hyd["qr"], 1000.0, 1e-3))

    conv_frac = cfg.get("convective", 0.0)
    if conv_frac > 0.0:
        conv = uniform_t(seed, col1, zero, "conv") < conv_frac
        wmax = 5.0 + 25.0 * uniform_t(seed, col1, zero, "wmax")
        zconv = 12000.0
        wconv = torch.where(zc < zconv, wmax[None, :] * torch.sin(np.pi * zc / zconv), zero_f)
        state["w"] = torch.where(conv[None, :], wconv, state["w"])
        qr_sc = 1e-4 + (5e-3 - 1e-4) * uniform_t(seed, cols, lev, "qr_sc")
        sc = conv[None, :] & (t < 273.15) & (zc < zconv)
        state["qr"] = torch.where(sc, qr_sc, state["qr"])
        state["nr"] = number(state["qr"], 1000.0, 1e-3)

    if cfg["aerosol"]:
        decay = torch.exp(-zc / 2000.0)
        state["nc"] = number(hyd["qc"], 1000.0, 10e-6)
        state["nwfa"] = 3e8 * torch.exp(normal_t(seed, cols, lev, "nwfa1", "nwfa2")) * decay
        state["nifa"] = 1e6 * torch.exp(normal_t(seed, cols, lev, "nifa1", "nifa2")) * decay
    if cfg["hail"]:
        rhog = 300.0 + 600.0 * uniform_t(seed, cols, lev, "rhog")
        state["qb"] = 1000.0 * state["qg"] / rhog
        state["ng"] = number(state["qg"], rhog, 2e-3)

    meta = dict(dt=cfg["dt"], itimestep=cfg["itimestep"], nz=nz, ncol=n, workload=workload,
                aerosol=cfg["aerosol"], hail=cfg["hail"])
    if cfg["ragged"]:
        meta["nlev"] = (30 + torch.floor(99.0 * uniform_t(seed, col1, zero, "nlev"))).to(torch.int32)
    state = {k: v.to(dtype).contiguous() for k, v in state.items()}
    return state, meta
