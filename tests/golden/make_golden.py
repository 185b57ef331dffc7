"""Write the golden vectors of the TEMPO column step and lookup tables.

Produced by the reference's own FP64 build (oracle/_ref/libtempo_ref.so,
compiled from the reference sources by `make -C oracle ref`, flags of
Makefile.gfortran:29) with the lookup tables built by the reference's own
table builder (`python -m oracle.ref_tables TABLE_DIR`).

    python tests/golden/make_golden.py TABLE_DIR

Each step fixture holds inputs, configuration and the reference outputs for
a small set of columns; the table fixture holds sampled table entries.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from oracle import ref  # noqa: E402
from tempo_amd import generator as gen  # noqa: E402
from tempo_amd import tables_io as tio  # noqa: E402

PLAIN = dict(aerosolaware=False, hailaware=False)
AERO = dict(aerosolaware=True, hailaware=False)
HAIL = dict(aerosolaware=False, hailaware=True)
BOTH = dict(aerosolaware=True, hailaware=True)


def mpas_profile():
    """The reference's 59-level MPAS test column (test/data/mpas_59lev_test.txt),
    converted as test/tests.F90:67-80 does."""
    rows = np.loadtxt(os.path.join("/root/reference", "test", "data", "mpas_59lev_test.txt"), skiprows=1)
    cols = ("klev qv qc qr qi qs qg ni nr nc nwfa nifa theta ng volg pressure w dz").split()
    d = {c: rows[:, i] for i, c in enumerate(cols)}
    t = d["theta"] * (d["pressure"] / 100000.) ** 0.286
    return d, t


def col(a):
    return np.ascontiguousarray(a.reshape(-1, 1), dtype=np.float64)


def run_case(name, state, dt, itimestep, flags, nlev=None, land=None, steps=1, series=False, need=()):
    """series: also store every step's 2-D diagnostics (ts_<name>, steps x
    ncol); need: 2-D diagnostics that must be non-zero in some column of
    some step, or the case does not exercise what it is for"""
    if ONLY and name not in ONLY:
        return
    ref.init(TABLES, **flags)
    inp = {k: v.copy() for k, v in state.items()}
    out = {k: v.copy() for k, v in state.items()}
    nz, ncol = state["p"].shape
    d2 = {k: np.zeros(ncol) for k in ref.DIAG2_FIELDS}
    d3 = {k: np.zeros((nz, ncol)) for k in ref.DIAG3_FIELDS}
    ts = {k: np.zeros((steps, ncol)) for k in ref.DIAG2_FIELDS}
    for s in range(steps):
        if nlev is None:
            d2, d3 = ref.run(out, dt, itimestep + s, land=land)
            for k in ts:
                ts[k][s] = d2[k]
        else:
            for c in range(ncol):
                n = int(nlev[c])
                sub = {k: np.ascontiguousarray(v[:n, c:c + 1]) for k, v in out.items()}
                e2, e3 = ref.run(sub, dt, itimestep + s, land=None if land is None else land[c:c + 1])
                for k, v in sub.items():
                    out[k][:n, c] = v[:, 0]
                for k in d2:
                    d2[k][c] = e2[k][0]
                for k in d3:
                    d3[k][:n, c] = e3[k][:, 0]
    # the C restatement on the same inputs: must be bit-identical, and its
    # table-access trace names the entries the fixture has to carry
    orc.init(TABLES, **flags)
    chk = {k: v.copy() for k, v in state.items()}
    with orc.Trace() as tr:
        for s in range(steps):
            c2, c3 = orc.run(chk, dt, itimestep + s, land=land, nlev=nlev)
    for k in chk:
        if not np.array_equal(chk[k], out[k]):
            raise SystemExit(f"{name}: C restatement differs from the reference on {k}")
    for k in need:
        if not (ts[k] != 0.0).any():
            raise SystemExit(f"{name}: diagnostic {k} is zero in every column and step")
    if CHECK_DIAGS:
        for k in CHECK_DIAGS:
            got, want = (c2[k], d2[k]) if k in c2 else (c3[k], d3[k])
            if not np.array_equal(got, want):
                raise SystemExit(f"{name}: C restatement differs from the reference on diagnostic {k}")
    meta = dict(name=name, dt=dt, itimestep=itimestep, steps=steps, flags=flags)
    arrays = {f"in_{k}": v for k, v in inp.items()}
    arrays.update({f"out_{k}": v for k, v in out.items()})
    arrays.update({f"d2_{k}": v for k, v in d2.items()})
    arrays.update({f"d3_{k}": v for k, v in d3.items()})
    if series:
        if nlev is not None:
            raise SystemExit(f"{name}: series are written for full columns only")
        arrays.update({f"ts_{k}": v for k, v in ts.items()})
    if nlev is not None:
        arrays["nlev"] = np.asarray(nlev, dtype=np.int32)
    if land is not None:
        arrays["land"] = np.asarray(land, dtype=np.int32)
    for kind in ("fz", "qs", "qg"):
        arrays[f"tab_{kind}_idx"] = tr.touched[kind].astype(np.int64)
        arrays[f"tab_{kind}_val"] = TABLES[kind][tr.touched[kind]]
    arrays["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), **arrays)
    print("wrote", name, {k: v.shape for k, v in list(inp.items())[:1]})


def check_inputs(rng, n):
    """points across the branches of the checks: absent, threshold-level and
    present masses, numbers at and beyond their limits, tendencies of both
    signs"""
    def mass():
        return np.where(rng.random(n) < 0.2, 0.0, 10.0 ** rng.uniform(-13, -2.5, n))
    def number(lo, hi):
        return np.where(rng.random(n) < 0.15, 0.0, 10.0 ** rng.uniform(lo, hi, n))
    def ten(q):
        return np.where(rng.random(n) < 0.5, 0.0, rng.normal(0.0, 1.0, n) * q / 60.0)
    rho = rng.uniform(0.3, 1.3, n)
    qc, qi, qs = mass(), mass(), mass()
    a = dict(rho=rho, l_qc=np.zeros(n, np.int32), qc1d=qc, nc1d=number(0, 10), ncsave=10.0 ** rng.uniform(7, 9, n),
             rc=np.zeros(n), nc=np.zeros(n), qcten=ten(qc), ncten=ten(np.full(n, 1e8)), ilamc=np.zeros(n),
             mvd_c=np.zeros(n), l_qi=np.zeros(n, np.int32), qi1d=qi, ni1d=number(0, 7), ri=np.zeros(n),
             ni=np.zeros(n), qiten=ten(qi), niten=ten(np.full(n, 1e5)), ilami=np.zeros(n), l_qs=np.zeros(n, np.int32),
             qs1d=qs, rs=np.zeros(n), qsten=ten(qs))
    return {k: np.ascontiguousarray(v) for k, v in a.items()}


def write_check_fixture():
    ref.init(TABLES, **BOTH)
    rng = np.random.default_rng(31)
    inp = check_inputs(rng, 4000)
    dt = 20.0
    out = {}
    for case, which, kw in (("cloud2", "cloud", dict(has_nc=True)), ("cloud_ncsave", "cloud", dict(has_nc=False, has_ncsave=True)),
                            ("cloud_land", "cloud", dict(has_nc=False, has_ncsave=False)), ("ice", "ice", {}),
                            ("snow", "snow", {})):
        a = {k: v.copy() for k, v in inp.items()}
        ref.check_and_update(which, a, dt, **kw)
        out.update({f"{case}__{k}": v for k, v in a.items()})
    np.savez_compressed(os.path.join(HERE, "check_and_update.npz"), dt=np.float64(dt),
                        **{f"in__{k}": v for k, v in inp.items()}, **out)
    print("wrote check_and_update")


def pick(state, idx):
    return {k: np.ascontiguousarray(v[:, idx]) for k, v in state.items()}


ONLY = ()
# diagnostics the C restatement must reproduce bit for bit in the next cases
CHECK_DIAGS = ()


def main():
    global TABLES, ONLY, CHECK_DIAGS
    if len(sys.argv) < 2:
        raise SystemExit("usage: make_golden.py TABLE_DIR [CASE ...]  (tables written by `python -m oracle.ref_tables`;"
                         " with CASE names only those fixtures are rewritten)")
    TABLES = tio.read_dir(sys.argv[1])
    ONLY = tuple(sys.argv[2:])

    # ---- workload samples (generator columns) ----
    st, m = gen.generate("plain", ncol=40)
    # append clear-sky columns: no hydrometeors, dry (early exit) and ice-supersaturated
    clear = pick(st, np.arange(8))
    for k in ("qc", "qr", "qi", "qs", "qg", "ni", "nr"):
        clear[k][:] = 0.0
    clear["qv"][:, :4] *= 0.1
    clear["qv"][:, 4:] = gen.rslf(clear["p"][:, 4:], clear["t"][:, 4:]) * 1.2
    st = {k: np.concatenate([st[k], clear[k]], axis=1) for k in st}
    run_case("plain", st, m["dt"], m["itimestep"], PLAIN)

    st, m = gen.generate("aerosol", ncol=32)
    run_case("aerosol", st, m["dt"], m["itimestep"], AERO)

    st, m = gen.generate("ragged", ncol=32)
    run_case("ragged", st, m["dt"], m["itimestep"], HAIL, nlev=m["nlev"])

    st, m = gen.generate("multistep", ncol=32)
    run_case("multistep", st, m["dt"], m["itimestep"], BOTH)

    st, m = gen.generate("adjoint", ncol=32)
    run_case("adjoint_shape", st, m["dt"], m["itimestep"], AERO)

    # first step: graupel number/volume and aerosols initialised by the step
    st, m = gen.generate("multistep", ncol=24, col0=100)
    for k in ("ng", "qb", "nwfa", "nifa"):
        st[k][:] = 0.0
    run_case("first_step", st, m["dt"], 1, BOTH)

    # theta/exner input, land mask with one-moment cloud number
    st, m = gen.generate("plain", ncol=24, col0=500)
    pii = (st["p"] / 1.0e5) ** 0.286
    th = st.pop("t") / pii
    st["th"], st["pii"] = np.ascontiguousarray(th), np.ascontiguousarray(pii)
    land = (np.arange(24) % 2).astype(np.int32)
    run_case("thpii_land", st, m["dt"], m["itimestep"], PLAIN, land=land)

    # option switches and timestep sweep (test/run_tempo_tests.F90:10-11)
    st, m = gen.generate("plain", ncol=16, col0=900)
    run_case("no_condensation", st, 30.0, 2, dict(PLAIN, cloud_condensation=False))
    run_case("micro_off", st, 30.0, 2, dict(PLAIN, turn_off_micro=True))
    run_case("ocean", st, 30.0, 2, dict(PLAIN, nc_opt=1))
    for dt in (1.0, 10.0, 60.0, 300.0, 600.0):
        run_case(f"dt{int(dt)}", st, dt, 2, PLAIN)

    # the reference's own single-column profile: graupel sedimentation test
    # (test/tests.F90:24-129, microphysics off) and a full-physics step
    d, t = mpas_profile()
    base = dict(t=col(t), p=col(d["pressure"]), dz=col(d["dz"]), w=col(np.zeros(59)), qv=col(d["qv"]),
                qc=col(np.zeros(59)), qr=col(np.zeros(59)), qi=col(np.zeros(59)), qs=col(np.zeros(59)),
                qg=col(d["qg"]), ni=col(np.zeros(59)), nr=col(np.zeros(59)), ng=col(d["ng"]),
                qb=col(d["volg"] * 1000.))
    run_case("mpas_graupel_sedi", base, 20.0, 1, dict(BOTH, turn_off_micro=True, mvd_g=True), steps=10)
    full = dict(t=col(t), p=col(d["pressure"]), dz=col(d["dz"]), w=col(d["w"]))
    for k in ("qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr", "nc", "nwfa", "nifa", "ng"):
        full[k] = col(d[k])
    full["qb"] = col(d["volg"] * 1000.)
    run_case("mpas_full", full, 20.0, 2, dict(BOTH, mvd_r=True, mvd_g=True, nc_mixing_ratio=True), steps=3)

    # semi-Lagrangian sedimentation (semi_sedi_flag): the reference's graupel
    # test with semi_sedi=.true. (test/run_tempo_tests.F90:29-31) at two of its
    # time steps, and full-physics columns with rain and graupel
    semi = dict(BOTH, turn_off_micro=True, mvd_g=True, semi_sedi=True)
    run_case("semi_mpas_graupel", base, 20.0, 1, semi, steps=10)
    run_case("semi_mpas_graupel_dt300", base, 300.0, 1, semi, steps=3)
    st, m = gen.generate("multistep", ncol=24, col0=300)
    run_case("semi_multistep", st, m["dt"], m["itimestep"], dict(BOTH, semi_sedi=True), steps=2)
    st, m = gen.generate("plain", ncol=16, col0=900)
    run_case("semi_plain_dt300", st, 300.0, 2, dict(PLAIN, semi_sedi=True))
    # cloud number from the trained network: one-moment cloud (ml_for_nc,
    # test/tests.F90:380-461) and boundary-layer cloud in the cloud radius
    # (ml_for_bl_nc with qc_bl/qcfrac_bl, test/tests.F90:464-558), on the
    # reference's column and on generated columns
    CHECK_DIAGS = ("re_cloud", "nc_mixing_ratio")
    one = {k: v for k, v in full.items() if k != "nc"}
    run_case("ml_nc_mpas_micro_off", one, 20.0, 1, dict(PLAIN, turn_off_micro=True, nc_mixing_ratio=True,
                                                         ml_for_nc=True), steps=3)
    run_case("ml_nc_mpas", one, 20.0, 1, dict(PLAIN, nc_mixing_ratio=True, ml_for_nc=True), steps=3)
    st, m = gen.generate("plain", ncol=16, col0=1300)
    run_case("ml_nc_plain", st, m["dt"], m["itimestep"], dict(PLAIN, nc_mixing_ratio=True, ml_for_nc=True),
             steps=2)
    qc_bl = np.zeros(59)
    qc_bl[:19] = d["qc"][15:34] * 0.2
    frac = np.where(qc_bl > 0.0, 0.35, 0.0)
    bl = dict(full, qc_bl=col(qc_bl), qcfrac_bl=col(frac))
    run_case("ml_bl_mpas_micro_off", bl, 20.0, 1, dict(PLAIN, turn_off_micro=True, nc_mixing_ratio=True,
                                                       ml_for_bl_nc=True), steps=3)
    run_case("ml_bl_mpas", bl, 20.0, 1, dict(BOTH, nc_mixing_ratio=True, ml_for_bl_nc=True), steps=2)
    st, m = gen.generate("aerosol", ncol=16, col0=700)
    rng = np.random.default_rng(11)
    nz, ncol = st["p"].shape
    qbl = np.where(rng.random((nz, ncol)) < 0.4, rng.uniform(1e-8, 1e-3, (nz, ncol)), 0.0)
    st["qc_bl"] = np.ascontiguousarray(qbl)
    st["qcfrac_bl"] = np.ascontiguousarray(np.where(qbl > 0.0, rng.uniform(0.05, 1.0, (nz, ncol)), 0.0))
    run_case("ml_bl_aerosol", st, m["dt"], m["itimestep"], dict(AERO, ml_for_bl_nc=True), steps=2)
    # reflectivity with melting snow and graupel below the melting level
    # (refl10cm_from_melting_flag, src/module_mp_tempo_diags.F90:59-395)
    CHECK_DIAGS = ("refl10cm",)
    run_case("melt_mpas", full, 20.0, 2, dict(BOTH, refl_from_melting=True), steps=2)
    st, m = gen.generate("multistep", ncol=32, col0=2100)
    run_case("melt_multistep", st, m["dt"], m["itimestep"], dict(BOTH, refl_from_melting=True), steps=2)
    st, m = gen.generate("ragged", ncol=24, col0=3300)
    run_case("melt_ragged", st, m["dt"], m["itimestep"], dict(HAIL, refl_from_melting=True), nlev=m["nlev"])
    CHECK_DIAGS = ()

    # the reference's own integration set-ups (test/tests.F90, driven by
    # test/run_tempo_tests.F90:19-40 after tempo_init with the default flags):
    # 1,200 s of sedimentation alone at every time step of its sweep, with the
    # surface precipitation of every step stored
    sweep = (1.0, 10.0, 20.0, 60.0, 120.0, 300.0, 600.0)
    zeros = col(np.zeros(59))
    snow = dict(base, qs=col(d["qs"]), qg=zeros, ng=zeros, qb=zeros)
    for dt in sweep:
        n = int(1200 / dt)
        # test_snow_sedimentation (test/tests.F90:132-214): the column's snow
        # starts 2.4 km up and does not reach the ground within 1,200 s
        run_case(f"ref_snow_sedi_dt{int(dt)}", snow, dt, 1, dict(BOTH, turn_off_micro=True), steps=n, series=True)
        # test_graupel_sedimentation, flux form and semi-Lagrangian (:24-129)
        run_case(f"ref_graupel_sedi_dt{int(dt)}", base, dt, 1, dict(BOTH, turn_off_micro=True, mvd_g=True),
                 steps=n, series=True, need=("graupel_precip",))
        run_case(f"ref_graupel_semi_dt{int(dt)}", base, dt, 1,
                 dict(BOTH, turn_off_micro=True, mvd_g=True, semi_sedi=True), steps=n, series=True,
                 need=("graupel_precip",))
    # test_cloud_number_aerosolaware / _non_aerosolaware / _ml (:217-461),
    # dt = 20 s over 1,200 s
    cn = dict(t=col(t), p=col(d["pressure"]), dz=col(d["dz"]), w=col(d["w"]), qv=col(d["qv"]), qc=col(d["qc"]),
              qr=col(d["qr"]), nr=col(d["nr"]), qi=col(d["qi"]), qs=col(d["qs"]), ni=col(d["ni"]), qg=zeros)
    run_case("ref_cloud_number_aero", dict(cn, nc=col(d["nc"])), 20.0, 1, dict(BOTH, turn_off_micro=True),
             steps=60, series=True)
    CHECK_DIAGS = ("nc_mixing_ratio",)
    run_case("ref_cloud_number_const", cn, 20.0, 1, dict(BOTH, turn_off_micro=True, nc_mixing_ratio=True),
             steps=60, series=True)
    run_case("ref_cloud_number_ml", cn, 20.0, 1,
             dict(BOTH, turn_off_micro=True, nc_mixing_ratio=True, ml_for_nc=True), steps=60, series=True)
    CHECK_DIAGS = ()
    # below-freezing surface with cloud down to it: freezing rain and ice,
    # snow and graupel reaching the ground, in both sedimentation forms
    st, m = gen.generate("cold", ncol=24, col0=0)
    allp = ("rain_precip", "ice_precip", "snow_precip", "graupel_precip", "frz_rain_precip")
    run_case("cold_surface", st, m["dt"], m["itimestep"], BOTH, steps=3, series=True, need=allp)
    run_case("cold_surface_semi", st, m["dt"], m["itimestep"], dict(BOTH, semi_sedi=True), steps=3, series=True,
             need=allp)
    st, m = gen.generate("cold", ncol=16, col0=5000)
    run_case("cold_surface_plain_dt60", {k: v for k, v in st.items() if k not in ("nc", "nwfa", "nifa", "ng", "qb")},
             60.0, 2, PLAIN, steps=2, series=True, need=("rain_precip", "frz_rain_precip", "snow_precip"))

    # tempo_aerosol_surface_emissions (src/module_mp_tempo_driver.F90:516-529)
    if not ONLY or "surface_emissions" in ONLY:
        rng = np.random.default_rng(77)
        nwfa = np.ascontiguousarray(rng.uniform(1e7, 1e9, (3, 6, 7)))  # (j, k, i): Fortran (7, 6, 3)
        nwfa2d = np.ascontiguousarray(rng.uniform(0.0, 5e4, (3, 7)))
        out = nwfa.copy()
        ref.surface_emissions(20.0, out, nwfa2d, 2)
        chk = nwfa.copy()
        orc.surface_emissions(20.0, chk, nwfa2d, 2)
        if not np.array_equal(chk, out):
            raise SystemExit("surface_emissions: C restatement differs from the reference")
        np.savez_compressed(os.path.join(HERE, "surface_emissions.npz"), nwfa=nwfa, nwfa2d=nwfa2d, out=out,
                            dt=np.float64(20.0), kts=np.int32(2))
        print("wrote surface_emissions")
    # the public cloud/ice/snow_check_and_update (src/module_mp_tempo_main.F90:18-20)
    if not ONLY or "check_and_update" in ONLY:
        write_check_fixture()
    if ONLY:
        return

    # ---- sampled lookup-table entries ----
    rng = np.random.default_rng(2024)
    out = {}
    for kind, shape, narr in (("qg", (37, 37, 9, 37, 37), 5), ("qs", (37, 9, 37, 37), 12),
                              ("fzr", (37, 37, 45, 55), 4), ("fzc", (37, 100, 45, 55), 2)):
        n1 = int(np.prod(shape))
        idx = rng.integers(0, n1, size=3000)
        idx = np.unique(np.concatenate([idx, [0, n1 - 1, n1 // 2]]))
        if kind == "fzr":
            flat = TABLES["fz"][:4 * n1]
        elif kind == "fzc":
            flat = TABLES["fz"][4 * tio.N_FZ_R:]
        else:
            flat = TABLES[kind]
        vals = np.stack([flat[a * n1 + idx] for a in range(narr)], axis=1)
        out[f"{kind}_idx"] = idx.astype(np.int64)
        out[f"{kind}_val"] = vals
    np.savez_compressed(os.path.join(HERE, "tables_sample.npz"), **out)
    print("wrote tables_sample")


if __name__ == "__main__":
    main()
