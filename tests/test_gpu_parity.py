"""Parity of the HIP path (through the C ABI) with the reference golden
vectors and the C oracle, on an MI355X.

Tolerances (stated per test):
  FP64 twin vs reference / oracle: relative 1e-12 per field, normalised by
      the field's largest magnitude (|gpu - ref|_inf / |ref|_inf).
  FP32 production path vs the FP64 twin, both run on the same
      FP32-representable inputs: relative L2 of every tendency <= 1e-4 over
      the points whose branch decisions agree, surface precipitation <= 1e-5
      over all columns.  A point counts as branch-flipped when its result
      moved by more than 1e-3 of its value (plus 1e-6 of the field's largest
      magnitude); the flipped fraction is reported and bounded, and compared
      with the reference's own single-precision build (-DSINGLE_PREC) on the
      same inputs when that build is present.
"""
import os
import tempfile

import numpy as np
import pytest

from conftest import ILL_CONDITIONED, golden_cases, load_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STATE_OUT = ("t", "th", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr", "nc", "nwfa", "nifa", "ng", "qb")
D2_MAP = dict(rain_precip="rain_precip", ice_precip="ice_liquid_equiv_precip",
              snow_precip="snow_liquid_equiv_precip", graupel_precip="graupel_liquid_equiv_precip",
              frozen_fraction="frozen_fraction", frz_rain_precip="frz_rain_precip",
              max_hail_sfc="max_hail_diameter_sfc", max_hail_column="max_hail_diameter_column")
D3_MAP = dict(refl10cm="refl10cm", re_cloud="re_cloud", re_ice="re_ice", re_snow="re_snow",
              mvd_r="rain_med_vol_diam", mvd_g="graupel_med_vol_diam", nc_mixing_ratio="cloud_number_mixing_ratio")
FLAG_TO_CFG = dict(aerosolaware="aerosolaware_flag", hailaware="hailaware_flag", turn_off_micro="turn_off_micro_flag",
                   cloud_condensation="cloud_condensation_flag", refl10cm="refl10cm_flag", re_cloud="re_cloud_flag",
                   re_ice="re_ice_flag", re_snow="re_snow_flag", max_hail="max_hail_diameter_flag",
                   mvd_r="rain_med_vol_diam_flag", mvd_g="graupel_med_vol_diam_flag",
                   nc_mixing_ratio="cloud_number_mixing_ratio_flag", semi_sedi="semi_sedi_flag",
                   ml_for_nc="ml_for_nc_flag", ml_for_bl_nc="ml_for_bl_nc_flag",
                   refl_from_melting="refl10cm_from_melting_flag")


def cfg_of(flags):
    cfg = {}
    for k, v in flags.items():
        if k == "nc_opt":
            cfg["single_moment_nc_opt"] = "ocean" if v == 1 else "land"
        elif k in FLAG_TO_CFG:
            cfg[FLAG_TO_CFG[k]] = v
    return cfg


@pytest.fixture(scope="module")
def hip():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from tempo_amd import host
    host.lib()  # fails loudly if the library is missing
    return host


_ctx_cache = {}


def context(hip, flags):
    key = tuple(sorted(cfg_of(flags).items()))
    if key not in _ctx_cache:
        _ctx_cache[key] = hip.TempoHip(**cfg_of(flags))
    return _ctx_cache[key]


@pytest.fixture(scope="module")
def gpu_tables(hip):
    ctx = context(hip, dict(aerosolaware=True, hailaware=True))
    return dict(ccn=__import__("tempo_amd.tables_io", fromlist=["x"]).read_ccn(),
                fz=ctx.export_table(hip.TABLE_FZ), qs=ctx.export_table(hip.TABLE_QS),
                qg=ctx.export_table(hip.TABLE_QG), build_ms=ctx.table_build_ms)


def to_dev(state, dtype):
    return {k: torch.tensor(v, dtype=dtype, device="cuda") for k, v in state.items()}


def gpu_run(ctx, case_inp, dt, itimestep, steps=1, dtype=torch.float64, nlev=None, land=None, series=None):
    """steps of the case on the GPU; series (dict) receives every step's 2-D
    diagnostics under the reference's names"""
    st = to_dev(case_inp, dtype)
    nl = None if nlev is None else torch.tensor(nlev, dtype=torch.int32, device="cuda")
    ld = None if land is None else torch.tensor(land, dtype=torch.int32, device="cuda")
    for s in range(steps):
        d = ctx.run(st, dt, itimestep + s, nlev=nl, land=ld)
        if series is not None:
            for k, n in D2_MAP.items():
                series.setdefault(k, []).append(d[n].double().cpu().numpy())
    torch.cuda.synchronize()
    return {k: v.double().cpu().numpy() for k, v in st.items()}, {k: v.double().cpu().numpy() for k, v in d.items()}


def rel_inf(a, b):
    scale = np.abs(b).max()
    if scale == 0:
        return float(np.abs(a).max())
    return float(np.abs(a - b).max() / scale)


def flipped(a, b, tol=1e-3, floor=1e-6):
    """points whose result moved by more than rounding can explain"""
    return np.abs(a - b) > tol * np.abs(b) + floor * np.abs(b).max()


def flip_fraction(a, b, tol=1e-3):
    return float(np.mean(flipped(a, b, tol)))


def test_gpu_tables_match_reference_builder(hip, gpu_tables):
    """FP64 GPU table build vs entries of the tables produced by the
    reference's own builder: identical summation order, so the entries match
    to the last bit except where the host libm rounds a factor differently."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "tables_sample.npz"))
    from tempo_amd import tables_io as tio
    sizes = dict(qg=tio.N_QG1, qs=tio.N_QS1, fzr=tio.N_FZ_R, fzc=tio.N_FZ_C)
    src = dict(qg=gpu_tables["qg"], qs=gpu_tables["qs"], fzr=gpu_tables["fz"][:4 * tio.N_FZ_R],
               fzc=gpu_tables["fz"][4 * tio.N_FZ_R:])
    report = {}
    for kind in ("qg", "qs", "fzr", "fzc"):
        idx, val = z[f"{kind}_idx"], z[f"{kind}_val"]
        n1 = sizes[kind]
        got = np.stack([src[kind][a * n1 + idx] for a in range(val.shape[1])], axis=1)
        rel = np.abs(got - val) / np.maximum(np.abs(val), 1e-300)
        report[kind] = (float(rel.max()), float(np.mean(got == val)))
        assert rel.max() <= 1e-13, (kind, rel.max())
    print("table parity (max rel, exact fraction):", report, "build ms", gpu_tables["build_ms"])


def test_gpu_tables_match_oracle_entries(hip, gpu_tables):
    from oracle import oracle as orc
    from tempo_amd import tables_io as tio
    orc.init(gpu_tables, aerosolaware=False, hailaware=False)
    rng = np.random.default_rng(5)
    qg = tio.view(gpu_tables, "qg")
    for _ in range(8):
        i, j, n3, k, m = (int(x) for x in rng.integers(1, [38, 38, 10, 38, 38]))
        ref = orc.qg_entry(i, j, n3, k, m)
        got = np.array([qg[n][i - 1, j - 1, n3 - 1, k - 1, m - 1] for n in tio.QG_NAMES])
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("name", [n for n in golden_cases() if n not in ILL_CONDITIONED])
def test_gpu_fp64_matches_reference_golden(hip, name):
    """FP64 twin vs the reference FP64 build on the golden columns, over all
    of each case's steps."""
    case = load_case(name)
    meta = case["meta"]
    ctx = context(hip, meta["flags"])
    series = {}
    st, d = gpu_run(ctx, case["inp"], meta["dt"], meta["itimestep"], meta["steps"], nlev=case["nlev"],
                    land=case["land"], series=series)
    has_q = np.zeros(case["inp"]["p"].shape[1], bool)
    for q in ("qc", "qr", "qi", "qs", "qg"):
        has_q |= (case["inp"][q] > 1e-12).any(axis=0)
    worst = {}
    for k, v in case["out"].items():
        worst[k] = rel_inf(st[k], v)
        assert worst[k] <= 1e-12, (name, k, worst[k])
    for k, v in case["d2"].items():
        g = d[D2_MAP[k]]
        if k == "frozen_fraction":  # undefined in the reference after an early exit
            g, v = g[has_q], v[has_q]
        worst[k] = rel_inf(g, v)
        assert worst[k] <= 1e-12, (name, k, worst[k])
    for k, v in case["d3"].items():
        worst[k] = rel_inf(d[D3_MAP[k]], v)
        assert worst[k] <= 1e-12, (name, k, worst[k])
    # the surface precipitation of every step, where the case stores it
    for k, v in case["ts"].items():
        got = np.array(series[k])
        if k == "frozen_fraction":
            got, v = got[:, has_q], v[:, has_q]
        worst["ts_" + k] = rel_inf(got, v)
        assert worst["ts_" + k] <= 1e-12, (name, "step series", k, worst["ts_" + k])
    print(name, "max rel", max(worst.values()))


@pytest.mark.parametrize("name", ILL_CONDITIONED)
def test_gpu_fp64_stepwise_on_ill_conditioned_case(hip, gpu_tables, name):
    """The ill-conditioned golden cases step by step: every step starts from
    the reference's own state (the oracle's, bit-identical to the reference
    on the case) and the FP64 twin's step must match the reference's step to
    1e-12 -- state, surface precipitation and 3-D diagnostics, except 3-D
    diagnostics at points where graupel volume sits at its clamp and the
    density index nint(10 qg/qb) is taken at an exact half (at most 2 points
    per step)."""
    from oracle import oracle as orc
    case = load_case(name)
    meta = case["meta"]
    ctx = context(hip, meta["flags"])
    orc.init(gpu_tables, **meta["flags"])
    st = {k: v.copy() for k, v in case["inp"].items()}
    worst = 0.0
    for s_ in range(meta["steps"]):
        prev = {k: v.copy() for k, v in st.items()}
        d2, d3 = orc.run(st, meta["dt"], meta["itimestep"] + s_)
        got, d = gpu_run(ctx, prev, meta["dt"], meta["itimestep"] + s_)
        for k in st:
            e = rel_inf(got[k], st[k])
            worst = max(worst, e)
            assert e <= 1e-12, (name, s_, k, e)
        for k, v in d2.items():
            e = rel_inf(d[D2_MAP[k]], v)
            assert e <= 1e-12, (name, s_, k, e)
        for k, v in d3.items():
            bad = np.abs(d[D3_MAP[k]] - v) > 1e-12 * max(np.abs(v).max(), 1e-300)
            assert bad.sum() <= 2, (name, s_, k, np.argwhere(bad).tolist())
    # and the oracle trajectory is the reference's
    for k, v in case["out"].items():
        np.testing.assert_array_equal(st[k], v, err_msg=k)
    print(name, meta["steps"], "steps, max rel per step", worst)


def reference_steps(gpu_tables, flags, st, dt, itimestep, steps=1, nlev=None):
    """The reference FP64 build (oracle/_ref/libtempo_ref.so, -O0, compiled
    from the reference sources) over ``steps`` steps of a copy of ``st``;
    ragged columns (nlev) run as columns of their own level counts, levels
    above them untouched.  Returns (state, d2, d3) of the last step."""
    from oracle import ref
    if not ref.available("o0"):
        pytest.skip("reference build not present")
    ref.init(gpu_tables, variant="o0", **flags)
    out = {k: v.copy() for k, v in st.items()}
    if nlev is None:
        for s in range(steps):
            d2, d3 = ref.run(out, dt, itimestep + s, variant="o0")
        return out, d2, d3
    nz, ncol = st["p"].shape
    d2 = {k: np.zeros(ncol) for k in ref.DIAG2_FIELDS}
    d3 = {k: np.zeros((nz, ncol)) for k in ref.DIAG3_FIELDS}
    nlev = np.asarray(nlev)
    for n in np.unique(nlev):
        cols = np.where(nlev == n)[0]
        sub = {k: np.ascontiguousarray(v[:n][:, cols]) for k, v in out.items()}
        for s in range(steps):
            e2, e3 = ref.run(sub, dt, itimestep + s, variant="o0")
        for k, v in sub.items():
            out[k][:n, cols] = v
        for k, v in e2.items():
            d2[k][cols] = v
        for k, v in e3.items():
            d3[k][:n, cols] = v
    return out, d2, d3


def check_against_reference(out, d, ref_out, d2, d3, inp, what, d3_names=None):
    """FP64 twin vs the reference build: state, 2-D and 3-D diagnostics to
    1e-12 relative (frozen_fraction over the columns holding hydrometeors:
    the reference leaves it undefined after an early exit)"""
    has_q = np.zeros(inp["p"].shape[1], bool)
    for q in ("qc", "qr", "qi", "qs", "qg"):
        has_q |= (inp[q] > 1e-12).any(axis=0)
    worst = {}
    for k in ref_out:
        worst[k] = rel_inf(out[k], ref_out[k])
    for k, v in d2.items():
        g = d[D2_MAP[k]]
        if k == "frozen_fraction":
            g, v = g[has_q], v[has_q]
        if g.size:
            worst[k] = rel_inf(g, v)
    for k in (d3_names or d3):
        worst[k] = rel_inf(d[D3_MAP[k]], d3[k])
    print(what, "FP64 vs reference build, max rel", max(worst.values()))
    for k, e in worst.items():
        assert e <= 1e-12, (what, k, e)


@pytest.fixture(scope="module")
def big_case(hip, gpu_tables):
    """2,048 aerosol- and hail-aware columns of 90 levels, one step of the
    reference build"""
    from tempo_amd import generator as gen
    st, m = gen.generate("multistep", ncol=2048, col0=50_000)
    ref, d2, d3 = reference_steps(gpu_tables, dict(aerosolaware=True, hailaware=True), st, m["dt"], m["itimestep"])
    return st, m, ref, d2, d3


def test_gpu_fp64_matches_reference_at_2048_columns(hip, big_case):
    st, m, ref, d2, d3 = big_case
    ctx = context(hip, dict(aerosolaware=True, hailaware=True))
    out, d = gpu_run(ctx, st, m["dt"], m["itimestep"])
    check_against_reference(out, d, ref, d2, d3, st, "2048 columns", ("refl10cm", "re_cloud", "re_ice", "re_snow"))


def test_gpu_fp64_matches_reference_on_the_plain_configuration(hip, gpu_tables):
    """The reference's CPU-runnable configuration at its full size: 1,024
    columns x 60 levels, dt 30 s, aerosol- and hail-aware off, two steps --
    the FP64 twin against the reference FP64 build on the same inputs"""
    from tempo_amd import generator as gen
    st, m = gen.generate("plain")
    assert st["p"].shape == (60, 1024) and m["dt"] == 30.0
    flags = dict(aerosolaware=False, hailaware=False)
    ref, d2, d3 = reference_steps(gpu_tables, flags, st, m["dt"], m["itimestep"], steps=2)
    out, d = gpu_run(context(hip, flags), st, m["dt"], m["itimestep"], steps=2)
    check_against_reference(out, d, ref, d2, d3, st, "1024 x 60 plain", ("refl10cm", "re_cloud", "re_ice", "re_snow"))


def test_gpu_oracle_is_the_reference_at_2048_columns(hip, gpu_tables, big_case):
    """the C restatement on the same columns, bit for bit"""
    from oracle import oracle as orc
    st, m, ref, d2, d3 = big_case
    orc.init(gpu_tables, aerosolaware=True, hailaware=True)
    o = {k: v.copy() for k, v in st.items()}
    e2, e3 = orc.run(o, m["dt"], m["itimestep"])
    for k in ref:
        np.testing.assert_array_equal(o[k], ref[k], err_msg=k)


def test_gpu_single_column_of_two_levels(hip, gpu_tables):
    """nz = 2, ncol = 1 (the smallest step the library takes): FP64 twin vs
    the reference build, and vs the oracle"""
    from oracle import oracle as orc
    from tempo_amd import generator as gen
    flags = dict(aerosolaware=True, hailaware=True)
    full, m = gen.generate("multistep", ncol=64, col0=3)
    nz = full["p"].shape[0]
    # the two levels of a column around its cloud
    has = (full["qc"] > 1e-12) | (full["qr"] > 1e-12) | (full["qs"] > 1e-12)
    c = int(np.argmax(has.sum(axis=0)))
    k0 = int(np.where(has[:, c])[0][0]) if has[:, c].any() else 0
    k0 = min(k0, nz - 2)
    st = {k: np.ascontiguousarray(v[k0:k0 + 2, c:c + 1]) for k, v in full.items()}
    ref, d2, d3 = reference_steps(gpu_tables, flags, st, m["dt"], m["itimestep"], steps=3)
    out, d = gpu_run(context(hip, flags), st, m["dt"], m["itimestep"], steps=3)
    check_against_reference(out, d, ref, d2, d3, st, "2 levels x 1 column")
    orc.init(gpu_tables, **flags)
    o = {k: v.copy() for k, v in st.items()}
    for s in range(3):
        orc.run(o, m["dt"], m["itimestep"] + s)
    for k in ref:
        np.testing.assert_array_equal(o[k], ref[k], err_msg=k)


def test_gpu_semi_sedi_fp64_matches_oracle(hip, gpu_tables):
    """Semi-Lagrangian rain and graupel sedimentation on 512 aerosol- and
    hail-aware columns at dt = 120 s (several semi-Lagrangian substeps in the
    convective columns): FP64 path vs the oracle."""
    from tempo_amd import generator as gen
    st, m = gen.generate("multistep", ncol=512, col0=70_000)
    flags = dict(aerosolaware=True, hailaware=True, semi_sedi=True)
    ref, d2, d3 = reference_steps(gpu_tables, flags, st, 120.0, m["itimestep"])
    out, d = gpu_run(context(hip, flags), st, 120.0, m["itimestep"])
    for k in ref:
        assert rel_inf(out[k], ref[k]) <= 1e-12, (k, rel_inf(out[k], ref[k]))
    for k, v in d2.items():
        assert rel_inf(d[D2_MAP[k]], v) <= 1e-12, k
    # The remap of the departure cells turns one-ulp differences of the fall
    # speeds (device pow vs the host libm in about 1e-3 of calls) into
    # 1e-14-relative differences of the remapped graupel; where graupel
    # volume sits at its clamp (qb = 20 qg), the density index
    # nint(10 qg/qb) + 1 is taken exactly at 0.5 and flips with them, moving
    # the 3-D diagnostics of that point.  Such points are counted, not failed.
    for k, v in d3.items():
        bad = np.abs(d[D3_MAP[k]] - v) > 1e-12 * max(np.abs(v).max(), 1e-300)
        assert bad.sum() <= 8, (k, int(bad.sum()), np.argwhere(bad)[:8].tolist())


TENDENCY_FIELDS = ("t", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr", "nc", "ng", "qb", "nwfa", "nifa")


def moved_fraction(inp, out, o64, tol=1e-3):
    """per field, the fraction of points whose FP32 result moved from the
    FP64 one by more than ``tol`` relative (a magnitude measure, used only
    to set our FP32 path beside the reference's own single-precision build,
    which records no branch signatures)"""
    return {k: flip_fraction(out[k], o64[k], tol) for k in TENDENCY_FIELDS if k in inp}


def check_fp32_report(rep, precip_all=True, tendencies=True, theirs=None, semi=False):
    """the FP32 tolerances (tempo_amd/branch_report.py defines the measures):
    every tendency within 1e-4 relative L2 over the points that took the FP64
    run's branches (tendencies) -- or, for a field whose error in the
    reference's own single-precision build on the same inputs and the same
    points (theirs) exceeds 1e-4, within twice that; surface precipitation within 1e-5
    relative L2 over all columns (precip_all), or else over the columns that
    did not move by more than 1e-4, and (flux form) every column that did
    moved with a branch flip at its lowest two levels or in its
    sedimentation"""
    if tendencies:
        for k, e in rep["fields"].items():
            tol = 1e-4 if theirs is None else max(1e-4, 2.0 * theirs[k]["rel_l2"])
            assert e["rel_l2"] <= tol, (k, e, None if theirs is None else theirs[k])
    p = rep["precip"]
    if precip_all:
        assert p["rel_l2_all"] <= 1e-5, p
    else:
        assert p["rel_l2"] <= 1e-5, p
        if not semi:
            assert p["moved_columns_unexplained"] == 0, p


def report_line(tag, rep):
    from tempo_amd import branch_report as br
    s = br.summary(rep)
    print(tag, f"points with a flipped branch {s['flipped_points']:.4f} (discrete outcomes "
          f"{s['flipped_points_discrete']:.4f}, comparisons {s['flipped_points_comparisons']:.4f}), "
          f"columns with flipped sedimentation {s['flipped_sedimentation']:.4f}")
    for k, e in s["fields"].items():
        print(f"  {k}: rel L2 {e['rel_l2']:.3e} (all points {e['rel_l2_all']:.3e}, "
              f"against unrounded FP64 {e['rel_l2_raw']:.3e})")
    p = s["precip"]
    print(f"  precip: rel L2 all columns {p['rel_l2_all']:.3e}, {p['nonzero_columns']} columns with precipitation, "
          f"{p['moved_columns']} moved > 1e-4 ({p['moved_columns_unexplained']} without a low-level flip)")
    for n in br.PRECIP + ("frz_rain_precip",):
        print(f"    {n}: rel L2 all columns {p[n]['rel_l2_all']:.3e}, {p[n]['nonzero_columns']} columns, "
              f"{p[n]['moved_columns']} moved")


def reference_sp_raw(gpu_tables, flags, rep, st, dt, itimestep, nlev=None):
    """the reference's own single-precision build (oracle/_ref, -DSINGLE_PREC)
    on the same FP32-representable inputs, per tendency: rel_l2 over the
    points where our FP32 step took the FP64 run's branches, against the
    FP64 output rounded to FP32 -- our rel_l2's measure on the same points
    -- and rel_l2_raw over all points against the unrounded FP64 output
    (ragged columns, nlev: each column run on its own levels, the points
    above its top left out); None without that build"""
    from oracle import ref
    from tempo_amd import branch_report as br
    if not ref.available("sp"):
        return None
    inp = {k: v.astype(np.float32).astype(np.float64) for k, v in st.items()}
    _, _, o64, _ = rep["outputs"]
    keep = rep["keep"]
    ref.init(gpu_tables, variant="sp", **flags)
    osp = {k: v.copy() for k, v in inp.items()}
    valid = np.ones(inp["p"].shape, bool)
    if nlev is None:
        ref.run(osp, dt, itimestep, variant="sp")
    else:
        valid = np.arange(inp["p"].shape[0])[:, None] < np.asarray(nlev)[None, :]
        for c in range(inp["p"].shape[1]):
            n = int(nlev[c])
            sub = {k: np.ascontiguousarray(v[:n, c:c + 1]) for k, v in osp.items()}
            ref.run(sub, dt, itimestep, variant="sp")
            for k, v in sub.items():
                osp[k][:n, c] = v[:, 0]
    out = {}
    for k in rep["fields"]:
        t_sp = (osp[k] - inp[k]) / dt
        t64r = (o64[k].astype(np.float32).astype(np.float64) - inp[k]) / dt
        out[k] = dict(rel_l2=br._rel(t_sp[keep], t64r[keep]),
                      rel_l2_raw=br._rel(t_sp[valid], ((o64[k] - inp[k]) / dt)[valid]))
    return out


def print_theirs(theirs):
    if theirs is not None:
        print("  reference FP32 build, same points, against FP64 rounded:",
              {k: f"{v['rel_l2']:.2e}" for k, v in theirs.items()})
        print("  reference FP32 build, all points, against unrounded FP64:",
              {k: f"{v['rel_l2_raw']:.2e}" for k, v in theirs.items()})


@pytest.mark.parametrize("dtype", ["fp32", "fp64"])
def test_gpu_signature_run_is_the_step(hip, big_case, dtype):
    """tempo_hip_branch_signature performs tempo_hip_run's step bit for bit,
    its signatures are deterministic, and a planted threshold crossing (cloud
    water just above and just below the r1 = 1e-12 presence threshold at one
    point) changes the signature of that point and no other point's"""
    from tempo_amd import branch_report as br
    st, m, _, _, _ = big_case
    sub = {k: np.ascontiguousarray(v[:, :256]) for k, v in st.items()}
    tdt = torch.float32 if dtype == "fp32" else torch.float64
    ctx = context(hip, dict(aerosolaware=True, hailaware=True))
    want, dw = gpu_run(ctx, sub, m["dt"], m["itimestep"], dtype=tdt)
    got, dg, sig, _ = br.signature_run(ctx, sub, m["dt"], m["itimestep"], tdt)
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    for k in dw:
        np.testing.assert_array_equal(dg[k], dw[k], err_msg=k)
    _, _, sig2, _ = br.signature_run(ctx, sub, m["dt"], m["itimestep"], tdt)
    np.testing.assert_array_equal(sig, sig2)
    assert (sig[:2] != 0).all()  # the level-local phases record every point; flux-form sedimentation none
    assert (sig[2] == 0).all()
    # a point with only cloud water, at the presence threshold
    k, c = 30, 17
    a = {kk: v.copy() for kk, v in sub.items()}
    for q in ("qc", "qr", "qi", "qs", "qg", "ni", "nr", "nc", "ng", "qb"):
        a[q][k, c] = 0.0
    b = {kk: v.copy() for kk, v in a.items()}
    a["qc"][k, c] = 1.5e-12
    b["qc"][k, c] = 0.5e-12
    _, _, sa, _ = br.signature_run(ctx, a, m["dt"], m["itimestep"], tdt)
    _, _, sb, _ = br.signature_run(ctx, b, m["dt"], m["itimestep"], tdt)
    diff = (sa != sb).any(axis=(0, 1))
    assert diff[k, c] and (sa[0, 1, k, c] != sb[0, 1, k, c])  # the presence flag is a discrete outcome
    assert diff.sum() == 1 or diff[:, c].sum() == diff.sum(), np.argwhere(diff)[:8].tolist()


@pytest.mark.parametrize("semi", [False, True])
def test_gpu_fp32_within_tolerance(hip, big_case, gpu_tables, semi):
    """FP32 production path vs the FP64 twin on the same FP32-representable
    inputs (2,048 aerosol- and hail-aware columns), flux-form and
    semi-Lagrangian sedimentation, with branch accounting.  The
    semi-Lagrangian surface flux has no low-mass cut-off, so a column whose
    lowest level sits at a presence threshold flips its surface flux with
    its branch; its precipitation is bounded over the columns that took the
    FP64 run's branches, and its flipped columns are reported."""
    from tempo_amd import branch_report as br
    st, m, _, _, _ = big_case
    ctx = context(hip, dict(aerosolaware=True, hailaware=True, semi_sedi=semi))
    rep = br.compare(ctx, st, m["dt"], m["itimestep"])
    flags = dict(aerosolaware=True, hailaware=True, semi_sedi=semi)
    theirs = reference_sp_raw(gpu_tables, flags, rep, st, m["dt"], m["itimestep"])
    report_line(f"FP32 vs FP64 (semi={semi}):", rep)
    print_theirs(theirs)
    check_fp32_report(rep, precip_all=not semi, theirs=theirs, semi=semi)
    assert rep["flipped_points"] <= 0.25, rep["flipped_points"]
    from oracle import ref
    if ref.available("sp"):
        # the reference's own single-precision build on the same inputs,
        # against the same FP64 run: points moved by more than 1e-3
        inp = {k: v.astype(np.float32).astype(np.float64) for k, v in st.items()}
        o32, _, o64, _ = rep["outputs"]
        ref.init(gpu_tables, variant="sp", aerosolaware=True, hailaware=True, semi_sedi=semi)
        osp = {k: v.copy() for k, v in inp.items()}
        ref.run(osp, m["dt"], m["itimestep"], variant="sp")
        ours, theirs = moved_fraction(inp, o32, o64), moved_fraction(inp, osp, o64)
        print("points moved > 1e-3, ours vs the reference FP32 build:", ours, theirs)
        for k in ours:
            assert ours[k] <= 2.0 * theirs[k] + 1e-4, (k, ours[k], theirs[k])


def test_gpu_fp32_uneven_columns(hip, gpu_tables):
    """FP32 against the FP64 twin on 2,048 columns of the uneven-column
    workload (30-128 active levels, hail-aware, convective columns,
    stretched layers), with branch accounting over each column's active
    levels: every tendency within 1e-4 relative L2 over the branch-agreeing
    points, or within twice the reference FP32 build's error on the same
    points (each column run on its own levels there)"""
    from tempo_amd import branch_report as br
    from tempo_amd import generator as gen
    st, m = gen.generate("ragged", ncol=2048, col0=7777)
    flags = dict(aerosolaware=False, hailaware=True)
    ctx = context(hip, flags)
    rep = br.compare(ctx, st, m["dt"], m["itimestep"], nlev=m["nlev"])
    theirs = reference_sp_raw(gpu_tables, flags, rep, st, m["dt"], m["itimestep"], nlev=m["nlev"])
    report_line("FP32 vs FP64, uneven columns:", rep)
    print_theirs(theirs)
    check_fp32_report(rep, theirs=theirs)
    assert rep["flipped_points"] <= 0.15, rep["flipped_points"]


@pytest.mark.parametrize("workload,aero", [("aerosol", True), ("plain", False)])
def test_gpu_fp32_bench_workload(hip, gpu_tables, workload, aero):
    """FP32 against the FP64 twin on the benchmark's FP32 report sample (the
    first 8,192 columns of its aerosol-aware workload) and on 8,192 columns
    of the one-moment-cloud workload: the seeded
    step (each tendency within 1e-4 relative L2 over the points that took the
    same branches), and the step after eight FP64 steps, when rain has
    reached the ground (surface precipitation within 1e-5 over all columns).
    Near saturation after those steps, qv - qvs cancels in the FP32 inputs
    themselves; there the tendencies are held to the reference's own
    single-precision build's error on the same state (or 1e-4)."""
    from tempo_amd import branch_report as br
    from tempo_amd import generator as gen
    st, m = gen.generate(workload, ncol=8192)
    flags = dict(aerosolaware=aero, hailaware=False)
    ctx = context(hip, flags)
    rep = br.compare(ctx, st, m["dt"], m["itimestep"])
    theirs = reference_sp_raw(gpu_tables, flags, rep, st, m["dt"], m["itimestep"])
    report_line(f"{workload}, seeded step:", rep)
    print_theirs(theirs)
    check_fp32_report(rep, theirs=theirs)
    assert rep["flipped_points"] <= 0.25, rep["flipped_points"]
    st, _ = gpu_run(ctx, st, m["dt"], m["itimestep"], steps=8)
    rep = br.compare(ctx, st, m["dt"], m["itimestep"] + 8)
    theirs = reference_sp_raw(gpu_tables, flags, rep, st, m["dt"], m["itimestep"] + 8)
    report_line(f"{workload} after 8 steps:", rep)
    print_theirs(theirs)
    assert rep["precip"]["nonzero_columns"] > 0
    check_fp32_report(rep, theirs=theirs)


@pytest.mark.parametrize("semi", [False, True])
def test_gpu_fp32_cold_surface(hip, gpu_tables, semi):
    """FP32 against the FP64 twin where every surface diagnostic is live:
    1,024 columns with a below-freezing surface and cloud down to it (rain,
    freezing rain, ice, snow and graupel reaching the ground)."""
    from tempo_amd import branch_report as br
    from tempo_amd import generator as gen
    st, m = gen.generate("cold", ncol=1024)
    ctx = context(hip, dict(aerosolaware=True, hailaware=True, semi_sedi=semi))
    rep = br.compare(ctx, st, m["dt"], m["itimestep"])
    flags = dict(aerosolaware=True, hailaware=True, semi_sedi=semi)
    theirs = reference_sp_raw(gpu_tables, flags, rep, st, m["dt"], m["itimestep"])
    report_line(f"cold surface (semi={semi}):", rep)
    print_theirs(theirs)
    p = rep["precip"]
    for n in br.PRECIP + ("frz_rain_precip",):
        assert p[n]["nonzero_columns"] > 0, n
        # each kind on its own, over all columns
        assert p[n]["rel_l2_all"] <= 1e-4, (n, p[n])
    check_fp32_report(rep, precip_all=not semi, theirs=theirs, semi=semi)


def ml_columns(ncol, col0, with_bl):
    """generated columns for the cloud-number network: one-moment cloud (no
    nc) or two-moment cloud with boundary-layer cloud water and fraction"""
    from tempo_amd import generator as gen
    if not with_bl:
        return gen.generate("plain", ncol=ncol, col0=col0)
    st, m = gen.generate("aerosol", ncol=ncol, col0=col0)
    rng = np.random.default_rng(col0)
    nz = st["p"].shape[0]
    qbl = np.where(rng.random((nz, ncol)) < 0.4, rng.uniform(1e-8, 1e-3, (nz, ncol)), 0.0)
    st["qc_bl"] = np.ascontiguousarray(qbl)
    st["qcfrac_bl"] = np.ascontiguousarray(np.where(qbl > 0.0, rng.uniform(0.05, 1.0, (nz, ncol)), 0.0))
    return st, m


ML_CASES = {"ml_for_nc": (False, dict(aerosolaware=False, hailaware=False, nc_mixing_ratio=True, ml_for_nc=True)),
            "ml_for_bl_nc": (True, dict(aerosolaware=True, hailaware=False, nc_mixing_ratio=True,
                                        ml_for_bl_nc=True))}


@pytest.mark.parametrize("which", sorted(ML_CASES))
def test_gpu_ml_cloud_number_matches_oracle(hip, gpu_tables, which):
    """Cloud number from the trained network (src/module_mp_tempo_ml.F90):
    the one-moment cloud's number at every level (ml_for_nc) and the
    boundary-layer cloud in the cloud radius (ml_for_bl_nc with qc_bl,
    qcfrac_bl), FP64 path vs the oracle on 512 columns, two steps; and the
    FP32 path vs the FP64 twin on the cloud diagnostics the network feeds."""
    with_bl, flags = ML_CASES[which]
    st, m = ml_columns(512, 4000, with_bl)
    ref, d2, d3 = reference_steps(gpu_tables, flags, st, m["dt"], m["itimestep"], steps=2)
    ctx = context(hip, flags)
    out, d = gpu_run(ctx, st, m["dt"], m["itimestep"], steps=2)
    for k in ref:
        assert rel_inf(out[k], ref[k]) <= 1e-12, (k, rel_inf(out[k], ref[k]))
    for k, v in d3.items():
        assert rel_inf(d[D3_MAP[k]], v) <= 1e-12, (k, rel_inf(d[D3_MAP[k]], v))
    assert (d["re_cloud"] > 0).sum() > 100
    # FP32: same FP32-representable inputs, one step, vs the FP64 twin.  Where
    # cloud freezes completely, the new cloud water is a rounding residual
    # near the r1 = 1e-12 presence threshold in single precision (the
    # reference's own -DSINGLE_PREC build leaves such residuals too), and the
    # cloud radius follows that presence; radius flips at those points are
    # the state's and are counted apart.
    inp = {k: v.astype(np.float32).astype(np.float64) for k, v in st.items()}
    o64, e64 = gpu_run(ctx, inp, m["dt"], m["itimestep"], dtype=torch.float64)
    o32, e32 = gpu_run(ctx, inp, m["dt"], m["itimestep"], dtype=torch.float32)
    qc_flip = (o32["qc"] > 1e-12) != (o64["qc"] > 1e-12)
    for k in ("re_cloud", "cloud_number_mixing_ratio"):
        flip = flipped(e32[k], e64[k])
        keep = ~flip
        err = np.linalg.norm((e32[k] - e64[k])[keep]) / max(np.linalg.norm(e64[k][keep]), 1e-300)
        other = float((flip & ~qc_flip).mean())
        print(which, k, "FP32 rel L2", err, "flipped", float(flip.mean()), "of which not at a cloud-presence flip",
              other, "cloud-presence flips", float(qc_flip.mean()))
        assert err <= 1e-4, (k, err)
        assert other <= 1e-3, (k, other)


def test_gpu_melting_reflectivity_matches_oracle(hip, gpu_tables):
    """Reflectivity of melting snow and graupel below the melting level
    (refl10cm_from_melting_flag, src/module_mp_tempo_diags.F90:59-395) on 512
    hail- and aerosol-aware columns: FP64 path vs the oracle; FP32 path vs the
    FP64 twin (dBZ)."""
    from tempo_amd import generator as gen
    st, m = gen.generate("multistep", ncol=512, col0=91_000)
    flags = dict(aerosolaware=True, hailaware=True, refl_from_melting=True)
    ref, d2, d3 = reference_steps(gpu_tables, flags, st, m["dt"], m["itimestep"])
    ctx = context(hip, flags)
    out, d = gpu_run(ctx, st, m["dt"], m["itimestep"])
    for k in ref:
        assert rel_inf(out[k], ref[k]) <= 1e-12, k
    assert rel_inf(d["refl10cm"], d3["refl10cm"]) <= 1e-12, rel_inf(d["refl10cm"], d3["refl10cm"])
    # the melting branch is reached
    _, _, e3 = reference_steps(gpu_tables, dict(flags, refl_from_melting=False), st, m["dt"], m["itimestep"])
    assert (np.abs(e3["refl10cm"] - d3["refl10cm"]) > 0.01).sum() > 100
    inp = {k: v.astype(np.float32).astype(np.float64) for k, v in st.items()}
    o64, e64 = gpu_run(ctx, inp, m["dt"], m["itimestep"], dtype=torch.float64)
    o32, e32 = gpu_run(ctx, inp, m["dt"], m["itimestep"], dtype=torch.float32)
    flip = np.abs(e32["refl10cm"] - e64["refl10cm"]) > 0.01
    keep = ~flip
    err = np.linalg.norm((e32["refl10cm"] - e64["refl10cm"])[keep]) / np.linalg.norm(e64["refl10cm"][keep])
    print("melting refl10cm FP32 rel L2", err, "points moved by more than 0.01 dBZ", flip.mean())
    assert err <= 1e-4, err
    assert flip.mean() <= 0.01, flip.mean()


CHECK_CASES = (("cloud2", "cloud", ("ncsave",)), ("cloud_ncsave", "cloud", ("nc1d",)), ("cloud_land", "cloud", ("nc1d", "ncsave")),
               ("ice", "ice", ()), ("snow", "snow", ()))


@pytest.mark.parametrize("case,which,absent", CHECK_CASES)
def test_gpu_check_and_update_matches_reference_fixture(hip, case, which, absent):
    """The public cloud/ice/snow_check_and_update entry points
    (src/module_mp_tempo_main.F90:18-20) on 4,000 points spanning their
    branches, against the reference routines' outputs: FP64 device and host
    arrays to 1e-12, FP32 to 1e-5 relative with identical presence flags."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "check_and_update.npz"))
    ctx = context(hip, dict(aerosolaware=True, hailaware=True))
    names = hip.TempoHip.CHECK_ARGS[which]
    inp = {k: z[f"in__{k}"] for k in names}
    for dtype in (torch.float64, torch.float32, "host"):
        if dtype == "host":
            a = {k: (None if k in absent else np.ascontiguousarray(v.copy())) for k, v in inp.items()}
        else:
            a = {k: (None if k in absent else torch.tensor(v, device="cuda",
                                                        dtype=torch.int32 if k.startswith("l_") else dtype))
                 for k, v in inp.items()}
        ctx.check_and_update(which, a, float(z["dt"]))
        torch.cuda.synchronize()
        out = {k: (a[k] if dtype == "host" else a[k].double().cpu().numpy()) for k in names if a[k] is not None}
        lk = names[1]
        # FP32: masses within rounding of the r1 presence threshold may decide
        # differently; values are compared where the decision agrees
        same = out[lk] == z[f"{case}__{lk}"]
        if dtype == torch.float32:
            assert same.mean() >= 0.995, (case, float(same.mean()))
        else:
            assert same.all(), (case, str(dtype))
        for k in out:
            if k in ("rho", "ncsave") or k.startswith("l_"):
                continue
            got, want = out[k][same], z[f"{case}__{k}"][same]
            tol = 1e-5 if dtype == torch.float32 else 1e-12
            err = np.abs(got - want) / np.maximum(np.abs(want), 1e-300)
            assert err.max() <= tol or rel_inf(got, want) <= tol, (case, k, str(dtype), float(err.max()))


def test_gpu_run_cfg_changes_between_steps(hip):
    """Flags changed after init (tempo_hip_set_run_cfg) act like a context
    initialised with them, as a host editing ty_tempo_cfgs between calls."""
    flags = dict(aerosolaware=False, hailaware=False)
    st, m = ml_columns(64, 900, False)
    want, dw = gpu_run(context(hip, dict(flags, ml_for_nc=True, nc_mixing_ratio=True)), st, m["dt"], m["itimestep"])
    ctx = hip.TempoHip(aerosolaware_flag=0, hailaware_flag=0)
    ctx.set_run_cfg(ml_for_nc_flag=1, cloud_number_mixing_ratio_flag=1)
    got, dg = gpu_run(ctx, st, m["dt"], m["itimestep"])
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    np.testing.assert_array_equal(dg["cloud_number_mixing_ratio"], dw["cloud_number_mixing_ratio"])
    with pytest.raises(hip.TempoError) as e:
        ctx.set_run_cfg(hailaware_flag=1)
    assert e.value.code == hip.E_ARG
    ctx.close()


def test_gpu_ragged_matches_reference(hip, gpu_tables):
    """512 columns of 30-128 active levels: the FP64 twin vs the reference
    build run on each column's own levels, and vs the oracle's ragged step"""
    from oracle import oracle as orc
    from tempo_amd import generator as gen
    st, m = gen.generate("ragged", ncol=512, col0=1000)
    flags = dict(aerosolaware=False, hailaware=True)
    ref, _, _ = reference_steps(gpu_tables, flags, st, m["dt"], m["itimestep"], nlev=m["nlev"])
    orc.init(gpu_tables, **flags)
    o = {k: v.copy() for k, v in st.items()}
    orc.run(o, m["dt"], m["itimestep"], nlev=m["nlev"])
    for k in ref:
        np.testing.assert_array_equal(o[k], ref[k], err_msg=k)
    ctx = context(hip, flags)
    out, _ = gpu_run(ctx, st, m["dt"], m["itimestep"], nlev=m["nlev"])
    for k in ref:
        assert rel_inf(out[k], ref[k]) <= 1e-12, k
    # levels above each column's top are untouched
    for k in ("qr", "t"):
        for c in range(0, 512, 37):
            n = m["nlev"][c]
            np.testing.assert_array_equal(out[k][n:, c], st[k][n:, c])


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_gpu_column_order_does_not_change_results(hip, dtype):
    """Every column is computed on its own: the uneven-column batch stored by
    decreasing active depth (the order a host may keep its columns in)
    gives each column the bits it gets in the generator's order"""
    from tempo_amd import generator as gen
    st, m = gen.generate("ragged", ncol=2048, col0=500)
    nlev = np.asarray(m["nlev"])
    order = np.argsort(-nlev, kind="stable")
    ctx = context(hip, dict(aerosolaware=False, hailaware=True))
    tdt = getattr(torch, dtype)
    a, da = gpu_run(ctx, st, m["dt"], m["itimestep"], nlev=nlev, dtype=tdt)
    b, db = gpu_run(ctx, {k: np.ascontiguousarray(v[:, order]) for k, v in st.items()}, m["dt"], m["itimestep"],
                    nlev=np.ascontiguousarray(nlev[order]), dtype=tdt)
    for k in a:
        np.testing.assert_array_equal(b[k], a[k][:, order], err_msg=k)
    for k in da:
        np.testing.assert_array_equal(db[k], da[k][..., order], err_msg=k)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gpu_host_path_equals_device_path(hip, prec):
    """tempo_hip_run_host(_prec) on host arrays gives the device path's step
    bit for bit, in FP64 and in FP32 (FP32-representable inputs), over
    transfers larger than one pinned chunk (100,000 x 90 columns: 72 MB per
    field)"""
    from tempo_amd import generator as gen
    st, m = gen.generate("aerosol", ncol=100_000, col0=77)
    tdt = torch.float64 if prec == "fp64" else torch.float32
    if prec == "fp32":
        st = {k: v.astype(np.float32).astype(np.float64) for k, v in st.items()}
    ctx = context(hip, dict(aerosolaware=True, hailaware=False))
    dev, dd = gpu_run(ctx, st, m["dt"], m["itimestep"], dtype=tdt)
    hst = {k: v.copy() for k, v in st.items()}
    hd = ctx.run_host(hst, m["dt"], m["itimestep"], precision=hip.FP64 if prec == "fp64" else hip.FP32)
    for k in st:
        np.testing.assert_array_equal(hst[k], dev[k], err_msg=k)
    for k in dd:
        np.testing.assert_array_equal(hd[k], dd[k], err_msg=k)


def _host_call(ctx, hip, st, m, diag, dims, nlev=None, precision=None, land=None):
    """tempo_hip_run_host through ctypes on numpy arrays with the given diag
    arrays (name -> float64 array), geometry and land mask (int32 per column)"""
    import ctypes
    f = hip.Fields()
    for name in hip.FIELD_NAMES:
        if name in st:
            setattr(f, name, st[name].ctypes.data)
    if land is not None:
        f.land = land.ctypes.data
    d = hip.Diags()
    for n, a in diag.items():
        setattr(d, n, a.ctypes.data)
    if precision is None:
        hip._check(hip.lib().tempo_hip_run_host(ctx._ctx, ctypes.byref(f), ctypes.byref(d), ctypes.byref(dims),
                                                float(m["dt"]), int(m["itimestep"])))
    else:
        hip._check(hip.lib().tempo_hip_run_host_prec(ctx._ctx, ctypes.byref(f), ctypes.byref(d), ctypes.byref(dims),
                                                     float(m["dt"]), int(m["itimestep"]), int(precision)))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("layout", ["j_one_slab", "j_slabs", "i_slabs", "j_slabs_locked", "i_slabs_locked"])
def test_gpu_host_path_padded_diagnostics_and_ragged_columns(hip, prec, layout):
    """tempo_hip_run_host(_prec) with padded rows (state ldk > ni; diagnostics
    dldk, d2ldj > ni, as a Fortran host's halo'd tile), ragged columns
    (nlev) and a land mask, stepped in one slab, in slabs of j rows (32 x 3 tile) and in
    slabs of i ranges (1024 x 1 tile), with pageable arrays and with the
    state in page-locked memory (moved by the copy engines directly; the
    diagnostics pageable, or also page-locked in the i-slab case): tile
    points equal the device path's, and every point the step does not write
    -- padding between rows, levels k >= nlev[c] -- keeps the caller's value
    (in FP32 too: those points are not converted)"""
    from tempo_amd import generator as gen
    ni, nj, spad, slabs = dict(j_one_slab=(32, 3, 0, 1), j_slabs=(32, 3, 2, 3), i_slabs=(1024, 1, 3, 3),
                               j_slabs_locked=(32, 3, 2, 3), i_slabs_locked=(1024, 1, 3, 3))[layout]
    locked = layout.endswith("locked")

    def alloc(shape):
        return hip.pinned_empty(shape) if locked else np.empty(shape)
    st, m = gen.generate("ragged", ncol=ni * nj, col0=123)
    if prec == "fp32":
        st = {k: v.astype(np.float32).astype(np.float64) for k, v in st.items()}
    nz, ncol = st["p"].shape
    nlev = m["nlev"]
    ctx = context(hip, dict(aerosolaware=False, hailaware=True))
    # a land mask (the one-moment cloud number of these columns depends on it)
    land = np.ascontiguousarray((np.arange(ncol) % 3 == 0).astype(np.int32))
    dev, dd = gpu_run(ctx, st, m["dt"], m["itimestep"], nlev=nlev, land=land,
                      dtype=torch.float64 if prec == "fp64" else torch.float32)
    pad = 5
    sentinel = -7.25
    # state in the Fortran (i, k, j) layout of an (ni + spad, nz, nj) array,
    # column c = i + j*ni; diagnostics with rows padded by `pad`
    ldk = ni + spad
    def fort(a):
        out = alloc((nj, nz, ldk))
        out[...] = sentinel
        out[:, :, :ni] = a.reshape(nz, nj, ni).transpose(1, 0, 2)
        return out
    hst = {k: fort(v) for k, v in st.items()}
    dalloc = alloc if layout == "i_slabs_locked" else np.empty
    ld2 = ni + pad
    d2 = {n: dalloc(ld2 * nj) for n in hip.DIAG2_NAMES}
    ldk3 = ni + pad
    d3 = {n: dalloc(ldk3 * nz * nj) for n in hip.DIAG3_NAMES}
    for a in list(d2.values()) + list(d3.values()):
        a[...] = sentinel
    nl = np.ascontiguousarray(nlev, dtype=np.int32)
    dims = hip.Dims(ni, nj, nz, ldk, ldk * nz, ldk3, ldk3 * nz, ld2, nl.ctypes.data)
    p = hip.FP64 if prec == "fp64" else hip.FP32
    ctx.tune(p, hip.TUNE_HOST_SLABS, slabs)
    try:
        _host_call(ctx, hip, hst, m, dict(d2, **d3), dims, precision=None if prec == "fp64" else hip.FP32,
                   land=land)
    finally:
        ctx.tune(p, hip.TUNE_HOST_SLABS, 0)
    for k in st:
        np.testing.assert_array_equal(hst[k][:, :, :ni].transpose(1, 0, 2).reshape(nz, ncol), dev[k], err_msg=k)
        assert (hst[k][:, :, ni:] == sentinel).all(), k
    i = np.arange(ncol) % ni
    j = np.arange(ncol) // ni
    for n in hip.DIAG2_NAMES:
        got = d2[n].reshape(nj, ld2)
        np.testing.assert_array_equal(got[j, i], dd[n], err_msg=n)
        assert (got[:, ni:] == sentinel).all(), n
    for n in hip.DIAG3_NAMES:
        got = d3[n].reshape(nj, nz, ldk3)
        assert (got[:, :, ni:] == sentinel).all(), n
        cols = got.transpose(1, 0, 2)[:, :, :ni].reshape(nz, ncol)
        above = np.arange(nz)[:, None] >= np.asarray(nlev)[None, :]
        np.testing.assert_array_equal(np.where(above, 0.0, cols), np.where(above, 0.0, dd[n]), err_msg=n)
        assert (cols[above] == sentinel).all(), n
    if slabs > 1:  # no substep record after a step in slabs
        with pytest.raises(hip.TempoError, match="no substep record"):
            ctx.substeps(ncol)


def test_gpu_host_entry_points_share_one_context(hip):
    """the host-array entry points in growing order of their staging needs
    (surface emissions, a check routine, the step) on one context give what
    fresh contexts give"""
    from tempo_amd import generator as gen
    st, m = gen.generate("aerosol", ncol=64, col0=4242)
    nz, ncol = st["p"].shape

    def sequence(ctx_a, ctx_b, ctx_c):
        nwfa = st["nwfa"].copy()
        nwfa2d = np.linspace(0.0, 1e4, ncol)
        ctx_a.surface_emissions_host(nwfa, nwfa2d, 20.0, k=0)
        rng = np.random.default_rng(3)
        arr = dict(rho=rng.uniform(0.5, 1.2, 500), l_qi=np.zeros(500, np.int32), qi1d=rng.uniform(0, 1e-4, 500),
                   ni1d=rng.uniform(0, 1e5, 500), ri=np.zeros(500), ni=np.zeros(500), qiten=np.zeros(500),
                   niten=np.zeros(500), ilami=np.zeros(500))
        ctx_b.check_and_update("ice", arr, 20.0)
        hst = {k: v.copy() for k, v in st.items()}
        hst["nwfa"] = nwfa
        d = ctx_c.run_host(hst, m["dt"], m["itimestep"])
        return nwfa, arr, hst, d

    shared = hip.TempoHip(aerosolaware_flag=1, hailaware_flag=0)
    got = sequence(shared, shared, shared)
    fresh = [hip.TempoHip(aerosolaware_flag=1, hailaware_flag=0) for _ in range(3)]
    want = sequence(*fresh)
    np.testing.assert_array_equal(got[0], want[0])
    for k in got[1]:
        np.testing.assert_array_equal(got[1][k], want[1][k], err_msg=k)
    for k in got[2]:
        np.testing.assert_array_equal(got[2][k], want[2][k], err_msg=k)
    for k in got[3]:
        np.testing.assert_array_equal(got[3][k], want[3][k], err_msg=k)
    # and again on the same context, now holding the largest staging
    again = sequence(shared, shared, shared)
    for k in got[2]:
        np.testing.assert_array_equal(again[2][k], want[2][k], err_msg=k)
    for c in [shared] + fresh:
        c.close()


def test_gpu_single_column_and_tiny_grid(hip, gpu_tables):
    """one column of the plain workload, and its lowest two levels: the FP64
    twin vs the reference build"""
    from tempo_amd import generator as gen
    flags = dict(aerosolaware=False, hailaware=False)
    st, m = gen.generate("plain", ncol=1, col0=3)
    ctx = context(hip, flags)
    for inp in (st, {k: np.ascontiguousarray(v[:2]) for k, v in st.items()}):
        out, d = gpu_run(ctx, inp, m["dt"], m["itimestep"])
        ref, d2, d3 = reference_steps(gpu_tables, flags, inp, m["dt"], m["itimestep"])
        check_against_reference(out, d, ref, d2, d3, inp, f"{inp['p'].shape[0]} levels x 1 column")


def test_gpu_fp64_matches_reference_above_128_levels(hip, gpu_tables):
    """64 columns of the aerosol workload interpolated onto 160 levels (layer
    depths scaled so that each column keeps its height): the sedimentation
    runs one lane per column there (the wavefront kernel holds at most 128
    levels); the FP64 twin vs the reference build"""
    from tempo_amd import generator as gen
    flags = dict(aerosolaware=True, hailaware=False)
    st, m = gen.generate("aerosol", ncol=64, col0=1234)
    nz0, nz = st["p"].shape[0], 160
    x_old, x_new = np.arange(nz0, dtype=np.float64), np.linspace(0.0, nz0 - 1.0, nz)
    deep = {k: np.ascontiguousarray(np.stack([np.interp(x_new, x_old, v[:, c]) for c in range(v.shape[1])], axis=1))
            for k, v in st.items()}
    deep["dz"] = deep["dz"] * (nz0 / nz)
    ctx = context(hip, flags)
    out, d = gpu_run(ctx, deep, m["dt"], m["itimestep"])
    ref, d2, d3 = reference_steps(gpu_tables, flags, deep, m["dt"], m["itimestep"])
    check_against_reference(out, d, ref, d2, d3, deep, f"{nz} levels x 64 columns")


def test_gpu_empty_batch_is_a_no_op(hip):
    """zero columns: the step entry points (device arrays in both precisions,
    host arrays, tangent-linear, adjoint) and the surface emissions return
    without error and touch nothing, as the reference's loops over an empty
    tile (src/module_mp_tempo_driver.F90:400-449)"""
    ctx = context(hip, dict(aerosolaware=False, hailaware=False))
    nz = 30
    names = ("t", "p", "w", "dz", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr")
    for dtype in (torch.float32, torch.float64):
        print("empty batch, device arrays,", dtype, flush=True)
        dev = {k: torch.empty((nz, 0), dtype=dtype, device="cuda") for k in names}
        d = ctx.run(dev, 20.0, 2)
        assert all(v.numel() == 0 for v in d.values())
    dev = {k: torch.empty((nz, 0), dtype=torch.float64, device="cuda") for k in names}
    print("empty batch, tangent-linear and adjoint", flush=True)
    ctx.tlm(dev, {k: torch.empty_like(v) for k, v in dev.items()}, 20.0, 2)
    xbar = ctx.adjoint(dev, {}, {}, 20.0, 2)
    assert all(v.numel() == 0 for v in xbar.values())
    print("empty batch, host arrays and emissions", flush=True)
    ctx.run_host({k: np.empty((nz, 0)) for k in names}, 20.0, 2)
    ctx.surface_emissions(torch.empty((nz, 0), dtype=torch.float64, device="cuda"),
                          torch.empty((0,), dtype=torch.float64, device="cuda"), 20.0)
    ctx.surface_emissions_host(np.empty((nz, 0)), np.empty((0,)), 20.0)
    torch.cuda.synchronize()


def test_gpu_error_codes(hip):
    from tempo_amd import generator as gen
    ctx = context(hip, dict(aerosolaware=False, hailaware=False))
    st, m = gen.generate("plain", ncol=4)
    dev = to_dev(st, torch.float64)
    bad = dict(dev)
    del bad["qv"]
    with pytest.raises(hip.TempoError) as e:
        ctx.run(bad, m["dt"], 2)
    assert e.value.code == hip.E_ARG
    with pytest.raises(hip.TempoError) as e:
        ctx.run(dev, -1.0, 2)
    assert e.value.code == hip.E_ARG
    # the adjoint covers the flux-form sedimentation only (the tangent-linear
    # step both forms)
    semi = hip.TempoHip(aerosolaware_flag=0, hailaware_flag=0, semi_sedi_flag=1)
    with pytest.raises(hip.TempoError) as e:
        semi.adjoint(dict(dev), {}, {}, m["dt"], 2)
    assert e.value.code == hip.E_UNSUPPORTED
    semi.close()
    # boundary-layer cloud without the network, in the step and its derivatives
    bl = dict(dev, qc_bl=torch.zeros_like(dev["qc"]), qcfrac_bl=torch.zeros_like(dev["qc"]))
    with pytest.raises(hip.TempoError) as e:
        ctx.run(bl, m["dt"], 2)
    assert e.value.code == hip.E_ARG
    with pytest.raises(hip.TempoError) as e:
        ctx.tlm(dict(bl), {k: torch.zeros_like(v) for k, v in dev.items()}, m["dt"], 2)
    assert e.value.code == hip.E_ARG
    with pytest.raises(hip.TempoError) as e:
        ctx.adjoint(dict(bl), {}, {}, m["dt"], 2)
    assert e.value.code == hip.E_ARG
    # the melting-layer reflectivity (binned integrals) has no derivative here
    other = hip.TempoHip(aerosolaware_flag=0, hailaware_flag=0, refl10cm_from_melting_flag=1)
    with pytest.raises(hip.TempoError) as e:
        other.tlm(dict(dev), {k: torch.zeros_like(v) for k, v in dev.items()}, m["dt"], 2)
    assert e.value.code == hip.E_UNSUPPORTED
    other.close()
    with tempfile.TemporaryDirectory() as d:
        with pytest.raises(hip.TempoError) as e:
            hip.TempoHip(table_dir=d)
        assert e.value.code == hip.E_TABLE


def test_gpu_table_file_roundtrip(hip):
    """GPU-built tables written in the reference file format and read back
    give the identical step."""
    from tempo_amd import generator as gen
    from tempo_amd.host import CCN_PATH
    built = context(hip, dict(aerosolaware=True, hailaware=True))
    st, m = gen.generate("multistep", ncol=256, col0=9)
    a, _ = gpu_run(built, st, m["dt"], m["itimestep"])
    with tempfile.TemporaryDirectory() as d:
        built.write_tables(d)
        sizes = {f: os.path.getsize(os.path.join(d, f)) for f in os.listdir(d)}
        assert sizes == {"freeze_water_data_tempo_v3": 254944800, "qr_acr_qs_data_tempo_v3": 43764192,
                         "qr_acr_qg_data_tempo_v3": 674697960}
        read = hip.TempoHip(table_dir=d, ccn_path=CCN_PATH, aerosolaware_flag=1, hailaware_flag=1)
        b, _ = gpu_run(read, st, m["dt"], m["itimestep"])
        read.close()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_gpu_launch_shapes_give_identical_results(hip, dtype):
    """Every tuning setting (sedimentation kernel, block sizes, point order,
    phase A's register budget, the CCN table in LDS) produces the same bits:
    the wavefront-per-column sedimentation performs
    the serial kernel's arithmetic level by level, and the point order only
    regroups independent points into wavefronts.  The 64-level case fills
    exactly one staging pass of the sedimentation kernel's blocks; the check
    form of its staging (sed_wave 2) also issues the second pass there, every
    lane loading the group's last element and writing nothing."""
    from tempo_amd import generator as gen
    tdt = getattr(torch, dtype)
    prec = hip.FP32 if dtype == "float32" else hip.FP64
    for wl, ncol, flags in (("ragged", 300, dict(aerosolaware=False, hailaware=True)),
                            ("multistep", 257, dict(aerosolaware=True, hailaware=True)),
                            ("adjoint", 256, dict(aerosolaware=True, hailaware=False))):
        st, m = gen.generate(wl, ncol=ncol, col0=123)
        ctx = context(hip, flags)
        nl = None if "nlev" not in m else torch.tensor(m["nlev"], dtype=torch.int32, device="cuda")
        results = []
        for sed, ba, bc, order, wa, cl in ((0, 256, 256, 0, 0, 0), (1, 256, 256, 0, 0, 0), (1, 64, 128, 0, 1, 0),
                                           (1, 128, 64, 1, 0, 1), (1, 256, 256, 3, 1, 1), (0, 64, 64, 2, 0, 0),
                                           (1, 128, 256, 2, 1, 0), (2, 256, 256, 0, 0, 1)):
            ctx.tune(prec, hip.TUNE_SED_WAVE, sed)
            ctx.tune(prec, hip.TUNE_BLOCK_A, ba)
            ctx.tune(prec, hip.TUNE_BLOCK_C, bc)
            ctx.tune(prec, hip.TUNE_ORDER, order)
            ctx.tune(prec, hip.TUNE_WAVES_A, wa)
            ctx.tune(prec, hip.TUNE_CCN_LDS, cl)
            dev = {k: torch.tensor(v, dtype=tdt, device="cuda") for k, v in st.items()}
            d = ctx.run(dev, m["dt"], m["itimestep"], nlev=nl)
            torch.cuda.synchronize()
            sub = ctx.substeps(ncol)
            results.append(({k: v.cpu().numpy() for k, v in dev.items()},
                            {k: v.cpu().numpy() for k, v in d.items()}, sub))
        ctx.tune(prec, hip.TUNE_SED_WAVE, 1)
        ctx.tune(prec, hip.TUNE_BLOCK_A, 256)
        ctx.tune(prec, hip.TUNE_BLOCK_C, 256)
        ctx.tune(prec, hip.TUNE_ORDER, 0)
        ctx.tune(prec, hip.TUNE_WAVES_A, 0)
        ctx.tune(prec, hip.TUNE_CCN_LDS, 0)
        base = results[0]
        for r in results[1:]:
            for k in base[0]:
                np.testing.assert_array_equal(r[0][k], base[0][k], err_msg=f"{wl} {k}")
            for k in base[1]:
                np.testing.assert_array_equal(r[1][k], base[1][k], err_msg=f"{wl} {k}")
            np.testing.assert_array_equal(r[2], base[2], err_msg=f"{wl} substeps")
        print(wl, dtype, "max substeps", base[2][0::2].max(), "identical over", len(results), "settings")


@pytest.mark.skip(reason="the fused step is switched off: this test stopped finishing (killed after 180 s without "
                         "output, FP32 aerosol case) once the final update's tendency loads changed its code "
                         "generation; the cause is not found")
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("workload", ["aerosol", "ragged", "multistep"])
def test_gpu_fused_step_equals_four_launches(hip, workload, dtype):
    """The persistent work-queued launch (k_step_fused) gives the four-launch
    step's state, diagnostics and substep counts bit for bit, on
    uniform, ragged and hail- and aerosol-aware columns"""
    from tempo_amd import generator as gen
    st, m = gen.generate(workload, ncol=1500, col0=4000)
    flags = dict(aerosolaware_flag=int(m["aerosol"]), hailaware_flag=int(m["hail"]))
    prec = hip.FP32 if dtype == torch.float32 else hip.FP64
    nl = None if "nlev" not in m else torch.tensor(m["nlev"], dtype=torch.int32, device="cuda")
    out = []
    for fused in (0, 1):
        ctx = hip.TempoHip(**flags)
        ctx.tune(prec, hip.TUNE_FUSED, fused)
        dev = to_dev(st, dtype)
        for i in range(2):
            d = ctx.run(dev, m["dt"], m["itimestep"] + i, nlev=nl)
        torch.cuda.synchronize()
        out.append((dev, d, ctx.substeps(st["p"].shape[1])))
        ctx.close()
    (a, da, ka), (b, db, kb) = out
    for k in a:
        assert torch.equal(a[k], b[k]), k
    for k in da:
        assert torch.equal(da[k], db[k]), k
    np.testing.assert_array_equal(ka, kb)


def test_gpu_tune_rejects_bad_values(hip):
    ctx = context(hip, dict(aerosolaware=False, hailaware=False))
    for key, val in ((hip.TUNE_SED_WAVE, 3), (hip.TUNE_BLOCK_A, 96), (hip.TUNE_BLOCK_C, 512), (hip.TUNE_ORDER, 4),
                     (hip.TUNE_HOST_SLABS, -1), (hip.TUNE_HOST_THREADS, 0), (hip.TUNE_HOST_THREADS, 65),
                     (hip.TUNE_WAVES_A, 2), (hip.TUNE_CCN_LDS, -1), (99, 1)):
        with pytest.raises(hip.TempoError) as e:
            ctx.tune(hip.FP32, key, val)
        assert e.value.code == hip.E_ARG
    with pytest.raises(hip.TempoError) as e:  # the fused step is switched off
        ctx.tune(hip.FP32, hip.TUNE_FUSED, 1)
    assert e.value.code == hip.E_UNSUPPORTED


def test_gpu_surface_emissions_matches_reference_fixture(hip):
    """tempo_hip_aerosol_surface_emissions (device, FP64 and FP32) and the
    host-array entry point against the reference's output
    (src/module_mp_tempo_driver.F90:516-529).  The fixture's nwfa is
    Fortran (ni, nk, nj); the C ABI takes it through ldk = ni, ldj = ni*nk."""
    import ctypes
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "surface_emissions.npz"))
    nj, nk, ni = z["nwfa"].shape
    dt, k = float(z["dt"]), int(z["kts"]) - 1
    ctx = context(hip, dict(aerosolaware=True, hailaware=False))
    dims = hip.Dims(ni, nj, nk, ni, ni * nk, ni, ni * nk, ni, None)
    for tdt, prec, tol in ((torch.float64, hip.FP64, 0.0), (torch.float32, hip.FP32, 1e-6)):
        nwfa = torch.tensor(z["nwfa"], dtype=tdt, device="cuda")
        n2 = torch.tensor(z["nwfa2d"], dtype=tdt, device="cuda")
        hip._check(hip.lib().tempo_hip_aerosol_surface_emissions(ctx._ctx, dt, nwfa.data_ptr(), n2.data_ptr(),
                                                                 ctypes.byref(dims), k, prec, None))
        torch.cuda.synchronize()
        got = nwfa.double().cpu().numpy()
        if tol == 0.0:
            np.testing.assert_array_equal(got, z["out"])
        else:
            np.testing.assert_allclose(got, z["out"], rtol=tol)
    host = z["nwfa"].copy()
    hip._check(hip.lib().tempo_hip_aerosol_surface_emissions_host(ctx._ctx, dt, host.ctypes.data,
                                                                  np.ascontiguousarray(z["nwfa2d"]).ctypes.data,
                                                                  ctypes.byref(dims), k))
    np.testing.assert_array_equal(host, z["out"])
    with pytest.raises(hip.TempoError) as e:
        ctx.surface_emissions(torch.zeros((4, 5), device="cuda"), torch.zeros(5, device="cuda"), dt, k=4)
    assert e.value.code == hip.E_ARG
