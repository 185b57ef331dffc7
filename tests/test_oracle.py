"""The C restatement (oracle/) against the reference's golden vectors, known
answers of the reference's unit tests, closed forms and -- where the
reference build is present -- the reference itself."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_cases, load_case
from oracle import oracle as orc
from oracle import ref

FLAG_KEYS = set(ref.DEFAULT_FLAGS)


def run_oracle(case, tables, series=None):
    """the case's steps on the oracle; series (dict) receives every step's
    2-D diagnostics"""
    orc.init(tables, **case["meta"]["flags"])
    st = {k: v.copy() for k, v in case["inp"].items()}
    for s in range(case["meta"]["steps"]):
        d2, d3 = orc.run(st, case["meta"]["dt"], case["meta"]["itimestep"] + s, land=case["land"],
                         nlev=case["nlev"])
        if series is not None:
            for k, v in d2.items():
                series.setdefault(k, []).append(v.copy())
    return st, d2, d3


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_reference_golden(name, golden_tables):
    """Bit-identical to the reference FP64 build on every golden case.

    The tables hold only the entries the case reads (NaN elsewhere), so an
    index error in a table lookup cannot go unnoticed."""
    case = load_case(name)
    series = {}
    st, d2, d3 = run_oracle(case, golden_tables, series)
    for k, v in case["out"].items():
        np.testing.assert_array_equal(st[k], v, err_msg=f"{name}: state {k}")
    # surface precipitation (and every other 2-D diagnostic) of every step
    for k, v in case["ts"].items():
        np.testing.assert_array_equal(np.array(series[k]), v, err_msg=f"{name}: step series of {k}")
    inp = case["inp"]
    has_q = np.zeros(inp["p"].shape[1], bool)
    for q in ("qc", "qr", "qi", "qs", "qg"):
        has_q |= (inp[q] > 1e-12).any(axis=0)
    for k, v in case["d2"].items():
        if k == "frozen_fraction":
            # undefined in the reference for columns that leave tempo_main
            # early (:418): its intent(out) diagnostics keep the previous
            # column's value; such columns carry no hydrometeors
            np.testing.assert_array_equal(d2[k][has_q], v[has_q], err_msg=f"{name}: {k}")
            continue
        np.testing.assert_array_equal(d2[k], v, err_msg=f"{name}: diag {k}")
    for k, v in case["d3"].items():
        np.testing.assert_array_equal(d3[k], v, err_msg=f"{name}: diag {k}")


def _max_rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def test_ill_conditioned_cases(golden_tables):
    """One in a thousand real powers moved by one ulp (the probe) moves the
    ill-conditioned cases' outcome by more than 1e-6 and every other case's
    by no more than 1e-12: the classification the GPU parity tests rely on"""
    from conftest import ILL_CONDITIONED
    for name in golden_cases():
        case = load_case(name)
        with orc.PerturbedPow(1000):
            st, d2, d3 = run_oracle(case, golden_tables)
        worst = max([_max_rel(st[k], v) for k, v in case["out"].items()] +
                    [_max_rel(d3[k], v) for k, v in case["d3"].items()])
        if name in ILL_CONDITIONED:
            assert worst > 1e-6, (name, worst)
        else:
            assert worst <= 1e-12, (name, worst)


def test_golden_cases_cover_paths():
    names = set(golden_cases())
    for need in ("plain", "aerosol", "ragged", "multistep", "first_step", "thpii_land", "micro_off",
                 "no_condensation", "dt300", "mpas_full", "mpas_graupel_sedi", "cold_surface", "cold_surface_semi",
                 "ref_cloud_number_aero", "ref_cloud_number_const", "ref_cloud_number_ml"):
        assert need in names
    # the reference's sedimentation sweep (test/run_tempo_tests.F90:10-40), both forms
    for dt in (1, 10, 20, 60, 120, 300, 600):
        for kind in ("snow_sedi", "graupel_sedi", "graupel_semi"):
            assert f"ref_{kind}_dt{dt}" in names


def test_golden_cases_exercise_every_surface_diagnostic():
    """every 2-D diagnostic, freezing rain and the ice, snow and graupel
    surface precipitation included, is non-zero in some golden case"""
    seen = set()
    for name in golden_cases():
        case = load_case(name)
        for src in (case["d2"], case["ts"]):
            seen.update(k for k, v in src.items() if (np.asarray(v) != 0.0).any())
    for k in ("rain_precip", "ice_precip", "snow_precip", "graupel_precip", "frz_rain_precip", "frozen_fraction",
              "max_hail_sfc", "max_hail_column"):
        assert k in seen, k


@pytest.fixture(scope="module")
def oracle_plain(golden_tables):
    orc.init(golden_tables, aerosolaware=False, hailaware=False)
    return orc


def test_gamma_known_answers(oracle_plain):
    # ccg(2,4) = gamma(8) = 5040 (test/unittests/test_tempo_init_suite.F90:121-128)
    consts = oracle_plain.get_array(8)
    ccg = consts[:75].reshape(15, 5)  # [n][a]
    assert ccg[3, 1] == 5040.0
    # every gamma constant against the closed form
    for n in range(15):
        for a, x in enumerate((n + 2.0, 3.0 + n + 2.0, 3.0 + n + 5.0, n + 1 + 2.0 + 1.0, 3.0 + n + 1 + 2.0 + 1.0)):
            assert ccg[n, a] == pytest.approx(math.gamma(x), rel=1e-14)
    # crg(n) = gamma(cre(n)) with bm_r=3, mu_r=0, bv_r=1: spot values
    crg = consts[82:95]
    assert crg[0] == pytest.approx(math.gamma(4.0), rel=1e-15)      # cre(1) = bm_r + 1
    assert crg[10] == pytest.approx(math.gamma(3.0), rel=1e-15)     # cre(11) = (bv_r+5+2mu_r)/2


def test_bins_known_answers(oracle_plain):
    bins = oracle_plain.get_array(7).reshape(11, 100)
    dc = bins[0]
    # dc(nbc) = 100 microns (test/unittests/test_tempo_init_suite.F90:131-138)
    assert dc[99] == pytest.approx(100.e-6, rel=1e-12)
    # log-spaced bins: geometric means of the edges, edges from d0 to the top diameter
    dr, dtr = bins[4], bins[5]
    edges = np.concatenate([[50e-6], 50e-6 + np.cumsum(dtr)])
    assert edges[-1] == pytest.approx(0.005, rel=1e-12)
    np.testing.assert_allclose(dr, np.sqrt(edges[:-1] * edges[1:]), rtol=1e-13)
    np.testing.assert_allclose(np.diff(np.log(edges)), np.log(0.005 / 50e-6) / 100, rtol=1e-9)


def test_gamma_p_known_answer_and_closed_form():
    from scipy.special import gammainc
    # calc_gamma_p(8,8) (test/unittests/test_tempo_utils_suite.F90:169-176)
    assert abs(orc.gamma_p(8.0, 8.0) - 0.54703919051300551) < 1e-6
    for a, x in ((2.0, 0.3), (2.0, 5.0), (3.5, 1.0), (8.0, 12.0), (2.0, 20.0)):
        assert orc.gamma_p(a, x) == pytest.approx(gammainc(a, x), rel=2e-6)


def test_saturation_brackets_and_ice_below_liquid():
    for t in (220.0, 250.0, 265.0, 272.0):
        assert orc.rsif(8e4, t) < orc.rslf(8e4, t)
    # over liquid at 0 C close to the textbook 3.8 g/kg at 1000 hPa
    assert orc.rslf(1.0e5, 273.15) == pytest.approx(3.8e-3, rel=0.02)
    # 15 % vapour-pressure cap (src/module_mp_tempo_utils.F90:302)
    p = 2000.0
    assert orc.rslf(p, 330.0) == pytest.approx(0.622 * 0.15 * p / (p - 0.15 * p), rel=1e-12)


def test_table_index_brackets(oracle_plain):
    """Bracketing known answers of test/unittests/test_tempo_main_suite.F90:47-214."""
    from tempo_amd.tables_io import DATA_DIR  # noqa: F401
    bins = oracle_plain.get_array(7).reshape(11, 100)
    t_nc = bins[10]
    r_dec = np.array([m * 10.0 ** e for e in range(-6, -1) for m in range(1, 10)] + [1e-2])
    ic, inn = orc.table_index("cloud", 1.2345e-4, 54.321e6)
    assert r_dec[ic - 1] <= 1.2345e-4 <= r_dec[ic]
    assert t_nc[inn - 2] <= 54.321e6 <= t_nc[inn - 1]
    (isn, _) = orc.table_index("snow", 2.3456e-4)
    assert r_dec[isn - 1] <= 2.3456e-4 <= r_dec[isn]
    tcs = np.array([-0.01, -5., -10., -15., -20., -25., -30., -35., -40.])
    (it, _) = orc.table_index("temperature", 273.15 - 23.4)
    assert tcs[it - 2] >= -23.4 >= tcs[it - 1]
    ir, _ = orc.table_index("rain", 1.2345e-4, 3.4567e8)
    assert r_dec[ir - 1] <= 1.2345e-4 <= r_dec[ir]
    ig, _ = orc.table_index("graupel", 1.2345e-4, 3.4567e8)
    assert r_dec[ig - 1] <= 1.2345e-4 <= r_dec[ig]
    ri_dec = np.array([m * 10.0 ** e for e in range(-10, -3) for m in range(1, 10)] + [1e-3])
    nti = np.array([m * 10.0 ** e for e in range(0, 6) for m in range(1, 10)] + [1e6])
    ii, ii1 = orc.table_index("ice", 1.2345e-6, 100000.)
    assert ri_dec[ii - 1] <= 1.2345e-6 <= ri_dec[ii]
    assert nti[ii1 - 1] <= 100000. <= nti[ii1]
    # decade search against brute force over many magnitudes
    rng = np.random.default_rng(3)
    for x in 10.0 ** rng.uniform(-6, -2, 500):
        (i, _) = orc.table_index("snow", x)
        j = np.searchsorted(r_dec, x, side="right")
        assert i == max(1, min(j, 37))


def test_table_entries_match_reference_builder(golden_tables):
    """Single table entries of the restatement against entries of the tables
    the reference's own builder produced (tests/golden/tables_sample.npz)."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "tables_sample.npz"))
    orc.init(golden_tables, aerosolaware=False, hailaware=False)
    rng = np.random.default_rng(11)

    def unravel(idx, shape):
        return [int(v) + 1 for v in np.unravel_index(idx, shape, order="F")]
    for kind, shape, fn, n in (("qg", (37, 37, 9, 37, 37), orc.qg_entry, 12),
                               ("qs", (37, 9, 37, 37), orc.qs_entry, 40),
                               ("fzr", (37, 37, 45, 55), orc.fz_rain_entry, 300),
                               ("fzc", (37, 100, 45, 55), orc.fz_cloud_entry, 300)):
        idx, val = z[f"{kind}_idx"], z[f"{kind}_val"]
        for p in rng.choice(len(idx), size=n, replace=False):
            got = fn(*unravel(idx[p], shape))
            np.testing.assert_allclose(got, val[p], rtol=1e-13, atol=0,
                                       err_msg=f"{kind} entry {unravel(idx[p], shape)}")


def test_sedimentation_only_conserves_water(golden_tables):
    """With microphysics off the step only sediments: column water (mass per
    area) changes only by the surface precipitation, up to the clamps to
    r1 and to the reference's ktop rule (neither active in this case)."""
    case = load_case("mpas_graupel_sedi")
    orc.init(golden_tables, **case["meta"]["flags"])
    st = {k: v.copy() for k, v in case["inp"].items()}
    dt = case["meta"]["dt"]
    rho = lambda s: 0.622 * s["p"] / (287.04 * s["t"] * (np.maximum(1e-10, s["qv"]) + 0.622))
    for step in range(3):
        before = (st["qg"] * rho(st) * st["dz"]).sum()
        d2, _ = orc.run(st, dt, case["meta"]["itimestep"] + step)
        after = (st["qg"] * rho(st) * st["dz"]).sum()
        assert after + d2["graupel_precip"][0] == pytest.approx(before, rel=1e-6)


@pytest.mark.skipif(not ref.available(), reason="reference build not present")
def test_oracle_matches_reference_build_directly(golden_tables):
    """Fresh columns (not in the fixtures) through both the reference build
    and the restatement: bit-identical."""
    from tempo_amd import generator as gen
    for wl, flags in (("plain", dict(aerosolaware=False, hailaware=False)),
                      ("multistep", dict(aerosolaware=True, hailaware=True))):
        st, m = gen.generate(wl, ncol=12, col0=7000)
        # constants, small tables and per-point helpers
        ref.init(golden_tables, **flags)
        orc.init(golden_tables, **flags)
        for w in (1, 2, 3, 4, 5, 6, 7, 8):
            np.testing.assert_array_equal(orc.get_array(w), ref.get_array(w), err_msg=f"array {w}")
        for rs, tc in ((1e-5, -0.1), (3e-4, -12.0), (2e-3, -35.0)):
            np.testing.assert_array_equal(orc.snow_moments(rs, tc), ref.snow_moments(rs, tc))
        for p, t in ((9e4, 290.0), (5e4, 250.0), (2e4, 215.0)):
            assert orc.rslf(p, t) == ref.rslf(p, t)
            assert orc.rsif(p, t) == ref.rsif(p, t)
        # a microphysics-off step reads no big table: compare it end to end
        ref.init(golden_tables, turn_off_micro=True, **flags)
        orc.init(golden_tables, turn_off_micro=True, **flags)
        a = {k: v.copy() for k, v in st.items()}
        b = {k: v.copy() for k, v in st.items()}
        ra = ref.run(a, m["dt"], m["itimestep"])
        rb = orc.run(b, m["dt"], m["itimestep"])
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_oracle_surface_emissions_matches_reference_fixture():
    """tempo_aerosol_surface_emissions (src/module_mp_tempo_driver.F90:516-529):
    the C restatement against the reference's output on a 7 x 6 x 3 block
    (kts = 2), and only level kts changes."""
    z = np.load(os.path.join(GOLDEN, "surface_emissions.npz"))
    out = z["nwfa"].copy()
    orc.surface_emissions(float(z["dt"]), out, z["nwfa2d"], int(z["kts"]))
    np.testing.assert_array_equal(out, z["out"])
    changed = np.nonzero((z["out"] != z["nwfa"]).any(axis=(0, 2)))[0]
    assert changed.tolist() == [int(z["kts"]) - 1]
