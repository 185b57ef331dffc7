"""Counter-based generator: any column subset regenerates bit-for-bit."""
import numpy as np

from tempo_amd import generator as gen


def test_subset_equals_slice():
    full, _ = gen.generate("aerosol", ncol=64)
    part, _ = gen.generate("aerosol", ncol=16, col0=20)
    for k in full:
        np.testing.assert_array_equal(part[k], full[k][:, 20:36], err_msg=k)


def test_profiles_follow_the_stated_shape():
    st, m = gen.generate("plain", ncol=256)
    assert st["t"].shape == (60, 256)
    assert st["t"].max() <= 300.0 and st["t"].min() >= 210.0
    assert np.all(np.diff(st["p"][:, 0]) < 0) and 9.7e4 < st["p"][0, 0] < 1.0e5
    for q in ("qc", "qr", "qi", "qs", "qg"):
        nz = st[q][st[q] > 0]
        assert 2e-5 < np.median(nz) < 5e-4
    assert "nc" not in st and "ng" not in st
    assert abs(st["w"].std() - 0.5) < 0.05


def test_ragged_and_hail_fields():
    st, m = gen.generate("ragged", ncol=4096)
    assert m["nlev"].min() >= 30 and m["nlev"].max() <= 128
    rho_g = 1000.0 * st["qg"] / np.where(st["qb"] > 0, st["qb"], 1.0)
    sel = st["qb"] > 0
    assert rho_g[sel].min() >= 299.9 and rho_g[sel].max() <= 900.1
    assert (st["w"].max(axis=0) > 5.0).mean() > 0.2


def test_device_generator_matches_numpy():
    """generate_torch reproduces generate(): identical counters, float math
    within a few ulp."""
    torch = __import__("pytest").importorskip("torch")
    for wl in ("plain", "aerosol", "ragged", "multistep", "cold"):
        a, ma = gen.generate(wl, ncol=200, col0=4321)
        b, mb = gen.generate_torch(wl, ncol=200, col0=4321, device="cpu")
        assert sorted(a) == sorted(b)
        for k in a:
            np.testing.assert_allclose(b[k].numpy(), a[k], rtol=1e-14, atol=0, err_msg=f"{wl} {k}")
        if "nlev" in ma:
            np.testing.assert_array_equal(mb["nlev"].numpy(), ma["nlev"])
    assert gen.uniform_t(7, torch.arange(5), torch.zeros(1, dtype=torch.int64), "rh").numpy().tolist() == \
        gen.uniform(7, np.arange(5), np.zeros(1), "rh").tolist()
