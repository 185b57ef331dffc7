"""The Fortran binding (tempo_amd/fortran/tempo_hip_mod.F90) driven like a
host model drives the reference: tempo_init + tempo_run on a 2-row tile of
arrays with halos.  Its results must equal the C ABI's on the same columns
bit for bit, and the halo must stay untouched."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN_CASE = os.path.join(ROOT, "tempo_amd", "fortran", "_build", "run_case")
FIELDS = ("t", "p", "w", "dz", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr", "nc", "nwfa", "nifa", "ng", "qb")
OUT = ("t", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr", "nc", "nwfa", "nifa", "ng", "qb")
D2 = ("rain_precip", "ice_liquid_equiv_precip", "snow_liquid_equiv_precip", "graupel_liquid_equiv_precip",
      "frozen_fraction", "frz_rain_precip", "max_hail_diameter_sfc", "max_hail_diameter_column")
D3 = ("refl10cm", "re_cloud", "re_ice", "re_snow")


@pytest.mark.parametrize("memory", ["pageable", "locked"])
def test_fortran_binding_matches_c_abi(tmp_path, memory):
    """pageable: the state in allocatable arrays; locked: in page-locked
    memory from tempo_hip_host_alloc, moved without staging copies (and the
    3-D helper tempo_alloc_page_locked checked for its bounds)"""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(RUN_CASE), "Fortran driver not built (make -C tempo_amd/fortran)"
    from tempo_amd import generator as gen
    from tempo_amd import host
    ni, nj = 20, 2
    st, m = gen.generate("multistep", ncol=ni * nj, col0=7000)
    nz = st["p"].shape[0]
    land = (np.arange(ni * nj) % 3 == 0).astype(np.int32)
    # Fortran order (i, k, j): column c = i + j*ni
    def fort(a):
        return np.ascontiguousarray(a.reshape(nz, nj, ni).transpose(1, 0, 2))  # [j][k][i]
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([ni, nj, nz, m["itimestep"]], dtype=np.int32).tofile(f)
        np.array([m["dt"]], dtype=np.float64).tofile(f)
        for k in FIELDS:
            fort(st[k]).tofile(f)
        land.tofile(f)
    print("run_case", ni, nj, nz, memory, flush=True)
    r = subprocess.run([RUN_CASE, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), host.CCN_PATH, memory],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.uint8)
    n3 = ni * nz * nj
    vals = raw[:-4].view(np.float64)
    halo_ok = int(raw[-4:].view(np.int32)[0])
    got, o = {}, 0
    for k in OUT:
        got[k] = vals[o:o + n3].reshape(nj, nz, ni).transpose(1, 0, 2).reshape(nz, ni * nj)
        o += n3
    for k in D2:
        got[k] = vals[o:o + ni * nj]
        o += ni * nj
    for k in D3:
        got[k] = vals[o:o + n3].reshape(nj, nz, ni).transpose(1, 0, 2).reshape(nz, ni * nj)
        o += n3
    assert o == vals.size
    assert halo_ok == 1

    ctx = host.TempoHip(aerosolaware_flag=1, hailaware_flag=1)
    dev = {k: torch.tensor(v, dtype=torch.float64, device="cuda") for k, v in st.items()}
    d = ctx.run(dev, m["dt"], m["itimestep"], land=torch.tensor(land))
    torch.cuda.synchronize()
    for k in OUT:
        np.testing.assert_array_equal(got[k], dev[k].cpu().numpy(), err_msg=k)
    for k in D2 + D3:
        np.testing.assert_array_equal(got[k], d[k].cpu().numpy(), err_msg=k)
