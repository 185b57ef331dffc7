"""tempo_hip_run_noalloc (static device scratch, no allocation) against
tempo_hip_run through the C ABI: bit-identical state and diagnostics on two
calls in a row, the second on a smaller batch, in both precisions; batches
beyond the static scratch are rejected with a clear error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _args(hip, state, diags, nlev=None):
    nz, ncol = state["p"].shape
    f = hip.Fields()
    for name in hip.FIELD_NAMES:
        if name in state:
            setattr(f, name, state[name].data_ptr())
    d = hip.Diags()
    for n, a in diags.items():
        setattr(d, n, a.data_ptr())
    dims = hip.Dims(ncol, 1, nz, ncol, nz * ncol, ncol, nz * ncol, ncol,
                    None if nlev is None else nlev.data_ptr())
    return f, d, dims


def _call(hip, ctx, fn, state, dt, itimestep, dtype, nlev=None):
    nz, ncol = state["p"].shape
    diags = {n: torch.zeros(ncol, dtype=dtype, device="cuda") for n in hip.DIAG2_NAMES}
    diags.update({n: torch.zeros((nz, ncol), dtype=dtype, device="cuda") for n in hip.DIAG3_NAMES})
    f, d, dims = _args(hip, state, diags, nlev)
    prec = hip.FP32 if dtype == torch.float32 else hip.FP64
    s = torch.cuda.current_stream().cuda_stream
    rc = fn(ctx._ctx, ctypes.byref(f), ctypes.byref(d), ctypes.byref(dims), float(dt), int(itimestep), prec,
            ctypes.c_void_p(s))
    torch.cuda.synchronize()
    return rc, diags


@pytest.fixture(scope="module")
def hip():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from tempo_amd import host
    L = host.lib()
    L.tempo_hip_run_noalloc.argtypes = L.tempo_hip_run.argtypes
    return host


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_equals_run_twice_then_smaller(hip, dtype):
    from tempo_amd import generator as gen
    ctx = hip.TempoHip(aerosolaware_flag=1, hailaware_flag=1)
    L = hip.lib()
    for ncol, col0, wl in ((4096, 0, "multistep"), (700, 9000, "ragged")):
        st, m = gen.generate(wl, ncol=ncol, col0=col0)
        if wl == "ragged":  # fewer levels than the first call, ragged columns
            st = {k: np.ascontiguousarray(v[:72]) for k, v in st.items()}
            nlev = torch.tensor(np.minimum(m["nlev"], 72), dtype=torch.int32, device="cuda")
        else:
            nlev = None
        print("noalloc case", wl, ncol, str(dtype), flush=True)
        a = {k: torch.tensor(v, dtype=dtype, device="cuda") for k, v in st.items()}
        b = {k: v.clone() for k, v in a.items()}
        rc, da = _call(hip, ctx, L.tempo_hip_run, a, m["dt"], m["itimestep"], dtype, nlev)
        assert rc == hip.OK, L.tempo_hip_last_error()
        rc, db = _call(hip, ctx, L.tempo_hip_run_noalloc, b, m["dt"], m["itimestep"], dtype, nlev)
        assert rc == hip.OK, L.tempo_hip_last_error()
        for k in a:
            assert torch.equal(a[k], b[k]), (wl, k)
        for k in da:
            assert torch.equal(da[k], db[k]), (wl, k)
        # the substep counts reported after a no-alloc step are that step's
        ks = ctx.substeps(ncol)
        assert ks.shape == (10, ncol) and ks[0].max() >= 1
    ctx.close()


def test_noalloc_rejects_batches_beyond_its_scratch(hip):
    ctx = hip.TempoHip(aerosolaware_flag=0, hailaware_flag=0)
    L = hip.lib()
    nz, ncol = 90, 20000  # 20,000 x 96 x 348 bytes in FP64: beyond 256 MiB
    st = {k: torch.zeros((nz, ncol), dtype=torch.float64, device="cuda")
          for k in ("t", "p", "w", "dz", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr")}
    rc, _ = _call(hip, ctx, L.tempo_hip_run_noalloc, st, 20.0, 2, torch.float64)
    assert rc == hip.E_ARG
    msg = L.tempo_hip_last_error().decode()
    assert "bytes of scratch" in msg, msg
    ctx.close()
