"""Multi-step runs captured in a HIP graph (torch.cuda.CUDAGraph over the
library's stream-ordered launches): every replay equals the eager steps bit
for bit, with eager work and synchronisations between the replays; and the
autotuner's choices leave results unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def hip():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from tempo_amd import host
    host.lib()
    return host


def test_graph_replays_equal_eager_steps(hip):
    from tempo_amd import generator as gen
    ctx = hip.TempoHip(aerosolaware_flag=0, hailaware_flag=1)
    st0, m = gen.generate_torch("ragged", ncol=4099, device="cuda", dtype=torch.float32)
    nlev = m["nlev"]
    nz, ncol = st0["p"].shape
    ctx.reserve(ncol, nz)
    nsteps = 10

    def run(state, diags):
        for i in range(nsteps):
            ctx.run(state, m["dt"], m["itimestep"] + i, nlev=nlev, diags=diags)

    def fresh_diags():
        d = {n: torch.zeros(ncol, device="cuda") for n in hip.DIAG2_NAMES}
        d.update({n: torch.zeros((nz, ncol), device="cuda") for n in ("refl10cm", "re_cloud", "re_ice", "re_snow")})
        return d
    eager = {k: v.clone() for k, v in st0.items()}
    ed = fresh_diags()
    run(eager, ed)
    torch.cuda.synchronize()

    state = {k: v.clone() for k, v in st0.items()}
    diags = fresh_diags()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run(state, diags)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run(state, diags)
    for replay in range(3):
        for k in state:
            state[k].copy_(st0[k])
        g.replay()
        torch.cuda.synchronize()
        for k in eager:
            assert torch.equal(state[k], eager[k]), (replay, k)
        for k in ed:
            assert torch.equal(diags[k], ed[k]), (replay, k)
        # eager work between replays
        other = {k: v.clone() for k, v in st0.items()}
        ctx.run(other, m["dt"], m["itimestep"], nlev=nlev)
        torch.cuda.synchronize()


def test_autotune_choice_keeps_results(hip):
    from tempo_amd import autotune
    from tempo_amd import generator as gen
    ctx = hip.TempoHip(aerosolaware_flag=1, hailaware_flag=0)
    st0, m = gen.generate_torch("aerosol", ncol=20000, device="cuda", dtype=torch.float32)
    a = {k: v.clone() for k, v in st0.items()}
    ctx.run(a, m["dt"], m["itimestep"])
    best = autotune.autotune(ctx, st0, m["dt"], m["itimestep"])
    print("autotuned", {k: best[k] for k in ("block_a", "block_c", "order", "sed_wave", "waves_a", "ccn_lds")},
          best["phase_a_ms"])
    b = {k: v.clone() for k, v in st0.items()}
    ctx.run(b, m["dt"], m["itimestep"])
    torch.cuda.synchronize()
    for k in a:
        assert torch.equal(a[k], b[k]), k
    assert best["block_a"] in (64, 128, 256) and best["block_c"] in (64, 128, 256) and best["sed_wave"] in (0, 1)
    assert best["waves_a"] in (0, 1) and best["ccn_lds"] in (0, 1)
