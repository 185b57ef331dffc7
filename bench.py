#!/usr/bin/env python3
"""Throughput of the TEMPO column microphysics step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|fp64]

Workload: 100,000 columns x 90 levels, aerosol-aware, one step of dt = 20 s
from the seeded inputs (synthesised in device memory by the seeded column
generator, tempo_amd/generator.py).  Each timed step starts from those
inputs, restored by a device copy outside the timed events; the same steps
run in place (the state evolving, itimestep advancing) are reported beside
it.  Every step runs the whole tempo_run replacement (tempo_hip_run:
pre-pass, level physics, sedimentation, final update and diagnostics
refl10cm, re_cloud, re_ice, re_snow and the 2-D precipitation and hail
fields).  Lookup tables are built on the GPU at init (reported, not timed).

Prints one JSON line with the metric, the roofline of the dominant kernel
(algorithmic HBM bytes over its event-timed average duration, with its VALU
and transcendental issue rates from the committed counter pass), the FP32
against FP64 report with branch accounting, the forward + adjoint timings
(16,384 x 64 single step; 12 steps x 262,144 x 90), the step on host
arrays with its PCIe transfers (tempo_hip_run_host_prec, FP64 and FP32) and
the reference CPU step timed on this host's cores over a bounded sample of
the same seeded step (oracle/cpu_baseline.py, run in a child process).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

HBM_PEAK_GBS = 8000.0
# VALU issue of the chip: 256 CUs x 4 SIMDs of 16 lanes at 2.4 GHz, so a
# wave64 VALU instruction (FP32 unpacked, as this library is built) holds
# its SIMD for 4 cycles
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4
SIMDS = 256 * 4
XCDS = 8
PHASES = ("k_prepass", "k_phase_a", "k_phase_b", "k_phase_c")
# 3-D diagnostics the default configuration computes
DIAG3 = ("refl10cm", "re_cloud", "re_ice", "re_snow")


def algorithmic_bytes(nz, ncol, width, fields_in, fields_out):
    """Algorithmic HBM bytes of one step, per kernel, for the aerosol-aware
    state.  The step's compulsory I/O ("step": every input field read once,
    every output field, 3-D and 2-D diagnostic written once) is split as:
    the pre-pass re-reads inputs counted in the level physics (0); the level
    physics reads every input field once; the final update writes every
    output and diagnostic once.  The sedimentation kernel touches no state
    field: its bytes are the scratch it must exchange in this three-kernel
    split -- point flags, layer depth, density and the nine raw fall speeds
    and the ten sedimenting masses/numbers and ten tendencies read, the ten
    tendencies written, all in the working precision."""
    lv = nz * ncol
    return {
        "k_prepass": 0,
        "k_phase_a": fields_in * width * lv,
        "k_phase_b": (4 + 41 * width) * lv,
        "k_phase_c": (fields_out + len(DIAG3)) * width * lv + 8 * width * ncol,
        "step": (fields_in + fields_out + len(DIAG3)) * width * lv + 8 * width * ncol,
    }


def cpu_baseline(ctx, cols, workers, repeats):
    """The reference -O2 build on the host cores, fed the GPU-built tables:
    the seeded step of the first ``cols`` columns, ``repeats`` times."""
    with tempfile.TemporaryDirectory(prefix="tempo_tables_") as d:
        ctx.write_tables(d)
        cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--tables", d, "--workload", "aerosol",
               "--cols", str(cols), "--steps", "1", "--repeats", str(repeats)]
        if workers:
            cmd += ["--workers", str(workers)]
        r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed:\n{r.stdout}\n{r.stderr}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    cb = {k: out[k] for k in ("value", "unit", "cores", "kind", "sample")}
    cb["cpu_model"] = out["cpu_model"]
    cb["affinity_cores"] = out["affinity_cores"]
    return cb, out


def adjoint_timing(steps):
    """Differentiable step, 16,384 columns x 64 levels, FP32: forward alone
    vs forward + adjoint of the loss (weighted surface precipitation +
    reflectivity at levels 5, 20, 40) through torch.autograd."""
    import torch
    from tempo_amd import generator as gen
    from tempo_amd import host
    from tempo_amd.autograd import tempo_step
    ctx = host.TempoHip(aerosolaware_flag=1, hailaware_flag=0)
    st, m = gen.generate_torch("adjoint", device="cuda", dtype=torch.float32)
    nz, ncol = st["p"].shape
    g = torch.Generator(device="cuda").manual_seed(3)
    aw = torch.rand(ncol, device="cuda", generator=g)
    bw = torch.rand((3, ncol), device="cuda", generator=g)
    xn = ("t", "qv", "qc", "qr", "qi", "qs", "qg", "nc", "nr")

    def fwd(grad):
        x = {k: (v.clone().requires_grad_(True) if grad and k in xn else v) for k, v in st.items()}
        out = tempo_step(ctx, m["dt"], m["itimestep"], x)
        p = out["rain_precip"] + out["snow_liquid_equiv_precip"] + out["graupel_liquid_equiv_precip"] + \
            out["ice_liquid_equiv_precip"]
        loss = (aw * p).sum() + (bw * out["refl10cm"][[4, 19, 39]]).sum()
        if grad:
            loss.backward()
        return loss

    def timed(grad):
        fwd(grad)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            with torch.set_grad_enabled(grad):
                fwd(grad)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps
    f_ms = timed(False)
    fa_ms = timed(True)
    return {"workload": f"{ncol} columns x {nz} levels, fp32, loss = weighted precipitation + dBZ at 3 levels",
            "forward_ms": round(f_ms, 3), "forward_adjoint_ms": round(fa_ms, 3), "ratio": round(fa_ms / f_ms, 2),
            "mcl_per_s_forward_adjoint": round(ncol * nz / (fa_ms * 1e-3) / 1e6, 2)}


def fp32_parity(ncol=8192, steps_before=8):
    """The FP32 path against the FP64 path with branch accounting
    (tempo_amd/branch_report.py) on the bench workload's first ncol columns:
    the seeded step, and the step after ``steps_before`` FP64 steps, when
    rain has reached the ground (the seeded step has no surface
    precipitation yet).  Per tendency: relative L2 over the points that took
    the FP64 run's branches against the FP64 result rounded to FP32, and
    over all points against the unrounded FP64 result; the fraction of points
    with a flipped branch; surface precipitation over all columns.  The FP64
    path matches the reference's FP64 build to 1e-12 (the parity tests)."""
    import torch
    from tempo_amd import branch_report as br
    from tempo_amd import generator as gen
    from tempo_amd import host
    ctx = host.TempoHip(aerosolaware_flag=1, hailaware_flag=0)
    st, m = gen.generate("aerosol", ncol=ncol)

    def brief(rep):
        s = br.summary(rep)
        return {"tendency_rel_l2": {k: float(f"{e['rel_l2']:.3g}") for k, e in s["fields"].items()},
                "tendency_rel_l2_all_points_unrounded": {k: float(f"{e['rel_l2_raw']:.3g}")
                                                         for k, e in s["fields"].items()},
                "flipped_points": float(f"{s['flipped_points']:.3g}"),
                "flipped_points_discrete": float(f"{s['flipped_points_discrete']:.3g}"),
                "flipped_sedimentation_columns": float(f"{s['flipped_sedimentation']:.3g}"),
                "precip_rel_l2_all_columns": float(f"{s['precip']['rel_l2_all']:.3g}"),
                "precip_columns": s["precip"]["nonzero_columns"]}
    seeded = brief(br.compare(ctx, st, m["dt"], m["itimestep"]))
    dev = {k: torch.tensor(v, dtype=torch.float64, device="cuda") for k, v in st.items()}
    for i in range(steps_before):
        ctx.run(dev, m["dt"], m["itimestep"] + i)
    torch.cuda.synchronize()
    later = brief(br.compare(ctx, {k: v.cpu().numpy() for k, v in dev.items()}, m["dt"],
                             m["itimestep"] + steps_before))
    return {"sample": f"{ncol} columns of the bench workload", "seeded_step": seeded,
            f"after_{steps_before}_steps": later, "tolerance": {"tendency_rel_l2": 1e-4, "precip_rel_l2": 1e-5}}


def host_path(ncol=100_000, reps=3):
    """The step on host arrays (tempo_hip_run_host_prec, the Fortran
    binding's path): the aerosol workload's FP64 host arrays copied to the
    device, stepped in FP64 and in FP32, and the updated fields and
    diagnostics copied back, PCIe transfers included; with the arrays in
    pageable memory (staged through pinned buffers by host threads) and in
    page-locked memory (tempo_hip_host_alloc: moved directly).  bytes_in /
    bytes_out are the bytes each step moves across the link."""
    import time
    import numpy as np
    from tempo_amd import generator as gen
    from tempo_amd import host
    ctx = host.TempoHip(aerosolaware_flag=1, hailaware_flag=0)
    st, m = gen.generate("aerosol", ncol=ncol)
    nz = st["p"].shape[0]
    out = {"workload": f"aerosol-aware step on host float64 arrays, {ncol} columns x {nz} levels, transfers included"}
    written = [k for k in st if k not in ("p", "w", "dz")]
    nd = len(host.DIAG2_NAMES) / nz + len(host.DIAG3_NAMES)
    up = 8 * ncol * nz * len(st)
    down = 8 * ncol * nz * (len(written) + nd)
    out["bytes_in"], out["bytes_out"] = up, int(down)
    for memory in ("pageable", "page_locked"):
        alloc = np.empty if memory == "pageable" else host.pinned_empty
        work = {k: alloc(v.shape) for k, v in st.items()}
        diags = {n: alloc(ncol) for n in host.DIAG2_NAMES}
        diags.update({n: alloc((nz, ncol)) for n in host.DIAG3_NAMES})
        for name, prec in (("fp64", host.FP64), ("fp32", host.FP32)):
            for k in work:
                np.copyto(work[k], st[k])
            ctx.run_host(work, m["dt"], m["itimestep"], precision=prec, diags=diags)
            ts = []
            for _ in range(reps):
                for k in work:
                    np.copyto(work[k], st[k])
                t0 = time.perf_counter()
                ctx.run_host(work, m["dt"], m["itimestep"], precision=prec, diags=diags)
                ts.append(time.perf_counter() - t0)
            ms = 1e3 * min(ts)
            key = name if memory == "pageable" else f"{name}_page_locked"
            out[key] = {"ms": round(ms, 2), "mcl_per_s": round(ncol * nz / (ms * 1e-3) / 1e6, 1),
                        "link_gb_per_s": round((up + down) / (ms * 1e-3) / 1e9, 1)}
        del work, diags
    ctx.close()
    return out


def ragged_graph(steps_per_graph=10, replays=3, tune=True):
    """Uneven columns: 131,072 columns of U{30..128} active levels, hail-aware,
    25 % convective, layers stretched from 25 m; 10 steps captured in one HIP
    graph (torch.cuda.CUDAGraph over the library's stream-ordered launches)
    and replayed -- with the columns in the generator's order, and, as
    ``ordered_by_depth``, the same columns stored by decreasing active depth
    (a host's choice of column order, made once before the graph; the
    library computes every column as in the other order)."""
    import torch
    from tempo_amd import autotune as at
    from tempo_amd import generator as gen
    from tempo_amd import host
    ctx = host.TempoHip(aerosolaware_flag=0, hailaware_flag=1)
    st0, m = gen.generate_torch("ragged", device="cuda", dtype=torch.float32)
    nlev = m["nlev"]
    nz, ncol = st0["p"].shape
    levels = int(nlev.sum())
    best = at.autotune(ctx, st0, m["dt"], m["itimestep"], nlev=nlev) if tune else None
    ctx.reserve(ncol, nz)
    diags = {n: torch.zeros(ncol, dtype=torch.float32, device="cuda") for n in host.DIAG2_NAMES}
    diags.update({n: torch.zeros((nz, ncol), dtype=torch.float32, device="cuda") for n in DIAG3})

    def graph_ms(src, nl):
        state = {k: v.clone() for k, v in src.items()}

        def ten_steps():
            for i in range(steps_per_graph):
                ctx.run(state, m["dt"], m["itimestep"] + i, nlev=nl, diags=diags)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            ten_steps()  # warm-up on the capture stream
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        for k in state:
            state[k].copy_(src[k])
        with torch.cuda.graph(g):
            ten_steps()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for k in state:
            state[k].copy_(src[k])
        g.replay()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(replays):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / replays

    ms = graph_ms(st0, nlev)
    order = torch.argsort(nlev.to(torch.int64), descending=True, stable=True)
    ms_sorted = graph_ms({k: v[:, order].contiguous() for k, v in st0.items()}, nlev[order].contiguous())
    rate = lambda t: round(levels * steps_per_graph / (t * 1e-3) / 1e6, 2)
    return {"workload": f"{ncol} columns of 30-128 active levels ({levels} column-levels), hail-aware, "
                        f"25 % convective, layers stretched from 25 m, {steps_per_graph} steps per HIP graph, fp32",
            "ms_per_graph": round(ms, 3), "mcl_per_s": rate(ms),
            "ordered_by_depth": {"ms_per_graph": round(ms_sorted, 3), "mcl_per_s": rate(ms_sorted)},
            "tuned": None if best is None else {k: best[k] for k in ("block_a", "block_c", "order", "sed_wave",
                                                                     "waves_a", "ccn_lds", "phase_a_ms")}}


def valu_rates(kernel, precision, kernel_ms):
    """VALU use of a kernel from the committed counter pass
    (profiles/counters_<kernel>_aerosol_<precision>.json, per-launch means):
      issue_frac  its VALU wave-instructions over this run's event-timed
                  duration, against the chip's issue rate (VALU_PEAK_WAVE_INSTS);
      busy_frac   SQ_ACTIVE_INST_VALU (quad-cycles a VALU instruction holds
                  its SIMD, transcendental ones longer) over the SIMD-cycles
                  of the launch (GRBM_GUI_ACTIVE per XCD x SIMDs), both from
                  the counter pass -- the share of the SIMDs' time spent
                  issuing VALU work;
      active_lanes  SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU): the
                  lanes that do useful work in an issued VALU instruction.
    None when there is no counter pass for the kernel."""
    path = os.path.join(HERE, "profiles", f"counters_{kernel}_aerosol_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        c = json.load(f)
    sec = kernel_ms * 1e-3
    busy = c["active_inst_valu"] * 4 / (SIMDS * c["gui_active"] / XCDS)
    return {"valu_wave_insts": c["valu_wave_insts"], "trans_wave_insts": c["trans_wave_insts"],
            "issue_frac": round(c["valu_wave_insts"] / sec / VALU_PEAK_WAVE_INSTS, 4),
            "busy_frac": round(busy, 4),
            "active_lanes": round(c["thread_cycles_valu"] / (64 * c["active_inst_valu"]), 4),
            "source": os.path.relpath(path, HERE)}


def multistep_adjoint(nsteps=12, ncol=262_144):
    """Multi-step adjoint: nsteps chained steps of the hail- and aerosol-aware
    workload (ncol x 90, FP32), every prognostic field differentiated, each
    step's inputs kept as its checkpoint (the autograd graph saves them) and
    the sedimentation substeps recomputed inside each step's adjoint."""
    import torch
    from tempo_amd import generator as gen
    from tempo_amd import host
    from tempo_amd.autograd import Y_NAMES, tempo_step
    ctx = host.TempoHip(aerosolaware_flag=1, hailaware_flag=1)
    st, m = gen.generate_torch("multistep", ncol=ncol, device="cuda", dtype=torch.float32)
    prog = [k for k in Y_NAMES if k in st]
    free0, total = torch.cuda.mem_get_info()

    def run(grad):
        x = {k: (v.clone().requires_grad_(grad and k in prog)) for k, v in st.items()}
        cur = dict(x)
        loss = 0.0
        for n in range(nsteps):
            out = tempo_step(ctx, m["dt"], m["itimestep"] + n, cur)
            loss = loss + out["rain_precip"].sum() + out["refl10cm"][[4, 19, 39]].sum()
            cur = {k: (out[k] if k in out else cur[k]) for k in cur}
        if grad:
            loss.backward()
        return loss

    with torch.no_grad():
        run(False)
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    with torch.no_grad():
        run(False)
    e[1].record()
    torch.cuda.reset_peak_memory_stats()
    e[2].record()
    run(True)
    e[3].record()
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    f_ms, fa_ms = e[0].elapsed_time(e[1]), e[2].elapsed_time(e[3])
    return {"workload": f"{nsteps} steps x {ncol} columns x {st['p'].shape[0]} levels, fp32, aerosol- and hail-aware",
            "forward_ms": round(f_ms, 2), "forward_adjoint_ms": round(fa_ms, 2), "ratio": round(fa_ms / f_ms, 2),
            "torch_peak_gib": round(torch.cuda.max_memory_allocated() / 2**30, 2),
            "device_in_use_gib": round((total - free1) / 2**30, 2)}


def copy_probe(gib):
    """A plain device copy of known size, for scaling rocprof's FETCH_SIZE and
    WRITE_SIZE counters."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": n * 4, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--ncol", type=int, default=100_000)
    ap.add_argument("--cpu-cols", type=int, default=65536, help="columns in the CPU baseline sample")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="seeded steps of the CPU baseline sample")
    ap.add_argument("--cpu-workers", type=int, default=0, help="0: every core this process may run on")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=float, default=0.0, help="run only a copy of this many GiB")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint timing")
    ap.add_argument("--adjoint-steps", type=int, default=3)
    ap.add_argument("--no-ragged", action="store_true", help="skip the uneven-column HIP-graph run")
    ap.add_argument("--no-host-path", action="store_true", help="skip the host-array (PCIe-inclusive) timing")
    ap.add_argument("--no-parity", action="store_true", help="skip the FP32-against-FP64 report")
    ap.add_argument("--multistep-adjoint", type=int, default=262_144,
                    help="columns of the 12-step chained adjoint timing (0: skip)")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the column step shards by columns with no per-step exchange, "
              "and this build runs on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
        return

    import torch
    from tempo_amd import generator as gen
    from tempo_amd import host

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    dtype = torch.float32 if a.precision == "fp32" else torch.float64
    width = 4 if dtype == torch.float32 else 8

    ctx = host.TempoHip(aerosolaware_flag=1, hailaware_flag=0)
    seeded, meta = gen.generate_torch("aerosol", ncol=a.ncol, device=dev, dtype=dtype)
    state = {k: v.clone() for k, v in seeded.items()}
    nz, ncol = state["p"].shape
    diags = {n: torch.zeros(ncol, dtype=dtype, device=dev) for n in host.DIAG2_NAMES}
    diags.update({n: torch.zeros((nz, ncol), dtype=dtype, device=dev) for n in DIAG3})
    ctx.reserve(ncol, nz)
    stream = torch.cuda.current_stream(dev)

    def restore():
        for k, v in seeded.items():
            state[k].copy_(v)

    # the workload's step: every timed step from the seeded inputs, restored
    # by a device copy between the events
    for _ in range(a.warmup):
        restore()
        ctx.run(state, meta["dt"], meta["itimestep"], diags=diags)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    for e0, e1 in ev:
        restore()
        e0.record(stream)
        ctx.run(state, meta["dt"], meta["itimestep"], diags=diags)
        e1.record(stream)
    torch.cuda.synchronize()
    ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / a.steps

    # per-kernel durations of the same seeded step: HIP events recorded by
    # the library around each launch, on the stream the kernels run on
    ctx.profile(True)
    for _ in range(a.steps):
        restore()
        ctx.run(state, meta["dt"], meta["itimestep"], diags=diags)
    phase_ms, runs = ctx.profile_read()
    ctx.profile(False)
    phase_ms = phase_ms / max(runs, 1)
    for v in state.values():
        if not torch.isfinite(v).all():
            raise RuntimeError("non-finite state after the timed steps")

    # the same number of steps in place, the state evolving
    restore()
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for i in range(a.steps):
        ctx.run(state, meta["dt"], meta["itimestep"] + i, diags=diags)
    t1.record(stream)
    torch.cuda.synchronize()
    ms_inplace = t0.elapsed_time(t1) / a.steps

    levels = nz * ncol
    value = levels / (ms * 1e-3) / 1e6
    fields_in = sum(1 for k in state if k != "th" and k != "pii")
    fields_out = sum(1 for k in state if k not in ("p", "w", "dz"))
    ab = algorithmic_bytes(nz, ncol, width, fields_in, fields_out)
    dom = int(max(range(4), key=lambda i: phase_ms[i]))
    dom_name = PHASES[dom]
    dom_bytes = ab[dom_name]
    achieved = dom_bytes / (phase_ms[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(HERE, "profiles", f"traffic_{dom_name}_aerosol_{a.precision}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

    valu = valu_rates(dom_name, a.precision, float(phase_ms[dom]))
    out = {
        "metric": "Mcolumn-levels/s per step (forward)",
        "value": round(value, 3),
        "unit": "Mcolumn-levels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if width == 4 else "f64",
        "data": "synthetic (seeded counter-based column generator, device-side)",
        "config": {"workload": f"aerosol-aware step from the seeded inputs, {ncol} columns x {nz} levels, "
                               f"dt {meta['dt']:g} s",
                   "columns": ncol, "levels": nz, "dt": meta["dt"], "precision": a.precision,
                   "table_build_ms": round(ctx.table_build_ms, 2)},
        "in_place": {"ms_per_step": round(ms_inplace, 4),
                     "value": round(levels / (ms_inplace * 1e-3) / 1e6, 3),
                     "note": f"{a.steps} consecutive steps of the evolving state, itimestep advancing"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": dom_name, "kernel_ms": round(float(phase_ms[dom]), 4),
                     "algorithmic_bytes": dom_bytes,
                     "step_algorithmic_bytes": ab["step"],
                     "step_achieved": round(ab["step"] / (ms * 1e-3) / 1e9, 2),
                     "valu": valu},
        "phase_ms": {PHASES[i]: round(float(phase_ms[i]), 4) for i in range(4)},
        "cpu_baseline": None,
    }
    if a.precision == "fp32" and not a.no_parity:
        out["fp32_parity"] = fp32_parity()
    if not a.no_adjoint:
        out["adjoint"] = adjoint_timing(a.adjoint_steps)
    if not a.no_ragged:
        out["ragged_graph"] = ragged_graph()
    if not a.no_host_path:
        out["host_path"] = host_path()
    if a.multistep_adjoint:
        out["multistep_adjoint"] = multistep_adjoint(ncol=a.multistep_adjoint)
    if not a.no_cpu_baseline:
        cb, raw = cpu_baseline(ctx, a.cpu_cols, a.cpu_workers, a.cpu_repeats)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(value / cb["value"], 2)
        out["cpu_baseline_wall_s"] = round(raw["wall_s"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
