"""Launch-shape autotuner for the column step.

Times the step's kernels under each candidate launch shape with the
library's own per-phase HIP events (tempo_hip_profile) and keeps the fastest
choice per phase:

  * level-local physics (phase A) and final update (phase C): columns per
    block, 64 / 128 / 256 (one lane per column at one level), and whether
    each block groups its points with microphysics to do into its first
    wavefronts;
  * sedimentation: one wavefront per column with LDS-staged column groups,
    or one lane per column;
  * then, at the chosen blocks and point order, phase A's register budget
    (compiled for its default waves per SIMD or one fewer) and whether its
    blocks stage the CCN activation table's slice in LDS.

Every candidate computes identical results (the kernels do the same
arithmetic per point), so the choice affects time only.  Results are cached
per (precision, nz, column-count bucket, flags) for the life of the process
and applied to the context with tempo_hip_tune.
"""
import itertools

from . import host

BLOCKS = (64, 128, 256)
SED = (1, 0)
ORDERS = (0, 3)  # point grouping off / on in both level-local kernels
WAVES_A = (0, 1)  # phase A at its default waves per SIMD / one fewer
CCN_LDS = (0, 1)  # CCN table slice read from global memory / staged in LDS

_cache = {}


def _key(ctx, precision, nz, ncol):
    bucket = 1 << max(0, int(ncol - 1).bit_length())
    c = ctx.cfg
    return (precision, nz, bucket, c.aerosolaware_flag, c.hailaware_flag)


def _phase_ms(ctx, state, dt, itimestep, nlev, reps):
    import torch
    work = {k: v.clone() for k, v in state.items()}
    ctx.run(work, dt, itimestep, nlev=nlev, diags=False)  # warm-up with this shape
    ctx.profile(True)
    for _ in range(reps):
        for k in work:
            work[k].copy_(state[k])
        ctx.run(work, dt, itimestep, nlev=nlev, diags=False)
    ms, runs = ctx.profile_read()
    ctx.profile(False)
    torch.cuda.synchronize()
    return ms / max(runs, 1)


def autotune(ctx, state, dt, itimestep, nlev=None, reps=3, verbose=False):
    """Pick and apply the fastest launch shapes for ``state``'s shape and
    precision; returns {"block_a": .., "block_c": .., "order": .., "sed_wave": .., "waves_a": ..,
    "ccn_lds": .., "ms": {...}, "phase_a_ms": {...}}."""
    import torch
    nz, ncol = state["p"].shape
    precision = host.FP32 if state["p"].dtype == torch.float32 else host.FP64
    key = _key(ctx, precision, nz, ncol)
    if key not in _cache:
        # phase A's time depends only on (block_a, order bit 0), phase C's on
        # (block_c, order bit 1) and sedimentation's on sed_wave, so one
        # sweep of (block, order) with equal blocks covers both
        timings = {}
        for b, order, sw in itertools.product(BLOCKS, ORDERS, SED):
            if sw == 0 and (b != BLOCKS[-1] or order != ORDERS[-1]):
                continue
            ctx.tune(precision, host.TUNE_BLOCK_A, b)
            ctx.tune(precision, host.TUNE_BLOCK_C, b)
            ctx.tune(precision, host.TUNE_ORDER, order)
            ctx.tune(precision, host.TUNE_SED_WAVE, sw)
            timings[(b, order, sw)] = _phase_ms(ctx, state, dt, itimestep, nlev, reps)
            if verbose:
                print("autotune", (b, order, sw), [round(float(x), 4) for x in timings[(b, order, sw)]])
        block_a, oa = min(((k[0], k[1] & 1) for k in timings),
                          key=lambda p: min(t[1] for k, t in timings.items() if (k[0], k[1] & 1) == p))
        block_c, oc = min(((k[0], k[1] & 2) for k in timings),
                          key=lambda p: min(t[3] for k, t in timings.items() if (k[0], k[1] & 2) == p))
        best_s = min(SED, key=lambda s: min(t[2] for k, t in timings.items() if k[2] == s))
        # phase A's register budget and CCN staging at the chosen shape
        ctx.tune(precision, host.TUNE_BLOCK_A, block_a)
        ctx.tune(precision, host.TUNE_BLOCK_C, block_c)
        ctx.tune(precision, host.TUNE_ORDER, oa | oc)
        ctx.tune(precision, host.TUNE_SED_WAVE, best_s)
        phase_a = {}
        for w, cl in itertools.product(WAVES_A, CCN_LDS):
            ctx.tune(precision, host.TUNE_WAVES_A, w)
            ctx.tune(precision, host.TUNE_CCN_LDS, cl)
            phase_a[(w, cl)] = float(_phase_ms(ctx, state, dt, itimestep, nlev, reps)[1])
            if verbose:
                print("autotune phase A", (w, cl), round(phase_a[(w, cl)], 4))
        waves_a, ccn_lds = min(phase_a, key=phase_a.get)
        _cache[key] = dict(block_a=block_a, block_c=block_c, order=oa | oc, sed_wave=best_s, waves_a=waves_a,
                           ccn_lds=ccn_lds,
                           ms={f"{k[0]}/{k[1]}/{k[2]}": [round(float(x), 4) for x in v] for k, v in timings.items()},
                           phase_a_ms={f"waves{k[0]}/ccn{k[1]}": round(v, 4) for k, v in phase_a.items()})
    best = _cache[key]
    ctx.tune(precision, host.TUNE_BLOCK_A, best["block_a"])
    ctx.tune(precision, host.TUNE_BLOCK_C, best["block_c"])
    ctx.tune(precision, host.TUNE_ORDER, best["order"])
    ctx.tune(precision, host.TUNE_SED_WAVE, best["sed_wave"])
    ctx.tune(precision, host.TUNE_WAVES_A, best["waves_a"])
    ctx.tune(precision, host.TUNE_CCN_LDS, best["ccn_lds"])
    return best
