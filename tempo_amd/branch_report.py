"""FP32 production step against the FP64 twin, with branch accounting.

Both precisions run the same step on the same FP32-representable inputs
through tempo_hip_branch_signature, which records per point two hashes of
the decisions the level-local physics and the semi-Lagrangian sedimentation
took -- their comparisons, and their discrete outcomes (table indices, nint,
presence flags, clamps, remap source cells) -- and
tempo_hip_substeps, which gives each column's sedimentation substep counts
and tops and, for the semi-Lagrangian rain and graupel sedimentation, hashes
of its decisions (the limited interface speeds, the monotone limiter, the
source cells of each level's remap, the clamps, the surface-flux cells, the
checks between substeps).  A point *flipped* a branch when a hash of either
class differs between the precisions, or its column's sedimentation record
differs.

Errors are measured on the step's tendency (output - input)/dt of each
prognostic field, as relative L2 norms against the FP64 twin's:
  rel_l2      over the points that took the same branches, against the FP64
              output rounded to FP32 -- the FP32 path's own error: a state
              held in FP32 cannot carry its tendency closer than half a unit
              of its last place over dt, so the best any FP32 step can
              return is the FP64 result rounded;
  rel_l2_all  the same over all points;
  rel_l2_raw  over all points against the unrounded FP64 output (includes
              that representation limit).
Surface precipitation is compared per column over all columns.  Columns
whose precipitation moved by more than 1e-4 of its value (plus 1e-7 of the
largest) are counted with whether a branch flipped at their two lowest
levels or in their sedimentation record.
"""
import numpy as np

TENDENCY_FIELDS = ("t", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr", "nc", "ng", "qb", "nwfa", "nifa")
PRECIP = ("rain_precip", "ice_liquid_equiv_precip", "snow_liquid_equiv_precip", "graupel_liquid_equiv_precip")


def _rel(a, b):
    """relative L2 of a against b; absolute when b is zero"""
    nb = float(np.linalg.norm(b))
    na = float(np.linalg.norm(a - b))
    return na / nb if nb > 0.0 else na


def signature_run(ctx, state, dt, itimestep, dtype, nlev=None):
    """one step of ``state`` (numpy float64 (nz, ncol), values taken as
    given) in ``dtype`` on the current device: (outputs, diagnostics,
    signatures (3, 2, nz, ncol) int32, sedimentation record (12, ncol)) as numpy"""
    import torch
    nz, ncol = state["p"].shape
    dev = {k: torch.tensor(v, dtype=dtype, device="cuda") for k, v in state.items()}
    sig = torch.zeros((3, 2, nz, ncol), dtype=torch.int32, device="cuda")
    nl = None if nlev is None else torch.as_tensor(nlev, dtype=torch.int32, device="cuda")
    d = ctx.run(dev, dt, itimestep, nlev=nl, signature=sig)
    torch.cuda.synchronize()
    out = {k: v.double().cpu().numpy() for k, v in dev.items()}
    diags = {k: v.double().cpu().numpy() for k, v in d.items()}
    return out, diags, sig.cpu().numpy(), ctx.substeps(ncol, decisions=True)


def _f32(a):
    return a.astype(np.float32).astype(np.float64)


def compare(ctx, state, dt, itimestep, nlev=None):
    """the report of the module docstring for one step of ``state`` (numpy
    float64 (nz, ncol); rounded to FP32 first so both precisions see the
    same numbers)"""
    import torch
    inp = {k: np.ascontiguousarray(_f32(v)) for k, v in state.items()}
    o64, d64, s64, k64 = signature_run(ctx, inp, dt, itimestep, torch.float64, nlev)
    o32, d32, s32, k32 = signature_run(ctx, inp, dt, itimestep, torch.float32, nlev)
    nz, ncol = inp["p"].shape
    valid = np.ones((nz, ncol), bool)
    if nlev is not None:
        valid = np.arange(nz)[:, None] < np.asarray(nlev)[None, :]
    flip_cmp = (s32[:, 0] != s64[:, 0]).any(axis=0) & valid
    flip_idx = (s32[:, 1] != s64[:, 1]).any(axis=0) & valid
    sed_flip = (k32 != k64).any(axis=0)
    flip = flip_cmp | flip_idx | (sed_flip[None, :] & valid)
    keep = valid & ~flip
    fields = {}
    for k in TENDENCY_FIELDS:
        if k not in inp:
            continue
        t64 = (o64[k] - inp[k]) / dt
        t64r = (_f32(o64[k]) - inp[k]) / dt
        t32 = (o32[k] - inp[k]) / dt
        fields[k] = dict(rel_l2=_rel(t32[keep], t64r[keep]), rel_l2_all=_rel(t32[valid], t64r[valid]),
                         rel_l2_raw=_rel(t32[valid], t64[valid]))
    p64 = sum(d64[n] for n in PRECIP)
    p32 = sum(d32[n] for n in PRECIP)
    low = flip[: min(2, nz)].any(axis=0) | sed_flip
    moved = np.abs(p32 - p64) > 1e-4 * np.abs(p64) + 1e-7 * np.abs(p64).max()
    precip = dict(rel_l2_all=_rel(p32, p64), rel_l2=_rel(p32[~moved], p64[~moved]),
                  nonzero_columns=int((p64 != 0.0).sum()), moved_columns=int(moved.sum()),
                  moved_columns_unexplained=int((moved & ~low).sum()))
    for n in PRECIP + ("frz_rain_precip",):
        mv = np.abs(d32[n] - d64[n]) > 1e-4 * np.abs(d64[n]) + 1e-7 * np.abs(d64[n]).max()
        precip[n] = dict(rel_l2_all=_rel(d32[n], d64[n]), rel_l2=_rel(d32[n][~mv], d64[n][~mv]),
                         nonzero_columns=int((d64[n] != 0.0).sum()), moved_columns=int(mv.sum()),
                         moved_columns_unexplained=int((mv & ~low).sum()))
    nv = float(valid.sum())
    return dict(fields=fields, precip=precip,
                flipped_points=float(flip.sum() / nv),                  # any branch
                flipped_points_discrete=float(flip_idx.sum() / nv),     # table indices, presence, clamps
                flipped_points_comparisons=float(flip_cmp.sum() / nv),  # `if` comparisons
                flipped_sedimentation=float(sed_flip.mean()),           # columns
                points=int(nv), columns=int(ncol), keep=keep,
                outputs=(o32, d32, o64, d64), signatures=(s32, s64), substeps=(k32, k64))


def summary(rep):
    """the report without its arrays (JSON-ready)"""
    return {k: v for k, v in rep.items() if k not in ("outputs", "signatures", "substeps", "keep")}
