"""FP32 production step against the FP64 twin, with branch accounting.

Both precisions run the same step on the same FP32-representable inputs
through tempo_hip_branch_signature, which records per point a hash of every
threshold decision the level-local physics took, and tempo_hip_substeps,
which gives each column's sedimentation substep counts and tops.  A point
*flipped* a branch when its two hashes (physics up to sedimentation, final
update) differ between the precisions; a column flipped when any of its
points did or its substeps or tops differ.

Reported per prognostic field: the relative L2 error of the step's tendency
(output - input)/dt over the points that took the same branches (filtered)
and over all points (unfiltered); the flipped fraction of points and of
columns; and the relative L2 error of the surface precipitation over all
columns and over the columns that took the same branches.
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
    signatures (2, nz, ncol) int32, substeps (10, ncol)) as numpy"""
    import torch
    nz, ncol = state["p"].shape
    dev = {k: torch.tensor(v, dtype=dtype, device="cuda") for k, v in state.items()}
    sig = torch.zeros((2, nz, ncol), dtype=torch.int32, device="cuda")
    nl = None if nlev is None else torch.as_tensor(nlev, dtype=torch.int32, device="cuda")
    d = ctx.run(dev, dt, itimestep, nlev=nl, signature=sig)
    torch.cuda.synchronize()
    out = {k: v.double().cpu().numpy() for k, v in dev.items()}
    diags = {k: v.double().cpu().numpy() for k, v in d.items()}
    return out, diags, sig.cpu().numpy(), ctx.substeps(ncol)


def compare(ctx, state, dt, itimestep, nlev=None):
    """the report of the module docstring for one step of ``state`` (numpy
    float64 (nz, ncol); rounded to FP32 first so both precisions see the
    same numbers)"""
    import torch
    inp = {k: np.ascontiguousarray(v.astype(np.float32).astype(np.float64)) for k, v in state.items()}
    o64, d64, s64, k64 = signature_run(ctx, inp, dt, itimestep, torch.float64, nlev)
    o32, d32, s32, k32 = signature_run(ctx, inp, dt, itimestep, torch.float32, nlev)
    nz, ncol = inp["p"].shape
    valid = np.ones((nz, ncol), bool)
    if nlev is not None:
        valid = np.arange(nz)[:, None] < np.asarray(nlev)[None, :]
    pt_flip = ((s32[0] != s64[0]) | (s32[1] != s64[1])) & valid
    col_flip = pt_flip.any(axis=0) | (k32 != k64).any(axis=0)
    keep = valid & ~pt_flip & ~col_flip[None, :]
    fields = {}
    for k in TENDENCY_FIELDS:
        if k not in inp:
            continue
        t64 = (o64[k] - inp[k]) / dt
        t32 = (o32[k] - inp[k]) / dt
        fields[k] = dict(rel_l2=_rel(t32[keep], t64[keep]), rel_l2_all=_rel(t32[valid], t64[valid]))
    p64 = sum(d64[n] for n in PRECIP)
    p32 = sum(d32[n] for n in PRECIP)
    keep_c = ~col_flip
    precip = dict(rel_l2_all=_rel(p32, p64), rel_l2=_rel(p32[keep_c], p64[keep_c]),
                  nonzero_columns=int((p64 != 0.0).sum()))
    for n in PRECIP + ("frz_rain_precip",):
        precip[n] = dict(rel_l2_all=_rel(d32[n], d64[n]), rel_l2=_rel(d32[n][keep_c], d64[n][keep_c]),
                         nonzero_columns=int((d64[n] != 0.0).sum()))
    return dict(fields=fields, precip=precip, flipped_points=float(pt_flip.sum() / valid.sum()),
                flipped_columns=float(col_flip.mean()), points=int(valid.sum()), columns=int(ncol),
                outputs=(o32, d32, o64, d64), signatures=(s32, s64), substeps=(k32, k64))


def summary(rep):
    """the report without its arrays (JSON-ready)"""
    return {k: v for k, v in rep.items() if k not in ("outputs", "signatures", "substeps")}
