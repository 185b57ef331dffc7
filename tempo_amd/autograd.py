"""Differentiable TEMPO step for torch: forward through tempo_hip_run,
backward through the adjoint kernels (tempo_hip_adjoint).

    from tempo_amd.autograd import tempo_step
    out = tempo_step(ctx, dt, itimestep, state)   # state: dict of (nz, ncol) tensors
    loss = (a * (out["rain_precip"] + out["snow_liquid_equiv_precip"] + ...)).sum() + (b * out["refl10cm"][levels]).sum()
    loss.backward()                               # grads on state["t"], state["qv"], ... (requires_grad inputs)

Differentiable inputs: t, qv, qc, qr, qi, qs, qg, nc, nr, ni, nwfa, nifa,
ng, qb (those with requires_grad get gradients, each costs one tangent pass
of the adjoint); p, w and dz enter as constants.  The outputs cover every
prognostic field, so steps chain into a multi-step adjoint, each step's
inputs being its checkpoint.  Differentiated
outputs: the updated t, qv, qc, qr, qi, qs, qg, ni, nr, nc, nwfa, nifa, ng,
qb, the four precipitation diagnostics and refl10cm.  Derivatives are
branch-frozen: every threshold keeps the branch the forward step took.
"""
import torch

X_NAMES = ("t", "qv", "qc", "qr", "qi", "qs", "qg", "nc", "nr", "ni", "nwfa", "nifa", "ng", "qb")
Y_NAMES = ("t", "qv", "qc", "qr", "qi", "qs", "qg", "ni", "nr", "nc", "nwfa", "nifa", "ng", "qb")
D_NAMES = ("rain_precip", "ice_liquid_equiv_precip", "snow_liquid_equiv_precip", "graupel_liquid_equiv_precip",
           "refl10cm")


class _TempoStep(torch.autograd.Function):
    @staticmethod
    def forward(fctx, tempo, dt, itimestep, keys, nlev, land, *tensors):
        inputs = dict(zip(keys, tensors))
        state = {k: v.detach().clone() for k, v in inputs.items()}
        diags = tempo.run(state, dt, itimestep, nlev=nlev, land=land)
        fctx.tempo, fctx.dt, fctx.itimestep, fctx.keys = tempo, dt, itimestep, keys
        fctx.nlev, fctx.land = nlev, land
        fctx.save_for_backward(*[inputs[k].detach() for k in keys])
        ys = [state[k] for k in Y_NAMES if k in state]
        ds = [diags[k] for k in D_NAMES]
        fctx.y_keys = [k for k in Y_NAMES if k in state]
        return tuple(ys + ds)

    @staticmethod
    def backward(fctx, *grads):
        saved = dict(zip(fctx.keys, fctx.saved_tensors))
        ny = len(fctx.y_keys)
        ybar = {k: g.contiguous() for k, g in zip(fctx.y_keys, grads[:ny]) if g is not None}
        dbar = {k: g.contiguous() for k, g in zip(D_NAMES, grads[ny:]) if g is not None}
        wrt = tuple(k for k, need in zip(fctx.keys, fctx.needs_input_grad[6:]) if need and k in X_NAMES)
        xbar = fctx.tempo.adjoint(saved, ybar, dbar, fctx.dt, fctx.itimestep, nlev=fctx.nlev, land=fctx.land,
                                  wrt=wrt) if wrt else {}
        out = [None] * 6 + [xbar.get(k) if k in X_NAMES else None for k in fctx.keys]
        return tuple(out)


def tempo_step(tempo, dt, itimestep, state, nlev=None, land=None):
    """One differentiable step; returns a dict of the outputs (new state
    fields and the diagnostics named in D_NAMES)."""
    keys = tuple(k for k in state if state[k] is not None)
    outs = _TempoStep.apply(tempo, dt, itimestep, keys, nlev, land, *[state[k] for k in keys])
    y_keys = [k for k in Y_NAMES if k in state]
    res = dict(zip(y_keys, outs[:len(y_keys)]))
    res.update(zip(D_NAMES, outs[len(y_keys):]))
    return res
