"""torch.autograd.Function wrappers over the engine's adjoints.

PyTorch is used here only as the differentiation front end: parameters and
losses are torch tensors, while the forward solves and the adjoints run in
the native engine (HIP kernels behind the C ABI).  Tensors are float64 and
shaped (count, B): one column per circuit instance.

A Function keeps a reference to the engine and reads the engine's state in
backward (the last Newton factors for the DC adjoint, the last transient
record for the transient adjoint), so call backward before the next forward
on the same engine.

BundleDCOperatingPoint does the same for a bundle of circuits (any
topologies): its tensors are per device over every instance (the layout of
Bundle.device_values), and the solution is packed instance after instance.
"""
import numpy as np
import torch


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


def _t(a, like):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=like.device, dtype=torch.float64)


class DCOperatingPoint(torch.autograd.Function):
    """x = DC operating point (n x B) as a function of the R, C, V and I
    values (each count x B; C has no effect in DC and receives a zero
    gradient)."""

    @staticmethod
    def forward(ctx, engine, R, C, V, I):
        for fam, val in (("R", R), ("C", C), ("V", V), ("I", I)):
            if val is not None and engine.family_count(fam):
                engine.set_family_values(fam, _np(val))
        x, _ = engine.dc_op()
        ctx.engine = engine
        ctx.like = R
        return _t(x, R)

    @staticmethod
    def backward(ctx, gx):
        g = ctx.engine.dc_adjoint(_np(gx))
        like = ctx.like
        return (None, _t(g["R"], like), _t(g["C"], like), _t(g["V"], like), _t(g["I"], like))


class TransientProbes(torch.autograd.Function):
    """Probe-node voltages ((nsteps+1) x nprobe x B) of a fixed-step
    transient (TRANOP operating point, backward Euler first step, then
    trapezoidal) as a function of the R and C values, with recomputation
    checkpoints every `ckpt_every` steps in the adjoint."""

    @staticmethod
    def forward(ctx, engine, nsteps, dt, probes, ckpt_every, R, C):
        engine.set_family_values("R", _np(R))
        engine.set_family_values("C", _np(C))
        out = engine.tran_forward(int(nsteps), float(dt), np.asarray(probes, np.int32),
                                  int(ckpt_every))
        ctx.engine = engine
        ctx.like = R
        return _t(out, R)

    @staticmethod
    def backward(ctx, gout):
        gR, gC = ctx.engine.tran_adjoint(_np(gout))
        return (None, None, None, None, None, _t(gR, ctx.like), _t(gC, ctx.like))


class BundleDCOperatingPoint(torch.autograd.Function):
    """Every instance's DC operating point (packed unknowns, total_n) as a
    function of the per-device values (R ohms, C farads, V and I dc), W and
    L (BSIM4v7 devices), each a tensor over all devices of the bundle.
    Forward: solve_dc's ladder (direct Newton, gmin stepping for instances
    that need it, MODEDCOP); backward: the bundle's DC adjoint (transposed
    solves with the last factors, dL/dW and dL/dL through BSIM4v7temp's
    size-dependent constants).  C receives a zero gradient (open in DC)."""

    @staticmethod
    def forward(ctx, bundle, value, w, l):
        bundle.set_device_values(_np(value), _np(w), _np(l))
        conv, _, _ = bundle.op_ladder(dcop=True)
        if not conv.all():
            raise RuntimeError("DC operating point failed for instances %s" % np.nonzero(~conv)[0][:10].tolist())
        ctx.bundle = bundle
        ctx.like = value
        return _t(bundle.solution(), value)

    @staticmethod
    def backward(ctx, gx):
        g = ctx.bundle.dc_adjoint(_np(gx))
        like = ctx.like
        return None, _t(g["value"], like), _t(g["w"], like), _t(g["l"], like)
