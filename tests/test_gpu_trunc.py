"""Device local-truncation-error step proposals on the GPU against the
reference's own compute_trunc (capacitor.cpp:242-250, dio_device.cpp:
278-291, mos1_device.cpp:352-365, all through ckt_terr.hpp:19-64), device
by device and step by step, on the reference's fixed-step sequence
(backward Euler, then trapezoidal order 2) with the proposals taken after
each step's Newton solve and before the states rotate, as
transient.cpp:478-485 takes them.

Tolerances: on a linear circuit the GPU states are bit-identical to the
reference's, so the proposals agree to the last bits of the square root
(1e-12 relative).  With diodes or MOS1 the states agree to the transient
bar (transcendental functions differ in the last ulps), and a proposal is
a second divided difference of the charge history, which amplifies that:
1e-6 relative there."""
import numpy as np
import pytest

import neospice_amd as na
import refbind as RB
from neospice_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.ref]

XMU = 0.5


def step_ctx(k, dt):
    """The context ref_tran_trunc uses at step k (0-based)."""
    t = (k + 1) * dt
    if k == 0:
        return na.make_ctx(na.MODETRAN | na.MODEINITTRAN, 1, 0, (1.0 / dt, -1.0 / dt, 0.0), 0.0, t,
                           delta=dt, delta_old1=dt, delta_old2=dt)
    om = 1.0 - XMU
    return na.make_ctx(na.MODETRAN | na.MODEINITPRED, 2, 0, (1.0 / (dt * om), -1.0 / (dt * om), 0.0),
                       XMU / om, t, delta=dt, delta_old1=dt, delta_old2=dt)


def charged_inverters(stages=3):
    """MOS1 inverters whose cards carry gate-oxide, overlap and junction
    charge, so that every MOS1 charge state moves."""
    c = na.Circuit()
    c.model(".model nch NMOS(LEVEL=1 VTO=0.5 KP=120u LAMBDA=0.02 TOX=4e-9 CGSO=3e-10 CGDO=3e-10 "
            "CGBO=1e-10 CBD=2e-15 CBS=2e-15 PB=0.8)")
    c.model(".model pch PMOS(LEVEL=1 VTO=-0.5 KP=40u LAMBDA=0.02 TOX=4e-9 CGSO=3e-10 CGDO=3e-10 "
            "CGBO=1e-10 CBD=4e-15 CBS=4e-15 PB=0.8)")
    vdd, nin = c.node("vdd"), c.node("in")
    c.V("vsup", vdd, na.GND, 1.8)
    c.V_pulse("vin", nin, na.GND, 0.0, 1.8, 0.0, 20e-12, 20e-12, 0.2e-9, 0.5e-9)
    drive = nin
    for k in range(stages):
        o = c.node("o%d" % k)
        c.M("mn%d" % k, o, drive, na.GND, na.GND, "nch", 1e-6, 0.18e-6)
        c.M("mp%d" % k, o, drive, vdd, vdd, "pch", 2e-6, 0.18e-6)
        c.C("cl%d" % k, o, na.GND, 2e-15)
        drive = o
    return c.finalize()


def check_trunc(ckt, dt, nsteps, rtol, B=2):
    _, tr_ref = RB.RefCircuit(ckt, probe=False).tran_trunc(dt, nsteps)
    eng = na.BatchEngine(ckt, batch=B, xmu=XMU)
    eng.tran_begin()
    moved = set()
    for k in range(nsteps):
        conv, _ = eng.tran_step(step_ctx(k, dt))
        assert conv.all(), k
        dt_min, idx, dev = eng.trunc(per_device=True)
        ref = tr_ref[k]
        # every device without a row proposes nothing in the reference either
        others = np.setdiff1d(np.arange(len(ref)), idx)
        assert (ref[others] == 1e30).all(), k
        for i in range(B):
            np.testing.assert_allclose(dev[:, i], ref[idx], rtol=rtol, atol=0, err_msg="step %d" % k)
            np.testing.assert_allclose(dt_min[i], ref.min(), rtol=rtol, atol=0)
        moved.update(int(d) for d in idx[ref[idx] < 1e29])
        eng.rotate_states()
    return idx, moved


def test_trunc_linear_rc_mesh():
    ckt, _ = W.diode_mesh(4, t_stop=40e-12, diodes=False)
    ckt.finalize()
    idx, moved = check_trunc(ckt, 2e-12, 10, rtol=1e-12)
    assert len(idx) == len(ckt.family["C"]) and moved


def test_trunc_diode_mesh():
    ckt, _ = W.diode_mesh(4, t_stop=40e-12)
    ckt.finalize()
    idx, moved = check_trunc(ckt, 2e-12, 10, rtol=1e-6)
    kinds = {ckt.devices[d][0] for d in moved}
    assert len(idx) == len(ckt.family["C"]) + len(ckt.family["D"]) and kinds == {"C", "D"}, kinds


def test_trunc_mos1_charges():
    ckt = charged_inverters(3)
    idx, moved = check_trunc(ckt, 5e-12, 16, rtol=1e-6)
    kinds = [ckt.devices[d][0] for d in moved]
    assert "M" in kinds and "C" in kinds, kinds


def bsim4_inverters(stages=3):
    """BSIM4v7 inverters driven by an input edge, so that every device's
    charges move from the first step (on a static circuit the second
    differences are roundoff and the proposals are not comparable)."""
    c = na.Circuit()
    c.model(".model nb nmos level=14 toxe=1.8e-9 vth0=0.4")
    c.model(".model pb pmos level=14 toxe=1.8e-9 vth0=-0.4")
    vdd, nin = c.node("vdd"), c.node("in")
    c.V("vsup", vdd, na.GND, 1.0)
    c.V_pulse("vin", nin, na.GND, 0.0, 1.0, 0.0, 20e-12, 20e-12, 0.2e-9, 0.5e-9)
    drive = nin
    for k in range(stages):
        o = c.node("o%d" % k)
        c.M("mn%d" % k, o, drive, na.GND, na.GND, "nb", 0.5e-6, 0.05e-6)
        c.M("mp%d" % k, o, drive, vdd, vdd, "pb", 1e-6, 0.05e-6)
        c.C("cl%d" % k, o, na.GND, 1e-15)
        drive = o
    return c.finalize()


def test_trunc_bsim4():
    ckt = bsim4_inverters(1)
    idx, moved = check_trunc(ckt, 2e-12, 10, rtol=1e-6)
    kinds = [ckt.devices[d][0] for d in moved]
    assert kinds.count("M") == 2 and "C" in kinds, kinds
