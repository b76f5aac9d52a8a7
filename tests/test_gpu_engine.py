"""GPU parity of the batched Newton engine on the linear RC power-grid mesh
(the batched-refactor workload) against the compiled reference: device
loads, DC operating point, and a transient replayed on the reference's own
accepted/rejected step sequence."""
import numpy as np
import pytest

import neospice_amd as na
import refbind as RB
from helpers import mesh
from neospice_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.ref]


def instance_values(info, B):
    R = W.mesh_instance_resistors(info, B)
    return R


def test_load_bit_exact_dc_and_tran():
    ckt, info = mesh(8)
    rc = RB.RefCircuit(ckt)
    eng = na.BatchEngine(ckt, batch=2)
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, ckt.n)
    eng.set_solution(np.repeat(x[:, None], 2, axis=1))
    # DC load
    eng.set_ctx(na.make_ctx(na.MODEDCOP | na.MODEINITJCT))
    rc.set_ctx(na.MODEDCOP | na.MODEINITJCT)
    v_ref, r_ref = rc.load(x)
    v, r = eng.load()
    assert np.array_equal(v[:, 0], v_ref) and np.array_equal(r[:, 0], r_ref)
    # transient load (BE predictor step, states seeded from a DC pass)
    dt = 1e-12
    ctx = na.be_ctx(dt, dt, first=False)
    eng.set_ctx(ctx)
    rc.set_ctx(ctx.mode, 1, 0, (ctx.ag0, ctx.ag1, 0.0), dt, dt, 0.0, dt)
    v_ref, r_ref = rc.load(x)
    v, r = eng.load()
    assert np.array_equal(v[:, 1], v_ref)
    assert np.array_equal(r[:, 1], r_ref)


@pytest.mark.parametrize("m,B", [(8, 3), (16, 2)])
def test_dc_op_batched(m, B):
    ckt, info = mesh(m)
    R = instance_values(info, B)
    eng = na.BatchEngine(ckt, batch=B)
    for k, d in enumerate(info["res_devs"]):
        eng.set_values(int(d), R[k])
    x, its = eng.dc_op()
    for i in range(B):
        rc = RB.RefCircuit(ckt, values={int(d): float(R[k, i]) for k, d in enumerate(info["res_devs"])})
        x_ref, its_ref, _ = rc.dc()
        assert its[i] == its_ref
        assert np.array_equal(x[:, i], x_ref), np.max(np.abs(x[:, i] - x_ref))


def replay(eng, calls):
    """Replays the reference's transient Newton calls: accepted calls rotate
    the state history, rejected ones restore the last accepted solution
    (transient.cpp:916-960, 1004)."""
    xs = []
    last = eng.solution()
    for c in calls:
        ctx = na.make_ctx(c["mode"], c["order"], c["method"], c["ag"], c["xmu_ratio"], c["time"],
                          delta=c["delta"], delta_old1=c["dold1"])
        conv, its = eng.tran_step(ctx)
        if c["accepted"]:
            assert conv.all()
            eng.rotate_states()
            last = eng.solution()
            xs.append((c["time"], last.copy()))
        else:
            eng.set_solution(last)
    return xs


@pytest.mark.parametrize("m", [6, 10])
def test_transient_replay(m):
    """Backward-Euler transient (xmu=0) of the mesh, adaptive in the
    reference; the engine replays the same step sequence and must produce
    the same waveforms bit for bit."""
    ckt, info = mesh(m)
    rc = RB.RefCircuit(ckt, options={"xmu": 0.0})
    t_ref, x_ref, its, _ = rc.tran(1e-12, 20e-12)
    calls = RB.split_newton_calls(*rc.records())
    assert sum(c["accepted"] for c in calls) == len(t_ref) - 1
    eng = na.BatchEngine(ckt, batch=1, xmu=0.0)
    x0, _ = eng.tran_begin()
    assert np.array_equal(x0[:, 0], x_ref[0])
    xs = replay(eng, calls)
    assert len(xs) == len(t_ref) - 1
    for k, (t, x) in enumerate(xs):
        assert t == t_ref[k + 1]
        assert np.array_equal(x[:, 0], x_ref[k + 1]), (k, np.max(np.abs(x[:, 0] - x_ref[k + 1])))


def test_pivot_groups_dc_full_mesh():
    """On the 64x64 mesh the reference's AMD-LU picks a different pivot row
    for some perturbed instances (instance 61 of the seeded set): the batch
    splits into pivot groups and every instance still matches its own
    reference solve bit for bit."""
    ckt, info = mesh(64)
    Rall = W.mesh_instance_resistors(info, 62)
    pick = [0, 61, 1, 2]
    R = Rall[:, pick]
    eng = na.BatchEngine(ckt, batch=len(pick))
    for k, d in enumerate(info["res_devs"]):
        eng.set_values(int(d), R[k])
    x, its = eng.dc_op()
    assert eng.stats()["pivot_groups"] >= 2
    for i in range(len(pick)):
        rc = RB.RefCircuit(ckt, values={int(d): float(R[k, i]) for k, d in enumerate(info["res_devs"])},
                           probe=False)
        x_ref, its_ref, _ = rc.dc()
        assert its[i] == its_ref
        assert np.array_equal(x[:, i], x_ref), (i, np.max(np.abs(x[:, i] - x_ref)))
