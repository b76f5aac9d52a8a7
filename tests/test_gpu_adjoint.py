"""GPU adjoints of the batched engine:

  * DC: dL/dR, dL/dV, dL/dI against the reference's own .sens
    (sens.cpp:57-186, forward differences with a 1e-4 relative step, so
    agreement is to that method's truncation error) and against central
    finite differences (h = 1e-6 |p|) of the engine's own operating point
    to 1e-5 relative;
  * transient (fixed step, backward Euler then trapezoidal): the forward
    waveforms against the reference's fixed-step run, and dL/dR, dL/dC
    against central finite differences to 1e-5 relative, with checkpoint
    intervals that do and do not divide the step count;
  * the torch.autograd.Function front end.

Finite differences run with tightened Newton tolerances so that the
forward's convergence noise stays far below the perturbation."""
import numpy as np
import pytest

import neospice_amd as na
import refbind as RB
from helpers import complex_step, dense_mna, dense_transient, mesh
from neospice_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.ref]

TIGHT = dict(reltol=1e-12, vntol=1e-14, abstol=1e-17)


def test_dc_adjoint_matches_sens():
    ckt, info = mesh(5)
    out = ckt.n - 2  # a node far from the source (last node; the V branch is ckt.n - 1)
    eng = na.BatchEngine(ckt, batch=1)
    eng.dc_op()
    dLdx = np.zeros((ckt.n, 1))
    dLdx[out, 0] = 1.0
    g = eng.dc_adjoint(dLdx)
    sens = RB.RefCircuit(ckt, probe=False).sens(out + 1)
    checked = 0
    for fam, par in (("R", "resistance"), ("V", "dc"), ("I", "dc")):
        gf = g[fam][:, 0]
        scale = np.abs(gf).max()
        for k, d in enumerate(ckt.family[fam]):
            ref = sens[ckt.devices[d][1].lower() + ":" + par]
            # .sens is a forward difference: truncation ~1e-4 relative, plus a
            # roundoff floor (its step is 1e-4 |p|) relative to the largest entry
            assert abs(gf[k] - ref) <= 2e-3 * abs(ref) + 2e-5 * scale, (fam, d, gf[k], ref)
            checked += 1
    assert checked == len(sens)


def dc_loss(eng, w):
    x, _ = eng.dc_op()
    return (w * x).sum(axis=0)


def central_fd(eng, fam, rows, base, loss, h=1e-6):
    """Per-instance central differences of loss() w.r.t. rows of a family."""
    out = []
    for k in rows:
        p = base.copy()
        dp = h * np.abs(p[k])
        p[k] = base[k] + dp
        eng.set_family_values(fam, p)
        lp = loss()
        p[k] = base[k] - dp
        eng.set_family_values(fam, p)
        lm = loss()
        out.append((lp - lm) / (2 * dp))
    eng.set_family_values(fam, base)
    return np.array(out)


def test_dc_adjoint_central_fd():
    """Diode network with a dc drive: dL/dR, dL/dV, dL/dI against central
    differences.  (The linear DC adjoint is checked exactly by the complex
    step below; on the linear mesh, whose gradients span five decades,
    finite differences sit on their roundoff floor.)"""
    B = 2
    c = na.Circuit()
    c.model(".model dd D(IS=1e-14 N=1)")
    m = 4
    nid = [[c.node("n%d_%d" % (i, j)) for j in range(m)] for i in range(m)]
    rng = np.random.default_rng(3)
    for i in range(m):
        for j in range(m):
            if j + 1 < m:
                c.R("rh%d_%d" % (i, j), nid[i][j], nid[i][j + 1], rng.uniform(5, 50))
            if i + 1 < m:
                c.R("rv%d_%d" % (i, j), nid[i][j], nid[i + 1][j], rng.uniform(5, 50))
            c.D("d%d_%d" % (i, j), nid[i][j], na.GND, "dd")
    c.I("i0", nid[m - 1][m - 1], na.GND, 1e-3)
    c.V("v0", nid[0][0], na.GND, 0.75)
    ckt = c.finalize()
    eng = na.BatchEngine(ckt, batch=B, **TIGHT)
    R0 = eng.family_values("R")
    R0 = R0 * (1 + 0.05 * np.random.default_rng(1).standard_normal(R0.shape))
    eng.set_family_values("R", R0)
    w = np.random.default_rng(2).standard_normal((ckt.n, B))
    dc_loss(eng, w)
    g = eng.dc_adjoint(w)
    for fam, rows in (("R", [0, 3, 7, len(R0) - 1]), ("V", [0]), ("I", [0])):
        base = eng.family_values(fam)
        fd = central_fd(eng, fam, rows, base, lambda: dc_loss(eng, w))
        for j, k in enumerate(rows):
            # central differences at h = 1e-6 |p|: 1e-5 relative above their
            # roundoff floor (~1e-16 cond(J) |L| / h, about 1e-6 of the largest entry)
            np.testing.assert_allclose(g[fam][k], fd[j], rtol=1e-5,
                                       atol=1e-6 * np.abs(g[fam]).max(), err_msg=f"{fam}{k}")


def diode_mesh(m):
    ckt, info = W.diode_mesh(m, t_stop=60e-12)
    ckt.finalize()
    return ckt


def test_tran_forward_matches_reference():
    ckt = diode_mesh(5)
    nsteps, dt = 30, 2e-12
    probes = np.arange(ckt.num_nodes, dtype=np.int32)
    eng = na.BatchEngine(ckt, batch=2)
    out = eng.tran_forward(nsteps, dt, probes, ckpt_every=7)
    xs_ref, _, _ = RB.RefCircuit(ckt, probe=False).tran_steps(dt, nsteps)
    ref = xs_ref[:, :ckt.num_nodes]
    for i in range(2):
        err = np.abs(out[:, :, i] - ref) / np.maximum(np.abs(ref), 1e-3)
        assert err.max() <= 1e-6, err.max()


def tran_loss(eng, nsteps, dt, probes, target, ckpt):
    v = eng.tran_forward(nsteps, dt, probes, ckpt_every=ckpt)
    return ((v - target) ** 2).sum(axis=(0, 1)), 2 * (v - target)


@pytest.mark.parametrize("linear", [True, False])
def test_tran_adjoint_central_fd(linear):
    B = 2
    if linear:
        ckt, _ = W.diode_mesh(4, t_stop=50e-12, diodes=False)
        ckt.finalize()
        nsteps, dt = 20, 2e-12
    else:
        ckt = diode_mesh(4)
        nsteps, dt = 24, 2e-12
    probes = np.array([1, 5, ckt.num_nodes - 1], np.int32)
    eng = na.BatchEngine(ckt, batch=B, **TIGHT)
    C0 = eng.family_values("C") * (1 + 0.1 * np.random.default_rng(4).random((len(ckt.family["C"]), B)))
    eng.set_family_values("C", C0)
    v0 = eng.tran_forward(nsteps, dt, probes, ckpt_every=nsteps)
    target = v0 + 0.01 * np.random.default_rng(5).standard_normal(v0.shape)
    _, dLdv = tran_loss(eng, nsteps, dt, probes, target, 5)
    gR, gC = eng.tran_adjoint(dLdv)
    # the same gradient with every step stored (no recomputation)
    tran_loss(eng, nsteps, dt, probes, target, 1)
    gR1, gC1 = eng.tran_adjoint(dLdv)
    np.testing.assert_allclose(gR1, gR, rtol=1e-9, atol=1e-12 * np.abs(gR).max())
    np.testing.assert_allclose(gC1, gC, rtol=1e-9, atol=1e-12 * np.abs(gC).max())
    loss = lambda: tran_loss(eng, nsteps, dt, probes, target, 7)[0]
    for fam, g in (("R", gR), ("C", gC)):
        base = eng.family_values(fam)
        rows = [0, 2, len(base) // 2, len(base) - 1]
        fd = central_fd(eng, fam, rows, base, loss)
        for j, k in enumerate(rows):
            np.testing.assert_allclose(g[k], fd[j], rtol=1e-5, atol=1e-6 * np.abs(g).max(),
                                       err_msg=f"{fam}{k}")


def test_torch_autograd_dc():
    torch = pytest.importorskip("torch")
    from neospice_amd.autograd import DCOperatingPoint
    ckt, _ = mesh(4)
    B = 2
    eng = na.BatchEngine(ckt, batch=B)
    vals = {f: torch.tensor(eng.family_values(f), dtype=torch.float64, requires_grad=True)
            for f in ("R", "C", "V", "I")}
    x = DCOperatingPoint.apply(eng, vals["R"], vals["C"], vals["V"], vals["I"])
    loss = (x[:ckt.num_nodes] ** 2).sum()
    loss.backward()
    g = eng.dc_adjoint(2 * x.detach().numpy() * (np.arange(ckt.n) < ckt.num_nodes)[:, None])
    for f in ("R", "V", "I"):
        np.testing.assert_allclose(vals[f].grad.numpy(), g[f], rtol=1e-12, atol=0)
    assert not vals["C"].grad.abs().max().item()


def test_torch_autograd_transient():
    """TransientProbes: the torch front end returns the engine's own
    transient-adjoint gradients (same checkpointing) and its forward the
    engine's probe waveforms."""
    torch = pytest.importorskip("torch")
    from neospice_amd.autograd import TransientProbes
    ckt = diode_mesh(4)
    B, nsteps, dt, ckpt = 2, 20, 2e-12, 5
    probes = np.array([1, 6, ckt.num_nodes - 1], np.int32)
    eng = na.BatchEngine(ckt, batch=B)
    R = torch.tensor(eng.family_values("R"), dtype=torch.float64, requires_grad=True)
    C = torch.tensor(eng.family_values("C") * 1.05, dtype=torch.float64, requires_grad=True)
    v = TransientProbes.apply(eng, nsteps, dt, probes, ckpt, R, C)
    target = torch.full_like(v, 0.2)
    loss = ((v - target) ** 2).sum()
    loss.backward()
    v_np = v.detach().numpy()
    eng2 = na.BatchEngine(ckt, batch=B)
    eng2.set_family_values("C", C.detach().numpy())
    v2 = eng2.tran_forward(nsteps, dt, probes, ckpt_every=ckpt)
    assert np.array_equal(v_np, v2)
    gR, gC = eng2.tran_adjoint(2.0 * (v2 - 0.2))
    assert np.array_equal(R.grad.numpy(), gR) and np.array_equal(C.grad.numpy(), gC)
    assert np.abs(gR).max() > 0 and np.abs(gC).max() > 0


@pytest.mark.parametrize("which", ["dc", "tran"])
def test_adjoint_linear_complex_step(which):
    if which == "dc":
        ckt, _ = mesh(4)
    else:
        # an RC mesh driven by a PWL source: time constants around the step,
        # so the waveforms, and dL/dC, are far above roundoff
        ckt, _ = W.diode_mesh(4, t_stop=30e-12, diodes=False)
        ckt.finalize()
    B = 1
    eng = na.BatchEngine(ckt, batch=B)
    vals = {f: eng.family_values(f)[:, 0] for f in ("R", "C", "V", "I")}
    vals["V"] = np.array([ckt.ops_devices()[d][4] if ckt.devices[d][0] == "V" else 0.0
                          for d in ckt.family["V"]])
    nsteps, dt = 12, 2e-12
    probes = np.arange(ckt.num_nodes, dtype=np.int32)
    wts = np.random.default_rng(7).standard_normal((nsteps + 1, len(probes)))
    if which == "dc":
        w = np.random.default_rng(8).standard_normal(ckt.n)
        eng.dc_op()
        g = eng.dc_adjoint(w[:, None])

        def loss(f, p):
            v = dict(vals)
            v[f] = p
            G, _, b = dense_mna(ckt, v["R"], v["C"], v["V"], v["I"])
            return w @ np.linalg.solve(G, b)
        fams = ("R", "V", "I")
    else:
        eng.tran_forward(nsteps, dt, probes, ckpt_every=5)
        gR, gC = eng.tran_adjoint(wts[:, :, None])
        g = {"R": gR, "C": gC}

        def loss(f, p):
            v = dict(vals)
            v[f] = p
            return (wts * dense_transient(ckt, v["R"], v["C"], v["V"], v["I"], nsteps, dt,
                                          probes)).sum()
        fams = ("R", "C")
    for f in fams:
        rows = np.arange(len(vals[f]))
        cs = complex_step(lambda p: loss(f, p), vals[f], rows)
        scale = np.abs(cs).max()
        np.testing.assert_allclose(g[f][:, 0], cs, rtol=1e-9, atol=1e-11 * scale, err_msg=f)


@pytest.mark.parametrize("vin", [0.42, 0.36])
def test_dc_adjoint_bsim4_central_fd(vin):
    """BSIM4v7 inverters with resistive loads at a DC bias: dL/dR and dL/dV
    through the BSIM4v7 Jacobian against central differences.  The
    transposed solve uses the factors of the last Newton iteration, whose
    load ran at the returned iterate with junction and FET limiting in
    force (newton.cpp:378-381); at a converged point the limiting leaves the
    voltages alone, so those are the Jacobian at the solution.  The second
    bias sits just below threshold, where the limiting acts most during the
    solve."""
    B = 2
    c = na.Circuit()
    c.model(".model nb nmos level=14 toxe=1.8e-9 vth0=0.4")
    c.model(".model pb pmos level=14 toxe=1.8e-9 vth0=-0.4")
    vdd, nin = c.node("vdd"), c.node("in")
    c.V("vsup", vdd, na.GND, 1.0)
    c.V("vin", nin, na.GND, vin)
    drive = nin
    rng = np.random.default_rng(4)
    for k in range(3):
        o, l = c.node("o%d" % k), c.node("l%d" % k)
        c.M("mn%d" % k, o, drive, na.GND, na.GND, "nb", 0.5e-6, 0.05e-6)
        c.M("mp%d" % k, o, drive, vdd, vdd, "pb", 1e-6, 0.05e-6)
        c.R("rl%d" % k, o, l, rng.uniform(2e3, 8e3))
        c.R("rg%d" % k, l, na.GND, rng.uniform(2e4, 8e4))
        drive = l
    ckt = c.finalize()
    eng = na.BatchEngine(ckt, batch=B, **TIGHT)
    w = np.random.default_rng(5).standard_normal((ckt.n, B))
    dc_loss(eng, w)
    g = eng.dc_adjoint(w)
    for fam, rows in (("R", [0, 2, 5]), ("V", [0, 1])):
        base = eng.family_values(fam)
        fd = central_fd(eng, fam, rows, base, lambda: dc_loss(eng, w))
        for j, k in enumerate(rows):
            np.testing.assert_allclose(g[fam][k], fd[j], rtol=1e-5,
                                       atol=1e-6 * np.abs(g[fam]).max(), err_msg=f"{fam}{k}")


def test_tran_adjoint_refuses_bsim4():
    c = na.Circuit()
    c.model(".model nb nmos level=14 toxe=1.8e-9 vth0=0.4")
    a = c.node("a")
    c.V("va", a, na.GND, 0.5)
    c.M("m1", a, a, na.GND, na.GND, "nb", 0.5e-6, 0.05e-6)
    ckt = c.finalize()
    eng = na.BatchEngine(ckt, batch=1)
    from neospice_amd._lib import NsaError
    # no forward record yet: a misuse-order error
    with pytest.raises(NsaError) as e0:
        eng.tran_adjoint(np.zeros((4, 1, 1)))
    assert e0.value.code == -1
    # the forward run itself is fine; the adjoint refuses the BSIM4v7 charges
    eng.tran_forward(3, 1e-12, np.array([0]))
    with pytest.raises(NsaError) as e1:
        eng.tran_adjoint(np.zeros((4, 1, 1)))
    assert e1.value.code == -5 and "BSIM4v7" in str(e1.value)
