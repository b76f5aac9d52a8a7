"""DC adjoint of the bundle (nsa_bundle_dc_adjoint) on the two-stage Miller
op-amp batch of the DC-adjoint configuration (8 BSIM4v7 transistors, RC
compensation, 10k / 1pF load, four input biases per instance):

  * the operating points (direct Newton, then gmin stepping for the
    instances that need it, in MODEDCOP) against the reference's same
    ladder (1e-9 relative, 1e-12 absolute), and those of direct Newton
    against the reference's solve_dc too (solve_dc tries true gmin stepping
    before gmin stepping, dc.cpp:118-150, which this build does not
    restate: its instances end at an operating point equal to within the
    Newton tolerance only);
  * the packed transposed solve: J^T lambda = g with J the reference's own
    load at the bundle's operating point (a residual check; bit-equal to the
    reference NeoSolver's solve_transposed on a linear circuit, whose load is
    exact);
  * dL/dW, dL/dL, dL/dR, dL/dC of L = sum over the biases of (Vout - 0.5)^2
    against central finite differences of the operating points (h = 1e-6
    relative) at 1e-5 relative plus the differences' own noise (measured by
    the second difference), at operating points taken to roundoff so that
    the differences are not Newton noise; dL/dC is zero in DC, and so is
    dL/dRc (no DC current in the compensation branch);
  * dVout/dR and dVout/dV against the reference's .sens (its forward
    differences, step 1e-4 |p|)."""
import numpy as np
import pytest
import scipy.sparse as sp

import neospice_amd as na
import refbind as RB
from neospice_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.ref]

MODEDCOP, MODEINITFLOAT = 0x10, 0x100
TIGHT = dict(reltol=1e-11, vntol=1e-14, abstol=1e-17)


def opamp(p, vin):
    c, info = W.opamp_miller(p["w"], p["l"], rc=p["r"][0], rl=p["r"][1], rb=p["r"][2], cc=p["c"][0],
                             cl=p["c"][1], vin=vin)
    c.finalize()
    return c, info


def batch(inst, biases=W.OPAMP_BIASES):
    cks, infos = [], []
    for p in inst:
        for v in biases:
            c, info = opamp(p, v)
            cks.append(c)
            infos.append(info)
    return cks, infos


def test_opamp_dc_matches_solve_dc():
    inst = W.opamp_instances(8)
    cks, infos = batch(inst)
    b = na.Bundle(cks)
    conv, its, stage = b.op_ladder(dcop=True)
    print("stages", np.bincount(stage + 1), "iterations", its.min(), its.max(), flush=True)
    assert conv.all()
    for i, c in enumerate(cks):
        x = b.solution(i)
        x_lad, _, st_ref = RB.RefCircuit(c, probe=False).dc_ladder(dcop=True)
        assert st_ref == stage[i], (i, st_ref, stage[i])
        tol = 1e-9 * np.maximum(np.abs(x_lad), 1e-3) + 1e-12
        assert (np.abs(x - x_lad) <= tol).all(), (i, float(np.max(np.abs(x - x_lad))))
        if stage[i] == 0:
            x_dc, _, _ = RB.RefCircuit(c, probe=False).dc()
            assert (np.abs(x - x_dc) <= tol).all(), (i, float(np.max(np.abs(x - x_dc))))


def test_transposed_packed_solve():
    inst = W.opamp_instances(2)
    cks, _ = batch(inst, biases=(0.4, 0.5))
    mesh, _ = W.rc_mesh(5)
    mesh.finalize()
    cks.append(mesh)
    b = na.Bundle(cks)
    conv, _, _ = b.op_ladder(dcop=True)
    assert conv.all()
    g = np.random.default_rng(7).standard_normal(b.total_n)
    lam = b.dc_adjoint(g)["lambda"]
    for i, c in enumerate(cks):
        sl = slice(b.xoff[i], b.xoff[i + 1])
        x, li, gi = b.solution(i), lam[sl], g[sl]
        rc = RB.RefCircuit(c, probe=False)
        rc.set_ctx(MODEDCOP | MODEINITFLOAT)
        vals, _ = rc.load(x)
        J = sp.csc_matrix((vals, np.asarray(c.row_idx), np.asarray(c.col_ptr)), shape=(c.n, c.n))
        res = J.T @ li - gi
        scale = abs(J).T @ np.abs(li) + np.abs(gi)
        print("instance", i, "max |J^T lam - g| / scale", float(np.max(np.abs(res) / scale)), flush=True)
        assert (np.abs(res) <= 1e-9 * scale).all(), i
        if c.is_linear:
            rs = RB.RefSolver(0, c.n, c.col_ptr, c.row_idx)
            assert not rs.numeric(vals)
            assert np.array_equal(li, rs.solve_transposed(gi.copy()))


def params_of(p):
    """The differentiable parameters of an instance: (kind, index, value)."""
    out = [("w", k, p["w"][k]) for k in range(8)] + [("l", k, p["l"][k]) for k in range(8)]
    out += [("r", k, p["r"][k]) for k in range(3)] + [("c", k, p["c"][k]) for k in range(2)]
    return out


def with_param(p, kind, k, v):
    q = {key: np.array(val, dtype=float) for key, val in p.items()}
    q[kind][k] = v
    return q


def loss_and_grad(inst):
    """L_i = sum_b (Vout_ib - 0.5)^2 and its adjoint gradient per parameter."""
    cks, infos = batch(inst)
    b = na.Bundle(cks, **TIGHT)
    conv, _, _ = b.op_ladder(dcop=True)
    assert conv.all()
    # two more Newton solves from the operating point: each ends one Newton
    # step further (the first iteration never converges), which takes the
    # points to roundoff, so that differences of h = 1e-6 |p| are not
    # Newton noise (the adjoint uses the factors of the last step)
    for _ in range(2):
        conv, _ = b.dc_op(MODEDCOP | MODEINITFLOAT, x0=b.solution())
        assert conv.all()
    nb = len(W.OPAMP_BIASES)
    dLdx = np.zeros(b.total_n)
    L = np.zeros(len(inst))
    for j, info in enumerate(infos):
        vout = b.solution(j)[info["out"]]
        L[j // nb] += (vout - 0.5) ** 2
        dLdx[b.xoff[j] + info["out"]] = 2.0 * (vout - 0.5)
    g = b.dc_adjoint(dLdx)
    doff = b.device_offsets()
    grads = []
    for i, p in enumerate(inst):
        gi = {}
        for kind, k, _ in params_of(p):
            tot = 0.0
            for j in range(i * nb, (i + 1) * nb):
                info = infos[j]
                if kind in ("w", "l"):
                    tot += g[kind][doff[j] + info["mos"][k]]
                elif kind == "r":
                    tot += g["value"][doff[j] + info["res"][k]]
                else:
                    tot += g["value"][doff[j] + info["cap"][k]]
            gi[(kind, k)] = tot
        grads.append(gi)
    return L, grads


def test_opamp_gradient_vs_central_fd():
    inst = W.opamp_instances(3)[1:]
    L0, grads = loss_and_grad(inst)
    # every perturbed instance (+h and -h per parameter) in one bundle
    pert, keys = [], []
    for i, p in enumerate(inst):
        for kind, k, v in params_of(p):
            h = 1e-6 * abs(v)
            pert += [with_param(p, kind, k, v + h), with_param(p, kind, k, v - h)]
            keys.append((i, kind, k, h))
    Lp, _ = loss_and_grad(pert)
    # the differences' own noise: the operating points carry roundoff
    # (~1e-13 of L for this high-gain loop).  The second differences
    # L(p+h) + L(p-h) - 2 L(p) sample it (their curvature term h^2 L'' is
    # below it at h = 1e-6 |p|); their largest over an instance's parameters
    # bounds the noise of that instance's loss evaluations.
    second = np.array([abs(Lp[2 * t] + Lp[2 * t + 1] - 2.0 * L0[i]) for t, (i, _, _, _) in enumerate(keys)])
    sigma = {i: max(second[[t for t, key in enumerate(keys) if key[0] == i]].max(), 1e-15 * abs(L0[i]))
             for i in range(len(inst))}
    worst = 0.0
    for t, (i, kind, k, h) in enumerate(keys):
        fd = (Lp[2 * t] - Lp[2 * t + 1]) / (2 * h)
        ad = grads[i][(kind, k)]
        if kind == "c":
            assert ad == 0.0
            assert abs(fd) <= 1e-6 * abs(L0[i]) / h
            continue
        if kind == "r" and k == 0:
            # rc is in series with cc: no DC current flows through it
            assert ad == 0.0 and abs(fd) <= 2.0 * sigma[i] / h
            continue
        noise = 2.0 * sigma[i] / h
        scale = max(abs(ad), abs(fd))
        err = abs(ad - fd) / scale if scale > 0 else 0.0
        worst = max(worst, err)
        print("instance", i, kind, k, "adjoint %.10e fd %.10e rel %.2e fd noise %.1e" % (ad, fd, err, noise),
              flush=True)
        assert abs(ad - fd) <= 1e-5 * scale + noise, (i, kind, k, ad, fd, noise)
    print("worst relative difference", worst)


def test_opamp_adjoint_vs_sens():
    inst = W.opamp_instances(2)
    cks, infos = batch(inst, biases=(0.45,))
    b = na.Bundle(cks)
    conv, _, _ = b.op_ladder(dcop=True)
    assert conv.all()
    doff = b.device_offsets()
    for j, (c, info) in enumerate(zip(cks, infos)):
        dLdx = np.zeros(b.total_n)
        dLdx[b.xoff[j] + info["out"]] = 1.0
        g = b.dc_adjoint(dLdx)
        sens = RB.RefCircuit(c, probe=False).sens(info["out"] + 1)
        checked = 0
        for d, op in enumerate(c.ops_devices()):
            kind, name = op[0], op[1].lower()
            key = {"R": "resistance", "V": "dc", "I": "dc"}.get(kind)
            if key is None:
                continue
            ref = sens[name + ":" + key]
            ad = g["value"][doff[j] + d]
            # .sens is a forward difference with step 1e-4 |p| between two
            # operating points solved at the default tolerances (vntol 1e-6
            # V): its truncation error (~1e-4 relative) plus the two
            # points' Newton residue, taken as 1e-8 V, over the step
            step = 1e-4 * abs(c.ops_devices()[d][-1])
            tol = 2e-3 * abs(ref) + 1e-8 / step
            print(name, "adjoint %.6e sens %.6e tol %.1e" % (ad, ref, tol), flush=True)
            assert abs(ad - ref) <= tol, (name, ad, ref)
            checked += 1
        assert checked == 5


def test_set_device_values_and_autograd():
    """New W, L and R values set on a built bundle give the operating points
    of circuits built with those values (bit for bit), and the torch front
    end returns the adjoint's gradients."""
    import torch
    from neospice_amd.autograd import BundleDCOperatingPoint
    inst = W.opamp_instances(2)
    cks, infos = batch(inst, biases=(0.4, 0.55))
    b = na.Bundle(cks)
    value, w, l = b.device_values()
    doff = b.device_offsets()
    # scale every transistor's W by 1.1, L by 0.9 and the load resistors by 1.05
    q = [dict(p) for p in inst]
    for p in q:
        p["w"] = p["w"] * 1.1
        p["l"] = p["l"] * 0.9
        p["r"] = p["r"] * 1.05
    cks2, _ = batch(q, biases=(0.4, 0.55))
    w2, l2, value2 = w * 1.1, l * 0.9, value.copy()
    for j, info in enumerate(infos):
        for k in info["res"]:
            value2[doff[j] + k] *= 1.05
    b.set_device_values(value2, w2, l2)
    conv, _, _ = b.op_ladder(dcop=True)
    ref = na.Bundle(cks2)
    conv2, _, _ = ref.op_ladder(dcop=True)
    assert conv.all() and conv2.all()
    assert np.array_equal(b.solution(), ref.solution())
    v_now, w_now, l_now = b.device_values()
    assert np.array_equal(v_now, value2) and np.array_equal(w_now, w2) and np.array_equal(l_now, l2)
    # torch: L = sum (Vout - 0.5)^2 over the circuits
    tv, tw, tl = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in (value2, w2, l2))
    x = BundleDCOperatingPoint.apply(b, tv, tw, tl)
    outs = torch.tensor([int(b.xoff[j] + info["out"]) for j, info in enumerate(infos)])
    loss = ((x[outs] - 0.5) ** 2).sum()
    loss.backward()
    dLdx = np.zeros(b.total_n)
    xs = b.solution()
    for j, o in enumerate(outs.tolist()):
        dLdx[o] = 2.0 * (xs[o] - 0.5)
    g = b.dc_adjoint(dLdx)
    assert np.array_equal(tw.grad.numpy(), g["w"]) and np.array_equal(tl.grad.numpy(), g["l"])
    assert np.array_equal(tv.grad.numpy(), g["value"])
    assert np.abs(tw.grad.numpy()).max() > 0
