"""Synthetic workloads of the benchmark configurations, generated with
deterministic seeds (PCG64, seed = 20261203 + config index; per-instance
streams from SeedSequence.spawn) rather than netlist random functions.
"""
import numpy as np

from . import GND, Circuit

BASE_SEED = 20261203


def rng(config_index):
    return np.random.Generator(np.random.PCG64(BASE_SEED + config_index))


def instance_rngs(config_index, count):
    ss = np.random.SeedSequence(BASE_SEED + config_index)
    return [np.random.Generator(np.random.PCG64(s)) for s in ss.spawn(count)]


def rc_mesh(m=64, config_index=1, r_range=(0.1, 10.0), c_range=(10e-15, 1e-12),
            sink_frac=0.05, sink_max=1e-3, vdd=1.0):
    """Linear power-grid mesh: m x m nodes, a resistor on every grid edge
    (log-uniform in r_range), a decap per node to ground (log-uniform in
    c_range), current sinks on sink_frac of the nodes (uniform [0, sink_max])
    and one voltage source at node (0, 0).

    Returns (circuit, info) where info holds the nominal resistor values and
    device indices used for per-instance perturbation."""
    g = rng(config_index)
    ckt = Circuit()
    nid = [[ckt.node("n%d_%d" % (i, j)) for j in range(m)] for i in range(m)]
    res_devs, res_vals = [], []
    for i in range(m):
        for j in range(m):
            for di, dj, tag in ((0, 1, "h"), (1, 0, "v")):
                ii, jj = i + di, j + dj
                if ii < m and jj < m:
                    r = float(np.exp(g.uniform(np.log(r_range[0]), np.log(r_range[1]))))
                    res_devs.append(ckt.R("r%s%d_%d" % (tag, i, j), nid[i][j], nid[ii][jj], r))
                    res_vals.append(r)
    cap_devs = []
    for i in range(m):
        for j in range(m):
            c = float(np.exp(g.uniform(np.log(c_range[0]), np.log(c_range[1]))))
            cap_devs.append(ckt.C("c%d_%d" % (i, j), nid[i][j], GND, c))
    nsink = int(round(sink_frac * m * m))
    sinks = g.choice(m * m, size=nsink, replace=False)
    sink_devs = []
    for k in sorted(sinks.tolist()):
        i, j = divmod(k, m)
        sink_devs.append(ckt.I("i%d_%d" % (i, j), nid[i][j], GND, float(g.uniform(0.0, sink_max))))
    ckt.V("vdd", nid[0][0], GND, vdd)
    info = {"res_devs": np.array(res_devs), "res_vals": np.array(res_vals),
            "cap_devs": np.array(cap_devs), "sink_devs": np.array(sink_devs), "m": m}
    return ckt, info


def mesh_instance_resistors(info, batch, config_index=1, sigma=0.10):
    """Per-instance Gaussian perturbation R * (1 + sigma * N(0, 1)), shape
    (num_resistors, batch)."""
    rs = instance_rngs(config_index, batch)
    nominal = info["res_vals"]
    out = np.empty((len(nominal), batch))
    for b, r in enumerate(rs):
        out[:, b] = nominal * (1.0 + sigma * r.standard_normal(len(nominal)))
    return out


def diode_mesh(m=32, config_index=5, r_range=(1.0, 100.0), c_range=(10e-15, 100e-15),
               n_pwl=16, v_max=1.2, t_stop=1e-9, diodes=True):
    """Nonlinear mesh: m x m nodes, a resistor on every grid edge
    (log-uniform in r_range), a capacitor and a diode (IS=1e-14, N=1,
    anode at the node, cathode at ground) per node, and a PWL voltage
    source at node (0, 0) with n_pwl seeded breakpoints in [0, v_max] over
    [0, t_stop]."""
    g = rng(config_index)
    ckt = Circuit()
    ckt.model(".model dmesh D(IS=1e-14 N=1)")
    nid = [[ckt.node("n%d_%d" % (i, j)) for j in range(m)] for i in range(m)]
    res_devs, res_vals = [], []
    for i in range(m):
        for j in range(m):
            for di, dj, tag in ((0, 1, "h"), (1, 0, "v")):
                ii, jj = i + di, j + dj
                if ii < m and jj < m:
                    r = float(np.exp(g.uniform(np.log(r_range[0]), np.log(r_range[1]))))
                    res_devs.append(ckt.R("r%s%d_%d" % (tag, i, j), nid[i][j], nid[ii][jj], r))
                    res_vals.append(r)
    for i in range(m):
        for j in range(m):
            c = float(np.exp(g.uniform(np.log(c_range[0]), np.log(c_range[1]))))
            ckt.C("c%d_%d" % (i, j), nid[i][j], GND, c)
            if diodes:
                ckt.D("d%d_%d" % (i, j), nid[i][j], GND, "dmesh")
    t = np.sort(g.uniform(0.0, t_stop, n_pwl - 1))
    t = np.concatenate([[0.0], t])
    v = g.uniform(0.0, v_max, n_pwl)
    v[0] = 0.0
    ckt.V_pwl("vin", nid[0][0], GND, t.tolist(), v.tolist())
    info = {"res_devs": np.array(res_devs), "res_vals": np.array(res_vals), "m": m,
            "pwl_t": t, "pwl_v": v}
    return ckt, info


def inverter_chain(stages=50, vdd=1.8, r_load=1e3, c_load=2e-15):
    """MOS1 inverter chain: LEVEL=1 NMOS/PMOS (VTO +-0.5, KP 120u/40u,
    LAMBDA 0.02, W/L 2u/0.18u and 1u/0.18u), a PULSE input 0 -> vdd (50 ps
    edges, 1 ns period) and,
    per stage, an RC load (series r_load to a node with c_load to ground)
    that drives the next stage: 2 nodes per stage, 2 + 2*stages + 2
    unknowns."""
    ckt = Circuit()
    ckt.model(".model nch NMOS(LEVEL=1 VTO=0.5 KP=120u LAMBDA=0.02)")
    ckt.model(".model pch PMOS(LEVEL=1 VTO=-0.5 KP=40u LAMBDA=0.02)")
    nvdd, nin = ckt.node("vdd"), ckt.node("in")
    ckt.V("vsup", nvdd, GND, vdd)
    ckt.V_pulse("vin", nin, GND, 0.0, vdd, 0.0, 50e-12, 50e-12, 0.45e-9, 1e-9)
    drive = nin
    for k in range(stages):
        o = ckt.node("o%d" % k)
        l = ckt.node("l%d" % k)
        ckt.M("mn%d" % k, o, drive, GND, GND, "nch", 2e-6, 0.18e-6)
        ckt.M("mp%d" % k, o, drive, nvdd, nvdd, "pch", 1e-6, 0.18e-6)
        ckt.R("rl%d" % k, o, l, r_load)
        ckt.C("cl%d" % k, l, GND, c_load)
        drive = l
    return ckt, {"stages": stages}


def ring_oscillator(stages=31, vdd=1.0, wn=0.5e-6, wp=1.0e-6, l=0.05e-6, toxe=1.8e-9, vth0=0.4,
                    mismatch_rng=None, sigma_vth0=0.03, sigma_toxe=0.01, kick=True):
    """BSIM4v7 ring oscillator: `stages` (odd) CMOS inverters in a loop,
    NMOS W/L = wn/l and PMOS wp/l on the default LEVEL=14 card with TOXE and
    VTH0 +-vth0, VDD = vdd.  With `mismatch_rng`, every transistor gets its
    own card with Gaussian relative mismatch of sigma_vth0 on VTH0 and
    sigma_toxe on TOXE.  A ring's only operating point is metastable (Newton
    does not find it for long rings, in the reference either), so `kick`
    adds an enable transistor that holds stage 0's input low until 20 ps:
    the operating point is then a settled chain, and the ring oscillates
    once the enable releases.  Unknowns: stages + vdd (+ enable node) +
    source branches."""
    ckt = Circuit()
    shared = mismatch_rng is None
    if shared:
        ckt.model(".model nmod nmos level=14 toxe=%.17g vth0=%.17g" % (toxe, vth0))
        ckt.model(".model pmod pmos level=14 toxe=%.17g vth0=%.17g" % (toxe, -vth0))
    nvdd = ckt.node("vdd")
    ckt.V("vsup", nvdd, GND, vdd)
    nodes = [ckt.node("n%d" % k) for k in range(stages)]
    for k in range(stages):
        a, o = nodes[k], nodes[(k + 1) % stages]
        names = []
        for pol, sign in (("n", 1.0), ("p", -1.0)):
            if shared:
                names.append(pol + "mod")
            else:
                dv = 1.0 + sigma_vth0 * mismatch_rng.standard_normal()
                dt = 1.0 + sigma_toxe * mismatch_rng.standard_normal()
                nm = "%s%d" % (pol, k)
                ckt.model(".model %s %smos level=14 toxe=%.17g vth0=%.17g" %
                          (nm, pol, toxe * dt, sign * vth0 * dv))
                names.append(nm)
        ckt.M("mn%d" % k, o, a, GND, GND, names[0], wn, l)
        ckt.M("mp%d" % k, o, a, nvdd, nvdd, names[1], wp, l)
    if kick:
        # enable: a wide NMOS holds stage 0's input low until its gate falls
        # (t = 20 ps), which gives the ring a well-defined operating point
        ne = ckt.node("en")
        ckt.V_pulse("ven", ne, GND, vdd, 0.0, 20e-12, 10e-12, 10e-12, 1.0, 2.0)
        if shared:
            ckt.M("mrst", nodes[0], ne, GND, GND, "nmod", 8 * wn, l)
        else:
            ckt.model(".model nrst nmos level=14 toxe=%.17g vth0=%.17g" % (toxe, vth0))
            ckt.M("mrst", nodes[0], ne, GND, GND, "nrst", 8 * wn, l)
    return ckt, {"stages": stages, "nodes": nodes}


def c3_stage_counts(count, config_index=3, lo=31, hi=201):
    """Stage counts of the uneven-size ring batch: uniform over the odd
    numbers lo..hi (seeded)."""
    odd = np.arange(lo, hi + 1, 2)
    return rng(config_index).choice(odd, size=count).astype(int)


def c3_rings(count=256, config_index=3, stages=None, **kw):
    """The uneven-size BSIM4v7 ring-oscillator batch: `count` rings with
    stage counts from c3_stage_counts (or `stages`), each with its own
    per-transistor mismatch (3 % VTH0, 1 % TOXE) from its own seeded stream.
    Returns (circuits, stage counts)."""
    if stages is None:
        stages = c3_stage_counts(count, config_index)
    streams = instance_rngs(config_index, len(stages))
    out = []
    for s, r in zip(stages, streams):
        ckt, _ = ring_oscillator(int(s), mismatch_rng=r, **kw)
        out.append(ckt)
    return out, np.asarray(stages)


# Two-stage Miller op-amp of the DC-adjoint configuration: transistor order
# M1..M8 (inputs, mirror load, tail, second stage, output sink, bias).
OPAMP_DEVICES = ("m1", "m2", "m3", "m4", "m5", "m6", "m7", "m8")
OPAMP_NMOS = (True, True, False, False, True, False, True, True)


def opamp_miller(w, l, rc=1e3, cc=0.5e-12, rl=10e3, cl=1e-12, rb=20e3, vin=0.5, vdd=1.0,
                 toxe=1.8e-9, vth0=0.4):
    """Two-stage Miller op-amp of 8 BSIM4v7 (LEVEL=14) transistors as a
    unity-gain buffer: NMOS input pair M1 (gate at the output) / M2 (gate at
    the input source `vin`), PMOS mirror load M3/M4, NMOS tail M5, PMOS
    common-source second stage M6 with NMOS sink M7, both current sources
    biased by the diode M8 fed through rb from VDD; Miller compensation rc
    in series with cc from the first-stage output to the output; load rl || cl.
    w, l: the eight transistor widths and lengths.  Returns (circuit, info)
    with the device indices of the transistors, resistors (rc, rl, rb) and
    capacitors (cc, cl) and the output node."""
    ckt = Circuit()
    ckt.model(".model nmod nmos level=14 toxe=%.17g vth0=%.17g" % (toxe, vth0))
    ckt.model(".model pmod pmos level=14 toxe=%.17g vth0=%.17g" % (toxe, -vth0))
    n = {k: ckt.node(k) for k in ("vdd", "inp", "out", "n1", "n2", "tail", "vb", "nc")}
    ckt.V("vsup", n["vdd"], GND, vdd)
    ckt.V("vin", n["inp"], GND, vin)
    pins = {  # d, g, s, b
        "m1": ("n1", "out", "tail", None), "m2": ("n2", "inp", "tail", None),
        "m3": ("n1", "n1", "vdd", "vdd"), "m4": ("n2", "n1", "vdd", "vdd"),
        "m5": ("tail", "vb", None, None), "m6": ("out", "n2", "vdd", "vdd"),
        "m7": ("out", "vb", None, None), "m8": ("vb", "vb", None, None),
    }
    node = lambda k: GND if k is None else n[k]
    mos = []
    for k, name in enumerate(OPAMP_DEVICES):
        dn, gn, sn, bn = pins[name]
        mos.append(ckt.M(name, node(dn), node(gn), node(sn), node(bn),
                         "nmod" if OPAMP_NMOS[k] else "pmod", float(w[k]), float(l[k])))
    res = [ckt.R("rc", n["n2"], n["nc"], rc), ckt.R("rl", n["out"], GND, rl),
           ckt.R("rb", n["vdd"], n["vb"], rb)]
    cap = [ckt.C("cc", n["nc"], n["out"], cc), ckt.C("cl", n["out"], GND, cl)]
    return ckt, {"mos": mos, "res": res, "cap": cap, "out": n["out"]}


OPAMP_BIASES = (0.3, 0.4, 0.5, 0.6)


def opamp_instances(count=1024, config_index=4):
    """Seeded op-amp sizes of the DC-adjoint batch: per instance eight W
    uniform in [0.2u, 5u], eight L uniform in [0.05u, 0.5u], and the
    resistors (rc, rl, rb) and capacitors (cc, cl) scaled by U[0.8, 1.2]."""
    out = []
    for r in instance_rngs(config_index, count):
        w = r.uniform(0.2e-6, 5e-6, 8)
        l = r.uniform(0.05e-6, 0.5e-6, 8)
        rs = np.array([1e3, 10e3, 20e3]) * r.uniform(0.8, 1.2, 3)
        cs = np.array([0.5e-12, 1e-12]) * r.uniform(0.8, 1.2, 2)
        out.append({"w": w, "l": l, "r": rs, "c": cs})
    return out
