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
