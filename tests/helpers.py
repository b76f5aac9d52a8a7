"""Shared builders for the tests."""
import numpy as np

import neospice_amd as na
from neospice_amd import workloads as W


def csc_from_entries(n, ent):
    ent = sorted(set(ent), key=lambda e: (e[1], e[0]))
    cp = np.zeros(n + 1, np.int32)
    for _, c in ent:
        cp[c + 1] += 1
    cp = np.cumsum(cp).astype(np.int32)
    ri = np.array([r for r, _ in ent], np.int32)
    return cp, ri


def random_sparse(seed, n, density=3, twins=0.2, small_diag=0.3, diag_frac=0.8):
    """Unsymmetric sparse test matrix with MNA-like +-1 couplings and some
    tiny diagonals (forces off-diagonal / Markowitz pivoting)."""
    rng = np.random.default_rng(seed)
    ent = set((i, i) for i in range(n) if rng.random() < diag_frac)
    for _ in range(density * n):
        ent.add((int(rng.integers(n)), int(rng.integers(n))))
    cp, ri = csc_from_entries(n, ent)
    cols = np.repeat(np.arange(n), np.diff(cp))
    vals = rng.standard_normal(len(ri))
    for k in range(len(ri)):
        if ri[k] != cols[k] and rng.random() < twins:
            vals[k] = rng.choice([1.0, -1.0])
        if ri[k] == cols[k] and rng.random() < small_diag:
            vals[k] *= 1e-5
    return cp, ri, vals


def mesh(m, config_index=1):
    ckt, info = W.rc_mesh(m, config_index=config_index)
    ckt.finalize()
    return ckt, info


def dense(n, cp, ri, vals):
    A = np.zeros((n, n))
    for c in range(n):
        for p in range(cp[c], cp[c + 1]):
            A[ri[p], c] += vals[p]
    return A


# ---------------------------------------------------------------------------
# Exact oracle for linear R/C/V/I circuits: the same discrete equations
# assembled densely in numpy and differentiated by the complex step
# (dL/dp = Im L(p + i h) / h, no subtraction, so every entry is exact to
# roundoff, including the ones below the finite-difference noise floor).


def pwl_value(t, ts, vs):
    """VSource::value_at for PWL (vsource.cpp:80-155)."""
    if t <= ts[0]:
        return vs[0]
    if t >= ts[-1]:
        return vs[-1]
    hi = int(np.searchsorted(ts, t, side="left"))
    t0, t1, v0, v1 = ts[hi - 1], ts[hi], vs[hi - 1], vs[hi]
    return v1 if t1 == t0 else v0 + (v1 - v0) * ((t - t0) / (t1 - t0))


def dense_mna(ckt, R, Cv, V, I, t=0.0):
    """G (n x n), capacitor incidence P (n x nC) and the rhs b at time t of
    a linear R/C/V/I circuit (V dc or PWL; dc values from V / I)."""
    n = ckt.n
    G = np.zeros((n, n), dtype=complex)
    b = np.zeros(n, dtype=complex)
    P = np.zeros((n, len(ckt.family["C"])))
    pos = {f: {int(d): k for k, d in enumerate(ckt.family[f])} for f in ("R", "C", "V", "I")}
    for d, op in enumerate(ckt.ops_devices()):
        kind, a, c = op[0], op[2], op[3]
        if kind == "R":
            g = 1.0 / R[pos["R"][d]]
            for (r, s_), sg in (((a, a), 1), ((a, c), -1), ((c, a), -1), ((c, c), 1)):
                if r >= 0 and s_ >= 0:
                    G[r, s_] += sg * g
        elif kind == "C":
            k = pos["C"][d]
            if a >= 0:
                P[a, k] = 1.0
            if c >= 0:
                P[c, k] = -1.0
        elif kind in ("V", "Vpwl"):
            br = ckt.branch[d]
            for node, sg in ((a, 1), (c, -1)):
                if node >= 0:
                    G[node, br] += sg
                    G[br, node] += sg
            b[br] += V[pos["V"][d]] if kind == "V" else pwl_value(t, op[4], op[5])
        elif kind == "I":
            val = I[pos["I"][d]]
            if a >= 0:
                b[a] -= val
            if c >= 0:
                b[c] += val
        else:
            raise ValueError("dense_mna: linear R/C/V/I circuits only")
    return G, P, b


def dense_transient(ckt, R, Cv, V, I, nsteps, dt, probes, xmu=0.5):
    """The engine's fixed-step scheme (backward Euler first, then
    trapezoidal, capacitor companion i_k = a0 q_k + a1 q_{k-1} - r i_{k-1})
    in dense complex arithmetic; probe values (nsteps+1) x nprobe."""
    G, P, b = dense_mna(ckt, R, Cv, V, I, 0.0)
    Cd = np.diag(Cv)
    x = np.linalg.solve(G, b)
    out = [x[probes]]
    i_prev = np.zeros(len(Cv), dtype=complex)
    for k in range(1, nsteps + 1):
        if k == 1:
            a0, a1, r = 1.0 / dt, -1.0 / dt, 0.0
        else:
            a0, a1, r = 1.0 / (dt * (1 - xmu)), -1.0 / (dt * (1 - xmu)), xmu / (1 - xmu)
        _, _, b = dense_mna(ckt, R, Cv, V, I, k * dt)
        J = G + a0 * P @ Cd @ P.T
        rhs = b - P @ (a1 * Cd @ (P.T @ x) - r * i_prev)
        xn = np.linalg.solve(J, rhs)
        i_prev = a0 * Cd @ (P.T @ xn) + a1 * Cd @ (P.T @ x) - r * i_prev
        x = xn
        out.append(x[probes])
    return np.array(out)


def complex_step(fn, base, rows, h=1e-20):
    out = []
    for k in rows:
        p = base.astype(complex)
        p[k] += 1j * h * abs(base[k])
        out.append(fn(p).imag / (h * abs(base[k])))
    return np.array(out)
