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
