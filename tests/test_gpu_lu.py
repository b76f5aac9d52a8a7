"""GPU parity of the solver seam (nsa_lu_*) against the compiled reference.

Refactor and solves are compared bit for bit: the kernels perform the
reference's own operation sequence per value, without FMA contraction.
"""
import numpy as np
import pytest

import neospice_amd as na
import refbind as RB
from helpers import mesh, random_sparse, dense

pytestmark = [pytest.mark.gpu, pytest.mark.ref]

KIND = {"markowitz": 0, "amdlu": 1}


def mesh_matrix(m, seed=0):
    """A transient-step Jacobian of the RC mesh (capacitor companion on)."""
    ckt, _ = mesh(m)
    rc = RB.RefCircuit(ckt)
    dt = 1e-12
    rc.set_ctx(0x1 | 0x2000, 1, 0, (1 / dt, -1 / dt, 0), dt, dt, 0.0, dt)
    rng = np.random.default_rng(seed)
    vals, rhs = rc.load(rng.uniform(0.9, 1.0, ckt.n))
    return ckt.n, ckt.col_ptr, ckt.row_idx, vals, rhs


CASES = [("mesh8",) + mesh_matrix(8)[:4], ("mesh16",) + mesh_matrix(16)[:4]] + [
    ("rand%d_%d" % (s, n),) + (n,) + random_sparse(s, n, diag_frac=0.95)
    for s, n in ((2, 80), (3, 40), (5, 40), (6, 20), (1, 20))
]


def nonsingular(case, engine):
    name, n, cp, ri, vals = case
    return not RB.RefSolver(KIND[engine], n, cp, ri).numeric(vals)


@pytest.mark.parametrize("engine", ["markowitz", "amdlu"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_numeric_solve_bit_exact(case, engine):
    name, n, cp, ri, vals = case
    rs = RB.RefSolver(KIND[engine], n, cp, ri)
    if rs.numeric(vals):
        pytest.skip("singular in the reference")
    s = na.Solver(n, cp, ri, batch=1, engine=engine)
    assert not s.numeric(vals).any()
    b = np.random.default_rng(7).standard_normal(n)
    x_ref = rs.solve(b)
    x = s.solve(b.reshape(n, 1))[:, 0]
    assert np.array_equal(x, x_ref), np.max(np.abs(x - x_ref))
    if engine == "markowitz":
        xt_ref = rs.solve_transposed(b)
        xt = s.solve_transposed(b.reshape(n, 1))[:, 0]
        assert np.array_equal(xt, xt_ref)
    else:
        # The reference has no transposed AMD-LU solve: compare with the
        # oracle's restated U^T / L^T substitution on the reference's own
        # factors (the oracle itself is pinned against dense A^T solves in
        # tests/test_oracle.py).
        d = rs.amd_dump()
        xt_orc = b.copy()
        P = RB.P_I32
        RB.orc().orc_gp_solve_t(n, *[d[k].ctypes.data_as(P if d[k].dtype == np.int32 else RB.P_F64)
                                     for k in ("q", "pinv", "Lp", "Li", "Lx", "Up", "Ui", "Ux")],
                                xt_orc.ctypes.data_as(RB.P_F64))
        xt = s.solve_transposed(b.reshape(n, 1))[:, 0]
        assert np.array_equal(xt, xt_orc), np.max(np.abs(xt - xt_orc))


@pytest.mark.parametrize("engine", ["markowitz", "amdlu"])
@pytest.mark.parametrize("case", CASES[:3], ids=lambda c: c[0])
def test_refactorize_batched(case, engine):
    """B instances with perturbed values: refactorize on the fixed order,
    each instance equal to the reference refactorize of its own values."""
    name, n, cp, ri, vals = case
    rs = RB.RefSolver(KIND[engine], n, cp, ri)
    if rs.numeric(vals):
        pytest.skip("singular in the reference")
    B = 5
    rng = np.random.default_rng(3)
    V = np.stack([vals * (1 + 1e-3 * rng.standard_normal(len(vals))) for _ in range(B)], axis=1)
    s = na.Solver(n, cp, ri, batch=B, engine=engine)
    assert not s.numeric(np.repeat(vals[:, None], B, axis=1)).any()
    # instances whose re-pivot changes the order move to a plan of their own
    st = s.refactorize(V)
    b = rng.standard_normal((n, B))
    x = s.solve(b)
    for i in range(B):
        r2 = RB.RefSolver(KIND[engine], n, cp, ri)
        assert not r2.numeric(vals)
        sing = r2.refactorize(V[:, i])
        assert bool(st[i]) == sing
        if sing:
            continue
        assert np.array_equal(x[:, i], r2.solve(b[:, i]))


def test_pivot_order_matches_reference():
    """The GPU plan's pivot order/fill is the reference's (Markowitz)."""
    name, n, cp, ri, vals = CASES[2]   # rand2_80: off-diagonal and Markowitz pivots
    rs = RB.RefSolver(0, n, cp, ri)
    assert not rs.numeric(vals)
    d = rs.neo_dump()
    s = na.Solver(n, cp, ri, batch=1, engine="markowitz")
    s.numeric(vals)
    st = s.structure()
    assert np.array_equal(st["row_perm"], d["i2e_row"] - 1)
    assert np.array_equal(st["col_perm"], d["i2e_col"] - 1)
    assert np.array_equal(st["seg_ptr"], d["col_ptr"])
    assert np.array_equal(st["rows"], d["rows"] - 1)


@pytest.mark.parametrize("engine", ["markowitz", "amdlu"])
def test_refactorize_batched_diverging_orders(engine):
    """One instance of a batch whose diagonal collapses: its reused order
    fails the threshold (Markowitz, factor.cpp:929-946) or the replay's
    growth check (AMD-LU, amd_lu_solver.cpp:214-271), so it re-pivots on
    the CPU and leaves the batch's shared plan for one of its own; every
    instance -- moved or not -- equals the reference refactorize of its own
    values, and a later solve uses each instance's own order."""
    name, n, cp, ri, vals = CASES[2]
    rs = RB.RefSolver(KIND[engine], n, cp, ri)
    if rs.numeric(vals):
        pytest.skip("singular in the reference")
    B = 4
    rng = np.random.default_rng(11)
    V = np.stack([vals * (1 + 1e-3 * rng.standard_normal(len(vals))) for _ in range(B)], axis=1)
    rows = np.asarray(ri)
    cols = np.repeat(np.arange(n), np.diff(np.asarray(cp)))
    diag = rows == cols
    V[diag, 2] *= 1e-9
    s = na.Solver(n, cp, ri, batch=B, engine=engine)
    assert not s.numeric(np.repeat(vals[:, None], B, axis=1)).any()
    st = s.refactorize(V)
    b = rng.standard_normal((n, B))
    x = s.solve(b)
    moved = 0
    for i in range(B):
        r2 = RB.RefSolver(KIND[engine], n, cp, ri)
        assert not r2.numeric(vals)
        sing = r2.refactorize(V[:, i])
        assert bool(st[i]) == sing, i
        if sing:
            continue
        assert np.array_equal(x[:, i], r2.solve(b[:, i])), i
        if i == 2:
            moved = 1
    assert moved


@pytest.mark.parametrize("engine", ["markowitz", "amdlu"])
@pytest.mark.parametrize("case", [CASES[1], CASES[2]], ids=lambda c: c[0])
@pytest.mark.parametrize("gmin", [0.0, 1e-9])
def test_equilibrate_matches_reference(case, engine, gmin):
    """ISolver::set_equilibrate: the radix-2 row/column scaling NeoSolver
    applies as the DC ladder's last fallback (neo_solver.cpp:52-102,
    204-225), on matrices whose rows and columns span twelve decades.
    numeric, refactorize and solve on a batch of three instances, each
    bit-equal to the reference solver of its own values (AMD-LU ignores the
    switch, as AmdLuSolver does)."""
    name, n, cp, ri, vals = case
    rng = np.random.default_rng(5)
    rows = np.asarray(ri)
    cols = np.repeat(np.arange(n), np.diff(np.asarray(cp)))
    B = 3
    V0 = np.empty((len(vals), B))
    V1 = np.empty((len(vals), B))
    for i in range(B):
        rs_ = 10.0 ** rng.uniform(-6, 6, n)
        cs_ = 10.0 ** rng.uniform(-6, 6, n)
        V0[:, i] = vals * rs_[rows] * cs_[cols]
        V1[:, i] = V0[:, i] * (1 + 1e-3 * rng.standard_normal(len(vals)))
    s = na.Solver(n, cp, ri, batch=B, engine=engine)
    s.set_equilibrate(True)
    st0 = s.numeric(V0, gmin)
    b = rng.standard_normal((n, B))
    x0 = s.solve(b)
    st1 = s.refactorize(V1, gmin)
    x1 = s.solve(b)
    for i in range(B):
        r = RB.RefSolver(KIND[engine], n, cp, ri)
        r.set_equilibrate(True)
        sing0 = r.numeric(V0[:, i], gmin)
        assert bool(st0[i]) == sing0, i
        if sing0:
            continue
        assert np.array_equal(x0[:, i], r.solve(b[:, i])), (i, np.max(np.abs(x0[:, i] - r.solve(b[:, i]))))
        sing1 = r.refactorize(V1[:, i], gmin)
        assert bool(st1[i]) == sing1, i
        if not sing1:
            assert np.array_equal(x1[:, i], r.solve(b[:, i])), i


def ac_pair(case, seed=9):
    """g (the case's values) and c (capacitance-like values on the same
    pattern) of an AC system g + j w c."""
    name, n, cp, ri, vals = case
    rng = np.random.default_rng(seed)
    c = np.abs(np.asarray(vals)) * rng.uniform(0.5, 2.0, len(vals)) * 1e-12
    return np.asarray(vals, np.float64), c


@pytest.mark.parametrize("engine", ["markowitz", "amdlu"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_complex_paths_bit_exact(case, engine):
    """ISolver's complex paths on a batch of three instances, each at its
    own frequency: numeric_complex (the complex Markowitz ordering on the
    host, the complex factor on the GPU), solve_complex, refactorize_complex
    at new frequencies, and a second numeric_complex (the reuse path):
    every instance's solutions bit-equal to a reference solver of its own
    values (AmdLuSolver delegates these to NeoSolver, so both engines
    must equal it)."""
    name, n, cp, ri, vals = case
    g, c = ac_pair(case)
    rng = np.random.default_rng(3)
    B = 3
    W0 = (1e3, 1e9, 1e13)
    W1 = (2e3, 3e10, 5e15)
    Z0 = np.stack([g + 1j * w * c for w in W0], axis=1)
    Z1 = np.stack([g * (1 + 1e-3 * rng.standard_normal(len(g))) + 1j * w * c for w in W1], axis=1)
    b = rng.standard_normal((n, B)) + 1j * rng.standard_normal((n, B))
    s = na.Solver(n, cp, ri, batch=B, engine=engine)
    st0 = s.numeric_complex(Z0)
    x0 = s.solve_complex(b)
    st1 = s.refactorize_complex(Z1)
    x1 = s.solve_complex(b)
    st2 = s.numeric_complex(Z1)
    x2 = s.solve_complex(b)
    checked = 0
    for i in range(B):
        r = RB.RefSolver(KIND[engine], n, cp, ri)
        sing0 = r.numeric_complex(Z0[:, i])
        assert bool(st0[i]) == sing0, i
        if sing0:
            continue
        assert np.array_equal(x0[:, i], r.solve_complex(b[:, i])), (i, np.max(np.abs(x0[:, i] - r.solve_complex(b[:, i]))))
        sing1 = r.refactorize_complex(Z1[:, i])
        assert bool(st1[i]) == sing1, i
        if not sing1:
            assert np.array_equal(x1[:, i], r.solve_complex(b[:, i])), i
        r2 = RB.RefSolver(KIND[engine], n, cp, ri)
        r2.numeric_complex(Z0[:, i])
        assert bool(st2[i]) == r2.numeric_complex(Z1[:, i]), i
        if not st2[i]:
            assert np.array_equal(x2[:, i], r2.solve_complex(b[:, i])), i
        checked += 1
    if checked == 0:
        pytest.skip("singular in the reference")


@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[2]], ids=lambda c: c[0])
def test_ac_sweep_matches_reference(case):
    """The AC analysis's sweep (ac.cpp:196-238) for two instances and 24
    frequencies in one launch: each instance ordered at the first frequency,
    every frequency factored on that order and solved -- bit-equal to the
    reference's own loop (numeric_complex at f0, then refactorize_complex
    and solve_complex per frequency, on a NeoSolver as ac.cpp makes)."""
    name, n, cp, ri, vals = case
    g, c = ac_pair(case)
    rng = np.random.default_rng(4)
    B = 2
    G = np.stack([g, g * (1 + 1e-2 * rng.standard_normal(len(g)))], axis=1)
    Cm = np.stack([c, c * (1 + 1e-2 * rng.standard_normal(len(c)))], axis=1)
    freqs = np.logspace(2, 14, 24)
    rhs = rng.standard_normal((n, B)) + 1j * rng.standard_normal((n, B))
    s = na.Solver(n, cp, ri, batch=B)
    x, sing = s.ac_sweep(G, Cm, freqs, rhs)
    assert x.shape == (24, n, B)
    for i in range(B):
        r = RB.RefSolver(0, n, cp, ri)
        for f, fr in enumerate(freqs):
            w = 2.0 * np.pi * fr
            ax = G[:, i] + 1j * (w * Cm[:, i])
            if f == 0:
                sing0 = r.numeric_complex(ax)
                assert bool(sing[0, i]) == sing0
                if sing0:
                    break
            else:
                assert bool(sing[f, i]) == r.refactorize_complex(ax), (i, f)
            xr = r.solve_complex(rhs[:, i])
            if not sing[f, i]:
                assert np.array_equal(x[f, :, i], xr), (i, f, np.max(np.abs(x[f, :, i] - xr)))
