"""Pins the oracle (oracle/oracle.c, the plain-C restatement of the
reference's numeric LU path) before it is trusted as the checker:

  * the reference's own known-answer tests (tests/golden/
    solver_known_answers.json, from test_neo_solver.cpp / test_amd_lu.cpp /
    test_amd.cpp): the restated Sparse-1.3 refactor+solve on the host
    library's Markowitz order, and the restated AMD-LU pipeline, recover the
    published solutions within the published tolerances, and the replay's
    growth check rejects the unstable case;
  * mathematics: dense solves of random matrices, the transposed-solve
    identity <A^T y, x> = <y, A x>, and a planted mistake that the dense
    check must catch;
  * the compiled reference (marked `ref`): bit-exact factor values and
    solutions of both restated solvers.

CPU only."""
import json
import os

import numpy as np
import pytest

import neospice_amd as na
import refbind as RB
from helpers import csc_from_entries, dense, random_sparse

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "solver_known_answers.json")))


def csc_values(n, entries):
    """CSC pattern and slot values of an entry list (duplicates add)."""
    acc = {}
    for r, c, v in entries:
        acc[(r, c)] = acc.get((r, c), 0.0) + v
    cp, ri = csc_from_entries(n, list(acc))
    cols = np.repeat(np.arange(n), np.diff(cp))
    vals = np.array([acc[(int(r), int(c))] for r, c in zip(ri, cols)])
    return cp, ri, vals


def s13_oracle(n, cp, ri, vals, gmin=0.0):
    """Oracle Sparse-1.3 on the host library's Markowitz pivot order."""
    st = na.pivot_structure("markowitz", n, cp, ri, vals, gmin)
    assert st is not None, "host ordering reports singular"
    o = RB.OracleS13.from_structure(st)
    assert o.refactor(cp, ri, vals, gmin) == 0
    return o


@pytest.mark.parametrize("case", GOLD["cases"], ids=lambda c: c["name"])
def test_known_answers(case):
    n, tol, gmin = case["n"], case["tol"], case["gmin"]
    cp, ri, v0 = csc_values(n, case["entries"])
    final = case["refactor"] if case["refactor"] is not None else case["entries"]
    _, _, v1 = csc_values(n, final)
    x_true = np.array(case["x"], dtype=float)
    b = np.array(case["b"], dtype=float)
    if "neo" in case["solvers"]:
        o = s13_oracle(n, cp, ri, v0, gmin)
        if case["refactor"] is not None:
            assert o.refactor(cp, ri, v1, gmin) == 0
        np.testing.assert_allclose(o.solve(b), x_true, rtol=0, atol=tol)
    if "amd" in case["solvers"]:
        g = RB.OracleGP(n, cp, ri, v0, gmin)
        assert not g.singular
        if case["refactor"] is not None:
            accepted = g.refactor(v1, gmin)
            if case["name"] == "amd_refactor_fallback_3":
                # test_amd_lu.cpp:356-400: the growth check must reject the
                # replay; the reference then re-pivots with a fresh factor
                assert not accepted
                g = RB.OracleGP(n, cp, ri, v1, gmin)
                assert not g.singular
            else:
                assert accepted
        np.testing.assert_allclose(g.solve(b), x_true, rtol=0, atol=tol)


@pytest.mark.parametrize("case", GOLD["amd_order"], ids=lambda c: c["name"])
def test_min_degree_order_properties(case):
    """test_amd.cpp:8-70: a valid permutation; the arrowhead hub goes last.
    Both the oracle and the host library's ordering (which must agree)."""
    n = case["n"]
    cp, ri = csc_from_entries(n, [tuple(e) for e in case["pattern"]])
    q = np.zeros(n, np.int32)
    RB.orc().orc_amd_order(n, cp.ctypes.data_as(RB.P_I32), ri.ctypes.data_as(RB.P_I32),
                           q.ctypes.data_as(RB.P_I32))
    assert sorted(q.tolist()) == list(range(n))
    if case["check"] == "hub_last":
        assert int(np.nonzero(q == 0)[0][0]) >= n - 2
    assert np.array_equal(na.min_degree_order(n, cp, ri), q)


def test_singular_detection():
    """test_amd_lu.cpp:442-453: an empty column is structurally singular."""
    cp, ri, v = csc_values(2, [[0, 0, 1.0], [1, 0, 1.0]])
    assert RB.OracleGP(2, cp, ri, v).singular
    assert na.pivot_structure("markowitz", 2, cp, ri, v) is None


RANDOM = [(2, 80), (3, 40), (5, 40), (6, 20), (1, 20), (11, 60)]


@pytest.mark.parametrize("seed,n", RANDOM)
def test_dense_solve_and_transpose_identity(seed, n):
    cp, ri, vals = random_sparse(seed, n, diag_frac=0.95)
    A = dense(n, cp, ri, vals)
    cond = np.linalg.cond(A)
    if cond > 1e10:
        pytest.skip("ill-conditioned draw")
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(n)
    y = rng.standard_normal(n)
    x_np = np.linalg.solve(A, b)
    xt_np = np.linalg.solve(A.T, b)
    # threshold pivoting (tolerance 1e-3) trades some growth for sparsity:
    # the error bound is cond * eps * growth, checked loosely here; a dropped
    # or wrong term misses by orders of magnitude more (planted below)
    rel = max(1e-6, 1e-13 * cond)
    tol = rel * max(1.0, np.abs(x_np).max())
    o = s13_oracle(n, cp, ri, vals)
    g = RB.OracleGP(n, cp, ri, vals)
    assert not g.singular
    for solve, solve_t in ((o.solve, o.solve_t), (g.solve, g.solve_t)):
        x = solve(b)
        xt = solve_t(b)
        assert np.abs(x - x_np).max() < tol
        assert np.abs(xt - xt_np).max() < rel * max(1.0, np.abs(xt_np).max())
        # <A^T u, v> = <u, A v> with u = A^-T y, v = A^-1 b:  <y, v> = <u, b>
        u = solve_t(y)
        assert abs(y @ x - u @ b) < rel * (np.abs(y).sum() * np.abs(x).max() + 1)


def test_planted_mistake_is_caught():
    """The dense check above detects a dropped update term: solving with one
    off-diagonal entry of A zeroed (what dropping its update would amount to)
    misses the dense solution by far more than the tolerance."""
    n = 40
    cp, ri, vals = random_sparse(3, n, diag_frac=0.95)
    A = dense(n, cp, ri, vals)
    b = np.random.default_rng(0).standard_normal(n)
    x_np = np.linalg.solve(A, b)
    cols = np.repeat(np.arange(n), np.diff(cp))
    k = int(np.nonzero(ri != cols)[0][0])
    bad = vals.copy()
    bad[k] = 0.0
    o = s13_oracle(n, cp, ri, bad)
    assert np.abs(o.solve(b) - x_np).max() > 1e-6


@pytest.mark.ref
@pytest.mark.parametrize("seed,n", RANDOM)
def test_oracle_bit_exact_vs_compiled_reference(seed, n):
    """Restated Sparse-1.3 and AMD-LU match the compiled reference bit for
    bit: factor values, solutions and (Sparse 1.3) transposed solutions."""
    cp, ri, vals = random_sparse(seed, n, diag_frac=0.95)
    b = np.random.default_rng(seed + 1).standard_normal(n)
    rs = RB.RefSolver(0, n, cp, ri)
    if rs.numeric(vals):
        pytest.skip("singular in the reference")
    d = rs.neo_dump()
    o = RB.OracleS13(n, d["i2e_row"], d["i2e_col"], d["col_ptr"], d["rows"])
    assert o.refactor(cp, ri, vals) == 0
    assert np.array_equal(o.vals, d["vals"])
    assert np.array_equal(o.solve(b), rs.solve(b))
    assert np.array_equal(o.solve_t(b), rs.solve_transposed(b))
    # refactorize on perturbed values
    v2 = vals * (1 + 1e-3 * np.random.default_rng(seed + 2).standard_normal(len(vals)))
    if not rs.refactorize(v2):
        assert o.refactor(cp, ri, v2) == 0
        assert np.array_equal(o.solve(b), rs.solve(b))
    ra = RB.RefSolver(1, n, cp, ri)
    if ra.numeric(vals):
        pytest.skip("AMD-LU singular in the reference")
    g = RB.OracleGP(n, cp, ri, vals)
    da = ra.amd_dump()
    assert np.array_equal(g.q, da["q"]) and np.array_equal(g.pinv, da["pinv"])
    lnz, unz = da["Lp"][-1], da["Up"][-1]
    assert np.array_equal(g.Lp, da["Lp"]) and np.array_equal(g.Up, da["Up"])
    assert np.array_equal(g.Li[:lnz], da["Li"]) and np.array_equal(g.Ui[:unz], da["Ui"])
    assert np.array_equal(g.Lx[:lnz], da["Lx"]) and np.array_equal(g.Ux[:unz], da["Ux"])
    assert np.array_equal(g.solve(b), ra.solve(b))
