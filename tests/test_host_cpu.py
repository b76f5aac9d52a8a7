"""Host-side logic of the library against the compiled reference (CPU only):
the Markowitz and AMD-LU pivot order and fill that the GPU replays, the
circuit's MNA pattern, load order and branch numbering, and the device
loads' matrix/rhs values on the host-built tables.  Marked `ref`: needs
oracle/_ref (built from the reference sources by oracle/Makefile)."""
import numpy as np
import pytest

import neospice_amd as na
import refbind as RB
from helpers import mesh, random_sparse

pytestmark = pytest.mark.ref

RANDOM = [(2, 80), (3, 40), (5, 40), (6, 20), (1, 20), (11, 60), (13, 120)]


def matrices():
    out = []
    for seed, n in RANDOM:
        cp, ri, v = random_sparse(seed, n, diag_frac=0.95)
        out.append(("rand%d_%d" % (seed, n), n, cp, ri, v))
    for m in (4, 8, 16):
        ckt, info = mesh(m)
        rc = RB.RefCircuit(ckt, probe=False)
        dt = 1e-12
        rc.set_ctx(na.MODETRAN | na.MODEINITPRED, 1, 0, (1 / dt, -1 / dt, 0), dt, dt, 0.0, dt)
        A, _ = rc.load(np.zeros(ckt.n))
        out.append(("mesh%d" % m, ckt.n, ckt.col_ptr, ckt.row_idx, A))
    return out


MATS = matrices()


@pytest.mark.parametrize("case", MATS, ids=lambda c: c[0])
def test_markowitz_order_and_fill_match_reference(case):
    """Sparse 1.3 ordering (mna_preorder, add_diag_gmin, order_and_factor):
    internal row/column translation and every column list, bit-exact."""
    name, n, cp, ri, v = case
    rs = RB.RefSolver(0, n, cp, ri)
    if rs.numeric(v):
        assert na.pivot_structure("markowitz", n, cp, ri, v) is None
        return
    d = rs.neo_dump()
    st = na.pivot_structure("markowitz", n, cp, ri, v)
    assert st is not None
    assert np.array_equal(st["row_perm"] + 1, d["i2e_row"])
    assert np.array_equal(st["col_perm"] + 1, d["i2e_col"])
    assert np.array_equal(st["seg_ptr"], d["col_ptr"])
    assert np.array_equal(st["rows"] + 1, d["rows"])


@pytest.mark.parametrize("case", MATS, ids=lambda c: c[0])
def test_amdlu_order_and_fill_match_reference(case):
    """Minimum-degree order and Gilbert-Peierls pivots/fill: q, the row of
    each pivot step, and per column the U rows in reach order and the L rows
    in storage order, as the reference's AmdLuSolver holds them."""
    name, n, cp, ri, v = case
    ra = RB.RefSolver(1, n, cp, ri)
    if ra.numeric(v):
        assert na.pivot_structure("amdlu", n, cp, ri, v) is None
        return
    d = ra.amd_dump()
    st = na.pivot_structure("amdlu", n, cp, ri, v)
    assert st is not None
    assert np.array_equal(st["col_perm"], d["q"])
    pinv = d["pinv"]
    row_perm = np.empty(n, np.int32)
    row_perm[pinv] = np.arange(n, dtype=np.int32)
    assert np.array_equal(st["row_perm"], row_perm)
    seg, nU, rows = st["seg_ptr"], st["nU"], st["rows"]
    for k in range(n):
        u_ref = d["Ui"][d["Up"][k]:d["Up"][k + 1]]
        assert u_ref[-1] == k                      # U keeps its diagonal last
        l_ref = d["Li"][d["Lp"][k]:d["Lp"][k + 1]]
        assert l_ref[0] == k                       # unit L diagonal first
        assert np.array_equal(rows[seg[k]:seg[k] + nU[k]], u_ref[:-1])
        assert rows[seg[k] + nU[k]] == k
        assert np.array_equal(rows[seg[k] + nU[k] + 1:seg[k + 1]], l_ref[1:])


@pytest.mark.parametrize("m", [4, 8, 16])
def test_circuit_pattern_load_order_and_branches(m):
    """Circuit::finalize: MNA pattern (CSC), device load order (ngspice
    load ranks, newest first within a rank) and branch numbering."""
    ckt, info = mesh(m)
    rc = RB.RefCircuit(ckt, probe=False)
    cp, ri = rc.pattern()
    assert np.array_equal(cp, ckt.col_ptr) and np.array_equal(ri, ckt.row_idx)
    assert np.array_equal(rc.load_order(), ckt.load_order)


def test_diode_and_mos_circuit_pattern():
    """Nonlinear families: pattern and load order of a small circuit with a
    diode and MOS1 devices match the reference's finalize."""
    c = na.Circuit()
    c.model(".model dn D(IS=1e-14 N=1)")
    c.model(".model nm NMOS(LEVEL=1 VTO=0.5 KP=120u LAMBDA=0.02)")
    c.model(".model pm PMOS(LEVEL=1 VTO=-0.5 KP=40u LAMBDA=0.02)")
    vdd, a, b, o = (c.node(x) for x in ("vdd", "a", "b", "o"))
    c.V("V1", vdd, na.GND, 1.8)
    c.V_pulse("V2", a, na.GND, 0.0, 1.8, 0.0, 50e-12, 50e-12, 0.45e-9, 1e-9)
    c.M("M1", o, a, na.GND, na.GND, "nm", 2e-6, 0.18e-6)
    c.M("M2", o, a, vdd, vdd, "pm", 1e-6, 0.18e-6)
    c.R("R1", o, b, 1e3)
    c.C("C1", b, na.GND, 1e-15)
    c.D("D1", b, na.GND, "dn")
    c.finalize()
    rc = RB.RefCircuit(c, probe=False)
    cp, ri = rc.pattern()
    assert np.array_equal(cp, c.col_ptr) and np.array_equal(ri, c.row_idx)
    assert np.array_equal(rc.load_order(), c.load_order)


def test_min_degree_order_many_dense_nodes():
    """More than 16 dense nodes (degree > max(10 sqrt(n), 16)) with tied
    degrees: the dense tail keeps the reference's std::sort order
    (amd.cpp:134-137), not a stable one."""
    import ctypes as C
    n = 400
    rng = np.random.default_rng(3)
    dense = rng.choice(n, 24, replace=False)
    cols = [set([j]) for j in range(n)]
    for d in dense:
        # every dense node touches the same number of others: tied degrees
        others = rng.choice(np.setdiff1d(np.arange(n), [d]), 260, replace=False)
        for o in others:
            cols[d].add(int(o))
            cols[int(o)].add(int(d))
    for j in range(n - 1):
        cols[j].add(j + 1)
        cols[j + 1].add(j)
    cp = np.zeros(n + 1, np.int32)
    ri = []
    for j in range(n):
        r = sorted(cols[j])
        ri.extend(r)
        cp[j + 1] = len(ri)
    ri = np.asarray(ri, np.int32)
    q_ref = np.zeros(n, np.int32)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    assert RB.ref().ref_amd_ordering(n, P(cp), P(ri), P(q_ref)) == n
    q = na.min_degree_order(n, cp, ri)
    assert np.array_equal(q, q_ref)


def complex_cases():
    """Complex matrices on the same patterns: g + j w c with c a second set
    of values on the pattern, at a low, a middle and a high frequency (the
    imaginary parts dominating moves the pivots)."""
    out = []
    rng = np.random.default_rng(7)
    for name, n, cp, ri, v in MATS:
        c = np.abs(np.asarray(v)) * rng.uniform(0.5, 2.0, len(v)) * 1e-12
        for w in (1e3, 1e12, 1e16):
            out.append(("%s_w%.0e" % (name, w), n, cp, ri, np.asarray(v) + 1j * w * c))
    return out


CMATS = complex_cases()


@pytest.mark.parametrize("case", CMATS, ids=lambda c: c[0])
def test_complex_markowitz_order_and_fill_match_reference(case):
    """NeoSolver::numeric_complex's ordering (neo_solver.cpp:228-245):
    pivots chosen on |re| + |im| with complex eliminations
    (factor.cpp:769-804); the internal row/column translation and every
    column list, bit-exact, and singular where the reference's ordering
    meets a singular matrix."""
    name, n, cp, ri, z = case
    rs = RB.RefSolver(0, n, cp, ri)
    if rs.numeric_complex(z):
        assert na.pivot_structure("markowitz", n, cp, ri, z) is None
        return
    d = rs.neo_dump()
    st = na.pivot_structure("markowitz", n, cp, ri, z)
    assert st is not None
    assert np.array_equal(st["row_perm"] + 1, d["i2e_row"])
    assert np.array_equal(st["col_perm"] + 1, d["i2e_col"])
    assert np.array_equal(st["seg_ptr"], d["col_ptr"])
    assert np.array_equal(st["rows"] + 1, d["rows"])
