"""The drop-in boundary end to end: the reference's own analyses with
neospice_amd behind its ISolver seam, against the unmodified reference.

oracle/_ref/libneospice_ref_gpu.so is the reference's src/ linked with
make_solver wrapped (oracle/solver_wrap.cpp): switched on, every solver the
reference's DC and transient analyses ask for is the Mi355xSolver adapter
(neospice_amd/integration/mi355x_solver.hpp) over nsa_lu_*.  The
reference's solve_dc, solve_transient and newton_solve then run unchanged,
with every factorization and solve on the GPU.  Simulator::run_dc and
run_transient are these two functions behind finalize and a save filter
(src/api/neospice.cpp:72-92).  Every result must equal the reference's own
CPU solvers' bit for bit: iteration counts, operating points, accepted time
points and waveforms.
"""
import numpy as np
import pytest

import refbind as RB
from neospice_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.ref]


def circuits():
    out = []
    mesh, _ = W.rc_mesh(20)                       # 401 unknowns, linear: AMD-LU (make_solver.cpp:169)
    out.append(("rc_mesh_20x20", mesh, 1e-12, 40e-12))
    chain, _ = W.inverter_chain(6)                # MOS1 + PULSE: Markowitz
    out.append(("mos1_chain_6", chain, 10e-12, 400e-12))
    dm, _ = W.diode_mesh(4, t_stop=100e-12)       # diodes + PWL
    out.append(("diode_mesh_4", dm, 2e-12, 100e-12))
    rings, _ = W.c3_rings(stages=[9])             # BSIM4v7
    out.append(("bsim4_ring_9", rings[0], 1e-12, 100e-12))
    for _, c, _, _ in out:
        c.finalize()
    return out


@pytest.fixture(scope="module")
def gpu_lib():
    if not RB.gpu_available():
        pytest.fail("oracle/_ref/libneospice_ref_gpu.so is missing: build it (python __graft_entry__.py build)")
    return RB.ref_gpu()


@pytest.mark.parametrize("case", range(4))
def test_reference_analyses_through_mi355x_solver(gpu_lib, case):
    name, ckt, tstep, tstop = circuits()[case]
    cpu = RB.RefCircuit(ckt, probe=False)
    x_cpu, it_cpu, _ = cpu.dc()
    t_cpu, w_cpu, nt_cpu, _ = cpu.tran(tstep, tstop)
    made0 = gpu_lib.ref_gpu_solvers_made()
    gpu_lib.ref_set_gpu_solver(1)
    try:
        g = RB.RefCircuit(ckt, probe=False, gpu=True)
        x_gpu, it_gpu, _ = g.dc()
        t_gpu, w_gpu, nt_gpu, _ = g.tran(tstep, tstop)
    finally:
        gpu_lib.ref_set_gpu_solver(0)
    made = gpu_lib.ref_gpu_solvers_made() - made0
    print(name, "n", ckt.n, "dc iterations", it_cpu, "points", len(t_cpu), "newton", nt_cpu,
          "mi355x solvers", made, flush=True)
    # one for solve_dc, two for solve_transient (its operating point and its steps)
    assert made == 3, made
    assert it_gpu == it_cpu
    assert np.array_equal(x_gpu, x_cpu), float(np.max(np.abs(x_gpu - x_cpu)))
    assert nt_gpu == nt_cpu
    assert np.array_equal(t_gpu, t_cpu)
    assert np.array_equal(w_gpu, w_cpu), float(np.max(np.abs(w_gpu - w_cpu)))


def test_newton_solve_through_mi355x_solver(gpu_lib):
    """newton_solve alone (newton.cpp:51-419) in MODEDCOP|MODEINITJCT from
    zeros, on the BSIM4v7 ring: the same iterate sequence's end."""
    _, ckt, _, _ = circuits()[3]
    cpu = RB.RefCircuit(ckt, probe=False)
    x0 = np.zeros(cpu.n)
    mode = 0x10 | 0x200
    it_cpu = cpu.L.ref_newton(cpu.h, x0.ctypes.data_as(RB.P_F64), mode)
    gpu_lib.ref_set_gpu_solver(1)
    try:
        g = RB.RefCircuit(ckt, probe=False, gpu=True)
        x1 = np.zeros(g.n)
        it_gpu = g.L.ref_newton(g.h, x1.ctypes.data_as(RB.P_F64), mode)
    finally:
        gpu_lib.ref_set_gpu_solver(0)
    assert it_gpu == it_cpu
    assert np.array_equal(x1, x0)


def test_mi355x_solver_complex_paths(gpu_lib):
    """The adapter's complex paths as the AC analysis drives a solver
    (ac.cpp:215-222): numeric_complex then solve_complex, refactorize_complex
    at another frequency then solve_complex, through Mi355xSolver (compiled
    against the reference's headers, linked into the reference) -- bit-equal
    to the reference's own NeoSolver on the AC system g + j w c of a
    145-unknown RC mesh."""
    import ctypes as C
    import neospice_amd as na
    mesh, _ = W.rc_mesh(12)
    mesh.finalize()
    rc = RB.RefCircuit(mesh, probe=False)
    x, _, _ = rc.dc()
    g, _ = rc.load(x)                                     # DC conductances
    dt = 1e-12
    rc.set_ctx(na.MODETRAN | na.MODEINITPRED, 1, 0, (1 / dt, -1 / dt, 0), dt, dt, 0.0, dt)
    a1, _ = rc.load(x)
    c = (np.asarray(a1) - np.asarray(g)) * dt             # capacitances on the pattern
    cp, ri = rc.pattern()
    n = mesh.n
    rng = np.random.default_rng(1)
    rhs = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z0 = np.asarray(g) + 1j * (2 * np.pi * 1e6) * c
    z1 = np.asarray(g) + 1j * (2 * np.pi * 1e10) * c

    def inter(z):
        out = np.empty(2 * len(z))
        out[0::2], out[1::2] = z.real, z.imag
        return np.ascontiguousarray(out)

    L = gpu_lib
    L.ref_mi355x_complex.restype = C.c_int
    L.ref_mi355x_complex.argtypes = [C.c_int] + [C.c_void_p] * 7 + [C.c_char_p]
    cpa = np.ascontiguousarray(cp, np.int32)
    ria = np.ascontiguousarray(ri, np.int32)
    a0, a1_, r = inter(z0), inter(z1), inter(rhs)
    x0 = np.zeros(2 * n)
    x1 = np.zeros(2 * n)
    err = C.create_string_buffer(256)
    rcode = L.ref_mi355x_complex(n, cpa.ctypes.data, ria.ctypes.data, a0.ctypes.data, a1_.ctypes.data,
                                 r.ctypes.data, x0.ctypes.data, x1.ctypes.data, err)
    assert rcode >= 0, err.value.decode()
    ref = RB.RefSolver(0, n, cp, ri)
    assert not ref.numeric_complex(z0)
    y0 = ref.solve_complex(rhs)
    assert bool(rcode) == ref.refactorize_complex(z1)
    y1 = ref.solve_complex(rhs)
    assert np.array_equal(x0[0::2] + 1j * x0[1::2], y0)
    assert np.array_equal(x1[0::2] + 1j * x1[1::2], y1)
