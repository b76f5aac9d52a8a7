"""ctypes wrapper of the compiled reference (oracle/_ref/libneospice_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, never by the product.  It replays a
neospice_amd.Circuit (the recorded builder calls) into the reference's own
Circuit through its typed API, so both sides see identical inputs; node k
of the product (0-based) is node "n<k+1>" of the reference, created in the
same order, so MNA indices coincide.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libneospice_ref.so")
# the same reference with neospice_amd behind its solver seam (solver_wrap.cpp)
REF_GPU_SO = os.path.join(HERE, "_ref", "libneospice_ref_gpu.so")
ORC_SO = os.path.join(HERE, "_ref", "liboracle.so")

P_I32 = C.POINTER(C.c_int32)
P_F64 = C.POINTER(C.c_double)
VP = C.c_void_p
_ref = None
_ref_gpu = None
_orc = None


def available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = _bind(C.CDLL(REF_SO))
    return _ref


def gpu_available():
    return os.path.exists(REF_GPU_SO)


def ref_gpu():
    """The reference whose make_solver can hand out the MI355X solver
    (ref_set_gpu_solver)."""
    global _ref_gpu
    if _ref_gpu is None:
        L = _bind(C.CDLL(REF_GPU_SO))
        L.ref_set_gpu_solver.restype = None
        L.ref_set_gpu_solver.argtypes = [C.c_int]
        L.ref_gpu_solvers_made.restype = C.c_int
        L.ref_gpu_solvers_made.argtypes = []
        _ref_gpu = L
    return _ref_gpu


def _bind(L):
    """Sets the result and argument types of the harness functions."""
    sig = {
        "ref_last_error": (C.c_char_p, []),
        "ref_ckt_new": (VP, [C.c_int]),
        "ref_ckt_free": (None, [VP]),
        "ref_ckt_model": (C.c_int, [VP, C.c_char_p]),
        "ref_ckt_r": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int, C.c_double]),
        "ref_ckt_c": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int, C.c_double]),
        "ref_ckt_v": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int, C.c_double]),
        "ref_ckt_i": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int, C.c_double]),
        "ref_ckt_v_pulse": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int] + [C.c_double] * 7),
        "ref_ckt_v_pwl": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int, C.c_int, P_F64, P_F64]),
        "ref_ckt_d": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int, C.c_char_p]),
        "ref_ckt_m": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                C.c_double, C.c_double]),
        "ref_bsim4_fields": (C.c_int, [VP, C.c_int, P_F64]),
        "ref_ckt_devnoncon": (C.c_int, [VP, P_F64]),
        "ref_ckt_ic": (C.c_int, [VP, C.c_int, C.c_double]),
        "ref_ckt_option": (C.c_int, [VP, C.c_char_p, C.c_double]),
        "ref_ckt_finalize": (C.c_int, [VP, C.c_int]),
        "ref_ckt_num_vars": (C.c_int, [VP]),
        "ref_ckt_num_nodes": (C.c_int, [VP]),
        "ref_ckt_num_states": (C.c_int, [VP]),
        "ref_ckt_nnz": (C.c_int, [VP]),
        "ref_ckt_is_linear": (C.c_int, [VP]),
        "ref_ckt_num_devices": (C.c_int, [VP]),
        "ref_ckt_pattern": (None, [VP, P_I32, P_I32]),
        "ref_ckt_load_order": (C.c_int, [VP, P_I32]),
        "ref_ckt_branch_index": (C.c_int, [VP, C.c_int]),
        "ref_ckt_dead_nodes": (C.c_int, [VP, P_I32]),
        "ref_ckt_state": (P_F64, [VP, C.c_int]),
        "ref_ckt_set_ctx": (None, [VP, C.c_int, C.c_int, C.c_int] + [C.c_double] * 7),
        "ref_ckt_load": (C.c_int, [VP, P_F64, P_F64, P_F64]),
        "ref_dc": (C.c_int, [VP, P_F64, P_F64]),
        "ref_tran": (C.c_int, [VP, C.c_double, C.c_double, C.c_int, P_F64, P_F64, P_F64,
                               P_I32]),
        "ref_tran_fixed": (C.c_int, [VP, C.c_double, C.c_int, P_F64, P_F64]),
        "ref_tran_num_records": (C.c_int, [VP]),
        "ref_tran_records": (None, [VP, P_I32, P_F64]),
        "ref_newton": (C.c_int, [VP, P_F64, C.c_int]),
        "ref_sens": (C.c_int, [VP, C.c_int, C.c_int, P_F64, C.c_char_p, C.c_int]),
        "ref_tran_steps": (C.c_int, [VP, C.c_double, C.c_int, P_F64, P_F64]),
        "ref_tran_trunc": (C.c_int, [VP, C.c_double, C.c_int, P_F64, P_F64]),
        "ref_num_devices": (C.c_int, [VP]),
        "ref_solver_new": (VP, [C.c_int, C.c_int, P_I32, P_I32]),
        "ref_solver_free": (None, [VP]),
        "ref_solver_numeric": (C.c_int, [VP, P_F64, C.c_double]),
        "ref_solver_set_equilibrate": (None, [VP, C.c_int]),
        "ref_ladder_method": (C.c_int, [VP, P_F64, C.c_int, C.c_int]),
        "ref_solver_refactorize": (C.c_int, [VP, P_F64, C.c_double]),
        "ref_solver_solve": (C.c_int, [VP, P_F64]),
        "ref_solver_solve_transposed": (C.c_int, [VP, P_F64]),
        "ref_solver_numeric_complex": (C.c_int, [VP, P_F64]),
        "ref_solver_refactorize_complex": (C.c_int, [VP, P_F64]),
        "ref_solver_solve_complex": (C.c_int, [VP, P_F64]),
        "ref_solver_fast_count": (C.c_int, [VP]),
        "ref_solver_fallback_count": (C.c_int, [VP]),
        "ref_neo_dump": (C.c_int, [VP, P_I32, P_I32, P_I32, P_I32, P_F64]),
        "ref_neo_fillins": (C.c_int, [VP]),
        "ref_amd_dump": (C.c_int64, [VP, P_I32, P_I32, P_I32, P_I32, P_F64, P_I32, P_I32, P_F64]),
        "ref_amd_ordering": (C.c_int, [C.c_int, P_I32, P_I32, P_I32]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    return L


def orc():
    """The plain-C restatement (oracle/oracle.c)."""
    global _orc
    if _orc is None:
        L = C.CDLL(ORC_SO)
        L.orc_amd_order.argtypes = [C.c_int, P_I32, P_I32, P_I32]
        L.orc_gp_factor.argtypes = [C.c_int, P_I32, P_I32, P_F64, C.c_double, C.c_double, P_I32,
                                    P_I32, P_I32, P_I32, P_F64, C.c_int, P_I32, P_I32, P_F64,
                                    C.c_int]
        L.orc_gp_refactor.argtypes = [C.c_int, P_I32, P_I32, P_F64, C.c_double, C.c_double, P_I32,
                                      P_I32, P_I32, P_I32, P_F64, P_I32, P_I32, P_F64]
        L.orc_gp_solve.argtypes = [C.c_int, P_I32, P_I32, P_I32, P_I32, P_F64, P_I32, P_I32,
                                   P_F64, P_F64]
        L.orc_gp_solve_t.argtypes = L.orc_gp_solve.argtypes
        L.orc_s13_refactor.argtypes = [C.POINTER(OrcS13), P_I32, P_I32, P_F64, C.c_double]
        L.orc_s13_solve.argtypes = [C.POINTER(OrcS13), P_F64, P_F64]
        L.orc_s13_solve_t.argtypes = [C.POINTER(OrcS13), P_F64, P_F64]
        _orc = L
    return _orc


class OrcS13(C.Structure):
    """orc_s13 (oracle.h): a Sparse-1.3 factored structure."""
    _fields_ = [("n", C.c_int), ("i2e_row", P_I32), ("i2e_col", P_I32), ("cptr", P_I32),
                ("rows", P_I32), ("vals", P_F64)]


class OracleGP:
    """The oracle's AMD-LU pipeline on one matrix: minimum-degree order,
    Gilbert-Peierls factor, replay, solves."""

    def __init__(self, n, col_ptr, row_idx, values, gmin=0.0, tol=1e-3):
        self.n = n
        self.cp, _ = i32(col_ptr)
        self.ri, _ = i32(row_idx)
        O = orc()
        self.q = np.zeros(n, np.int32)
        O.orc_amd_order(n, self.cp.ctypes.data_as(P_I32), self.ri.ctypes.data_as(P_I32),
                        self.q.ctypes.data_as(P_I32))
        cap = n * n + 1
        self.pinv = np.zeros(n, np.int32)
        self.Lp = np.zeros(n + 1, np.int32)
        self.Up = np.zeros(n + 1, np.int32)
        self.Li = np.zeros(cap, np.int32)
        self.Ui = np.zeros(cap, np.int32)
        self.Lx = np.zeros(cap)
        self.Ux = np.zeros(cap)
        self.tol = tol
        v, pv = f64(values)
        self.singular = O.orc_gp_factor(
            n, self.cp.ctypes.data_as(P_I32), self.ri.ctypes.data_as(P_I32), pv, gmin, tol,
            self.q.ctypes.data_as(P_I32), self.pinv.ctypes.data_as(P_I32),
            self.Lp.ctypes.data_as(P_I32), self.Li.ctypes.data_as(P_I32),
            self.Lx.ctypes.data_as(P_F64), cap, self.Up.ctypes.data_as(P_I32),
            self.Ui.ctypes.data_as(P_I32), self.Ux.ctypes.data_as(P_F64), cap) != 0

    def _args(self):
        return [a.ctypes.data_as(P_I32 if a.dtype == np.int32 else P_F64)
                for a in (self.q, self.pinv, self.Lp, self.Li, self.Lx, self.Up, self.Ui, self.Ux)]

    def refactor(self, values, gmin=0.0):
        """True when the replay is accepted by the growth check."""
        v, pv = f64(values)
        a = self._args()
        return orc().orc_gp_refactor(self.n, self.cp.ctypes.data_as(P_I32),
                                     self.ri.ctypes.data_as(P_I32), pv, gmin, self.tol,
                                     *a) == 1

    def solve(self, b):
        x, px = f64(np.array(b, dtype=np.float64))
        orc().orc_gp_solve(self.n, *self._args(), px)
        return x

    def solve_t(self, b):
        x, px = f64(np.array(b, dtype=np.float64))
        orc().orc_gp_solve_t(self.n, *self._args(), px)
        return x


class OracleS13:
    """The oracle's Sparse-1.3 refactor and solves on a fixed pivot order
    and fill given in the internal numbering: i2e_row/i2e_col (1-based
    external index of internal row/column i+1), the column lists cptr/rows
    (1-based internal rows, ascending)."""

    def __init__(self, n, i2e_row, i2e_col, cptr, rows):
        self.n = n
        self.keep = [i32(i2e_row)[0], i32(i2e_col)[0], i32(cptr)[0], i32(rows)[0]]
        self.vals = np.zeros(len(self.keep[3]))
        self.s = OrcS13(n, *[k.ctypes.data_as(P_I32) for k in self.keep],
                        self.vals.ctypes.data_as(P_F64))

    @classmethod
    def from_structure(cls, st):
        """From a pivot-space structure (row_perm, col_perm, seg_ptr, rows;
        0-based, column segments in ascending row order)."""
        return cls(len(st["row_perm"]), np.asarray(st["row_perm"]) + 1, np.asarray(st["col_perm"]) + 1,
                   st["seg_ptr"], np.asarray(st["rows"]) + 1)

    def refactor(self, col_ptr, row_idx, values, gmin=0.0):
        """0 ok, else the 1-based internal step of an exact zero pivot."""
        cp, pcp = i32(col_ptr)
        ri, pri = i32(row_idx)
        v, pv = f64(values)
        return orc().orc_s13_refactor(C.byref(self.s), pcp, pri, pv, gmin)

    def solve(self, b):
        b, pb = f64(np.array(b, dtype=np.float64))
        x = np.zeros(self.n)
        orc().orc_s13_solve(C.byref(self.s), pb, x.ctypes.data_as(P_F64))
        return x

    def solve_t(self, b):
        b, pb = f64(np.array(b, dtype=np.float64))
        x = np.zeros(self.n)
        orc().orc_s13_solve_t(C.byref(self.s), pb, x.ctypes.data_as(P_F64))
        return x


def i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(P_I32)


def f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(P_F64)


def _chk(rc):
    if rc < 0:
        msg = ref().ref_last_error().decode()
        if not msg and _ref_gpu is not None:
            msg = _ref_gpu.ref_last_error().decode()
        raise RuntimeError("reference: " + msg)
    return rc


class RefCircuit:
    """A reference Circuit built from a neospice_amd.Circuit's recorded calls.
    `values` optionally overrides device primary values {device index: value}."""

    def __init__(self, ckt, values=None, options=None, probe=True, gpu=False):
        L = ref_gpu() if gpu else ref()
        self.L = L
        nn = len(ckt.nodes)
        self.h = C.c_void_p(L.ref_ckt_new(nn))
        values = values or {}
        d = 0
        for op in ckt.ops:
            k = op[0]
            node = lambda x: x + 1  # product 0-based (GND=-1) -> reference "n<k+1>" (0=GND)
            if k == "node":
                continue
            if k == "model":
                _chk(L.ref_ckt_model(self.h, op[1].encode()))
                continue
            v = values.get(d)
            if k in ("R", "C", "V", "I"):
                f = {"R": L.ref_ckt_r, "C": L.ref_ckt_c, "V": L.ref_ckt_v, "I": L.ref_ckt_i}[k]
                _chk(f(self.h, op[1].encode(), node(op[2]), node(op[3]), op[4] if v is None else v))
            elif k == "Vpulse":
                _chk(L.ref_ckt_v_pulse(self.h, op[1].encode(), node(op[2]), node(op[3]), *op[4]))
            elif k == "Vpwl":
                t, pt = f64(op[4])
                vv, pv = f64(op[5])
                _chk(L.ref_ckt_v_pwl(self.h, op[1].encode(), node(op[2]), node(op[3]), len(t), pt, pv))
            elif k == "D":
                _chk(L.ref_ckt_d(self.h, op[1].encode(), node(op[2]), node(op[3]), op[4].encode()))
            elif k == "M":
                _chk(L.ref_ckt_m(self.h, op[1].encode(), node(op[2]), node(op[3]), node(op[4]),
                                 node(op[5]), op[6].encode(), op[7], op[8]))
            d += 1
        for key, val in (options or {}).items():
            _chk(L.ref_ckt_option(self.h, key.encode(), float(val)))
        _chk(L.ref_ckt_finalize(self.h, 1 if probe else 0))
        self.n = L.ref_ckt_num_vars(self.h)
        self.num_nodes = L.ref_ckt_num_nodes(self.h)
        self.nnz = L.ref_ckt_nnz(self.h)
        self.ndev = L.ref_ckt_num_devices(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_ckt_free(self.h)
            self.h = None

    def pattern(self):
        cp = np.zeros(self.n + 1, np.int32)
        ri = np.zeros(self.nnz, np.int32)
        self.L.ref_ckt_pattern(self.h, i32(cp)[1], i32(ri)[1])
        return cp, ri

    def load_order(self):
        o = np.zeros(self.ndev, np.int32)
        k = self.L.ref_ckt_load_order(self.h, o.ctypes.data_as(P_I32))
        return o[:k]

    def dc(self):
        x = np.zeros(self.n)
        sec = C.c_double()
        its = _chk(self.L.ref_dc(self.h, x.ctypes.data_as(P_F64), C.byref(sec)))
        return x, its, sec.value

    def dc_ladder(self, dcop=False):
        """The transient preamble's operating point from zeros: direct
        Newton, gmin stepping, source stepping, pseudo-transient
        (ref_dc_ladder_op), in MODETRANOP or, with dcop, MODEDCOP.
        Returns (solution, iterations of the method that converged, stage)."""
        x = np.zeros(self.n)
        st = C.c_int()
        its = _chk(self.L.ref_dc_ladder_op(self.h, x.ctypes.data_as(P_F64), C.byref(st),
                                           0x10 if dcop else 0x20))
        return x, its, st.value

    def ladder_method(self, method, dcop=False):
        """source stepping ("source") or true gmin stepping ("true_gmin")
        alone from zeros on a fresh solver; returns (solution, iterations),
        iterations negative when it failed."""
        x = np.zeros(self.n)
        its = self.L.ref_ladder_method(self.h, x.ctypes.data_as(P_F64), {"source": 1, "true_gmin": 2}[method],
                                       0x10 if dcop else 0x20)
        return x, its

    def gmin_stepping(self):
        """gmin_stepping alone from zeros on a fresh solver; returns
        (solution, iterations), iterations negative when it failed."""
        x = np.zeros(self.n)
        its = self.L.ref_gmin_stepping(self.h, x.ctypes.data_as(P_F64))
        return x, its

    def tran(self, tstep, tstop, max_points=200000):
        t = np.zeros(max_points)
        s = np.zeros(max_points * self.n)
        sec = C.c_double()
        its = C.c_int32()
        k = _chk(self.L.ref_tran(self.h, tstep, tstop, max_points, t.ctypes.data_as(P_F64),
                                 s.ctypes.data_as(P_F64), C.byref(sec), C.byref(its)))
        return t[:k].copy(), s[:k * self.n].reshape(k, self.n).copy(), its.value, sec.value

    def tran_fixed(self, dt, nsteps):
        """Fixed-step backward-Euler stepping through the reference's own
        newton_solve / rotate_state; returns (final solution, Newton
        iterations of the stepping loop, seconds)."""
        x = np.zeros(self.n)
        sec = C.c_double()
        its = _chk(self.L.ref_tran_fixed(self.h, dt, nsteps, x.ctypes.data_as(P_F64), C.byref(sec)))
        return x, its, sec.value

    def tran_steps(self, dt, nsteps):
        """Fixed-step transient with the controller's constant-step order
        sequence (backward Euler, then order 2 of the circuit's method):
        all solutions ((nsteps+1) x n), Newton iterations, seconds."""
        xs = np.zeros((nsteps + 1) * self.n)
        sec = C.c_double()
        its = _chk(self.L.ref_tran_steps(self.h, dt, nsteps, xs.ctypes.data_as(P_F64),
                                         C.byref(sec)))
        return xs.reshape(nsteps + 1, self.n), its, sec.value

    def tran_trunc(self, dt, nsteps):
        """The fixed-step transient of tran_steps, recording after each step
        (before the states rotate) every device's compute_trunc with
        CKTdeltaOld[0..2] = dt: solutions ((nsteps+1) x n) and
        (nsteps x num devices) step proposals (1e30: none)."""
        nd = self.L.ref_num_devices(self.h)
        xs = np.zeros((nsteps + 1) * self.n)
        tr = np.zeros(nsteps * max(nd, 1))
        _chk(self.L.ref_tran_trunc(self.h, dt, nsteps, xs.ctypes.data_as(P_F64),
                                   tr.ctypes.data_as(P_F64)))
        return xs.reshape(nsteps + 1, self.n), tr[:nsteps * nd].reshape(nsteps, nd)

    def sens(self, out_node, max_out=4096):
        """.sens of v(out_node) (sens.cpp:57-186): {element:parameter: value}."""
        vals = np.zeros(max_out)
        stride = 64
        names = C.create_string_buffer(max_out * stride)
        k = _chk(self.L.ref_sens(self.h, out_node, max_out, vals.ctypes.data_as(P_F64), names,
                                 stride))
        raw = names.raw
        out = {}
        for i in range(k):
            nm = raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode()
            out[nm] = float(vals[i])
        return out

    def records(self):
        k = self.L.ref_tran_num_records(self.h)
        ints = np.zeros(4 * k, np.int32)
        dbl = np.zeros(10 * k)
        self.L.ref_tran_records(self.h, ints.ctypes.data_as(P_I32), dbl.ctypes.data_as(P_F64))
        return ints.reshape(k, 4), dbl.reshape(k, 10)

    def set_ctx(self, mode, order=1, method=0, ag=(0, 0, 0), delta=0.0, dold1=0.0, xmu_ratio=1.0,
                time=0.0):
        self.L.ref_ckt_set_ctx(self.h, mode, order, method, ag[0], ag[1], ag[2], delta, dold1,
                               xmu_ratio, time)

    def load(self, x):
        x, px = f64(x)
        vals = np.zeros(self.nnz)
        rhs = np.zeros(self.n)
        _chk(self.L.ref_ckt_load(self.h, px, vals.ctypes.data_as(P_F64), rhs.ctypes.data_as(P_F64)))
        return vals, rhs

    def state(self, which, count):
        p = self.L.ref_ckt_state(self.h, which)
        return np.ctypeslib.as_array(p, shape=(count,)).copy()


class RefSolver:
    """A reference NeoSolver (kind 0) or AmdLuSolver (kind 1) on a CSC pattern."""

    def __init__(self, kind, n, col_ptr, row_idx):
        self.L = ref()
        self.n = n
        self.cp, pcp = i32(col_ptr)
        self.ri, pri = i32(row_idx)
        self.h = C.c_void_p(self.L.ref_solver_new(kind, n, pcp, pri))
        self.kind = kind

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_solver_free(self.h)
            self.h = None

    def set_equilibrate(self, on=True):
        self.L.ref_solver_set_equilibrate(self.h, 1 if on else 0)

    def numeric(self, vals, gmin=0.0):
        v, pv = f64(vals)
        return bool(_chk(self.L.ref_solver_numeric(self.h, pv, gmin)))

    def refactorize(self, vals, gmin=0.0):
        v, pv = f64(vals)
        return bool(_chk(self.L.ref_solver_refactorize(self.h, pv, gmin)))

    def solve(self, b):
        b, pb = f64(np.array(b, dtype=np.float64))
        _chk(self.L.ref_solver_solve(self.h, pb))
        return b

    def solve_transposed(self, b):
        b, pb = f64(np.array(b, dtype=np.float64))
        _chk(self.L.ref_solver_solve_transposed(self.h, pb))
        return b

    # ---- complex paths: complex numpy arrays (slot order / unknowns)
    @staticmethod
    def _inter(z):
        z = np.asarray(z, np.complex128).ravel()
        out = np.empty(2 * len(z), np.float64)
        out[0::2] = z.real
        out[1::2] = z.imag
        return out

    def numeric_complex(self, ax):
        """True when the ordering met a singular matrix (the reference's
        numeric_complex returns nothing; its matrix's error state says so)."""
        a, pa = f64(self._inter(ax))
        return bool(_chk(self.L.ref_solver_numeric_complex(self.h, pa)))

    def refactorize_complex(self, ax):
        a, pa = f64(self._inter(ax))
        r = self.L.ref_solver_refactorize_complex(self.h, pa)
        if r < 0:
            _chk(r)
        return bool(r)

    def solve_complex(self, b):
        r, pr = f64(self._inter(b))
        _chk(self.L.ref_solver_solve_complex(self.h, pr))
        return r[0::2] + 1j * r[1::2]

    def neo_dump(self):
        n = self.n
        i2r = np.zeros(n, np.int32)
        i2c = np.zeros(n, np.int32)
        cnt = self.L.ref_neo_dump(self.h, None, None, None, None, None)
        cp = np.zeros(n + 1, np.int32)
        rows = np.zeros(cnt, np.int32)
        vals = np.zeros(cnt)
        self.L.ref_neo_dump(self.h, i2r.ctypes.data_as(P_I32), i2c.ctypes.data_as(P_I32),
                            cp.ctypes.data_as(P_I32), rows.ctypes.data_as(P_I32),
                            vals.ctypes.data_as(P_F64))
        return {"i2e_row": i2r, "i2e_col": i2c, "col_ptr": cp, "rows": rows, "vals": vals}

    def amd_dump(self):
        n = self.n
        sz = self.L.ref_amd_dump(self.h, *([None] * 8))
        lnz, unz = sz >> 32, sz & 0xFFFFFFFF
        out = {k: np.zeros(n, np.int32) for k in ("q", "pinv")}
        out["Lp"] = np.zeros(n + 1, np.int32)
        out["Up"] = np.zeros(n + 1, np.int32)
        out["Li"] = np.zeros(lnz, np.int32)
        out["Ui"] = np.zeros(unz, np.int32)
        out["Lx"] = np.zeros(lnz)
        out["Ux"] = np.zeros(unz)
        args = [out[k] for k in ("q", "pinv", "Lp", "Li", "Lx", "Up", "Ui", "Ux")]
        self.L.ref_amd_dump(self.h, *[a.ctypes.data_as(P_I32 if a.dtype == np.int32 else P_F64)
                                      for a in args])
        return out


def split_newton_calls(ints, dbls):
    """Groups probe records into Newton calls: a transient call starts at a
    record whose mode carries INITTRAN or INITPRED (newton.cpp:396-402).
    Returns a list of dicts with the call's first-iteration context, the
    number of iterations and whether the step was accepted."""
    calls = []
    for k in range(len(ints)):
        mode = int(ints[k, 0])
        if mode & 0x1 and mode & (0x1000 | 0x2000):
            calls.append({"first": k, "iters": 0, "mode": mode, "order": int(ints[k, 1]),
                          "method": int(ints[k, 2]), "ag": tuple(dbls[k, 0:3]),
                          "delta": float(dbls[k, 3]), "dold1": float(dbls[k, 5]),
                          "xmu_ratio": float(dbls[k, 7]),
                          "time": float(dbls[k, 8]), "acc0": int(ints[k, 3])})
        if calls and mode & 0x1:
            calls[-1]["iters"] += 1
    for i, c in enumerate(calls):
        # the run ends on an accepted step (transient.cpp:873, 965-1004)
        c["accepted"] = calls[i + 1]["acc0"] > c["acc0"] if i + 1 < len(calls) else True
    return calls
