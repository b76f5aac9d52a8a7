"""neospice_amd: the neospice Newton hot path on MI355X.

Python mirror of the reference's typed Circuit API (src/core/circuit.hpp:
179-226) and of the analyses that drive Newton (DC operating point,
transient), over the C ABI of libneospice_amd.so.  Batched: one circuit
topology, B instances with per-instance device values.

    ckt = Circuit()
    a, b = ckt.node("a"), ckt.node("b")
    ckt.V("v1", a, GND, 1.0); ckt.R("r1", a, b, 1e3); ckt.C("c1", b, GND, 1e-9)
    eng = BatchEngine(ckt, batch=256)
    x = eng.dc_op()                       # (n, B)
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import Ctx, Options, check, f64p, i32p, as_f64, as_i32, lib

GND = -1

# CKTmode bits (ngspice cktdefs.h; src/core/newton.cpp:109-115)
MODETRAN, MODEAC, MODEDCOP, MODETRANOP, MODEDCTRANCURVE = 0x1, 0x2, 0x10, 0x20, 0x40
MODEINITFLOAT, MODEINITJCT, MODEINITFIX = 0x100, 0x200, 0x400
MODEINITSMSIG, MODEINITTRAN, MODEINITPRED, MODEUIC = 0x800, 0x1000, 0x2000, 0x10000

ENGINES = {"auto": 0, "markowitz": 1, "amdlu": 2}


def default_options(**kw):
    o = Options()
    lib().nsa_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "engine" and isinstance(v, str):
            v = ENGINES[v]
        if k == "method" and isinstance(v, str):
            v = 1 if v == "gear" else 0
        setattr(o, k, v)
    return o


class Circuit:
    """Typed circuit builder (records the calls; builds the native circuit on finalize)."""

    def __init__(self):
        self.ops = []
        self.nodes = {}
        self.devices = []     # (kind, name)
        self._native = None

    # -- builders (names and argument order follow circuit.hpp:179-226) --
    def node(self, name):
        if name in ("0", "gnd", "GND"):
            return GND
        if name not in self.nodes:
            self.nodes[name] = len(self.nodes)
            self.ops.append(("node", name))
        return self.nodes[name]

    def _dev(self, kind, name, *args):
        self._check_open()
        self.ops.append((kind, name) + tuple(args))
        self.devices.append((kind, name))
        return len(self.devices) - 1

    def R(self, name, a, b, ohms):
        return self._dev("R", name, a, b, float(ohms))

    def C(self, name, a, b, farads):
        return self._dev("C", name, a, b, float(farads))

    def V(self, name, p, n, dc):
        return self._dev("V", name, p, n, float(dc))

    def V_pulse(self, name, p, n, v1, v2, td=0.0, tr=-1.0, tf=-1.0, pw=-1.0, per=-1.0):
        return self._dev("Vpulse", name, p, n, (v1, v2, td, tr, tf, pw, per))

    def V_pwl(self, name, p, n, times, values):
        return self._dev("Vpwl", name, p, n, tuple(map(float, times)), tuple(map(float, values)))

    def I(self, name, p, n, dc):
        return self._dev("I", name, p, n, float(dc))

    def model(self, text):
        self._check_open()
        self.ops.append(("model", text))

    def D(self, name, a, k, model):
        return self._dev("D", name, a, k, model)

    def M(self, name, d, g, s, b, model, w, l):
        return self._dev("M", name, d, g, s, b, model, float(w), float(l))

    def ops_devices(self):
        """The recorded device calls, one per device index (kind, name, pins...)."""
        return [op for op in self.ops if op[0] not in ("node", "model")]

    def _check_open(self):
        if self._native is not None:
            raise RuntimeError("Cannot add to a finalized circuit")

    # -- native --
    def finalize(self):
        if self._native is not None:
            return self
        L = lib()
        h = L.nsa_circuit_create()
        self._native = h
        idx = C.c_int32()
        dev = C.c_int32()
        for op in self.ops:
            k = op[0]
            if k == "node":
                check(L.nsa_circuit_node(h, op[1].encode(), C.byref(idx)))
            elif k == "model":
                check(L.nsa_circuit_model(h, op[1].encode()))
            elif k in ("R", "C", "V", "I"):
                f = {"R": L.nsa_circuit_r, "C": L.nsa_circuit_c, "V": L.nsa_circuit_v,
                     "I": L.nsa_circuit_i}[k]
                check(f(h, op[1].encode(), op[2], op[3], op[4], C.byref(dev)))
            elif k == "Vpulse":
                p = as_f64(op[4])
                check(L.nsa_circuit_v_pulse(h, op[1].encode(), op[2], op[3], f64p(p), C.byref(dev)))
            elif k == "Vpwl":
                t, v = as_f64(op[4]), as_f64(op[5])
                check(L.nsa_circuit_v_pwl(h, op[1].encode(), op[2], op[3], len(t), f64p(t), f64p(v),
                                          C.byref(dev)))
            elif k == "D":
                check(L.nsa_circuit_d(h, op[1].encode(), op[2], op[3], op[4].encode(), C.byref(dev)))
            elif k == "M":
                check(L.nsa_circuit_m(h, op[1].encode(), op[2], op[3], op[4], op[5], op[6].encode(),
                                      op[7], op[8], C.byref(dev)))
        check(L.nsa_circuit_finalize(h))
        info = np.zeros(5, np.int32)
        check(L.nsa_circuit_info(h, i32p(info)))
        self.n, self.num_nodes, self.nnz, self.num_states, linear = map(int, info)
        self.is_linear = bool(linear)
        self.col_ptr = np.zeros(self.n + 1, np.int32)
        self.row_idx = np.zeros(self.nnz, np.int32)
        check(L.nsa_circuit_pattern(h, i32p(self.col_ptr), i32p(self.row_idx)))
        self.load_order = np.zeros(len(self.devices), np.int32)
        check(L.nsa_circuit_load_order(h, i32p(self.load_order)))
        self.branch = np.zeros(len(self.devices), np.int32)
        for d in range(len(self.devices)):
            check(L.nsa_circuit_branch(h, d, C.byref(dev)))
            self.branch[d] = dev.value
        # device indices of each family in family (load) order: the row order
        # of the engine's per-family value and gradient arrays
        kinds = {"R": ("R",), "C": ("C",), "V": ("V", "Vpulse", "Vpwl"), "I": ("I",),
                 "D": ("D",), "M": ("M",)}
        self.family = {f: np.array([d for d in self.load_order if self.devices[d][0] in ks],
                                   np.int32) for f, ks in kinds.items()}
        return self

    @property
    def native(self):
        if self._native is None:
            self.finalize()
        return self._native

    def device_index(self, name):
        for i, (_, nm) in enumerate(self.devices):
            if nm == name:
                return i
        raise KeyError(name)

    def __del__(self):
        if getattr(self, "_native", None) is not None:
            try:
                lib().nsa_circuit_destroy(self._native)
            except Exception:
                pass


FAMILIES = {"R": 0, "C": 1, "V": 2, "I": 3}


def make_ctx(mode, order=1, method=0, ag=(0.0, 0.0, 0.0), xmu_ratio=1.0, time=0.0, src_fact=1.0,
             delta=0.0, delta_old1=0.0, delta_old2=None):
    c = Ctx()
    c.mode, c.order, c.method = int(mode), int(order), int(method)
    c.ag0, c.ag1, c.ag2 = map(float, ag)
    c.xmu_ratio, c.time, c.src_fact = float(xmu_ratio), float(time), float(src_fact)
    c.delta, c.delta_old1 = float(delta), float(delta_old1)
    # CKTdeltaOld[2]; a constant-step history unless given
    c.delta_old2 = float(delta_old1 if delta_old2 is None else delta_old2)
    return c


def be_ctx(dt, time, first):
    """Context of a fixed backward-Euler step (fill_integrator_context,
    transient.cpp:398-447): delta = deltaOld[1] = dt."""
    return make_ctx(MODETRAN | (MODEINITTRAN if first else MODEINITPRED), 1, 0,
                    (1.0 / dt, -1.0 / dt, 0.0), 0.0, time, 1.0, dt, dt)


class BatchEngine:
    """B instances of one circuit on the GPU (nsa_engine_*)."""

    def __init__(self, circuit, batch=1, **opts):
        circuit.finalize()
        self.circuit = circuit
        self.B = int(batch)
        self.opts = default_options(**opts)
        h = C.c_void_p()
        check(lib().nsa_engine_create(circuit.native, self.B, C.byref(self.opts), C.byref(h)))
        self.h = h
        self.n = circuit.n

    def close(self):
        if getattr(self, "h", None):
            lib().nsa_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_values(self, dev, vals):
        if isinstance(dev, str):
            dev = self.circuit.device_index(dev)
        v = as_f64(np.broadcast_to(np.asarray(vals, np.float64), (self.B,)))
        check(lib().nsa_engine_set_values(self.h, int(dev), f64p(v)))

    def set_model_values(self, dev, param, vals):
        """Per-instance value of a BSIM4v7 card parameter of device `dev`."""
        if isinstance(dev, str):
            dev = self.circuit.device_index(dev)
        v = as_f64(np.broadcast_to(np.asarray(vals, np.float64), (self.B,)))
        check(lib().nsa_engine_set_model_values(self.h, int(dev), param.encode(), f64p(v)))

    def set_ctx(self, ctx):
        if isinstance(ctx, Ctx):
            check(lib().nsa_engine_set_ctx(self.h, C.byref(ctx), 0))
        else:
            arr = (Ctx * self.B)(*ctx)
            check(lib().nsa_engine_set_ctx(self.h, arr, 1))

    def newton(self, itl):
        conv = np.zeros(self.B, np.int32)
        its = np.zeros(self.B, np.int32)
        check(lib().nsa_engine_newton(self.h, int(itl), i32p(conv), i32p(its)))
        return conv.astype(bool), its

    def solution(self):
        x = np.zeros((self.n, self.B))
        check(lib().nsa_engine_get_solution(self.h, f64p(x)))
        return x

    def set_solution(self, x):
        x = as_f64(np.broadcast_to(x, (self.n, self.B)) if np.ndim(x) else np.full((self.n, self.B), x))
        check(lib().nsa_engine_set_solution(self.h, f64p(x)))

    def state(self, which=0):
        s = np.zeros((self.circuit.num_states, self.B))
        check(lib().nsa_engine_get_state(self.h, int(which), f64p(s)))
        return s

    def set_state(self, which, s):
        s = as_f64(s)
        check(lib().nsa_engine_set_state(self.h, int(which), f64p(s)))

    def load(self):
        vals = np.zeros((self.circuit.nnz, self.B))
        rhs = np.zeros((self.n, self.B))
        check(lib().nsa_engine_load(self.h, f64p(vals), f64p(rhs)))
        return vals, rhs

    def solve_transposed(self, b):
        b = as_f64(np.array(b, dtype=np.float64).reshape(self.n, self.B))
        check(lib().nsa_engine_solve_transposed(self.h, f64p(b)))
        return b

    def seed_states(self):
        check(lib().nsa_engine_seed_states(self.h))

    def rotate_states(self):
        check(lib().nsa_engine_rotate_states(self.h))

    def trunc(self, per_device=False):
        """Device truncation-error step proposal after a converged step,
        before rotate_states (nsa_engine_trunc): the minimum over devices
        per instance (B,), and with per_device the devices' circuit indices
        and their values (ndev x B)."""
        B = self.B
        dt = np.zeros(B)
        if not per_device:
            check(lib().nsa_engine_trunc(self.h, f64p(dt), None))
            return dt
        nd = check(lib().nsa_engine_trunc_devices(self.h, None))
        idx = np.zeros(max(nd, 1), np.int32)
        check(lib().nsa_engine_trunc_devices(self.h, i32p(idx)))
        dev = np.zeros((max(nd, 1), B))
        check(lib().nsa_engine_trunc(self.h, f64p(dt), f64p(dev)))
        return dt, idx[:nd], dev[:nd]

    def reset_solver(self):
        check(lib().nsa_engine_reset_solver(self.h))

    @property
    def solver_name(self):
        return lib().nsa_engine_solver_name(self.h).decode()

    def stats(self):
        s = np.zeros(8, np.int64)
        check(lib().nsa_engine_stats(self.h, s.ctypes.data_as(_lib.P_I64)))
        return {"cpu_pivot": int(s[0]), "gpu_refactor": int(s[1]), "nnz_lu": int(s[2]),
                "levels_refactor": int(s[3]), "levels_solve": int(s[4]), "gpu_solve": int(s[5]),
                "loads": int(s[6]), "pivot_groups": int(s[7])}

    def set_timing(self, on=True):
        check(lib().nsa_engine_set_timing(self.h, int(bool(on))))

    def timing(self):
        """Accumulated HIP-event milliseconds: load+gather, refactor, solve."""
        t = np.zeros(5)
        check(lib().nsa_engine_timing(self.h, f64p(t)))
        return {"load_ms": float(t[0]), "refactor_ms": float(t[1]), "solve_ms": float(t[2]),
                "bsim4_ms": float(t[3]), "loads": int(t[4])}

    # -- parameters and adjoints --
    def family_count(self, family):
        c = C.c_int32()
        check(lib().nsa_engine_family_count(self.h, FAMILIES[family], C.byref(c)))
        return c.value

    def set_family_values(self, family, vals):
        """All per-instance values of a family: (count, B) array (R ohms, C
        farads, V / I dc)."""
        n = self.family_count(family)
        v = as_f64(np.broadcast_to(np.asarray(vals, np.float64), (n, self.B)))
        check(lib().nsa_engine_set_family_values(self.h, FAMILIES[family], f64p(v)))

    def family_values(self, family):
        v = np.zeros((self.family_count(family), self.B))
        check(lib().nsa_engine_get_family_values(self.h, FAMILIES[family], f64p(v)))
        return v

    def dc_adjoint(self, dLdx):
        """Gradients of a loss through the operating point of the last Newton
        solve, given dL/dx (n x B): dict with lambda and dL/dR, dL/dC, dL/dV,
        dL/dI (count x B)."""
        g = as_f64(np.asarray(dLdx, np.float64).reshape(self.n, self.B))
        out = {"lambda": np.zeros((self.n, self.B))}
        for f in ("R", "C", "V", "I"):
            out[f] = np.zeros((max(1, self.family_count(f)), self.B))
        check(lib().nsa_engine_dc_adjoint(self.h, f64p(g), f64p(out["lambda"]), f64p(out["R"]),
                                          f64p(out["C"]), f64p(out["V"]), f64p(out["I"])))
        for f in ("R", "C", "V", "I"):
            out[f] = out[f][:self.family_count(f)]
        return out

    def tran_forward(self, nsteps, dt, probes, ckpt_every=25, x0=0.0):
        """Fixed-step transient (TRANOP operating point, then nsteps steps:
        backward Euler first, then the options' method); returns the probe
        voltages, (nsteps+1, nprobe, B)."""
        self.set_solution(x0)
        p = as_i32(probes)
        out = np.zeros((nsteps + 1, len(p), self.B))
        check(lib().nsa_engine_tran_forward(self.h, int(nsteps), float(dt), i32p(p), len(p),
                                            int(ckpt_every), f64p(out)))
        return out

    def tran_adjoint(self, dLdprobe):
        """dL/dR and dL/dC (count x B) of the last tran_forward, given dL/d
        probe values shaped like its output."""
        g = as_f64(np.asarray(dLdprobe, np.float64))
        gR = np.zeros((max(1, self.family_count("R")), self.B))
        gC = np.zeros((max(1, self.family_count("C")), self.B))
        check(lib().nsa_engine_tran_adjoint(self.h, f64p(g), f64p(gR), f64p(gC)))
        return gR[:self.family_count("R")], gC[:self.family_count("C")]

    def adjoint_timing(self):
        t = np.zeros(2)
        check(lib().nsa_engine_adjoint_timing(self.h, f64p(t)))
        return {"forward_ms": float(t[0]), "adjoint_ms": float(t[1])}

    # -- analyses --
    def dc_op(self, mode=MODEDCOP | MODEINITJCT, x0=0.0):
        """solve_dc's direct Newton attempt (dc.cpp:104-107)."""
        self.set_solution(x0)
        self.set_ctx(make_ctx(mode, 1, int(self.opts.method), (0, 0, 0), 1.0, 0.0,
                              self.opts.src_fact))
        conv, its = self.newton(self.opts.itl1)
        if not conv.all():
            raise RuntimeError("DC operating point failed to converge for instances %s"
                               % np.nonzero(~conv)[0][:10].tolist())
        return self.solution(), its

    def tran_begin(self, x0=0.0):
        """DC preamble of solve_transient (transient.cpp:710-802): TRANOP Newton on
        its own solver, a fresh solver for the transient, state1..3 = state0."""
        x, its = self.dc_op(MODETRANOP | MODEINITJCT, x0)
        self.reset_solver()
        self.seed_states()
        return x, its

    def tran_step(self, ctx, itl=None):
        """One transient Newton solve; the caller accepts (rotate) or rejects."""
        self.set_ctx(ctx)
        return self.newton(self.opts.itl4 if itl is None else itl)


class Solver:
    """The ISolver seam (nsa_lu_*) for a batch of matrices on one pattern."""

    def __init__(self, n, col_ptr, row_idx, batch=1, is_linear=True, engine="auto"):
        self.n, self.B = int(n), int(batch)
        self.nnz = int(col_ptr[n])
        cp, ri = as_i32(col_ptr), as_i32(row_idx)
        h = C.c_void_p()
        check(lib().nsa_lu_create(self.B, self.n, i32p(cp), i32p(ri), int(bool(is_linear)),
                                  ENGINES[engine] if isinstance(engine, str) else int(engine),
                                  C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().nsa_lu_destroy(self.h)
            self.h = None

    def _vals(self, values):
        return as_f64(np.asarray(values, np.float64).reshape(self.nnz, self.B))

    def numeric(self, values, diag_gmin=0.0):
        st = np.zeros(self.B, np.int32)
        check(lib().nsa_lu_numeric(self.h, f64p(self._vals(values)), float(diag_gmin), i32p(st)))
        return st.astype(bool)

    def refactorize(self, values, diag_gmin=0.0):
        st = np.zeros(self.B, np.int32)
        check(lib().nsa_lu_refactorize(self.h, f64p(self._vals(values)), float(diag_gmin), i32p(st)))
        return st.astype(bool)

    def solve(self, rhs):
        b = as_f64(np.array(rhs, np.float64).reshape(self.n, self.B))
        check(lib().nsa_lu_solve(self.h, f64p(b)))
        return b

    def solve_transposed(self, rhs):
        b = as_f64(np.array(rhs, np.float64).reshape(self.n, self.B))
        check(lib().nsa_lu_solve_transposed(self.h, f64p(b)))
        return b

    def solve_noalloc(self, rhs):
        """solve() through nsa_lu_solve_noalloc: no device allocation (the
        library's static device array stages the vectors)."""
        b = as_f64(np.array(rhs, np.float64).reshape(self.n, self.B))
        check(lib().nsa_lu_solve_noalloc(self.h, f64p(b)))
        return b

    # ---- the complex (AC) paths: complex arrays in, complex arrays out
    def _cvals(self, ax, rows):
        a = np.asarray(ax, np.complex128).reshape(rows, self.B)
        out = np.empty((2 * rows, self.B), np.float64)
        out[0::2] = a.real
        out[1::2] = a.imag
        return as_f64(out)

    def numeric_complex(self, ax):
        """ISolver::numeric_complex for every instance: ax (nnz, B) complex,
        CSC slot order.  Returns the per-instance singular flags."""
        st = np.zeros(self.B, np.int32)
        check(lib().nsa_lu_numeric_complex(self.h, f64p(self._cvals(ax, self.nnz)), i32p(st)))
        return st.astype(bool)

    def refactorize_complex(self, ax):
        st = np.zeros(self.B, np.int32)
        check(lib().nsa_lu_refactorize_complex(self.h, f64p(self._cvals(ax, self.nnz)), i32p(st)))
        return st.astype(bool)

    def solve_complex(self, rhs):
        """rhs (n, B) complex; returns the solutions (n, B)."""
        b = self._cvals(rhs, self.n)
        check(lib().nsa_lu_solve_complex(self.h, f64p(b)))
        return b[0::2] + 1j * b[1::2]

    def ac_sweep(self, g, c, freqs, rhs):
        """The AC analysis's sweep for every instance and frequency: matrix
        g + j 2 pi f c (g, c: (nnz, B) real), ordered at freqs[0], solved for
        rhs ((n, B) complex).  Returns (x, singular): x (nf, n, B) complex
        (a view of the library's instance-major output), singular (nf, B)."""
        freqs = as_f64(np.asarray(freqs, np.float64).ravel())
        nf = len(freqs)
        # instance-major blocks: a batch of the reference's own arrays
        gg = as_f64(np.asarray(g, np.float64).reshape(self.nnz, self.B).T)
        cc = as_f64(np.asarray(c, np.float64).reshape(self.nnz, self.B).T)
        rb = np.ascontiguousarray(np.asarray(rhs, np.complex128).reshape(self.n, self.B).T)
        out = np.empty((self.B, nf, self.n), np.complex128)
        st = np.zeros((self.B, nf), np.int32)
        check(lib().nsa_lu_ac_sweep(self.h, f64p(gg), f64p(cc), f64p(freqs), nf, f64p(rb.view(np.float64)),
                                    f64p(out.view(np.float64)), i32p(st)))
        return out.transpose(1, 2, 0), st.T.astype(bool)

    def ac_stats(self):
        """Seconds of the last ac_sweep: host orderings, device work with the
        inputs resident, the whole call inside the library."""
        o = np.zeros(3, np.float64)
        check(lib().nsa_lu_ac_stats(self.h, f64p(o)))
        return {"order_s": o[0], "device_s": o[1], "total_s": o[2]}

    @property
    def name(self):
        return lib().nsa_lu_name(self.h).decode()

    def set_equilibrate(self, on=True):
        """ISolver::set_equilibrate: radix-2 row/column scaling of the next
        factorizations and their solves (Markowitz solvers only, as in the
        reference)."""
        check(lib().nsa_lu_set_equilibrate(self.h, int(bool(on))))

    def info(self):
        o = np.zeros(4, np.int32)
        check(lib().nsa_lu_info(self.h, i32p(o)))
        return {"n": int(o[0]), "nnz_lu": int(o[1]), "levels_refactor": int(o[2]),
                "levels_solve": int(o[3])}

    def structure(self):
        info = self.info()
        rp = np.zeros(self.n, np.int32)
        cp = np.zeros(self.n, np.int32)
        seg = np.zeros(self.n + 1, np.int32)
        rows = np.zeros(info["nnz_lu"], np.int32)
        check(lib().nsa_lu_structure(self.h, i32p(rp), i32p(cp), i32p(seg), i32p(rows)))
        return {"row_perm": rp, "col_perm": cp, "seg_ptr": seg, "rows": rows}


def pivot_structure(engine, n, col_ptr, row_idx, values, diag_gmin=0.0):
    """Host-only: pivot order and fill of a fresh factorization (nsa_pivot_structure).
    Returns None when singular."""
    eng = ENGINES[engine] if isinstance(engine, str) else int(engine)
    cp, ri = as_i32(col_ptr), as_i32(row_idx)
    if np.iscomplexobj(values):
        # the complex Markowitz ordering (NeoSolver::numeric_complex)
        z = np.asarray(values, np.complex128).ravel()
        v = np.empty(2 * len(z), np.float64)
        v[0::2], v[1::2] = z.real, z.imag
        eng = 3
    else:
        v = as_f64(values)
    nlu = C.c_int32()
    rc = check(lib().nsa_pivot_structure(eng, n, i32p(cp), i32p(ri), f64p(v), float(diag_gmin),
                                         C.byref(nlu), None, None, None, None, None))
    if rc == 1:
        return None
    out = {"row_perm": np.zeros(n, np.int32), "col_perm": np.zeros(n, np.int32),
           "seg_ptr": np.zeros(n + 1, np.int32), "nU": np.zeros(n, np.int32),
           "rows": np.zeros(nlu.value, np.int32)}
    check(lib().nsa_pivot_structure(eng, n, i32p(cp), i32p(ri), f64p(v), float(diag_gmin),
                                    C.byref(nlu), i32p(out["row_perm"]), i32p(out["col_perm"]),
                                    i32p(out["seg_ptr"]), i32p(out["nU"]), i32p(out["rows"])))
    return out


def min_degree_order(n, col_ptr, row_idx):
    perm = np.zeros(n, np.int32)
    check(lib().nsa_min_degree_order(n, i32p(as_i32(col_ptr)), i32p(as_i32(row_idx)), i32p(perm)))
    return perm


class Bundle:
    """Instances of different circuits (topologies and sizes may differ)
    advanced together on the GPU (nsa_bundle_*): the batched form of the
    reference's one-circuit-at-a-time .step / Monte Carlo sweep
    (run_step_sweep, src/api/neospice.cpp:229-334).  Solutions are packed
    instance after instance; `solution(i)` returns instance i's unknowns."""

    def __init__(self, circuits, **opts):
        self.circuits = list(circuits)
        for c in self.circuits:
            c.finalize()
        self.N = len(self.circuits)
        self.opts = default_options(**opts)
        arr = (C.c_void_p * self.N)(*[c.native for c in self.circuits])
        h = C.c_void_p()
        check(lib().nsa_bundle_create(arr, self.N, C.byref(self.opts), C.byref(h)))
        self.h = h
        info = np.zeros(4, np.int32)
        check(lib().nsa_bundle_info(self.h, i32p(info)))
        self.total_n, self.total_nnz, self.num_states = int(info[1]), int(info[2]), int(info[3])
        self.xoff = np.zeros(self.N + 1, np.int32)
        self.aoff = np.zeros(self.N + 1, np.int32)
        check(lib().nsa_bundle_offsets(self.h, i32p(self.xoff), i32p(self.aoff)))

    def close(self):
        if getattr(self, "h", None):
            lib().nsa_bundle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def solution(self, i=None):
        x = np.zeros(self.total_n)
        check(lib().nsa_bundle_get_solution(self.h, f64p(x)))
        if i is None:
            return x
        return x[self.xoff[i]:self.xoff[i + 1]].copy()

    def set_solution(self, x):
        x = as_f64(np.broadcast_to(np.asarray(x, np.float64), (self.total_n,)))
        check(lib().nsa_bundle_set_solution(self.h, f64p(x)))

    def dc_op(self, mode=MODEDCOP | MODEINITJCT, x0=0.0):
        """Direct Newton attempt of every instance's operating point."""
        self.set_solution(x0)
        conv = np.zeros(self.N, np.int32)
        its = np.zeros(self.N, np.int32)
        check(lib().nsa_bundle_dc(self.h, int(mode), i32p(conv), i32p(its)))
        return conv.astype(bool), its

    def op_ladder(self, x0=0.0, skip_direct=False, dcop=False, method="gmin"):
        """The operating-point ladder per instance: direct Newton in
        MODETRANOP (MODEDCOP with dcop, as solve_dc), then for the instances
        that failed true gmin stepping (dcop only), gmin stepping and source
        stepping.  Returns (converged, iterations, stage) per instance, stage
        0 direct, 1 gmin, 2 source, 4 true gmin stepping, -1 none.  With
        skip_direct, `method` ("gmin", "source", "true_gmin") runs alone on
        every instance."""
        self.set_solution(x0)
        conv, its, st = (np.zeros(self.N, np.int32) for _ in range(3))
        m = {"gmin": 0, "source": 1, "true_gmin": 2}[method]
        flags = int(bool(skip_direct)) | (2 if dcop else 0) | (m << 4)
        check(lib().nsa_bundle_dc_ladder(self.h, flags, i32p(conv), i32p(its), i32p(st)))
        return conv.astype(bool), its, st

    def set_device_values(self, value=None, w=None, l=None):
        """New per-device values (arrays over all devices, layout of
        dc_adjoint's outputs; None leaves a kind alone): value for R, C, V
        and I; w and l for BSIM4v7 devices."""
        nd = int(self.device_offsets()[-1])
        arrs = [None if a is None else as_f64(np.asarray(a, np.float64).reshape(nd)) for a in (value, w, l)]
        ptr = [None if a is None else f64p(a) for a in arrs]
        check(lib().nsa_bundle_set_device_values(self.h, *ptr))

    def device_values(self):
        """The current per-device values (R, C, V, I), W and L (BSIM4v7):
        arrays over all devices, zero where a kind does not apply."""
        nd = int(self.device_offsets()[-1])
        value, w, l = np.zeros(max(nd, 1)), np.zeros(max(nd, 1)), np.zeros(max(nd, 1))
        check(lib().nsa_bundle_get_device_values(self.h, f64p(value), f64p(w), f64p(l)))
        return value[:nd], w[:nd], l[:nd]

    def device_offsets(self):
        d = np.zeros(self.N + 1, np.int32)
        check(lib().nsa_bundle_device_offsets(self.h, i32p(d)))
        return d

    def dc_adjoint(self, dLdx):
        """Gradients of a loss through every instance's operating point,
        given dL/dx (total unknowns, packed): dict with lambda (packed) and,
        per device (instance i's device d at device_offsets()[i] + d),
        value (dL/dR, dL/dC = 0, dL/dV, dL/dI), w and l (BSIM4v7 dL/dW, dL/dL)."""
        g = as_f64(np.asarray(dLdx, np.float64).reshape(self.total_n))
        nd = int(self.device_offsets()[-1])
        out = {"lambda": np.zeros(self.total_n), "value": np.zeros(max(nd, 1)), "w": np.zeros(max(nd, 1)),
               "l": np.zeros(max(nd, 1))}
        check(lib().nsa_bundle_dc_adjoint(self.h, f64p(g), f64p(out["lambda"]), f64p(out["value"]),
                                          f64p(out["w"]), f64p(out["l"])))
        for k in ("value", "w", "l"):
            out[k] = out[k][:nd]
        return out

    def tran_begin(self, tstep, tstop, probes=None, max_points=1):
        """solve_transient's preamble for every instance; probes: one list
        of unknown indices per instance (or None)."""
        if probes is not None:
            ptr = np.zeros(self.N + 1, np.int32)
            flat = []
            for i, p in enumerate(probes):
                flat.extend(int(v) for v in p)
                ptr[i + 1] = len(flat)
            flat = as_i32(flat if flat else [0])
            self._nprobe = np.diff(ptr)
            check(lib().nsa_bundle_tran_begin(self.h, float(tstep), float(tstop), i32p(ptr), i32p(flat),
                                              int(max_points)))
        else:
            self._nprobe = np.zeros(self.N, np.int32)
            check(lib().nsa_bundle_tran_begin(self.h, float(tstep), float(tstop), None, None, 1))

    def tran_advance(self, t_until):
        """Adaptive steps until every instance reaches t_until; True when all
        instances are finished."""
        r = C.c_int32()
        rc = check(lib().nsa_bundle_tran_advance(self.h, float(t_until), C.byref(r)))
        self.last_rounds = r.value
        return rc == 1

    def tran(self, tstep, tstop, probes=None, max_points=1):
        self.tran_begin(tstep, tstop, probes, max_points)
        self.tran_advance(tstop)

    def tran_status(self, i):
        o = np.zeros(8, np.int64)
        check(lib().nsa_bundle_tran_status(self.h, int(i), o.ctypes.data_as(_lib.P_I64)))
        return {"accepted": int(o[0]), "rejected": int(o[1]), "failed": bool(o[2]), "finished": bool(o[3]),
                "newton_iters": int(o[4]), "all_iters": int(o[5]), "loop_passes": int(o[6]),
                "dc_iters": int(o[7])}

    def tran_steps(self, i):
        """Accepted steps of instance i: dict of arrays t, dt, order, method, iters."""
        n = self.tran_status(i)["accepted"]
        t, dt = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        order, method, iters = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        check(lib().nsa_bundle_tran_steps(self.h, int(i), f64p(t), f64p(dt), i32p(order), i32p(method),
                                          i32p(iters)))
        return {"t": t[:n], "dt": dt[:n], "order": order[:n], "method": method[:n], "iters": iters[:n]}

    def tran_probes(self, i):
        """Recorded probe rows of instance i: (rows, nprobe) array (row 0: t = 0)."""
        npb = int(self._nprobe[i])
        rows = C.c_int32()
        check(lib().nsa_bundle_tran_probes(self.h, int(i), C.byref(rows), None))
        out = np.zeros((max(rows.value, 1), max(npb, 1)))
        check(lib().nsa_bundle_tran_probes(self.h, int(i), C.byref(rows), f64p(out)))
        return out[:rows.value, :npb]

    def stats(self):
        s = np.zeros(8, np.int64)
        check(lib().nsa_bundle_stats(self.h, s.ctypes.data_as(_lib.P_I64)))
        return {"iterations": int(s[0]), "cpu_pivot": int(s[1]), "plans": int(s[2]),
                "graph_launches": int(s[3]), "host_iterations": int(s[4]),
                "max_refactor_levels": int(s[5]), "max_solve_levels": int(s[6]), "max_nnz_lu": int(s[7])}

    def set_timing(self, on=True):
        check(lib().nsa_bundle_set_timing(self.h, int(bool(on))))

    def timing(self):
        t = np.zeros(5)
        check(lib().nsa_bundle_timing(self.h, f64p(t)))
        return {"load_ms": float(t[0]), "lu_ms": float(t[1]), "post_ms": float(t[2]),
                "bsim4_ms": float(t[3]), "iterations_timed": int(t[4])}

    def set_graph(self, on=True):
        check(lib().nsa_bundle_set_graph(self.h, int(bool(on))))

    def tick_stats(self):
        """Device-controlled transient: ticks, host looks between batches,
        host seconds enqueueing ticks and waiting for the device."""
        t = np.zeros(4)
        check(lib().nsa_bundle_tick_stats(self.h, f64p(t)))
        return {"ticks": int(t[0]), "batches": int(t[1]), "enqueue_s": float(t[2]), "wait_s": float(t[3])}

    def set_tuning(self, on=True):
        """The packed LU's schedule chosen by timing at a transient's start
        (on by default; the same results under every schedule)."""
        check(lib().nsa_bundle_set_tuning(self.h, int(bool(on))))

    def tuning(self):
        """The last tuning: the chosen schedule and every candidate's time."""
        k = check(lib().nsa_bundle_tuning(self.h, None, 0))
        if k <= 0:
            return None
        o = np.zeros(k, np.float64)
        check(lib().nsa_bundle_tuning(self.h, f64p(o), k))
        nc = int(o[1])
        cands = [{"min_serial_run": int(o[2 + 3 * c]), "stage": bool(o[3 + 3 * c]), "us": float(o[4 + 3 * c])}
                 for c in range(nc)]
        return {"chosen": cands[int(o[0])], "candidates": cands}

    def set_device_control(self, on=True):
        """True (default): the transient's step controller runs on the device
        (every instance on its own step sequence, no host round trip per
        step); False: on the host between batched Newton rounds."""
        check(lib().nsa_bundle_set_device_control(self.h, int(bool(on))))

    def plan_info(self, i):
        """Pivot structure of instance i: plan id, sizes, levels and the
        operation counts of one refactor and one forward + backward solve."""
        o = np.zeros(9, np.int64)
        check(lib().nsa_bundle_plan_info(self.h, int(i), o.ctypes.data_as(_lib.P_I64)))
        keys = ("plan", "n", "nnz_a", "nnz_lu", "levels", "bwd_levels", "flops_refactor", "flops_solve",
                "plan_words")
        return {k: int(v) for k, v in zip(keys, o)}
