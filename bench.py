#!/usr/bin/env python3
"""Benchmark: batched Newton iterations/s of the uneven-size BSIM4v7
ring-oscillator batch, against the same kernels at batch 1.

Headline workload: 256 BSIM4v7 ring oscillators (LEVEL=14 default cards,
TOXE 1.8 nm, VTH0 +-0.4 V, W/L 0.5u/0.05u N and 1u/0.05u P, VDD 1 V) whose
stage counts are drawn uniformly from the odd numbers 31..201, every
transistor with its own Gaussian mismatch (3 % VTH0, 1 % TOXE); 2 ns of
adaptive trapezoidal transient with the step capped at 1 ps, every
instance with its own step controller, in one bundle (all 256 circuits in
every launch, the Newton iteration replayed from a captured HIP graph).

A bench step is one window of tstop / K of that transient for all 256
instances, so the K timed steps cover the whole 2 ns; the warm-up steps run
on a separate bundle of the same circuits.  value = sum over instances of
Newton iterations / wall time of the K windows (the operating point before
them is reported apart).

Also measured in the same run: the batch-1 GPU path (the same kernels, one
instance per bundle) on a stratified sample of the instances, the parity of
several instances against the reference's own adaptive run (oracle/_ref),
HIP-event times of every packed-LU launch over a replay of the whole
transient with its roofline, the reference CPU path on the host cores (one
core and all cores), the linear RC-mesh batch, and the second half of the
metric, adjoint gradient time / forward time.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def profile_value(name, key):
    """A figure from a committed rocprofv3 summary under profiles/ (a
    markdown table row `| key | value |`), or None when it is absent."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 2 and cells[0] == key:
                return float(cells[1])
    except (OSError, ValueError):
        pass
    return None


def measured_traffic():
    """HBM bytes per batched refactor of this workload from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE scaled by a plain copy's counters), or
    None when it is absent."""
    path = os.path.join(ROOT, "profiles", "k_refactor_pmc_rc_mesh_64x64_b256.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_refactor:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def lu_traffic(eng, ckt, B):
    """Algorithmic bytes and flops per refactor and per solve (per batch)."""
    st = eng.stats()
    nnz_lu = st["nnz_lu"]
    n, nnz_a = ckt.n, ckt.nnz
    refactor_bytes = 8.0 * (nnz_a + nnz_lu) * B
    solve_bytes = (8.0 * nnz_lu + 16.0 * n) * B
    return refactor_bytes, solve_bytes, st


def refactor_flops(eng1, ckt, B):
    """F_refactor = sum_j sum_{r in U(:,j)} (1 + 2 |L(:,r)|_offdiag) + n over the
    pivot structure the host fixes for this matrix (same pattern and values)."""
    import neospice_amd as na
    vals, _ = eng1.load()
    s = na.pivot_structure("amdlu", ckt.n, ckt.col_ptr, ckt.row_idx, vals[:, 0])
    seg, nU, rows = s["seg_ptr"], s["nU"], s["rows"]
    nl = np.array([seg[k + 1] - seg[k] - nU[k] - 1 for k in range(ckt.n)])
    f = ckt.n
    for k in range(ckt.n):
        f += int(np.sum(1 + 2 * nl[rows[seg[k]:seg[k] + nU[k]]]))
    return float(f) * B


def run_steps(eng, na, dt, first_step, count, t0_step):
    iters = 0
    for k in range(count):
        step = t0_step + k
        ctx = na.be_ctx(dt, (step + 1) * dt, first=(step == 0))
        conv, its = eng.tran_step(ctx)
        if not conv.all():
            raise RuntimeError("Newton failed at step %d for %d instances" % (step, (~conv).sum()))
        eng.rotate_states()
        iters += int(its.sum())
    return iters


def mesh_cpu_baseline(ckt, info, R, dt, nsteps, budget_s, parity_steps, x_gpu0):
    """CPU-baseline leg: the reference's own CPU Newton path (oracle/_ref) on
    whole instances of the same workload, one instance per host thread
    (ctypes releases the GIL).  The same leg checks instance 0 of the GPU run
    against the reference after the same step sequence."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refbind as RB
    if not RB.available():
        return None, None
    res_devs = info["res_devs"]
    vals0 = {int(d): float(R[k, 0]) for k, d in enumerate(res_devs)}
    rc = RB.RefCircuit(ckt, values=vals0, options={"xmu": 0.0}, probe=False)
    x_ref, _, _ = rc.tran_fixed(dt, parity_steps)
    parity = float(np.max(np.abs(x_gpu0 - x_ref)))
    cores = min(os.cpu_count() or 1, 16)
    # one circuit per host thread, built before the timed window (building
    # goes through Python and holds the GIL; the stepping runs in C)
    circuits = []
    for tid in range(cores):
        vals = {int(d): float(R[k, tid % R.shape[1]]) for k, d in enumerate(res_devs)}
        circuits.append(RB.RefCircuit(ckt, values=vals, options={"xmu": 0.0}, probe=False))
    def leg(nthreads, seconds):
        tot = [[0, 0.0, 0] for _ in range(nthreads)]   # iterations, stepping seconds, runs
        deadline = time.perf_counter() + seconds

        def worker(tid):
            while time.perf_counter() < deadline:
                _, its, secs = circuits[tid].tran_fixed(dt, nsteps)
                tot[tid][0] += its
                tot[tid][1] += secs
                tot[tid][2] += 1

        t0 = time.perf_counter()
        th = [threading.Thread(target=worker, args=(t,)) for t in range(nthreads)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        wall = time.perf_counter() - t0
        rate = sum(it / sec for it, sec, _ in tot if sec > 0)
        return rate, sum(it for it, _, _ in tot), sum(r for _, _, r in tot), wall

    rate1, _, _, _ = leg(1, budget_s * 0.25)
    rate, iters, runs, wall = leg(cores, budget_s * 0.75)
    cpu = {"value": round(rate, 2), "unit": "Newton iters/s", "cores": cores,
           "kind": "reference",
           "value_1core": round(rate1, 2),
           "cpu_model": cpu_model(),
           "sample": "%d runs of %d BE steps (%d Newton iterations) of the 64x64 mesh, one "
                     "instance per host thread on %d threads within %.1f s (plus a 1-thread leg); "
                     "rate = sum over threads of iterations / stepping time; reference CPU path: "
                     "newton_solve with AMD-LU replay refactor/solve and device loads" %
                     (runs, nsteps, iters, cores, wall)}
    return cpu, parity


def rc_mesh(na, W, args, B=256, dt=1e-12):
    """The linear batched-refactor workload: 64x64 RC mesh (4096 nodes + 1
    source branch), branch R log-uniform [0.1, 10] ohm, one decap per node
    log-uniform [10 fF, 1 pF], 5% of nodes with DC current sinks in [0, 1 mA],
    1.0 V at a corner; 256 instances with per-instance +-10% Gaussian R
    perturbation; backward-Euler steps (xmu = 0) of 1 ps.  A step is one
    transient timestep of all instances (newton_solve: device load + gather,
    refactor on the fixed AMD-LU order, solve, convergence)."""
    import torch
    ckt, info = W.rc_mesh(args.mesh)
    ckt.finalize()
    R = W.mesh_instance_resistors(info, B)
    log("bench: mesh %dx%d n=%d nnz=%d, batch %d" % (args.mesh, args.mesh, ckt.n, ckt.nnz, B))
    eng = na.BatchEngine(ckt, batch=B, xmu=0.0)
    for k, d in enumerate(info["res_devs"]):
        eng.set_values(int(d), R[k])
    eng.tran_begin()
    warm = 3
    run_steps(eng, na, dt, True, warm, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = run_steps(eng, na, dt, False, args.mesh_steps, warm)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = iters / wall
    log("bench: mesh batched %d steps, %d instance-iterations in %.3f s -> %.1f iters/s"
        % (args.mesh_steps, iters, wall, value))
    nph = 5
    eng.set_timing(True)
    it_ph = run_steps(eng, na, dt, False, nph, warm + args.mesh_steps)
    torch.cuda.synchronize()
    timing = eng.timing()
    eng.set_timing(False)
    refactor_bytes, solve_bytes, st = lu_traffic(eng, ckt, B)
    x_gpu0 = eng.solution()[:, 0].copy()
    total_steps = warm + args.mesh_steps + nph

    e1 = na.BatchEngine(ckt, batch=1, xmu=0.0)
    for k, d in enumerate(info["res_devs"]):
        e1.set_values(int(d), R[k, :1])
    e1.tran_begin()
    run_steps(e1, na, dt, True, warm, 0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    it1 = run_steps(e1, na, dt, False, args.mesh_batch1_steps, warm)
    torch.cuda.synchronize()
    w1 = time.perf_counter() - t1
    batch1_value = it1 / w1

    nit = max(1, it_ph // B)
    ms_refactor = timing["refactor_ms"] / nit
    ms_solve = timing["solve_ms"] / nit
    ms_load = timing["load_ms"] / nit
    fl = refactor_flops(e1, ckt, B)
    achieved = refactor_bytes / (ms_refactor * 1e-3) / 1e9
    traffic = measured_traffic() if (args.mesh == 64 and B == 256) else None
    roofline = {"bound": "hbm", "kernel": "k_refactor", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "launch_ms": round(ms_refactor, 4), "bytes_per_launch": refactor_bytes,
                "flops_per_launch": fl,
                "fp64_TFLOPs": round(fl / (ms_refactor * 1e-3) / 1e12, 4),
                "fp64_frac": round(fl / (ms_refactor * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 6),
                "traffic_source": ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a plain "
                                   "copy (profiles/k_refactor_pmc_rc_mesh_64x64_b256.md)" if traffic else None)}
    phases = {"load_gather_ms": ms_load, "refactor_ms": ms_refactor, "solve_ms": ms_solve,
              "solve_GBs": solve_bytes / (ms_solve * 1e-3) / 1e9,
              "levels_refactor": st["levels_refactor"], "levels_solve": st["levels_solve"],
              "nnz_lu": st["nnz_lu"], "pivot_groups": st["pivot_groups"]}
    cpu, parity = None, None
    if not args.no_cpu:
        try:
            cpu, parity = mesh_cpu_baseline(ckt, info, R, dt, 10, args.cpu_seconds * 0.5, total_steps, x_gpu0)
        except Exception as e:  # the baseline is reported, not required
            log("bench: mesh cpu baseline unavailable:", e)
    return {"workload": "rc_mesh_%dx%d_batched_refactor_be_1ps_b%d" % (args.mesh, args.mesh, B),
            "value": round(value, 2), "unit": "Newton iters/s", "steps": args.mesh_steps,
            "ms_per_step": round(wall / args.mesh_steps * 1e3, 4),
            "unknowns": ckt.n, "nnz": ckt.nnz, "solver": eng.solver_name,
            "batch1_gpu_value": round(batch1_value, 2),
            "speedup_vs_batch1_gpu": round(value / batch1_value, 2),
            "parity_max_abs_diff_vs_reference": parity,
            "phases": phases, "roofline": roofline, "cpu_baseline": cpu}



def ac_sweep(na, W, args, m=16, B=256, nf=64, cpu=True):
    """Batched AC (the (f)2 row): the AC analysis's frequency sweep of a
    16x16 RC mesh (257 unknowns), 256 instances with per-instance +-10 %
    Gaussian R, 64 log-spaced frequencies 1 kHz .. 1 THz, unit AC drive on
    the source's branch.  Instance i's matrix at frequency f is
    g_i + j 2 pi f c_i, g the conductances and c the capacitances on the
    pattern (both from the library's own loads: a DC load, and a transient
    load with ag0 = 1 whose difference is c); each instance ordered at the
    first frequency, every (instance, frequency) factored on that order and
    solved in one launch (ac.cpp:196-238).  A step is one whole sweep; value =
    frequency points (instances x frequencies) per second, wall time of the
    call with its host orderings and host<->device copies."""
    import torch
    ckt, info = W.rc_mesh(m)
    ckt.finalize()
    R = W.mesh_instance_resistors(info, B)
    eng = na.BatchEngine(ckt, batch=B, xmu=0.0)
    for k, d in enumerate(info["res_devs"]):
        eng.set_values(int(d), R[k])
    eng.tran_begin()
    eng.set_ctx(na.make_ctx(na.MODEDCOP))
    g, _ = eng.load()
    eng.set_ctx(na.make_ctx(na.MODETRAN | na.MODEINITPRED, 1, 0, (1.0, -1.0, 0.0)))
    a, _ = eng.load()
    c = a - g
    n = ckt.n
    rhs = np.zeros((n, B), np.complex128)
    rhs[n - 1, :] = 1.0                         # the voltage source's branch row: unit AC drive
    freqs = np.logspace(3, 12, nf)
    s = na.Solver(n, ckt.col_ptr, ckt.row_idx, batch=B)
    x, sing = s.ac_sweep(g, c, freqs, rhs)      # warm-up (allocations, plans)
    torch.cuda.synchronize()
    reps = 3
    dev = order = inside = 0.0
    t0 = time.perf_counter()
    for _ in range(reps):
        x, sing = s.ac_sweep(g, c, freqs, rhs)
        st = s.ac_stats()
        dev += st["device_s"]
        order += st["order_s"]
        inside += st["total_s"]
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    dev, order, inside = dev / reps, order / reps, inside / reps
    value = B * nf / dev
    log("bench: AC sweep %d instances x %d frequencies, n=%d: device %.3f ms (%.0f points/s), host orderings "
        "%.2f ms, library call %.2f ms, Python call %.2f ms" % (B, nf, n, dev * 1e3, value, order * 1e3,
                                                                  inside * 1e3, wall * 1e3))
    res = {"workload": "rc_mesh_%dx%d_ac_sweep_b%d_f%d" % (m, m, B, nf), "value": round(value, 2),
           "unit": "frequency points/s", "device_ms_per_sweep": round(dev * 1e3, 3),
           "host_order_ms": round(order * 1e3, 3), "library_call_ms": round(inside * 1e3, 3),
           "python_call_ms": round(wall * 1e3, 3),
           "points_per_s_end_to_end": round(B * nf / wall, 2), "unknowns": n,
           "nnz": ckt.nnz, "instances": B, "frequencies": nf, "singular_points": int(sing.sum()),
           "note": "value: device work with the matrices' g, c resident (build, factor, solve of every point); "
                   "the call also orders each instance on the host (at the first frequency) and copies "
                   "host buffers both ways"}
    if cpu:
        # the reference's own loop (ac.cpp:215-222) on a NeoSolver per
        # instance, for a bounded sample of the instances; also the parity
        # check of those instances, bit for bit
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import refbind as RB
        sample = [0, B // 2, B - 1]
        t0 = time.perf_counter()
        same = True
        for i in sample:
            r = RB.RefSolver(0, n, ckt.col_ptr, ckt.row_idx)
            for f, fr in enumerate(freqs):
                w = 2.0 * np.pi * fr
                ax = g[:, i] + 1j * (w * c[:, i])
                if f == 0:
                    r.numeric_complex(ax)
                else:
                    r.refactorize_complex(ax)
                y = r.solve_complex(rhs[:, i])
                same = same and bool(np.array_equal(y, x[f, :, i]))
        tc = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(len(sample) * nf / tc, 2), "unit": "frequency points/s", "cores": 1,
                               "kind": "reference", "sample": "%d instances x %d frequencies, NeoSolver loop of "
                               "ac.cpp on one core" % (len(sample), nf)}
        res["parity_bit_equal"] = same
        log("bench: AC sweep CPU reference %.0f points/s (1 core), sample bit-equal %s"
            % (res["cpu_baseline"]["value"], same))
    return res


def adjoint_c4(na, W, count=1024):
    """DC-operating-point adjoint of the two-stage Miller op-amp batch: 1024
    seeded instances (8 BSIM4v7 transistors with W in [0.2u, 5u] and L in
    [0.05u, 0.5u], RC compensation and the 10k / 1pF load scaled by
    U[0.8, 1.2]) at 4 input biases 0.3..0.6 V, 4096 circuits in one bundle;
    loss = sum over the biases of (Vout - 0.5)^2; gradient w.r.t. every W,
    L, R and C through the torch front end (BundleDCOperatingPoint).
    forward = solve_dc's ladder of all 4096 circuits; grad = the backward
    (transposed solves, R/V/I gradients, dual-number BSIM4v7 W and L
    gradients).  Timed after one warm-up forward and backward (pivot orders
    kept from it)."""
    import torch
    from neospice_amd.autograd import BundleDCOperatingPoint
    inst = W.opamp_instances(count)
    t0 = time.perf_counter()
    cks, outs = [], []
    for p in inst:
        for v in W.OPAMP_BIASES:
            c, info = W.opamp_miller(p["w"], p["l"], rc=p["r"][0], rl=p["r"][1], rb=p["r"][2], cc=p["c"][0],
                                     cl=p["c"][1], vin=v)
            c.finalize()
            cks.append(c)
            outs.append(info["out"])
    b = na.Bundle(cks)
    build_s = time.perf_counter() - t0
    out_idx = torch.tensor([int(b.xoff[j] + o) for j, o in enumerate(outs)])
    value, w, l = b.device_values()

    def step():
        tv, tw, tl = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in (value, w, l))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x = BundleDCOperatingPoint.apply(b, tv, tw, tl)
        loss = ((x[out_idx] - 0.5) ** 2).sum()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        return t2 - t1, t3 - t2, float(loss.detach()), tw.grad

    cold_f, cold_g, _, _ = step()
    it0 = b.stats()["iterations"]
    fwd, grad, loss, gw = step()
    its = b.stats()["iterations"] - it0
    log("bench: C4 forward %.4f s (cold %.4f), gradient %.4f s (cold %.4f), ratio %.3f" %
        (fwd, cold_f, grad, cold_g, grad / fwd))
    return {"workload": "miller_opamp_bsim4v7_%dx%d_biases_dc_adjoint" % (count, len(W.OPAMP_BIASES)),
            "circuits": len(cks), "unknowns_total": b.total_n, "params": int(2 * 8 * count + 5 * count),
            "forward_ms": round(fwd * 1e3, 3), "grad_ms": round(grad * 1e3, 3),
            "grad_over_forward": round(grad / fwd, 4),
            "forward_newton_iterations": int(its), "loss": loss,
            "cold_forward_ms": round(cold_f * 1e3, 3), "cold_grad_ms": round(cold_g * 1e3, 3),
            "build_s": round(build_s, 3),
            "note": "forward = DC ladder (MODEDCOP) of every circuit from zero; grad = transposed packed "
                    "solves + dL/dR,V,I + dual-number dL/dW,dL/dL kernels; both through torch.autograd"}


def adjoint_c5(na, W, mesh=32, batch=128, nsteps=500, dt=2e-12, ckpt=25, nprobe=64):
    """Transient adjoint of the diode-clamped RC mesh (32x32, diode IS=1e-14
    N=1 per node, 16-point PWL drive): 500 trapezoidal steps of 2 ps over
    128 instances with seeded per-instance +-10% R and C perturbation; loss
    = L2 error of 64 seeded probe waveforms against a seeded target;
    gradient w.r.t. every R and C with recomputation checkpoints every 25
    steps.  grad time includes the checkpoint recomputation."""
    ckt, _ = W.diode_mesh(mesh)
    ckt.finalize()
    eng = na.BatchEngine(ckt, batch=batch)
    R, C = eng.family_values("R"), eng.family_values("C")
    for i, r in enumerate(W.instance_rngs(5, batch)):
        R[:, i] *= 1 + 0.1 * r.standard_normal(R.shape[0])
        C[:, i] *= 1 + 0.1 * r.standard_normal(C.shape[0])
    eng.set_family_values("R", R)
    eng.set_family_values("C", C)
    g = np.random.default_rng(W.BASE_SEED + 5)
    probes = np.sort(g.choice(ckt.num_nodes, nprobe, replace=False)).astype(np.int32)
    target = g.uniform(0.0, 0.8, (nsteps + 1, nprobe, 1))
    st0 = eng.stats()
    t0 = time.perf_counter()
    v = eng.tran_forward(nsteps, dt, probes, ckpt_every=ckpt)
    t1 = time.perf_counter()
    its = eng.stats()["gpu_refactor"] + eng.stats()["cpu_pivot"] - st0["gpu_refactor"] - st0["cpu_pivot"]
    eng.tran_adjoint(2.0 * (v - target))
    t2 = time.perf_counter()
    fwd, grad = t1 - t0, t2 - t1
    log("bench: adjoint forward %.3f s, gradient %.3f s, ratio %.3f" % (fwd, grad, grad / fwd))
    return {"workload": "diode_mesh_%dx%d_trap_%dx%gps_b%d" % (mesh, mesh, nsteps, dt * 1e12, batch),
            "unknowns": ckt.n, "nnz": ckt.nnz, "instances": batch, "steps": nsteps,
            "probes": nprobe, "checkpoint_every": ckpt, "params": int(R.shape[0] + C.shape[0]),
            "forward_ms": round(fwd * 1e3, 2), "grad_ms": round(grad * 1e3, 2),
            "grad_over_forward": round(grad / fwd, 3),
            "batched_factorizations_forward": int(its),
            "note": "grad = checkpoint recomputation + backward sweep (one Jacobian load, refactor "
                    "and transposed solve per step)"}



# ------------------------------------------------------------------ C3

C3_TSTEP = 1e-12
C3_TSTOP = 2e-9               # the workload's transient; --tstop shortens it for profiler runs
C3_MAX_POINTS = 4000          # accepted points per instance (2 ns / 1 ps + breakpoint steps)


def c3_probes(c):
    n = c.n
    return sorted({1, n // 4, n // 2, (3 * n) // 4, n - 2, n - 1})


def c3_run(na, cks, probes, windows, timing=False):
    """One adaptive transient of the whole batch in `windows` equal windows:
    (bundle, seconds of the operating point, seconds of the windows,
    transient Newton iterations)."""
    import torch
    b = na.Bundle(cks)
    t0 = time.perf_counter()
    b.tran_begin(C3_TSTEP, C3_TSTOP, probes=probes, max_points=C3_MAX_POINTS)
    t1 = time.perf_counter()
    it0 = b.stats()["iterations"]
    if timing:
        b.set_timing(True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    done = False
    for k in range(windows):
        done = b.tran_advance(C3_TSTOP * (k + 1) / windows)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    if not done:
        raise RuntimeError("C3: the transient did not finish")
    return b, t1 - t0, t3 - t2, b.stats()["iterations"] - it0


def c3_check_status(b, N):
    acc = [b.tran_status(i) for i in range(N)]
    bad = [i for i, a in enumerate(acc) if a["failed"] or not a["finished"]]
    if bad:
        raise RuntimeError("C3: instances %s failed or did not finish" % bad[:8])
    return acc


def c3_reference(cks, pick, probes, b, budget_s):
    """The reference's own adaptive transient (solve_transient through
    oracle/_ref) of the picked instances: their parity against the batched
    run, and the one-core CPU rate of those runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refbind as RB
    if not RB.available():
        return None, None
    rows, its_tot, sec_tot = [], 0, 0.0
    t_start = time.perf_counter()
    for i in pick:
        if rows and time.perf_counter() - t_start > budget_s:
            break
        rc = RB.RefCircuit(cks[i], probe=False)
        t_ref, x_ref, its, sec = rc.tran(C3_TSTEP, C3_TSTOP, max_points=C3_MAX_POINTS)
        its_tot += its
        sec_tot += sec
        st = b.tran_status(i)
        steps = b.tran_steps(i)
        row = {"instance": int(i), "unknowns": cks[i].n, "accepted": st["accepted"],
               "accepted_reference": len(t_ref) - 1, "reference_newton_iters": its}
        if st["accepted"] == len(t_ref) - 1:
            row["max_rel_t"] = float(np.max(np.abs(steps["t"] - t_ref[1:]) / t_ref[1:]))
            x = b.tran_probes(i)
            xr = x_ref[:, probes[i]]
            row["max_rel_wave"] = float(np.max(np.abs(x - xr) / np.maximum(np.abs(xr), 1e-3)))
        rows.append(row)
        log("bench: C3 parity", row)
    ok = all(r["accepted"] == r["accepted_reference"] and r.get("max_rel_wave", 1.0) <= 1e-6 for r in rows)
    parity = {"instances": rows, "within_1e-6": bool(ok),
              "note": "each instance's accepted steps and probed waveforms against the reference's own "
                      "adaptive solve_transient of that ring (rel err vs max(|x|, 1e-3))"}
    rate1 = its_tot / sec_tot if sec_tot > 0 else None
    return parity, (rate1, its_tot, sec_tot, len(rows))


def c3_cpu_all_cores(cks, seconds):
    """The reference's adaptive transient of whole instances, one per host
    thread (ctypes releases the GIL), instances taken round robin."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refbind as RB
    cores = min(os.cpu_count() or 1, 16)
    tot = [[0, 0.0, 0] for _ in range(cores)]
    deadline = time.perf_counter() + seconds

    def worker(tid):
        i = tid
        while time.perf_counter() < deadline:
            rc = RB.RefCircuit(cks[i % len(cks)], probe=False)
            _, _, its, sec = rc.tran(C3_TSTEP, C3_TSTOP, max_points=C3_MAX_POINTS)
            tot[tid][0] += its
            tot[tid][1] += sec
            tot[tid][2] += 1
            i += cores

    t0 = time.perf_counter()
    th = [threading.Thread(target=worker, args=(t,)) for t in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    rate = sum(it / sec for it, sec, _ in tot if sec > 0)
    return rate, cores, sum(r for _, _, r in tot), sum(it for it, _, _ in tot), wall


def c3_headline(na, W, args):
    import torch
    global C3_TSTOP
    C3_TSTOP = args.tstop
    N = args.instances
    cks, stages = W.c3_rings(N)
    for c in cks:
        c.finalize()
    probes = [c3_probes(c) for c in cks]
    log("bench: C3 %d rings, stages %d..%d (mean %.1f)" % (N, stages.min(), stages.max(), stages.mean()))
    K = max(1, args.steps)
    # warm-up: the first W windows on a bundle of their own
    if args.warmup > 0:
        bw = na.Bundle(cks)
        bw.tran_begin(C3_TSTEP, C3_TSTOP, probes=None)
        for k in range(args.warmup):
            bw.tran_advance(C3_TSTOP * (k + 1) / K)
        bw.close()
    b, dc_s, wall, iters = c3_run(na, cks, probes, K)
    acc = c3_check_status(b, N)
    value = iters / wall
    st = b.stats()
    ticks = b.tick_stats()
    log("bench: C3 batched: %d windows, %d Newton iterations in %.3f s -> %.0f iters/s (operating point %.3f s)"
        % (K, iters, wall, value, dc_s))
    log("bench: C3 ticks", ticks)
    info = [b.plan_info(i) for i in range(N)]

    # HIP events around every packed-LU launch over a replay of the whole
    # transient (the kernels launched one by one on the bundle's stream)
    bt, _, wall_t, _ = c3_run(na, cks, probes, K, timing=True)
    tim = bt.timing()
    its_i = np.array([bt.tran_status(i)["all_iters"] for i in range(N)], dtype=np.float64)
    bt.set_timing(False)
    launches = max(1, tim["iterations_timed"])
    lu_ms = tim["lu_ms"] / launches
    # algorithmic bytes of one refactor + forward/backward solve of an
    # instance: A read, L+U written by the refactor and read by the solve,
    # rhs in and x out; summed over the instance-iterations of the replay
    per_inst = np.array([8.0 * (p["nnz_a"] + 2 * p["nnz_lu"]) + 16.0 * p["n"] for p in info])
    flops_inst = np.array([float(p["flops_refactor"] + p["flops_solve"]) for p in info])
    # plan indices count once per launch for each distinct plan in use
    plan_words, plan_its = {}, {}
    for i, p in enumerate(info):
        plan_words[p["plan"]] = p["plan_words"]
        plan_its[p["plan"]] = max(plan_its.get(p["plan"], 0.0), its_i[i])
    index_bytes = sum(4.0 * w * min(1.0, plan_its[q] / launches) for q, w in plan_words.items())
    bytes_launch = float(np.dot(its_i, per_inst)) / launches + index_bytes
    flops_launch = float(np.dot(its_i, flops_inst)) / launches
    achieved = bytes_launch / (lu_ms * 1e-3) / 1e9
    # the launch lasts as long as its longest chain of dependent levels
    crit = max(p["levels"] + p["bwd_levels"] for p in info)
    traffic = profile_value("k_small_lu_pmc_c3_rings_256.md", "traffic_bytes_per_launch") if N == 256 else None
    roofline = {"bound": "hbm", "kernel": "k_small_lu", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "launch_ms": round(lu_ms, 5), "launches": int(launches),
                "bytes_per_launch": round(bytes_launch, 1),
                "index_bytes_per_launch": round(index_bytes, 1),
                "flops_per_launch": round(flops_launch, 1),
                "fp64_TFLOPs": round(flops_launch / (lu_ms * 1e-3) / 1e12, 5),
                "fp64_frac": round(flops_launch / (lu_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 7),
                "critical_path_levels": int(crit),
                "ns_per_level": round(lu_ms * 1e6 / crit, 2),
                "timing": "HIP events on the bundle stream around each k_small_lu launch, %d launches over "
                          "a replay of the whole 2 ns transient" % launches,
                "traffic_source": ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the same "
                                   "counters on a plain copy of known size (profiles/k_small_lu_pmc_c3_rings_256.md)"
                                   if traffic is not None else None)}
    kernels = {"iteration_ms": {"load_gather": tim["load_ms"] / launches, "k_bsim4": tim["bsim4_ms"] / launches,
                                "k_small_lu": lu_ms, "convergence": tim["post_ms"] / launches},
               "replay_wall_s": round(wall_t, 4)}
    log("bench: C3 k_small_lu %.4f ms/launch, %.1f GB/s, kernels %s" % (lu_ms, achieved, kernels))

    # batch-1 GPU path: the same kernels, one instance per bundle, on a
    # sample stratified by stage count
    order = np.argsort(stages, kind="stable")
    S = min(args.batch1_sample, N)
    pick = [int(order[int(round(q))]) for q in np.linspace(0, N - 1, max(S, 1))]
    it1, w1, same = 0, 0.0, True
    b1_rows = []
    for i in (pick if S > 0 else []):
        b1, _, w, it = c3_run(na, [cks[i]], [probes[i]], K)
        c3_check_status(b1, 1)
        it1 += it
        w1 += w
        eq = bool(np.array_equal(b1.tran_probes(0), b.tran_probes(i)))
        same &= eq
        b1_rows.append({"instance": i, "stages": int(stages[i]), "iters": it, "s": round(w, 4),
                        "bit_equal_to_batched": eq})
        b1.close()
    batch1_value = it1 / w1 if w1 > 0 else None
    log("bench: C3 batch-1 sample %s -> %s iters/s" % (b1_rows, batch1_value))

    parity, cpu1 = None, None
    cpu = None
    if not args.no_cpu:
        try:
            parity, cpu1 = c3_reference(cks, pick[::max(1, len(pick) // 4)][:4] + [pick[-1]],
                                        probes, b, args.cpu_seconds * 0.4)
            if cpu1 is not None:
                rate_all, cores, runs, its_all, wall_all = c3_cpu_all_cores(cks, args.cpu_seconds * 0.6)
                cpu = {"value": round(rate_all, 2), "unit": "Newton iters/s", "cores": cores, "kind": "reference",
                       "value_1core": round(cpu1[0], 2), "cpu_model": cpu_model(),
                       "sample": "reference solve_transient (oracle/_ref) of whole 2 ns instances of this batch: "
                                 "%d runs (%d Newton iterations) one instance per host thread on %d threads "
                                 "within %.1f s; value_1core: %d instances (%d iterations) in sequence on one "
                                 "core; rate = sum over threads of iterations / solve_transient time"
                                 % (runs, its_all, cores, wall_all, cpu1[3], cpu1[1])}
                log("bench: C3 cpu baseline", cpu)
        except Exception as e:  # the baseline is reported, not required
            log("bench: C3 reference leg unavailable:", e)

    accepted = np.array([a["accepted"] for a in acc])
    config = {"workload": "bsim4v7_rings_%d_uneven_31_201_tran_%gps_adaptive_trap_1ps" % (N, C3_TSTOP * 1e12),
              "instances": N, "stages_min": int(stages.min()), "stages_max": int(stages.max()),
              "stages_mean": round(float(stages.mean()), 2),
              "unknowns_total": b.total_n, "nnz_total": b.total_nnz,
              "tstop": C3_TSTOP, "max_step": C3_TSTEP, "integration": "trapezoidal, adaptive (device LTE)",
              "step": "one window of tstop/K of every instance's adaptive transient",
              "pivot_plans": st["plans"], "lu_schedule": b.tuning()}
    out = {"value": value, "wall": wall, "K": K, "config": config,
           "accepted_steps": {"min": int(accepted.min()), "max": int(accepted.max()),
                              "mean": round(float(accepted.mean()), 2),
                              "rejected": int(sum(a["rejected"] for a in acc))},
           "operating_point_s": round(dc_s, 4),
           "ticks": {"count": ticks["ticks"], "host_looks": ticks["batches"],
                     "us_per_tick": round(wall / max(1, ticks["ticks"]) * 1e6, 2),
                     "host_enqueue_s": round(ticks["enqueue_s"], 4), "host_wait_s": round(ticks["wait_s"], 4)},
           "batch1_gpu_value": round(batch1_value, 2) if batch1_value else None,
           "speedup_vs_batch1_gpu": round(value / batch1_value, 3) if batch1_value else None,
           "batch1_sample": b1_rows,
           "batch1_bit_equal_to_batched": bool(same),
           "parity": parity, "roofline": roofline, "kernels": kernels, "cpu_baseline": cpu}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed windows; together they cover the whole 2 ns transient")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--batch1-sample", type=int, default=8,
                    help="instances timed one per bundle for the batch-1 rate (0: skip)")
    ap.add_argument("--tstop", type=float, default=2e-9,
                    help="transient length (shorter only for profiler runs)")
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="only the ring batch")
    ap.add_argument("--ac-only", action="store_true", help="only the AC-sweep leg (its JSON)")
    ap.add_argument("--mesh", type=int, default=64)
    ap.add_argument("--mesh-steps", type=int, default=100)
    ap.add_argument("--mesh-batch1-steps", type=int, default=20)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of known size (for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(instances shard across GPUs with no exchange)" % args.gpus)
        sys.exit(2)

    import torch
    if args.calibrate_copy:
        n = 1 << 25  # 256 MiB of float64 read and written per copy
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 3}))
        return
    import neospice_amd as na
    from neospice_amd import workloads as W

    if args.ac_only:
        print(json.dumps(ac_sweep(na, W, args, cpu=not args.no_cpu)))
        return
    # the mesh leg first: its memory-bound refactor runs ~2x slower in a
    # process whose device heap the ring leg's many bundle allocations have
    # already churned (the ring leg is not affected by the order)
    mesh = None if (args.no_mesh or args.headline_only) else rc_mesh(na, W, args)
    c3 = c3_headline(na, W, args)
    ac = None if args.headline_only else ac_sweep(na, W, args, cpu=not args.no_cpu)
    adjoint = None
    if not (args.no_adjoint or args.headline_only):
        adjoint = {"dc_opamp": adjoint_c4(na, W), "transient_diode_mesh": adjoint_c5(na, W, nsteps=args.adjoint_steps)}
    out = {
        "metric": "batched Newton iters/s vs batch-1 GPU @1e-9 rel; adjoint grad/forward time",
        "value": round(c3["value"], 2),
        "unit": "Newton iters/s",
        "n_gpus": 1,
        "steps": c3["K"],
        "warmup": args.warmup,
        "ms_per_step": round(c3["wall"] / c3["K"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded BSIM4v7 rings: stage counts and per-transistor mismatch)",
        "config": c3["config"],
    }
    for k in ("batch1_gpu_value", "speedup_vs_batch1_gpu", "batch1_bit_equal_to_batched", "batch1_sample",
              "accepted_steps", "operating_point_s", "ticks", "parity", "roofline", "kernels", "cpu_baseline"):
        out[k] = c3[k]
    out["adjoint"] = adjoint
    out["rc_mesh"] = mesh
    out["ac_sweep"] = ac
    print(json.dumps(out))


if __name__ == "__main__":
    main()
