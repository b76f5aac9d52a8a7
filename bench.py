#!/usr/bin/env python3
"""Benchmark: batched Newton iterations/s of the linear RC power-grid mesh.

Workload (the batched-refactor configuration): 64x64 RC mesh (4096 nodes +
1 source branch), branch R log-uniform [0.1, 10] ohm, one decap per node
log-uniform [10 fF, 1 pF], 5% of nodes with DC current sinks in [0, 1 mA],
1.0 V at a corner; 256 instances with per-instance +-10% Gaussian R
perturbation; backward-Euler transient (xmu = 0) at dt = 1 ps.

A bench step is one transient timestep of all 256 instances: the full
newton_solve (device load + gather, GPU refactor on the fixed AMD-LU order,
GPU solve, convergence) -- two Newton iterations per step for this linear
circuit.  value = sum over instances of Newton iterations / wall time.

Also measured in the same run: the same kernels at batch 1 (the plain
batch-1 GPU path), HIP-event times of the load, refactor and solve phases
with their roofline, the reference CPU path (compiled neospice,
oracle/_ref) on the host cores on a bounded sample, and the second half of
the metric, adjoint gradient time / forward time, on the transient-adjoint
workload (32x32 diode-clamped RC mesh, 128 instances, 500 trapezoidal
steps, checkpoints every 25 steps; see adjoint_c5).
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


def ring_load_traffic():
    """HBM bytes per k_bsim4 launch on the 117-stage, 256-instance ring batch
    from the committed PMC summary, or None when it is absent."""
    path = os.path.join(ROOT, "profiles", "k_bsim4_bsim4v7_ring_117_b256.md")
    try:
        for line in open(path):
            if line.startswith("| traffic_bytes_per_launch |"):
                return float(line.split("|")[2])
    except (OSError, ValueError):
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


def cpu_baseline(ckt, info, R, dt, nsteps, budget_s, parity_steps, x_gpu0):
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
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cpu = {"value": round(rate, 2), "unit": "Newton iters/s", "cores": cores,
           "kind": "reference",
           "value_1core": round(rate1, 2),
           "cpu_model": model,
           "sample": "%d runs of %d BE steps (%d Newton iterations) of the 64x64 mesh, one "
                     "instance per host thread on %d threads within %.1f s (plus a 1-thread leg); "
                     "rate = sum over threads of iterations / stepping time; reference CPU path: "
                     "newton_solve with AMD-LU replay refactor/solve and device loads" %
                     (runs, nsteps, iters, cores, wall)}
    return cpu, parity


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


def bsim4_ring(na, W, stages=117, batch=256, steps=30, warmup=3, b1_steps=10, dt=1e-12,
               cpu_steps=10, seed=31):
    """The BSIM4v7 ring-oscillator batch: `batch` instances of a `stages`-stage
    ring (LEVEL=14 cards, TOXE 1.8 nm, VTH0 +-0.4 V, W/L 0.5u/0.05u and
    1u/0.05u, VDD 1 V) with per-transistor, per-instance Gaussian mismatch of
    3% on VTH0 and 1% on TOXE, fixed 1 ps backward-Euler steps (the adaptive
    controller's cap).  Batched and batch-1 Newton iterations/s, and the
    reference CPU path on instance 0 (its own fixed-step loop) for parity and
    as the CPU baseline."""
    import re
    import torch
    rng = np.random.default_rng(seed)
    ckt, info = W.ring_oscillator(stages, mismatch_rng=np.random.default_rng(seed + 1))
    ckt.finalize()
    texts = {op[1].split()[1].lower(): op[1] for op in ckt.ops if op[0] == "model"}
    devs = ckt.ops_devices()
    mdevs = [i for i, op in enumerate(devs) if op[0] == "M" and op[1] != "mrst"]
    mism = {}
    for dv in mdevs:
        text = texts[devs[dv][6].lower()]
        for par, sig in (("vth0", 0.03), ("toxe", 0.01)):
            base = float(re.search(par + r"=([-0-9.e+]+)", text).group(1))
            v = base * (1.0 + sig * rng.standard_normal(batch))
            v[0] = base          # instance 0 keeps the circuit's own cards
            mism[(dv, par)] = v

    def make(B):
        eng = na.BatchEngine(ckt, batch=B, xmu=0.0, itl1=400)
        for (dv, par), v in mism.items():
            eng.set_model_values(dv, par, v[:B])
        eng.tran_begin()
        return eng

    eng = make(batch)
    log("bench: ring %d stages n=%d nnz=%d, batch %d, DC done" % (stages, ckt.n, ckt.nnz, batch))
    run_steps(eng, na, dt, True, warmup, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = run_steps(eng, na, dt, False, steps, warmup)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    eng.set_timing(True)
    it_ph = run_steps(eng, na, dt, False, 3, warmup + steps)
    torch.cuda.synchronize()
    timing = eng.timing()
    eng.set_timing(False)
    nit = max(1, it_ph // batch)
    x0 = eng.solution()[:, 0].copy()
    total = warmup + steps + 3
    e1 = make(1)
    run_steps(e1, na, dt, True, warmup, 0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    it1 = run_steps(e1, na, dt, False, b1_steps, warmup)
    torch.cuda.synchronize()
    w1 = time.perf_counter() - t1
    # the BSIM4v7 load kernel against HBM: algorithmic bytes per device-instance
    # of a backward-Euler corrector load with per-instance cards -- the
    # 110-double constant record (880 B), 3 node voltages (24 B), 4 state
    # voltages in (32 B), 3 charges of state1 in (24 B), 13 state slots out
    # (104 B), 10 carried instance fields in and out (160 B), 20 stamp sites
    # out (160 B): 1384 B
    b4_bytes = 1384.0 * (len(mdevs) + 1) * batch
    b4_ms = timing["bsim4_ms"] / max(1, timing["loads"])
    b4_gbs = b4_bytes / (b4_ms * 1e-3) / 1e9
    out = {"workload": "bsim4v7_ring_%d_stages_be_1ps_b%d" % (stages, batch), "instances": batch,
           "unknowns": ckt.n, "nnz": ckt.nnz, "devices_bsim4": len(mdevs) + 1,
           "value": round(iters / wall, 2), "unit": "Newton iters/s",
           "ms_per_step": round(wall / steps * 1e3, 4),
           "batch1_gpu_value": round(it1 / w1, 2),
           "speedup_vs_batch1_gpu": round((iters / wall) / (it1 / w1), 2),
           "phases_ms_per_iter": {"load_gather": timing["load_ms"] / nit,
                                  "refactor": timing["refactor_ms"] / nit,
                                  "solve": timing["solve_ms"] / nit},
           "roofline": {"bound": "hbm", "kernel": "k_bsim4", "achieved": round(b4_gbs, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(b4_gbs / HBM_PEAK_GBS, 6),
                        "traffic": (ring_load_traffic() if (stages == 117 and batch == 256) else None),
                        "launch_ms": round(b4_ms, 5), "bytes_per_launch": b4_bytes}}
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import refbind as RB
        rc = RB.RefCircuit(ckt, probe=False)
        xr, its_r, sec = rc.tran_fixed(dt, total)
        out["parity_max_abs_diff_vs_reference"] = float(np.max(np.abs(x0 - xr)))
        rc2 = RB.RefCircuit(ckt, probe=False)
        _, its_c, sec_c = rc2.tran_fixed(dt, cpu_steps)
        out["cpu_baseline"] = {"value": round(its_c / sec_c, 2), "unit": "Newton iters/s", "cores": 1,
                               "kind": "reference",
                               "sample": "instance 0, %d fixed 1 ps BE steps on one core" % cpu_steps}
    except Exception as e:  # reported, not required
        log("bench: ring reference leg unavailable:", e)
    log("bench: ring", out)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--mesh", type=int, default=64)
    ap.add_argument("--batch1-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-ring", action="store_true")
    ap.add_argument("--ring-stages", type=int, default=117)
    ap.add_argument("--ring-only", action="store_true",
                    help="only the BSIM4v7 ring-oscillator leg (its JSON object)")
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
    if args.ring_only:
        print(json.dumps(bsim4_ring(na, W, stages=args.ring_stages, steps=args.steps,
                                    warmup=args.warmup, b1_steps=args.batch1_steps)))
        return

    dt = 1e-12
    B = args.batch
    ckt, info = W.rc_mesh(args.mesh)
    ckt.finalize()
    R = W.mesh_instance_resistors(info, B)
    log("bench: mesh %dx%d n=%d nnz=%d, batch %d" % (args.mesh, args.mesh, ckt.n, ckt.nnz, B))

    eng = na.BatchEngine(ckt, batch=B, xmu=0.0)
    for k, d in enumerate(info["res_devs"]):
        eng.set_values(int(d), R[k])
    eng.tran_begin()
    log("bench: DC preamble done, solver", eng.solver_name)
    run_steps(eng, na, dt, True, args.warmup, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = run_steps(eng, na, dt, False, args.steps, args.warmup)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = iters / wall
    log("bench: batched %d steps, %d instance-iterations in %.3f s -> %.1f iters/s"
        % (args.steps, iters, wall, value))
    # phase breakdown on a few more steps, HIP events on the engine's stream
    nph = max(1, min(5, args.steps))
    eng.set_timing(True)
    it_ph = run_steps(eng, na, dt, False, nph, args.warmup + args.steps)
    torch.cuda.synchronize()
    timing = eng.timing()
    eng.set_timing(False)
    refactor_bytes, solve_bytes, st = lu_traffic(eng, ckt, B)
    x_gpu0 = eng.solution()[:, 0].copy()
    total_steps = args.warmup + args.steps + nph

    # batch-1 GPU path: the same kernels, one instance
    e1 = na.BatchEngine(ckt, batch=1, xmu=0.0)
    for k, d in enumerate(info["res_devs"]):
        e1.set_values(int(d), R[k, :1])
    e1.tran_begin()
    run_steps(e1, na, dt, True, args.warmup, 0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    it1 = run_steps(e1, na, dt, False, args.batch1_steps, args.warmup)
    torch.cuda.synchronize()
    w1 = time.perf_counter() - t1
    batch1_value = it1 / w1
    log("bench: batch-1 %d steps, %d iterations in %.3f s -> %.1f iters/s"
        % (args.batch1_steps, it1, w1, batch1_value))

    nit = max(1, it_ph // B)      # batched Newton iterations in the phase run
    ms_refactor = timing["refactor_ms"] / nit
    ms_solve = timing["solve_ms"] / nit
    ms_load = timing["load_ms"] / nit
    fl = refactor_flops(e1, ckt, B)
    phases = {
        "load_gather_ms": ms_load, "refactor_ms": ms_refactor, "solve_ms": ms_solve,
        "refactor_GBs": refactor_bytes / (ms_refactor * 1e-3) / 1e9,
        "refactor_GFLOPs": fl / (ms_refactor * 1e-3) / 1e9,
        "refactor_flops": fl,
        "solve_GBs": solve_bytes / (ms_solve * 1e-3) / 1e9,
        "levels_refactor": st["levels_refactor"], "levels_solve": st["levels_solve"],
        "nnz_lu": st["nnz_lu"], "pivot_groups": st["pivot_groups"],
    }
    achieved = refactor_bytes / (ms_refactor * 1e-3) / 1e9
    # rocprof PMC traffic applies to the default workload only
    traffic = measured_traffic() if (args.mesh == 64 and B == 256) else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "kernel": "k_refactor",
                "launch_ms": round(ms_refactor, 4),
                "bytes_per_launch": refactor_bytes,
                "flops_per_launch": fl,
                "fp64_TFLOPs": round(fl / (ms_refactor * 1e-3) / 1e12, 4),
                "fp64_frac": round(fl / (ms_refactor * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 6),
                "traffic_GBs": (round(traffic / (ms_refactor * 1e-3) / 1e9, 1) if traffic else None)}

    cpu, parity = None, None
    if not args.no_cpu:
        try:
            cpu, parity = cpu_baseline(ckt, info, R, dt, 10, args.cpu_seconds, total_steps, x_gpu0)
            log("bench: instance 0 vs reference CPU path, max |dx| =", parity)
        except Exception as e:  # the baseline is reported, not required
            log("bench: cpu baseline unavailable:", e)

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_c5(na, W, nsteps=args.adjoint_steps)

    ring = None
    if not args.no_ring:
        ring = bsim4_ring(na, W, stages=args.ring_stages)

    out = {
        "metric": "batched Newton iters/s vs batch-1 GPU @1e-9 rel; adjoint grad/forward time",
        "value": round(value, 2),
        "unit": "Newton iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 64x64 RC mesh, per-instance R perturbation)",
        "config": {"workload": "rc_mesh_64x64_batched_refactor_be", "instances": B,
                   "unknowns": ckt.n, "nnz": ckt.nnz, "dt": dt, "integration": "backward Euler",
                   "solver": eng.solver_name},
        "batch1_gpu_value": round(batch1_value, 2),
        "speedup_vs_batch1_gpu": round(value / batch1_value, 2),
        "parity_max_abs_diff_vs_reference": parity,
        "phases": phases,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "bsim4_ring": ring,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
