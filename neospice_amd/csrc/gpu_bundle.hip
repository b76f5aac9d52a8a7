#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "bundle.hpp"
#include "common.hpp"
#include "gpu_bundle.hpp"

namespace nsa {

namespace {

constexpr int kW = 64;

// newton_solve's locals at entry, in the caller's mode (newton.cpp:117-131)
__device__ inline NtInst fresh_newton(int mode, int itl) {
    NtInst s{};
    s.saved_mode = mode;
    s.itl = itl;
    s.kind = SLU_SKIP;
    return s;
}

__global__ __launch_bounds__(kW) void k_nt_begin(NtInst* __restrict__ nt, const int32_t* __restrict__ xoff,
                                                const int32_t* __restrict__ nn,
                                                const double* __restrict__ x,
                                                double* __restrict__ xold,
                                                DevCtx* __restrict__ ctx,
                                                int32_t* __restrict__ kind, int32_t* __restrict__ fail,
                                                int32_t* __restrict__ devnc, int32_t* __restrict__ done,
                                                unsigned long long* __restrict__ tmin, unsigned long long big,
                                                int32_t* __restrict__ spec, int ni, RotSide rs) {
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= ni) {
        // a side workgroup: the lane-strided state rotation the previous
        // tick's accepted steps owe (rotate_state, circuit.cpp:328-334), one
        // thread per lane over its slots g, g + lanes, ... (all loads issued
        // before the stores)
        const long g = (long)(i - ni) * kW + lane;
        if (g >= rs.lanes || rs.acc[rs.lane_inst[g]] != 1) return;
        constexpr int kU = 8;
        for (int j0 = 0; j0 < rs.slots; j0 += kU) {
            double v0[kU], v1[kU], v2[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const long k = (long)min(j0 + u, rs.slots - 1) * rs.lanes + g;
                v0[u] = rs.s0[k];
                v1[u] = rs.s1[k];
                v2[u] = rs.s2[k];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (j0 + u >= rs.slots) break;
                const long k = (long)(j0 + u) * rs.lanes + g;
                rs.s3[k] = v2[u];
                rs.s2[k] = v1[u];
                rs.s1[k] = v0[u];
            }
        }
        return;
    }
    NtInst s = nt[i];
    if (s.done || s.stalled) {
        if (lane == 0) {
            kind[i] = SLU_SKIP;
            fail[i] = -1;
            done[i] = 1;
            if (spec) spec[i] = 0;
        }
        return;
    }
    const int o = xoff[i], n = nn[i];
    for (int k = lane; k < n; k += kW) xold[o + k] = x[o + k];
    if (lane == 0) {
        devnc[i] = 0;
        // ngspice niiter.c:107-110: MODEINITJCT forces a reorder
        if (ctx[i].mode & MODEINITJCT) s.always_reorder = 1;
        const int k = (s.force_numeric || s.always_reorder) ? SLU_VERIFY : SLU_REFACTOR;
        s.kind = k;
        kind[i] = k;
        fail[i] = 0;
        done[i] = 0;
        nt[i] = s;
        ctx[i].iter = s.iter;
        if (tmin) {
            // the truncation minima of this iteration's loads
            tmin[i] = big;
            tmin[ni + i] = big;
        }
        // iteration 0 never ends a converged solve (newton.cpp:318-319):
        // the devices' step proposals are needed from iteration 1 on
        if (spec) spec[i] = s.iter >= 1 ? 1 : 0;
    }
}

__global__ __launch_bounds__(kW) void k_nt_post(NtInst* __restrict__ nt, const int32_t* __restrict__ xoff,
                                               const int32_t* __restrict__ nn,
                                               const int32_t* __restrict__ nnodes,
                                               double* __restrict__ x,
                                               const double* __restrict__ xold,
                                               DevCtx* __restrict__ ctx,
                                               const int32_t* __restrict__ kind,
                                               const int32_t* __restrict__ fail,
                                               const int32_t* __restrict__ devnc,
                                               int32_t* __restrict__ done, NtConvTol tol,
                                               const TranDev* __restrict__ tc, int32_t* __restrict__ att) {
    const int i = blockIdx.x, lane = threadIdx.x;
    NtInst s = nt[i];
    const int kd = kind[i];
    // the transient's tick: instances whose Newton solve has ended take their
    // step decision (att)
    auto tick_end = [&](const NtInst& sf) {
        if (!tc || lane != 0) return;
        att[i] = (tc[i].phase == TP_NEWTON && sf.done) ? 1 : 0;
    };
    if (s.done || s.stalled || kd == SLU_SKIP) {
        tick_end(s);
        return;
    }
    const int f = fail[i];
    const int iter = s.iter;
    int mode = ctx[i].mode;
    bool finish = false, restore = false;
    if (f == FAIL_NOPLAN) {
        // a fresh solver's first factorization: the host orders this instance
        s.stalled = 1;
        if (lane == 0) nt[i] = s;
        tick_end(s);
        return;
    }
    if (f != FAIL_NONE) {
        if (kd == SLU_REFACTOR) {
            // singular refactorize(): numeric() next iteration, nothing solved
            // (newton.cpp:229-234)
            s.force_numeric = 1;
        } else if (kd == SLU_VERIFY) {
            // the order no longer holds: the host re-pivots this instance
            s.stalled = 1;
            if (lane == 0) nt[i] = s;
            tick_end(s);
            return;
        } else {
            // singular numeric(): this Newton attempt fails (newton.cpp:220-224)
            s.converged = 0;
            finish = true;
        }
        s.iter = iter + 1;
    } else {
        if (kd != SLU_REFACTOR) s.force_numeric = 0;  // newton.cpp:226
        // convergence test (newton.cpp:267-292)
        const int o = xoff[i], n = nn[i], nnod = nnodes[i];
        bool bad = false;
        for (int k = lane; k < n; k += kW) {
            const double vn = x[o + k], vo = xold[o + k];
            if (isnan(vn) || isinf(vn)) {
                bad = true;
            } else {
                const double diff = fabs(vn - vo);
                const double t = tol.reltol * fmax(fabs(vn), fabs(vo)) + (k < nnod ? tol.vntol : tol.abstol);
                if (diff > t) bad = true;
            }
        }
        bool conv = !__any(bad);
        if (iter == 0) conv = false;        // newton.cpp:318-319
        if (devnc[i]) conv = false;         // device_converged (newton.cpp:321-334)
        // init-mode cascade (newton.cpp:370-411)
        const int m = mode;
        if (m & MODEINITFLOAT) {
            if (conv) { restore = true; finish = true; s.converged = 1; }
        } else if (m & MODEINITJCT) {
            mode = (m & ~INITF_MASK) | MODEINITFIX;
            s.force_numeric = 1;
            s.always_reorder = 0;
        } else if (m & MODEINITFIX) {
            if (conv) mode = (m & ~INITF_MASK) | MODEINITFLOAT;
        } else if (m & MODEINITSMSIG) {
            mode = (m & ~INITF_MASK) | MODEINITFLOAT;
        } else if (m & MODEINITTRAN) {
            mode = (m & ~INITF_MASK) | MODEINITFLOAT;
            s.force_numeric = 1;
        } else if (m & MODEINITPRED) {
            mode = (m & ~INITF_MASK) | MODEINITFLOAT;
        } else if (conv) {
            restore = true; finish = true; s.converged = 1;
        }
        s.iter = iter + 1;
        if (restore) {
            // converged: the previous iterate is returned (newton.cpp:378-381)
            for (int k = lane; k < n; k += kW) x[o + k] = xold[o + k];
        }
    }
    if (!finish && s.iter >= s.itl) {
        s.converged = 0;
        finish = true;
    }
    if (finish) {
        s.done = 1;
        mode = s.saved_mode;  // newton.cpp:224, 380, 417
    }
    if (lane == 0) {
        nt[i] = s;
        ctx[i].mode = mode;
        done[i] = s.done;
    }
    tick_end(s);
}

__global__ __launch_bounds__(kW) void k_tr_ring(const int32_t* __restrict__ xoff,
                                               const int32_t* __restrict__ nnodes,
                                               const double* __restrict__ x,
                                               const double* __restrict__ h1,
                                               const double* __restrict__ h2, double reltol,
                                               double vntol, const int32_t* __restrict__ sel,
                                               int32_t* __restrict__ ring) {
    const int i = blockIdx.x, lane = threadIdx.x;
    if (!sel[i]) return;
    const int o = xoff[i], nn = nnodes[i];
    bool found = false;
    for (int k = lane; k < nn; k += kW) {
        const double sol = x[o + k], prev = x[o + k], prev2 = h1[o + k], prev3 = h2[o + k];
        const double d2c = sol - 2.0 * prev + prev2;
        const double d2p = prev - 2.0 * prev2 + prev3;
        const double t = reltol * fabs(sol) + vntol;
        if (d2c * d2p < 0.0 && fabs(d2c) > t && fabs(d2p) > t) {
            const double d1c = sol - prev, d1p = prev - prev2;
            double d1max = fmax(fabs(d1c), fabs(d1p));
            if (d1max < 1e-30) d1max = 1e-30;
            const double ratio = fmax(fabs(d2c), fabs(d2p)) / d1max;
            if (ratio > 0.5) found = true;
        }
    }
    const bool any = __any(found);
    if (lane == 0) ring[i] = any ? 1 : 0;
}

__global__ __launch_bounds__(kW) void k_tr_hist(const int32_t* __restrict__ xoff,
                                               const int32_t* __restrict__ nn, double* __restrict__ x,
                                               double* __restrict__ h1, double* __restrict__ h2,
                                               double* __restrict__ h3,
                                               const int32_t* __restrict__ acc,
                                               const int32_t* __restrict__ pptr,
                                               const int32_t* __restrict__ probe,
                                               const int32_t* __restrict__ prow,
                                               double* __restrict__ pvals, int max_points) {
    const int i = blockIdx.x, lane = threadIdx.x;
    const int a = acc[i];
    if (!a) return;
    const int o = xoff[i], n = nn[i];
    if (a == 1 || a == 3) {
        if (a == 1)
        for (int k = lane; k < n; k += kW) {
            const double v = x[o + k];
            h3[o + k] = h2[o + k];
            h2[o + k] = h1[o + k];
            h1[o + k] = v;
        }
        if (pvals) {
            const int r = prow[i];
            if (r < max_points) {
                const int p0 = pptr[i], np = pptr[i + 1] - p0;
                for (int q = lane; q < np; q += kW)
                    pvals[(size_t)p0 * max_points + (size_t)r * np + q] = x[o + probe[p0 + q]];
            }
        }
    } else {
        for (int k = lane; k < n; k += kW) x[o + k] = h1[o + k];
    }
}

__global__ void k_tr_states(long start, long ns, long ns_lane, int lanes, const int32_t* __restrict__ lane_inst,
                            const int32_t* __restrict__ slot_inst, const int32_t* __restrict__ acc,
                            const double* __restrict__ s0, double* __restrict__ s1,
                            double* __restrict__ s2, double* __restrict__ s3) {
    const long k = start + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int inst = (k < ns_lane) ? lane_inst[k % lanes] : slot_inst[k - ns_lane];
    if (acc[inst] != 1) return;
    s3[k] = s2[k];
    s2[k] = s1[k];
    s1[k] = s0[k];
}

__global__ void k_masked_copy(long ns, long ns_lane, int lanes, const int32_t* __restrict__ lane_inst,
                              const int32_t* __restrict__ slot_inst, const int32_t* __restrict__ mask,
                              const double* __restrict__ src, double* __restrict__ dst, int zero) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int inst = (k < ns_lane) ? lane_inst[k % lanes] : slot_inst[k - ns_lane];
    if (!mask[inst]) return;
    dst[k] = zero ? 0.0 : src[k];
}

__global__ void k_copy_f64(double* __restrict__ dst, const double* __restrict__ src, long count) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) dst[k] = src[k];
}

// ---- the transient's tick (one thread per instance)

// check_ringing (as k_tr_ring), the step decision (tc_decide) and the
// history rotation / restore (as k_tr_hist) of one instance whose Newton
// solve ended in this tick.
__global__ __launch_bounds__(kW) void k_tc_step(TranDev* __restrict__ tc, const double* __restrict__ bpt,
                                               const int32_t* __restrict__ bph, NtInst* __restrict__ nt,
                                               DevCtx* __restrict__ ctx, const int32_t* __restrict__ att,
                                               const unsigned long long* __restrict__ tmin,
                                               const int32_t* __restrict__ xoff, const int32_t* __restrict__ nn,
                                               const int32_t* __restrict__ nnodes, double* __restrict__ x,
                                               double* __restrict__ h1, double* __restrict__ h2,
                                               double* __restrict__ h3, double reltol, double vntol,
                                               int32_t* __restrict__ acc, TranStep* __restrict__ log, int max_log,
                                               const int32_t* __restrict__ pptr, const int32_t* __restrict__ probe,
                                               double* __restrict__ pvals, int max_points) {
    const int i = blockIdx.x, lane = threadIdx.x, ni = gridDim.x;
    if (!att[i]) {
        if (lane == 0) acc[i] = 0;
        return;
    }
    const int o = xoff[i], n = nn[i], nnod = nnodes[i];
    // ringing on the post-solve buffers: sol = sol_prev = x, sol_prev2 = h1,
    // sol_prev3 = h2 (timestep.cpp:191-229, transient.cpp:1039-1048)
    bool found = false;
    for (int k = lane; k < nnod; k += kW) {
        const double sol = x[o + k], prev = x[o + k], prev2 = h1[o + k], prev3 = h2[o + k];
        const double d2c = sol - 2.0 * prev + prev2;
        const double d2p = prev - 2.0 * prev2 + prev3;
        const double t = reltol * fabs(sol) + vntol;
        if (d2c * d2p < 0.0 && fabs(d2c) > t && fabs(d2p) > t) {
            const double d1c = sol - prev, d1p = prev - prev2;
            double d1max = fmax(fabs(d1c), fabs(d1p));
            if (d1max < 1e-30) d1max = 1e-30;
            const double ratio = fmax(fabs(d2c), fabs(d2p)) / d1max;
            if (ratio > 0.5) found = true;
        }
    }
    const int ring = __any(found) ? 1 : 0;
    TranDev c = tc[i];
    const NtInst s = nt[i];
    const DevCtx xc = ctx[i];
    TranStep rec{};
    const int a = tc_decide(c, bpt, bph, s.converged, s.iter, __longlong_as_double((long long)tmin[i]),
                            __longlong_as_double((long long)tmin[ni + i]), ring, xc.order, xc.method, rec);
    const int step_row = c.step_count;
    // the loop head of the instance's next step (every lane takes the same
    // decisions; lane 0 writes them)
    DevCtx xn;
    const bool next = tc_attempt(c, bpt, xn);
    c.phase = next ? TP_NEWTON : TP_IDLE;
    if (lane == 0) {
        if (a == 1 && step_row - 1 < max_log) log[(size_t)i * max_log + step_row - 1] = rec;
        acc[i] = a;
        if (next) {
            ctx[i] = xn;
            nt[i] = fresh_newton(xn.mode, c.itl);
        }
        tc[i] = c;
    }
    if (a == 1) {
        for (int k = lane; k < n; k += kW) {
            const double v = x[o + k];
            h3[o + k] = h2[o + k];
            h2[o + k] = h1[o + k];
            h1[o + k] = v;
        }
        const int r = step_row;
        if (pvals && r < max_points) {
            const int p0 = pptr[i], np = pptr[i + 1] - p0;
            for (int q = lane; q < np; q += kW)
                pvals[(size_t)p0 * max_points + (size_t)r * np + q] = x[o + probe[p0 + q]];
        }
    } else {
        for (int k = lane; k < n; k += kW) x[o + k] = h1[o + k];
    }
}

__global__ void k_tc_resume(TranDev* __restrict__ tc, const double* __restrict__ bpt, NtInst* __restrict__ nt,
                            DevCtx* __restrict__ ctx, int ni, double t_until) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    TranDev c = tc[i];
    c.t_until = t_until;
    if (!c.finished && c.phase == TP_IDLE) {
        // the loop head of the instance's next step
        DevCtx xn;
        if (tc_attempt(c, bpt, xn)) {
            ctx[i] = xn;
            nt[i] = fresh_newton(xn.mode, c.itl);
            c.phase = TP_NEWTON;
        }
    }
    tc[i] = c;
}

__global__ __launch_bounds__(256) void k_tc_status(const TranDev* __restrict__ tc, const NtInst* __restrict__ nt,
                                                   int ni, TcStatus* __restrict__ out) {
    __shared__ int s_act[256], s_st[256], s_ms[256];
    __shared__ double s_est[256];
    const int t = threadIdx.x;
    int act = 0, st = 0, ms = 0;
    double est = 0.0;
    for (int i = t; i < ni; i += 256) {
        const TranDev& c = tc[i];
        ms = c.step_count > ms ? c.step_count : ms;
        if (c.phase == TP_IDLE) continue;
        ++act;
        if (c.phase == TP_NEWTON && nt[i].stalled && !nt[i].done) ++st;
        const double end = c.t_until < c.tstop ? c.t_until : c.tstop;
        const double e = c.dt > 0.0 ? (end - c.time) / c.dt : 0.0;
        est = e > est ? e : est;
    }
    s_act[t] = act;
    s_st[t] = st;
    s_est[t] = est;
    s_ms[t] = ms;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            s_act[t] += s_act[t + o];
            s_st[t] += s_st[t + o];
            s_ms[t] = s_ms[t + o] > s_ms[t] ? s_ms[t + o] : s_ms[t];
            s_est[t] = s_est[t + o] > s_est[t] ? s_est[t + o] : s_est[t];
        }
        __syncthreads();
    }
    if (t == 0) {
        out->active = s_act[0];
        out->stalled = s_st[0];
        out->max_steps = s_ms[0];
        out->pad = 0;
        out->est_steps = s_est[0];
    }
}

}  // namespace

void launch_nt_begin(NtInst* nt, int ni, const int32_t* xoff, const int32_t* n, const double* x,
                     double* xold, DevCtx* ctx, int32_t* kind, int32_t* fail,
                     int32_t* devnc, int32_t* done, hipStream_t s, unsigned long long* tmin,
                     unsigned long long big, int32_t* spec, const RotSide* rot) {
    if (ni <= 0) return;
    RotSide rs{};
    int extra = 0;
    if (rot && rot->slots > 0 && rot->lanes > 0) {
        rs = *rot;
        extra = (rs.lanes + kW - 1) / kW;
    }
    hipLaunchKernelGGL(k_nt_begin, dim3(ni + extra), dim3(kW), 0, s, nt, xoff, n, x, xold, ctx, kind, fail,
                       devnc, done, tmin, big, spec, ni, rs);
}

void launch_nt_post(NtInst* nt, int ni, const int32_t* xoff, const int32_t* n,
                    const int32_t* num_nodes, double* x, const double* xold, DevCtx* ctx,
                    const int32_t* kind, const int32_t* fail, const int32_t* devnc, int32_t* done,
                    NtConvTol tol, hipStream_t s, const TranDev* tc, int32_t* att) {
    if (ni <= 0) return;
    hipLaunchKernelGGL(k_nt_post, dim3(ni), dim3(kW), 0, s, nt, xoff, n, num_nodes, x, xold, ctx,
                       kind, fail, devnc, done, tol, tc, att);
}

void launch_tr_ring(int ni, const int32_t* xoff, const int32_t* num_nodes, const double* x,
                    const double* h1, const double* h2, double reltol, double vntol,
                    const int32_t* sel, int32_t* ring, hipStream_t s) {
    if (ni <= 0) return;
    hipLaunchKernelGGL(k_tr_ring, dim3(ni), dim3(kW), 0, s, xoff, num_nodes, x, h1, h2, reltol,
                       vntol, sel, ring);
}

void launch_tr_hist(int ni, const int32_t* xoff, const int32_t* n, double* x, double* h1,
                    double* h2, double* h3, const int32_t* acc, const int32_t* pptr,
                    const int32_t* probe, const int32_t* prow, double* pvals, int max_points,
                    hipStream_t s) {
    if (ni <= 0) return;
    hipLaunchKernelGGL(k_tr_hist, dim3(ni), dim3(kW), 0, s, xoff, n, x, h1, h2, h3, acc, pptr, probe,
                       prow, pvals, max_points);
}

void launch_tr_states(long ns, long ns_lane, int lanes, const int32_t* lane_inst,
                      const int32_t* slot_inst, const int32_t* acc, double* s0, double* s1,
                      double* s2, double* s3, hipStream_t s, long start, long end) {
    if (end < 0) end = ns;
    if (end <= start) return;
    const int tpb = 256;
    hipLaunchKernelGGL(k_tr_states, dim3((unsigned)((end - start + tpb - 1) / tpb)), dim3(tpb), 0, s, start, end,
                       ns_lane, lanes > 0 ? lanes : 1, lane_inst, slot_inst, acc, s0, s1, s2, s3);
}

void launch_masked_copy(long ns, long ns_lane, int lanes, const int32_t* lane_inst,
                        const int32_t* slot_inst, const int32_t* mask, const double* src, double* dst,
                        hipStream_t s) {
    if (ns <= 0) return;
    const int tpb = 256;
    hipLaunchKernelGGL(k_masked_copy, dim3((unsigned)((ns + tpb - 1) / tpb)), dim3(tpb), 0, s, ns, ns_lane,
                       lanes > 0 ? lanes : 1, lane_inst, slot_inst, mask, src, dst, src ? 0 : 1);
}

void launch_tc_step(TranDev* tc, const double* bpt, const int32_t* bph, NtInst* nt,
                    DevCtx* ctx, const int32_t* att, const unsigned long long* tmin, int ni,
                    const int32_t* xoff, const int32_t* n, const int32_t* num_nodes, double* x, double* h1,
                    double* h2, double* h3, double reltol, double vntol, int32_t* acc, TranStep* log,
                    int max_log, const int32_t* pptr, const int32_t* probe, double* pvals, int max_points,
                    hipStream_t s) {
    if (ni <= 0) return;
    hipLaunchKernelGGL(k_tc_step, dim3(ni), dim3(kW), 0, s, tc, bpt, bph, nt, ctx, att, tmin, xoff, n, num_nodes,
                       x, h1, h2, h3, reltol, vntol, acc, log, max_log, pptr, probe, pvals, max_points);
}

void launch_tc_resume(TranDev* tc, const double* bpt, NtInst* nt, DevCtx* ctx, int ni, double t_until,
                      hipStream_t s) {
    if (ni <= 0) return;
    hipLaunchKernelGGL(k_tc_resume, dim3((ni + kW - 1) / kW), dim3(kW), 0, s, tc, bpt, nt, ctx, ni, t_until);
}

void launch_tc_status(const TranDev* tc, const NtInst* nt, int ni, TcStatus* out, hipStream_t s) {
    hipLaunchKernelGGL(k_tc_status, dim3(1), dim3(256), 0, s, tc, nt, ni, out);
}

void launch_copy_f64(double* dst, const double* src, long count, hipStream_t s) {
    if (count <= 0) return;
    const int tpb = 256;
    hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)((count + tpb - 1) / tpb)), dim3(tpb), 0, s, dst,
                       src, count);
}


namespace {
__device__ inline double at_abs(const double* v, int i) { return i >= 0 ? v[i] : 0.0; }

__global__ void k_bundle_grad_lin(int nr, const int32_t* __restrict__ r_nodes,
                                  const double* __restrict__ ohms, int nv,
                                  const int32_t* __restrict__ v_branch, const int32_t* __restrict__ v_inst,
                                  int ni, const int32_t* __restrict__ i_nodes,
                                  const int32_t* __restrict__ i_inst, const DevCtx* __restrict__ ctx,
                                  const double* __restrict__ x, const double* __restrict__ lam,
                                  double* __restrict__ gR, double* __restrict__ gV, double* __restrict__ gI) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nr) {
        const int a = r_nodes[2 * k], b = r_nodes[2 * k + 1];
        const double R = ohms[k];
        gR[k] = (at_abs(lam, a) - at_abs(lam, b)) * (at_abs(x, a) - at_abs(x, b)) / (R * R);
        return;
    }
    const long kv = k - nr;
    if (kv < nv) {
        gV[kv] = lam[v_branch[kv]] * ctx[v_inst[kv]].src_fact;
        return;
    }
    const long ki = kv - nv;
    if (ki < ni)
        gI[ki] = (at_abs(lam, i_nodes[2 * ki + 1]) - at_abs(lam, i_nodes[2 * ki])) * ctx[i_inst[ki]].src_fact;
}
}  // namespace

void launch_bundle_grad_lin(int nr, const int32_t* r_nodes, const double* ohms, int nv,
                            const int32_t* v_branch, const int32_t* v_inst, int ni,
                            const int32_t* i_nodes, const int32_t* i_inst, const DevCtx* ctx,
                            const double* x, const double* lam, double* gR, double* gV, double* gI,
                            hipStream_t s) {
    const long total = (long)nr + nv + ni;
    if (!total) return;
    hipLaunchKernelGGL(k_bundle_grad_lin, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, nr, r_nodes,
                       ohms, nv, v_branch, v_inst, ni, i_nodes, i_inst, ctx, x, lam, gR, gV, gI);
    NSA_HIP(hipGetLastError());
}

}  // namespace nsa
