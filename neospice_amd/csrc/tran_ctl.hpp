// Per-instance adaptive step control of solve_transient (src/core/
// transient.cpp:694-1066) with its TimeStepController (src/core/timestep.cpp,
// timestep.hpp:9-85), as plain data and functions that compile for the host
// and the device.  The same code runs in two places: on the host between
// batched Newton rounds, and in the tick kernels of gpu_bundle.hip, where
// every instance walks its own transient without the host.
//
// Breakpoints: the controller's two sets (breakpoints, src_breakpoints) hold
// the same times here, because only the sources add any, all before the
// first step; they are one sorted array per instance (times and HARD
// flags) consumed from the front as time passes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gpu_dev.hpp"

namespace nsa {

// solve_transient's constants (transient.cpp:45-78).
constexpr double kInitialStopDivisor = 100.0;
constexpr double kInitialStepDivisor = 10.0;
constexpr double kMinTimeStepRatio = 1e-11;
constexpr int kMaxTransientIterations = 500000;
constexpr int kNoRecentBreakpoint = 1000;
constexpr double kNewtonFailureDtFactor = 8.0;
constexpr double kDeviceLteRejectRatio = 0.9;
constexpr double kMaxDtGrowthFactor = 2.0;
constexpr int kRingingMinStepsAfterBp = 5;
constexpr int kLteMinStepCount = 2;
constexpr double kRestartStepScale = 0.1;  // SimOptions::restart_step_scale

// One accepted transient step of an instance.
struct TranStep {
    double t, dt;
    int32_t order, method, iters, pad;
};

// Where an instance is in the loop body.
enum TranPhase : int32_t {
    TP_READY = 0,   // the next attempt is set up at the loop head
    TP_NEWTON = 1,  // an attempt's Newton iterations are running
    TP_IDLE = 2,    // finished, failed, or at the caller's window end
};

// Everything the loop body reads or writes for one instance.
struct TranDev {
    // TimeStepController
    double ctrl_dt, time, tstop, proposed_dt, prev_dt, min_break;
    // loop locals of solve_transient
    double dt, saved_delta, prev_prev_dt, dt_min, dt_max;
    double t_until;      // the caller's window end
    double xmu, src_fact;
    double soft_scale;   // sqrt(restart_step_scale), for soft breakpoints
    int64_t newton_iters;  // iterations of converged steps (TransientResult)
    int64_t all_iters;     // every iteration run, including failed and rejected steps
    int32_t rejected, order, crossed_src_bp, last_bp_hard, ringing, ringing_cooldown;
    int32_t step_count, steps_after_bp, order_promote_cooldown, total_iterations;
    int32_t integrate_method, use_gear, tried_delmin, finished, failed;
    int32_t phase;
    int32_t bp_head, bp_end;  // pending breakpoints [bp_head, bp_end) of the packed arrays
    int32_t itl;              // Newton iteration limit of a step (itl4)
    int32_t pad;
};

// std::min / std::max, argument order included.
__host__ __device__ inline double tc_min(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline double tc_max(double a, double b) { return (a < b) ? b : a; }

// TimeStepController::next_breakpoint_gap.
__host__ __device__ inline double tc_next_bp_gap(const TranDev& c, const double* bpt) {
    if (c.bp_head < c.bp_end) {
        const double bp = bpt[c.bp_head];
        if (bp > c.time + 1e-18) return bp - c.time;
    }
    return c.tstop - c.time;
}

// TimeStepController::clamp_to_breakpoint.
__host__ __device__ inline double tc_clamp_bp(const TranDev& c, const double* bpt, double dt) {
    const double t_next = c.time + dt;
    if (c.bp_head < c.bp_end) {
        const double bp = bpt[c.bp_head];
        if (t_next > bp - 1e-18) return bp - c.time;
        if (t_next > bp - 0.1 * dt) return bp - c.time;
    }
    return dt;
}

// TimeStepController::advance: time moves on, crossed breakpoints drop out.
__host__ __device__ inline void tc_advance(TranDev& c, const double* bpt, const int32_t* bph, double dt) {
    c.time += dt;
    c.crossed_src_bp = 0;
    c.last_bp_hard = 0;
    while (c.bp_head < c.bp_end && bpt[c.bp_head] <= c.time + 1e-18) {
        if (bph[c.bp_head]) c.last_bp_hard = 1;
        ++c.bp_head;
        c.crossed_src_bp = 1;
    }
}

// The loop head (transient.cpp:873-903) and fill_integrator_context
// (transient.cpp:398-447): false when the instance takes no step now
// (finished, failed, or at its window end); else the step's context in x.
__host__ __device__ inline bool tc_attempt(TranDev& c, const double* bpt, DevCtx& x) {
    if (c.finished) return false;
    if (!(c.time < c.tstop - 1e-18)) {
        c.finished = 1;
        return false;
    }
    if (c.time >= c.t_until) return false;
    if (++c.total_iterations > kMaxTransientIterations) {
        c.failed = c.finished = 1;
        return false;
    }
    // clamp to bounds and breakpoints (transient.cpp:894-903)
    double dt = tc_min(c.dt, c.dt_max);
    dt = tc_max(dt, c.dt_min);
    const double unclamped = dt;
    dt = tc_clamp_bp(c, bpt, dt);
    if (dt < unclamped - 1e-18) c.saved_delta = unclamped;
    if (c.time + dt > c.tstop) dt = c.tstop - c.time;
    if (dt < 1e-20) {
        c.finished = 1;
        return false;
    }
    c.dt = dt;
    const bool first = c.step_count == 0;
    x = DevCtx{};
    x.order = c.order;
    x.delta = dt;
    x.time = c.time + dt;
    x.delta_old1 = first ? dt : c.prev_dt;
    x.delta_old2 = c.prev_prev_dt;
    x.mode = MODETRAN | (first ? MODEINITTRAN : MODEINITPRED);
    x.method = c.integrate_method;
    x.src_fact = c.src_fact;
    if (c.order == 1) {
        x.ag0 = 1.0 / dt;
        x.ag1 = -1.0 / dt;
        x.ag2 = 0.0;
        x.xmu_ratio = 0.0;
    } else if (c.integrate_method != 1) {
        const double one_minus_xmu = 1.0 - c.xmu;
        x.ag0 = 1.0 / (dt * one_minus_xmu);
        x.ag1 = -1.0 / (dt * one_minus_xmu);
        x.ag2 = 0.0;
        x.xmu_ratio = c.xmu / one_minus_xmu;
    } else {
        const double h_old = c.prev_dt;
        if (h_old > 0.0) {
            const double sum = dt + h_old;
            x.ag0 = (2.0 * dt + h_old) / (dt * sum);
            x.ag1 = -(dt + h_old) / (dt * h_old);
            x.ag2 = dt / (h_old * sum);
        } else {
            x.ag0 = 1.0 / dt;
            x.ag1 = -1.0 / dt;
            x.ag2 = 0.0;
        }
        x.xmu_ratio = 0.0;
    }
    return true;
}

// The loop body after an attempt's Newton solve: the failure path
// (transient.cpp:917-948), evaluate_lte on the devices' proposals
// (transient.cpp:469-513: tmin_k at the step's order, tmin_2 at order 2),
// and the accepted path (transient.cpp:976-1063) with breakpoint restarts,
// order promotion and the ringing test.  Returns 1 when the step is
// accepted (rec filled), 2 when the solution must go back to the last
// accepted one.
__host__ __device__ inline int tc_decide(TranDev& c, const double* bpt, const int32_t* bph, int conv,
                                         int its, double tmin_k, double tmin_2, int ring, int order_used,
                                         int method_used, TranStep& rec) {
    c.all_iters += its;
    double dt = c.dt;
    if (!conv) {
        dt /= kNewtonFailureDtFactor;
        c.order = 1;
        c.order_promote_cooldown = 5;
        if (dt < c.dt_min) {
            if (!c.tried_delmin) {
                dt = c.dt_min;
                c.tried_delmin = 1;
            } else {
                c.failed = c.finished = 1;
            }
        }
        c.dt = dt;
        return 2;
    }
    c.newton_iters += its;
    double device_dt = 1e30;
    bool accepted = true;
    if (c.step_count >= kLteMinStepCount) {
        device_dt = tc_min(device_dt, tmin_k);
        device_dt = tc_min(kMaxDtGrowthFactor * dt, device_dt);
        if (device_dt < dt * kDeviceLteRejectRatio && dt > c.dt_min * 1.01) {
            accepted = false;
            ++c.rejected;
            c.ctrl_dt = dt;
        }
    }
    if (!accepted) {
        double proposed = c.proposed_dt;
        if (device_dt < proposed) proposed = device_dt;
        c.dt = tc_max(proposed, c.dt_min);
        return 2;
    }
    const double accepted_dt = dt;
    c.prev_prev_dt = c.prev_dt;
    c.prev_dt = dt;
    tc_advance(c, bpt, bph, dt);
    c.ctrl_dt = dt;
    ++c.step_count;
    c.tried_delmin = 0;
    if (c.steps_after_bp < kNoRecentBreakpoint) ++c.steps_after_bp;
    rec.t = c.time;
    rec.dt = accepted_dt;
    rec.order = order_used;
    rec.method = method_used;
    rec.iters = its;
    rec.pad = 0;
    if (c.crossed_src_bp && c.step_count > 2) {
        c.steps_after_bp = 0;
        c.order = 1;
        const double bp_gap = tc_next_bp_gap(c, bpt);
        double scale = kRestartStepScale;
        if (!c.last_bp_hard) scale = c.soft_scale;
        dt = tc_min(dt, scale * tc_min(c.saved_delta, bp_gap));
        dt = tc_max(dt, c.dt_min * 2.0);
    }
    if (c.order_promote_cooldown > 0) --c.order_promote_cooldown;
    if (c.order == 1 && c.step_count >= 2 && c.order_promote_cooldown == 0) {
        if (tmin_2 > 1.05 * dt) c.order = 2;
    }
    // detect_and_handle_ringing (transient.cpp:672-688)
    if (!(c.step_count < 3 || c.use_gear || c.steps_after_bp < kRingingMinStepsAfterBp)) {
        c.ringing = ring != 0;
        if (c.ringing) c.ringing_cooldown = 3;
        if (c.ringing_cooldown > 0 && !c.ringing) --c.ringing_cooldown;
        c.integrate_method = (c.ringing || c.ringing_cooldown > 0) ? 1 : 0;
    }
    if (c.step_count >= kLteMinStepCount) {
        double proposed = device_dt;
        proposed = tc_min(proposed, kMaxDtGrowthFactor * dt);
        dt = tc_max(proposed, c.dt_min);
    }
    c.dt = dt;
    return 1;
}

}  // namespace nsa
