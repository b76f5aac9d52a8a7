// Kernels of the bundle engine: newton_solve's per-iteration bookkeeping
// with per-instance state, and the per-instance parts of the transient
// loop (history rotation, ringing test, state rotation, probes).  One wave
// per instance unless stated; every instance's unknowns are contiguous
// (xoff[i] .. xoff[i] + n[i]).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nsa {

struct NtInst;
struct DevCtx;
struct TranDev;
struct TranStep;

// Summary of a bundle's transient for the host between batches of ticks.
struct TcStatus {
    int32_t active;    // instances not idle (taking steps in the current window)
    int32_t stalled;   // instances waiting for the host to order their matrix
    int32_t max_steps; // most accepted steps of any instance
    int32_t pad;
    double est_steps;  // max over active instances of (window end - time) / dt
};

struct NtConvTol {
    double reltol, vntol, abstol;
};

// Lane-strided states to rotate (s3 = s2, s2 = s1, s1 = s0): slot j of
// lane g at j * lanes + g, j < slots, lane g owned by instance
// lane_inst[g]; only lanes whose instance has acc == 1.
struct RotSide {
    int slots;
    int lanes;
    const int32_t* lane_inst;
    const int32_t* acc;
    const double* s0;
    double *s1, *s2, *s3;
};

// Start of an iteration (newton.cpp:138, 207-210): old solution, cleared
// device non-convergence, the LU call kind (numeric() when forced or in
// MODEINITJCT, else refactorize()), and the completion flags the loads read.
// ctx[i].iter gets the iteration's index.  With tmin (a transient's tick),
// every instance that runs an iteration has its truncation minima (tmin[i],
// tmin[ni + i]) reset to `big` for this iteration's loads; with spec,
// spec[i] = 1 where this iteration could end a converged solve (iteration
// 1 on), 0 elsewhere.  rot (optional): a state rotation the same launch
// carries in extra workgroups.
void launch_nt_begin(NtInst* nt, int ni, const int32_t* xoff, const int32_t* n, const double* x,
                     double* xold, DevCtx* ctx, int32_t* kind, int32_t* fail,
                     int32_t* devnc, int32_t* done, hipStream_t s, unsigned long long* tmin = nullptr,
                     unsigned long long big = 0, int32_t* spec = nullptr, const RotSide* rot = nullptr);
// End of an iteration (newton.cpp:211-412): LU outcome (singular
// refactorize: forced numeric and retry; singular numeric: the attempt
// fails; a numeric() whose order no longer holds stalls the instance for
// the host), convergence test, iteration-0 rule, device convergence, the
// init-mode cascade, the return of the previous iterate on convergence and
// the iteration limit.  With tc (a transient's tick): att[i] = 1 for the
// instances whose Newton solve has ended, 0 for the others.
void launch_nt_post(NtInst* nt, int ni, const int32_t* xoff, const int32_t* n,
                    const int32_t* num_nodes, double* x, const double* xold, DevCtx* ctx,
                    const int32_t* kind, const int32_t* fail, const int32_t* devnc, int32_t* done,
                    NtConvTol tol, hipStream_t s, const TranDev* tc = nullptr, int32_t* att = nullptr);

// TimeStepController::check_ringing (timestep.cpp:191-229) on the buffers
// solve_transient passes after its history rotation (transient.cpp:1039-1048):
// sol = x, sol_prev = x, sol_prev2 = h1, sol_prev3 = h2.  ring[i] = 1 when
// an instance rings; only instances with sel[i] != 0 are tested.
void launch_tr_ring(int ni, const int32_t* xoff, const int32_t* num_nodes, const double* x,
                    const double* h1, const double* h2, double reltol, double vntol,
                    const int32_t* sel, int32_t* ring, hipStream_t s);
// After the step decisions: acc[i] == 1 (accepted) rotates the solution
// history (h3 = h2, h2 = h1, h1 = x) and records the probes at row
// prow[i]; acc[i] == 2 (rejected / failed) restores x = h1 (sol_prev).
void launch_tr_hist(int ni, const int32_t* xoff, const int32_t* n, double* x, double* h1,
                    double* h2, double* h3, const int32_t* acc, const int32_t* pptr,
                    const int32_t* probe, const int32_t* prow, double* pvals, int max_points,
                    hipStream_t s);
// Circuit::rotate_state (circuit.cpp:328-334) for accepted instances:
// s3 = s2, s2 = s1, s1 = s0 per state slot.  Slots [0, ns_lane) belong to
// the lane-strided family (instance lane_inst[slot % lanes]); the others to
// slot_inst[slot - ns_lane].  Only slots [start, end) (end < 0: ns).
void launch_tr_states(long ns, long ns_lane, int lanes, const int32_t* lane_inst,
                      const int32_t* slot_inst, const int32_t* acc, double* s0, double* s1,
                      double* s2, double* s3, hipStream_t s, long start = 0, long end = -1);
// dst[k] = src[k] (src null: 0) for the slots k of the instances with
// mask[inst] != 0; slot owners as for launch_tr_states.
void launch_masked_copy(long ns, long ns_lane, int lanes, const int32_t* lane_inst,
                        const int32_t* slot_inst, const int32_t* mask, const double* src, double* dst,
                        hipStream_t s);
// DC-adjoint gradients of the linear families of a packed batch (absolute
// unknown indices, -1 = ground): resistors dL/dR = (lam_a - lam_b)(x_a -
// x_b) / R^2; voltage sources (rhs[branch] = V src_fact, vsource.cpp:193-212)
// dL/dV = lam_branch src_fact; current sources (rhs[p] -= I, rhs[n] += I,
// isource.cpp:157-174) dL/dI = (lam_n - lam_p) src_fact, src_fact from the
// owning instance's context.
void launch_bundle_grad_lin(int nr, const int32_t* r_nodes, const double* ohms, int nv,
                            const int32_t* v_branch, const int32_t* v_inst, int ni,
                            const int32_t* i_nodes, const int32_t* i_inst, const DevCtx* ctx,
                            const double* x, const double* lam, double* gR, double* gV, double* gI,
                            hipStream_t s);
// ---- the transient's tick: one Newton iteration of every instance, each
// at its own step.  A tick is
//   nt_begin, carrying the BSIM4v7 state rotation owed by the previous
//     tick's accepted steps (lane region, masked by acc), and the other
//     families' loads;
//   the BSIM4v7 load and the gathers;
//   the LU, carrying the BSIM4v7 step proposals of the iterations that
//     could end a solve (masked by spec), and nt_post with the end-of-solve
//     flags (att);
//   the other families' step proposals (masked by att), tc_step, and the
//     other families' state rotation (masked by acc).
//
// For every instance with att[i]: check_ringing, the step decision
// (tc_decide) -- acc[i] = 1 accepted (the step logged at row step_count - 1
// of log, up to max_log), 2 restore -- the solution history (accepted steps
// rotate it, h3 = h2, h2 = h1, h1 = x, and record the probes at row
// step_count; the others restore x = h1), then the loop head of the
// instance's next step (tc_attempt): its context in ctx and a fresh Newton
// state in nt, or IDLE.  acc[i] = 0 for the rest.
void launch_tc_step(TranDev* tc, const double* bpt, const int32_t* bph, NtInst* nt,
                    DevCtx* ctx, const int32_t* att, const unsigned long long* tmin, int ni,
                    const int32_t* xoff, const int32_t* n, const int32_t* num_nodes, double* x, double* h1,
                    double* h2, double* h3, double reltol, double vntol, int32_t* acc, TranStep* log,
                    int max_log, const int32_t* pptr, const int32_t* probe, double* pvals, int max_points,
                    hipStream_t s);
// A new window end for every instance; idle instances that are not
// finished pass the loop head again (as in tc_step).
void launch_tc_resume(TranDev* tc, const double* bpt, NtInst* nt, DevCtx* ctx, int ni, double t_until,
                      hipStream_t s);
// The status summary (one workgroup).
void launch_tc_status(const TranDev* tc, const NtInst* nt, int ni, TcStatus* out, hipStream_t s);

// dst[i] = src[i] for i < count.
void launch_copy_f64(double* dst, const double* src, long count, hipStream_t s);

}  // namespace nsa
