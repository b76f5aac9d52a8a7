/* neospice_amd: MI355X (gfx950) Newton hot path for neospice, C ABI.
 *
 * Two seams of the reference are exposed, both for a batch of B circuit
 * instances that share one topology (B = 1 is the plain drop-in):
 *
 *   nsa_lu_*      the solver interface ISolver (src/core/solver_iface.hpp:16-41):
 *                 symbolic / numeric / refactorize / solve, plus the
 *                 transposed solve the adjoint needs (SparseMatrix::
 *                 solve_transposed, src/solver/solve.cpp:102-141).
 *   nsa_engine_*  one Newton-iteration pipeline over a finalized circuit:
 *                 Device::evaluate for every device (src/devices/device.hpp:85-86),
 *                 the NumericMatrix accumulation, factor, solve and the
 *                 NIiter bookkeeping of newton_solve (src/core/newton.cpp:51-419).
 *   nsa_circuit_* the typed Circuit builders and finalize() the engine consumes
 *                 (src/core/circuit.hpp:179-273).
 *
 * Conventions
 *   - Every function returns an int status: 0 = OK, 1 = singular (the
 *     `true` of ISolver::numeric/refactorize), negative = error:
 *     -1 misuse order (std::logic_error in the reference), -2 bad argument
 *     or size (std::invalid_argument), -3 runtime failure
 *     (std::runtime_error), -4 HIP error, -5 a reference feature this
 *     build does not cover.  nsa_last_error() returns the message of the
 *     calling thread's last error.  No C++ exception crosses the ABI.
 *   - Batched arrays are unknown-major with the instance index fastest:
 *     value(i, inst) = a[i * B + inst].  Matrix values follow the slot
 *     order of the reference's NumericMatrix (CSC, sorted by column then
 *     row).  Node indices are 0-based MNA indices, -1 = ground.
 *   - The caller owns every host buffer; handles own their device memory.
 *   - A handle is used by one host thread at a time (the reference keeps
 *     its analysis contexts thread_local, src/core/newton.cpp:22).
 */
#ifndef NEOSPICE_AMD_H
#define NEOSPICE_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nsa_circuit nsa_circuit;
typedef struct nsa_engine nsa_engine;
typedef struct nsa_lu nsa_lu;

/* Mirrors SimOptions (src/core/types.hpp:24-66) fields used on this path. */
typedef struct {
    double abstol, reltol, vntol, trtol, chgtol, gmin, diag_gmin, xmu, src_fact;
    int32_t itl1, itl4, method, engine; /* engine: 0 auto, 1 markowitz, 2 amdlu */
} nsa_options;

/* Mirrors the IntegratorCtx fields device loads read (types.hpp:73-94). */
typedef struct {
    int32_t mode, order, method, pad;
    double ag0, ag1, ag2, xmu_ratio, time, src_fact;
    double delta, delta_old1; /* CKTdelta and CKTdeltaOld[1] (predictor ratio) */
    double delta_old2;        /* CKTdeltaOld[2] (order-2 truncation error) */
} nsa_ctx;

const char* nsa_last_error(void);
void nsa_default_options(nsa_options* opt);

/* ---- circuit (Circuit typed API, src/api/circuit_typed.cpp) ---- */
nsa_circuit* nsa_circuit_create(void);
void nsa_circuit_destroy(nsa_circuit* c);
int nsa_circuit_node(nsa_circuit* c, const char* name, int32_t* index);
int nsa_circuit_r(nsa_circuit* c, const char* name, int32_t a, int32_t b, double ohms, int32_t* dev);
int nsa_circuit_c(nsa_circuit* c, const char* name, int32_t a, int32_t b, double farads, int32_t* dev);
int nsa_circuit_v(nsa_circuit* c, const char* name, int32_t p, int32_t n, double dc, int32_t* dev);
int nsa_circuit_v_pulse(nsa_circuit* c, const char* name, int32_t p, int32_t n,
                        const double pulse[7], int32_t* dev);
int nsa_circuit_v_pwl(nsa_circuit* c, const char* name, int32_t p, int32_t n, int32_t npts,
                      const double* t, const double* v, int32_t* dev);
int nsa_circuit_i(nsa_circuit* c, const char* name, int32_t p, int32_t n, double dc, int32_t* dev);
int nsa_circuit_model(nsa_circuit* c, const char* text);
int nsa_circuit_d(nsa_circuit* c, const char* name, int32_t a, int32_t k, const char* model, int32_t* dev);
int nsa_circuit_m(nsa_circuit* c, const char* name, int32_t d, int32_t g, int32_t s, int32_t b,
                  const char* model, double w, double l, int32_t* dev);
/* The constants the BSIM4v7 load of device `dev` (a LEVEL=14 M device) uses,
 * as the reference holds them after BSIM4v7temp in its model, size-dependent
 * (pParam) and instance records (bsim4v7_temp.cpp:98-2317): NSA_BSIM4_NPARAM
 * doubles, in the order documented in neospice_amd/csrc/bsim4.hpp (B4Param). */
#define NSA_BSIM4_NPARAM 110
int nsa_circuit_bsim4_params(const nsa_circuit* c, int32_t dev, double* out);
/* The same record (val) with its derivatives with respect to the
 * instance's W (dw) and L (dl), NSA_BSIM4_NPARAM doubles each: the
 * size-dependent path of BSIM4v7temp (bsim4v7_temp.cpp:416-419 onwards)
 * differentiated in forward mode, for the W/L adjoint. */
int nsa_circuit_bsim4_param_derivs(const nsa_circuit* c, int32_t dev, double* val, double* dw,
                                   double* dl);
/* Circuit::finalize (src/core/circuit.cpp:201-283). */
int nsa_circuit_finalize(nsa_circuit* c);
/* n (unknowns), nodes, nnz, states, is_linear */
int nsa_circuit_info(const nsa_circuit* c, int32_t out[5]);
int nsa_circuit_pattern(const nsa_circuit* c, int32_t* col_ptr, int32_t* row_idx);
int nsa_circuit_load_order(const nsa_circuit* c, int32_t* order);
int nsa_circuit_branch(const nsa_circuit* c, int32_t dev, int32_t* branch);

/* ---- batched Newton engine ---- */
int nsa_engine_create(nsa_circuit* c, int32_t batch, const nsa_options* opt, nsa_engine** out);
void nsa_engine_destroy(nsa_engine* e);
/* Per-instance primary value of device `dev` (R ohms, C farads, V/I dc), B values. */
int nsa_engine_set_values(nsa_engine* e, int32_t dev, const double* vals);
/* Per-instance value (vals[B]) of card parameter `param` of BSIM4v7 device
 * `dev` -- the .step / Monte-Carlo form of a per-instance model card (the
 * reference evaluates one card per device; instance i here is the reference
 * circuit whose card for `dev` gives `param` = vals[i]).  `param` is one of
 * the card value parameters the BSIM4v7 build covers; repeated calls
 * accumulate per instance. */
int nsa_engine_set_model_values(nsa_engine* e, int32_t dev, const char* param, const double* vals);
/* per_instance = 0: ctx[0] for all instances; 1: ctx[inst]. */
int nsa_engine_set_ctx(nsa_engine* e, const nsa_ctx* ctx, int32_t per_instance);
/* newton_solve with iteration limit itl; converged/iterations per instance. */
int nsa_engine_newton(nsa_engine* e, int32_t itl, int32_t* converged, int32_t* iterations);
/* Host copies, n x B (solution) / states x B (state slot `which` 0..3). */
int nsa_engine_get_solution(nsa_engine* e, double* x);
int nsa_engine_set_solution(nsa_engine* e, const double* x);
int nsa_engine_get_state(nsa_engine* e, int32_t which, double* s);
int nsa_engine_set_state(nsa_engine* e, int32_t which, const double* s);
/* state1..3 = state0 (transient.cpp:791-802) / rotate (circuit.cpp:328-334). */
int nsa_engine_seed_states(nsa_engine* e);
int nsa_engine_rotate_states(nsa_engine* e);
/* Device local-truncation-error step proposal for the current context and
 * states, after a step converged and before the states rotate: every
 * device's compute_trunc (capacitor.cpp:242-250, dio_device.cpp:278-291,
 * mos1_device.cpp:352-365; ckt_terr.hpp:19-64) and their minimum, as transient.cpp:478-485 and
 * convergence.cpp:948-951 take it.  dt_min: B values (1e30 when no device
 * proposes a step); dt_dev: NULL or nsa_engine_trunc_devices() x B,
 * instance fastest.  Orders above 2 are not kept (four state buffers). */
int nsa_engine_trunc(nsa_engine* e, double* dt_min, double* dt_dev);
/* Devices with a truncation row (capacitors, diodes, then MOS1), as circuit device
 * indices in dt_dev's row order; dev_index may be NULL.  Returns the count. */
int nsa_engine_trunc_devices(nsa_engine* e, int32_t* dev_index);
/* Fresh solver on the same pattern (a new make_solver + symbolic). */
int nsa_engine_reset_solver(nsa_engine* e);
/* One device load at the current solution: matrix values nnz x B, rhs n x B. */
int nsa_engine_load(nsa_engine* e, double* values, double* rhs);
/* A^T x = b on the factors of the last Newton iteration (b in place, n x B). */
int nsa_engine_solve_transposed(nsa_engine* e, double* b);
/* Name of the engine in use ("markowitz" / "amdlu"), pivot statistics. */
/* ---- parameters and adjoints ----
 * No counterpart in the reference, whose only sensitivity analysis is the
 * forward-difference .sens (src/core/sens.cpp:57-186: one extra DC solve per
 * R, V and I); its batched / Monte Carlo and adjoint execution is listed as
 * planned (ROADMAP.md:128-141).  Family codes match the device kinds. */
enum { NSA_FAMILY_R = 0, NSA_FAMILY_C = 1, NSA_FAMILY_V = 2, NSA_FAMILY_I = 3 };
int nsa_engine_family_count(nsa_engine* e, int32_t family, int32_t* count);
/* All per-instance values of a family at once: count x B, instance fastest
 * (R ohms, C farads, V / I dc value). */
int nsa_engine_set_family_values(nsa_engine* e, int32_t family, const double* vals);
int nsa_engine_get_family_values(nsa_engine* e, int32_t family, double* vals);
/* DC adjoint on the factors of the last Newton solve: J^T lambda = dL/dx
 * (n x B), then dL/dR, dL/dC (zero), dL/dV, dL/dI (count x B each; any
 * output may be null). */
int nsa_engine_dc_adjoint(nsa_engine* e, const double* dLdx, double* lambda, double* gR,
                          double* gC, double* gV, double* gI);
/* Fixed-step transient from the current solution: TRANOP operating point,
 * then nsteps steps of dt (backward Euler first, then the options' method);
 * probe node voltages out ((nsteps+1) x nprobe x B); a checkpoint every
 * ckpt_every steps for nsa_engine_tran_adjoint. */
int nsa_engine_tran_forward(nsa_engine* e, int32_t nsteps, double dt, const int32_t* probes,
                            int32_t nprobe, int32_t ckpt_every, double* probe_out);
/* Discrete adjoint of the last tran_forward: dL/d(probe values) in
 * ((nsteps+1) x nprobe x B); dL/dR and dL/dC out (count x B). */
int nsa_engine_tran_adjoint(nsa_engine* e, const double* dLdprobe, double* gR, double* gC);
/* Wall milliseconds of the last tran_forward and tran_adjoint. */
int nsa_engine_adjoint_timing(nsa_engine* e, double out[2]);
const char* nsa_engine_solver_name(nsa_engine* e);
/* out: cpu pivots, gpu refactors, nnz(L+U), refactor levels, solve levels
 * (of the first pivot group), gpu solves, device loads, pivot groups. */
int nsa_engine_stats(nsa_engine* e, int64_t out[8]);
/* HIP-event timing on the engine's stream since timing was enabled:
 * accumulated milliseconds of device load+gather, GPU refactor, GPU solve and
 * the BSIM4v7 load kernel alone, then the number of loads timed. */
int nsa_engine_set_timing(nsa_engine* e, int32_t on);
int nsa_engine_timing(nsa_engine* e, double out[5]);

/* ---- bundles: instances of different topologies in one batch ----
 * The reference runs a Monte Carlo / .step sweep one circuit at a time, each
 * a fresh Circuit with its own solver (run_step_sweep,
 * src/api/neospice.cpp:229-334; Simulator::run_dc / run_transient,
 * src/api/neospice.hpp:50-93).  A bundle holds one finalized circuit per
 * instance (instances may differ in topology and size) and advances all of
 * them together: solutions are packed instance after instance
 * (x[xoff[i] + k], xoff from nsa_bundle_offsets).  Circuits must outlive the
 * bundle.  Every instance is factored by Sparse 1.3, as make_solver chooses
 * for nonlinear circuits and for linear ones under 256 unknowns
 * (src/core/make_solver.cpp:156-176); other circuits are refused (-5). */
typedef struct nsa_bundle nsa_bundle;
int nsa_bundle_create(nsa_circuit* const* circuits, int32_t count, const nsa_options* opt,
                      nsa_bundle** out);
void nsa_bundle_destroy(nsa_bundle* b);
/* instances, total unknowns, total matrix slots, total state slots */
int nsa_bundle_info(nsa_bundle* b, int32_t out[4]);
/* xoff, aoff: count+1 entries each (either may be NULL) */
int nsa_bundle_offsets(nsa_bundle* b, int32_t* xoff, int32_t* aoff);
/* newton_solve's direct DC attempt (src/core/dc.cpp:104-107) on every
 * instance in `mode` (MODEDCOP or MODETRANOP with MODEINITJCT) from the
 * current solution; converged / iterations per instance. */
int nsa_bundle_dc(nsa_bundle* b, int32_t mode, int32_t* converged, int32_t* iterations);
/* The DC operating-point ladder per instance, from the current solution:
 * the transient preamble's (compute_dc_operating_point,
 * src/core/transient.cpp:172-238: direct Newton in MODETRANOP, then gmin
 * stepping and source stepping, src/core/convergence.cpp:216-346, 472-582)
 * or, with flags bit 1, solve_dc's (src/core/dc.cpp:104-184: MODEDCOP,
 * with true gmin stepping, convergence.cpp:348-470, before gmin stepping).
 * Every instance follows its own schedule.  stage per instance: 0 direct,
 * 1 gmin stepping, 2 source stepping, 4 true gmin stepping, -1 not
 * converged (pseudo-transient continuation and the later fallbacks are not
 * restated).  flags bit 0: one method alone on every instance (no direct
 * attempt), bits 4-5 naming it: 0 gmin, 1 source, 2 true gmin stepping. */
int nsa_bundle_dc_ladder(nsa_bundle* b, int32_t flags, int32_t* converged, int32_t* iterations,
                         int32_t* stage);
int nsa_bundle_get_solution(nsa_bundle* b, double* x);
int nsa_bundle_set_solution(nsa_bundle* b, const double* x);
/* solve_transient (src/core/transient.cpp:694-1066) for every instance:
 * tran_begin runs the MODETRANOP operating point from zeros, the state
 * seeding, source defaults, breakpoints and the first step (-3 if an
 * instance's operating point does not converge); tran_advance runs batched
 * adaptive steps until every instance has reached t_until (its own step
 * controller: LTE accept/reject, order promotion, breakpoints, ringing and
 * Newton-failure step cuts).  Returns 1 when every instance is finished.
 * probes: per-instance unknown indices (probe_ptr[count+1]) recorded at t=0
 * and after every accepted step, up to max_points rows per instance. */
int nsa_bundle_tran_begin(nsa_bundle* b, double tstep, double tstop, const int32_t* probe_ptr,
                          const int32_t* probes, int32_t max_points);
int nsa_bundle_tran_advance(nsa_bundle* b, double t_until, int32_t* rounds);
/* accepted steps, rejected steps, failed, finished, Newton iterations of
 * converged steps (TransientResult::status.iterations), all iterations run,
 * loop passes, DC iterations */
int nsa_bundle_tran_status(nsa_bundle* b, int32_t inst, int64_t out[8]);
/* accepted steps of an instance: end time, step, order, integration method,
 * Newton iterations (any output may be NULL); returns the count */
int nsa_bundle_tran_steps(nsa_bundle* b, int32_t inst, double* t, double* dt, int32_t* order,
                          int32_t* method, int32_t* iters);
/* recorded probe rows of an instance (rows x nprobe), rows out */
int nsa_bundle_tran_probes(nsa_bundle* b, int32_t inst, int32_t* rows, double* out);
/* iterations, CPU pivot computations, plans, graph launches, host-assisted
 * iterations, max refactor levels, max solve levels, max nnz(L+U) */
int nsa_bundle_stats(nsa_bundle* b, int64_t out[8]);
/* HIP-event timing of device-driven iterations (loads+gathers, LU,
 * convergence/control, the BSIM4v7 load alone; count) */
int nsa_bundle_set_timing(nsa_bundle* b, int32_t on);
int nsa_bundle_timing(nsa_bundle* b, double out[5]);
/* 1 (default): device-driven iterations replay a captured HIP graph */
int nsa_bundle_set_graph(nsa_bundle* b, int32_t on);
/* 1 (default): the transient's step controller (solve_transient's loop body
 * with its TimeStepController, src/core/transient.cpp:873-1063) runs on the
 * device, every instance on its own step sequence; 0: on the host between
 * batched Newton rounds.  Both give the same results per instance. */
int nsa_bundle_set_device_control(nsa_bundle* b, int32_t on);
/* 1 (default): at a transient's start the packed LU's schedule (the
 * shortest run of one-row levels solved as a wave recurrence, LDS staging
 * of the plan) is chosen by timing every candidate on the batch's own
 * matrices; every schedule gives the same results.  nsa_bundle_tuning
 * writes up to max_out values -- the chosen candidate, the candidate count,
 * then per candidate (min serial run, staging, microseconds per launch,
 * as timed when the decision was made: a process reuses a decision for the
 * same plans and batch size) -- and returns how many there are
 * (no reference counterpart: the autotuner of this engine). */
int nsa_bundle_set_tuning(nsa_bundle* b, int32_t on);
int nsa_bundle_tuning(nsa_bundle* b, double* out, int32_t max_out);
/* device-controlled transient: ticks run, host looks between batches of
 * ticks, host seconds spent enqueueing ticks and waiting for the device
 * (no reference counterpart: statistics of this engine) */
int nsa_bundle_tick_stats(nsa_bundle* b, double out[4]);
/* pivot structure of an instance: plan index (-1 before its first
   factorization), n, nnz(A), nnz(L+U), levels of a factor-and-solve launch
   (the refactor with the forward substitution folded in, then the
   backward substitution), and
   the operation counts of one refactor (SparseMatrix::factor's left-looking
   loop, src/solver/factor.cpp:1010-1026) and of one forward + backward
   solve (src/solver/solve.cpp:6-47), and the int32 plan words such a
   launch reads */
int nsa_bundle_plan_info(nsa_bundle* b, int32_t inst, int64_t out[9]);
/* Device offsets (count + 1): instance i's device d is entry doff[i] + d of
 * the per-device outputs below. */
int nsa_bundle_device_offsets(nsa_bundle* b, int32_t* doff);
/* DC adjoint at every instance's last operating point (the quantities of
 * .sens, src/core/sens.cpp:105-186, by the adjoint method): lambda (total
 * unknowns, packed) solves J^T lambda = dL/dx with the factors of the last
 * Newton iteration; g_value = dL/d(value) per device for R (ohms), C (zero
 * in DC), V and I; g_w, g_l = dL/dW and dL/dL of BSIM4v7 devices through the
 * size-dependent parameters (bsim4v7_temp.cpp:416-419 onwards).  Outputs may
 * be NULL. */
int nsa_bundle_dc_adjoint(nsa_bundle* b, const double* dLdx, double* lambda, double* g_value, double* g_w,
                          double* g_l);
/* New per-device values for later solves, in the layout of
 * nsa_bundle_dc_adjoint (NULL arrays are left alone): value sets R, C, V and
 * I (Resistor::set_resistance and its siblings, as .sens perturbs them,
 * src/core/sens.cpp:74-130); w and l set BSIM4v7 W and L (their constants
 * are recomputed: BSIM4v7temp, bsim4v7_temp.cpp:416-419 onwards). */
int nsa_bundle_set_device_values(nsa_bundle* b, const double* value, const double* w, const double* l);
/* The current values in the same layout (zero where a kind does not apply). */
int nsa_bundle_get_device_values(nsa_bundle* b, double* value, double* w, double* l);

/* ---- solver seam (ISolver) ---- */
int nsa_lu_create(int32_t batch, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                  int32_t is_linear, int32_t engine, nsa_lu** out);
void nsa_lu_destroy(nsa_lu* s);
/* values: host, nnz x B.  status: per instance 0 ok / 1 singular.  Returns
 * 1 if any instance is singular. */
int nsa_lu_numeric(nsa_lu* s, const double* values, double diag_gmin, int32_t* status);
int nsa_lu_refactorize(nsa_lu* s, const double* values, double diag_gmin, int32_t* status);
/* ISolver::set_equilibrate (src/core/solver_iface.hpp:33-36): radix-2
 * row/column equilibration of the following numeric/refactorize and the
 * solves on their factors (NeoSolver, neo_solver.cpp:52-102, 204-225).
 * Honoured by Markowitz solvers only, as in the reference (AmdLuSolver and
 * the AMD-LU auto-fallback keep the interface's no-op). */
int nsa_lu_set_equilibrate(nsa_lu* s, int32_t on);
/* The complex (AC) paths of ISolver (src/core/solver_iface.hpp:23-28), the
 * reference's NeoSolver ones whatever the engine (AmdLuSolver delegates
 * them, amd_lu_solver.cpp:347-349).  Host arrays in the real calls' layout,
 * instance fastest, a complex value's real and imaginary parts in rows 2k
 * and 2k + 1: ax (2 nnz) x B, rhs (2 n) x B.
 * numeric_complex (neo_solver.cpp:228-245): every instance ordered by the
 * complex Markowitz search on its own values (reusing its order on a later
 * call when the order still holds, factor.cpp:929-947) and factored;
 * refactorize_complex (neo_solver.cpp:247-258) factors on the orders held
 * (factor_complex, factor.cpp:1054-1141); solve_complex (solve.cpp:49-100)
 * solves in place.  status per instance: 0 ok, 1 singular; the factor calls
 * return 1 if any instance is singular. */
int nsa_lu_numeric_complex(nsa_lu* s, const double* ax, int32_t* status);
int nsa_lu_refactorize_complex(nsa_lu* s, const double* ax, int32_t* status);
int nsa_lu_solve_complex(nsa_lu* s, double* rhs);
/* The AC analysis's frequency sweep (src/core/ac.cpp:196-238) for every
 * instance and frequency in one launch: instance i's matrix at frequency f
 * is g_i + j (2 pi f) c_i, ordered at freqs[0] (numeric_complex) and
 * factored on that order at every frequency (refactorize_complex), then
 * solved for rhs_i.  Host arrays, instance-major (the layout of a batch of
 * the reference's own arrays, and a complex array's memory): g, c: B blocks
 * of nnz (CSC slot order); rhs: B blocks of 2 n ((re, im) per unknown);
 * out: B nf blocks of 2 n, block i nf + f; status: B nf entries, same
 * order.  Per-frequency device stamps (ac_stamp_freq) are the caller's: g
 * and c hold the frequency-independent parts.  After a sweep the next
 * solve needs a refactorize_complex. */
int nsa_lu_ac_sweep(nsa_lu* s, const double* g, const double* c, const double* freqs, int32_t nf,
                    const double* rhs, double* out, int32_t* status);
/* Seconds of the last nsa_lu_ac_sweep: out[0] the host orderings, out[1]
 * the device work with inputs resident (matrix build, factor and solve, by
 * HIP events), out[2] the whole call inside the library (no reference
 * counterpart: statistics of this engine). */
int nsa_lu_ac_stats(nsa_lu* s, double out[3]);
/* rhs: host, n x B, overwritten with the solution. */
int nsa_lu_solve(nsa_lu* s, double* rhs);
int nsa_lu_solve_transposed(nsa_lu* s, double* rhs);
/* nsa_lu_solve without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops).  The staging vector lives in a 16 MiB device array the
 * library holds on each GPU until the process ends, so n x batch may not
 * exceed 2097152 values (larger: NSA_ERR_ARGUMENT).  Calls on one GPU share
 * that array: calls that overlap on one GPU race. */
int nsa_lu_solve_noalloc(nsa_lu* s, double* rhs);
const char* nsa_lu_name(nsa_lu* s);
/* Pivot order and fill: n, nnz(L+U), levels of refactor, levels of solve. */
int nsa_lu_info(nsa_lu* s, int32_t out[4]);
/* Pivot-space structure: row_perm (n), col_perm (n), seg_ptr (n+1), rows (nnz(L+U)). */
int nsa_lu_structure(nsa_lu* s, int32_t* row_perm, int32_t* col_perm, int32_t* seg_ptr,
                     int32_t* rows);

/* ---- host-only pivot computations (no GPU needed) ----
 * The pivot order and fill pattern a fresh numeric factorization of
 * `values` fixes: engine 1 = Sparse 1.3 Markowitz (NeoSolver::numeric,
 * src/core/neo_solver.cpp:148-174), engine 2 = min-degree + Gilbert-Peierls
 * (AmdLuSolver::symbolic + factor_, src/core/amd_lu_solver.cpp:11-197),
 * engine 3 = Sparse 1.3 Markowitz on a complex matrix (NeoSolver::
 * numeric_complex, neo_solver.cpp:228-245; values 2 nnz, (re, im) per
 * slot, diag_gmin 0).
 * Pivot space k = 0..n-1: b'[k] = b[row_perm[k]], x[col_perm[k]] = y[k];
 * column k holds nU[k] U rows (processing order), the diagonal, then L rows,
 * at rows[seg_ptr[k] .. seg_ptr[k+1]).  Pass rows = NULL to get *nnz_lu only.
 * Returns 1 when the matrix is singular. */
int nsa_pivot_structure(int32_t engine, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                        const double* values, double diag_gmin, int32_t* nnz_lu,
                        int32_t* row_perm, int32_t* col_perm, int32_t* seg_ptr, int32_t* nU,
                        int32_t* rows);
/* amd_ordering (src/core/amd.cpp:14-143). */
int nsa_min_degree_order(int32_t n, const int32_t* col_ptr, const int32_t* row_idx, int32_t* perm);

#ifdef __cplusplus
}
#endif
#endif
