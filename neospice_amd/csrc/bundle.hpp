// A bundle: many independent circuit instances, each with its own topology
// and values (a Monte Carlo / .step sweep over circuits of uneven size),
// advanced together on one GPU.
//
// The reference runs such a sweep one circuit at a time (run_step_sweep,
// src/api/neospice.cpp:229-334): every instance is its own Circuit with its
// own solver.  Here every per-instance array is packed instance after
// instance (unknowns at xoff[i], matrix slots at aoff[i], states in the
// families' regions), and one launch per phase covers every instance:
//   * device loads: one lane per device of every instance (BSIM4v7 with its
//     slots and constants strided by lane, so a wave's loads coalesce);
//   * gathers: one lane per matrix slot / rhs row of the packed system;
//   * LU: one wave per instance on its own plan (small_lu), the factor in LDS;
//   * Newton control: newton_solve's per-iteration bookkeeping
//     (src/core/newton.cpp:136-412) in a kernel with per-instance state, so
//     an iteration is a fixed sequence of launches captured in a HIP graph.
//   * transient step control: solve_transient's loop body per instance
//     (tran_ctl.hpp) in kernels, so that in a "tick" every instance runs one
//     Newton iteration of its own current step, and an instance whose
//     Newton solve ends takes its step decision and starts its next step in
//     the same tick, without waiting for the others or for the host.
// The host keeps what the reference decides from values on the CPU: the
// Markowitz pivot orders (one Sparse 1.3 state per instance).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <set>
#include <vector>

#include "bsim4.hpp"
#include "circuit.hpp"
#include "engine.hpp"
#include "gpu_bundle.hpp"
#include "gpu_dev.hpp"
#include "lu_cpu.hpp"
#include "small_lu.hpp"
#include "tran_ctl.hpp"

namespace nsa {

// The method of the DC ladder that converged for an instance.
enum LadderStage : int32_t {
    LADDER_NONE = -1,
    LADDER_DIRECT = 0,
    LADDER_GMIN = 1,       // gmin stepping (convergence.cpp:216-346)
    LADDER_SOURCE = 2,     // source stepping (convergence.cpp:472-582)
    LADDER_TRUE_GMIN = 4,  // true gmin stepping (convergence.cpp:348-470)
};

// Per-instance Newton state on the device (newton_solve's locals).
struct NtInst {
    int32_t saved_mode;      // the caller's mode, restored on exit (newton.cpp:117, 417)
    int32_t done;            // finished (converged, failed) or not taking part
    int32_t converged;
    int32_t iter;            // iterations so far
    int32_t itl;             // iteration limit
    int32_t force_numeric;   // newton.cpp:125
    int32_t always_reorder;  // newton.cpp:131
    int32_t kind;            // LU call of the current iteration (SluKind)
    int32_t stalled;         // a numeric() whose pivot order failed: the host re-pivots
    int32_t pad[3];
};

class Bundle {
public:
    // One circuit per instance (the same Circuit may appear more than once:
    // instances are copies).  Every circuit must be finalized.
    Bundle(const std::vector<const Circuit*>& circuits, const Options& opt);
    ~Bundle();
    Bundle(const Bundle&) = delete;
    Bundle& operator=(const Bundle&) = delete;

    int size() const { return NI_; }
    int total_n() const { return tot_n_; }
    int total_nnz() const { return tot_nnz_; }
    const std::vector<int32_t>& xoff() const { return xoff_; }
    const std::vector<int32_t>& aoff() const { return aoff_; }
    hipStream_t stream() const { return stream_; }

    // Integrator context per instance (host copy kept).
    void set_ctx(const std::vector<DevCtx>& c);
    const std::vector<DevCtx>& ctx() const { return h_ctx_; }

    // newton_solve on every instance with active[i] != 0 (null: all), from
    // the current solution with the mode in its context.
    void newton(int itl, const int32_t* active, std::vector<int32_t>& converged,
                std::vector<int32_t>& iterations);
    // Direct Newton attempt of the DC operating point (dc.cpp:104-107 /
    // transient.cpp:195-205) for every instance, from the current solution.
    void dc_op(int mode, std::vector<int32_t>& converged, std::vector<int32_t>& iterations);
    // The operating-point ladder, per instance: direct Newton in firstmode,
    // then for the instances that failed the continuation methods of the
    // reference's ladder in its order -- solve_dc's (MODEDCOP, dc.cpp:
    // 104-184): true gmin stepping, gmin stepping, source stepping; the
    // transient preamble's (MODETRANOP, compute_dc_operating_point,
    // transient.cpp:172-238): gmin stepping, source stepping -- each
    // instance on its own schedule (diag_gmin, device gmin and src_fact per
    // instance).  stage[i] is the method that converged (LadderStage).
    // Pseudo-transient continuation, OPtran, the node-classified guess and
    // the equilibrated retry that follow in the reference are not
    // restated.  alone (a LadderStage other than LADDER_NONE): that method
    // alone for every instance, without the direct attempt.
    void op_ladder(int firstmode, int continuemode, std::vector<int32_t>& converged,
                   std::vector<int32_t>& iterations, std::vector<int32_t>& stage,
                   int alone = LADDER_NONE);
    const std::vector<int32_t>& dc_stage() const { return dc_stage_; }
    // A fresh solver for every instance (make_solver + symbolic).
    void reset_solver();

    void get_solution(double* out) const;
    void set_solution(const double* in);
    void get_state(int which, double* out) const;  // packed states (family regions)
    int num_states() const { return ns_tot_; }

    // ---- adaptive transient (solve_transient, transient.cpp:694-1066)
    // DC operating point (MODETRANOP), fresh solver, state seeding, source
    // defaults and breakpoints, initial step.  probes: per-instance list of
    // unknown indices recorded at t = 0 and every accepted step
    // (probe_ptr[NI+1] into probes; may be null), up to max_points rows.
    void tran_begin(double tstep, double tstop, const int32_t* probe_ptr, const int32_t* probes,
                    int max_points);
    // Runs steps until every instance has reached t_until (or tstop, or
    // failed).  Returns the number of ticks (device control) or batched
    // rounds (host control).
    int tran_advance(double t_until);
    bool tran_done() const;
    // controller state of instance i (current after tran_advance returns)
    const TranDev& tran_state(int i) const { return tcd_[i]; }
    // accepted steps of instance i
    std::vector<TranStep> tran_steps(int i) const;
    // true (default): the step controller runs on the device in ticks; false:
    // on the host between batched Newton rounds (the same decisions; every
    // instance's results are bit-identical either way).
    bool device_control = true;
    // The packed LU's schedule chosen by timing (tune_lu): at a transient's
    // start when on.  tune_report: chosen candidate, then per candidate
    // (min_serial_run, stage, microseconds per launch).
    bool tuning = true;
    std::vector<double> tune_report;
    void tune_lu();
    // probe rows of instance i: (count, values row-major count x nprobe_i)
    void tran_probes(int i, int* count, double* out) const;
    int tran_probe_count(int i) const { return pptr_.empty() ? 0 : pptr_[i + 1] - pptr_[i]; }

    // ---- DC adjoint (sens.cpp's quantities, by the adjoint instead of
    // finite differences): at every instance's last converged operating
    // point, lambda = J^-T dL/dx through the stored factors of the last
    // iteration, then dL/dp = -lambda^T dF/dp.  Outputs are per circuit
    // device (instance i's device d at device_offset(i) + d): g_value is
    // dL/d(value) of R (ohms), C (farads: zero in DC), V and I sources;
    // g_w / g_l are dL/dW and dL/dL of BSIM4v7 devices (zero elsewhere).
    // Any output may be null.
    void dc_adjoint(const double* dLdx, double* lambda, double* g_value, double* g_w, double* g_l);
    // New device values for later solves, per circuit device as for
    // dc_adjoint (null arrays leave theirs): value sets R (ohms), C (farads),
    // V (dc) and I (dc); w and l set the W and L of BSIM4v7 devices (their
    // constant records are recomputed).  Entries of other devices are
    // ignored.
    void set_device_values(const double* value, const double* w, const double* l);
    void get_device_values(double* value, double* w, double* l);
    int device_offset(int i) const { return doff_[i]; }
    int total_devices() const { return doff_[NI_]; }

    // ---- statistics
    long n_cpu_pivot = 0, n_iterations = 0, n_graph_launches = 0, n_host_iterations = 0;
    long n_ticks = 0;  // transient ticks run under device control
    long n_tick_batches = 0;     // host looks between batches of ticks
    double tick_enqueue_s = 0.0;  // host time spent enqueueing ticks
    double tick_wait_s = 0.0;     // host time spent waiting for the device between batches
    int num_plans() const { return plans_.count(); }
    const SmallLuPlans& plans() const { return plans_; }
    int plan_of(int i) const { return plan_[i]; }
    // HIP-event timing of the phases of device-driven iterations
    void set_timing(bool on);
    double ms_load = 0.0, ms_lu = 0.0, ms_post = 0.0, ms_bsim4 = 0.0;
    long n_timed = 0;
    bool use_graph = true;

private:
    void build_tables();
    void launch_begin();
    void launch_loads();
    void launch_loads_small();
    void launch_loads_b4_gather();
    void launch_lu(const int32_t* d_kind, int mode);
    void launch_post(bool tick = false);
    void iteration_device();
    void capture_graph();
    void host_lu_post(std::vector<NtInst>& nt, const std::vector<int>& repivot, bool others);
    void collect_timing();
    void set_gmin(const std::vector<double>& g);
    void set_dev_gmin(const std::vector<double>& g);
    void ladder_gmin(int firstmode, int continuemode, bool device, std::vector<int32_t>& converged,
                     std::vector<int32_t>& iterations, std::vector<int32_t>& stage);
    void ladder_source(int base_mode, std::vector<int32_t>& converged, std::vector<int32_t>& iterations,
                       std::vector<int32_t>& stage);
    void masked_copy_state(const std::vector<int32_t>& mask, double* const* src, double* const* dst);
    std::vector<double> instance_matrix(int i);
    void upload_inst();
    void ensure_adjoint();
    void refresh_size_derivs();
    int tran_advance_host(double t_until);
    int tran_advance_device(double t_until);
    void launch_tick(bool timed);
    void flush_rotations();
    void tick_device();
    void capture_tick_graph();
    bool graph_stale(const int32_t* words, int nlu, int n, int hot) const;
    void tran_log_reserve(int rows);
    void handle_stalls();

    std::vector<const Circuit*> ckt_;
    int NI_ = 0, tot_n_ = 0, tot_nnz_ = 0, ns_tot_ = 0;
    Options opt_;
    hipStream_t stream_ = nullptr;
    std::vector<void*> allocs_;
    std::vector<int32_t> xoff_, aoff_, n_, nnodes_;
    std::vector<int32_t> doff_;  // NI+1 device offsets (circuit device order)
    // adjoint: family tables with the owners of their entries (global device positions)
    bool adj_ready_ = false, size_derivs_dirty_ = true;
    std::vector<int32_t> r_owner_, v_owner_, i_owner_, c_owner_, b4_owner_;
    std::vector<double> b4_w_, b4_l_;  // per lane: the W and L its record was computed with
    int32_t *d_r_nodes_ = nullptr, *d_v_branch_ = nullptr, *d_v_inst_ = nullptr, *d_i_nodes_ = nullptr,
            *d_i_inst_ = nullptr;
    double *d_lam_ = nullptr, *d_gR_ = nullptr, *d_gV_ = nullptr, *d_gI_ = nullptr, *d_gW_ = nullptr,
           *d_gL_ = nullptr, *d_b4_dW_ = nullptr, *d_b4_dL_ = nullptr;
    // device vectors
    double *d_x_ = nullptr, *d_xold_ = nullptr, *d_rhs_ = nullptr, *d_A_ = nullptr, *d_LU_ = nullptr;
    double* d_s_[4] = {nullptr, nullptr, nullptr, nullptr};
    double *d_cm_ = nullptr, *d_cr_ = nullptr;
    DevCtx* d_ctx_ = nullptr;
    std::vector<DevCtx> h_ctx_;
    DevCtx* h_ctx_pin_ = nullptr;
    NtInst* d_nt_ = nullptr;
    NtInst* h_nt_pin_ = nullptr;
    int32_t *d_done_ = nullptr, *d_devnc_ = nullptr, *d_kind_ = nullptr, *d_fail_ = nullptr,
            *d_fail_col_ = nullptr, *d_kind_host_ = nullptr;
    int32_t* h_kind_pin_ = nullptr;
    int32_t *d_ixoff_ = nullptr, *d_in_ = nullptr, *d_inn_ = nullptr;  // per instance xoff, n, num_nodes
    // families
    CapTable cap_{};
    VsrcTable vsrc_{};
    IsrcTable isrc_{};
    DioTable dio_{};
    Mos1Table mos_{};
    B4Table b4_{};
    TruncTable trunc1_{}, trunc4_{};
    int nR_ = 0, r_site_ = 0, n_const_ = 0, const_site_ = 0, n_msites_ = 0, n_rsites_ = 0;
    double *d_r_ohms_ = nullptr, *d_const_vals_ = nullptr;
    int32_t *d_mptr_ = nullptr, *d_msites_ = nullptr, *d_rptr_d_ = nullptr, *d_rsites_d_ = nullptr,
            *d_rptr_u_ = nullptr, *d_rsites_u_ = nullptr;
    DevTol tol_{};
    GatherPacked gm_{}, gr_{};  // row lists of the matrix and rhs gathers
    std::vector<double> h_pulse_;  // 7 per V source, for source-default resolution
    // LU
    SmallLuPlans plans_;
    std::vector<int32_t> plan_;                       // per instance, -1 = no factorization yet
    std::vector<std::unique_ptr<Sparse13>> s13_;      // per instance CPU pivot state
    std::vector<SluInst> h_inst_;
    SluInst* d_inst_ = nullptr;
    bool inst_dirty_ = true;
    // graph of one device-driven iteration
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    // graph of one transient tick, and the launch shape it was captured with
    hipGraph_t tick_graph_ = nullptr;
    hipGraphExec_t tick_exec_ = nullptr;
    const int32_t* tick_words_ = nullptr;
    int tick_nlu_ = 0, tick_n_ = 0, tick_hot_ = 0;
    TranStep* tick_log_ = nullptr;
    bool timing_ = false;
    hipEvent_t ev_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool timed_pending_ = false;
    const int32_t* graph_words_ = nullptr;
    int graph_nlu_ = 0, graph_n_ = 0, graph_hot_ = 0;
    int b4_lanes_ = 0;
    long ns_lane_ = 0;
    int32_t* d_slot_inst_ = nullptr;
    int32_t* d_xinst_ = nullptr;   // owner instance of every unknown
    double* d_gmin_ = nullptr;     // diag_gmin per instance
    double* d_devgmin_ = nullptr;  // device gmin per instance (true gmin stepping)
    double* h_gmin_pin_ = nullptr;
    int32_t* d_mask_ = nullptr;
    int32_t* h_mask_pin_ = nullptr;
    double* d_lad_[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<int32_t> dc_stage_;
    // transient: controllers (host mirror and device copy), breakpoints,
    // step logs
    std::vector<TranDev> tcd_;
    TranDev* d_tc_ = nullptr;
    int32_t* d_spec_ = nullptr;                 // per instance: this iteration could end a converged solve

    bool tc_on_device_ = false;                 // d_tc_ holds the current controllers
    std::vector<double> bpt_;                   // breakpoint times, packed per instance
    std::vector<int32_t> bph_;                  // their HARD flags
    double* d_bpt_ = nullptr;
    int32_t* d_bph_ = nullptr;
    std::vector<std::vector<TranStep>> hsteps_;  // host-control step logs
    TranStep* d_log_ = nullptr;                 // device-control step logs, log_rows_ per instance
    int log_rows_ = 0;
    bool log_on_device_ = false;
    TcStatus* d_tcs_ = nullptr;
    TcStatus* h_tcs_pin_ = nullptr;
    double tstep_ = 0.0, tstop_ = 0.0;
    unsigned long long* d_tmin_ = nullptr;   // 2 x NI: order-k proposals, order-2 proposals
    int32_t* d_ring_ = nullptr;              // NI ringing flags
    double *d_h1_ = nullptr, *d_h2_ = nullptr, *d_h3_ = nullptr;  // sol_prev, sol_prev2, sol_prev3
    int32_t* d_acc_ = nullptr;               // NI: 1 accept, 2 restore, 3 record only, 0 nothing
    int32_t* d_att_ = nullptr;               // NI: attempted a step this round
    int32_t* h_acc_pin_ = nullptr;
    unsigned long long* h_tmin_pin_ = nullptr;
    int32_t* h_ring_pin_ = nullptr;
    std::vector<int32_t> pptr_;              // probes
    int32_t *d_probe_ = nullptr, *d_prow_ = nullptr, *d_pptr_ = nullptr;
    double* d_pvals_ = nullptr;
    int max_points_ = 0;
    std::vector<int32_t> prow_;
    std::vector<int32_t> dc_iterations_;
public:
    const std::vector<int32_t>& dc_iterations() const { return dc_iterations_; }
private:
    std::vector<void*> tran_allocs_;
};

}  // namespace nsa
