// Batched Newton engine: B instances of one finalized Circuit advance
// together through the reference's NIiter loop (src/core/newton.cpp:51-419)
// with per-instance init-mode state, convergence and completion.  Device
// loads, the gather, LU and the convergence test run on the GPU; the
// per-instance mode machine and the numeric()/refactorize() policy run on
// the host from a few bytes per instance read back each iteration.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <map>
#include <string>
#include <vector>

#include "batch_lu.hpp"
#include "circuit.hpp"
#include "bsim4.hpp"
#include "gpu_dev.hpp"

namespace nsa {

struct Options {
    double abstol = 1e-12, reltol = 1e-3, vntol = 1e-6, trtol = 7.0, chgtol = 1e-14;
    double gmin = 1e-12, diag_gmin = 0.0, xmu = 0.5, src_fact = 1.0;
    int itl1 = 100, itl4 = 50;
    int method = 0;  // 0 trap, 1 gear
    int engine = ENGINE_AUTO;
};

struct NewtonOutcome {
    std::vector<int32_t> converged;   // per instance
    std::vector<int32_t> iterations;  // per instance
    long total_iterations = 0;        // sum over instances
};

class BatchEngine {
public:
    BatchEngine(const Circuit& ckt, int batch, const Options& opt);
    ~BatchEngine();
    BatchEngine(const BatchEngine&) = delete;
    BatchEngine& operator=(const BatchEngine&) = delete;

    int batch() const { return B_; }
    int n() const { return n_; }
    hipStream_t stream() const { return stream_; }
    const Circuit& circuit() const { return ckt_; }
    BatchLu& lu() { return *lu_; }

    // Per-instance override of a device's primary value (R ohms, C farads,
    // V/I dc); host array of B values.
    void set_device_values(int dev, const double* vals);
    // Per-instance value of a BSIM4v7 card parameter of device dev (one of
    // the card's value parameters): the device's constants are recomputed
    // per instance with the parameter overridden by vals[inst].
    void set_model_values(int dev, const std::string& param, const double* vals);
    // Integrator context: one for all instances, or per instance.
    void set_ctx(const DevCtx& c);
    void set_ctx(const std::vector<DevCtx>& c);
    const std::vector<DevCtx>& ctx() const { return h_ctx_; }

    // newton_solve on the current solution, with the mode in the context.
    // itl = iteration limit.
    NewtonOutcome newton(int itl);

    // Solution and state access (host arrays are unknown-major: v[i*B+inst]).
    void get_solution(double* out) const;
    void set_solution(const double* in);
    void get_state(int which, double* out) const;
    void set_state(int which, const double* in);
    double* d_solution() { return d_x_; }
    double* d_matrix() { return d_A_; }
    double* d_rhs() { return d_rhs_; }
    double* d_state(int which) { return d_s_[which]; }

    // Transient bookkeeping (transient.cpp:791-802, circuit.cpp:328-334).
    void seed_states_from_state0();  // state1..3 = state0
    void rotate_states();            // state3 <- state2 <- state1 <- state0 (copy)

    // Device local-truncation-error step proposal (compute_trunc of every
    // device, transient.cpp:478-485 and convergence.cpp:948-951) for the
    // current context and states, called after a step converges and before
    // rotate_states.  h_min (B): the minimum over the devices per instance
    // (1e30: no device proposes); h_dev (optional, trunc_devices() x B,
    // instance fastest): each device's value.
    void trunc(double* h_min, double* h_dev);
    const std::vector<int32_t>& trunc_devices() const { return trunc_dev_; }

    // A new solver on the same pattern, as a fresh make_solver + symbolic.
    void reset_solver();

    // One load + gather at the current solution without factoring (for
    // stamp-level parity checks).  Matrix and rhs end up in d_matrix/d_rhs.
    void load_only();

    // ---- parameters and adjoints (adjoint.cpp)
    int family_count(DevKind k) const { return (int)ckt_.family(k).size(); }
    // All per-instance values of a family (R ohms, C farads, V/I dc) at
    // once: count x B, device-major with the instance fastest.
    void set_family_values(DevKind k, const double* h_vals);
    void get_family_values(DevKind k, double* h_vals) const;
    // DC operating-point adjoint on the factors of the last newton() call:
    // J^T lambda = dL/dx, then dL/dp = -lambda^T dF/dp for every R (ohms),
    // C (farads: zero in DC), V and I (dc value).  Host arrays; any output
    // may be null.
    void dc_adjoint(const double* h_dLdx, double* h_lambda, double* h_gR, double* h_gC,
                    double* h_gV, double* h_gI);
    // Fixed-step transient from the current solution as the initial guess:
    // the TRANOP operating point, then nsteps steps of dt (backward Euler
    // first, then the options' method at order 2).  Records the probe nodes'
    // voltages ((nsteps+1) x nprobe x B) and keeps a checkpoint of the full
    // engine state every ckpt_every steps for tran_adjoint.
    void tran_forward(int nsteps, double dt, const int32_t* probes, int nprobe, int ckpt_every,
                      double* h_probes);
    // Discrete adjoint of the last tran_forward: given dL/d(probe values),
    // accumulates dL/dR and dL/dC (count x B).  Segments between checkpoints
    // are recomputed forward, then swept backward with one Jacobian load,
    // refactor and transposed solve per step.
    void tran_adjoint(const double* h_dLdprobe, double* h_gR, double* h_gC);
    // Wall-clock milliseconds of the last tran_forward / tran_adjoint.
    double ms_forward = 0.0, ms_adjoint = 0.0;

    // Timing breakdown of the last newton() call, milliseconds.
    double ms_load = 0, ms_factor = 0, ms_solve = 0;
    bool time_phases = false;
    // HIP-event timing of device load + gather (this engine's stream) and of
    // the solver's refactor / solve; accumulated since set_timing(true).
    void set_timing(bool on);
    double ms_load_gpu = 0.0;
    double ms_bsim4_gpu = 0.0;  // the BSIM4v7 load kernel alone (HIP events)
    long n_loads = 0;

private:
    void build_tables();
    void ensure_adjoint();
    double* family_values_ptr(DevKind k) const;
    void factor_current();
    void jacobian_at(const double* d_x, const DevCtx& c);
    DevCtx step_ctx(int k, double dt) const;
    void save_checkpoint(int slot);
    void restore_checkpoint(int slot);
    void load(const int32_t* d_done);

    const Circuit& ckt_;
    int B_, n_, nnz_, ns_;
    Options opt_;
    hipStream_t stream_ = nullptr;
    std::unique_ptr<BatchLu> lu_;
    std::vector<void*> allocs_;
    double *d_x_ = nullptr, *d_xold_ = nullptr, *d_rhs_ = nullptr, *d_A_ = nullptr;
    double* d_s_[4] = {nullptr, nullptr, nullptr, nullptr};
    TruncTable trunc_{};
    std::vector<int32_t> trunc_dev_;          // circuit device index per trunc row
    unsigned long long* d_tmin_ = nullptr;    // B
    double* d_tdev_ = nullptr;                // trunc rows x B
    double *d_cm_ = nullptr, *d_cr_ = nullptr;
    DevCtx* d_ctx_ = nullptr;
    int32_t *d_done_ = nullptr, *d_nonconv_ = nullptr;
    std::vector<DevCtx> h_ctx_;
    // Pinned staging for the per-iteration host->device flags and contexts
    // (4 x B int32: live, not-live, restore, done) so that asynchronous
    // copies never read host memory that is rewritten or freed.
    int32_t* h_flags_ = nullptr;
    DevCtx* h_ctx_pin_ = nullptr;
    void upload_ctx();
    void upload_flags(int slot, const std::vector<int32_t>& v, int32_t* dst);
    bool r_dirty_ = false;
    // family tables
    CapTable cap_{};
    VsrcTable vsrc_{};
    IsrcTable isrc_{};
    DioTable dio_{};
    Mos1Table mos_{};
    B4Table b4_{};
    DeviceEnv env_{};
    std::vector<B4Param> b4_host_;                              // shared records
    std::vector<std::vector<std::map<std::string, double>>> b4_over_;  // per device, instance
    double* d_b4_inst_ = nullptr;
    DevTol tol_{};
    // per-instance device non-convergence of the current iteration
    // (CKTnoncon of the loads, device_converged after the solve)
    int32_t* d_devnc_ = nullptr;
    int nR_ = 0, r_site_ = 0;
    double* d_r_ohms_ = nullptr;
    double *d_c_val_ = nullptr, *d_v_val_ = nullptr, *d_i_val_ = nullptr;
    int const_site_ = 0, n_const_ = 0;
    double* d_const_vals_ = nullptr;
    // gathers
    int32_t *d_mptr_ = nullptr, *d_msites_ = nullptr;
    int32_t *d_rptr_d_ = nullptr, *d_rsites_d_ = nullptr, *d_rptr_u_ = nullptr,
            *d_rsites_u_ = nullptr;
    int n_msites_ = 0, n_rsites_ = 0;
    std::vector<int32_t> dev_site_;  // first matrix site per device (insertion index)
    std::vector<int32_t> dev_rsite_; // first rhs site per device
    bool timing_ = false;
    hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
    // adjoint tables and buffers (allocated on first use)
    bool adj_ready_ = false;
    int32_t *d_r_nodes_ = nullptr, *d_v_branch_ = nullptr, *d_i_nodes_ = nullptr;
    int32_t *d_cinc_ptr_ = nullptr, *d_cinc_ent_ = nullptr;
    double *d_lam_ = nullptr, *d_carry_ = nullptr, *d_ibar_ = nullptr, *d_q_ = nullptr;
    double *d_gR_ = nullptr, *d_gC_ = nullptr, *d_gV_ = nullptr, *d_gI_ = nullptr;
    // transient record of the last tran_forward
    int tr_steps_ = 0, tr_ckpt_ = 0, tr_nprobe_ = 0;
    double tr_dt_ = 0.0;
    std::vector<DevCtx> tr_ctx_;       // per step 0..nsteps (0 = TRANOP)
    int32_t* d_probes_ = nullptr;
    double* d_probe_vals_ = nullptr;   // (nsteps+1) x nprobe x B
    double* d_ckpt_ = nullptr;         // checkpoints: x, s0..s3, MOS1 aux
    size_t ckpt_stride_ = 0;
    double* d_segx_ = nullptr;         // (ckpt+1) x n x B solutions of one segment
    double* d_dlp_ = nullptr;          // dL/d(probe values)
    std::vector<void*> tr_allocs_;
    void free_tran();
};

}  // namespace nsa
