// Batched Newton engine: B instances of one finalized Circuit advance
// together through the reference's NIiter loop (src/core/newton.cpp:51-419)
// with per-instance init-mode state, convergence and completion.  Device
// loads, the gather, LU and the convergence test run on the GPU; the
// per-instance mode machine and the numeric()/refactorize() policy run on
// the host from a few bytes per instance read back each iteration.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "batch_lu.hpp"
#include "circuit.hpp"
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

    // A new solver on the same pattern, as a fresh make_solver + symbolic.
    void reset_solver();

    // One load + gather at the current solution without factoring (for
    // stamp-level parity checks).  Matrix and rhs end up in d_matrix/d_rhs.
    void load_only();

    // Timing breakdown of the last newton() call, milliseconds.
    double ms_load = 0, ms_factor = 0, ms_solve = 0;
    bool time_phases = false;
    // HIP-event timing of device load + gather (this engine's stream) and of
    // the solver's refactor / solve; accumulated since set_timing(true).
    void set_timing(bool on);
    double ms_load_gpu = 0.0;
    long n_loads = 0;

private:
    void build_tables();
    void load(const int32_t* d_done);

    const Circuit& ckt_;
    int B_, n_, nnz_, ns_;
    Options opt_;
    hipStream_t stream_ = nullptr;
    std::unique_ptr<BatchLu> lu_;
    std::vector<void*> allocs_;
    double *d_x_ = nullptr, *d_xold_ = nullptr, *d_rhs_ = nullptr, *d_A_ = nullptr;
    double* d_s_[4] = {nullptr, nullptr, nullptr, nullptr};
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
    hipEvent_t ev_[2] = {nullptr, nullptr};
};

}  // namespace nsa
