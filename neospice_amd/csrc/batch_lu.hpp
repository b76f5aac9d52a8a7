// The solver seam: the reference's ISolver contract (src/core/solver_iface.hpp:16-41)
// for a batch of B instances that share one sparsity pattern.
//
//   symbolic(pattern)       -> choose the engine like make_solver
//                              (make_solver.cpp:156-176) and keep the pattern
//   numeric(A, gmin)        -> pivot order fixed on the CPU from real values,
//                              factors computed on the GPU; later calls
//                              verify on the GPU that the order still holds
//   refactorize(A, gmin)    -> GPU refactor on the fixed order
//   solve(b) / solve_t(b)   -> GPU forward/back substitution (in place)
//
// Every instance behaves like its own reference solver: pivot orders are
// value-dependent, so instances are partitioned into structure groups, one
// GPU plan per distinct pivot order.  An instance whose order changes (a
// failed threshold check on numeric(), or a rejected AMD-LU replay) is
// re-pivoted on the CPU from its own values and moved to the group of its
// new order.
//
// A and b live in device memory, instance-interleaved (v[i * B + inst]).
// Per-instance outcomes are reported in status[] (0 ok, 1 singular).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "gpu_lu.hpp"
#include "lu_cpu.hpp"
#include "lu_structure.hpp"

namespace nsa {

enum Engine : int32_t { ENGINE_AUTO = 0, ENGINE_MARKOWITZ = 1, ENGINE_AMDLU = 2 };

class BatchLu {
public:
    BatchLu(int batch, hipStream_t stream);
    ~BatchLu();

    void symbolic(int n, const int32_t* col_ptr, const int32_t* row_idx, bool is_linear,
                  int engine = ENGINE_AUTO);
    // mask (host, B entries, may be null = all): instances whose call is a
    // numeric(); the others are refactorize().  Returns the number of
    // singular instances; status[inst] = 1 for each.
    int factor(const double* dA, double diag_gmin, const int32_t* numeric_mask, int32_t* status);
    int numeric(const double* dA, double diag_gmin, int32_t* status) {
        return factor(dA, diag_gmin, nullptr, status);
    }
    int refactorize(const double* dA, double diag_gmin, int32_t* status);
    void solve(double* db);
    void solve_t(double* db);
    // Row/column equilibration (ISolver::set_equilibrate; NeoSolver's
    // radix-2 scaling, neo_solver.cpp:52-102 and 204-225).  Honoured where
    // the reference's solver would honour it: on Markowitz (NeoSolver)
    // solvers; AMD-LU and the AMD-LU auto-fallback ignore it, as
    // AmdLuSolver and AutoFallbackSolver inherit the interface's no-op.
    void set_equilibrate(bool on) { equil_ = on; }
    bool equilibrating() const { return equil_ && honor_equil_; }

    // the engine of instance 0 (AMD-LU auto-selection can fall back to
    // Markowitz per instance, as AutoFallbackSolver does per solver)
    const char* name() const {
        const LuForm f = iform_.empty() ? form_ : iform_[0];
        return f == FORM_S13 ? "markowitz" : "amdlu";
    }
    bool factored() const { return !groups_.empty() && factored_; }
    int num_groups() const { return (int)groups_.size(); }
    // The plan of structure group g (group 0 holds the first instance's order).
    GpuLu* gpu(int g = 0) { return groups_.empty() ? nullptr : groups_[g].lu.get(); }
    const LuStructure* structure(int g = 0) const {
        return groups_.empty() ? nullptr : &groups_[g].lu->structure();
    }
    int group_of(int inst) const { return group_of_[inst]; }
    int n() const { return n_; }
    int batch() const { return B_; }
    // counters for tests and the bench
    long n_cpu_pivot = 0;     // CPU pivot computations
    long n_gpu_refactor = 0;  // GPU refactors of the batch (each = all levels of every group)
    long n_gpu_solve = 0;
    long n_regroup = 0;       // instances moved to another pivot order
    // HIP-event timing of the GPU refactor / solve on this object's stream
    // (enabled by set_timing); accumulated milliseconds.
    void set_timing(bool on);
    void collect_timing();
    double ms_refactor = 0.0, ms_solve = 0.0;

private:
    struct Group {
        std::unique_ptr<GpuLu> lu;
        std::vector<int32_t> inst;
        std::unique_ptr<Sparse13> s13;  // Markowitz: CPU state after the numeric() that fixed the order
    };
    std::vector<double> instance_values(const double* dA, int inst);
    bool cpu_order(LuForm form, const std::vector<double>& vals, double gmin, const Sparse13* from,
                   LuStructure& out, std::unique_ptr<Sparse13>& state);
    int order_fresh(const double* dA, double gmin, std::vector<int32_t>& kind, int32_t* status);
    int find_group(const LuStructure& s) const;
    void rebuild_group(int g);
    void run_gpu(const double* dA, double gmin, const std::vector<int32_t>& kind,
                 const std::vector<int>& which, std::vector<int32_t>& fail);

    const double* equilibrate(const double* dA, double& gmin);

    int B_;
    hipStream_t stream_;
    bool equil_ = false, honor_equil_ = false;
    double *d_eqA_ = nullptr, *d_dr_ = nullptr, *d_dc_ = nullptr;
    int32_t *d_erows_ = nullptr, *d_ecols_ = nullptr;
    int n_ = 0;
    std::vector<int32_t> acp_, ari_;
    LuForm form_ = FORM_S13;      // the engine make_solver's policy picks
    bool auto_amd_ = false;       // AMD-LU by the auto policy: a singular first factorization falls back
    std::vector<LuForm> iform_;   // per instance: the engine factoring it
    std::vector<int32_t> q_;
    std::vector<Group> groups_;
    std::vector<int32_t> group_of_;
    bool factored_ = false;
    int32_t *d_fail_ = nullptr, *d_fail_col_ = nullptr, *d_kind_ = nullptr;
    std::vector<int32_t> h_kind_;
    // pinned staging for every host<->device transfer of the factor driver
    int32_t *h_kind_pin_ = nullptr, *h_fail_pin_ = nullptr;
    double* h_vals_pin_ = nullptr;   // one instance's matrix values (nnz)
    double* d_vals_ = nullptr;
    bool timing_ = false, pend_rf_ = false, pend_sv_ = false;
    hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
    // Pivot groups run concurrently: group 0 on the main stream, group g on
    // side stream g-1, forked from and joined back into the main stream.
    std::vector<hipStream_t> side_;
    std::vector<hipEvent_t> join_ev_;
    hipEvent_t fork_ev_ = nullptr;
    template <class F>
    void for_groups(const std::vector<int>& which, F&& launch);
    std::vector<int> all_groups() const {
        std::vector<int> w(groups_.size());
        for (size_t g = 0; g < w.size(); ++g) w[g] = (int)g;
        return w;
    }
};

}  // namespace nsa
