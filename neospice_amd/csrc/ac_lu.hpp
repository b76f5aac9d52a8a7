// Batched complex (AC) factorization and solves: ISolver's complex methods
// (numeric_complex, refactorize_complex, solve_complex;
// src/core/solver_iface.hpp:23-28) for a batch of independent solvers, and
// the AC analysis's frequency sweep (src/core/ac.cpp:196-238) over every
// instance and frequency at once.
//
// The reference's complex paths are NeoSolver's whatever the engine
// (AmdLuSolver delegates them, amd_lu_solver.cpp:347-349): a complex
// Markowitz ordering on the host per instance (Sparse13::numeric_complex),
// then the complex packed LU on the GPU for the values and the solves.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "lu_cpu.hpp"
#include "small_lu.hpp"

namespace nsa {

class AcLu {
public:
    AcLu(int n, const int32_t* col_ptr, const int32_t* row_idx, int batch, hipStream_t stream);
    ~AcLu();
    AcLu(const AcLu&) = delete;
    AcLu& operator=(const AcLu&) = delete;

    // Host arrays, instance-major: ax holds batch blocks of 2 nnz values
    // ((re, im) per CSC slot, as ISolver's ax); rhs batch blocks of 2 n.
    // numeric_complex orders every instance from its own values (the first
    // call afresh, later ones through the order's reuse test, as a
    // NeoSolver's) and factors it; refactorize_complex factors on the
    // orders held.  status[i] = 1 for a singular instance; both return the
    // number of singular instances.
    int numeric_complex(const double* ax, int32_t* status);
    int refactorize_complex(const double* ax, int32_t* status);
    void solve_complex(double* rhs);
    // The sweep of ac.cpp for every instance: ax_i(f) = g_i + j (2 pi f) c_i,
    // ordered at freqs[0] (numeric_complex), every frequency factored on that
    // order and solved for rhs_i.  g, c: batch blocks of nnz; rhs: batch
    // blocks of 2 n; out: (instance, frequency) blocks of 2 n, instance-major.
    int sweep(const double* g, const double* c, const double* freqs, int nf, const double* rhs, double* out,
              int32_t* status);

    // seconds of the last sweep: host orderings, device work (matrix build,
    // factor and solve, by HIP events), whole call
    double last_order_s = 0.0, last_device_s = 0.0, last_total_s = 0.0;

    int batch() const { return B_; }
    int plans() const { return plans_.count(); }
    long n_cpu_order = 0;  // host orderings

private:
    int order_one(const double* ax, int inst, LuStructure& st);
    void order(const std::vector<int>& which, const std::function<const double*(int)>& ax,
               std::vector<int32_t>& singular);
    void factor_gpu(const double* ax, std::vector<int32_t>& st);
    void upload_inst(const std::vector<SluInst>& inst);
    template <class T>
    T* grow(T*& p, size_t& cap, size_t count);

    int n_, nnz_, B_;
    hipStream_t stream_;
    std::vector<int32_t> acp_, ari_;
    std::vector<std::unique_ptr<Sparse13>> s13_;
    std::vector<int> plan_of_;
    bool ordered_ = false;   // numeric_complex (or a sweep) fixed the orders
    bool factored_ = false;  // the factor slabs hold the current orders' factors
    SmallLuPlans plans_;
    std::vector<SluInst> solve_inst_;  // the descriptors of the last factor
    double *d_ax_ = nullptr, *d_lu_ = nullptr, *d_v_ = nullptr, *d_g_ = nullptr, *d_c_ = nullptr, *d_w_ = nullptr,
           *d_r_ = nullptr;
    size_t cap_ax_ = 0, cap_lu_ = 0, cap_v_ = 0, cap_g_ = 0, cap_c_ = 0, cap_w_ = 0, cap_r_ = 0;
    SluInst* d_inst_ = nullptr;
    size_t cap_inst_ = 0;
    int32_t* d_fail_ = nullptr;
    size_t cap_fail_ = 0;
};

}  // namespace nsa
