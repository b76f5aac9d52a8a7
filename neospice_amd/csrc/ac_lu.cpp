#include "ac_lu.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "parallel.hpp"

namespace nsa {

AcLu::AcLu(int n, const int32_t* col_ptr, const int32_t* row_idx, int batch, hipStream_t stream)
    : n_(n), nnz_(col_ptr[n]), B_(batch), stream_(stream), acp_(col_ptr, col_ptr + n + 1),
      ari_(row_idx, row_idx + col_ptr[n]), s13_(batch), plan_of_(batch, -1) {
    if (batch <= 0) throw Error(NSA_ERR_ARGUMENT, "AcLu: batch must be positive");
}

AcLu::~AcLu() {
    for (double* p : {d_ax_, d_lu_, d_v_, d_g_, d_c_, d_w_, d_r_})
        if (p) (void)hipFree(p);
    if (d_inst_) (void)hipFree(d_inst_);
    if (d_fail_) (void)hipFree(d_fail_);
}

template <class T>
T* AcLu::grow(T*& p, size_t& cap, size_t count) {
    count = std::max<size_t>(count, 1);
    if (count > cap) {
        // launches in flight may still use the old buffer
        NSA_HIP(hipStreamSynchronize(stream_));
        if (p) NSA_HIP(hipFree(p));
        NSA_HIP(hipMalloc(&p, sizeof(T) * count));
        cap = count;
    }
    return p;
}

// NeoSolver::numeric_complex of instance inst on its own matrix: the
// complex Markowitz ordering (a fresh one on the first call, the reuse
// test on an existing order after) fixes its structure.  Instances are
// independent, so a batch orders them on host threads; their plans are
// added in instance order afterwards.
int AcLu::order_one(const double* ax, int inst, LuStructure& st) {
    if (!s13_[inst]) s13_[inst] = std::make_unique<Sparse13>(n_, acp_.data(), ari_.data());
    if (s13_[inst]->numeric_complex(ax)) return 1;
    st = s13_[inst]->structure();
    if (!SmallLuPlans::fits(st))
        throw Error(NSA_ERR_UNSUPPORTED, "complex LU: the factor exceeds the packed kernel's limits");
    return 0;
}

void AcLu::order(const std::vector<int>& which, const std::function<const double*(int)>& ax,
                 std::vector<int32_t>& singular) {
    std::vector<LuStructure> st(which.size());
    std::vector<int32_t> sing(which.size(), 0);
    parallel_for((int)which.size(), [&](int k) { sing[k] = order_one(ax(which[k]), which[k], st[k]); });
    n_cpu_order += (long)which.size();
    for (size_t k = 0; k < which.size(); ++k) {
        const int i = which[k];
        singular[i] = sing[k];
        plan_of_[i] = sing[k] ? -1 : plans_.add(st[k]);
    }
}

void AcLu::upload_inst(const std::vector<SluInst>& inst) {
    grow(d_inst_, cap_inst_, inst.size());
    grow(d_fail_, cap_fail_, inst.size());
    NSA_HIP(hipMemcpyAsync(d_inst_, inst.data(), sizeof(SluInst) * inst.size(), hipMemcpyHostToDevice, stream_));
}

int AcLu::numeric_complex(const double* ax, int32_t* status) {
    std::vector<int32_t> host(B_, 0), st(B_, 0);
    std::vector<int> all(B_);
    for (int i = 0; i < B_; ++i) all[i] = i;
    order(all, [&](int i) { return ax + (size_t)2 * nnz_ * i; }, host);
    ordered_ = true;
    factor_gpu(ax, st);
    int k = 0;
    for (int i = 0; i < B_; ++i) {
        const int32_t s = (host[i] || st[i]) ? 1 : 0;
        if (status) status[i] = s;
        k += s;
    }
    return k;
}

int AcLu::refactorize_complex(const double* ax, int32_t* status) {
    if (!ordered_) throw Error(NSA_ERR_LOGIC, "refactorize_complex: numeric_complex() not called");
    // an instance whose ordering failed is ordered again, as the
    // reference's factor() orders a matrix that still needs ordering
    std::vector<int32_t> host(B_, 0), st(B_, 0);
    std::vector<int> redo;
    for (int i = 0; i < B_; ++i)
        if (plan_of_[i] < 0) redo.push_back(i);
    order(redo, [&](int i) { return ax + (size_t)2 * nnz_ * i; }, host);
    factor_gpu(ax, st);
    int k = 0;
    for (int i = 0; i < B_; ++i) {
        const int32_t s = (host[i] || st[i]) ? 1 : 0;
        if (status) status[i] = s;
        k += s;
    }
    return k;
}

// The complex factor of every instance on its plan, into the factor slabs;
// st[i] = 1 where it fails (an exact zero pivot, or no order).
void AcLu::factor_gpu(const double* ax, std::vector<int32_t>& st) {
    plans_.sync(stream_);
    const int max_nlu = plans_.max_nlu();
    std::vector<SluInst> inst(B_);
    for (int i = 0; i < B_; ++i) inst[i] = SluInst{plan_of_[i], i * nnz_, i * max_nlu, i * n_};
    grow(d_ax_, cap_ax_, (size_t)2 * nnz_ * B_);
    grow(d_lu_, cap_lu_, (size_t)2 * std::max(1, max_nlu) * B_);
    upload_inst(inst);
    NSA_HIP(hipMemcpyAsync(d_ax_, ax, sizeof(double) * 2 * nnz_ * B_, hipMemcpyHostToDevice, stream_));
    launch_small_lu_complex(plans_, d_inst_, B_, SLU_FACTOR, d_ax_, d_fail_, nullptr, d_lu_, nullptr, nullptr,
                            stream_);
    std::vector<int32_t> fail(B_);
    NSA_HIP(hipMemcpyAsync(fail.data(), d_fail_, sizeof(int32_t) * B_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    for (int i = 0; i < B_; ++i) st[i] = fail[i] != FAIL_NONE ? 1 : 0;
    solve_inst_ = inst;
    factored_ = true;
}

void AcLu::solve_complex(double* rhs) {
    if (!factored_) throw Error(NSA_ERR_LOGIC, "solve_complex: numeric_complex() not called");
    grow(d_v_, cap_v_, (size_t)2 * n_ * B_);
    upload_inst(solve_inst_);
    NSA_HIP(hipMemcpyAsync(d_v_, rhs, sizeof(double) * 2 * n_ * B_, hipMemcpyHostToDevice, stream_));
    launch_small_lu_complex(plans_, d_inst_, B_, SLU_SOLVE, nullptr, nullptr, nullptr, d_lu_, d_v_, d_v_, stream_);
    NSA_HIP(hipMemcpyAsync(rhs, d_v_, sizeof(double) * 2 * n_ * B_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

int AcLu::sweep(const double* g, const double* c, const double* freqs, int nf, const double* rhs, double* out,
                int32_t* status) {
    if (nf <= 0) throw Error(NSA_ERR_ARGUMENT, "AC sweep: no frequencies");
    const auto t_start = std::chrono::steady_clock::now();
    // omega = 2 pi f as ac.cpp:197 forms it
    std::vector<double> w(nf);
    for (int f = 0; f < nf; ++f) w[f] = 2.0 * M_PI * freqs[f];
    // every instance ordered at the first frequency
    std::vector<int32_t> host(B_, 0);
    std::vector<double> ax0((size_t)2 * nnz_ * B_);
    std::vector<int> all(B_);
    for (int i = 0; i < B_; ++i) {
        all[i] = i;
        for (int k = 0; k < nnz_; ++k) {
            ax0[(size_t)2 * nnz_ * i + 2 * k] = g[(size_t)i * nnz_ + k];
            ax0[(size_t)2 * nnz_ * i + 2 * k + 1] = w[0] * c[(size_t)i * nnz_ + k];
        }
    }
    order(all, [&](int i) { return ax0.data() + (size_t)2 * nnz_ * i; }, host);
    ordered_ = true;
    last_order_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    plans_.sync(stream_);
    const long blocks = (long)B_ * nf;
    std::vector<SluInst> inst(blocks);
    for (int i = 0; i < B_; ++i)
        for (int f = 0; f < nf; ++f) {
            const long b = (long)i * nf + f;
            inst[b] = SluInst{plan_of_[i], (int32_t)(b * nnz_), 0, (int32_t)(b * n_)};
        }
    if (blocks * std::max(nnz_, n_) > 0x7fffffffL)
        throw Error(NSA_ERR_ARGUMENT, "AC sweep: instances x frequencies too large for 32-bit offsets");
    grow(d_g_, cap_g_, (size_t)nnz_ * B_);
    grow(d_c_, cap_c_, (size_t)nnz_ * B_);
    grow(d_w_, cap_w_, (size_t)nf);
    grow(d_ax_, cap_ax_, (size_t)2 * nnz_ * blocks);
    grow(d_v_, cap_v_, (size_t)2 * n_ * blocks);
    grow(d_r_, cap_r_, (size_t)2 * n_ * B_);
    upload_inst(inst);
    NSA_HIP(hipMemcpyAsync(d_g_, g, sizeof(double) * nnz_ * B_, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipMemcpyAsync(d_c_, c, sizeof(double) * nnz_ * B_, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipMemcpyAsync(d_w_, w.data(), sizeof(double) * nf, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipMemcpyAsync(d_r_, rhs, sizeof(double) * 2 * n_ * B_, hipMemcpyHostToDevice, stream_));
    hipEvent_t ev[2];
    NSA_HIP(hipEventCreate(&ev[0]));
    NSA_HIP(hipEventCreate(&ev[1]));
    NSA_HIP(hipEventRecord(ev[0], stream_));
    launch_ac_build(d_g_, d_c_, d_w_, B_, nf, nnz_, d_ax_, stream_);
    launch_ac_rhs(d_r_, B_, nf, n_, d_v_, stream_);
    launch_small_lu_complex(plans_, d_inst_, (int)blocks, SLU_FACTOR | SLU_SOLVE, d_ax_, d_fail_, nullptr, nullptr,
                            d_v_, d_v_, stream_);
    NSA_HIP(hipEventRecord(ev[1], stream_));
    std::vector<int32_t> fail(blocks);
    NSA_HIP(hipMemcpyAsync(out, d_v_, sizeof(double) * 2 * n_ * blocks, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipMemcpyAsync(fail.data(), d_fail_, sizeof(int32_t) * blocks, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    float ms = 0.0f;
    NSA_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    NSA_HIP(hipEventDestroy(ev[0]));
    NSA_HIP(hipEventDestroy(ev[1]));
    last_device_s = ms * 1e-3;
    last_total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    // the factor slabs hold factors of orders the sweep may have replaced:
    // solves need a refactorize_complex first
    factored_ = false;
    int k = 0;
    for (long b = 0; b < blocks; ++b) {
        const int32_t s = (host[b / nf] || fail[b] != FAIL_NONE) ? 1 : 0;
        if (status) status[b] = s;
        k += s;
    }
    return k;
}

}  // namespace nsa
