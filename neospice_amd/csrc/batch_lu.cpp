#include "batch_lu.hpp"

#include <algorithm>
#include <cstring>

#include "common.hpp"
#include "gpu_dev.hpp"
#include "gpu_equil.hpp"
#include "parallel.hpp"

namespace nsa {

// make_solver.cpp:37 -- AMD-LU is auto-selected for linear circuits with at
// least this many unknowns.
static constexpr int kAmdLuAutoThreshold = 256;

// Per-instance call kinds handed to the refactor kernel.
enum : int32_t { KIND_REFACTOR = 0, KIND_VERIFY = 1, KIND_TRUSTED = 2, KIND_SKIP = 3 };

BatchLu::BatchLu(int batch, hipStream_t stream) : B_(batch), stream_(stream) {
    if (batch <= 0) throw Error(NSA_ERR_ARGUMENT, "BatchLu: batch must be positive");
    NSA_HIP(hipMalloc(&d_fail_, sizeof(int32_t) * B_));
    NSA_HIP(hipMalloc(&d_fail_col_, sizeof(int32_t) * B_));
    NSA_HIP(hipMalloc(&d_kind_, sizeof(int32_t) * B_));
    h_kind_.assign(B_, -1);
    group_of_.assign(B_, -1);
    NSA_HIP(hipHostMalloc((void**)&h_kind_pin_, sizeof(int32_t) * B_, hipHostMallocDefault));
    NSA_HIP(hipHostMalloc((void**)&h_fail_pin_, sizeof(int32_t) * B_, hipHostMallocDefault));
}

BatchLu::~BatchLu() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    groups_.clear();
    for (auto st : side_) (void)hipStreamDestroy(st);
    for (auto e : join_ev_) (void)hipEventDestroy(e);
    if (fork_ev_) (void)hipEventDestroy(fork_ev_);
    for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
    (void)hipFree(d_fail_);
    (void)hipFree(d_fail_col_);
    (void)hipFree(d_kind_);
    (void)hipFree(d_vals_);
    for (void* p : {(void*)d_eqA_, (void*)d_dr_, (void*)d_dc_, (void*)d_erows_, (void*)d_ecols_})
        if (p) (void)hipFree(p);
    (void)hipHostFree(h_kind_pin_);
    (void)hipHostFree(h_fail_pin_);
    if (h_vals_pin_) (void)hipHostFree(h_vals_pin_);
}

void BatchLu::symbolic(int n, const int32_t* col_ptr, const int32_t* row_idx, bool is_linear,
                       int engine) {
    n_ = n;
    acp_.assign(col_ptr, col_ptr + n + 1);
    ari_.assign(row_idx, row_idx + col_ptr[n]);
    groups_.clear();
    group_of_.assign(B_, -1);
    factored_ = false;
    auto_amd_ = false;
    if (engine == ENGINE_AMDLU) {
        form_ = FORM_GP;
    } else if (engine == ENGINE_MARKOWITZ) {
        form_ = FORM_S13;
    } else {
        form_ = (n >= kAmdLuAutoThreshold && is_linear) ? FORM_GP : FORM_S13;
        auto_amd_ = form_ == FORM_GP;
    }
    honor_equil_ = form_ == FORM_S13;
    iform_.assign(B_, form_);
    for (void* p : {(void*)d_eqA_, (void*)d_dr_, (void*)d_dc_, (void*)d_erows_, (void*)d_ecols_})
        if (p) NSA_HIP(hipFree(p));
    d_eqA_ = d_dr_ = d_dc_ = nullptr;
    d_erows_ = d_ecols_ = nullptr;
    if (form_ == FORM_GP) {
        q_ = min_degree_order(n, acp_.data(), ari_.data());
        if ((int)q_.size() != n) {
            q_.resize(n);
            for (int i = 0; i < n; ++i) q_[i] = i;
        }
    }
}

std::vector<double> BatchLu::instance_values(const double* dA, int inst) {
    const int nnz = acp_[n_];
    if (!d_vals_) {
        NSA_HIP(hipMalloc(&d_vals_, sizeof(double) * std::max(1, nnz)));
        NSA_HIP(hipHostMalloc((void**)&h_vals_pin_, sizeof(double) * std::max(1, nnz),
                              hipHostMallocDefault));
    }
    launch_take_instance(dA, B_, inst, nnz, d_vals_, stream_);
    NSA_HIP(hipMemcpyAsync(h_vals_pin_, d_vals_, sizeof(double) * nnz, hipMemcpyDeviceToHost,
                           stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    return std::vector<double>(h_vals_pin_, h_vals_pin_ + nnz);
}

// CPU pivot computation for one instance's values.  Markowitz continues
// from `from` (the state of the instance's previous numeric(), reuse path)
// or starts fresh; AMD-LU always re-pivots from scratch (factor_).  Returns
// true when singular.
bool BatchLu::cpu_order(LuForm form, const std::vector<double>& vals, double gmin, const Sparse13* from,
                        LuStructure& out, std::unique_ptr<Sparse13>& state) {
    if (form == FORM_S13) {
        state = from ? from->clone() : std::make_unique<Sparse13>(n_, acp_.data(), ari_.data());
        if (state->numeric(vals.data(), gmin) != 0) return true;
        out = state->structure();
        return false;
    }
    state.reset();
    return !gp_factor_structure(n_, acp_.data(), ari_.data(), vals.data(), gmin, q_, out);
}

int BatchLu::find_group(const LuStructure& s) const {
    for (size_t g = 0; g < groups_.size(); ++g)
        if (groups_[g].lu && groups_[g].lu->structure().same_as(s)) return (int)g;
    return -1;
}

void BatchLu::rebuild_group(int g) {
    Group& G = groups_[g];
    LuStructure s = G.lu->structure();
    std::sort(G.inst.begin(), G.inst.end());
    const bool all = (int)G.inst.size() == B_;
    G.lu.reset();
    G.lu = std::make_unique<GpuLu>(s, (int)G.inst.size(), stream_, all ? nullptr : &G.inst, B_);
}

template <class F>
void BatchLu::for_groups(const std::vector<int>& which, F&& launch) {
    if (which.size() <= 1) {
        for (int g : which) launch(g, stream_);
        return;
    }
    const size_t nside = which.size() - 1;
    if (!fork_ev_) NSA_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
    while (side_.size() < nside) {
        hipStream_t st;
        NSA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        side_.push_back(st);
        hipEvent_t e;
        NSA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        join_ev_.push_back(e);
    }
    NSA_HIP(hipEventRecord(fork_ev_, stream_));
    for (size_t i = 1; i < which.size(); ++i) {
        NSA_HIP(hipStreamWaitEvent(side_[i - 1], fork_ev_, 0));
        launch(which[i], side_[i - 1]);
        NSA_HIP(hipEventRecord(join_ev_[i - 1], side_[i - 1]));
    }
    launch(which[0], stream_);
    for (size_t i = 1; i < which.size(); ++i) NSA_HIP(hipStreamWaitEvent(stream_, join_ev_[i - 1], 0));
}

void BatchLu::run_gpu(const double* dA, double gmin, const std::vector<int32_t>& kind,
                      const std::vector<int>& which, std::vector<int32_t>& fail) {
    launch_fill_i32(d_fail_, FAIL_NONE, B_, stream_);
    launch_fill_i32(d_fail_col_, 0x7fffffff, B_, stream_);
    if (kind != h_kind_) {
        // the stream is idle here: every run ends with a synchronize
        h_kind_ = kind;
        std::memcpy(h_kind_pin_, kind.data(), sizeof(int32_t) * B_);
        NSA_HIP(hipMemcpyAsync(d_kind_, h_kind_pin_, sizeof(int32_t) * B_,
                               hipMemcpyHostToDevice, stream_));
    }
    if (timing_) NSA_HIP(hipEventRecord(ev_[0], stream_));
    for_groups(which, [&](int g, hipStream_t st) {
        groups_[g].lu->refactor(dA, gmin, d_kind_, d_fail_, d_fail_col_, st);
    });
    if (timing_) { NSA_HIP(hipEventRecord(ev_[1], stream_)); pend_rf_ = true; }
    ++n_gpu_refactor;
    fail.resize(B_);
    NSA_HIP(hipMemcpyAsync(h_fail_pin_, d_fail_, sizeof(int32_t) * B_, hipMemcpyDeviceToHost,
                           stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(fail.data(), h_fail_pin_, sizeof(int32_t) * B_);
    for (int i = 0; i < B_; ++i)
        if (fail[i] == FAIL_TIMEOUT)
            throw Error(NSA_ERR_RUNTIME, "GPU refactor: a dependency wait ran out of its bound");
    collect_timing();
}

// The equilibrated system of dA: D_r A D_c with the scaled diag_gmin on
// the diagonal slots (gmin is then 0 for the factorization, as NeoSolver
// skips add_diag_gmin when equilibrating, neo_solver.cpp:160-166).
const double* BatchLu::equilibrate(const double* dA, double& gmin) {
    const int nnz = acp_[n_];
    if (!d_eqA_) {
        std::vector<int32_t> cols(nnz);
        for (int j = 0; j < n_; ++j)
            for (int p = acp_[j]; p < acp_[j + 1]; ++p) cols[p] = j;
        NSA_HIP(hipMalloc(&d_eqA_, sizeof(double) * std::max(1, nnz) * B_));
        NSA_HIP(hipMalloc(&d_dr_, sizeof(double) * n_ * B_));
        NSA_HIP(hipMalloc(&d_dc_, sizeof(double) * n_ * B_));
        NSA_HIP(hipMalloc(&d_erows_, sizeof(int32_t) * std::max(1, nnz)));
        NSA_HIP(hipMalloc(&d_ecols_, sizeof(int32_t) * std::max(1, nnz)));
        NSA_HIP(hipMemcpyAsync(d_erows_, ari_.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, stream_));
        NSA_HIP(hipMemcpyAsync(d_ecols_, cols.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, stream_));
        NSA_HIP(hipStreamSynchronize(stream_));
    }
    launch_equilibrate(n_, nnz, d_erows_, d_ecols_, dA, gmin, B_, d_eqA_, d_dr_, d_dc_, stream_);
    gmin = 0.0;
    return d_eqA_;
}

// Every instance without a factorization is ordered on the CPU from its own
// values, as the first numeric() of its own solver does (a refactorize()
// there is a numeric(): NeoSolver throws and newton_solve retries with
// numeric(), newton.cpp:235-242; AmdLuSolver runs factor_,
// amd_lu_solver.cpp:280-301).  Under the auto policy a singular AMD-LU
// order falls back to Markowitz for that instance (AutoFallbackSolver,
// make_solver.cpp:85-101).  Instances whose structures are identical share
// a group (one GPU plan); those are TRUSTED in this call's factorization,
// and singular ones keep no group.  Returns the number of singular ones.
int BatchLu::order_fresh(const double* dA, double gmin, std::vector<int32_t>& kind, int32_t* status) {
    std::vector<int> fresh;
    for (int i = 0; i < B_; ++i)
        if (group_of_[i] < 0) fresh.push_back(i);
    if (fresh.empty()) return 0;
    const int nnz = acp_[n_];
    std::vector<std::vector<double>> vals(fresh.size());
    if (fresh.size() == 1) {
        vals[0] = instance_values(dA, fresh[0]);
    } else {
        std::vector<double> all((size_t)nnz * B_);
        NSA_HIP(hipMemcpyAsync(all.data(), dA, sizeof(double) * all.size(), hipMemcpyDeviceToHost, stream_));
        NSA_HIP(hipStreamSynchronize(stream_));
        for (size_t t = 0; t < fresh.size(); ++t) {
            vals[t].resize(nnz);
            for (int k = 0; k < nnz; ++k) vals[t][k] = all[(size_t)k * B_ + fresh[t]];
        }
    }
    std::vector<LuStructure> st(fresh.size());
    std::vector<std::unique_ptr<Sparse13>> s13(fresh.size());
    std::vector<char> sing(fresh.size(), 0);
    std::vector<LuForm> form(fresh.size());
    parallel_for((int)fresh.size(), [&](int t) {
        LuForm f = iform_[fresh[t]];
        bool sg = cpu_order(f, vals[t], gmin, nullptr, st[t], s13[t]);
        if (sg && f == FORM_GP && auto_amd_) {
            f = FORM_S13;
            sg = cpu_order(f, vals[t], gmin, nullptr, st[t], s13[t]);
        }
        sing[t] = sg ? 1 : 0;
        form[t] = f;
    });
    n_cpu_pivot += (long)fresh.size();
    int nsing = 0;
    std::vector<int> touched;
    for (size_t t = 0; t < fresh.size(); ++t) {
        const int i = fresh[t];
        // the auto policy's fallback is decided at an instance's first
        // factorization only (AutoFallbackSolver commits after it)
        iform_[i] = form[t];
        if (sing[t]) {
            status[i] = 1;
            ++nsing;
            kind[i] = KIND_SKIP;
            continue;
        }
        int g = find_group(st[t]);
        if (g < 0) {
            Group G;
            std::vector<int32_t> one{i};
            G.lu = std::make_unique<GpuLu>(st[t], 1, stream_, &one, B_);
            G.s13 = std::move(s13[t]);
            groups_.push_back(std::move(G));
            g = (int)groups_.size() - 1;
        }
        groups_[g].inst.push_back(i);
        group_of_[i] = g;
        kind[i] = KIND_TRUSTED;  // its order was just computed from its own values
        if (std::find(touched.begin(), touched.end(), g) == touched.end()) touched.push_back(g);
    }
    auto_amd_ = false;
    for (int g : touched) rebuild_group(g);
    return nsing;
}

int BatchLu::factor(const double* dA, double gmin, const int32_t* numeric_mask, int32_t* status) {
    if (acp_.empty()) throw Error(NSA_ERR_LOGIC, "BatchLu: symbolic() not called");
    if (equilibrating()) dA = equilibrate(dA, gmin);
    std::vector<int32_t> kind(B_, KIND_VERIFY);
    if (numeric_mask)
        for (int i = 0; i < B_; ++i) kind[i] = numeric_mask[i] ? KIND_VERIFY : KIND_REFACTOR;
    std::fill(status, status + B_, 0);
    // instances without a factorization: a fresh solver's, or one whose
    // last numeric() was singular
    int nsing = order_fresh(dA, gmin, kind, status);
    if (groups_.empty()) return nsing;
    std::vector<int> all_groups(groups_.size());
    for (size_t g = 0; g < groups_.size(); ++g) all_groups[g] = (int)g;
    std::vector<int32_t> fail;
    run_gpu(dA, gmin, kind, all_groups, fail);
    std::vector<char> changed(groups_.size(), 0);
    std::vector<int32_t> rerun_kind(B_, KIND_SKIP);
    bool moved = false;
    for (int i = 0; i < B_; ++i) {
        const int32_t f = fail[i];
        if (f == FAIL_NONE) continue;
        const int k = kind[i];
        const LuForm fi = iform_[i];
        // Sparse 1.3 refactorize: exact zero pivot is reported as singular
        // (factor.cpp:997, 1025, 1045); newton_solve then asks for numeric().
        if (fi == FORM_S13 && k == KIND_REFACTOR) { status[i] = 1; ++nsing; continue; }
        // A fresh AMD-LU factorization would be singular, or an exact zero
        // on an order just computed for these values.
        if ((fi == FORM_GP && k == KIND_VERIFY && f == FAIL_ZERO_PIVOT) || k == KIND_TRUSTED) {
            status[i] = 1;
            ++nsing;
            continue;
        }
        // numeric() whose order no longer holds, or an AMD-LU replay rejected
        // by its growth check (answered with a full factor_, amd_lu_solver.cpp:291-300):
        // re-pivot this instance on the CPU from its own values.
        const int g = group_of_[i];
        auto vals = instance_values(dA, i);
        LuStructure s;
        std::unique_ptr<Sparse13> st;
        ++n_cpu_pivot;
        if (cpu_order(fi, vals, gmin, groups_[g].s13.get(), s, st)) {
            status[i] = 1;
            ++nsing;
            continue;
        }
        if (groups_[g].lu->structure().same_as(s)) {
            // Same order: the GPU values are the re-pivoting factorization's.
            continue;
        }
        int dst = find_group(s);
        if (dst < 0) {
            Group G;
            std::vector<int32_t> one{i};
            G.lu = std::make_unique<GpuLu>(s, 1, stream_, &one, B_);
            G.s13 = std::move(st);
            groups_.push_back(std::move(G));
            changed.push_back(1);
            dst = (int)groups_.size() - 1;
        }
        auto& src_inst = groups_[g].inst;
        src_inst.erase(std::find(src_inst.begin(), src_inst.end(), i));
        groups_[dst].inst.push_back(i);
        group_of_[i] = dst;
        changed[g] = 1;
        changed[dst] = 1;
        rerun_kind[i] = KIND_TRUSTED;
        ++n_regroup;
        moved = true;
    }
    if (moved) {
        // Drop emptied groups, rebuild the plans whose membership changed and
        // recompute their members: moved instances on their new order, the
        // others exactly as before (same order, same values).
        std::vector<Group> keep;
        std::vector<char> keep_changed;
        for (size_t g = 0; g < groups_.size(); ++g) {
            if (groups_[g].inst.empty()) continue;
            keep.push_back(std::move(groups_[g]));
            keep_changed.push_back(changed[g]);
        }
        groups_ = std::move(keep);
        std::vector<int> redo;
        for (size_t g = 0; g < groups_.size(); ++g) {
            for (int32_t i : groups_[g].inst) group_of_[i] = (int)g;
            if (keep_changed[g]) { rebuild_group((int)g); redo.push_back((int)g); }
        }
        for (int g : redo)
            for (int32_t i : groups_[g].inst)
                if (rerun_kind[i] == KIND_SKIP && !status[i]) rerun_kind[i] = KIND_TRUSTED;
        run_gpu(dA, gmin, rerun_kind, redo, fail);
        for (int i = 0; i < B_; ++i)
            if (rerun_kind[i] != KIND_SKIP && fail[i] != FAIL_NONE && !status[i]) {
                status[i] = 1;
                ++nsing;
            }
    }
    factored_ = true;
    return nsing;
}

int BatchLu::refactorize(const double* dA, double gmin, int32_t* status) {
    // NeoSolver::refactorize before numeric() is a logic_error
    // (neo_solver.cpp:178-180); AmdLuSolver falls through to factor_.
    if (groups_.empty() && form_ == FORM_S13)
        throw Error(NSA_ERR_LOGIC, "refactorize: numeric() not called");
    std::vector<int32_t> mask(B_, 0);
    return factor(dA, gmin, mask.data(), status);
}

void BatchLu::solve(double* db) {
    if (!factored()) throw Error(NSA_ERR_LOGIC, "BatchLu::solve: numeric() not called");
    // equilibrated factors are of D_r A D_c: solve for D_r b, recover x = D_c y
    const bool eq = equilibrating() && d_dr_;
    if (eq) launch_scale_vec(n_, B_, d_dr_, db, stream_);
    if (timing_) NSA_HIP(hipEventRecord(ev_[2], stream_));
    for_groups(all_groups(), [&](int g, hipStream_t st) { groups_[g].lu->solve(db, false, st); });
    if (timing_) { NSA_HIP(hipEventRecord(ev_[3], stream_)); pend_sv_ = true; }
    if (eq) launch_scale_vec(n_, B_, d_dc_, db, stream_);
    ++n_gpu_solve;
}

void BatchLu::solve_t(double* db) {
    if (!factored()) throw Error(NSA_ERR_LOGIC, "BatchLu::solve_t: numeric() not called");
    // A^T x = b from the factors of M = D_r A D_c: x = D_r M^-T (D_c b)
    const bool eq = equilibrating() && d_dr_;
    if (eq) launch_scale_vec(n_, B_, d_dc_, db, stream_);
    for_groups(all_groups(), [&](int g, hipStream_t st) { groups_[g].lu->solve(db, true, st); });
    if (eq) launch_scale_vec(n_, B_, d_dr_, db, stream_);
}

void BatchLu::set_timing(bool on) {
    if (on && !ev_[0])
        for (auto& e : ev_) NSA_HIP(hipEventCreate(&e));
    timing_ = on;
    ms_refactor = ms_solve = 0.0;
    pend_rf_ = pend_sv_ = false;
}

// Reads the elapsed time of the last recorded refactor / solve once the
// stream has passed them (callers synchronize first).
void BatchLu::collect_timing() {
    float ms = 0.f;
    if (pend_rf_) {
        NSA_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        ms_refactor += ms;
        pend_rf_ = false;
    }
    if (pend_sv_) {
        NSA_HIP(hipEventElapsedTime(&ms, ev_[2], ev_[3]));
        ms_solve += ms;
        pend_sv_ = false;
    }
}

}  // namespace nsa
