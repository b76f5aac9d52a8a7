// Parameter access and the discrete adjoints of the batched engine.
//
// DC:  at a converged operating point F(x, p) = 0, dL/dp = -lambda^T dF/dp
//      with J^T lambda = dL/dx; J's factors are those of the last Newton
//      iteration (newton_solve returns the iterate that iteration loaded,
//      newton.cpp:378-381), so no extra factorization is needed.
//
// Transient (fixed step, backward Euler then trapezoidal): step k solves
//   G_k(x_k) = f(x_k, p) + P i_k - b(t_k) = 0,
//   i_k = a0 C P^T x_k + a1 C P^T x_{k-1} - r i_{k-1}
// (capacitor companion model, capacitor.cpp:69-85; r = xmu_ratio at order 2,
// 0 at order 1).  Sweeping k = N..1 with adjoints xbar (node voltages) and
// ibar (capacitor currents):
//   xbar_k += a0 P C ibar_k;  lambda_k = J_k^-T xbar_k;
//   dL/dR += (P_R^T lambda_k)(P_R^T x_k) / R^2;
//   dL/dC += (ibar_k - P^T lambda_k) (a0 P^T x_k + a1 P^T x_{k-1});
//   xbar_{k-1} = a1 P C (ibar_k - P^T lambda_k) + dL/dx_{k-1};
//   ibar_{k-1} = r (P^T lambda_k - ibar_k);
// and finally the DC step at x_0.  Device families must be charge-free
// apart from linear capacitors (diodes without CJO/CJSW/TT, MOS1 without
// gate or junction capacitance); others are refused.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "common.hpp"
#include "engine.hpp"
#include "gpu_adjoint.hpp"

namespace nsa {

namespace {

template <class T>
T* alloc(std::vector<void*>& allocs, size_t count) {
    T* p = nullptr;
    NSA_HIP(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)));
    allocs.push_back(p);
    return p;
}

template <class T>
T* upload(std::vector<void*>& allocs, const std::vector<T>& v) {
    T* p = alloc<T>(allocs, v.size());
    if (!v.empty()) NSA_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

}  // namespace

double* BatchEngine::family_values_ptr(DevKind k) const {
    switch (k) {
        case DK_R: return d_r_ohms_;
        case DK_C: return d_c_val_;
        case DK_V: return d_v_val_;
        case DK_I: return d_i_val_;
        default: throw Error(NSA_ERR_ARGUMENT, "family values: only R, C, V and I carry a value");
    }
}

void BatchEngine::set_family_values(DevKind k, const double* h_vals) {
    double* dst = family_values_ptr(k);
    const size_t cnt = (size_t)family_count(k) * B_;
    if (!cnt) return;
    std::vector<double> v(h_vals, h_vals + cnt);
    if (k == DK_R) {
        for (auto& x : v) if (!std::isfinite(x) || x == 0.0) x = 1e-3;  // as set_device_values
        r_dirty_ = true;
    }
    NSA_HIP(hipMemcpyAsync(dst, v.data(), sizeof(double) * cnt, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void BatchEngine::get_family_values(DevKind k, double* h_vals) const {
    const size_t cnt = (size_t)family_count(k) * B_;
    NSA_HIP(hipStreamSynchronize(stream_));  // values are written on the engine's stream
    if (cnt) NSA_HIP(hipMemcpy(h_vals, family_values_ptr(k), sizeof(double) * cnt, hipMemcpyDeviceToHost));
}

void BatchEngine::ensure_adjoint() {
    if (adj_ready_) return;
    const auto& devs = ckt_.devices();
    std::vector<int32_t> rn, vb, in;
    for (int d : ckt_.family(DK_R)) { rn.push_back(devs[d].node[0]); rn.push_back(devs[d].node[1]); }
    for (int d : ckt_.family(DK_V)) vb.push_back(devs[d].branch);
    for (int d : ckt_.family(DK_I)) { in.push_back(devs[d].node[0]); in.push_back(devs[d].node[1]); }
    // node -> capacitor incidence, capacitors in family order
    std::vector<std::vector<int32_t>> inc(n_);
    const auto& fc = ckt_.family(DK_C);
    for (size_t c = 0; c < fc.size(); ++c) {
        const Device& dv = devs[fc[c]];
        if (dv.node[0] >= 0) inc[dv.node[0]].push_back((int32_t)(2 * c));
        if (dv.node[1] >= 0) inc[dv.node[1]].push_back((int32_t)(2 * c + 1));
    }
    std::vector<int32_t> ptr(n_ + 1, 0), ent;
    for (int i = 0; i < n_; ++i) {
        ptr[i + 1] = ptr[i] + (int32_t)inc[i].size();
        ent.insert(ent.end(), inc[i].begin(), inc[i].end());
    }
    d_r_nodes_ = upload(allocs_, rn);
    d_v_branch_ = upload(allocs_, vb);
    d_i_nodes_ = upload(allocs_, in);
    d_cinc_ptr_ = upload(allocs_, ptr);
    d_cinc_ent_ = upload(allocs_, ent);
    const size_t nB = (size_t)n_ * B_;
    d_lam_ = alloc<double>(allocs_, nB);
    d_carry_ = alloc<double>(allocs_, nB);
    // at least one row per family: the clears and copies below use max(1, count)
    auto rows = [](size_t k) { return std::max<size_t>(1, k); };
    d_ibar_ = alloc<double>(allocs_, rows(fc.size()) * B_);
    d_q_ = alloc<double>(allocs_, rows(fc.size()) * B_);
    d_gR_ = alloc<double>(allocs_, rows(family_count(DK_R)) * B_);
    d_gC_ = alloc<double>(allocs_, rows(fc.size()) * B_);
    d_gV_ = alloc<double>(allocs_, rows(family_count(DK_V)) * B_);
    d_gI_ = alloc<double>(allocs_, rows(family_count(DK_I)) * B_);
    adj_ready_ = true;
}

void BatchEngine::dc_adjoint(const double* h_dLdx, double* h_lambda, double* h_gR, double* h_gC,
                             double* h_gV, double* h_gI) {
    if (!lu_->factored())
        throw Error(NSA_ERR_LOGIC, "dc_adjoint: no factors (run a Newton solve first)");
    ensure_adjoint();
    const size_t nB = (size_t)n_ * B_;
    NSA_HIP(hipMemcpyAsync(d_lam_, h_dLdx, sizeof(double) * nB, hipMemcpyHostToDevice, stream_));
    lu_->solve_t(d_lam_);
    const int nR = family_count(DK_R), nC = family_count(DK_C), nV = family_count(DK_V),
              nI = family_count(DK_I);
    NSA_HIP(hipMemsetAsync(d_gR_, 0, sizeof(double) * std::max(1, nR) * B_, stream_));
    NSA_HIP(hipMemsetAsync(d_gV_, 0, sizeof(double) * std::max(1, nV) * B_, stream_));
    NSA_HIP(hipMemsetAsync(d_gI_, 0, sizeof(double) * std::max(1, nI) * B_, stream_));
    launch_grad_res(nR, d_r_nodes_, d_r_ohms_, d_x_, d_lam_, d_gR_, B_, stream_);
    launch_grad_src(nV, d_v_branch_, nI, d_i_nodes_, d_ctx_, d_lam_, d_gV_, d_gI_, B_, stream_);
    NSA_HIP(hipStreamSynchronize(stream_));
    if (h_lambda) NSA_HIP(hipMemcpy(h_lambda, d_lam_, sizeof(double) * nB, hipMemcpyDeviceToHost));
    NSA_HIP(hipStreamSynchronize(stream_));
    if (h_gR && nR) NSA_HIP(hipMemcpy(h_gR, d_gR_, sizeof(double) * nR * B_, hipMemcpyDeviceToHost));
    if (h_gC && nC) std::memset(h_gC, 0, sizeof(double) * nC * B_);  // capacitors are open in DC
    if (h_gV && nV) NSA_HIP(hipMemcpy(h_gV, d_gV_, sizeof(double) * nV * B_, hipMemcpyDeviceToHost));
    if (h_gI && nI) NSA_HIP(hipMemcpy(h_gI, d_gI_, sizeof(double) * nI * B_, hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------
// fixed-step transient

// Integrator context of step k (fill_integrator_context,
// transient.cpp:398-447, at a constant step): k = 0 is the TRANOP operating
// point, k = 1 backward Euler, later steps order 2 of the options' method.
DevCtx BatchEngine::step_ctx(int k, double dt) const {
    DevCtx c{};
    c.method = opt_.method;
    c.src_fact = opt_.src_fact;
    c.delta = dt;
    c.delta_old1 = dt;
    if (k == 0) {
        c.mode = MODETRANOP | MODEINITJCT;
        c.order = 1;
        c.xmu_ratio = 1.0;
        c.delta = c.delta_old1 = 0.0;
        return c;
    }
    c.time = k * dt;
    c.mode = MODETRAN | (k == 1 ? MODEINITTRAN : MODEINITPRED);
    c.order = (k == 1) ? 1 : 2;
    if (c.order == 1) {
        c.ag0 = 1.0 / dt;
        c.ag1 = -1.0 / dt;
        c.ag2 = 0.0;
        c.xmu_ratio = 0.0;
    } else if (opt_.method == 0) {
        const double one_minus_xmu = 1.0 - opt_.xmu;
        c.ag0 = 1.0 / (dt * one_minus_xmu);
        c.ag1 = -1.0 / (dt * one_minus_xmu);
        c.ag2 = 0.0;
        c.xmu_ratio = opt_.xmu / one_minus_xmu;
    } else {
        const double h_old = dt, sum = dt + h_old;
        c.ag0 = (2.0 * dt + h_old) / (dt * sum);
        c.ag1 = -(dt + h_old) / (dt * h_old);
        c.ag2 = dt / (h_old * sum);
        c.xmu_ratio = 0.0;
    }
    return c;
}

void BatchEngine::free_tran() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : tr_allocs_) (void)hipFree(p);
    tr_allocs_.clear();
    d_probes_ = nullptr;
    d_probe_vals_ = d_ckpt_ = d_segx_ = d_dlp_ = nullptr;
    tr_steps_ = 0;
}

void BatchEngine::save_checkpoint(int slot) {
    double* base = d_ckpt_ + (size_t)slot * ckpt_stride_;
    const size_t nB = (size_t)n_ * B_, sB = (size_t)ns_ * B_;
    NSA_HIP(hipMemcpyAsync(base, d_x_, sizeof(double) * nB, hipMemcpyDeviceToDevice, stream_));
    for (int k = 0; k < 4; ++k)
        if (sB)
            NSA_HIP(hipMemcpyAsync(base + nB + k * sB, d_s_[k], sizeof(double) * sB,
                                   hipMemcpyDeviceToDevice, stream_));
    const size_t aB = (size_t)mos_.count * kMos1Aux * B_;
    if (aB)
        NSA_HIP(hipMemcpyAsync(base + nB + 4 * sB, mos_.aux, sizeof(double) * aB,
                               hipMemcpyDeviceToDevice, stream_));
}

void BatchEngine::restore_checkpoint(int slot) {
    const double* base = d_ckpt_ + (size_t)slot * ckpt_stride_;
    const size_t nB = (size_t)n_ * B_, sB = (size_t)ns_ * B_;
    NSA_HIP(hipMemcpyAsync(d_x_, base, sizeof(double) * nB, hipMemcpyDeviceToDevice, stream_));
    for (int k = 0; k < 4; ++k)
        if (sB)
            NSA_HIP(hipMemcpyAsync(d_s_[k], base + nB + k * sB, sizeof(double) * sB,
                                   hipMemcpyDeviceToDevice, stream_));
    const size_t aB = (size_t)mos_.count * kMos1Aux * B_;
    if (aB)
        NSA_HIP(hipMemcpyAsync(mos_.aux, base + nB + 4 * sB, sizeof(double) * aB,
                               hipMemcpyDeviceToDevice, stream_));
}

void BatchEngine::tran_forward(int nsteps, double dt, const int32_t* probes, int nprobe,
                               int ckpt_every, double* h_probes) {
    if (nsteps < 1 || !(dt > 0.0) || ckpt_every < 1 || nprobe < 0)
        throw Error(NSA_ERR_ARGUMENT, "tran_forward: need nsteps >= 1, dt > 0, ckpt_every >= 1");
    for (int k = 0; k < nprobe; ++k)
        if (probes[k] < 0 || probes[k] >= n_)
            throw Error(NSA_ERR_ARGUMENT, "tran_forward: probe index out of range");
    ensure_adjoint();
    free_tran();
    const auto t0 = clk::now();
    tr_steps_ = nsteps;
    tr_dt_ = dt;
    tr_ckpt_ = ckpt_every;
    tr_nprobe_ = nprobe;
    tr_ctx_.resize(nsteps + 1);
    for (int k = 0; k <= nsteps; ++k) tr_ctx_[k] = step_ctx(k, dt);
    const size_t nB = (size_t)n_ * B_, sB = (size_t)ns_ * B_;
    ckpt_stride_ = nB + 4 * sB + (size_t)mos_.count * kMos1Aux * B_;
    const int nckpt = (nsteps + ckpt_every - 1) / ckpt_every;
    d_probes_ = upload(tr_allocs_, std::vector<int32_t>(probes, probes + nprobe));
    d_probe_vals_ = alloc<double>(tr_allocs_, (size_t)(nsteps + 1) * nprobe * B_);
    d_ckpt_ = alloc<double>(tr_allocs_, (size_t)nckpt * ckpt_stride_);
    d_segx_ = alloc<double>(tr_allocs_, (size_t)(ckpt_every + 1) * nB);
    d_dlp_ = alloc<double>(tr_allocs_, (size_t)(nsteps + 1) * nprobe * B_);
    // operating point (transient.cpp:710-760), then a fresh solver and
    // state1..3 = state0 (:791-802)
    set_ctx(tr_ctx_[0]);
    NewtonOutcome r = newton(opt_.itl1);
    for (int i = 0; i < B_; ++i)
        if (!r.converged[i])
            throw Error(NSA_ERR_RUNTIME, "tran_forward: operating point did not converge");
    reset_solver();
    seed_states_from_state0();
    launch_probe_gather(nprobe, d_probes_, d_x_, d_probe_vals_, B_, stream_);
    save_checkpoint(0);
    for (int k = 1; k <= nsteps; ++k) {
        set_ctx(tr_ctx_[k]);
        r = newton(opt_.itl4);
        for (int i = 0; i < B_; ++i)
            if (!r.converged[i])
                throw Error(NSA_ERR_RUNTIME,
                            "tran_forward: Newton did not converge at step " + std::to_string(k));
        rotate_states();
        launch_probe_gather(nprobe, d_probes_, d_x_, d_probe_vals_ + (size_t)k * nprobe * B_, B_,
                            stream_);
        if (k % ckpt_every == 0 && k < nsteps) save_checkpoint(k / ckpt_every);
    }
    NSA_HIP(hipStreamSynchronize(stream_));
    if (h_probes)
        NSA_HIP(hipMemcpy(h_probes, d_probe_vals_, sizeof(double) * (nsteps + 1) * nprobe * B_,
                          hipMemcpyDeviceToHost));
    ms_forward = ms_since(t0);
}

// Factor the matrix in d_A_ with the current pivot order, re-pivoting the
// instances whose replay is rejected (as a singular refactorize does in
// newton.cpp:229-234).
void BatchEngine::factor_current() {
    std::vector<int32_t> status(B_, 0);
    int nsing = lu_->factored() ? lu_->refactorize(d_A_, opt_.diag_gmin, status.data())
                                : lu_->numeric(d_A_, opt_.diag_gmin, status.data());
    if (nsing) {
        std::vector<int32_t> mask(B_);
        for (int i = 0; i < B_; ++i) mask[i] = status[i] ? 1 : 0;
        nsing = lu_->factor(d_A_, opt_.diag_gmin, mask.data(), status.data());
        if (nsing) throw Error(NSA_ERR_RUNTIME, "adjoint: singular Jacobian");
    }
}

// Device load at d_x with the step's integration coefficients and no
// limiting, then factorization: J = dG/dx at that point.
void BatchEngine::jacobian_at(const double* d_x, const DevCtx& c) {
    DevCtx a = c;
    a.mode = (c.mode & ~INITF_MASK) | MODEINITFLOAT | MODEADJOINT;
    set_ctx(a);
    if (d_x != d_x_)
        NSA_HIP(hipMemcpyAsync(d_x_, d_x, sizeof(double) * n_ * B_, hipMemcpyDeviceToDevice,
                               stream_));
    launch_fill_i32(d_devnc_, 0, B_, stream_);
    load(nullptr);
    factor_current();
}

void BatchEngine::tran_adjoint(const double* h_dLdprobe, double* h_gR, double* h_gC) {
    if (!tr_steps_) throw Error(NSA_ERR_LOGIC, "tran_adjoint: no tran_forward record");
    if (opt_.method != 0)
        throw Error(NSA_ERR_UNSUPPORTED, "tran_adjoint: only backward Euler / trapezoidal steps");
    for (int d : ckt_.family(DK_D)) {
        const DioParam p = diode_param(ckt_.models().at(ckt_.devices()[d].model),
                                       ckt_.devices()[d], DeviceEnv{});
        if (p.tJctCap != 0.0 || p.tJctSWCap != 0.0 || p.tTransitTime != 0.0)
            throw Error(NSA_ERR_UNSUPPORTED, "tran_adjoint: diodes with charge storage");
    }
    if (!ckt_.family(DK_B4).empty())
        throw Error(NSA_ERR_UNSUPPORTED, "tran_adjoint: BSIM4v7 devices (charge storage)");
    for (int d : ckt_.family(DK_M1)) {
        const Mos1Param p = mos1_param(ckt_.models().at(ckt_.devices()[d].model),
                                       ckt_.devices()[d], DeviceEnv{});
        if (p.OxideCap != 0.0 || p.GateSourceOverlapCap != 0.0 || p.GateDrainOverlapCap != 0.0 ||
            p.GateBulkOverlapCap != 0.0 || p.Cbd != 0.0 || p.Cbs != 0.0 || p.Cbdsw != 0.0 ||
            p.Cbssw != 0.0)
            throw Error(NSA_ERR_UNSUPPORTED, "tran_adjoint: MOS1 with charge storage");
    }
    const auto t0 = clk::now();
    const int N = tr_steps_, K = tr_ckpt_, np = tr_nprobe_;
    const size_t nB = (size_t)n_ * B_;
    const int nR = family_count(DK_R), nC = family_count(DK_C);
    NSA_HIP(hipMemcpyAsync(d_dlp_, h_dLdprobe, sizeof(double) * (N + 1) * np * B_,
                           hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipMemsetAsync(d_gR_, 0, sizeof(double) * std::max(1, nR) * B_, stream_));
    NSA_HIP(hipMemsetAsync(d_gC_, 0, sizeof(double) * std::max(1, nC) * B_, stream_));
    NSA_HIP(hipMemsetAsync(d_ibar_, 0, sizeof(double) * std::max(1, nC) * B_, stream_));
    NSA_HIP(hipMemsetAsync(d_carry_, 0, sizeof(double) * nB, stream_));
    const int nseg = (N + K - 1) / K;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int c = seg * K, e = std::min(N, c + K);
        // recompute the segment forward from its checkpoint
        restore_checkpoint(seg);
        NSA_HIP(hipMemcpyAsync(d_segx_, d_x_, sizeof(double) * nB, hipMemcpyDeviceToDevice,
                               stream_));
        for (int k = c + 1; k <= e; ++k) {
            set_ctx(tr_ctx_[k]);
            NewtonOutcome r = newton(opt_.itl4);
            for (int i = 0; i < B_; ++i)
                if (!r.converged[i])
                    throw Error(NSA_ERR_RUNTIME, "tran_adjoint: recomputed step did not converge");
            rotate_states();
            NSA_HIP(hipMemcpyAsync(d_segx_ + (size_t)(k - c) * nB, d_x_, sizeof(double) * nB,
                                   hipMemcpyDeviceToDevice, stream_));
        }
        // backward sweep over the segment
        for (int k = e; k >= c + 1; --k) {
            const double* xk = d_segx_ + (size_t)(k - c) * nB;
            const double* xkm1 = d_segx_ + (size_t)(k - c - 1) * nB;
            // xbar_k = carry + dL/dx_k + a0 P C ibar_k
            NSA_HIP(hipMemcpyAsync(d_lam_, d_carry_, sizeof(double) * nB,
                                   hipMemcpyDeviceToDevice, stream_));
            launch_probe_scatter(np, d_probes_, d_dlp_ + (size_t)k * np * B_, d_lam_, B_, stream_);
            jacobian_at(xk, tr_ctx_[k]);  // also sets this step's coefficients
            launch_cap_gather(n_, d_cinc_ptr_, d_cinc_ent_, d_c_val_, d_ibar_, d_ctx_, 0, 1, 1,
                              d_lam_, B_, stream_);
            lu_->solve_t(d_lam_);
            launch_grad_res(nR, d_r_nodes_, d_r_ohms_, xk, d_lam_, d_gR_, B_, stream_);
            launch_cap_back(nC, cap_.nodes, d_c_val_, d_ctx_, xk, xkm1, d_lam_, d_ibar_, d_q_,
                            d_gC_, B_, stream_);
            launch_cap_gather(n_, d_cinc_ptr_, d_cinc_ent_, d_c_val_, d_q_, d_ctx_, 1, 0, 0,
                              d_carry_, B_, stream_);
        }
    }
    // operating point: G_0(x_0) = 0 with the capacitors open
    restore_checkpoint(0);
    NSA_HIP(hipMemcpyAsync(d_lam_, d_carry_, sizeof(double) * nB, hipMemcpyDeviceToDevice,
                           stream_));
    launch_probe_scatter(np, d_probes_, d_dlp_, d_lam_, B_, stream_);
    jacobian_at(d_x_, tr_ctx_[0]);
    lu_->solve_t(d_lam_);
    launch_grad_res(nR, d_r_nodes_, d_r_ohms_, d_x_, d_lam_, d_gR_, B_, stream_);
    NSA_HIP(hipStreamSynchronize(stream_));
    if (h_gR && nR) NSA_HIP(hipMemcpy(h_gR, d_gR_, sizeof(double) * nR * B_, hipMemcpyDeviceToHost));
    if (h_gC && nC) NSA_HIP(hipMemcpy(h_gC, d_gC_, sizeof(double) * nC * B_, hipMemcpyDeviceToHost));
    ms_adjoint = ms_since(t0);
}

}  // namespace nsa
