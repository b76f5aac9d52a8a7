#include "engine.hpp"

#include "bsim4.hpp"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "devices.hpp"

namespace nsa {

namespace {

template <class T>
T* dalloc(std::vector<void*>& allocs, size_t count) {
    T* p = nullptr;
    NSA_HIP(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)));
    allocs.push_back(p);
    return p;
}

template <class T>
T* dupload(std::vector<void*>& allocs, const std::vector<T>& v) {
    T* p = dalloc<T>(allocs, v.size());
    if (!v.empty()) NSA_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

// Replicates one value per device into a device-major, instance-fastest array.
std::vector<double> replicate(const std::vector<double>& per_dev, int B) {
    std::vector<double> out(per_dev.size() * (size_t)B);
    for (size_t d = 0; d < per_dev.size(); ++d)
        std::fill(out.begin() + d * B, out.begin() + (d + 1) * B, per_dev[d]);
    return out;
}

}  // namespace

BatchEngine::BatchEngine(const Circuit& ckt, int batch, const Options& opt)
    : ckt_(ckt), B_(batch), opt_(opt) {
    if (!ckt.finalized()) throw Error(NSA_ERR_LOGIC, "BatchEngine: circuit not finalized");
    if (batch <= 0) throw Error(NSA_ERR_ARGUMENT, "BatchEngine: batch must be positive");
    n_ = ckt.num_vars();
    nnz_ = ckt.nnz();
    ns_ = ckt.num_states();
    NSA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const size_t nB = (size_t)n_ * B_;
    d_x_ = dalloc<double>(allocs_, nB);
    d_xold_ = dalloc<double>(allocs_, nB);
    d_rhs_ = dalloc<double>(allocs_, nB);
    d_A_ = dalloc<double>(allocs_, (size_t)nnz_ * B_);
    for (int k = 0; k < 4; ++k) d_s_[k] = dalloc<double>(allocs_, (size_t)ns_ * B_);
    NSA_HIP(hipMemsetAsync(d_x_, 0, nB * sizeof(double), stream_));
    for (int k = 0; k < 4; ++k)
        NSA_HIP(hipMemsetAsync(d_s_[k], 0, (size_t)ns_ * B_ * sizeof(double), stream_));
    d_ctx_ = dalloc<DevCtx>(allocs_, B_);
    d_done_ = dalloc<int32_t>(allocs_, B_);
    d_nonconv_ = dalloc<int32_t>(allocs_, B_);
    d_devnc_ = dalloc<int32_t>(allocs_, B_);
    NSA_HIP(hipMemsetAsync(d_devnc_, 0, sizeof(int32_t) * B_, stream_));
    tol_.gmin = opt_.gmin;
    tol_.reltol = opt_.reltol;
    tol_.abstol = opt_.abstol;
    NSA_HIP(hipHostMalloc((void**)&h_flags_, sizeof(int32_t) * 4 * B_, hipHostMallocDefault));
    NSA_HIP(hipHostMalloc((void**)&h_ctx_pin_, sizeof(DevCtx) * B_, hipHostMallocDefault));
    DevCtx c{};
    c.mode = MODEDCOP | MODEINITJCT;
    c.order = 1;
    c.xmu_ratio = 1.0;
    c.src_fact = opt_.src_fact;
    set_ctx(c);
    build_tables();
    lu_ = std::make_unique<BatchLu>(B_, stream_);
    lu_->symbolic(n_, ckt.col_ptr().data(), ckt.row_idx().data(), ckt.is_linear(), opt_.engine);
    NSA_HIP(hipStreamSynchronize(stream_));
}

BatchEngine::~BatchEngine() {
    free_tran();
    lu_.reset();
    for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : allocs_) (void)hipFree(p);
    if (h_flags_) (void)hipHostFree(h_flags_);
    if (h_ctx_pin_) (void)hipHostFree(h_ctx_pin_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void BatchEngine::upload_ctx() {
    // the previous context copy may still read the staging area
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(h_ctx_pin_, h_ctx_.data(), sizeof(DevCtx) * B_);
    NSA_HIP(hipMemcpyAsync(d_ctx_, h_ctx_pin_, sizeof(DevCtx) * B_, hipMemcpyHostToDevice, stream_));
}

// Callers guarantee that the stream has passed the previous copy from `slot`.
void BatchEngine::upload_flags(int slot, const std::vector<int32_t>& v, int32_t* dst) {
    int32_t* h = h_flags_ + (size_t)slot * B_;
    std::memcpy(h, v.data(), sizeof(int32_t) * B_);
    NSA_HIP(hipMemcpyAsync(dst, h, sizeof(int32_t) * B_, hipMemcpyHostToDevice, stream_));
}

void BatchEngine::build_tables() {
    const auto& devs = ckt_.devices();
    dev_site_.assign(devs.size(), -1);
    dev_rsite_.assign(devs.size(), -1);
    // ---- matrix sites: R block, C block, constant block (V +-1, dead pins)
    int site = 0, rsite = 0;
    const auto& fr = ckt_.family(DK_R);
    const auto& fc = ckt_.family(DK_C);
    const auto& fv = ckt_.family(DK_V);
    const auto& fi = ckt_.family(DK_I);
    nR_ = (int)fr.size();
    r_site_ = site;
    std::vector<double> r_ohms;
    for (int d : fr) { dev_site_[d] = site; site += 4; r_ohms.push_back(devs[d].value); }
    cap_.count = (int)fc.size();
    cap_.site = site;
    cap_.rsite = rsite;
    std::vector<int32_t> cap_nodes, cap_state;
    std::vector<double> cap_val;
    for (int d : fc) {
        dev_site_[d] = site; site += 4;
        dev_rsite_[d] = rsite; rsite += 2;
        cap_nodes.push_back(devs[d].node[0]);
        cap_nodes.push_back(devs[d].node[1]);
        cap_state.push_back(devs[d].state_base);
        cap_val.push_back(devs[d].value);
    }
    // diodes: 7 matrix sites, 2 rhs sites (DIOload stamp order)
    DeviceEnv env;
    env.gmin = opt_.gmin;
    env.reltol = opt_.reltol;
    env.abstol = opt_.abstol;
    env.vntol = opt_.vntol;
    const auto& fd = ckt_.family(DK_D);
    dio_.count = (int)fd.size();
    dio_.site = site;
    dio_.rsite = rsite;
    std::vector<int32_t> dio_nodes, dio_state;
    std::vector<DioParam> dio_par;
    for (int d : fd) {
        dev_site_[d] = site; site += 7;
        dev_rsite_[d] = rsite; rsite += 2;
        dio_nodes.push_back(devs[d].node[0]);
        dio_nodes.push_back(devs[d].node[1]);
        dio_state.push_back(devs[d].state_base);
        dio_par.push_back(diode_param(ckt_.models().at(devs[d].model), devs[d], env));
    }
    // MOS1: 22 matrix sites, 4 rhs sites (MOS1load stamp order)
    const auto& fm = ckt_.family(DK_M1);
    mos_.count = (int)fm.size();
    mos_.site = site;
    mos_.rsite = rsite;
    std::vector<int32_t> mos_nodes, mos_state;
    std::vector<Mos1Param> mos_par;
    for (int d : fm) {
        dev_site_[d] = site; site += 22;
        dev_rsite_[d] = rsite; rsite += 4;
        for (int k = 0; k < 4; ++k) mos_nodes.push_back(devs[d].node[k]);
        mos_state.push_back(devs[d].state_base);
        mos_par.push_back(mos1_param(ckt_.models().at(devs[d].model), devs[d], env));
    }
    // BSIM4v7: 16 matrix sites, 4 rhs sites (BSIM4v7load stamp order)
    const auto& fb = ckt_.family(DK_B4);
    b4_.count = (int)fb.size();
    b4_.site = site;
    b4_.rsite = rsite;
    std::vector<int32_t> b4_nodes, b4_state;
    std::vector<B4Param> b4_par;
    for (int d : fb) {
        dev_site_[d] = site; site += 16;
        dev_rsite_[d] = rsite; rsite += 4;
        for (int k = 0; k < 4; ++k) b4_nodes.push_back(devs[d].node[k]);
        b4_state.push_back(devs[d].state_base);
        b4_par.push_back(bsim4_param(ckt_.models().at(devs[d].model), devs[d], env));
    }
    const_site_ = site;
    std::vector<double> const_vals;
    for (int d : fv) {
        dev_site_[d] = site; site += 4;
        const_vals.insert(const_vals.end(), {1.0, -1.0, 1.0, -1.0});
    }
    const double pin = (opt_.gmin > 0.0) ? opt_.gmin : 1e-12;  // newton.cpp:80
    std::vector<int32_t> pin_site;
    for (size_t k = 0; k < ckt_.dead_nodes().size(); ++k) {
        pin_site.push_back(site++);
        const_vals.push_back(pin);
    }
    n_const_ = (int)const_vals.size();
    n_msites_ = site;
    // ---- rhs sites: C (2, above), V (1), I (2)
    vsrc_.count = (int)fv.size();
    vsrc_.rsite = rsite;
    std::vector<int32_t> vw, vdg, pwl_ptr{0};
    std::vector<double> vdc, vpulse, pwl_t, pwl_v;
    for (int d : fv) {
        dev_rsite_[d] = rsite; rsite += 1;
        vw.push_back(devs[d].wave);
        vdg.push_back(0);
        vdc.push_back(devs[d].value);
        for (int k = 0; k < 7; ++k) vpulse.push_back(devs[d].pulse[k]);
        pwl_t.insert(pwl_t.end(), devs[d].pwl_t.begin(), devs[d].pwl_t.end());
        pwl_v.insert(pwl_v.end(), devs[d].pwl_v.begin(), devs[d].pwl_v.end());
        pwl_ptr.push_back((int32_t)pwl_t.size());
    }
    isrc_.count = (int)fi.size();
    isrc_.rsite = rsite;
    std::vector<double> ival;
    for (int d : fi) { dev_rsite_[d] = rsite; rsite += 2; ival.push_back(devs[d].value); }
    n_rsites_ = rsite;
    // ---- gathers in load order
    std::vector<std::vector<int32_t>> slot_sites(nnz_), rows_d(n_), rows_u(n_);
    for (int d : ckt_.load_order()) {
        const Device& dv = devs[d];
        for (size_t k = 0; k < dv.mat_slot.size(); ++k)
            if (dv.mat_slot[k] >= 0) slot_sites[dv.mat_slot[k]].push_back(dev_site_[d] + (int)k);
        // UCB devices stamp the circuit-global one-based rhs, which
        // newton_solve adds to rhs after all loads (newton.cpp:166-168)
        auto& rows = (dv.kind == DK_D || dv.kind == DK_M1 || dv.kind == DK_B4) ? rows_u : rows_d;
        for (size_t k = 0; k < dv.rhs_row.size(); ++k)
            if (dv.rhs_row[k] >= 0) rows[dv.rhs_row[k]].push_back(dev_rsite_[d] + (int)k);
    }
    for (size_t k = 0; k < ckt_.dead_nodes().size(); ++k) {
        const int nd = ckt_.dead_nodes()[k];
        slot_sites[ckt_.offset(nd, nd)].push_back(pin_site[k]);
    }
    auto csr = [](const std::vector<std::vector<int32_t>>& L, std::vector<int32_t>& ptr,
                  std::vector<int32_t>& idx) {
        ptr.assign(L.size() + 1, 0);
        for (size_t i = 0; i < L.size(); ++i) ptr[i + 1] = ptr[i] + (int32_t)L[i].size();
        idx.clear();
        for (auto& l : L) idx.insert(idx.end(), l.begin(), l.end());
    };
    std::vector<int32_t> p, ix;
    csr(slot_sites, p, ix);
    d_mptr_ = dupload(allocs_, p);
    d_msites_ = dupload(allocs_, ix);
    csr(rows_d, p, ix);
    d_rptr_d_ = dupload(allocs_, p);
    d_rsites_d_ = dupload(allocs_, ix);
    csr(rows_u, p, ix);
    d_rptr_u_ = dupload(allocs_, p);
    d_rsites_u_ = dupload(allocs_, ix);
    // ---- device arrays
    d_cm_ = dalloc<double>(allocs_, (size_t)n_msites_ * B_);
    d_cr_ = dalloc<double>(allocs_, (size_t)n_rsites_ * B_);
    NSA_HIP(hipMemset(d_cm_, 0, (size_t)n_msites_ * B_ * sizeof(double)));
    NSA_HIP(hipMemset(d_cr_, 0, (size_t)n_rsites_ * B_ * sizeof(double)));
    d_r_ohms_ = dupload(allocs_, replicate(r_ohms, B_));
    cap_.nodes = dupload(allocs_, cap_nodes);
    cap_.state = dupload(allocs_, cap_state);
    d_c_val_ = dupload(allocs_, replicate(cap_val, B_));
    cap_.farads = d_c_val_;
    vsrc_.wave = dupload(allocs_, vw);
    vsrc_.dc_given = dupload(allocs_, vdg);
    d_v_val_ = dupload(allocs_, replicate(vdc, B_));
    vsrc_.dc = d_v_val_;
    vsrc_.pulse = dupload(allocs_, vpulse);
    vsrc_.pwl_ptr = dupload(allocs_, pwl_ptr);
    vsrc_.pwl_t = dupload(allocs_, pwl_t);
    vsrc_.pwl_v = dupload(allocs_, pwl_v);
    d_i_val_ = dupload(allocs_, replicate(ival, B_));
    isrc_.dc = d_i_val_;
    d_const_vals_ = dupload(allocs_, const_vals);
    dio_.nodes = dupload(allocs_, dio_nodes);
    dio_.state = dupload(allocs_, dio_state);
    dio_.par = dupload(allocs_, dio_par);
    mos_.nodes = dupload(allocs_, mos_nodes);
    mos_.state = dupload(allocs_, mos_state);
    mos_.par = dupload(allocs_, mos_par);
    const size_t n_aux = (size_t)std::max(1, mos_.count) * kMos1Aux * B_;
    mos_.aux = dalloc<double>(allocs_, n_aux);
    NSA_HIP(hipMemset(mos_.aux, 0, n_aux * sizeof(double)));
    b4_.nodes = dupload(allocs_, b4_nodes);
    b4_.state = dupload(allocs_, b4_state);
    b4_.par = dupload(allocs_, b4_par);
    b4_host_ = b4_par;
    env_ = env;
    const size_t n_b4aux = (size_t)std::max(1, b4_.count) * kB4Aux * B_;
    b4_.aux = dalloc<double>(allocs_, n_b4aux);
    NSA_HIP(hipMemset(b4_.aux, 0, n_b4aux * sizeof(double)));
    // truncation-error rows: the capacitors, the diodes, the MOS1, then the
    // BSIM4v7 devices
    std::vector<int32_t> tbase, tkind;
    for (int d : fc) { trunc_dev_.push_back(d); tbase.push_back(devs[d].state_base); tkind.push_back(0); }
    for (int d : fd) { trunc_dev_.push_back(d); tbase.push_back(devs[d].state_base); tkind.push_back(2); }
    for (int d : fm) { trunc_dev_.push_back(d); tbase.push_back(devs[d].state_base); tkind.push_back(1); }
    for (int d : fb) { trunc_dev_.push_back(d); tbase.push_back(devs[d].state_base); tkind.push_back(3); }
    trunc_.ndev = (int)trunc_dev_.size();
    trunc_.base = dupload(allocs_, tbase);
    trunc_.kind = dupload(allocs_, tkind);
    d_tmin_ = dalloc<unsigned long long>(allocs_, (size_t)B_);
    d_tdev_ = dalloc<double>(allocs_, (size_t)std::max(1, trunc_.ndev) * B_);
    // constant sites are written once
    launch_res_sites(nR_, d_r_ohms_, d_cm_, r_site_, B_, stream_);
    launch_const_sites(n_const_, d_const_vals_, d_cm_, const_site_, B_, stream_);
}

void BatchEngine::set_device_values(int dev, const double* vals) {
    const Device& d = ckt_.devices().at(dev);
    double* dst = nullptr;
    switch (d.kind) {
        case DK_R: dst = d_r_ohms_; break;
        case DK_C: dst = d_c_val_; break;
        case DK_V: dst = d_v_val_; break;
        case DK_I: dst = d_i_val_; break;
        default: throw Error(NSA_ERR_ARGUMENT, "set_device_values: unsupported device kind");
    }
    std::vector<double> v(vals, vals + B_);
    if (d.kind == DK_R) {
        for (auto& x : v) if (!std::isfinite(x) || x == 0.0) x = 1e-3;
        r_dirty_ = true;  // conductance sites are recomputed before the next load
    }
    NSA_HIP(hipMemcpyAsync(dst + (size_t)d.family_index * B_, v.data(), sizeof(double) * B_,
                           hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void BatchEngine::set_model_values(int dev, const std::string& param, const double* vals) {
    const auto& devs = ckt_.devices();
    if (dev < 0 || dev >= (int)devs.size() || devs[dev].kind != DK_B4)
        throw Error(NSA_ERR_ARGUMENT, "set_model_values: not a BSIM4v7 device");
    const Device& d = devs[dev];
    const size_t per_dev = (size_t)kB4Fields * B_;
    if (!d_b4_inst_) {
        // switch the table to per-instance records, each a copy of the shared one
        d_b4_inst_ = dalloc<double>(allocs_, per_dev * b4_.count);
        std::vector<double> all(per_dev * b4_.count);
        for (int k = 0; k < b4_.count; ++k) {
            const double* f = reinterpret_cast<const double*>(&b4_host_[k]);
            for (int j = 0; j < kB4Fields; ++j)
                for (int i = 0; i < B_; ++i) all[k * per_dev + (size_t)j * B_ + i] = f[j];
        }
        NSA_HIP(hipMemcpy(d_b4_inst_, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
        b4_over_.assign(b4_.count, std::vector<std::map<std::string, double>>(B_));
        b4_.par_inst = d_b4_inst_;
    }
    std::string key = param;
    for (auto& ch : key) ch = (char)std::tolower((unsigned char)ch);
    // every instance's card is built from a copy of its overrides; they are
    // kept only when no instance's card is refused (an unsupported parameter
    // or value), so a refused call leaves the device as it was
    std::vector<double> col(per_dev);
    std::vector<std::map<std::string, double>> over = b4_over_[d.family_index];
    for (int i = 0; i < B_; ++i) {
        over[i][key] = vals[i];
        ModelCard card = ckt_.models().at(d.model);
        for (const auto& kv : over[i]) card.p[kv.first] = kv.second;
        const B4Param p = bsim4_param(card, d, env_);
        const double* f = reinterpret_cast<const double*>(&p);
        for (int j = 0; j < kB4Fields; ++j) col[(size_t)j * B_ + i] = f[j];
    }
    NSA_HIP(hipMemcpy(d_b4_inst_ + (size_t)d.family_index * per_dev, col.data(),
                      per_dev * sizeof(double), hipMemcpyHostToDevice));
    b4_over_[d.family_index] = std::move(over);
}

void BatchEngine::set_ctx(const DevCtx& c) {
    set_ctx(std::vector<DevCtx>(B_, c));
}

void BatchEngine::set_ctx(const std::vector<DevCtx>& c) {
    if ((int)c.size() != B_) throw Error(NSA_ERR_ARGUMENT, "set_ctx: need one context per instance");
    h_ctx_ = c;
    upload_ctx();
}

void BatchEngine::load(const int32_t* d_done) {
    if (r_dirty_) {
        launch_res_sites(nR_, d_r_ohms_, d_cm_, r_site_, B_, stream_);
        r_dirty_ = false;
    }
    if (timing_) NSA_HIP(hipEventRecord(ev_[0], stream_));
    launch_cap(cap_, d_ctx_, d_x_, d_s_[0], d_s_[1], d_s_[2], d_done, d_cm_, d_cr_, B_, stream_);
    launch_vsrc(vsrc_, d_ctx_, d_done, d_cr_, B_, stream_);
    launch_isrc(isrc_, d_ctx_, d_done, d_cr_, B_, stream_);
    launch_diode(dio_, d_ctx_, d_x_, d_s_[0], d_s_[1], d_s_[2], d_done, d_cm_, d_cr_, d_devnc_,
                 tol_, B_, stream_);
    launch_mos1(mos_, d_ctx_, d_x_, d_s_[0], d_s_[1], d_s_[2], d_done, d_cm_, d_cr_, d_devnc_,
                tol_, B_, stream_);
    if (timing_ && b4_.count) NSA_HIP(hipEventRecord(ev_[2], stream_));
    launch_bsim4(b4_, d_ctx_, d_x_, d_s_[0], d_s_[1], d_s_[2], d_done, d_cm_, d_cr_, d_devnc_,
                tol_, B_, stream_);
    if (timing_ && b4_.count) NSA_HIP(hipEventRecord(ev_[3], stream_));
    launch_gather(nnz_, d_mptr_, d_msites_, nullptr, nullptr, d_cm_, nullptr, d_A_, B_, stream_);
    launch_gather(n_, d_rptr_d_, d_rsites_d_, d_rptr_u_, d_rsites_u_, d_cr_, d_cr_, d_rhs_, B_,
                  stream_);
    if (timing_) {
        NSA_HIP(hipEventRecord(ev_[1], stream_));
        NSA_HIP(hipEventSynchronize(ev_[1]));
        float ms = 0.f;
        NSA_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        ms_load_gpu += ms;
        ++n_loads;
        if (b4_.count) {
            NSA_HIP(hipEventElapsedTime(&ms, ev_[2], ev_[3]));
            ms_bsim4_gpu += ms;
        }
    }
}

void BatchEngine::set_timing(bool on) {
    if (on && !ev_[0])
        for (auto& e : ev_) NSA_HIP(hipEventCreate(&e));
    timing_ = on;
    ms_load_gpu = 0.0;
    ms_bsim4_gpu = 0.0;
    n_loads = 0;
    lu_->set_timing(on);
}

void BatchEngine::reset_solver() {
    lu_ = std::make_unique<BatchLu>(B_, stream_);
    lu_->symbolic(n_, ckt_.col_ptr().data(), ckt_.row_idx().data(), ckt_.is_linear(), opt_.engine);
    lu_->set_timing(timing_);
}

void BatchEngine::load_only() {
    load(nullptr);
    NSA_HIP(hipStreamSynchronize(stream_));
}

NewtonOutcome BatchEngine::newton(int itl) {
    NewtonOutcome out;
    out.converged.assign(B_, 0);
    out.iterations.assign(B_, 0);
    std::vector<int32_t> done(B_, 0), nonconv(B_, 0), status(B_, 0), numeric_mask(B_, 0);
    std::vector<int32_t> force_numeric(B_, 0), always_reorder(B_, 0), skip(B_, 0);
    std::vector<int32_t> saved_mode(B_);
    for (int i = 0; i < B_; ++i) saved_mode[i] = h_ctx_[i].mode;
    std::vector<DevCtx> ctx = h_ctx_;
    NSA_HIP(hipMemsetAsync(d_done_, 0, sizeof(int32_t) * B_, stream_));
    const int num_nodes = ckt_.num_nodes();
    ms_load = ms_factor = ms_solve = 0;
    using clk = std::chrono::steady_clock;
    for (int iter = 0; iter < itl; ++iter) {
        int active = 0;
        for (int i = 0; i < B_; ++i) active += !done[i];
        if (!active) break;
        auto t0 = clk::now();
        // old_solution = solution (newton.cpp:138)
        NSA_HIP(hipMemcpyAsync(d_xold_, d_x_, sizeof(double) * n_ * B_, hipMemcpyDeviceToDevice,
                               stream_));
        launch_fill_i32(d_devnc_, 0, B_, stream_);
        load(d_done_);
        if (time_phases) NSA_HIP(hipStreamSynchronize(stream_));
        auto t1 = clk::now();
        // numeric() when forced or in MODEINITJCT, else refactorize()
        // (newton.cpp:210-243)
        for (int i = 0; i < B_; ++i) {
            if (done[i]) { numeric_mask[i] = 0; continue; }
            if (ctx[i].mode & MODEINITJCT) always_reorder[i] = 1;
            numeric_mask[i] = (force_numeric[i] || always_reorder[i]) ? 1 : 0;
        }
        const bool any_num = std::any_of(numeric_mask.begin(), numeric_mask.end(),
                                         [](int32_t v) { return v != 0; });
        int nsing;
        if (!lu_->factored() || any_num)
            nsing = lu_->factor(d_A_, opt_.diag_gmin, numeric_mask.data(), status.data());
        else
            nsing = lu_->refactorize(d_A_, opt_.diag_gmin, status.data());
        auto t2 = clk::now();
        for (int i = 0; i < B_; ++i) {
            skip[i] = 0;
            if (done[i] || !nsing || !status[i]) continue;
            if (numeric_mask[i]) {
                // singular numeric(): this Newton attempt fails
                done[i] = 1;
                out.converged[i] = 0;
                out.iterations[i] = iter + 1;
                ctx[i].mode = saved_mode[i];
            } else {
                // singular refactorize(): force numeric and retry (newton.cpp:229-234)
                force_numeric[i] = 1;
                skip[i] = 1;
                out.iterations[i] = iter + 1;
            }
        }
        for (int i = 0; i < B_; ++i)
            if (!done[i] && numeric_mask[i] && !skip[i]) force_numeric[i] = 0;
        // solve; proposed solution replaces the solution of live instances
        lu_->solve(d_rhs_);
        std::vector<int32_t> live(B_), not_live(B_);
        for (int i = 0; i < B_; ++i) {
            live[i] = (!done[i] && !skip[i]) ? 1 : 0;
            not_live[i] = 1 - live[i];
        }
        upload_flags(0, live, d_nonconv_);
        launch_commit(n_, d_rhs_, d_x_, d_nonconv_, 1, B_, stream_);
        // convergence (newton.cpp:267-292); d_done_ doubles as "not live"
        upload_flags(1, not_live, d_done_);
        launch_fill_i32(d_nonconv_, 0, B_, stream_);
        launch_conv(n_, num_nodes, d_x_, d_xold_, opt_.reltol, opt_.vntol, opt_.abstol, d_done_,
                    d_nonconv_, B_, stream_);
        // device convergence (newton.cpp:321-334): the loads' CKTnoncon and
        // MOS1's current test on the proposed solution
        launch_mos1_conv(mos_, d_x_, d_s_[0], d_done_, d_devnc_, tol_, B_, stream_);
        launch_or_i32(d_nonconv_, d_devnc_, B_, stream_);
        NSA_HIP(hipMemcpyAsync(nonconv.data(), d_nonconv_, sizeof(int32_t) * B_,
                               hipMemcpyDeviceToHost, stream_));
        NSA_HIP(hipStreamSynchronize(stream_));
        lu_->collect_timing();
        auto t3 = clk::now();
        std::vector<int32_t> restore(B_, 0);
        bool any_restore = false;
        for (int i = 0; i < B_; ++i) {
            if (done[i] || skip[i]) continue;
            out.iterations[i] = iter + 1;
            bool conv = nonconv[i] == 0;
            if (iter == 0) conv = false;  // newton.cpp:318-319
            const int m = ctx[i].mode;
            // init-mode cascade (newton.cpp:370-411)
            if (m & MODEINITFLOAT) {
                if (conv) { restore[i] = 1; done[i] = 1; out.converged[i] = 1; }
            } else if (m & MODEINITJCT) {
                ctx[i].mode = (m & ~INITF_MASK) | MODEINITFIX;
                force_numeric[i] = 1;
                always_reorder[i] = 0;
            } else if (m & MODEINITFIX) {
                if (conv) ctx[i].mode = (m & ~INITF_MASK) | MODEINITFLOAT;
            } else if (m & MODEINITSMSIG) {
                ctx[i].mode = (m & ~INITF_MASK) | MODEINITFLOAT;
            } else if (m & MODEINITTRAN) {
                ctx[i].mode = (m & ~INITF_MASK) | MODEINITFLOAT;
                force_numeric[i] = 1;
            } else if (m & MODEINITPRED) {
                ctx[i].mode = (m & ~INITF_MASK) | MODEINITFLOAT;
            } else if (conv) {
                restore[i] = 1; done[i] = 1; out.converged[i] = 1;
            }
            any_restore |= restore[i] != 0;
        }
        if (any_restore) {
            // converged: return the previous iterate (newton.cpp:378-381)
            upload_flags(2, restore, d_nonconv_);
            launch_commit(n_, d_xold_, d_x_, d_nonconv_, 1, B_, stream_);
        }
        // mode changes for the next iteration's loads
        for (int i = 0; i < B_; ++i) h_ctx_[i].mode = ctx[i].mode;
        std::memcpy(h_ctx_pin_, h_ctx_.data(), sizeof(DevCtx) * B_);  // stream synced above
        NSA_HIP(hipMemcpyAsync(d_ctx_, h_ctx_pin_, sizeof(DevCtx) * B_, hipMemcpyHostToDevice,
                               stream_));
        upload_flags(3, done, d_done_);
        ms_load += std::chrono::duration<double, std::milli>(t1 - t0).count();
        ms_factor += std::chrono::duration<double, std::milli>(t2 - t1).count();
        ms_solve += std::chrono::duration<double, std::milli>(t3 - t2).count();
    }
    // restore the caller's mode (newton.cpp:224, 380, 417)
    for (int i = 0; i < B_; ++i) h_ctx_[i].mode = saved_mode[i];
    upload_ctx();
    NSA_HIP(hipStreamSynchronize(stream_));
    for (int i = 0; i < B_; ++i) out.total_iterations += out.iterations[i];
    return out;
}

// Host reads are ordered on the engine's (non-blocking) stream, behind any
// rotation, seeding or solve still queued there.
void BatchEngine::get_solution(double* out) const {
    NSA_HIP(hipMemcpyAsync(out, d_x_, sizeof(double) * n_ * B_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void BatchEngine::set_solution(const double* in) {
    NSA_HIP(hipMemcpyAsync(d_x_, in, sizeof(double) * n_ * B_, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void BatchEngine::get_state(int which, double* out) const {
    if (!ns_) return;
    NSA_HIP(hipMemcpyAsync(out, d_s_[which], sizeof(double) * ns_ * B_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void BatchEngine::set_state(int which, const double* in) {
    if (!ns_) return;
    NSA_HIP(hipMemcpyAsync(d_s_[which], in, sizeof(double) * ns_ * B_, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void BatchEngine::trunc(double* h_min, double* h_dev) {
    const double big = 1e30;
    unsigned long long big_bits;
    std::memcpy(&big_bits, &big, sizeof(big));
    launch_fill_u64(d_tmin_, big_bits, B_, stream_);
    const TruncOpts o{opt_.abstol, opt_.reltol, opt_.chgtol, opt_.trtol};
    launch_trunc(trunc_, d_ctx_, d_s_[0], d_s_[1], d_s_[2], d_s_[3], B_, o, h_dev ? d_tdev_ : nullptr,
                 d_tmin_, stream_);
    NSA_HIP(hipGetLastError());
    std::vector<unsigned long long> bits(B_);
    NSA_HIP(hipMemcpyAsync(bits.data(), d_tmin_, sizeof(unsigned long long) * B_, hipMemcpyDeviceToHost,
                           stream_));
    if (h_dev && trunc_.ndev)
        NSA_HIP(hipMemcpyAsync(h_dev, d_tdev_, sizeof(double) * trunc_.ndev * B_, hipMemcpyDeviceToHost,
                               stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    for (int i = 0; i < B_; ++i) std::memcpy(&h_min[i], &bits[i], sizeof(double));
}

void BatchEngine::seed_states_from_state0() {
    for (int k = 1; k < 4; ++k)
        if (ns_)
            NSA_HIP(hipMemcpyAsync(d_s_[k], d_s_[0], sizeof(double) * ns_ * B_,
                                   hipMemcpyDeviceToDevice, stream_));
}

void BatchEngine::rotate_states() {
    std::swap(d_s_[3], d_s_[2]);
    std::swap(d_s_[2], d_s_[1]);
    if (ns_)
        NSA_HIP(hipMemcpyAsync(d_s_[1], d_s_[0], sizeof(double) * ns_ * B_,
                               hipMemcpyDeviceToDevice, stream_));
}

}  // namespace nsa
