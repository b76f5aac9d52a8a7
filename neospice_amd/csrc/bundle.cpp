#include "bundle.hpp"
#include "parallel.hpp"

#include <chrono>

#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <atomic>
#include <exception>
#include <thread>

#include "bsim4.hpp"
#include "common.hpp"
#include "devices.hpp"
#include "gpu_bundle.hpp"

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

void csr(const std::vector<std::vector<int32_t>>& L, std::vector<int32_t>& ptr, std::vector<int32_t>& idx) {
    ptr.assign(L.size() + 1, 0);
    for (size_t i = 0; i < L.size(); ++i) ptr[i + 1] = ptr[i] + (int32_t)L[i].size();
    idx.clear();
    idx.reserve(ptr.back());
    for (auto& l : L) idx.insert(idx.end(), l.begin(), l.end());
}


}  // namespace

Bundle::Bundle(const std::vector<const Circuit*>& circuits, const Options& opt)
    : ckt_(circuits), opt_(opt) {
    NI_ = (int)circuits.size();
    if (NI_ <= 0) throw Error(NSA_ERR_ARGUMENT, "Bundle: no instances");
    for (const Circuit* c : circuits) {
        if (!c) throw Error(NSA_ERR_ARGUMENT, "Bundle: null circuit");
        if (!c->finalized()) throw Error(NSA_ERR_LOGIC, "Bundle: circuit not finalized");
        if (c->family(DK_B4).empty() && c->family(DK_M1).empty() && c->family(DK_D).empty() &&
            c->num_vars() >= 256 && c->is_linear() && opt.engine != ENGINE_MARKOWITZ)
            throw Error(NSA_ERR_UNSUPPORTED,
                        "Bundle: linear circuits of 256 or more unknowns use AMD-LU (BatchEngine)");
        if (opt.engine == ENGINE_AMDLU)
            throw Error(NSA_ERR_UNSUPPORTED, "Bundle: instances are factored with Sparse 1.3");
    }
    NSA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    xoff_.assign(NI_ + 1, 0);
    aoff_.assign(NI_ + 1, 0);
    doff_.assign(NI_ + 1, 0);
    n_.resize(NI_);
    nnodes_.resize(NI_);
    for (int i = 0; i < NI_; ++i) {
        n_[i] = ckt_[i]->num_vars();
        nnodes_[i] = ckt_[i]->num_nodes();
        xoff_[i + 1] = xoff_[i] + n_[i];
        aoff_[i + 1] = aoff_[i] + ckt_[i]->nnz();
        doff_[i + 1] = doff_[i] + (int32_t)ckt_[i]->devices().size();
    }
    tot_n_ = xoff_[NI_];
    tot_nnz_ = aoff_[NI_];
    d_x_ = dalloc<double>(allocs_, tot_n_);
    d_xold_ = dalloc<double>(allocs_, tot_n_);
    d_rhs_ = dalloc<double>(allocs_, tot_n_);
    d_A_ = dalloc<double>(allocs_, tot_nnz_);
    d_LU_ = dalloc<double>(allocs_, (size_t)NI_ * kSluMaxLu);
    NSA_HIP(hipMemset(d_x_, 0, sizeof(double) * tot_n_));
    d_ctx_ = dalloc<DevCtx>(allocs_, NI_);
    d_nt_ = dalloc<NtInst>(allocs_, NI_);
    d_done_ = dalloc<int32_t>(allocs_, NI_);
    d_devnc_ = dalloc<int32_t>(allocs_, NI_);
    d_kind_ = dalloc<int32_t>(allocs_, NI_);
    d_kind_host_ = dalloc<int32_t>(allocs_, NI_);
    d_fail_ = dalloc<int32_t>(allocs_, NI_);
    d_fail_col_ = dalloc<int32_t>(allocs_, NI_);
    NSA_HIP(hipMemset(d_done_, 0, sizeof(int32_t) * NI_));
    NSA_HIP(hipMemset(d_devnc_, 0, sizeof(int32_t) * NI_));
    NSA_HIP(hipHostMalloc((void**)&h_ctx_pin_, sizeof(DevCtx) * NI_, hipHostMallocDefault));
    NSA_HIP(hipHostMalloc((void**)&h_nt_pin_, sizeof(NtInst) * NI_, hipHostMallocDefault));
    // kinds, fail codes, and the modes and completion flags of attempts the host ends
    NSA_HIP(hipHostMalloc((void**)&h_kind_pin_, sizeof(int32_t) * 4 * NI_, hipHostMallocDefault));
    std::vector<int32_t> nn(n_.begin(), n_.end()), nd(nnodes_.begin(), nnodes_.end());
    d_ixoff_ = dupload(allocs_, std::vector<int32_t>(xoff_.begin(), xoff_.end()));
    {
        std::vector<int32_t> xo(tot_n_);
        for (int i = 0; i < NI_; ++i)
            for (int k = xoff_[i]; k < xoff_[i + 1]; ++k) xo[k] = i;
        d_xinst_ = dupload(allocs_, xo);
    }
    d_gmin_ = dupload(allocs_, std::vector<double>(NI_, opt_.diag_gmin));
    d_devgmin_ = dupload(allocs_, std::vector<double>(NI_, opt_.gmin));
    d_mask_ = dalloc<int32_t>(allocs_, 3 * (size_t)NI_);
    NSA_HIP(hipHostMalloc((void**)&h_mask_pin_, sizeof(int32_t) * 3 * NI_, hipHostMallocDefault));
    NSA_HIP(hipHostMalloc((void**)&h_gmin_pin_, sizeof(double) * NI_, hipHostMallocDefault));
    d_in_ = dupload(allocs_, nn);
    d_inn_ = dupload(allocs_, nd);
    tol_.gmin = opt_.gmin;
    tol_.gmin_inst = d_devgmin_;
    tol_.reltol = opt_.reltol;
    tol_.abstol = opt_.abstol;
    build_tables();
    plan_.assign(NI_, -1);
    s13_.resize(NI_);
    h_inst_.resize(NI_);
    for (int i = 0; i < NI_; ++i) h_inst_[i] = SluInst{-1, aoff_[i], i * kSluMaxLu, xoff_[i]};
    d_inst_ = dalloc<SluInst>(allocs_, NI_);
    upload_inst();
    DevCtx c{};
    c.mode = MODEDCOP | MODEINITJCT;
    c.order = 1;
    c.xmu_ratio = 1.0;
    c.src_fact = opt_.src_fact;
    set_ctx(std::vector<DevCtx>(NI_, c));
    NSA_HIP(hipStreamSynchronize(stream_));
}

Bundle::~Bundle() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (tick_exec_) (void)hipGraphExecDestroy(tick_exec_);
    if (tick_graph_) (void)hipGraphDestroy(tick_graph_);
    for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
    for (void* p : tran_allocs_) (void)hipFree(p);
    for (void* p : allocs_) (void)hipFree(p);
    if (h_ctx_pin_) (void)hipHostFree(h_ctx_pin_);
    if (h_nt_pin_) (void)hipHostFree(h_nt_pin_);
    if (h_kind_pin_) (void)hipHostFree(h_kind_pin_);
    if (h_mask_pin_) (void)hipHostFree(h_mask_pin_);
    if (h_gmin_pin_) (void)hipHostFree(h_gmin_pin_);
    if (h_acc_pin_) (void)hipHostFree(h_acc_pin_);
    if (h_tmin_pin_) (void)hipHostFree(h_tmin_pin_);
    if (h_ring_pin_) (void)hipHostFree(h_ring_pin_);
    if (h_tcs_pin_) (void)hipHostFree(h_tcs_pin_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

// Packed tables.  BSIM4v7 devices of every instance form one lane-strided
// batch (slot k of lane g at k * G + g, in the site, state, constant and
// auxiliary arrays, which it heads); every other family is a batch of one
// with its devices' sites and states packed after it.
void Bundle::build_tables() {
    DeviceEnv env;
    env.gmin = opt_.gmin;
    env.reltol = opt_.reltol;
    env.abstol = opt_.abstol;
    env.vntol = opt_.vntol;
    int G4 = 0;
    for (const Circuit* c : ckt_) G4 += (int)c->family(DK_B4).size();
    const int nsB4 = state_count(DK_B4);
    // per (instance, device): first site / rhs site / state, and the stride
    // of the device's local slots (G4 for BSIM4v7, 1 otherwise)
    std::vector<std::vector<int32_t>> site(NI_), rsite(NI_), sbase(NI_), stride(NI_);
    for (int i = 0; i < NI_; ++i) {
        const size_t nd = ckt_[i]->devices().size();
        site[i].assign(nd, -1);
        rsite[i].assign(nd, -1);
        sbase[i].assign(nd, -1);
        stride[i].assign(nd, 1);
    }
    // ---- BSIM4v7, lane-strided
    std::vector<int32_t> b4_pin(4 * (size_t)G4), b4_inst(G4);
    std::vector<double> b4_par((size_t)kB4Fields * G4);
    {
        int g = 0;
        for (int i = 0; i < NI_; ++i) {
            const Circuit& c = *ckt_[i];
            for (int d : c.family(DK_B4)) {
                const Device& dv = c.devices()[d];
                for (int k = 0; k < 4; ++k)
                    b4_pin[(size_t)k * G4 + g] = dv.node[k] < 0 ? -1 : xoff_[i] + dv.node[k];
                b4_inst[g] = i;
                const B4Param p = bsim4_param(c.models().at(dv.model), dv, env);
                const double* f = reinterpret_cast<const double*>(&p);
                for (int q = 0; q < kB4Fields; ++q) b4_par[(size_t)q * G4 + g] = f[q];
                site[i][d] = g;
                rsite[i][d] = g;
                sbase[i][d] = g;
                stride[i][d] = G4;
                ++g;
            }
        }
    }
    int ms = 16 * G4, rs = 4 * G4, st = nsB4 * G4;
    const long ns_lane = st;
    std::vector<int32_t> slot_inst;
    // ---- resistors
    std::vector<double> r_ohms;
    r_site_ = ms;
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        for (int d : c.family(DK_R)) { site[i][d] = ms; ms += 4; r_ohms.push_back(c.devices()[d].value); }
    }
    nR_ = (int)r_ohms.size();
    // ---- capacitors
    std::vector<int32_t> cap_nodes, cap_state, cap_inst;
    std::vector<double> cap_val;
    cap_.site = ms;
    cap_.rsite = rs;
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        for (int d : c.family(DK_C)) {
            const Device& dv = c.devices()[d];
            site[i][d] = ms; ms += 4;
            rsite[i][d] = rs; rs += 2;
            sbase[i][d] = st;
            for (int k = 0; k < state_count(DK_C); ++k) slot_inst.push_back(i);
            cap_state.push_back(st); st += state_count(DK_C);
            for (int k = 0; k < 2; ++k) cap_nodes.push_back(dv.node[k] < 0 ? -1 : xoff_[i] + dv.node[k]);
            cap_val.push_back(dv.value);
            cap_inst.push_back(i);
        }
    }
    cap_.count = (int)cap_val.size();
    // ---- diodes
    std::vector<int32_t> dio_nodes, dio_state, dio_inst;
    std::vector<DioParam> dio_par;
    dio_.site = ms;
    dio_.rsite = rs;
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        for (int d : c.family(DK_D)) {
            const Device& dv = c.devices()[d];
            site[i][d] = ms; ms += 7;
            rsite[i][d] = rs; rs += 2;
            sbase[i][d] = st;
            for (int k = 0; k < state_count(DK_D); ++k) slot_inst.push_back(i);
            dio_state.push_back(st); st += state_count(DK_D);
            for (int k = 0; k < 2; ++k) dio_nodes.push_back(dv.node[k] < 0 ? -1 : xoff_[i] + dv.node[k]);
            dio_par.push_back(diode_param(c.models().at(dv.model), dv, env));
            dio_inst.push_back(i);
        }
    }
    dio_.count = (int)dio_par.size();
    // ---- MOS1
    std::vector<int32_t> mos_nodes, mos_state, mos_inst;
    std::vector<Mos1Param> mos_par;
    mos_.site = ms;
    mos_.rsite = rs;
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        for (int d : c.family(DK_M1)) {
            const Device& dv = c.devices()[d];
            site[i][d] = ms; ms += 22;
            rsite[i][d] = rs; rs += 4;
            sbase[i][d] = st;
            for (int k = 0; k < state_count(DK_M1); ++k) slot_inst.push_back(i);
            mos_state.push_back(st); st += state_count(DK_M1);
            for (int k = 0; k < 4; ++k) mos_nodes.push_back(dv.node[k] < 0 ? -1 : xoff_[i] + dv.node[k]);
            mos_par.push_back(mos1_param(c.models().at(dv.model), dv, env));
            mos_inst.push_back(i);
        }
    }
    mos_.count = (int)mos_par.size();
    // ---- constant matrix sites: V source +-1, dead-node pins
    const_site_ = ms;
    std::vector<double> const_vals;
    const double pin = (opt_.gmin > 0.0) ? opt_.gmin : 1e-12;  // newton.cpp:80
    std::vector<std::vector<int32_t>> pin_site(NI_);
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        for (int d : c.family(DK_V)) {
            site[i][d] = ms; ms += 4;
            const_vals.insert(const_vals.end(), {1.0, -1.0, 1.0, -1.0});
        }
        for (size_t k = 0; k < c.dead_nodes().size(); ++k) {
            pin_site[i].push_back(ms++);
            const_vals.push_back(pin);
        }
    }
    n_const_ = (int)const_vals.size();
    n_msites_ = ms;
    // ---- voltage and current sources (rhs sites)
    std::vector<int32_t> vw, vdg, pwl_ptr{0}, v_inst, i_inst;
    std::vector<double> vdc, pwl_t, pwl_v, ival;
    vsrc_.rsite = rs;
    h_pulse_.clear();
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        for (int d : c.family(DK_V)) {
            const Device& dv = c.devices()[d];
            rsite[i][d] = rs; rs += 1;
            vw.push_back(dv.wave);
            vdg.push_back(0);
            vdc.push_back(dv.value);
            for (int k = 0; k < 7; ++k) h_pulse_.push_back(dv.pulse[k]);
            pwl_t.insert(pwl_t.end(), dv.pwl_t.begin(), dv.pwl_t.end());
            pwl_v.insert(pwl_v.end(), dv.pwl_v.begin(), dv.pwl_v.end());
            pwl_ptr.push_back((int32_t)pwl_t.size());
            v_inst.push_back(i);
        }
    }
    vsrc_.count = (int)vdc.size();
    isrc_.rsite = rs;
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        for (int d : c.family(DK_I)) {
            rsite[i][d] = rs; rs += 2;
            ival.push_back(c.devices()[d].value);
            i_inst.push_back(i);
        }
    }
    isrc_.count = (int)ival.size();
    n_rsites_ = rs;
    ns_tot_ = st;
    // ---- gathers in each instance's load order
    std::vector<std::vector<int32_t>> slot_sites(tot_nnz_), rows_d(tot_n_), rows_u(tot_n_);
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        const auto& devs = c.devices();
        for (int d : c.load_order()) {
            const Device& dv = devs[d];
            const int sd = stride[i][d];
            for (size_t k = 0; k < dv.mat_slot.size(); ++k)
                if (dv.mat_slot[k] >= 0)
                    slot_sites[aoff_[i] + dv.mat_slot[k]].push_back(site[i][d] + (int)k * sd);
            auto& rows = (dv.kind == DK_D || dv.kind == DK_M1 || dv.kind == DK_B4) ? rows_u : rows_d;
            for (size_t k = 0; k < dv.rhs_row.size(); ++k)
                if (dv.rhs_row[k] >= 0) rows[xoff_[i] + dv.rhs_row[k]].push_back(rsite[i][d] + (int)k * sd);
        }
        for (size_t k = 0; k < c.dead_nodes().size(); ++k) {
            const int nd = c.dead_nodes()[k];
            slot_sites[aoff_[i] + c.offset(nd, nd)].push_back(pin_site[i][k]);
        }
    }
    // every index the kernels will follow stays inside its array
    auto in_range = [](const std::vector<int32_t>& v, long lo, long hi, const char* what) {
        for (int32_t k : v)
            if (k < lo || k >= hi) throw Error(NSA_ERR_RUNTIME, std::string("Bundle tables: ") + what);
    };
    for (const auto& l : slot_sites) in_range(l, 0, n_msites_, "matrix site");
    for (const auto& l : rows_d) in_range(l, 0, n_rsites_, "rhs site");
    for (const auto& l : rows_u) in_range(l, 0, n_rsites_, "rhs site");
    in_range(b4_pin, -1, tot_n_, "BSIM4v7 terminal");
    in_range(cap_nodes, -1, tot_n_, "capacitor node");
    in_range(dio_nodes, -1, tot_n_, "diode node");
    in_range(mos_nodes, -1, tot_n_, "MOS1 node");
    in_range(cap_state, 0, ns_tot_ - state_count(DK_C) + 1, "capacitor state");
    in_range(dio_state, 0, ns_tot_ - state_count(DK_D) + 1, "diode state");
    in_range(mos_state, 0, ns_tot_ - state_count(DK_M1) + 1, "MOS1 state");
    if ((long)slot_inst.size() != ns_tot_ - ns_lane) throw Error(NSA_ERR_RUNTIME, "Bundle tables: state owners");
    if (r_site_ + 4L * nR_ > const_site_ || cap_.site + 4L * cap_.count > dio_.site ||
        dio_.site + 7L * dio_.count > mos_.site || mos_.site + 22L * mos_.count > const_site_ ||
        const_site_ + (long)n_const_ != n_msites_)
        throw Error(NSA_ERR_RUNTIME, "Bundle tables: matrix site blocks");
    if (cap_.rsite + 2L * cap_.count > dio_.rsite || dio_.rsite + 2L * dio_.count > mos_.rsite ||
        mos_.rsite + 4L * mos_.count > vsrc_.rsite || vsrc_.rsite + (long)vsrc_.count > isrc_.rsite ||
        isrc_.rsite + 2L * isrc_.count != n_rsites_)
        throw Error(NSA_ERR_RUNTIME, "Bundle tables: rhs site blocks");
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
    // short / long row lists of the two gathers (a row is long when either
    // of its site lists exceeds 32 sites)
    {
        std::vector<int32_t> ms, ml, rs_, rl;
        for (int r = 0; r < tot_nnz_; ++r) (slot_sites[r].size() > 32 ? ml : ms).push_back(r);
        for (int r = 0; r < tot_n_; ++r)
            (rows_d[r].size() > 32 || rows_u[r].size() > 32 ? rl : rs_).push_back(r);
        gm_.nshort = (int)ms.size();
        gm_.nlong = (int)ml.size();
        gm_.short_rows = dupload(allocs_, ms);
        gm_.long_rows = dupload(allocs_, ml);
        gr_.nshort = (int)rs_.size();
        gr_.nlong = (int)rl.size();
        gr_.short_rows = dupload(allocs_, rs_);
        gr_.long_rows = dupload(allocs_, rl);
    }
    // ---- device arrays
    d_cm_ = dalloc<double>(allocs_, n_msites_);
    d_cr_ = dalloc<double>(allocs_, n_rsites_);
    NSA_HIP(hipMemset(d_cm_, 0, sizeof(double) * std::max(1, n_msites_)));
    NSA_HIP(hipMemset(d_cr_, 0, sizeof(double) * std::max(1, n_rsites_)));
    for (int k = 0; k < 4; ++k) {
        d_s_[k] = dalloc<double>(allocs_, ns_tot_);
        NSA_HIP(hipMemset(d_s_[k], 0, sizeof(double) * std::max(1, ns_tot_)));
    }
    d_r_ohms_ = dupload(allocs_, r_ohms);
    cap_.nodes = dupload(allocs_, cap_nodes);
    cap_.state = dupload(allocs_, cap_state);
    cap_.farads = dupload(allocs_, cap_val);
    cap_.inst_of = dupload(allocs_, cap_inst);
    vsrc_.wave = dupload(allocs_, vw);
    vsrc_.dc_given = dupload(allocs_, vdg);
    vsrc_.dc = dupload(allocs_, vdc);
    vsrc_.pulse = dupload(allocs_, h_pulse_);
    vsrc_.pwl_ptr = dupload(allocs_, pwl_ptr);
    vsrc_.pwl_t = dupload(allocs_, pwl_t);
    vsrc_.pwl_v = dupload(allocs_, pwl_v);
    vsrc_.inst_of = dupload(allocs_, v_inst);
    isrc_.dc = dupload(allocs_, ival);
    isrc_.inst_of = dupload(allocs_, i_inst);
    d_const_vals_ = dupload(allocs_, const_vals);
    dio_.nodes = dupload(allocs_, dio_nodes);
    dio_.state = dupload(allocs_, dio_state);
    dio_.par = dupload(allocs_, dio_par);
    dio_.inst_of = dupload(allocs_, dio_inst);
    mos_.nodes = dupload(allocs_, mos_nodes);
    mos_.state = dupload(allocs_, mos_state);
    mos_.par = dupload(allocs_, mos_par);
    mos_.inst_of = dupload(allocs_, mos_inst);
    mos_.aux = dalloc<double>(allocs_, (size_t)std::max(1, mos_.count) * kMos1Aux);
    NSA_HIP(hipMemset(mos_.aux, 0, sizeof(double) * std::max(1, mos_.count) * kMos1Aux));
    // BSIM4v7 as one device of G4 lanes
    b4_.count = G4 > 0 ? 1 : 0;
    b4_.nodes = dupload(allocs_, std::vector<int32_t>{-1, -1, -1, -1});
    b4_.state = dupload(allocs_, std::vector<int32_t>{0});
    b4_.par = nullptr;
    b4_.par_inst = dupload(allocs_, b4_par);
    b4_.aux = dalloc<double>(allocs_, (size_t)kB4Aux * std::max(1, G4));
    NSA_HIP(hipMemset(b4_.aux, 0, sizeof(double) * kB4Aux * std::max(1, G4)));
    b4_.site = 0;
    b4_.rsite = 0;
    b4_.inst_of = dupload(allocs_, b4_inst);
    b4_.pin = dupload(allocs_, b4_pin);
    b4_lanes_ = G4;
    // truncation-error rows: capacitors, diodes, MOS1 (batch of one), BSIM4v7 (lanes)
    std::vector<int32_t> tb, tk, ti;
    for (size_t k = 0; k < cap_state.size(); ++k) { tb.push_back(cap_state[k]); tk.push_back(0); ti.push_back(cap_inst[k]); }
    for (size_t k = 0; k < dio_state.size(); ++k) { tb.push_back(dio_state[k]); tk.push_back(2); ti.push_back(dio_inst[k]); }
    for (size_t k = 0; k < mos_state.size(); ++k) { tb.push_back(mos_state[k]); tk.push_back(1); ti.push_back(mos_inst[k]); }
    trunc1_.ndev = (int)tb.size();
    trunc1_.base = dupload(allocs_, tb);
    trunc1_.kind = dupload(allocs_, tk);
    trunc1_.inst_of = dupload(allocs_, ti);
    trunc4_.ndev = G4 > 0 ? 1 : 0;
    trunc4_.base = b4_.state;
    trunc4_.kind = dupload(allocs_, std::vector<int32_t>{3});
    trunc4_.inst_of = b4_.inst_of;
    ns_lane_ = ns_lane;
    d_slot_inst_ = dupload(allocs_, slot_inst);
    // constant sites
    launch_res_sites(nR_, d_r_ohms_, d_cm_, r_site_, 1, stream_);
    launch_const_sites(n_const_, d_const_vals_, d_cm_, const_site_, 1, stream_);
}

void Bundle::upload_inst() {
    for (int i = 0; i < NI_; ++i) h_inst_[i].plan = plan_[i];
    NSA_HIP(hipMemcpyAsync(d_inst_, h_inst_.data(), sizeof(SluInst) * NI_, hipMemcpyHostToDevice,
                           stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    inst_dirty_ = false;
}

void Bundle::set_ctx(const std::vector<DevCtx>& c) {
    if ((int)c.size() != NI_) throw Error(NSA_ERR_ARGUMENT, "Bundle::set_ctx: one context per instance");
    h_ctx_ = c;
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(h_ctx_pin_, h_ctx_.data(), sizeof(DevCtx) * NI_);
    NSA_HIP(hipMemcpyAsync(d_ctx_, h_ctx_pin_, sizeof(DevCtx) * NI_, hipMemcpyHostToDevice, stream_));
}

void Bundle::reset_solver() {
    plan_.assign(NI_, -1);
    for (auto& s : s13_) s.reset();
    inst_dirty_ = true;
}

void Bundle::get_solution(double* out) const {
    NSA_HIP(hipMemcpyAsync(out, d_x_, sizeof(double) * tot_n_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void Bundle::set_solution(const double* in) {
    NSA_HIP(hipMemcpyAsync(d_x_, in, sizeof(double) * tot_n_, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void Bundle::get_state(int which, double* out) const {
    if (which < 0 || which > 3) throw Error(NSA_ERR_ARGUMENT, "Bundle::get_state: which in 0..3");
    if (!ns_tot_) return;
    NSA_HIP(hipMemcpyAsync(out, d_s_[which], sizeof(double) * ns_tot_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void Bundle::set_timing(bool on) {
    if (on && !ev_[0])
        for (auto& e : ev_) NSA_HIP(hipEventCreate(&e));
    timing_ = on;
    ms_load = ms_lu = ms_post = ms_bsim4 = 0.0;
    n_timed = 0;
}

// ---------------------------------------------------------------- iteration

namespace {
unsigned long long bits_1e30() {
    const double big = 1e30;
    unsigned long long b;
    std::memcpy(&b, &big, sizeof(big));
    return b;
}
}  // namespace

void Bundle::launch_begin() {
    launch_nt_begin(d_nt_, NI_, d_ixoff_, d_in_, d_x_, d_xold_, d_ctx_, d_kind_, d_fail_, d_devnc_,
                    d_done_, stream_);
}

// Loads of the families other than BSIM4v7.
void Bundle::launch_loads_small() {
    double* s0 = d_s_[0];
    double* s1 = d_s_[1];
    double* s2 = d_s_[2];
    launch_cap(cap_, d_ctx_, d_x_, s0, s1, s2, d_done_, d_cm_, d_cr_, 1, stream_);
    // with BSIM4v7 devices present the voltage sources ride in the BSIM4v7
    // load's launch (launch_loads_b4_gather)
    if (!b4_.count) launch_vsrc(vsrc_, d_ctx_, d_done_, d_cr_, 1, stream_);
    launch_isrc(isrc_, d_ctx_, d_done_, d_cr_, 1, stream_);
    launch_diode(dio_, d_ctx_, d_x_, s0, s1, s2, d_done_, d_cm_, d_cr_, d_devnc_, tol_, 1, stream_);
    launch_mos1(mos_, d_ctx_, d_x_, s0, s1, s2, d_done_, d_cm_, d_cr_, d_devnc_, tol_, 1, stream_);
}

// The BSIM4v7 load, then the matrix and rhs gathers.
void Bundle::launch_loads_b4_gather() {
    if (timing_ && b4_.count) NSA_HIP(hipEventRecord(ev_[1], stream_));
    const VsrcSide vs{vsrc_, d_ctx_, d_done_, d_cr_, 1};
    launch_bsim4(b4_, d_ctx_, d_x_, d_s_[0], d_s_[1], d_s_[2], d_done_, d_cm_, d_cr_, d_devnc_, tol_, b4_lanes_,
                 stream_, &vs);
    if (timing_ && b4_.count) NSA_HIP(hipEventRecord(ev_[2], stream_));
    GatherPacked m = gm_, r = gr_;
    m.ptr = d_mptr_; m.sites = d_msites_; m.contrib = d_cm_; m.ptr2 = nullptr; m.sites2 = nullptr;
    m.contrib2 = nullptr; m.out = d_A_;
    r.ptr = d_rptr_d_; r.sites = d_rsites_d_; r.contrib = d_cr_; r.ptr2 = d_rptr_u_;
    r.sites2 = d_rsites_u_; r.contrib2 = d_cr_; r.out = d_rhs_;
    launch_gather_packed(m, r, stream_);
}

void Bundle::launch_loads() {
    launch_loads_small();
    launch_loads_b4_gather();
}

void Bundle::launch_lu(const int32_t* d_kind, int mode) {
    launch_small_lu(plans_, d_inst_, NI_, mode, d_A_, opt_.diag_gmin, d_kind, d_fail_, d_fail_col_,
                    d_LU_, d_rhs_, d_x_, stream_, d_gmin_);
}

void Bundle::launch_post(bool tick) {
    // MOS1Device::conv_test on the proposed solution (skipped where nothing was solved)
    launch_mos1_conv(mos_, d_x_, d_s_[0], d_fail_, d_devnc_, tol_, 1, stream_);
    launch_nt_post(d_nt_, NI_, d_ixoff_, d_in_, d_inn_, d_x_, d_xold_, d_ctx_, d_kind_, d_fail_,
                   d_devnc_, d_done_, NtConvTol{opt_.reltol, opt_.vntol, opt_.abstol}, stream_,
                   tick ? d_tc_ : nullptr, tick ? d_att_ : nullptr);
}

void Bundle::capture_graph() {
    if (graph_exec_) { NSA_HIP(hipGraphExecDestroy(graph_exec_)); graph_exec_ = nullptr; }
    if (graph_) { NSA_HIP(hipGraphDestroy(graph_)); graph_ = nullptr; }
    NSA_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    launch_begin();
    launch_loads();
    launch_lu(d_kind_, SLU_FACTOR | SLU_SOLVE);
    launch_post();
    NSA_HIP(hipStreamEndCapture(stream_, &graph_));
    NSA_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
    graph_words_ = plans_.d_words();
    graph_nlu_ = plans_.max_nlu();
    graph_n_ = plans_.max_n();
    graph_hot_ = plans_.max_hot();
}

// Whether a graph captured with these plan words and launch-shape limits
// would launch the packed LU differently now (its LDS size and staging
// follow the largest factor, vector and hot block of the plans).
bool Bundle::graph_stale(const int32_t* words, int nlu, int n, int hot) const {
    return words != plans_.d_words() || nlu != plans_.max_nlu() || n != plans_.max_n() ||
           hot != plans_.max_hot();
}

void Bundle::iteration_device() {
    if (inst_dirty_) upload_inst();
    if (timing_) {
        NSA_HIP(hipEventRecord(ev_[0], stream_));
        launch_begin();
        launch_loads();
        NSA_HIP(hipEventRecord(ev_[3], stream_));
        launch_lu(d_kind_, SLU_FACTOR | SLU_SOLVE);
        NSA_HIP(hipEventRecord(ev_[4], stream_));
        launch_post();
        NSA_HIP(hipEventRecord(ev_[5], stream_));
        timed_pending_ = true;
        return;
    }
    if (use_graph) {
        if (!graph_exec_ || graph_stale(graph_words_, graph_nlu_, graph_n_, graph_hot_)) capture_graph();
        NSA_HIP(hipGraphLaunch(graph_exec_, stream_));
        ++n_graph_launches;
        return;
    }
    launch_begin();
    launch_loads();
    launch_lu(d_kind_, SLU_FACTOR | SLU_SOLVE);
    launch_post();
}

std::vector<double> Bundle::instance_matrix(int i) {
    std::vector<double> v(aoff_[i + 1] - aoff_[i]);
    NSA_HIP(hipMemcpyAsync(v.data(), d_A_ + aoff_[i], sizeof(double) * v.size(), hipMemcpyDeviceToHost,
                           stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    return v;
}

// Host part of an iteration for the instances the device cannot finish on
// its own: the first factorization of a fresh solver (called right after the
// iteration's loads, before any LU) and numeric() calls whose reused pivot
// order failed the threshold test (stalled after the LU).  Each instance in
// `repivot` is re-pivoted on the CPU from its own matrix as
// NeoSolver::numeric does (a fresh ordering, or the reuse path on the
// instance's own state), then factored and solved on the GPU on its new
// plan; with `others`, every other running instance takes its LU call of
// this iteration too.  The iteration then ends on the device as usual.
// `nt` is the host copy of the device state (current on entry).
void Bundle::host_lu_post(std::vector<NtInst>& nt, const std::vector<int>& repivot, bool others) {
    std::vector<double> A(tot_nnz_);
    NSA_HIP(hipMemcpyAsync(A.data(), d_A_, sizeof(double) * tot_nnz_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    std::vector<int32_t> sing(repivot.size(), 0);
    std::vector<LuStructure> st(repivot.size());
    parallel_for((int)repivot.size(), [&](int t) {
        const int i = repivot[t];
        const Circuit& c = *ckt_[i];
        if (!s13_[i])
            s13_[i] = std::make_unique<Sparse13>(c.num_vars(), c.col_ptr().data(), c.row_idx().data());
        if (s13_[i]->numeric(A.data() + aoff_[i], opt_.diag_gmin) != 0) {
            sing[t] = 1;
            return;
        }
        st[t] = s13_[i]->structure();
    });
    n_cpu_pivot += (long)repivot.size();
    std::vector<int32_t> kinds(NI_, SLU_SKIP), fails(NI_, -1);
    if (others)
        for (int i = 0; i < NI_; ++i)
            if (!nt[i].done && !nt[i].stalled && plan_[i] >= 0) {
                kinds[i] = nt[i].kind;
                fails[i] = 0;
            }
    std::vector<int> failed;
    for (size_t t = 0; t < repivot.size(); ++t) {
        const int i = repivot[t];
        NtInst& s = nt[i];
        if (sing[t]) {
            // singular numeric(): the attempt fails (newton.cpp:220-224)
            s.iter += 1;
            s.converged = 0;
            s.done = 1;
            s.stalled = 0;
            kinds[i] = SLU_SKIP;
            failed.push_back(i);
            continue;
        }
        if (!SmallLuPlans::fits(st[t]))
            throw Error(NSA_ERR_UNSUPPORTED, "Bundle: an instance's factor exceeds the packed LU limits");
        plan_[i] = plans_.add(st[t]);
        s.stalled = 0;
        s.kind = SLU_TRUSTED;
        kinds[i] = SLU_TRUSTED;
        fails[i] = 0;
    }
    plans_.sync(stream_);
    upload_inst();
    std::memcpy(h_nt_pin_, nt.data(), sizeof(NtInst) * NI_);
    NSA_HIP(hipMemcpyAsync(d_nt_, h_nt_pin_, sizeof(NtInst) * NI_, hipMemcpyHostToDevice, stream_));
    std::memcpy(h_kind_pin_, kinds.data(), sizeof(int32_t) * NI_);
    std::memcpy(h_kind_pin_ + NI_, fails.data(), sizeof(int32_t) * NI_);
    NSA_HIP(hipMemcpyAsync(d_kind_, h_kind_pin_, sizeof(int32_t) * NI_, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipMemcpyAsync(d_fail_, h_kind_pin_ + NI_, sizeof(int32_t) * NI_, hipMemcpyHostToDevice,
                           stream_));
    // a failed attempt leaves the caller's mode in the context and its
    // completion flag set
    for (int i : failed) {
        h_kind_pin_[2 * NI_ + i] = nt[i].saved_mode;
        h_kind_pin_[3 * NI_ + i] = nt[i].done;
        NSA_HIP(hipMemcpyAsync(&d_ctx_[i].mode, h_kind_pin_ + 2 * NI_ + i, sizeof(int32_t),
                               hipMemcpyHostToDevice, stream_));
        NSA_HIP(hipMemcpyAsync(d_done_ + i, h_kind_pin_ + 3 * NI_ + i, sizeof(int32_t), hipMemcpyHostToDevice,
                               stream_));
    }
    launch_lu(d_kind_, SLU_FACTOR | SLU_SOLVE);
    launch_post();
    NSA_HIP(hipStreamSynchronize(stream_));
    ++n_host_iterations;
}

void Bundle::collect_timing() {
    if (!timed_pending_) return;
    float ms = 0.f;
    NSA_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[3]));
    ms_load += ms;
    NSA_HIP(hipEventElapsedTime(&ms, ev_[3], ev_[4]));
    ms_lu += ms;
    NSA_HIP(hipEventElapsedTime(&ms, ev_[4], ev_[5]));
    ms_post += ms;
    if (b4_.count) {
        NSA_HIP(hipEventElapsedTime(&ms, ev_[1], ev_[2]));
        ms_bsim4 += ms;
    }
    ++n_timed;
    timed_pending_ = false;
}

void Bundle::newton(int itl, const int32_t* active, std::vector<int32_t>& converged,
                    std::vector<int32_t>& iterations) {
    std::vector<NtInst> nt(NI_);
    for (int i = 0; i < NI_; ++i) {
        NtInst s{};
        s.saved_mode = h_ctx_[i].mode;
        s.done = (active && !active[i]) ? 1 : 0;
        s.itl = itl;
        s.kind = SLU_SKIP;
        nt[i] = s;
    }
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(h_nt_pin_, nt.data(), sizeof(NtInst) * NI_);
    NSA_HIP(hipMemcpyAsync(d_nt_, h_nt_pin_, sizeof(NtInst) * NI_, hipMemcpyHostToDevice, stream_));
    auto readback = [&]() {
        NSA_HIP(hipMemcpyAsync(h_nt_pin_, d_nt_, sizeof(NtInst) * NI_, hipMemcpyDeviceToHost, stream_));
        NSA_HIP(hipStreamSynchronize(stream_));
        collect_timing();
        std::memcpy(nt.data(), h_nt_pin_, sizeof(NtInst) * NI_);
    };
    bool first = true;
    // every round advances each running instance by at least one iteration
    for (int round = 0; round <= itl + 1; ++round) {
        std::vector<int> fresh;
        for (int i = 0; i < NI_; ++i)
            if (!nt[i].done && plan_[i] < 0) fresh.push_back(i);
        if (!fresh.empty()) {
            // this iteration's loads, then the CPU orderings and every LU
            if (inst_dirty_) upload_inst();
            launch_begin();
            launch_loads();
            readback();
            host_lu_post(nt, fresh, true);
            readback();
        } else {
            // two iterations before the first look (iteration 0 never
            // converges); one at a time while HIP events time the phases
            const int k = (first && !timing_) ? 2 : 1;
            for (int r = 0; r < k; ++r) iteration_device();
            readback();
        }
        first = false;
        std::vector<int> stalled;
        for (int i = 0; i < NI_; ++i)
            if (!nt[i].done && nt[i].stalled) stalled.push_back(i);
        if (!stalled.empty()) {
            host_lu_post(nt, stalled, false);
            readback();
        }
        bool all = true;
        for (int i = 0; i < NI_; ++i) all &= nt[i].done != 0;
        if (all) break;
    }
    converged.assign(NI_, 0);
    iterations.assign(NI_, 0);
    for (int i = 0; i < NI_; ++i) {
        if (active && !active[i]) continue;
        if (!nt[i].done) throw Error(NSA_ERR_RUNTIME, "Bundle::newton: an instance did not finish");
        converged[i] = nt[i].converged;
        iterations[i] = nt[i].iter;
        n_iterations += nt[i].iter;
    }
}

// ---------------------------------------------------------------- DC ladder

void Bundle::set_dev_gmin(const std::vector<double>& g) {
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(h_gmin_pin_, g.data(), sizeof(double) * NI_);
    NSA_HIP(hipMemcpyAsync(d_devgmin_, h_gmin_pin_, sizeof(double) * NI_, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

void Bundle::set_gmin(const std::vector<double>& g) {
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(h_gmin_pin_, g.data(), sizeof(double) * NI_);
    NSA_HIP(hipMemcpyAsync(d_gmin_, h_gmin_pin_, sizeof(double) * NI_, hipMemcpyHostToDevice, stream_));
}

// Copies, for the instances with mask[i] != 0, the solution and states 0-2
// from the buffers `src` (x, s0, s1, s2; null = zeros) into `dst`.
void Bundle::masked_copy_state(const std::vector<int32_t>& mask, double* const* src, double* const* dst) {
    bool any = false;
    for (int32_t m : mask) any |= m != 0;
    if (!any) return;
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(h_mask_pin_, mask.data(), sizeof(int32_t) * NI_);
    NSA_HIP(hipMemcpyAsync(d_mask_, h_mask_pin_, sizeof(int32_t) * NI_, hipMemcpyHostToDevice, stream_));
    launch_masked_copy(tot_n_, 0, 1, d_xinst_, d_xinst_, d_mask_, src ? src[0] : nullptr, dst[0], stream_);
    for (int k = 1; k < 4; ++k)
        launch_masked_copy(ns_tot_, ns_lane_, b4_lanes_, b4_.inst_of, d_slot_inst_, d_mask_,
                           src ? src[k] : nullptr, dst[k], stream_);
    NSA_HIP(hipStreamSynchronize(stream_));
}

void Bundle::op_ladder(int firstmode, int continuemode, std::vector<int32_t>& converged,
                       std::vector<int32_t>& iterations, std::vector<int32_t>& stage,
                       int alone) {
    const bool skip_direct = alone != LADDER_NONE;
    // 1. direct Newton (dc.cpp:104-107, transient.cpp:195-205)
    if (skip_direct) {
        converged.assign(NI_, 0);
        iterations.assign(NI_, 0);
        std::vector<DevCtx> c = h_ctx_;
        for (auto& x : c) x.mode = firstmode;
        set_ctx(c);
    } else {
        dc_op(firstmode, converged, iterations);
    }
    stage.assign(NI_, LADDER_NONE);
    for (int i = 0; i < NI_; ++i)
        if (converged[i]) stage[i] = LADDER_DIRECT;
    if (!d_lad_[0])
        for (int k = 0; k < 8; ++k)
            d_lad_[k] = dalloc<double>(allocs_, k % 4 == 0 ? tot_n_ : std::max(1, ns_tot_));
    // 2. for the instances still failing, the continuation methods in the
    //    reference's order: solve_dc (MODEDCOP) tries true gmin stepping
    //    first (dc.cpp:120-145); both ladders then gmin stepping and source
    //    stepping (dc.cpp:146-184, transient.cpp:206-222)
    if (alone == LADDER_GMIN || alone == LADDER_TRUE_GMIN) {
        ladder_gmin(firstmode, continuemode, alone == LADDER_TRUE_GMIN, converged, iterations, stage);
        return;
    }
    if (alone == LADDER_SOURCE) {
        ladder_source(firstmode & ~INITF_MASK, converged, iterations, stage);
        return;
    }
    if (skip_direct) throw Error(NSA_ERR_ARGUMENT, "Bundle::op_ladder: unknown method");
    if (firstmode & MODEDCOP) ladder_gmin(firstmode, continuemode, true, converged, iterations, stage);
    ladder_gmin(firstmode, continuemode, false, converged, iterations, stage);
    ladder_source(firstmode & ~INITF_MASK, converged, iterations, stage);
}

// gmin stepping (convergence.cpp:216-346) or, with device, true gmin
// stepping (convergence.cpp:348-470) for every instance whose stage is
// still LADDER_NONE, each on its own schedule: the stepped conductance is
// the diagonal gmin (gmin stepping) or the devices' gmin (true gmin
// stepping), from 1e-3 down to max(gmin, gshunt) with ngspice's factor
// adaptation; then one more solve at the target (diag_gmin = gshunt, or
// the original device gmin).  An instance that fails goes back to its
// entry solution and states.
void Bundle::ladder_gmin(int firstmode, int continuemode, bool device, std::vector<int32_t>& converged,
                         std::vector<int32_t>& iterations, std::vector<int32_t>& stage) {
    std::vector<int32_t> failing(NI_, 0);
    bool any = false;
    for (int i = 0; i < NI_; ++i) {
        failing[i] = stage[i] == LADDER_NONE;
        any |= failing[i] != 0;
    }
    if (!any) return;
    double* cur[4] = {d_x_, d_s_[0], d_s_[1], d_s_[2]};
    double* entry[4] = {d_lad_[0], d_lad_[1], d_lad_[2], d_lad_[3]};
    double* saved[4] = {d_lad_[4], d_lad_[5], d_lad_[6], d_lad_[7]};
    masked_copy_state(failing, cur, entry);
    masked_copy_state(failing, nullptr, cur);   // zero initial guess, clear_state
    masked_copy_state(failing, cur, saved);
    const double gmin_factor = 10.0;
    const double gtarget = std::max(opt_.gmin, 0.0);  // max(gmin, gshunt)
    const int itl2 = 50;                               // SimOptions::itl2
    const int step_itl = std::max(itl2, 100);
    enum { L_IDLE = 0, L_STEP, L_FINAL, L_DONE, L_FAILED };
    std::vector<int> phase(NI_, L_IDLE), mode(NI_, firstmode), total(NI_, 0);
    std::vector<double> factor(NI_, gmin_factor), old_gmin(NI_, 1e-2), dg(NI_, 1e-2 / gmin_factor);
    for (int i = 0; i < NI_; ++i)
        if (failing[i]) phase[i] = L_STEP;
    std::vector<int32_t> conv, its;
    for (int round = 0; round < 100000; ++round) {
        std::vector<int32_t> active(NI_, 0);
        std::vector<double> g(NI_, opt_.diag_gmin), dev(NI_, opt_.gmin);
        std::vector<DevCtx> ctx = h_ctx_;
        bool live = false;
        for (int i = 0; i < NI_; ++i) {
            if (phase[i] != L_STEP && phase[i] != L_FINAL) continue;
            active[i] = 1;
            live = true;
            ctx[i].mode = mode[i];
            if (device) {
                if (phase[i] == L_STEP) dev[i] = dg[i];  // the final solve at the original gmin
            } else {
                g[i] = (phase[i] == L_STEP) ? dg[i] : 0.0;  // the final solve at gshunt
            }
        }
        if (!live) break;
        set_ctx(ctx);
        set_gmin(g);
        set_dev_gmin(dev);
        // both kinds of solve run the same iteration limit (max(itl2, 100) and
        // SimOptions::max_iter, 100 by default)
        newton(std::max(step_itl, opt_.itl1), active.data(), conv, its);
        std::vector<int32_t> save(NI_, 0), back(NI_, 0), to_entry(NI_, 0);
        for (int i = 0; i < NI_; ++i) {
            if (!active[i]) continue;
            total[i] += its[i];
            if (phase[i] == L_FINAL) {
                if (conv[i]) {
                    phase[i] = L_DONE;
                } else {
                    phase[i] = L_FAILED;
                    to_entry[i] = 1;
                }
                continue;
            }
            if (conv[i]) {
                mode[i] = continuemode;
                if (dg[i] <= gtarget) {
                    phase[i] = L_FINAL;
                } else {
                    save[i] = 1;
                    // ngspice cktop.c:187-194
                    if (its[i] <= itl2 / 4) {
                        factor[i] *= std::sqrt(factor[i]);
                        if (factor[i] > gmin_factor) factor[i] = gmin_factor;
                    }
                    if (its[i] > (3 * itl2 / 4)) factor[i] = std::max(std::sqrt(factor[i]), 1.00005);
                    old_gmin[i] = dg[i];
                    if (dg[i] < factor[i] * gtarget) {
                        factor[i] = dg[i] / gtarget;
                        dg[i] = gtarget;
                    } else {
                        dg[i] /= factor[i];
                    }
                }
            } else if (factor[i] < 1.00005) {
                phase[i] = L_FAILED;
                to_entry[i] = 1;
            } else {
                factor[i] = std::sqrt(std::sqrt(factor[i]));
                dg[i] = old_gmin[i] / factor[i];
                back[i] = 1;
            }
        }
        masked_copy_state(save, cur, saved);
        masked_copy_state(back, saved, cur);
        masked_copy_state(to_entry, entry, cur);
    }
    set_gmin(std::vector<double>(NI_, opt_.diag_gmin));
    set_dev_gmin(std::vector<double>(NI_, opt_.gmin));
    for (int i = 0; i < NI_; ++i) {
        if (!failing[i] || phase[i] != L_DONE) continue;
        converged[i] = 1;
        iterations[i] = total[i];
        stage[i] = device ? LADDER_TRUE_GMIN : LADDER_GMIN;
    }
}

// Source stepping (convergence.cpp:472-582) for every instance whose stage
// is still LADDER_NONE: from zeros, a solve with every independent source
// at zero (src_fact 0; if that fails, diagonal-gmin steps 1e10 gmin ...
// gmin at zero sources), then src_fact raised to 1 in adaptive steps
// (0.001 to start, x1.5 after an easy solve, /2 after a hard one, /10 and
// a retry from the last accepted point after a failure).  base_mode is
// the analysis's mode without its init bits.
void Bundle::ladder_source(int base_mode, std::vector<int32_t>& converged, std::vector<int32_t>& iterations,
                           std::vector<int32_t>& stage) {
    std::vector<int32_t> want(NI_, 0);
    bool any = false;
    for (int i = 0; i < NI_; ++i) {
        want[i] = stage[i] == LADDER_NONE;
        any |= want[i] != 0;
    }
    if (!any) return;
    double* cur[4] = {d_x_, d_s_[0], d_s_[1], d_s_[2]};
    double* acc[4] = {d_lad_[4], d_lad_[5], d_lad_[6], d_lad_[7]};
    masked_copy_state(want, nullptr, cur);   // zero solution, clear_state
    const int max_iter = opt_.itl1;           // SimOptions::max_iter
    const double min_step = 1e-7;
    enum { S_ZERO = 0, S_ZG, S_STEP, S_DONE, S_FAILED, S_IDLE };
    std::vector<int> phase(NI_, S_IDLE), mode(NI_, base_mode | MODEINITJCT), total(NI_, 0), zg_i(NI_, 0);
    std::vector<int> first_its(NI_, 0);
    std::vector<double> fraction(NI_, 0.0), step(NI_, 0.001), next(NI_, 0.0), src(NI_, 0.0), zdiag(NI_, 0.0);
    const double zg0 = std::max(opt_.gmin, 0.0) == 0.0 ? opt_.gmin : std::max(opt_.gmin, 0.0);
    for (int i = 0; i < NI_; ++i)
        if (want[i]) phase[i] = S_ZERO;
    std::vector<int32_t> conv, its;
    for (int round = 0; round < 1000000; ++round) {
        std::vector<int32_t> active(NI_, 0);
        std::vector<double> g(NI_, opt_.diag_gmin);
        std::vector<DevCtx> ctx = h_ctx_;
        bool live = false;
        for (int i = 0; i < NI_; ++i) {
            if (phase[i] != S_ZERO && phase[i] != S_ZG && phase[i] != S_STEP) continue;
            active[i] = 1;
            live = true;
            ctx[i].mode = mode[i];
            ctx[i].src_fact = src[i];
            if (phase[i] == S_ZG) g[i] = zdiag[i];
        }
        if (!live) break;
        set_ctx(ctx);
        set_gmin(g);
        newton(max_iter, active.data(), conv, its);
        std::vector<int32_t> save(NI_, 0), back(NI_, 0), zero(NI_, 0);
        for (int i = 0; i < NI_; ++i) {
            if (!active[i]) continue;
            if (phase[i] == S_ZERO) {
                first_its[i] = its[i];
                if (conv[i]) {
                    total[i] += its[i];
                    save[i] = 1;
                    mode[i] = base_mode | MODEINITFLOAT;
                    phase[i] = S_STEP;
                } else {
                    // gmin stepping at zero sources from 1e10 x gmin
                    double d = zg0;
                    for (int k = 0; k < 10; ++k) d *= 10.0;
                    zdiag[i] = d;
                    zg_i[i] = 0;
                    zero[i] = 1;
                    mode[i] = base_mode | MODEINITJCT;
                    phase[i] = S_ZG;
                }
                continue;
            }
            if (phase[i] == S_ZG) {
                if (!conv[i]) {
                    phase[i] = S_FAILED;
                    continue;
                }
                total[i] += its[i];
                zdiag[i] /= 10.0;
                mode[i] = base_mode | MODEINITFLOAT;
                if (zg_i[i]++ == 10) {
                    // the failed zero-source solve's iterations count too
                    total[i] += first_its[i];
                    save[i] = 1;
                    phase[i] = S_STEP;
                }
                continue;
            }
            // S_STEP: the solve at src[i] = next[i]
            if (conv[i]) {
                total[i] += its[i];
                fraction[i] = next[i];
                save[i] = 1;
                if (its[i] < max_iter / 4) step[i] = std::min(0.1, step[i] * 1.5);
                else if (its[i] > max_iter / 2) step[i] = std::max(min_step, step[i] * 0.5);
                if (!(fraction[i] < 1.0)) phase[i] = S_DONE;
            } else {
                if (step[i] * (1.0 - fraction[i]) < 1e-8) {
                    phase[i] = S_FAILED;
                    continue;
                }
                back[i] = 1;
                step[i] *= 0.1;
                if (step[i] > 0.01) step[i] = 0.01;
                if (step[i] < min_step) phase[i] = S_FAILED;
            }
        }
        masked_copy_state(zero, nullptr, cur);
        masked_copy_state(save, cur, acc);
        masked_copy_state(back, acc, cur);
        // the next solve of every stepping instance starts from its accepted point
        for (int i = 0; i < NI_; ++i) {
            if (phase[i] != S_STEP) continue;
            next[i] = std::min(1.0, fraction[i] + step[i]);
            src[i] = next[i];
        }
        std::vector<int32_t> from_acc(NI_, 0);
        for (int i = 0; i < NI_; ++i) from_acc[i] = (phase[i] == S_STEP && !back[i] && !save[i]) ? 1 : 0;
        masked_copy_state(from_acc, acc, cur);
    }
    set_gmin(std::vector<double>(NI_, opt_.diag_gmin));
    std::vector<DevCtx> ctx = h_ctx_;
    for (int i = 0; i < NI_; ++i) {
        ctx[i].src_fact = opt_.src_fact;
        if (!want[i] || phase[i] != S_DONE) continue;
        converged[i] = 1;
        iterations[i] = total[i];
        stage[i] = LADDER_SOURCE;
    }
    set_ctx(ctx);
}

void Bundle::dc_op(int mode, std::vector<int32_t>& converged, std::vector<int32_t>& iterations) {
    std::vector<DevCtx> c(NI_);
    for (int i = 0; i < NI_; ++i) {
        DevCtx x{};
        x.mode = mode;
        x.order = 1;
        x.method = opt_.method;
        x.xmu_ratio = 1.0;
        x.src_fact = opt_.src_fact;
        c[i] = x;
    }
    set_ctx(c);
    newton(opt_.itl1, nullptr, converged, iterations);
}


// ---------------------------------------------------------------- transient

namespace {

// The controller's breakpoint sets while they are built (TimeStepController::
// add_breakpoint with CKTminBreak, src/core/timestep.cpp).
struct BpSets {
    double time = 0.0, tstop = 0.0, min_break = 0.0;
    std::set<double> breakpoints;
    std::map<double, bool> src;  // time -> HARD
};

void add_source_breakpoint(BpSets& c, double t, bool hard) {
    if (t > c.time && t <= c.tstop) {
        // CKTminBreak: a breakpoint this close to an existing one only promotes it
        if (c.min_break > 0.0) {
            auto lb = c.src.lower_bound(t - c.min_break);
            if (lb != c.src.end() && lb->first <= t + c.min_break) {
                if (hard) lb->second = true;
                return;
            }
        }
        c.breakpoints.insert(t);
        auto it = c.src.find(t);
        if (it != c.src.end()) {
            if (hard) it->second = true;
        } else {
            c.src.emplace(t, hard);
        }
    }
}

// VSource::get_breakpoints (vsource.cpp:233-264) of a PULSE (resolved
// parameters) or PWL source.
void source_breakpoints(const Device& d, const double* pulse, double tstop, std::vector<double>& out) {
    out.clear();
    if (d.wave == WAVE_PWL) {
        for (double tp : d.pwl_t)
            if (tp > 0.0 && tp <= tstop) out.push_back(tp);
        return;
    }
    if (d.wave != WAVE_PULSE) return;
    const double td = pulse[2], tr = pulse[3], tf = pulse[4], pw = pulse[5], per = pulse[6];
    if (per <= 0.0) return;
    const int max_periods = static_cast<int>((tstop - td) / per) + 2;
    for (int k = 0; k < max_periods; ++k) {
        const double t0 = td + k * per;
        if (t0 > tstop) break;
        const double edges[] = {t0, t0 + tr, t0 + tr + pw, t0 + tr + pw + tf};
        for (double e : edges)
            if (e > 0.0 && e <= tstop) out.push_back(e);
    }
}

}  // namespace

// The packed LU's schedule by timing (the autotuner): every candidate
// schedule (the serial-phase threshold and LDS staging, SluOpts) rebuilds
// the plans and times a few factor-and-solve launches of the whole batch on
// its current matrices into scratch buffers; the fastest is kept.  Every
// schedule computes the same results, so the choice changes speed only.
// A process-wide cache keyed by the plans and the batch size reuses a
// decision (the bench's warm-up and timed bundles share one).
void Bundle::tune_lu() {
    if (NI_ <= 0 || plans_.count() == 0) return;
    static std::map<uint64_t, std::pair<int, std::vector<double>>> cache;
    const SluOpts cand[] = {{2, true}, {4, true}, {8, true}, {32, true}, {1 << 30, true},
                            {2, false}, {4, false}, {8, false}, {32, false}, {1 << 30, false}};
    const int nc = (int)(sizeof(cand) / sizeof(cand[0]));
    plans_.rebuild(SluOpts{});
    uint64_t key = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
    for (int32_t w : plans_.host_words()) mix((uint32_t)w);
    for (int32_t w : plans_.host_offsets()) mix((uint32_t)w);
    mix((uint64_t)NI_);
    std::vector<double> us(nc, -1.0);
    int best = -1;
    auto it = cache.find(key);
    if (it != cache.end()) {
        best = it->second.first;
        us = it->second.second;
    } else {
        double *lu = nullptr, *x = nullptr;
        int32_t* fail = nullptr;
        NSA_HIP(hipMalloc(&lu, sizeof(double) * (size_t)NI_ * kSluMaxLu));
        NSA_HIP(hipMalloc(&x, sizeof(double) * std::max<size_t>(1, (size_t)tot_n_)));
        NSA_HIP(hipMalloc(&fail, sizeof(int32_t) * NI_));
        hipEvent_t e0, e1;
        NSA_HIP(hipEventCreate(&e0));
        NSA_HIP(hipEventCreate(&e1));
        constexpr int kReps = 5;
        for (int c = 0; c < nc; ++c) {
            plans_.rebuild(cand[c]);
            plans_.sync(stream_);
            double bestc = 1e30;
            for (int r = 0; r <= kReps; ++r) {
                NSA_HIP(hipEventRecord(e0, stream_));
                launch_small_lu(plans_, d_inst_, NI_, SLU_FACTOR | SLU_SOLVE, d_A_, opt_.diag_gmin, nullptr, fail,
                                nullptr, lu, d_rhs_, x, stream_, d_gmin_);
                NSA_HIP(hipEventRecord(e1, stream_));
                NSA_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                NSA_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) bestc = std::min(bestc, 1e3 * (double)ms);  // the first launch warms up
            }
            us[c] = bestc;
            if (best < 0 || us[c] < us[best]) best = c;
        }
        NSA_HIP(hipEventDestroy(e0));
        NSA_HIP(hipEventDestroy(e1));
        NSA_HIP(hipFree(lu));
        NSA_HIP(hipFree(x));
        NSA_HIP(hipFree(fail));
        cache[key] = {best, us};
    }
    plans_.rebuild(cand[best]);
    plans_.sync(stream_);
    tune_report.assign(1, (double)best);
    tune_report.push_back((double)nc);
    for (int c = 0; c < nc; ++c) {
        tune_report.push_back((double)cand[c].min_serial_run);
        tune_report.push_back(cand[c].stage ? 1.0 : 0.0);
        tune_report.push_back(us[c]);
    }
}

void Bundle::tran_begin(double tstep, double tstop, const int32_t* probe_ptr, const int32_t* probes,
                        int max_points) {
    if (!(tstep > 0.0) || !(tstop > 0.0))
        throw Error(NSA_ERR_ARGUMENT, "Bundle::tran_begin: tstep and tstop must be positive");
    tstep_ = tstep;
    tstop_ = tstop;
    NSA_HIP(hipStreamSynchronize(stream_));
    for (void* p : tran_allocs_) NSA_HIP(hipFree(p));
    tran_allocs_.clear();
    // the tick graph names the buffers freed above
    if (tick_exec_) { NSA_HIP(hipGraphExecDestroy(tick_exec_)); tick_exec_ = nullptr; }
    if (tick_graph_) { NSA_HIP(hipGraphDestroy(tick_graph_)); tick_graph_ = nullptr; }
    tcd_.clear();
    tc_on_device_ = false;
    log_on_device_ = false;
    d_log_ = nullptr;
    d_spec_ = nullptr;
    log_rows_ = 0;
    // 1. DC operating point in MODETRANOP from zeros (transient.cpp:710-726, 195-205)
    std::vector<double> z(tot_n_, 0.0);
    set_solution(z.data());
    std::vector<int32_t> conv, its;
    op_ladder(MODETRANOP | MODEINITJCT, MODETRANOP | MODEINITFLOAT, conv, its, dc_stage_);
    dc_iterations_ = its;
    for (int i = 0; i < NI_; ++i)
        if (!conv[i])
            throw Error(NSA_ERR_RUNTIME, "Bundle::tran_begin: DC operating point did not converge for instance " +
                                             std::to_string(i));
    // the packed LU's schedule, timed on the operating point's matrices
    if (tuning) {
        if (inst_dirty_) upload_inst();
        tune_lu();
    }
    // 2. a fresh solver for the transient (transient.cpp:728-730); state
    //    history seeded from the operating point (transient.cpp:791-802)
    reset_solver();
    for (int k = 1; k < 4; ++k) launch_copy_f64(d_s_[k], d_s_[0], ns_tot_, stream_);
    // 3. PULSE defaults (VSource::resolve_defaults, vsource.cpp:51-58)
    std::vector<double> pulse = h_pulse_;
    for (size_t v = 0; v < pulse.size() / 7; ++v) {
        double* p = &pulse[7 * v];
        if (p[3] <= 0) p[3] = tstep;
        if (p[4] <= 0) p[4] = tstep;
        if (p[5] <= 0) p[5] = tstop;
        if (p[6] <= 0) p[6] = tstop;
    }
    if (!pulse.empty())
        NSA_HIP(hipMemcpyAsync(const_cast<double*>(vsrc_.pulse), pulse.data(), sizeof(double) * pulse.size(),
                               hipMemcpyHostToDevice, stream_));
    // 4. history ring, LTE and ringing buffers, probes
    auto tal = [&](size_t bytes) {
        void* p = nullptr;
        NSA_HIP(hipMalloc(&p, std::max<size_t>(bytes, 8)));
        tran_allocs_.push_back(p);
        return p;
    };
    d_h1_ = (double*)tal(sizeof(double) * tot_n_);
    d_h2_ = (double*)tal(sizeof(double) * tot_n_);
    d_h3_ = (double*)tal(sizeof(double) * tot_n_);
    launch_copy_f64(d_h1_, d_x_, tot_n_, stream_);
    launch_copy_f64(d_h2_, d_x_, tot_n_, stream_);
    launch_copy_f64(d_h3_, d_x_, tot_n_, stream_);
    d_tmin_ = (unsigned long long*)tal(sizeof(unsigned long long) * 2 * NI_);
    d_ring_ = (int32_t*)tal(sizeof(int32_t) * NI_);
    d_acc_ = (int32_t*)tal(sizeof(int32_t) * NI_);
    d_att_ = (int32_t*)tal(sizeof(int32_t) * NI_);
    d_prow_ = (int32_t*)tal(sizeof(int32_t) * NI_);
    d_tc_ = (TranDev*)tal(sizeof(TranDev) * NI_);
    d_spec_ = (int32_t*)tal(sizeof(int32_t) * NI_);
    d_tcs_ = (TcStatus*)tal(sizeof(TcStatus));
    if (!h_acc_pin_) {
        NSA_HIP(hipHostMalloc((void**)&h_acc_pin_, sizeof(int32_t) * 3 * NI_, hipHostMallocDefault));
        NSA_HIP(hipHostMalloc((void**)&h_tmin_pin_, sizeof(unsigned long long) * 2 * NI_, hipHostMallocDefault));
        NSA_HIP(hipHostMalloc((void**)&h_ring_pin_, sizeof(int32_t) * NI_, hipHostMallocDefault));
        NSA_HIP(hipHostMalloc((void**)&h_tcs_pin_, sizeof(TcStatus), hipHostMallocDefault));
    }
    pptr_.assign(NI_ + 1, 0);
    std::vector<int32_t> pr;
    if (probe_ptr) {
        for (int i = 0; i <= NI_; ++i) pptr_[i] = probe_ptr[i];
        pr.assign(probes, probes + pptr_[NI_]);
        for (int i = 0; i < NI_; ++i)
            for (int q = pptr_[i]; q < pptr_[i + 1]; ++q)
                if (pr[q] < 0 || pr[q] >= n_[i]) throw Error(NSA_ERR_ARGUMENT, "Bundle::tran_begin: probe index");
    }
    max_points_ = std::max(1, max_points);
    d_pptr_ = (int32_t*)tal(sizeof(int32_t) * (NI_ + 1));
    NSA_HIP(hipMemcpyAsync(d_pptr_, pptr_.data(), sizeof(int32_t) * (NI_ + 1), hipMemcpyHostToDevice, stream_));
    d_probe_ = (int32_t*)tal(sizeof(int32_t) * std::max<size_t>(1, pr.size()));
    if (!pr.empty())
        NSA_HIP(hipMemcpyAsync(d_probe_, pr.data(), sizeof(int32_t) * pr.size(), hipMemcpyHostToDevice, stream_));
    d_pvals_ = pr.empty() ? nullptr : (double*)tal(sizeof(double) * pr.size() * max_points_);
    // the t = 0 point (transient.cpp:805)
    prow_.assign(NI_, 0);
    std::vector<int32_t> rec(NI_, 3);
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(h_acc_pin_, rec.data(), sizeof(int32_t) * NI_);
    std::memcpy(h_acc_pin_ + NI_, prow_.data(), sizeof(int32_t) * NI_);
    NSA_HIP(hipMemcpyAsync(d_acc_, h_acc_pin_, sizeof(int32_t) * NI_, hipMemcpyHostToDevice, stream_));
    NSA_HIP(hipMemcpyAsync(d_prow_, h_acc_pin_ + NI_, sizeof(int32_t) * NI_, hipMemcpyHostToDevice, stream_));
    launch_tr_hist(NI_, d_ixoff_, d_in_, d_x_, d_h1_, d_h2_, d_h3_, d_acc_, d_pptr_, d_probe_, d_prow_,
                   d_pvals_, max_points_, stream_);
    // 5. step controllers (transient.cpp:826-903)
    tcd_.assign(NI_, TranDev{});
    hsteps_.assign(NI_, {});
    bpt_.clear();
    bph_.clear();
    const double max_step = std::min(tstep, tstop / 50.0);
    std::vector<double> bps;
    int vcur = 0;
    for (int i = 0; i < NI_; ++i) {
        TranDev& c = tcd_[i];
        c.order = 1;
        c.last_bp_hard = 1;
        c.steps_after_bp = kNoRecentBreakpoint;
        c.dt_min = max_step * kMinTimeStepRatio;
        c.dt_max = max_step;
        c.ctrl_dt = tstep;
        c.tstop = tstop;
        c.time = 0.0;
        c.proposed_dt = tstep;
        c.min_break = (max_step > 0.0) ? 5e-5 * max_step : 0.0;
        c.xmu = opt_.xmu;
        c.src_fact = opt_.src_fact;
        c.soft_scale = std::sqrt(kRestartStepScale);
        c.itl = opt_.itl4;
        c.phase = TP_IDLE;
        // collect_breakpoints (transient.cpp:135-151): sources in device order
        BpSets b;
        b.time = c.time;
        b.tstop = tstop;
        b.min_break = c.min_break;
        const Circuit& ck = *ckt_[i];
        const auto& fv = ck.family(DK_V);
        std::vector<int> vfam_pos(ck.devices().size(), -1);
        for (size_t k = 0; k < fv.size(); ++k) vfam_pos[fv[k]] = (int)k;
        for (size_t d = 0; d < ck.devices().size(); ++d) {
            const Device& dv = ck.devices()[d];
            if (dv.kind != DK_V) continue;
            source_breakpoints(dv, &pulse[7 * (size_t)(vcur + vfam_pos[d])], tstop, bps);
            for (double bp : bps) add_source_breakpoint(b, bp, true);
        }
        vcur += (int)fv.size();
        // both sets hold the same times (sources are the only producers)
        if (b.breakpoints.size() != b.src.size())
            throw Error(NSA_ERR_LOGIC, "Bundle::tran_begin: breakpoint sets differ");
        c.bp_head = (int32_t)bpt_.size();
        for (const auto& kv : b.src) {
            if (!b.breakpoints.count(kv.first))
                throw Error(NSA_ERR_LOGIC, "Bundle::tran_begin: breakpoint sets differ");
            bpt_.push_back(kv.first);
            bph_.push_back(kv.second ? 1 : 0);
        }
        c.bp_end = (int32_t)bpt_.size();
        c.dt = std::min(tstop / kInitialStopDivisor, tstep) / kInitialStepDivisor;
        c.use_gear = opt_.method == 1;
        c.integrate_method = c.use_gear ? 1 : 0;
        c.saved_delta = c.dt;
    }
    for (int i = 0; i < NI_; ++i) {
        TranDev& c = tcd_[i];
        const double bp_gap = tc_next_bp_gap(c, bpt_.data());
        if (bp_gap < tstop) {
            const double init_save_delta = tstop / 50.0;
            c.dt = std::min(c.dt, 0.1 * std::min(init_save_delta, bp_gap));
            c.dt /= 10;
            c.dt = std::max(c.dt, c.dt_min * 2.0);
        }
        c.saved_delta = c.dt;
        c.prev_prev_dt = c.dt;
    }
    d_bpt_ = (double*)tal(sizeof(double) * std::max<size_t>(1, bpt_.size()));
    d_bph_ = (int32_t*)tal(sizeof(int32_t) * std::max<size_t>(1, bph_.size()));
    if (!bpt_.empty()) {
        NSA_HIP(hipMemcpyAsync(d_bpt_, bpt_.data(), sizeof(double) * bpt_.size(), hipMemcpyHostToDevice, stream_));
        NSA_HIP(hipMemcpyAsync(d_bph_, bph_.data(), sizeof(int32_t) * bph_.size(), hipMemcpyHostToDevice, stream_));
    }
    NSA_HIP(hipStreamSynchronize(stream_));
}

bool Bundle::tran_done() const {
    for (const auto& c : tcd_)
        if (!c.finished) return false;
    return true;
}

std::vector<TranStep> Bundle::tran_steps(int i) const {
    if (i < 0 || i >= NI_) throw Error(NSA_ERR_ARGUMENT, "Bundle::tran_steps: instance");
    if (tcd_.empty()) return {};
    if (!log_on_device_) return hsteps_[i];
    const int rows = std::min(tcd_[i].step_count, log_rows_);
    std::vector<TranStep> out(rows);
    if (rows)
        NSA_HIP(hipMemcpyAsync(out.data(), d_log_ + (size_t)i * log_rows_, sizeof(TranStep) * rows,
                               hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    return out;
}

int Bundle::tran_advance(double t_until) {
    if (tcd_.empty()) throw Error(NSA_ERR_LOGIC, "Bundle::tran_advance: tran_begin not called");
    return device_control ? tran_advance_device(t_until) : tran_advance_host(t_until);
}

// Host control: every unfinished instance attempts one step per round; the
// round's Newton solves run together, then the host takes every decision.
int Bundle::tran_advance_host(double t_until) {
    if (log_on_device_)
        throw Error(NSA_ERR_LOGIC, "Bundle::tran_advance: a transient under device control stays there");
    if (tc_on_device_) {
        NSA_HIP(hipMemcpyAsync(tcd_.data(), d_tc_, sizeof(TranDev) * NI_, hipMemcpyDeviceToHost, stream_));
        NSA_HIP(hipStreamSynchronize(stream_));
        tc_on_device_ = false;
    }
    int rounds = 0;
    std::vector<int32_t> conv, its;
    const TruncOpts to{opt_.abstol, opt_.reltol, opt_.chgtol, opt_.trtol};
    const double big = 1e30;
    unsigned long long big_bits;
    std::memcpy(&big_bits, &big, sizeof(big));
    for (auto& c : tcd_) c.t_until = t_until;
    for (;;) {
        std::vector<int32_t> att(NI_, 0);
        std::vector<DevCtx> ctx = h_ctx_;
        bool any = false;
        for (int i = 0; i < NI_; ++i) {
            DevCtx x;
            if (!tc_attempt(tcd_[i], bpt_.data(), x)) continue;
            ctx[i] = x;
            att[i] = 1;
            any = true;
        }
        if (!any) break;
        set_ctx(ctx);
        newton(opt_.itl4, att.data(), conv, its);
        // every device's compute_trunc at the step's order and at order 2
        // (evaluate_lte, transient.cpp:469-485; order promotion :946-958),
        // and the ringing test, for the instances that attempted a step
        launch_fill_u64(d_tmin_, big_bits, 2 * NI_, stream_);
        launch_trunc(trunc1_, d_ctx_, d_s_[0], d_s_[1], d_s_[2], d_s_[3], 1, to, nullptr, d_tmin_, stream_, 0,
                     d_tmin_ + NI_);
        launch_trunc(trunc4_, d_ctx_, d_s_[0], d_s_[1], d_s_[2], d_s_[3], b4_lanes_, to, nullptr, d_tmin_,
                     stream_, 0, d_tmin_ + NI_);
        std::memcpy(h_acc_pin_ + 2 * NI_, att.data(), sizeof(int32_t) * NI_);
        NSA_HIP(hipMemcpyAsync(d_att_, h_acc_pin_ + 2 * NI_, sizeof(int32_t) * NI_, hipMemcpyHostToDevice,
                               stream_));
        launch_tr_ring(NI_, d_ixoff_, d_inn_, d_x_, d_h1_, d_h2_, opt_.reltol, opt_.vntol, d_att_, d_ring_,
                       stream_);
        NSA_HIP(hipMemcpyAsync(h_tmin_pin_, d_tmin_, sizeof(unsigned long long) * 2 * NI_,
                               hipMemcpyDeviceToHost, stream_));
        NSA_HIP(hipMemcpyAsync(h_ring_pin_, d_ring_, sizeof(int32_t) * NI_, hipMemcpyDeviceToHost, stream_));
        NSA_HIP(hipStreamSynchronize(stream_));
        std::vector<int32_t> acc(NI_, 0);
        for (int i = 0; i < NI_; ++i) {
            if (!att[i]) continue;
            double tmin_k, tmin_2;
            std::memcpy(&tmin_k, &h_tmin_pin_[i], sizeof(double));
            std::memcpy(&tmin_2, &h_tmin_pin_[NI_ + i], sizeof(double));
            TranStep rec{};
            acc[i] = tc_decide(tcd_[i], bpt_.data(), bph_.data(), conv[i], its[i], tmin_k, tmin_2,
                               h_ring_pin_[i], h_ctx_[i].order, h_ctx_[i].method, rec);
            if (acc[i] == 1) {
                hsteps_[i].push_back(rec);
                prow_[i] = tcd_[i].step_count;
            }
        }
        // rotate accepted instances, restore the others (transient.cpp:919-975, 1004, 1027-1037)
        std::memcpy(h_acc_pin_, acc.data(), sizeof(int32_t) * NI_);
        std::memcpy(h_acc_pin_ + NI_, prow_.data(), sizeof(int32_t) * NI_);
        NSA_HIP(hipMemcpyAsync(d_acc_, h_acc_pin_, sizeof(int32_t) * NI_, hipMemcpyHostToDevice, stream_));
        NSA_HIP(hipMemcpyAsync(d_prow_, h_acc_pin_ + NI_, sizeof(int32_t) * NI_, hipMemcpyHostToDevice,
                               stream_));
        launch_tr_states(ns_tot_, ns_lane_, b4_lanes_, b4_.inst_of, d_slot_inst_, d_acc_, d_s_[0], d_s_[1],
                         d_s_[2], d_s_[3], stream_);
        launch_tr_hist(NI_, d_ixoff_, d_in_, d_x_, d_h1_, d_h2_, d_h3_, d_acc_, d_pptr_, d_probe_, d_prow_,
                       d_pvals_, max_points_, stream_);
        ++rounds;
    }
    NSA_HIP(hipStreamSynchronize(stream_));
    return rounds;
}

// One transient tick.  Two launches carry independent work in extra
// workgroups, on SIMDs their own one-wave-per-instance grids leave idle:
// nt_begin carries the BSIM4v7 (lane-strided) state rotation the previous
// tick's accepted steps owe (acc), and the packed LU carries the BSIM4v7
// step proposals (compute_trunc) of the iterations that could end a
// converged solve (spec).  After the LU come the other families'
// proposals (att), the step decisions and their history (tc_step), and the
// other families' rotation (acc).  With `timed`, HIP events bracket the
// iteration's phases (loads, LU, end of iteration).
void Bundle::launch_tick(bool timed) {
    const TruncOpts to{opt_.abstol, opt_.reltol, opt_.chgtol, opt_.trtol};
    const RotSide rs{b4_lanes_ > 0 ? (int)(ns_lane_ / b4_lanes_) : 0, b4_lanes_, b4_.inst_of, d_acc_, d_s_[0],
                     d_s_[1], d_s_[2], d_s_[3]};
    if (timed) NSA_HIP(hipEventRecord(ev_[0], stream_));
    launch_nt_begin(d_nt_, NI_, d_ixoff_, d_in_, d_x_, d_xold_, d_ctx_, d_kind_, d_fail_, d_devnc_, d_done_,
                    stream_, d_tmin_, bits_1e30(), d_spec_, &rs);
    launch_loads_small();
    launch_loads_b4_gather();
    const SluSide side{trunc4_, d_ctx_, d_s_[0], d_s_[1], d_s_[2], d_s_[3], b4_lanes_, to, d_tmin_, NI_, d_spec_};
    // a timed tick runs the BSIM4v7 step proposals as a launch of their own,
    // so that the LU launch's events (and counters) cover the LU alone
    if (timed)
        launch_trunc(trunc4_, d_ctx_, d_s_[0], d_s_[1], d_s_[2], d_s_[3], b4_lanes_, to, nullptr, d_tmin_, stream_,
                     0, d_tmin_ + NI_, d_spec_);
    if (timed) NSA_HIP(hipEventRecord(ev_[3], stream_));
    launch_small_lu(plans_, d_inst_, NI_, SLU_FACTOR | SLU_SOLVE, d_A_, opt_.diag_gmin, d_kind_, d_fail_,
                    d_fail_col_, d_LU_, d_rhs_, d_x_, stream_, d_gmin_, timed ? nullptr : &side);
    if (timed) NSA_HIP(hipEventRecord(ev_[4], stream_));
    launch_post(true);
    if (timed) NSA_HIP(hipEventRecord(ev_[5], stream_));
    launch_trunc(trunc1_, d_ctx_, d_s_[0], d_s_[1], d_s_[2], d_s_[3], 1, to, nullptr, d_tmin_, stream_, 0,
                 d_tmin_ + NI_, d_att_);
    launch_tc_step(d_tc_, d_bpt_, d_bph_, d_nt_, d_ctx_, d_att_, d_tmin_, NI_, d_ixoff_, d_in_, d_inn_, d_x_,
                   d_h1_, d_h2_, d_h3_, opt_.reltol, opt_.vntol, d_acc_, d_log_, log_rows_, d_pptr_, d_probe_,
                   d_pvals_, max_points_, stream_);
    launch_tr_states(ns_tot_, ns_lane_, b4_lanes_, b4_.inst_of, d_slot_inst_, d_acc_, d_s_[0], d_s_[1],
                     d_s_[2], d_s_[3], stream_, ns_lane_);
}

// The BSIM4v7 rotations the last tick's accepted steps still owe, so the
// states are current when control returns to the caller; acc is cleared
// so that the next tick owes none.
void Bundle::flush_rotations() {
    launch_tr_states(ns_tot_, ns_lane_, b4_lanes_, b4_.inst_of, d_slot_inst_, d_acc_, d_s_[0], d_s_[1], d_s_[2],
                     d_s_[3], stream_, 0, ns_lane_);
    launch_fill_i32(d_acc_, 0, NI_, stream_);
}

void Bundle::capture_tick_graph() {
    if (tick_exec_) { NSA_HIP(hipGraphExecDestroy(tick_exec_)); tick_exec_ = nullptr; }
    if (tick_graph_) { NSA_HIP(hipGraphDestroy(tick_graph_)); tick_graph_ = nullptr; }
    NSA_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    launch_tick(false);
    NSA_HIP(hipStreamEndCapture(stream_, &tick_graph_));
    NSA_HIP(hipGraphInstantiate(&tick_exec_, tick_graph_, nullptr, nullptr, 0));
    tick_words_ = plans_.d_words();
    tick_nlu_ = plans_.max_nlu();
    tick_n_ = plans_.max_n();
    tick_hot_ = plans_.max_hot();
    tick_log_ = d_log_;
}

// One tick: every instance in a Newton solve runs one iteration, and every
// instance whose solve ended takes its step decision and sets up its next
// step.
void Bundle::tick_device() {
    if (inst_dirty_) upload_inst();
    if (timing_) {
        launch_tick(true);
        timed_pending_ = true;
    } else if (use_graph) {
        if (!tick_exec_ || graph_stale(tick_words_, tick_nlu_, tick_n_, tick_hot_) || tick_log_ != d_log_)
            capture_tick_graph();
        NSA_HIP(hipGraphLaunch(tick_exec_, stream_));
        ++n_graph_launches;
    } else {
        launch_tick(false);
    }
    ++n_ticks;
}

// Room in the device step logs for `rows` accepted steps per instance
// (stream-ordered; the logs keep what they hold).
void Bundle::tran_log_reserve(int rows) {
    if (rows <= log_rows_) return;
    const int cap = std::max(rows, std::max(2 * log_rows_, 1024));
    TranStep* nl = nullptr;
    NSA_HIP(hipStreamSynchronize(stream_));
    NSA_HIP(hipMalloc(&nl, sizeof(TranStep) * (size_t)cap * NI_));
    if (d_log_ && log_rows_)
        NSA_HIP(hipMemcpy2DAsync(nl, sizeof(TranStep) * cap, d_log_, sizeof(TranStep) * log_rows_,
                                 sizeof(TranStep) * log_rows_, NI_, hipMemcpyDeviceToDevice, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    for (auto it = tran_allocs_.begin(); it != tran_allocs_.end(); ++it)
        if (*it == d_log_) {
            NSA_HIP(hipFree(d_log_));
            tran_allocs_.erase(it);
            break;
        }
    tran_allocs_.push_back(nl);
    d_log_ = nl;
    log_rows_ = cap;
}

// Instances waiting in a Newton iteration for the host: their matrices are
// ordered on the CPU (a fresh solver's first factorization, or a reused
// order that failed its threshold test) and their iteration ends on the
// device; no tick is in flight.
void Bundle::handle_stalls() {
    std::vector<NtInst> nt(NI_);
    NSA_HIP(hipMemcpyAsync(h_nt_pin_, d_nt_, sizeof(NtInst) * NI_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(nt.data(), h_nt_pin_, sizeof(NtInst) * NI_);
    std::vector<int> stalled;
    for (int i = 0; i < NI_; ++i)
        if (!nt[i].done && nt[i].stalled) stalled.push_back(i);
    if (stalled.empty()) return;
    if (inst_dirty_) upload_inst();
    host_lu_post(nt, stalled, false);
}

// Device control: ticks in batches sized from the work left in the window;
// between batches the host reads a status summary and serves stalled
// instances.
int Bundle::tran_advance_device(double t_until) {
    if (!log_on_device_ && !hsteps_.empty()) {
        // steps taken under host control move into the device logs
        int most = 0;
        for (const auto& v : hsteps_) most = std::max(most, (int)v.size());
        tran_log_reserve(most + 64);
        for (int i = 0; i < NI_; ++i)
            if (!hsteps_[i].empty())
                NSA_HIP(hipMemcpyAsync(d_log_ + (size_t)i * log_rows_, hsteps_[i].data(),
                                       sizeof(TranStep) * hsteps_[i].size(), hipMemcpyHostToDevice, stream_));
        NSA_HIP(hipStreamSynchronize(stream_));
        log_on_device_ = true;
    }
    if (!tc_on_device_) {
        // ticks rotate the BSIM4v7 states the previous tick's acc names: none
        // are owed on entry (host control rotates at once)
        launch_fill_i32(d_acc_, 0, NI_, stream_);
        NSA_HIP(hipMemcpyAsync(d_tc_, tcd_.data(), sizeof(TranDev) * NI_, hipMemcpyHostToDevice, stream_));
        NSA_HIP(hipStreamSynchronize(stream_));
        tc_on_device_ = true;
    }
    int64_t it0 = 0;
    for (const auto& c : tcd_) it0 += c.all_iters;
    launch_tc_resume(d_tc_, d_bpt_, d_nt_, d_ctx_, NI_, t_until, stream_);
    int ticks = 0;
    for (;;) {
        launch_tc_status(d_tc_, d_nt_, NI_, d_tcs_, stream_);
        NSA_HIP(hipMemcpyAsync(h_tcs_pin_, d_tcs_, sizeof(TcStatus), hipMemcpyDeviceToHost, stream_));
        const auto tw = std::chrono::steady_clock::now();
        NSA_HIP(hipStreamSynchronize(stream_));
        tick_wait_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - tw).count();
        ++n_tick_batches;
        collect_timing();
        const TcStatus st = *h_tcs_pin_;
        if (st.stalled > 0) {
            handle_stalls();
            continue;
        }
        if (st.active == 0) break;
        // at least two iterations per step (iteration 0 never converges)
        int batch = 64;
        if (st.est_steps < 32.0) batch = std::max(2, 2 * (int)std::ceil(st.est_steps));
        // a fresh solver's instances stall in their first iteration
        if (timing_ || std::find(plan_.begin(), plan_.end(), -1) != plan_.end()) batch = 1;
        // a tick accepts at most one step per instance
        tran_log_reserve(st.max_steps + batch + 1);
        const auto te = std::chrono::steady_clock::now();
        for (int k = 0; k < batch; ++k) tick_device();
        tick_enqueue_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - te).count();
        ticks += batch;
    }
    flush_rotations();
    NSA_HIP(hipMemcpyAsync(tcd_.data(), d_tc_, sizeof(TranDev) * NI_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipMemcpyAsync(h_ctx_pin_, d_ctx_, sizeof(DevCtx) * NI_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    std::memcpy(h_ctx_.data(), h_ctx_pin_, sizeof(DevCtx) * NI_);
    int64_t it1 = 0;
    for (const auto& c : tcd_) it1 += c.all_iters;
    n_iterations += (long)(it1 - it0);
    return ticks;
}

void Bundle::tran_probes(int i, int* count, double* out) const {
    if (i < 0 || i >= NI_) throw Error(NSA_ERR_ARGUMENT, "Bundle::tran_probes: instance");
    const int np = tran_probe_count(i);
    const int rows = std::min(max_points_, (tcd_.empty() ? 0 : tcd_[i].step_count) + 1);
    *count = rows;
    if (!np || !out || !d_pvals_) return;
    NSA_HIP(hipMemcpyAsync(out, d_pvals_ + (size_t)pptr_[i] * max_points_, sizeof(double) * rows * np,
                           hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------- DC adjoint

void Bundle::ensure_adjoint() {
    if (adj_ready_) return;
    std::vector<int32_t> rn, vb, vi, in, ii;
    auto abs_node = [&](int i, int node) { return node < 0 ? -1 : xoff_[i] + node; };
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        const auto& devs = c.devices();
        for (int d : c.family(DK_R)) {
            rn.push_back(abs_node(i, devs[d].node[0]));
            rn.push_back(abs_node(i, devs[d].node[1]));
            r_owner_.push_back(doff_[i] + d);
        }
        for (int d : c.family(DK_V)) {
            vb.push_back(xoff_[i] + devs[d].branch);
            vi.push_back(i);
            v_owner_.push_back(doff_[i] + d);
        }
        for (int d : c.family(DK_I)) {
            in.push_back(abs_node(i, devs[d].node[0]));
            in.push_back(abs_node(i, devs[d].node[1]));
            ii.push_back(i);
            i_owner_.push_back(doff_[i] + d);
        }
        for (int d : c.family(DK_C)) c_owner_.push_back(doff_[i] + d);
        for (int d : c.family(DK_B4)) b4_owner_.push_back(doff_[i] + d);
    }
    if ((int)r_owner_.size() != nR_ || (int)v_owner_.size() != vsrc_.count || (int)i_owner_.size() != isrc_.count ||
        (int)b4_owner_.size() != b4_lanes_)
        throw Error(NSA_ERR_RUNTIME, "Bundle adjoint: family tables");
    d_r_nodes_ = dupload(allocs_, rn);
    d_v_branch_ = dupload(allocs_, vb);
    d_v_inst_ = dupload(allocs_, vi);
    d_i_nodes_ = dupload(allocs_, in);
    d_i_inst_ = dupload(allocs_, ii);
    d_lam_ = dalloc<double>(allocs_, tot_n_);
    d_gR_ = dalloc<double>(allocs_, nR_);
    d_gV_ = dalloc<double>(allocs_, vsrc_.count);
    d_gI_ = dalloc<double>(allocs_, isrc_.count);
    d_gW_ = dalloc<double>(allocs_, b4_lanes_);
    d_gL_ = dalloc<double>(allocs_, b4_lanes_);
    const size_t nf = (size_t)kB4Fields * std::max(1, b4_lanes_);
    d_b4_dW_ = dalloc<double>(allocs_, nf);
    d_b4_dL_ = dalloc<double>(allocs_, nf);
    adj_ready_ = true;
    size_derivs_dirty_ = true;
}

// d(record)/dW and d(record)/dL of every BSIM4v7 lane at its current size,
// in the lane order and field-major layout of the load's records.
void Bundle::refresh_size_derivs() {
    if (!size_derivs_dirty_) return;
    const int G4 = b4_lanes_;
    std::vector<double> dW((size_t)kB4Fields * std::max(1, G4)), dL(dW.size());
    const DeviceEnv env{};
    std::vector<double> val(kB4Fields), w(kB4Fields), l(kB4Fields);
    int g = 0;
    for (int i = 0; i < NI_; ++i) {
        const Circuit& c = *ckt_[i];
        for (int d : c.family(DK_B4)) {
            Device dv = c.devices()[d];
            if (!b4_w_.empty()) {
                dv.w = b4_w_[g];
                dv.l = b4_l_[g];
            }
            bsim4_param_derivs(c.models().at(dv.model), dv, env, val.data(), w.data(), l.data());
            for (int q = 0; q < kB4Fields; ++q) {
                dW[(size_t)q * G4 + g] = w[q];
                dL[(size_t)q * G4 + g] = l[q];
            }
            ++g;
        }
    }
    if (g != G4) throw Error(NSA_ERR_RUNTIME, "Bundle adjoint: BSIM4v7 lanes");
    NSA_HIP(hipMemcpy(d_b4_dW_, dW.data(), sizeof(double) * dW.size(), hipMemcpyHostToDevice));
    NSA_HIP(hipMemcpy(d_b4_dL_, dL.data(), sizeof(double) * dL.size(), hipMemcpyHostToDevice));
    size_derivs_dirty_ = false;
}

void Bundle::dc_adjoint(const double* dLdx, double* lambda, double* g_value, double* g_w, double* g_l) {
    for (int i = 0; i < NI_; ++i)
        if (plan_[i] < 0) throw Error(NSA_ERR_LOGIC, "Bundle::dc_adjoint: no factors (solve the operating point first)");
    ensure_adjoint();
    refresh_size_derivs();
    NSA_HIP(hipStreamSynchronize(stream_));
    NSA_HIP(hipMemcpyAsync(d_lam_, dLdx, sizeof(double) * tot_n_, hipMemcpyHostToDevice, stream_));
    // J^T lambda = dL/dx with the factors of each instance's last iteration
    launch_small_lu(plans_, d_inst_, NI_, SLU_SOLVE_T, nullptr, 0.0, nullptr, nullptr, nullptr, d_LU_, d_lam_,
                    d_lam_, stream_);
    launch_bundle_grad_lin(nR_, d_r_nodes_, d_r_ohms_, vsrc_.count, d_v_branch_, d_v_inst_, isrc_.count,
                           d_i_nodes_, d_i_inst_, d_ctx_, d_x_, d_lam_, d_gR_, d_gV_, d_gI_, stream_);
    if (b4_lanes_) {
        launch_bsim4_grad(b4_, d_ctx_, d_x_, d_b4_dW_, d_lam_, d_gW_, tol_, b4_lanes_, stream_);
        launch_bsim4_grad(b4_, d_ctx_, d_x_, d_b4_dL_, d_lam_, d_gL_, tol_, b4_lanes_, stream_);
    }
    std::vector<double> gR(nR_), gV(vsrc_.count), gI(isrc_.count), gW(b4_lanes_), gL(b4_lanes_);
    auto down = [&](std::vector<double>& h, const double* dptr) {
        if (!h.empty())
            NSA_HIP(hipMemcpyAsync(h.data(), dptr, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream_));
    };
    down(gR, d_gR_);
    down(gV, d_gV_);
    down(gI, d_gI_);
    down(gW, d_gW_);
    down(gL, d_gL_);
    if (lambda) NSA_HIP(hipMemcpyAsync(lambda, d_lam_, sizeof(double) * tot_n_, hipMemcpyDeviceToHost, stream_));
    NSA_HIP(hipStreamSynchronize(stream_));
    const size_t nd = (size_t)doff_[NI_];
    if (g_value) {
        std::fill(g_value, g_value + nd, 0.0);
        for (size_t k = 0; k < gR.size(); ++k) g_value[r_owner_[k]] = gR[k];
        for (size_t k = 0; k < gV.size(); ++k) g_value[v_owner_[k]] = gV[k];
        for (size_t k = 0; k < gI.size(); ++k) g_value[i_owner_[k]] = gI[k];
    }
    if (g_w) {
        std::fill(g_w, g_w + nd, 0.0);
        for (size_t k = 0; k < gW.size(); ++k) g_w[b4_owner_[k]] = gW[k];
    }
    if (g_l) {
        std::fill(g_l, g_l + nd, 0.0);
        for (size_t k = 0; k < gL.size(); ++k) g_l[b4_owner_[k]] = gL[k];
    }
}

void Bundle::set_device_values(const double* value, const double* w, const double* l) {
    ensure_adjoint();
    NSA_HIP(hipStreamSynchronize(stream_));
    if (value) {
        auto put = [&](const std::vector<int32_t>& owner, const double* dptr, const char* what) {
            if (owner.empty()) return;
            std::vector<double> v(owner.size());
            for (size_t k = 0; k < owner.size(); ++k) {
                v[k] = value[owner[k]];
                if (!std::isfinite(v[k])) throw Error(NSA_ERR_ARGUMENT, std::string("set_device_values: ") + what);
            }
            NSA_HIP(hipMemcpy(const_cast<double*>(dptr), v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
        };
        for (int32_t o : r_owner_)
            if (value[o] == 0.0) throw Error(NSA_ERR_ARGUMENT, "set_device_values: zero resistance");
        put(r_owner_, d_r_ohms_, "resistance");
        put(c_owner_, cap_.farads, "capacitance");
        put(v_owner_, vsrc_.dc, "source value");
        put(i_owner_, isrc_.dc, "source value");
        // the resistors' conductances are constant matrix sites
        launch_res_sites(nR_, d_r_ohms_, d_cm_, r_site_, 1, stream_);
    }
    if ((w || l) && b4_lanes_) {
        const int G4 = b4_lanes_;
        if (b4_w_.empty()) {
            b4_w_.resize(G4);
            b4_l_.resize(G4);
            int g = 0;
            for (int i = 0; i < NI_; ++i)
                for (int d : ckt_[i]->family(DK_B4)) {
                    b4_w_[g] = ckt_[i]->devices()[d].w;
                    b4_l_[g] = ckt_[i]->devices()[d].l;
                    ++g;
                }
        }
        std::vector<double> par((size_t)kB4Fields * G4);
        NSA_HIP(hipMemcpy(par.data(), b4_.par_inst, sizeof(double) * par.size(), hipMemcpyDeviceToHost));
        const DeviceEnv env{};
        bool changed = false;
        int g = 0;
        for (int i = 0; i < NI_; ++i) {
            const Circuit& c = *ckt_[i];
            for (int d : c.family(DK_B4)) {
                const double nw = w ? w[doff_[i] + d] : b4_w_[g];
                const double nl = l ? l[doff_[i] + d] : b4_l_[g];
                if (nw != b4_w_[g] || nl != b4_l_[g]) {
                    Device dv = c.devices()[d];
                    dv.w = nw;
                    dv.l = nl;
                    const B4Param p = bsim4_param(c.models().at(dv.model), dv, env);
                    const double* f = reinterpret_cast<const double*>(&p);
                    for (int q = 0; q < kB4Fields; ++q) par[(size_t)q * G4 + g] = f[q];
                    b4_w_[g] = nw;
                    b4_l_[g] = nl;
                    changed = true;
                }
                ++g;
            }
        }
        if (changed) {
            NSA_HIP(hipMemcpy(const_cast<double*>(b4_.par_inst), par.data(), sizeof(double) * par.size(),
                              hipMemcpyHostToDevice));
            // the size derivatives follow the new sizes
            size_derivs_dirty_ = true;
        }
    }
}

void Bundle::get_device_values(double* value, double* w, double* l) {
    ensure_adjoint();
    NSA_HIP(hipStreamSynchronize(stream_));
    const size_t nd = (size_t)doff_[NI_];
    if (value) {
        std::fill(value, value + nd, 0.0);
        auto get = [&](const std::vector<int32_t>& owner, const double* dptr) {
            if (owner.empty()) return;
            std::vector<double> v(owner.size());
            NSA_HIP(hipMemcpy(v.data(), dptr, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < owner.size(); ++k) value[owner[k]] = v[k];
        };
        get(r_owner_, d_r_ohms_);
        get(c_owner_, cap_.farads);
        get(v_owner_, vsrc_.dc);
        get(i_owner_, isrc_.dc);
    }
    if (w) std::fill(w, w + nd, 0.0);
    if (l) std::fill(l, l + nd, 0.0);
    int g = 0;
    for (int i = 0; i < NI_; ++i)
        for (int d : ckt_[i]->family(DK_B4)) {
            const Device& dv = ckt_[i]->devices()[d];
            if (w) w[doff_[i] + d] = b4_w_.empty() ? dv.w : b4_w_[g];
            if (l) l[doff_[i] + d] = b4_l_.empty() ? dv.l : b4_l_[g];
            ++g;
        }
}

}  // namespace nsa
