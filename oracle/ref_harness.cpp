// Test-only harness over the compiled reference (neospice).  Linked into
// oracle/_ref/libneospice_ref.so together with the reference's own objects.
//
// It exposes, through a plain C ABI for the Python tests and the CPU-baseline
// leg of bench.py:
//   * circuit construction through the reference's typed Circuit API
//     (Circuit::R/C/V/I/D/M, circuit_typed.cpp) with nodes "n1".."nN"
//     created in order, so node k maps to MNA index k-1;
//   * the finalized pattern, load order and branch indices;
//   * one Newton-iteration device load exactly as newton_solve stages it
//     (src/core/newton.cpp:141-175);
//   * solve_dc / solve_transient runs, with a probe device that records the
//     IntegratorCtx of every Newton iteration and every accepted step;
//   * raw dumps of the Sparse-1.3 (NeoSolver) and AMD-LU factorizations,
//     including pivot order and fill pattern.
//
// This file is the only place that reaches into private reference members:
// the two solver headers are included with `private` mapped to `public`.
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <functional>
#include <iostream>
#include <map>
#include <optional>
#include <set>
#include <span>
#include <sstream>
#include <string_view>
#include <unordered_map>
#include <unordered_set>
#include <utility>

// Standard headers are all included above, so the access override below
// only reaches the reference's own classes.
#define private public
#define protected public
#include "solver/matrix.hpp"
#include "core/neo_solver.hpp"
#include "core/amd_lu_solver.hpp"
#include "devices/bsim4v7/bsim4v7_device.hpp"
#undef private
#undef protected

#include "core/circuit.hpp"
#include "core/dc.hpp"
#include "core/convergence.hpp"
#include "core/newton.hpp"
#include "core/transient.hpp"
#include "core/sens.hpp"
#include "core/solver_iface.hpp"
#include "core/amd.hpp"
#include "devices/device.hpp"
#include "devices/vsource.hpp"
#include "devices/isource.hpp"
#include "devices/capacitor.hpp"

using namespace neospice;

namespace {

struct CtxRecord {
    int mode, order, method, accepted_before;
    double ag0, ag1, ag2, delta, dold0, dold1, dold2, xmu_ratio, time;
};

// A stamp-free device loaded last (rank 1000, circuit.cpp:57) that observes
// the integrator context of every Newton iteration.  Its single state slot
// makes Circuit::rotate_state() (called once per accepted transient step,
// transient.cpp:1004) re-bind it, which marks step acceptance.
class ProbeDevice : public Device {
public:
    explicit ProbeDevice(std::vector<CtxRecord>* recs, int* accepts)
        : Device("zprobe"), recs_(recs), accepts_(accepts) {}
    void stamp_pattern(SparsityBuilder&) const override {}
    void assign_offsets(const SparsityPattern&) override {}
    void evaluate(const std::vector<double>&, NumericMatrix&,
                  std::vector<double>&) override {
        if (!recording) return;
        const IntegratorCtx* c = tls_integrator_ctx;
        if (!c) return;
        CtxRecord r{};
        r.mode = c->mode; r.order = c->order; r.method = c->integrate_method;
        r.accepted_before = *accepts_;
        r.ag0 = c->ag[0]; r.ag1 = c->ag[1]; r.ag2 = c->ag[2];
        r.delta = c->delta; r.dold0 = c->delta_old[0]; r.dold1 = c->delta_old[1];
        r.dold2 = c->delta_old[2]; r.xmu_ratio = c->xmu_ratio; r.time = c->current_time;
        recs_->push_back(r);
    }
    int32_t state_vars() const override { return 1; }
    void set_state_ptrs(double*, double*, double*, double*, int32_t) override {
        if (recording) ++*accepts_;
    }
    bool recording = false;
private:
    std::vector<CtxRecord>* recs_;
    int* accepts_;
};

struct RefCkt {
    Circuit ckt;
    std::vector<NodeId> nodes;   // nodes[k] for k = 1..N (index 0 = GND)
    std::vector<CtxRecord> recs;
    int accepts = 0;
    ProbeDevice* probe = nullptr;
    std::string err;
};

NodeId nid(RefCkt* h, int k) {
    if (k <= 0) return GND;
    if (k >= static_cast<int>(h->nodes.size()))
        throw std::out_of_range("node id out of range");
    return h->nodes[k];
}

thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_ckt_new(int n_nodes) {
    auto* h = new RefCkt();
    h->nodes.push_back(GND);
    for (int k = 1; k <= n_nodes; ++k)
        h->nodes.push_back(h->ckt.node("n" + std::to_string(k)));
    return h;
}

void ref_ckt_free(void* p) { delete static_cast<RefCkt*>(p); }

int ref_ckt_model(void* p, const char* text) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] { h->ckt.model(text); return 0; });
}

int ref_ckt_r(void* p, const char* name, int a, int b, double v) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] { return (int)h->ckt.R(name, nid(h, a), nid(h, b), v); });
}

int ref_ckt_c(void* p, const char* name, int a, int b, double v) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] { return (int)h->ckt.C(name, nid(h, a), nid(h, b), v); });
}

int ref_ckt_v(void* p, const char* name, int a, int b, double dc) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] { return (int)h->ckt.V(name, nid(h, a), nid(h, b), dc); });
}

int ref_ckt_i(void* p, const char* name, int a, int b, double dc) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] { return (int)h->ckt.I(name, nid(h, a), nid(h, b), dc); });
}

// PULSE source: the DC value is left unset, as for "V1 a b PULSE(...)".
int ref_ckt_v_pulse(void* p, const char* name, int a, int b, double v1,
                    double v2, double td, double tr, double tf, double pw,
                    double per) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        int id = (int)h->ckt.V(name, nid(h, a), nid(h, b), 0.0);
        auto* vs = dynamic_cast<VSource*>(h->ckt.devices()[id].get());
        PulseParams pp;
        pp.v1 = v1; pp.v2 = v2; pp.td = td; pp.tr = tr; pp.tf = tf;
        pp.pw = pw; pp.per = per;
        vs->set_pulse(pp);
        return id;
    });
}

int ref_ckt_v_pwl(void* p, const char* name, int a, int b, int npts,
                  const double* t, const double* v) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        int id = (int)h->ckt.V(name, nid(h, a), nid(h, b), 0.0);
        auto* vs = dynamic_cast<VSource*>(h->ckt.devices()[id].get());
        PwlParams pw;
        for (int i = 0; i < npts; ++i) pw.points.emplace_back(t[i], v[i]);
        vs->set_pwl(pw);
        return id;
    });
}

int ref_ckt_d(void* p, const char* name, int a, int k, const char* model) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] { return (int)h->ckt.D(name, nid(h, a), nid(h, k), model); });
}

int ref_ckt_m(void* p, const char* name, int d, int g, int s, int b,
              const char* model, double w, double l) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        return (int)h->ckt.M(name, nid(h, d), nid(h, g), nid(h, s), nid(h, b),
                             model, w, l);
    });
}

int ref_ckt_ic(void* p, int node, double v) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] { h->ckt.ic[nid(h, node)] = v; return 0; });
}

// Options used by the configs: xmu (0 = backward Euler), gmin, reltol,
// abstol, vntol, trtol, itl4, method (0 trap / 1 gear).
int ref_ckt_option(void* p, const char* key, double v) {
    auto* h = static_cast<RefCkt*>(p);
    auto& o = h->ckt.options;
    std::string k(key);
    if (k == "xmu") o.xmu = v;
    else if (k == "gmin") o.gmin = v;
    else if (k == "reltol") o.reltol = v;
    else if (k == "abstol") o.abstol = v;
    else if (k == "vntol") o.vntol = v;
    else if (k == "trtol") o.trtol = v;
    else if (k == "itl4") o.itl4 = (int)v;
    else if (k == "itl1") o.itl1 = (int)v;
    else if (k == "method") o.method = (v != 0.0) ? "gear" : "trap";
    else if (k == "verbose") o.verbose = v != 0.0;
    else { g_err = "unknown option " + k; return -1; }
    return 0;
}

// Adds the context probe (must precede finalize) and finalizes.
int ref_ckt_finalize(void* p, int with_probe) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        if (with_probe) {
            auto pr = std::make_unique<ProbeDevice>(&h->recs, &h->accepts);
            h->probe = pr.get();
            h->ckt.add_dev(std::move(pr));
        }
        h->ckt.finalize();
        return 0;
    });
}

int ref_ckt_num_vars(void* p) { return static_cast<RefCkt*>(p)->ckt.num_vars(); }
int ref_ckt_num_nodes(void* p) { return static_cast<RefCkt*>(p)->ckt.num_nodes(); }
int ref_ckt_num_states(void* p) { return static_cast<RefCkt*>(p)->ckt.num_states(); }
int ref_ckt_nnz(void* p) { return static_cast<RefCkt*>(p)->ckt.pattern().nnz(); }
int ref_ckt_is_linear(void* p) { return static_cast<RefCkt*>(p)->ckt.is_linear() ? 1 : 0; }
int ref_ckt_num_devices(void* p) { return (int)static_cast<RefCkt*>(p)->ckt.devices().size(); }

void ref_ckt_pattern(void* p, int32_t* col_ptr, int32_t* row_idx) {
    auto csc = static_cast<RefCkt*>(p)->ckt.pattern().to_csc();
    std::copy(csc.col_ptr.begin(), csc.col_ptr.end(), col_ptr);
    std::copy(csc.row_idx.begin(), csc.row_idx.end(), row_idx);
}

// Device indices (insertion order) in load order.
int ref_ckt_load_order(void* p, int32_t* out) {
    auto* h = static_cast<RefCkt*>(p);
    const auto& devs = h->ckt.devices();
    int k = 0;
    for (Device* d : h->ckt.device_load_order()) {
        for (size_t i = 0; i < devs.size(); ++i)
            if (devs[i].get() == d) { out[k++] = (int)i; break; }
    }
    return k;
}

int ref_ckt_branch_index(void* p, int dev) {
    auto* h = static_cast<RefCkt*>(p);
    return h->ckt.devices()[dev]->branch_index();
}

int ref_ckt_dead_nodes(void* p, int32_t* out) {
    auto* h = static_cast<RefCkt*>(p);
    const auto& d = h->ckt.dead_node_indices();
    for (size_t i = 0; i < d.size(); ++i) out[i] = d[i];
    return (int)d.size();
}

double* ref_ckt_state(void* p, int which) {
    auto* h = static_cast<RefCkt*>(p);
    switch (which) {
        case 0: return h->ckt.state0();
        case 1: return h->ckt.state1();
        case 2: return h->ckt.state2();
        default: return h->ckt.state3();
    }
}

// Sets the integrator context fields a device load reads.
void ref_ckt_set_ctx(void* p, int mode, int order, int method, double ag0,
                     double ag1, double ag2, double delta, double dold1,
                     double xmu_ratio, double time) {
    auto* h = static_cast<RefCkt*>(p);
    auto& c = h->ckt.integrator_ctx;
    c.mode = mode; c.order = order; c.integrate_method = method;
    c.ag[0] = ag0; c.ag[1] = ag1; c.ag[2] = ag2;
    c.delta = delta; c.delta_old[0] = delta; c.delta_old[1] = dold1;
    c.xmu_ratio = xmu_ratio; c.current_time = time;
    c.options = &h->ckt.options;
    for (auto& dev : h->ckt.devices()) {
        if (auto* vs = dynamic_cast<VSource*>(dev.get())) vs->set_time(time);
        else if (auto* is = dynamic_cast<ISource*>(dev.get())) is->set_time(time);
        else if (auto* cp = dynamic_cast<Capacitor*>(dev.get())) {
            if (mode & 0x1) { cp->set_transient(delta); cp->set_integrator_order(order); }
        }
    }
}

// One Newton-iteration load at `solution`, staged as newton.cpp:141-175:
// values (nnz) and rhs (n) out.  Devices read/write the circuit's states.
int ref_ckt_load(void* p, const double* solution, double* values, double* rhs) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        const int n = ckt.num_vars();
        std::vector<double> sol(solution, solution + n);
        NumericMatrix mat(ckt.pattern());
        std::vector<double> r(n, 0.0);
        std::vector<double> one_sol(n + 1, 0.0), one_rhs(n + 1, 0.0);
        std::copy(sol.begin(), sol.end(), one_sol.begin() + 1);
        OneBasedEvalArrays arrays{one_sol.data(), one_rhs.data(), n + 1};
        tls_integrator_ctx = &ckt.integrator_ctx;
        tls_one_based_eval_arrays = &arrays;
        for (Device* dev : ckt.device_load_order()) dev->evaluate(sol, mat, r);
        tls_one_based_eval_arrays = nullptr;
        tls_integrator_ctx = nullptr;
        for (int i = 0; i < n; ++i) r[i] += one_rhs[i + 1];
        const double pin = (ckt.options.gmin > 0.0) ? ckt.options.gmin : 1e-12;
        for (int32_t idx : ckt.dead_node_indices())
            mat.add(ckt.pattern().offset(idx, idx), pin);
        std::copy(mat.data(), mat.data() + mat.nnz(), values);
        std::copy(r.begin(), r.end(), rhs);
        return 0;
    });
}

// Simulator::run_dc path (api/neospice.cpp:72-77 -> solve_dc).
// Writes the full MNA solution (n) and returns the Newton iteration count.
int ref_dc(void* p, double* solution_out, double* seconds) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.finalize_if_needed();
        auto t0 = std::chrono::steady_clock::now();
        DCResult r = solve_dc(ckt);
        auto t1 = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
        const auto* op = ckt.operating_point();
        if (!op) throw std::runtime_error("no operating point");
        std::copy(op->begin(), op->end(), solution_out);
        return r.status.iterations;
    });
}

// Simulator::run_transient path.  Output points (time + full solution of
// every node and branch) are copied out up to max_points; the probe records
// are kept in the handle (ref_tran_records).  Returns the point count.
int ref_tran(void* p, double tstep, double tstop, int max_points,
             double* times, double* sols, double* seconds, int* newton_iters) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.finalize_if_needed();
        h->recs.clear();
        h->accepts = 0;
        if (h->probe) h->probe->recording = true;
        auto t0 = std::chrono::steady_clock::now();
        TransientResult r = solve_transient(ckt, tstep, tstop);
        auto t1 = std::chrono::steady_clock::now();
        if (h->probe) h->probe->recording = false;
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
        if (newton_iters) *newton_iters = r.status.iterations;
        const int n = ckt.num_vars();
        const int nn = ckt.num_nodes();
        const int npts = (int)r.time.size();
        for (int k = 0; k < npts && k < max_points; ++k) {
            times[k] = r.time[k];
            double* row = sols + (size_t)k * n;
            for (int i = 0; i < n; ++i) row[i] = 0.0;
            for (int i = 0; i < nn; ++i) row[i] = r.voltages_dense[i][k];
            const auto& devs = ckt.devices();
            for (size_t d = 0; d < devs.size(); ++d) {
                int br = devs[d]->branch_index();
                if (br >= 0 && br < n) row[br] = r.currents_dense[d][k];
            }
        }
        return npts;
    });
}

// The accepted-step path of solve_transient (transient.cpp:710-802 preamble,
// then per step :907-915 and :1004) on a fixed backward-Euler step sequence:
// DC operating point in MODETRANOP on its own solver, a fresh transient
// solver and NewtonWorkspace, state1..3 = state0, then for every step the
// integrator context of an order-1 step, newton_solve with itl4, and
// rotate_state.  Times the stepping loop; returns the Newton iterations of
// the loop (negative on a non-converged step).  solution_out (n) receives
// the final solution.
int ref_tran_fixed(void* p, double dt, int nsteps, double* solution_out, double* seconds) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.finalize_if_needed();
        const int n = ckt.num_vars();
        ckt.integrator_ctx.options = &ckt.options;
        std::vector<double> sol(n, 0.0);
        {
            auto dc_solver = make_solver(n, ckt.is_linear());
            dc_solver->symbolic(ckt.pattern());
            ckt.integrator_ctx.mode = 0x20 | 0x200;  // MODETRANOP | MODEINITJCT
            NewtonResult r = newton_solve(ckt, *dc_solver, sol, ckt.options);
            if (!r.converged) throw std::runtime_error("DC preamble did not converge");
        }
        auto solver = make_solver(n, ckt.is_linear());
        solver->symbolic(ckt.pattern());
        NewtonWorkspace ws(ckt.pattern());
        const int ns = ckt.num_states();
        std::copy_n(ckt.state0(), ns, ckt.state1());
        std::copy_n(ckt.state0(), ns, ckt.state2());
        std::copy_n(ckt.state0(), ns, ckt.state3());
        SimOptions topts = ckt.options;
        topts.max_iter = ckt.options.itl4;
        int iters = 0;
        auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < nsteps; ++k) {
            auto& c = ckt.integrator_ctx;
            const double t = (k + 1) * dt;
            c.order = 1;
            c.delta = dt;
            c.current_time = t;
            c.delta_old[0] = dt;
            c.delta_old[1] = dt;
            c.mode = 0x1 | (k == 0 ? 0x1000 : 0x2000);
            c.ag[0] = 1.0 / dt;
            c.ag[1] = -1.0 / dt;
            c.ag[2] = 0.0;
            c.xmu_ratio = 0.0;
            for (auto& dev : ckt.devices()) {
                if (auto* cp = dynamic_cast<Capacitor*>(dev.get())) {
                    cp->set_transient(dt);
                    cp->set_integrator_order(1);
                } else if (auto* vs = dynamic_cast<VSource*>(dev.get())) {
                    vs->set_time(t);
                } else if (auto* is = dynamic_cast<ISource*>(dev.get())) {
                    is->set_time(t);
                }
            }
            NewtonResult r = newton_solve(ckt, *solver, sol, topts, ws);
            if (!r.converged) return -1;
            iters += r.iterations;
            ckt.rotate_state();
        }
        auto t1 = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
        if (solution_out) std::copy(sol.begin(), sol.end(), solution_out);
        return iters;
    });
}

// The same fixed-step sequence with the step order of solve_transient's
// controller at a constant step: backward Euler for the first step, then
// order 2 of the circuit's integration method (trapezoidal with the
// options' xmu, fill_integrator_context, transient.cpp:398-447).  Writes
// every accepted solution ((nsteps+1) x n, row 0 = operating point).
// Fixed steps of dt (backward Euler, then order 2 of the circuit's method).
// With trunc_out (nsteps x num devices), each step also records every
// device's compute_trunc after its Newton solve and before the states
// rotate (where transient.cpp:478-485 evaluates it), with the constant-step
// history CKTdeltaOld[0..2] = dt.
static int fixed_steps(void* p, double dt, int nsteps, double* solutions_out, double* seconds,
                       double* trunc_out) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.finalize_if_needed();
        const int n = ckt.num_vars();
        ckt.integrator_ctx.options = &ckt.options;
        std::vector<double> sol(n, 0.0);
        {
            auto dc_solver = make_solver(n, ckt.is_linear());
            dc_solver->symbolic(ckt.pattern());
            ckt.integrator_ctx.mode = 0x20 | 0x200;  // MODETRANOP | MODEINITJCT
            NewtonResult r = newton_solve(ckt, *dc_solver, sol, ckt.options);
            if (!r.converged) throw std::runtime_error("DC preamble did not converge");
        }
        if (solutions_out) std::copy(sol.begin(), sol.end(), solutions_out);
        auto solver = make_solver(n, ckt.is_linear());
        solver->symbolic(ckt.pattern());
        NewtonWorkspace ws(ckt.pattern());
        const int ns = ckt.num_states();
        std::copy_n(ckt.state0(), ns, ckt.state1());
        std::copy_n(ckt.state0(), ns, ckt.state2());
        std::copy_n(ckt.state0(), ns, ckt.state3());
        SimOptions topts = ckt.options;
        topts.max_iter = ckt.options.itl4;
        const bool gear = ckt.integrator_ctx.integrate_method == 1;
        int iters = 0;
        auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < nsteps; ++k) {
            auto& c = ckt.integrator_ctx;
            const double t = (k + 1) * dt;
            const int order = (k == 0) ? 1 : 2;
            c.order = order;
            c.delta = dt;
            c.current_time = t;
            c.delta_old[0] = dt;
            c.delta_old[1] = dt;
            c.mode = 0x1 | (k == 0 ? 0x1000 : 0x2000);
            if (order == 1) {
                c.ag[0] = 1.0 / dt;
                c.ag[1] = -1.0 / dt;
                c.ag[2] = 0.0;
                c.xmu_ratio = 0.0;
            } else if (!gear) {
                const double om = 1.0 - ckt.options.xmu;
                c.ag[0] = 1.0 / (dt * om);
                c.ag[1] = -1.0 / (dt * om);
                c.ag[2] = 0.0;
                c.xmu_ratio = ckt.options.xmu / om;
            } else {
                const double h_old = dt, sum = dt + h_old;
                c.ag[0] = (2.0 * dt + h_old) / (dt * sum);
                c.ag[1] = -(dt + h_old) / (dt * h_old);
                c.ag[2] = dt / (h_old * sum);
                c.xmu_ratio = 0.0;
            }
            for (auto& dev : ckt.devices()) {
                if (auto* cp = dynamic_cast<Capacitor*>(dev.get())) {
                    cp->set_transient(dt);
                    cp->set_integrator_order(order);
                } else if (auto* vs = dynamic_cast<VSource*>(dev.get())) {
                    vs->set_time(t);
                } else if (auto* is = dynamic_cast<ISource*>(dev.get())) {
                    is->set_time(t);
                }
            }
            NewtonResult r = newton_solve(ckt, *solver, sol, topts, ws);
            if (!r.converged) return -1;
            iters += r.iterations;
            if (trunc_out) {
                c.delta_old[2] = dt;
                const size_t nd = ckt.devices().size();
                for (size_t i = 0; i < nd; ++i)
                    trunc_out[(size_t)k * nd + i] = ckt.devices()[i]->compute_trunc(c, ckt.options);
            }
            ckt.rotate_state();
            if (solutions_out) std::copy(sol.begin(), sol.end(), solutions_out + (size_t)(k + 1) * n);
        }
        auto t1 = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
        return iters;
    });
}

int ref_tran_steps(void* p, double dt, int nsteps, double* solutions_out, double* seconds) {
    return fixed_steps(p, dt, nsteps, solutions_out, seconds, nullptr);
}

int ref_tran_trunc(void* p, double dt, int nsteps, double* solutions_out, double* trunc_out) {
    return fixed_steps(p, dt, nsteps, solutions_out, nullptr, trunc_out);
}

// The BSIM4v7 constants of device `dev` after BSIM4v7temp (the device must
// have been evaluated once), in the order of the product's B4Param record:
// model, junction capacitances, size-dependent (pParam), instance fields.
int ref_bsim4_fields(void* p, int dev, double* out) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        auto* b = dynamic_cast<BSIM4v7Device*>(h->ckt.devices().at(dev).get());
        if (!b || !b->temp_done_) throw std::runtime_error("not an evaluated BSIM4v7 device");
        const auto* m = b->model_;
        const auto& i = b->inst_;
        const auto* q = i.pParam;
        const double nf = i.BSIM4v7nf;
        const double v[] = {
            (double)m->BSIM4v7type, m->BSIM4v7vtm, m->BSIM4v7coxe, m->BSIM4v7coxp, m->BSIM4v7toxe,
            m->BSIM4v7toxp, m->BSIM4v7factor1, m->BSIM4v7vcrit, m->BSIM4v7bdos, m->BSIM4v7ados,
            h->ckt.options.temp / m->BSIM4v7tnom - 1.0,
            m->BSIM4v7PhiBS, m->BSIM4v7PhiBSWS, m->BSIM4v7PhiBSWGS, m->BSIM4v7PhiBD,
            m->BSIM4v7PhiBSWD, m->BSIM4v7PhiBSWGD,
            m->BSIM4v7SbulkJctBotGradingCoeff, m->BSIM4v7SbulkJctSideGradingCoeff,
            m->BSIM4v7SbulkJctGateSideGradingCoeff, m->BSIM4v7DbulkJctBotGradingCoeff,
            m->BSIM4v7DbulkJctSideGradingCoeff, m->BSIM4v7DbulkJctGateSideGradingCoeff,
            m->BSIM4v7SunitAreaTempJctCap * i.BSIM4v7Aseff,
            m->BSIM4v7SunitLengthSidewallTempJctCap * i.BSIM4v7Pseff,
            m->BSIM4v7SunitLengthGateSidewallTempJctCap * q->BSIM4v7weffCJ * nf,
            m->BSIM4v7DunitAreaTempJctCap * i.BSIM4v7Adeff,
            m->BSIM4v7DunitLengthSidewallTempJctCap * i.BSIM4v7Pdeff,
            m->BSIM4v7DunitLengthGateSidewallTempJctCap * q->BSIM4v7weffCJ * nf,
            q->BSIM4v7leff, q->BSIM4v7weff, q->BSIM4v7leffCV, q->BSIM4v7weffCV, q->BSIM4v7phi,
            q->BSIM4v7sqrtPhi, q->BSIM4v7Xdep0, q->BSIM4v7vbi, q->BSIM4v7k1, q->BSIM4v7k1ox,
            q->BSIM4v7k3, q->BSIM4v7k3b,
            q->BSIM4v7dvt0, q->BSIM4v7dvt1, q->BSIM4v7dvt2, q->BSIM4v7dvt0w, q->BSIM4v7dvt1w,
            q->BSIM4v7dvt2w, q->BSIM4v7lpe0, q->BSIM4v7lpeb, q->BSIM4v7kt1, q->BSIM4v7kt1l,
            q->BSIM4v7kt2, q->BSIM4v7w0,
            q->BSIM4v7theta0vb0, q->BSIM4v7etab, q->BSIM4v7nfactor, q->BSIM4v7cdsc,
            q->BSIM4v7cdscb, q->BSIM4v7cdscd, q->BSIM4v7cit, q->BSIM4v7mstar, q->BSIM4v7voffcbn,
            q->BSIM4v7cdep0,
            q->BSIM4v7dwg, q->BSIM4v7dwb, q->BSIM4v7prwg, q->BSIM4v7prwb, q->BSIM4v7rds0,
            q->BSIM4v7rdswmin, q->BSIM4v7xj, q->BSIM4v7a0, q->BSIM4v7ags, q->BSIM4v7b0,
            q->BSIM4v7b1, q->BSIM4v7keta,
            q->BSIM4v7ua, q->BSIM4v7ub, q->BSIM4v7uc, q->BSIM4v7ud, q->BSIM4v7a2, q->BSIM4v7delta,
            q->BSIM4v7pvag, q->BSIM4v7pclm, q->BSIM4v7litl, q->BSIM4v7thetaRout,
            q->BSIM4v7pdiblb, q->BSIM4v7pditsd,
            q->BSIM4v7pscbe1, q->BSIM4v7pscbe2, q->BSIM4v7noff, q->BSIM4v7voffcv,
            q->BSIM4v7abulkCVfactor, q->BSIM4v7acde, q->BSIM4v7ldeb, q->BSIM4v7moin,
            q->BSIM4v7cgdo, q->BSIM4v7cgso, q->BSIM4v7cgbo, q->BSIM4v7cgdl, q->BSIM4v7cgsl,
            q->BSIM4v7ckappad, q->BSIM4v7ckappas,
            i.BSIM4v7vth0, i.BSIM4v7k2ox, i.BSIM4v7eta0, i.BSIM4v7vbsc, i.BSIM4v7u0temp,
            i.BSIM4v7vsattemp, i.BSIM4v7vtfbphi2, i.BSIM4v7vfbzb};
        static_assert(sizeof(v) / sizeof(v[0]) == 110, "field count");
        std::memcpy(out, v, sizeof(v));
        return 0;
    });
}

// Devices whose device_converged(solution) is false after the last load.
int ref_ckt_devnoncon(void* p, const double* solution) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        std::vector<double> sol(solution, solution + h->ckt.num_vars());
        int bad = 0;
        for (Device* dev : h->ckt.device_load_order()) bad += dev->device_converged(sol) ? 0 : 1;
        return bad;
    });
}

int ref_num_devices(void* p) { return (int)static_cast<RefCkt*>(p)->ckt.devices().size(); }

int ref_tran_num_records(void* p) { return (int)static_cast<RefCkt*>(p)->recs.size(); }

// Record layout: 4 ints (mode, order, method, accepted_before) and
// 10 doubles (ag0, ag1, ag2, delta, dold0, dold1, dold2, xmu_ratio, time, 0).
void ref_tran_records(void* p, int32_t* ints, double* dbls) {
    auto* h = static_cast<RefCkt*>(p);
    for (size_t i = 0; i < h->recs.size(); ++i) {
        const auto& r = h->recs[i];
        ints[4 * i + 0] = r.mode; ints[4 * i + 1] = r.order;
        ints[4 * i + 2] = r.method; ints[4 * i + 3] = r.accepted_before;
        double* d = dbls + 10 * i;
        d[0] = r.ag0; d[1] = r.ag1; d[2] = r.ag2; d[3] = r.delta; d[4] = r.dold0;
        d[5] = r.dold1; d[6] = r.dold2; d[7] = r.xmu_ratio; d[8] = r.time; d[9] = 0;
    }
}

// newton_solve on the circuit as-is (mode taken from integrator_ctx), with
// a fresh solver from make_solver.  Returns iterations (negative: failed).
int ref_newton(void* p, double* solution, int mode) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.integrator_ctx.mode = mode;
        ckt.integrator_ctx.options = &ckt.options;
        auto solver = make_solver(ckt.num_vars(), ckt.is_linear());
        solver->symbolic(ckt.pattern());
        std::vector<double> sol(solution, solution + ckt.num_vars());
        NewtonResult r = newton_solve(ckt, *solver, sol, ckt.options);
        std::copy(sol.begin(), sol.end(), solution);
        return r.converged ? r.iterations : -r.iterations;
    });
}

// .sens (sens.cpp:105-186) for output "v(n<k>)".  Writes sensitivities in the
// order the reference reports them; returns their count.
int ref_sens(void* p, int out_node, int max_out, double* vals, char* names,
             int name_stride) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.finalize_if_needed();
        SensResult r = solve_sens(ckt, "v(n" + std::to_string(out_node) + ")");
        int k = 0;
        for (const auto& e : r.entries) {
            if (k >= max_out) break;
            vals[k] = e.sensitivity;
            std::string name = e.element + ":" + e.parameter;
            std::strncpy(names + (size_t)k * name_stride, name.c_str(), name_stride - 1);
            names[(size_t)k * name_stride + name_stride - 1] = 0;
            ++k;
        }
        return k;
    });
}

// ---------------------------------------------------------------------------
// Solver-level dumps on a raw CSC pattern (col_ptr/row_idx, values in CSC
// order, exactly the NumericMatrix slot order).
// ---------------------------------------------------------------------------

struct RefSolver {
    std::unique_ptr<SparsityPattern> pat;
    std::unique_ptr<NumericMatrix> mat;
    NeoSolver neo;
    AmdLuSolver amd;
    int kind;  // 0 neo, 1 amd
};

void* ref_solver_new(int kind, int n, const int32_t* col_ptr, const int32_t* row_idx) {
    auto* s = new RefSolver();
    std::vector<std::pair<int32_t, int32_t>> e;
    for (int j = 0; j < n; ++j)
        for (int p = col_ptr[j]; p < col_ptr[j + 1]; ++p) e.emplace_back(row_idx[p], j);
    s->pat = std::make_unique<SparsityPattern>(n, std::move(e));
    s->mat = std::make_unique<NumericMatrix>(*s->pat);
    s->kind = kind;
    if (kind == 0) s->neo.symbolic(*s->pat); else s->amd.symbolic(*s->pat);
    return s;
}

void ref_solver_free(void* p) { delete static_cast<RefSolver*>(p); }

static void load_vals(RefSolver* s, const double* v) {
    std::copy(v, v + s->mat->nnz(), s->mat->data());
}

// ISolver::set_equilibrate on the NeoSolver (a no-op on AmdLuSolver).
void ref_solver_set_equilibrate(void* p, int on) {
    auto* s = static_cast<RefSolver*>(p);
    if (s->kind == 0) s->neo.set_equilibrate(on != 0); else s->amd.set_equilibrate(on != 0);
}

int ref_solver_numeric(void* p, const double* values, double gmin) {
    auto* s = static_cast<RefSolver*>(p);
    return guarded([&] {
        load_vals(s, values);
        return s->kind == 0 ? (int)s->neo.numeric(*s->pat, *s->mat, gmin)
                            : (int)s->amd.numeric(*s->pat, *s->mat, gmin);
    });
}

int ref_solver_refactorize(void* p, const double* values, double gmin) {
    auto* s = static_cast<RefSolver*>(p);
    return guarded([&] {
        load_vals(s, values);
        return s->kind == 0 ? (int)s->neo.refactorize(*s->mat, gmin)
                            : (int)s->amd.refactorize(*s->mat, gmin);
    });
}

int ref_solver_solve(void* p, double* rhs) {
    auto* s = static_cast<RefSolver*>(p);
    return guarded([&] {
        std::vector<double> r(rhs, rhs + s->pat->size());
        if (s->kind == 0) s->neo.solve(r); else s->amd.solve(r);
        std::copy(r.begin(), r.end(), rhs);
        return 0;
    });
}

// Transposed solve on the Sparse-1.3 factors (solve.cpp:102-141).
int ref_solver_solve_transposed(void* p, double* rhs) {
    auto* s = static_cast<RefSolver*>(p);
    return guarded([&] {
        if (s->kind != 0) throw std::logic_error("transposed solve: markowitz only");
        const int n = s->pat->size();
        std::vector<double> r1(n + 1, 0.0), x1(n + 1, 0.0);
        for (int i = 0; i < n; ++i) r1[i + 1] = rhs[i];
        s->neo.matrix_->solve_transposed(r1.data(), x1.data());
        for (int i = 0; i < n; ++i) rhs[i] = x1[i + 1];
        return 0;
    });
}

// ISolver's complex paths (neo_solver.cpp:228-279; AmdLuSolver delegates
// them to a NeoSolver): ax is 2 nnz values, (re, im) per CSC slot; rhs 2 n.
int ref_solver_numeric_complex(void* p, const double* ax) {
    auto* s = static_cast<RefSolver*>(p);
    return guarded([&] {
        std::vector<double> a(ax, ax + 2 * s->mat->nnz());
        if (s->kind == 0) s->neo.numeric_complex(*s->pat, a); else s->amd.numeric_complex(*s->pat, a);
        // numeric_complex reports nothing; its matrix's error state says
        // whether the ordering met a singular matrix
        const auto& m = s->kind == 0 ? *s->neo.matrix_ : *s->amd.complex_solver_->matrix_;
        return m.error_ == neospice::solver::SparseError::Singular ? 1 : 0;
    });
}

int ref_solver_refactorize_complex(void* p, const double* ax) {
    auto* s = static_cast<RefSolver*>(p);
    return guarded([&] {
        std::vector<double> a(ax, ax + 2 * s->mat->nnz());
        return s->kind == 0 ? (int)s->neo.refactorize_complex(a) : (int)s->amd.refactorize_complex(a);
    });
}

int ref_solver_solve_complex(void* p, double* rhs) {
    auto* s = static_cast<RefSolver*>(p);
    return guarded([&] {
        std::vector<double> r(rhs, rhs + 2 * s->pat->size());
        if (s->kind == 0) s->neo.solve_complex(r); else s->amd.solve_complex(r);
        std::copy(r.begin(), r.end(), rhs);
        return 0;
    });
}

int ref_solver_fast_count(void* p) { return static_cast<RefSolver*>(p)->amd.refactor_fast_count(); }
int ref_solver_fallback_count(void* p) { return static_cast<RefSolver*>(p)->amd.refactor_fallback_count(); }

// Sparse-1.3 factor dump (internal 1-based indices).  int_to_ext_row/col
// (n entries each, external 1-based), and for every internal column the
// linked elements (row, value) in list order: returns the element count and
// fills col_ptr (n+1), rows, vals when non-null.  Diagonal values are the
// stored reciprocal pivots; L is below, U (unit) above.
int ref_neo_dump(void* p, int32_t* i2e_row, int32_t* i2e_col, int32_t* col_ptr,
                 int32_t* rows, double* vals) {
    auto* s = static_cast<RefSolver*>(p);
    auto& m = *s->neo.matrix_;
    const int n = m.size_;
    int k = 0;
    if (col_ptr) col_ptr[0] = 0;
    for (int j = 1; j <= n; ++j) {
        if (i2e_row) i2e_row[j - 1] = m.int_to_ext_row_[j];
        if (i2e_col) i2e_col[j - 1] = m.int_to_ext_col_[j];
        for (auto* e = m.first_in_col_[j]; e; e = e->NextInCol) {
            if (rows) rows[k] = e->Row;
            if (vals) vals[k] = e->Real;
            ++k;
        }
        if (col_ptr) col_ptr[j] = k;
    }
    return k;
}

int ref_neo_fillins(void* p) { return static_cast<RefSolver*>(p)->neo.matrix_->fillins_; }

// AMD-LU dump: q (n), pinv (n), Lp/Up (n+1), Li/Ui, Lx/Ux.  Returns
// nnz(L) * 2^20 + nnz(U) when arrays are null (for sizing).
int64_t ref_amd_dump(void* p, int32_t* q, int32_t* pinv, int32_t* Lp, int32_t* Li,
                     double* Lx, int32_t* Up, int32_t* Ui, double* Ux) {
    auto* s = static_cast<RefSolver*>(p);
    auto& a = s->amd;
    if (!q) return ((int64_t)a.Li_.size() << 32) | (int64_t)a.Ui_.size();
    std::copy(a.q_.begin(), a.q_.end(), q);
    std::copy(a.pinv_.begin(), a.pinv_.end(), pinv);
    std::copy(a.Lp_.begin(), a.Lp_.end(), Lp);
    std::copy(a.Li_.begin(), a.Li_.end(), Li);
    std::copy(a.Lx_.begin(), a.Lx_.end(), Lx);
    std::copy(a.Up_.begin(), a.Up_.end(), Up);
    std::copy(a.Ui_.begin(), a.Ui_.end(), Ui);
    std::copy(a.Ux_.begin(), a.Ux_.end(), Ux);
    return 0;
}

// amd_ordering (amd.cpp:14-143) on a CSC pattern.
int ref_amd_ordering(int n, const int32_t* col_ptr, const int32_t* row_idx, int32_t* perm) {
    auto q = amd_ordering(n, col_ptr, row_idx);
    std::copy(q.begin(), q.end(), perm);
    return (int)q.size();
}


// The transient preamble's operating-point ladder (compute_dc_operating_point,
// transient.cpp:172-238): direct Newton in MODETRANOP from the current
// solution, then gmin stepping, source stepping and pseudo-transient on one
// solver.  stage_out: 0..3 for the method that converged, -1 if none did.
// Returns the iterations of the converged method (total over its solves).
// op: 0x20 (MODETRANOP, compute_dc_operating_point) or 0x10 (MODEDCOP).
int ref_dc_ladder_op(void* p, double* solution, int* stage_out, int op) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.finalize_if_needed();
        ckt.integrator_ctx.options = &ckt.options;
        auto solver = make_solver(ckt.num_vars(), ckt.is_linear());
        solver->symbolic(ckt.pattern());
        std::vector<double> sol(solution, solution + ckt.num_vars());
        const int jct = op | 0x200, flt = op | 0x100;
        *stage_out = -1;
        ckt.integrator_ctx.mode = jct;
        NewtonResult r = newton_solve(ckt, *solver, sol, ckt.options);
        // solve_dc tries true gmin stepping before gmin stepping (dc.cpp:120-145)
        if (!r.converged && op == 0x10) {
            r = true_gmin_stepping(ckt, *solver, sol, ckt.options, jct, flt);
            if (r.converged) *stage_out = 4;
        }
        if (r.converged) {
            if (*stage_out < 0) *stage_out = 0;
        } else {
            r = gmin_stepping(ckt, *solver, sol, ckt.options, jct, flt);
            if (r.converged) {
                *stage_out = 1;
            } else {
                ckt.integrator_ctx.mode = jct;
                r = source_stepping(ckt, *solver, sol, ckt.options);
                if (r.converged) {
                    *stage_out = 2;
                } else {
                    ckt.integrator_ctx.mode = jct;
                    r = pseudo_transient(ckt, *solver, sol, ckt.options);
                    if (r.converged) *stage_out = 3;
                }
            }
        }
        std::copy(sol.begin(), sol.end(), solution);
        return r.iterations;
    });
}


// gmin_stepping (convergence.cpp:216-346) alone, from the current solution,
// in MODETRANOP (first MODEINITJCT, then MODEINITFLOAT), on a fresh solver.
// Returns its iterations, negative when it did not converge.
int ref_dc_ladder(void* p, double* solution, int* stage_out) {
    return ref_dc_ladder_op(p, solution, stage_out, 0x20);
}

// One continuation method alone from the given solution on a fresh
// solver, in MODETRANOP (op 0x20) or MODEDCOP (0x10): 1 source stepping
// (convergence.cpp:472-582), 2 true gmin stepping (:348-470).  Returns its
// iterations, negative when it did not converge.
int ref_ladder_method(void* p, double* solution, int method, int op) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.finalize_if_needed();
        ckt.integrator_ctx.options = &ckt.options;
        auto solver = make_solver(ckt.num_vars(), ckt.is_linear());
        solver->symbolic(ckt.pattern());
        std::vector<double> sol(solution, solution + ckt.num_vars());
        NewtonResult r;
        ckt.integrator_ctx.mode = op | 0x200;
        if (method == 1) r = source_stepping(ckt, *solver, sol, ckt.options);
        else r = true_gmin_stepping(ckt, *solver, sol, ckt.options, op | 0x200, op | 0x100);
        std::copy(sol.begin(), sol.end(), solution);
        return r.converged ? r.iterations : -r.iterations;
    });
}

int ref_gmin_stepping(void* p, double* solution) {
    auto* h = static_cast<RefCkt*>(p);
    return guarded([&] {
        Circuit& ckt = h->ckt;
        ckt.finalize_if_needed();
        ckt.integrator_ctx.options = &ckt.options;
        auto solver = make_solver(ckt.num_vars(), ckt.is_linear());
        solver->symbolic(ckt.pattern());
        std::vector<double> sol(solution, solution + ckt.num_vars());
        NewtonResult r = gmin_stepping(ckt, *solver, sol, ckt.options, 0x20 | 0x200, 0x20 | 0x100);
        std::copy(sol.begin(), sol.end(), solution);
        return r.converged ? r.iterations : -r.iterations;
    });
}

}  // extern "C"
