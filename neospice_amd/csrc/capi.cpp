// extern "C" boundary (include/neospice_amd.h).  Every entry point catches
// C++ exceptions and turns them into status codes.
#include <cstring>
#include <memory>
#include <string>

#include "../../include/neospice_amd.h"
#include "ac_lu.hpp"
#include "batch_lu.hpp"
#include "circuit.hpp"
#include "common.hpp"
#include "bsim4.hpp"
#include "engine.hpp"
#include "lu_cpu.hpp"
#include "bundle.hpp"
#include "noalloc.hpp"

struct nsa_circuit {
    nsa::Circuit c;
};

struct nsa_engine {
    nsa_circuit* ckt;
    std::unique_ptr<nsa::BatchEngine> e;
};

struct nsa_bundle {
    std::unique_ptr<nsa::Bundle> b;
};

struct nsa_lu {
    hipStream_t stream = nullptr;
    std::unique_ptr<nsa::BatchLu> lu;
    std::unique_ptr<nsa::AcLu> ac;  // the complex paths, made on first use
    std::vector<int32_t> cp, ri;
    double* dA = nullptr;
    double* db = nullptr;
    int n = 0, nnz = 0, B = 0;
    nsa::AcLu& complex_lu() {
        if (!ac) ac = std::make_unique<nsa::AcLu>(n, cp.data(), ri.data(), B, stream);
        return *ac;
    }
    ~nsa_lu() {
        ac.reset();
        lu.reset();
        if (dA) (void)hipFree(dA);
        if (db) (void)hipFree(db);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        return f();
    } catch (const nsa::Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return nsa::NSA_ERR_ARGUMENT;
    } catch (const std::logic_error& e) {
        g_err = e.what();
        return nsa::NSA_ERR_LOGIC;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nsa::NSA_ERR_RUNTIME;
    }
}

nsa::Options to_opts(const nsa_options* o) {
    nsa::Options r;
    if (!o) return r;
    r.abstol = o->abstol; r.reltol = o->reltol; r.vntol = o->vntol; r.trtol = o->trtol;
    r.chgtol = o->chgtol; r.gmin = o->gmin; r.diag_gmin = o->diag_gmin; r.xmu = o->xmu;
    r.src_fact = o->src_fact; r.itl1 = o->itl1; r.itl4 = o->itl4; r.method = o->method;
    r.engine = o->engine;
    return r;
}

int set_dev(int32_t* dev, int32_t v) {
    if (dev) *dev = v;
    return 0;
}
}  // namespace


// A CSC pattern as the reference's SparsityPattern::to_csc produces it:
// n >= 1, col_ptr[0] = 0, non-decreasing, rows in [0, n) strictly
// increasing within a column.
static void check_csc(int32_t n, const int32_t* col_ptr, const int32_t* row_idx) {
    if (n <= 0) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "pattern: n must be positive");
    if (!col_ptr || !row_idx) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "pattern: null pointer");
    if (col_ptr[0] != 0) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "pattern: col_ptr[0] != 0");
    for (int32_t c = 0; c < n; ++c) {
        if (col_ptr[c + 1] < col_ptr[c]) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "pattern: col_ptr decreasing");
        for (int32_t p = col_ptr[c]; p < col_ptr[c + 1]; ++p) {
            if (row_idx[p] < 0 || row_idx[p] >= n)
                throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "pattern: row index out of range");
            if (p > col_ptr[c] && row_idx[p] <= row_idx[p - 1])
                throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "pattern: rows not strictly increasing in a column");
        }
    }
}

extern "C" {

const char* nsa_last_error(void) { return g_err.c_str(); }

void nsa_default_options(nsa_options* o) {
    nsa::Options d;
    o->abstol = d.abstol; o->reltol = d.reltol; o->vntol = d.vntol; o->trtol = d.trtol;
    o->chgtol = d.chgtol; o->gmin = d.gmin; o->diag_gmin = d.diag_gmin; o->xmu = d.xmu;
    o->src_fact = d.src_fact; o->itl1 = d.itl1; o->itl4 = d.itl4; o->method = d.method;
    o->engine = d.engine;
}

nsa_circuit* nsa_circuit_create(void) { return new nsa_circuit(); }
void nsa_circuit_destroy(nsa_circuit* c) { delete c; }

int nsa_circuit_node(nsa_circuit* c, const char* name, int32_t* index) {
    return guard([&] { *index = c->c.node(name); return 0; });
}
int nsa_circuit_r(nsa_circuit* c, const char* name, int32_t a, int32_t b, double v, int32_t* dev) {
    return guard([&] { return set_dev(dev, c->c.R(name, a, b, v)); });
}
int nsa_circuit_c(nsa_circuit* c, const char* name, int32_t a, int32_t b, double v, int32_t* dev) {
    return guard([&] { return set_dev(dev, c->c.C(name, a, b, v)); });
}
int nsa_circuit_v(nsa_circuit* c, const char* name, int32_t p, int32_t n, double v, int32_t* dev) {
    return guard([&] { return set_dev(dev, c->c.V(name, p, n, v)); });
}
int nsa_circuit_v_pulse(nsa_circuit* c, const char* name, int32_t p, int32_t n,
                        const double pulse[7], int32_t* dev) {
    return guard([&] { return set_dev(dev, c->c.V_pulse(name, p, n, pulse)); });
}
int nsa_circuit_v_pwl(nsa_circuit* c, const char* name, int32_t p, int32_t n, int32_t npts,
                      const double* t, const double* v, int32_t* dev) {
    return guard([&] {
        return set_dev(dev, c->c.V_pwl(name, p, n, std::vector<double>(t, t + npts),
                                        std::vector<double>(v, v + npts)));
    });
}
int nsa_circuit_i(nsa_circuit* c, const char* name, int32_t p, int32_t n, double v, int32_t* dev) {
    return guard([&] { return set_dev(dev, c->c.I(name, p, n, v)); });
}
int nsa_circuit_model(nsa_circuit* c, const char* text) {
    return guard([&] { c->c.model(text); return 0; });
}
int nsa_circuit_d(nsa_circuit* c, const char* name, int32_t a, int32_t k, const char* model,
                  int32_t* dev) {
    return guard([&] { return set_dev(dev, c->c.D(name, a, k, model)); });
}
int nsa_circuit_m(nsa_circuit* c, const char* name, int32_t d, int32_t g, int32_t s, int32_t b,
                  const char* model, double w, double l, int32_t* dev) {
    return guard([&] { return set_dev(dev, c->c.M(name, d, g, s, b, model, w, l)); });
}
int nsa_circuit_bsim4_params(const nsa_circuit* c, int32_t dev, double* out) {
    return guard([&] {
        static_assert(sizeof(nsa::B4Param) == NSA_BSIM4_NPARAM * sizeof(double) + 8,
                      "B4Param layout");
        if (!out) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "out is NULL");
        const auto& devs = c->c.devices();
        if (dev < 0 || dev >= (int32_t)devs.size() || devs[dev].kind != nsa::DK_B4)
            throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "not a BSIM4v7 device");
        const nsa::B4Param p =
            nsa::bsim4_param(c->c.models().at(devs[dev].model), devs[dev], nsa::DeviceEnv{});
        std::memcpy(out, &p, NSA_BSIM4_NPARAM * sizeof(double));
        return 0;
    });
}
int nsa_circuit_bsim4_param_derivs(const nsa_circuit* c, int32_t dev, double* val, double* dw,
                                   double* dl) {
    return guard([&] {
        if (!val || !dw || !dl) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "output is NULL");
        const auto& devs = c->c.devices();
        if (dev < 0 || dev >= (int32_t)devs.size() || devs[dev].kind != nsa::DK_B4)
            throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "not a BSIM4v7 device");
        nsa::bsim4_param_derivs(c->c.models().at(devs[dev].model), devs[dev], nsa::DeviceEnv{}, val, dw, dl);
        return 0;
    });
}
int nsa_circuit_finalize(nsa_circuit* c) {
    return guard([&] { c->c.finalize(); return 0; });
}
int nsa_circuit_info(const nsa_circuit* c, int32_t out[5]) {
    return guard([&] {
        if (!c->c.finalized()) throw nsa::Error(nsa::NSA_ERR_LOGIC, "circuit not finalized");
        out[0] = c->c.num_vars();
        out[1] = c->c.num_nodes();
        out[2] = c->c.nnz();
        out[3] = c->c.num_states();
        out[4] = c->c.is_linear() ? 1 : 0;
        return 0;
    });
}
int nsa_circuit_pattern(const nsa_circuit* c, int32_t* col_ptr, int32_t* row_idx) {
    return guard([&] {
        std::memcpy(col_ptr, c->c.col_ptr().data(), sizeof(int32_t) * c->c.col_ptr().size());
        std::memcpy(row_idx, c->c.row_idx().data(), sizeof(int32_t) * c->c.row_idx().size());
        return 0;
    });
}
int nsa_circuit_load_order(const nsa_circuit* c, int32_t* order) {
    return guard([&] {
        std::memcpy(order, c->c.load_order().data(), sizeof(int32_t) * c->c.load_order().size());
        return 0;
    });
}
int nsa_circuit_branch(const nsa_circuit* c, int32_t dev, int32_t* branch) {
    return guard([&] { *branch = c->c.devices().at(dev).branch; return 0; });
}

int nsa_engine_create(nsa_circuit* c, int32_t batch, const nsa_options* opt, nsa_engine** out) {
    return guard([&] {
        auto e = std::make_unique<nsa_engine>();
        e->ckt = c;
        e->e = std::make_unique<nsa::BatchEngine>(c->c, batch, to_opts(opt));
        *out = e.release();
        return 0;
    });
}
void nsa_engine_destroy(nsa_engine* e) { delete e; }

int nsa_engine_set_values(nsa_engine* e, int32_t dev, const double* vals) {
    return guard([&] { e->e->set_device_values(dev, vals); return 0; });
}
int nsa_engine_set_model_values(nsa_engine* e, int32_t dev, const char* param, const double* vals) {
    return guard([&] {
        if (!param || !vals) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "param/vals is NULL");
        e->e->set_model_values(dev, param, vals);
        return 0;
    });
}
int nsa_engine_set_ctx(nsa_engine* e, const nsa_ctx* ctx, int32_t per_instance) {
    return guard([&] {
        static_assert(sizeof(nsa_ctx) == sizeof(nsa::DevCtx), "ctx layout");
        const int B = e->e->batch();
        std::vector<nsa::DevCtx> v(B);
        for (int i = 0; i < B; ++i)
            std::memcpy(&v[i], per_instance ? &ctx[i] : &ctx[0], sizeof(nsa::DevCtx));
        e->e->set_ctx(v);
        return 0;
    });
}
int nsa_engine_newton(nsa_engine* e, int32_t itl, int32_t* converged, int32_t* iterations) {
    return guard([&] {
        auto r = e->e->newton(itl);
        const int B = e->e->batch();
        if (converged) std::memcpy(converged, r.converged.data(), sizeof(int32_t) * B);
        if (iterations) std::memcpy(iterations, r.iterations.data(), sizeof(int32_t) * B);
        return 0;
    });
}
int nsa_engine_get_solution(nsa_engine* e, double* x) {
    return guard([&] { e->e->get_solution(x); return 0; });
}
int nsa_engine_set_solution(nsa_engine* e, const double* x) {
    return guard([&] { e->e->set_solution(x); return 0; });
}
int nsa_engine_get_state(nsa_engine* e, int32_t which, double* s) {
    return guard([&] {
        if (which < 0 || which > 3) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "state index 0..3");
        e->e->get_state(which, s);
        return 0;
    });
}
int nsa_engine_set_state(nsa_engine* e, int32_t which, const double* s) {
    return guard([&] {
        if (which < 0 || which > 3) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "state index 0..3");
        e->e->set_state(which, s);
        return 0;
    });
}
int nsa_engine_seed_states(nsa_engine* e) {
    return guard([&] { e->e->seed_states_from_state0(); return 0; });
}
int nsa_engine_rotate_states(nsa_engine* e) {
    return guard([&] { e->e->rotate_states(); return 0; });
}
int nsa_engine_trunc(nsa_engine* e, double* dt_min, double* dt_dev) {
    return guard([&] {
        if (!dt_min) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_engine_trunc: dt_min is NULL");
        e->e->trunc(dt_min, dt_dev);
        return 0;
    });
}
int nsa_engine_trunc_devices(nsa_engine* e, int32_t* dev_index) {
    return guard([&] {
        const auto& v = e->e->trunc_devices();
        if (dev_index) std::copy(v.begin(), v.end(), dev_index);
        return (int)v.size();
    });
}
int nsa_engine_reset_solver(nsa_engine* e) {
    return guard([&] { e->e->reset_solver(); return 0; });
}
// The engine's work runs on its own non-blocking stream, so every copy here
// is ordered on that stream and waited for before returning.
int nsa_engine_load(nsa_engine* e, double* values, double* rhs) {
    return guard([&] {
        e->e->load_only();
        const auto& c = e->e->circuit();
        const size_t B = e->e->batch();
        hipStream_t st = e->e->stream();
        NSA_HIP(hipMemcpyAsync(values, e->e->d_matrix(), sizeof(double) * c.nnz() * B,
                               hipMemcpyDeviceToHost, st));
        NSA_HIP(hipMemcpyAsync(rhs, e->e->d_rhs(), sizeof(double) * c.num_vars() * B,
                               hipMemcpyDeviceToHost, st));
        NSA_HIP(hipStreamSynchronize(st));
        return 0;
    });
}
int nsa_engine_solve_transposed(nsa_engine* e, double* b) {
    return guard([&] {
        if (!b) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_engine_solve_transposed: b is NULL");
        const size_t bytes = sizeof(double) * e->e->n() * e->e->batch();
        double* d = e->e->d_rhs();
        hipStream_t st = e->e->stream();
        NSA_HIP(hipMemcpyAsync(d, b, bytes, hipMemcpyHostToDevice, st));
        e->e->lu().solve_t(d);
        NSA_HIP(hipMemcpyAsync(b, d, bytes, hipMemcpyDeviceToHost, st));
        NSA_HIP(hipStreamSynchronize(st));
        return 0;
    });
}
namespace {
nsa::DevKind value_family(int32_t family) {
    if (family < NSA_FAMILY_R || family > NSA_FAMILY_I)
        throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "family must be NSA_FAMILY_R/C/V/I");
    return (nsa::DevKind)family;
}
}  // namespace
int nsa_engine_family_count(nsa_engine* e, int32_t family, int32_t* count) {
    return guard([&] { *count = e->e->family_count(value_family(family)); return 0; });
}
int nsa_engine_set_family_values(nsa_engine* e, int32_t family, const double* vals) {
    return guard([&] { e->e->set_family_values(value_family(family), vals); return 0; });
}
int nsa_engine_get_family_values(nsa_engine* e, int32_t family, double* vals) {
    return guard([&] { e->e->get_family_values(value_family(family), vals); return 0; });
}
int nsa_engine_dc_adjoint(nsa_engine* e, const double* dLdx, double* lambda, double* gR,
                          double* gC, double* gV, double* gI) {
    return guard([&] {
        if (!dLdx) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "dc_adjoint: dLdx is null");
        e->e->dc_adjoint(dLdx, lambda, gR, gC, gV, gI);
        return 0;
    });
}
int nsa_engine_tran_forward(nsa_engine* e, int32_t nsteps, double dt, const int32_t* probes,
                            int32_t nprobe, int32_t ckpt_every, double* probe_out) {
    return guard([&] {
        if (nprobe > 0 && !probes) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "tran_forward: probes is null");
        e->e->tran_forward(nsteps, dt, probes, nprobe, ckpt_every, probe_out);
        return 0;
    });
}
int nsa_engine_tran_adjoint(nsa_engine* e, const double* dLdprobe, double* gR, double* gC) {
    return guard([&] {
        if (!dLdprobe) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "tran_adjoint: dLdprobe is null");
        e->e->tran_adjoint(dLdprobe, gR, gC);
        return 0;
    });
}
int nsa_engine_adjoint_timing(nsa_engine* e, double out[2]) {
    return guard([&] {
        out[0] = e->e->ms_forward;
        out[1] = e->e->ms_adjoint;
        return 0;
    });
}
const char* nsa_engine_solver_name(nsa_engine* e) { return e->e->lu().name(); }
int nsa_engine_stats(nsa_engine* e, int64_t out[8]) {
    return guard([&] {
        auto& lu = e->e->lu();
        std::memset(out, 0, sizeof(int64_t) * 8);
        out[0] = lu.n_cpu_pivot;
        out[1] = lu.n_gpu_refactor;
        if (lu.structure()) {
            out[2] = lu.structure()->nlu();
            out[3] = lu.gpu()->nlevels_refactor();
            out[4] = lu.gpu()->nlevels_solve();
        }
        out[5] = lu.n_gpu_solve;
        out[6] = e->e->n_loads;
        out[7] = lu.num_groups();
        return 0;
    });
}

int nsa_engine_set_timing(nsa_engine* e, int32_t on) {
    return guard([&] { e->e->set_timing(on != 0); return 0; });
}

int nsa_engine_timing(nsa_engine* e, double out[5]) {
    return guard([&] {
        out[0] = e->e->ms_load_gpu;
        out[1] = e->e->lu().ms_refactor;
        out[2] = e->e->lu().ms_solve;
        out[3] = e->e->ms_bsim4_gpu;
        out[4] = (double)e->e->n_loads;
        return 0;
    });
}

int nsa_lu_create(int32_t batch, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                  int32_t is_linear, int32_t engine, nsa_lu** out) {
    return guard([&] {
        check_csc(n, col_ptr, row_idx);
        if (batch <= 0 || !out) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_lu_create: batch/out");
        auto s = std::make_unique<nsa_lu>();
        NSA_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->n = n;
        s->nnz = col_ptr[n];
        s->B = batch;
        s->cp.assign(col_ptr, col_ptr + n + 1);
        s->ri.assign(row_idx, row_idx + col_ptr[n]);
        s->lu = std::make_unique<nsa::BatchLu>(batch, s->stream);
        s->lu->symbolic(n, col_ptr, row_idx, is_linear != 0, engine);
        NSA_HIP(hipMalloc(&s->dA, sizeof(double) * std::max(1, s->nnz) * batch));
        NSA_HIP(hipMalloc(&s->db, sizeof(double) * std::max(1, n) * batch));
        *out = s.release();
        return 0;
    });
}
void nsa_lu_destroy(nsa_lu* s) { delete s; }

static int lu_factor(nsa_lu* s, const double* values, double gmin, int32_t* status, bool numeric) {
    NSA_HIP(hipMemcpyAsync(s->dA, values, sizeof(double) * s->nnz * s->B, hipMemcpyHostToDevice,
                           s->stream));
    std::vector<int32_t> st(s->B, 0);
    const int nsing = numeric ? s->lu->numeric(s->dA, gmin, st.data())
                              : s->lu->refactorize(s->dA, gmin, st.data());
    if (status) std::memcpy(status, st.data(), sizeof(int32_t) * s->B);
    return nsing ? 1 : 0;
}
int nsa_lu_numeric(nsa_lu* s, const double* values, double gmin, int32_t* status) {
    return guard([&] { return lu_factor(s, values, gmin, status, true); });
}
int nsa_lu_refactorize(nsa_lu* s, const double* values, double gmin, int32_t* status) {
    return guard([&] { return lu_factor(s, values, gmin, status, false); });
}
static int lu_solve(nsa_lu* s, double* rhs, bool t) {
    const size_t bytes = sizeof(double) * s->n * s->B;
    NSA_HIP(hipMemcpyAsync(s->db, rhs, bytes, hipMemcpyHostToDevice, s->stream));
    if (t) s->lu->solve_t(s->db); else s->lu->solve(s->db);
    NSA_HIP(hipMemcpyAsync(rhs, s->db, bytes, hipMemcpyDeviceToHost, s->stream));
    NSA_HIP(hipStreamSynchronize(s->stream));
    return 0;
}
int nsa_lu_set_equilibrate(nsa_lu* s, int32_t on) {
    return guard([&] {
        s->lu->set_equilibrate(on != 0);
        return 0;
    });
}
int nsa_lu_solve(nsa_lu* s, double* rhs) {
    return guard([&] { return lu_solve(s, rhs, false); });
}
int nsa_lu_solve_transposed(nsa_lu* s, double* rhs) {
    return guard([&] { return lu_solve(s, rhs, true); });
}
int nsa_lu_solve_noalloc(nsa_lu* s, double* rhs) {
    return guard([&] {
        const size_t bytes = sizeof(double) * (size_t)s->n * s->B;
        if (bytes > nsa::kNoAllocScratchBytes)
            throw nsa::Error(nsa::NSA_ERR_ARGUMENT,
                             "nsa_lu_solve_noalloc: n x batch = " + std::to_string((size_t)s->n * s->B) +
                                 " values exceed the " + std::to_string(nsa::kNoAllocScratchBytes >> 20) +
                                 " MiB device scratch");
        if (!s->lu->factored()) throw nsa::Error(nsa::NSA_ERR_LOGIC, "nsa_lu_solve_noalloc: numeric() not called");
        // the staging vector is the library's static device array (its
        // address on this device), not the handle's buffer
        double* d = nsa::noalloc_scratch();
        NSA_HIP(hipMemcpyAsync(d, rhs, bytes, hipMemcpyHostToDevice, s->stream));
        s->lu->solve(d);
        NSA_HIP(hipMemcpyAsync(rhs, d, bytes, hipMemcpyDeviceToHost, s->stream));
        NSA_HIP(hipStreamSynchronize(s->stream));
        return 0;
    });
}
// The complex paths.  Host arrays in the real calls' layout, instance
// fastest: value row r of instance i at r * B + i, a complex value's real
// and imaginary parts in rows 2k and 2k + 1.
static std::vector<double> to_blocks(const double* v, int rows, int B) {
    std::vector<double> out((size_t)rows * B);
    for (int i = 0; i < B; ++i)
        for (int r = 0; r < rows; ++r) out[(size_t)i * rows + r] = v[(size_t)r * B + i];
    return out;
}
static void from_blocks(const std::vector<double>& b, int rows, int B, double* v) {
    for (int i = 0; i < B; ++i)
        for (int r = 0; r < rows; ++r) v[(size_t)r * B + i] = b[(size_t)i * rows + r];
}
int nsa_lu_numeric_complex(nsa_lu* s, const double* ax, int32_t* status) {
    return guard([&] {
        if (!ax) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_lu_numeric_complex: ax");
        const auto blk = to_blocks(ax, 2 * s->nnz, s->B);
        return s->complex_lu().numeric_complex(blk.data(), status) ? 1 : 0;
    });
}
int nsa_lu_refactorize_complex(nsa_lu* s, const double* ax, int32_t* status) {
    return guard([&] {
        if (!ax) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_lu_refactorize_complex: ax");
        const auto blk = to_blocks(ax, 2 * s->nnz, s->B);
        return s->complex_lu().refactorize_complex(blk.data(), status) ? 1 : 0;
    });
}
int nsa_lu_solve_complex(nsa_lu* s, double* rhs) {
    return guard([&] {
        if (!rhs) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_lu_solve_complex: rhs");
        auto blk = to_blocks(rhs, 2 * s->n, s->B);
        s->complex_lu().solve_complex(blk.data());
        from_blocks(blk, 2 * s->n, s->B, rhs);
        return 0;
    });
}
int nsa_lu_ac_sweep(nsa_lu* s, const double* g, const double* c, const double* freqs, int32_t nf,
                    const double* rhs, double* out, int32_t* status) {
    return guard([&] {
        if (!g || !c || !freqs || !rhs || !out || nf <= 0)
            throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_lu_ac_sweep: arguments");
        return s->complex_lu().sweep(g, c, freqs, nf, rhs, out, status) ? 1 : 0;
    });
}
int nsa_lu_ac_stats(nsa_lu* s, double out[3]) {
    return guard([&] {
        if (!out) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_lu_ac_stats: out");
        const nsa::AcLu& a = s->complex_lu();
        out[0] = a.last_order_s;
        out[1] = a.last_device_s;
        out[2] = a.last_total_s;
        return 0;
    });
}
const char* nsa_lu_name(nsa_lu* s) { return s->lu->name(); }
int nsa_lu_info(nsa_lu* s, int32_t out[4]) {
    return guard([&] {
        if (!s->lu->structure()) throw nsa::Error(nsa::NSA_ERR_LOGIC, "numeric() not called");
        out[0] = s->n;
        out[1] = s->lu->structure()->nlu();
        out[2] = s->lu->gpu()->nlevels_refactor();
        out[3] = s->lu->gpu()->nlevels_solve();
        return 0;
    });
}
int nsa_lu_structure(nsa_lu* s, int32_t* row_perm, int32_t* col_perm, int32_t* seg_ptr,
                     int32_t* rows) {
    return guard([&] {
        const auto* st = s->lu->structure();
        if (!st) throw nsa::Error(nsa::NSA_ERR_LOGIC, "numeric() not called");
        std::memcpy(row_perm, st->row_perm.data(), sizeof(int32_t) * st->n);
        std::memcpy(col_perm, st->col_perm.data(), sizeof(int32_t) * st->n);
        std::memcpy(seg_ptr, st->cptr.data(), sizeof(int32_t) * (st->n + 1));
        std::memcpy(rows, st->rows.data(), sizeof(int32_t) * st->nlu());
        return 0;
    });
}

int nsa_pivot_structure(int32_t engine, int32_t n, const int32_t* col_ptr, const int32_t* row_idx,
                        const double* values, double diag_gmin, int32_t* nnz_lu,
                        int32_t* row_perm, int32_t* col_perm, int32_t* seg_ptr, int32_t* nU,
                        int32_t* rows) {
    return guard([&] {
        check_csc(n, col_ptr, row_idx);
        if (!values || !nnz_lu) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_pivot_structure: null pointer");
        nsa::LuStructure st;
        if (engine == nsa::ENGINE_AMDLU) {
            auto q = nsa::min_degree_order(n, col_ptr, row_idx);
            if (!nsa::gp_factor_structure(n, col_ptr, row_idx, values, diag_gmin, q, st)) return 1;
        } else if (engine == nsa::ENGINE_MARKOWITZ) {
            nsa::Sparse13 s13(n, col_ptr, row_idx);
            if (s13.numeric(values, diag_gmin) != 0) return 1;
            st = s13.structure();
        } else if (engine == 3) {
            if (diag_gmin != 0.0)
                throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_pivot_structure: no diag_gmin on a complex matrix");
            nsa::Sparse13 s13(n, col_ptr, row_idx);
            if (s13.numeric_complex(values) != 0) return 1;
            st = s13.structure();
        } else {
            throw nsa::Error(nsa::NSA_ERR_ARGUMENT,
                             "engine must be 1 (markowitz), 2 (amdlu) or 3 (complex markowitz)");
        }
        if (nnz_lu) *nnz_lu = st.nlu();
        if (rows) {
            std::memcpy(row_perm, st.row_perm.data(), sizeof(int32_t) * n);
            std::memcpy(col_perm, st.col_perm.data(), sizeof(int32_t) * n);
            std::memcpy(seg_ptr, st.cptr.data(), sizeof(int32_t) * (n + 1));
            std::memcpy(nU, st.nU.data(), sizeof(int32_t) * n);
            std::memcpy(rows, st.rows.data(), sizeof(int32_t) * st.nlu());
        }
        return 0;
    });
}

int nsa_min_degree_order(int32_t n, const int32_t* col_ptr, const int32_t* row_idx, int32_t* perm) {
    return guard([&] {
        check_csc(n, col_ptr, row_idx);
        if (!perm) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_min_degree_order: null perm");
        auto q = nsa::min_degree_order(n, col_ptr, row_idx);
        std::memcpy(perm, q.data(), sizeof(int32_t) * q.size());
        return 0;
    });
}

/* ---- bundles of circuits of different topologies ---- */

int nsa_bundle_create(nsa_circuit* const* circuits, int32_t count, const nsa_options* opt,
                      nsa_bundle** out) {
    return guard([&] {
        if (!circuits || count <= 0 || !out)
            throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_bundle_create: need count > 0 circuits");
        std::vector<const nsa::Circuit*> v(count);
        for (int i = 0; i < count; ++i) {
            if (!circuits[i]) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_bundle_create: null circuit");
            v[i] = &circuits[i]->c;
        }
        auto h = std::make_unique<nsa_bundle>();
        h->b = std::make_unique<nsa::Bundle>(v, to_opts(opt));
        *out = h.release();
        return 0;
    });
}
void nsa_bundle_destroy(nsa_bundle* b) { delete b; }

int nsa_bundle_info(nsa_bundle* b, int32_t out[4]) {
    return guard([&] {
        out[0] = b->b->size();
        out[1] = b->b->total_n();
        out[2] = b->b->total_nnz();
        out[3] = b->b->num_states();
        return 0;
    });
}
int nsa_bundle_offsets(nsa_bundle* b, int32_t* xoff, int32_t* aoff) {
    return guard([&] {
        const int n = b->b->size() + 1;
        if (xoff) std::memcpy(xoff, b->b->xoff().data(), sizeof(int32_t) * n);
        if (aoff) std::memcpy(aoff, b->b->aoff().data(), sizeof(int32_t) * n);
        return 0;
    });
}
int nsa_bundle_dc(nsa_bundle* b, int32_t mode, int32_t* converged, int32_t* iterations) {
    return guard([&] {
        std::vector<int32_t> c, it;
        b->b->dc_op(mode, c, it);
        const int n = b->b->size();
        if (converged) std::memcpy(converged, c.data(), sizeof(int32_t) * n);
        if (iterations) std::memcpy(iterations, it.data(), sizeof(int32_t) * n);
        return 0;
    });
}
int nsa_bundle_get_solution(nsa_bundle* b, double* x) {
    return guard([&] {
        if (!x) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_bundle_get_solution: x is NULL");
        b->b->get_solution(x);
        return 0;
    });
}
int nsa_bundle_set_solution(nsa_bundle* b, const double* x) {
    return guard([&] {
        if (!x) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_bundle_set_solution: x is NULL");
        b->b->set_solution(x);
        return 0;
    });
}
int nsa_bundle_tran_begin(nsa_bundle* b, double tstep, double tstop, const int32_t* probe_ptr,
                          const int32_t* probes, int32_t max_points) {
    return guard([&] {
        if (probe_ptr && !probes && probe_ptr[b->b->size()] > 0)
            throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "nsa_bundle_tran_begin: probes is NULL");
        b->b->tran_begin(tstep, tstop, probe_ptr, probes, max_points);
        return 0;
    });
}
int nsa_bundle_tran_advance(nsa_bundle* b, double t_until, int32_t* rounds) {
    return guard([&] {
        const int r = b->b->tran_advance(t_until);
        if (rounds) *rounds = r;
        return b->b->tran_done() ? 1 : 0;
    });
}
int nsa_bundle_tran_status(nsa_bundle* b, int32_t inst, int64_t out[8]) {
    return guard([&] {
        if (inst < 0 || inst >= b->b->size()) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "instance out of range");
        const nsa::TranDev& c = b->b->tran_state(inst);
        out[0] = c.step_count;
        out[1] = c.rejected;
        out[2] = c.failed ? 1 : 0;
        out[3] = c.finished ? 1 : 0;
        out[4] = c.newton_iters;
        out[5] = c.all_iters;
        out[6] = c.total_iterations;
        out[7] = b->b->dc_iterations().empty() ? 0 : b->b->dc_iterations()[inst];
        return 0;
    });
}
int nsa_bundle_tran_steps(nsa_bundle* b, int32_t inst, double* t, double* dt, int32_t* order,
                          int32_t* method, int32_t* iters) {
    return guard([&] {
        if (inst < 0 || inst >= b->b->size()) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "instance out of range");
        const std::vector<nsa::TranStep> st = b->b->tran_steps(inst);
        for (size_t k = 0; k < st.size(); ++k) {
            if (t) t[k] = st[k].t;
            if (dt) dt[k] = st[k].dt;
            if (order) order[k] = st[k].order;
            if (method) method[k] = st[k].method;
            if (iters) iters[k] = st[k].iters;
        }
        return (int)st.size();
    });
}
int nsa_bundle_tran_probes(nsa_bundle* b, int32_t inst, int32_t* rows, double* out) {
    return guard([&] {
        int r = 0;
        b->b->tran_probes(inst, &r, out);
        if (rows) *rows = r;
        return 0;
    });
}
int nsa_bundle_stats(nsa_bundle* b, int64_t out[8]) {
    return guard([&] {
        const nsa::Bundle& x = *b->b;
        out[0] = x.n_iterations;
        out[1] = x.n_cpu_pivot;
        out[2] = x.num_plans();
        out[3] = x.n_graph_launches;
        out[4] = x.n_host_iterations;
        int64_t lev = 0, slev = 0;
        for (int p = 0; p < x.num_plans(); ++p) {
            lev = std::max<int64_t>(lev, x.plans().levels(p));
            slev = std::max<int64_t>(slev, x.plans().solve_levels(p));
        }
        out[5] = lev;
        out[6] = slev;
        out[7] = x.plans().max_nlu();
        return 0;
    });
}
int nsa_bundle_set_timing(nsa_bundle* b, int32_t on) {
    return guard([&] { b->b->set_timing(on != 0); return 0; });
}
int nsa_bundle_timing(nsa_bundle* b, double out[5]) {
    return guard([&] {
        const nsa::Bundle& x = *b->b;
        out[0] = x.ms_load;
        out[1] = x.ms_lu;
        out[2] = x.ms_post;
        out[3] = x.ms_bsim4;
        out[4] = (double)x.n_timed;
        return 0;
    });
}
int nsa_bundle_set_graph(nsa_bundle* b, int32_t on) {
    return guard([&] { b->b->use_graph = on != 0; return 0; });
}
int nsa_bundle_tick_stats(nsa_bundle* b, double out[4]) {
    return guard([&] {
        const nsa::Bundle& x = *b->b;
        out[0] = (double)x.n_ticks;
        out[1] = (double)x.n_tick_batches;
        out[2] = x.tick_enqueue_s;
        out[3] = x.tick_wait_s;
        return 0;
    });
}
int nsa_bundle_set_tuning(nsa_bundle* b, int32_t on) {
    return guard([&] {
        b->b->tuning = on != 0;
        return 0;
    });
}
int nsa_bundle_tuning(nsa_bundle* b, double* out, int32_t max_out) {
    return guard([&] {
        const auto& r = b->b->tune_report;
        if (out)
            for (int k = 0; k < std::min<int>(max_out, (int)r.size()); ++k) out[k] = r[k];
        return (int)r.size();
    });
}
int nsa_bundle_set_device_control(nsa_bundle* b, int32_t on) {
    return guard([&] { b->b->device_control = on != 0; return 0; });
}
int nsa_bundle_dc_ladder(nsa_bundle* b, int32_t flags, int32_t* converged, int32_t* iterations,
                         int32_t* stage) {
    return guard([&] {
        std::vector<int32_t> c, it, st;
        // MODETRANOP (solve_transient's preamble) or MODEDCOP (solve_dc), | MODEINITJCT / MODEINITFLOAT
        const int op = (flags & 2) ? 0x10 : 0x20;
        // bit 0: one method alone (bits 4-5: 0 gmin, 1 source, 2 true gmin stepping)
        int alone = nsa::LADDER_NONE;
        if (flags & 1) {
            const int m = (flags >> 4) & 3;
            alone = m == 0 ? nsa::LADDER_GMIN : m == 1 ? nsa::LADDER_SOURCE : nsa::LADDER_TRUE_GMIN;
        }
        b->b->op_ladder(op | 0x200, op | 0x100, c, it, st, alone);
        const int n = b->b->size();
        if (converged) std::memcpy(converged, c.data(), sizeof(int32_t) * n);
        if (iterations) std::memcpy(iterations, it.data(), sizeof(int32_t) * n);
        if (stage) std::memcpy(stage, st.data(), sizeof(int32_t) * n);
        return 0;
    });
}
int nsa_bundle_device_offsets(nsa_bundle* b, int32_t* doff) {
    return guard([&] {
        for (int i = 0; i <= b->b->size(); ++i) doff[i] = i < b->b->size() ? b->b->device_offset(i) : b->b->total_devices();
        return 0;
    });
}
int nsa_bundle_set_device_values(nsa_bundle* b, const double* value, const double* w, const double* l) {
    return guard([&] {
        b->b->set_device_values(value, w, l);
        return 0;
    });
}
int nsa_bundle_get_device_values(nsa_bundle* b, double* value, double* w, double* l) {
    return guard([&] {
        b->b->get_device_values(value, w, l);
        return 0;
    });
}
int nsa_bundle_dc_adjoint(nsa_bundle* b, const double* dLdx, double* lambda, double* g_value, double* g_w,
                          double* g_l) {
    return guard([&] {
        if (!dLdx) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "dLdx is NULL");
        b->b->dc_adjoint(dLdx, lambda, g_value, g_w, g_l);
        return 0;
    });
}
int nsa_bundle_plan_info(nsa_bundle* b, int32_t inst, int64_t out[9]) {
    return guard([&] {
        if (inst < 0 || inst >= b->b->size()) throw nsa::Error(nsa::NSA_ERR_ARGUMENT, "instance out of range");
        const int p = b->b->plan_of(inst);
        for (int k = 0; k < 9; ++k) out[k] = 0;
        out[0] = p;
        if (p < 0) return 0;
        const nsa::LuStructure& s = b->b->plans().structure(p);
        out[1] = s.n;
        out[2] = s.nnz_a;
        out[3] = s.nlu();
        out[4] = b->b->plans().levels(p);
        out[5] = b->b->plans().bwd_levels(p);
        // left-looking operation counts of the Sparse 1.3 factor (factor.cpp:1010-1026)
        // and of the forward and backward substitutions (solve.cpp:6-47)
        int64_t f = s.n, off = 0;
        for (int j = 0; j < s.n; ++j) {
            off += s.cptr[j + 1] - s.cptr[j] - 1;
            for (int q = s.cptr[j]; q < s.cptr[j] + s.nU[j]; ++q) {
                const int r = s.rows[q];
                f += 1 + 2 * (s.cptr[r + 1] - s.cptr[r] - s.nU[r] - 1);
            }
        }
        out[6] = f;
        out[7] = 2 * off + s.n;
        // plan words a factor-and-solve launch reads: slot positions, gmin
        // positions, column segments and U counts, rows and columns of the
        // positions, both permutations, and the level / task / backward block
        const int32_t* w = b->b->plans().host_words().data() + b->b->plans().host_offsets()[p];
        out[8] = (int64_t)s.nnz_a + 5 * s.n + 1 + 2 * (int64_t)s.nlu() + w[nsa::SH_HOTLEN];
        return 0;
    });
}

}  // extern "C"
