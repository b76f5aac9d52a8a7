// TEST INFRASTRUCTURE: the compiled reference with neospice_amd behind its
// solver seam.  libneospice_ref_gpu.so is the reference's own src/ (the
// same objects as libneospice_ref.so) linked with
// --wrap=neospice::make_solver(int, bool), so that every DC and transient
// analysis of the reference (dc.cpp:80, 546; transient.cpp:713, 726) asks
// this function for its solver.  When switched on it returns the adapter a
// maintainer would add (neospice_amd/integration/mi355x_solver.hpp), and
// the reference's unmodified newton_solve, DC ladder, solve_transient and
// Simulator run on the MI355X; otherwise the reference's own make_solver.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <memory>
#include <vector>

#include "core/solver_iface.hpp"
#include "mi355x_solver.hpp"

namespace {
std::atomic<int> g_gpu{0};
std::atomic<int> g_made{0};
}  // namespace

extern "C" std::unique_ptr<neospice::ISolver> __real__ZN8neospice11make_solverEib(int num_vars, bool is_linear);

extern "C" std::unique_ptr<neospice::ISolver> __wrap__ZN8neospice11make_solverEib(int num_vars, bool is_linear) {
    if (g_gpu.load()) {
        ++g_made;
        return std::make_unique<neospice::Mi355xSolver>(0, is_linear);
    }
    return __real__ZN8neospice11make_solverEib(num_vars, is_linear);
}

extern "C" {
// 1: analyses from now on solve through Mi355xSolver; 0: the reference's own solvers.
void ref_set_gpu_solver(int on) { g_gpu.store(on ? 1 : 0); }
// Mi355xSolvers made so far.
int ref_gpu_solvers_made() { return g_made.load(); }
}

// The adapter's complex paths on a CSC pattern, as the AC analysis drives a
// solver (ac.cpp:215-222): numeric_complex(ax0) and solve_complex(rhs) into
// x0, then refactorize_complex(ax1) and solve_complex(rhs) into x1 (2 n
// each).  Returns refactorize_complex's singular flag, or -1 with the
// message in *err (256 bytes) on an exception.
extern "C" int ref_mi355x_complex(int n, const int32_t* col_ptr, const int32_t* row_idx, const double* ax0,
                                  const double* ax1, const double* rhs, double* x0, double* x1, char* err) {
    try {
        std::vector<std::pair<int32_t, int32_t>> e;
        for (int j = 0; j < n; ++j)
            for (int p = col_ptr[j]; p < col_ptr[j + 1]; ++p) e.emplace_back(row_idx[p], j);
        neospice::SparsityPattern pat(n, std::move(e));
        const int nnz = col_ptr[n];
        neospice::Mi355xSolver s(0, true);
        s.symbolic(pat);
        s.numeric_complex(pat, std::vector<double>(ax0, ax0 + 2 * nnz));
        std::vector<double> r(rhs, rhs + 2 * n);
        s.solve_complex(r);
        std::copy(r.begin(), r.end(), x0);
        const bool sing = s.refactorize_complex(std::vector<double>(ax1, ax1 + 2 * nnz));
        r.assign(rhs, rhs + 2 * n);
        s.solve_complex(r);
        std::copy(r.begin(), r.end(), x1);
        return sing ? 1 : 0;
    } catch (const std::exception& ex) {
        std::snprintf(err, 256, "%s", ex.what());
        return -1;
    }
}
