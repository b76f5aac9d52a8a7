// The drop-in for neospice's linear-solver seam: ISolver
// (src/core/solver_iface.hpp:16-41) over this library's C-ABI (nsa_lu_*),
// for one circuit.  A maintainer adds this header to neospice and returns a
// Mi355xSolver from make_solver (src/core/make_solver.cpp:156-176) when the
// GPU is asked for; newton_solve, the DC ladder and solve_transient then
// factor and solve on the MI355X with the reference's own pivot orders,
// fill and values, bit for bit.
//
// Conventions are NeoSolver's: numeric()/refactorize() return true on a
// singular matrix; misuse ordering throws std::logic_error, a size mismatch
// std::invalid_argument, other failures std::runtime_error
// (neo_solver.cpp:152, 170-202).  engine 0 is make_solver's own policy
// (AMD-LU iff n >= 256 and linear, with the AutoFallbackSolver's switch to
// Markowitz on a singular first factorization; make_solver.cpp:71-176).
// The complex (AC) paths are NeoSolver's, as AmdLuSolver's delegate to one
// (amd_lu_solver.cpp:347-349): the complex Markowitz ordering on the host,
// the complex factor and solves on the MI355X (nsa_lu_*_complex).
#pragma once
#include <neospice_amd.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "core/solver_iface.hpp"

namespace neospice {

class Mi355xSolver final : public ISolver {
public:
    explicit Mi355xSolver(int engine, bool is_linear) : engine_(engine), is_linear_(is_linear) {}
    ~Mi355xSolver() override {
        if (h_) nsa_lu_destroy(h_);
    }
    Mi355xSolver(const Mi355xSolver&) = delete;
    Mi355xSolver& operator=(const Mi355xSolver&) = delete;

    void symbolic(const SparsityPattern& p) override {
        const auto csc = p.to_csc();
        if (h_) nsa_lu_destroy(h_);
        h_ = nullptr;
        n_ = p.size();
        nnz_ = (int)csc.row_idx.size();
        factored_ = false;
        complex_factored_ = false;
        check(nsa_lu_create(1, n_, csc.col_ptr.data(), csc.row_idx.data(), is_linear_ ? 1 : 0, engine_, &h_));
    }
    bool numeric(const SparsityPattern&, const NumericMatrix& m, double diag_gmin) override {
        if (!h_) throw std::logic_error("Mi355xSolver::numeric: symbolic() not called");
        int32_t st = 0;
        const bool singular = check(nsa_lu_numeric(h_, m.data(), diag_gmin, &st)) == 1;
        factored_ = !singular;
        return singular;
    }
    bool refactorize(const NumericMatrix& m, double diag_gmin) override {
        if (!h_ || !factored_) throw std::logic_error("Mi355xSolver::refactorize: numeric() not called");
        int32_t st = 0;
        return check(nsa_lu_refactorize(h_, m.data(), diag_gmin, &st)) == 1;
    }
    void solve(std::vector<double>& rhs) override {
        if (!h_ || !factored_) throw std::logic_error("Mi355xSolver::solve: numeric() not called");
        if ((int)rhs.size() != n_) throw std::invalid_argument("Mi355xSolver::solve: rhs size");
        check(nsa_lu_solve(h_, rhs.data()));
    }
    // A^T x = b on the same factors (the adjoint's solve).
    void solve_transposed(std::vector<double>& rhs) {
        if (!h_ || !factored_) throw std::logic_error("Mi355xSolver::solve_transposed: numeric() not called");
        if ((int)rhs.size() != n_) throw std::invalid_argument("Mi355xSolver::solve_transposed: rhs size");
        check(nsa_lu_solve_transposed(h_, rhs.data()));
    }
    // Honoured on Markowitz solvers, as NeoSolver does (the DC ladder's last
    // fallback, dc.cpp:367-380).
    void set_equilibrate(bool on) override {
        if (!h_) throw std::logic_error("Mi355xSolver::set_equilibrate: symbolic() not called");
        check(nsa_lu_set_equilibrate(h_, on ? 1 : 0));
    }
    // ax: 2 nnz values, (re, im) per CSC slot.  Like NeoSolver's, it reports
    // no singularity, and the real factors are gone after it
    // (neo_solver.cpp:228-245).
    void numeric_complex(const SparsityPattern&, const std::vector<double>& ax) override {
        if (!h_) throw std::logic_error("Mi355xSolver::numeric_complex: symbolic() not called");
        if ((int)ax.size() != 2 * nnz_) throw std::invalid_argument("Mi355xSolver::numeric_complex: ax size mismatch");
        int32_t st = 0;
        check(nsa_lu_numeric_complex(h_, ax.data(), &st));
        complex_factored_ = true;
        factored_ = false;
    }
    bool refactorize_complex(const std::vector<double>& ax) override {
        if (!h_) throw std::logic_error("Mi355xSolver::refactorize_complex: symbolic() not called");
        if (!complex_factored_)
            throw std::logic_error("Mi355xSolver::refactorize_complex: numeric_complex() not called");
        if ((int)ax.size() != 2 * nnz_)
            throw std::invalid_argument("Mi355xSolver::refactorize_complex: ax size mismatch");
        int32_t st = 0;
        return check(nsa_lu_refactorize_complex(h_, ax.data(), &st)) == 1;
    }
    void solve_complex(std::vector<double>& rhs) override {
        if (!h_) throw std::logic_error("Mi355xSolver::solve_complex: symbolic() not called");
        if (!complex_factored_) throw std::logic_error("Mi355xSolver::solve_complex: numeric_complex() not called");
        if ((int)rhs.size() != 2 * n_) throw std::invalid_argument("Mi355xSolver::solve_complex: rhs size must be 2*n");
        check(nsa_lu_solve_complex(h_, rhs.data()));
    }
    const char* name() const override { return h_ ? nsa_lu_name(h_) : "mi355x"; }

private:
    static int check(int rc) {
        if (rc >= 0) return rc;
        const std::string msg = nsa_last_error();
        if (rc == -1) throw std::logic_error(msg);       // misuse order
        if (rc == -2) throw std::invalid_argument(msg);  // bad argument or size
        throw std::runtime_error(msg);
    }
    nsa_lu* h_ = nullptr;
    int n_ = 0, nnz_ = 0;
    bool complex_factored_ = false;
    int engine_;
    bool is_linear_;
    bool factored_ = false;
};

}  // namespace neospice
