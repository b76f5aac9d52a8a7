// CPU symbolic/numeric factorizations that fix the pivot order and fill
// pattern the GPU then replays.  These run on the host once per circuit
// (and again only when a pivot stops being acceptable), exactly where the
// reference makes its value-dependent pivot choices.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "lu_structure.hpp"

namespace nsa {

// Exact minimum-degree column ordering on the symmetrized pattern, with the
// reference's dense-node postponement (src/core/amd.cpp:14-143).
std::vector<int32_t> min_degree_order(int n, const int32_t* col_ptr, const int32_t* row_idx);

// Gilbert-Peierls left-looking LU with threshold partial pivoting and
// diagonal preference (AmdLuSolver::factor_, src/core/amd_lu_solver.cpp:74-197).
// Returns false when singular.  On success `out` holds the structure.
bool gp_factor_structure(int n, const int32_t* col_ptr, const int32_t* row_idx,
                         const double* values, double diag_gmin, const std::vector<int32_t>& q,
                         LuStructure& out);

// Sparse 1.3 Markowitz ordering and factorization, as NeoSolver::numeric
// runs it on a fresh matrix: MNA twin pre-ordering (utils.cpp:8-89),
// add_diag_gmin, then order_and_factor with rel 1e-3, abs 1e-13 and
// diagonal pivoting (factor.cpp:911-978).  Returns 0 ok, 1 singular.
// On success `out` holds the structure (pivot order + fill).
class Sparse13 {
public:
    explicit Sparse13(int n, const int32_t* col_ptr, const int32_t* row_idx);
    ~Sparse13();
    Sparse13& operator=(const Sparse13&) = delete;

    // First call: full ordering.  Later calls reuse the order with the
    // threshold check and re-pivot from the first failing step
    // (factor.cpp:929-947).
    int numeric(const double* values, double diag_gmin);
    // NeoSolver::numeric_complex (neo_solver.cpp:228-245): ax holds (real,
    // imaginary) per slot; the same ordering on element magnitudes |re| +
    // |im| with complex eliminations (factor.cpp:769-804), no diag_gmin.
    // The matrix stays complex from then on.
    int numeric_complex(const double* ax);
    LuStructure structure() const;
    int fillins() const;
    // Deep copy of the factorization state (pivot order, fill, counts).
    std::unique_ptr<Sparse13> clone() const;

private:
    Sparse13(const Sparse13& o);
    struct Impl;
    Impl* impl_;
    int n_;
    std::vector<int32_t> acp_, ari_;
};

}  // namespace nsa
