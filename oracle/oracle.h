/* CPU restatement of the reference's numeric LU path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU-baseline
 * leg of bench.py call this as the checker.  The product never links it.
 *
 * Every routine cites the reference function it restates (paths relative
 * to the reference repository root).  Index conventions:
 *   - CSC patterns are 0-based: col_ptr[n+1], row_idx[nnz], values in the
 *     same slot order as the reference's NumericMatrix (sorted by col, row).
 *   - "Sparse-1.3 plans" use the reference's internal 1-based indices.
 */
#ifndef NEOSPICE_ORACLE_H
#define NEOSPICE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Minimum-degree ordering, src/core/amd.cpp:14-143.  perm[k] = column
 * eliminated at step k.  Returns n (0 for n <= 0). */
int orc_amd_order(int n, const int32_t* col_ptr, const int32_t* row_idx,
                  int32_t* perm);

/* Gilbert-Peierls LU with threshold partial pivoting,
 * src/core/amd_lu_solver.cpp:74-197 (factor_).  q is the column order from
 * orc_amd_order.  Outputs: pinv[n], Lp/Up[n+1], Li/Lx (cap lcap), Ui/Ux (cap
 * ucap).  L row indices are returned in pivot-step space (as the reference
 * remaps them at :181).  Returns 0 ok, 1 singular, -1 capacity exceeded. */
int orc_gp_factor(int n, const int32_t* col_ptr, const int32_t* row_idx,
                  const double* values, double diag_gmin, double pivot_tol,
                  const int32_t* q, int32_t* pinv, int32_t* Lp, int32_t* Li,
                  double* Lx, int lcap, int32_t* Up, int32_t* Ui, double* Ux,
                  int ucap);

/* Fixed-structure replay, src/core/amd_lu_solver.cpp:214-271.  Overwrites
 * Lx/Ux in place.  Returns 1 when the replay succeeded, 0 when the pivot
 * growth check (|ukk| < tol * colmax, or ukk == 0) rejects it. */
int orc_gp_refactor(int n, const int32_t* col_ptr, const int32_t* row_idx,
                    const double* values, double diag_gmin, double pivot_tol,
                    const int32_t* q, const int32_t* pinv, const int32_t* Lp,
                    const int32_t* Li, double* Lx, const int32_t* Up,
                    const int32_t* Ui, double* Ux);

/* Solve on GP factors, src/core/amd_lu_solver.cpp:303-344.  rhs in place. */
void orc_gp_solve(int n, const int32_t* q, const int32_t* pinv,
                  const int32_t* Lp, const int32_t* Li, const double* Lx,
                  const int32_t* Up, const int32_t* Ui, const double* Ux,
                  double* rhs);

/* Transposed solve on GP factors: A^T x = b with P A Q = L U, so
 * x = P^T L^-T U^-T Q^T b.  The reference has no transposed AMD-LU solve;
 * this follows from the factorization algebra (U^T forward by rows of U,
 * unit L^T backward). */
void orc_gp_solve_t(int n, const int32_t* q, const int32_t* pinv,
                    const int32_t* Lp, const int32_t* Li, const double* Lx,
                    const int32_t* Up, const int32_t* Ui, const double* Ux,
                    double* rhs);

/* A Sparse-1.3 factored structure in the reference's internal numbering:
 * for internal column j (1..n) the elements of the linked column list, in
 * list order (ascending internal row), at cptr[j-1]..cptr[j]-1 of rows[]
 * and vals[].  diag values hold reciprocal pivots (factor.cpp:998/1026). */
typedef struct {
    int n;
    const int32_t* i2e_row; /* n, external 1-based row of internal row i */
    const int32_t* i2e_col; /* n, external 1-based col of internal col j */
    const int32_t* cptr;    /* n+1 */
    const int32_t* rows;    /* internal 1-based rows */
    double* vals;
} orc_s13;

/* Left-looking refactor on the fixed order, src/solver/factor.cpp:980-1052
 * (SparseMatrix::factor) after NeoSolver::load_real (neo_solver.cpp:45-50)
 * and add_diag_gmin (utils.cpp:91-97).  A is the pattern CSC in external
 * 0-based indexing.  Returns 0 ok, or the 1-based internal step of an exact
 * zero pivot (zero_pivot, factor.cpp:18-22). */
int orc_s13_refactor(orc_s13* f, const int32_t* col_ptr, const int32_t* row_idx,
                     const double* values, double diag_gmin);

/* SparseMatrix::solve, src/solver/solve.cpp:6-47.  rhs (0-based, n) in,
 * x out (0-based, n). */
void orc_s13_solve(const orc_s13* f, const double* rhs, double* x);

/* SparseMatrix::solve_transposed, src/solver/solve.cpp:102-141. */
void orc_s13_solve_t(const orc_s13* f, const double* rhs, double* x);

#ifdef __cplusplus
}
#endif
#endif
