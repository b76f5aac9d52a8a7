// Fixed-pivot-order LU structures and the GPU execution plan built from them.
//
// Both reference solvers end a numeric factorization with a fixed pivot
// order and fill pattern that later refactorizations replay:
//   * Sparse 1.3 (NeoSolver): SparseMatrix::factor, src/solver/factor.cpp:980-1052.
//     Left-looking over internal columns; U(r,j) = x_r * (1/pivot_r); L kept
//     unnormalised; the diagonal stores 1/pivot.                 (FORM_S13)
//   * AMD-LU (AmdLuSolver::refactor_replay_, src/core/amd_lu_solver.cpp:214-271).
//     Left-looking in the recorded reach order; unit L = x/ukk; U stores the
//     pivot explicitly; pivot-growth check.                      (FORM_GP)
//
// LuStructure expresses either one in "pivot space" (k = 0..n-1): column k is
// one contiguous segment [cptr[k], cptr[k+1]) of value positions holding
// its U entries in the reference's processing order, then the diagonal,
// then its L entries.  Values of a batch are stored position-major with the
// instance index fastest (value(pos, inst) at pos * B + inst), so lanes that
// work on consecutive instances touch consecutive addresses.
#pragma once
#include <cstdint>
#include <vector>

namespace nsa {

enum LuForm : int32_t { FORM_GP = 0, FORM_S13 = 1 };

struct LuStructure {
    int32_t n = 0;
    int32_t nnz_a = 0;
    LuForm form = FORM_GP;
    std::vector<int32_t> cptr;      // n+1: segment of column k
    std::vector<int32_t> nU;        // n: U entries before the diagonal
    std::vector<int32_t> rows;      // nlu: pivot-space row of each position
    std::vector<int32_t> a_pos;     // nnz_a: position receiving A slot p
    std::vector<int32_t> gmin_pos;  // n: position receiving diag_gmin in column k, or -1
    std::vector<int32_t> row_perm;  // n: b'[k] = b[row_perm[k]]
    std::vector<int32_t> col_perm;  // n: x[col_perm[k]] = y[k]
    // FORM_GP only: the pivot candidates of column k (rows not pivotal
    // before step k that are in its reach) in the reach (xi) order, used to
    // re-check that a fresh GP factorization would choose the same pivot
    // (AmdLuSolver::factor_, amd_lu_solver.cpp:145-160).  Stored as
    // positions into the segment of column k.
    std::vector<int32_t> cand_ptr;  // n+1
    std::vector<int32_t> cand_pos;
    std::vector<int32_t> diag_cand; // n: position of original row q[k] among candidates, or -1

    int32_t nlu() const { return cptr.empty() ? 0 : cptr[n]; }
    int32_t dpos(int k) const { return cptr[k] + nU[k]; }

    bool same_as(const LuStructure& o) const;
};

// Builds the structure of an AMD-LU (Gilbert-Peierls) factorization.
// Inputs follow AmdLuSolver's members: q (column order), pinv, Lp/Li (rows
// already in pivot-step space), Up/Ui; cand lists come from the factor run.
LuStructure structure_from_gp(int n, const int32_t* a_col_ptr, const int32_t* a_row_idx,
                              const std::vector<int32_t>& q, const std::vector<int32_t>& pinv,
                              const std::vector<int32_t>& Lp, const std::vector<int32_t>& Li,
                              const std::vector<int32_t>& Up, const std::vector<int32_t>& Ui,
                              const std::vector<int32_t>& cand_ptr,
                              const std::vector<int32_t>& cand_rows);

// Builds the structure of a Sparse-1.3 factorization from its internal
// (1-based) translation arrays and the internal column lists (rows in list
// order, which is ascending internal row).
LuStructure structure_from_s13(int n, const int32_t* a_col_ptr, const int32_t* a_row_idx,
                               const std::vector<int32_t>& i2e_row,
                               const std::vector<int32_t>& i2e_col,
                               const std::vector<int32_t>& col_ptr1,
                               const std::vector<int32_t>& rows1);

// A level-scheduled sparse triangular program in "row-gather" form.  Every
// reference solve loop in scatter form (x[t] -= v * L(t,c) for c in order)
// becomes, for target t, a sequential chain over the same c in the same
// order, so each target's subtraction sequence is the reference's own.
enum RowOp : int32_t {
    OP_PLAIN = 0,        // acc = y[t] - sum(LU[pos] * y[src]); y[t] = acc
    OP_SKIPZERO = 1,     // as PLAIN, but a term is skipped when y[src] == 0
    OP_MUL_DIAG = 2,     // ... then y[t] = acc * LU[dpos]
    OP_DIV_DIAG = 4,     // ... then y[t] = acc / LU[dpos]
    OP_MUL_NZ = 8,       // ... then y[t] = (acc != 0) ? acc * LU[dpos] : acc
};

struct TriProgram {
    int32_t op = OP_PLAIN;
    std::vector<int32_t> level_ptr;  // nlevels+1 into order
    std::vector<int32_t> order;      // targets (pivot-space) in level order
    std::vector<int32_t> term_ptr;   // n+1, indexed by target
    std::vector<int32_t> term_pos;   // LU positions
    std::vector<int32_t> term_src;   // source index (pivot space)
    std::vector<int32_t> dpos;       // n, diagonal position per target (or -1)
    // Column sweep over the densest end of the program (0 = none): the last
    // sweep_n targets of an ascending program or the first sweep_n of a
    // descending one, in processing order (sweep_tgt).  sweep_pos[s * n_s + r]
    // is the LU position of the term that step s (finalizing sweep_tgt[s])
    // applies to sweep_tgt[r], or -1.  Ascending: those targets keep only
    // their terms from before the segment in the level program, without the
    // diagonal operation (partial[t] = 1); descending: they leave it.
    int32_t sweep_n = 0;
    bool sweep_first = false;
    std::vector<int32_t> sweep_tgt, sweep_pos;
    std::vector<uint8_t> partial;
    int32_t nlevels() const { return (int32_t)level_ptr.size() - 1; }
};

// Refactor schedule: columns grouped by dependency level.
struct RefactorSchedule {
    std::vector<int32_t> level_ptr;
    std::vector<int32_t> cols;
    std::vector<int32_t> level_maxlen;  // longest segment in each level
    std::vector<int32_t> a_src;         // nlu: A slot feeding each position, or -1 (fill)
    int32_t nlevels() const { return (int32_t)level_ptr.size() - 1; }
};

RefactorSchedule build_refactor_schedule(const LuStructure& s);
// sweep_max > 0 plans a column sweep of up to sweep_max targets for the
// plain-solve programs whose per-target term order is the processing order
// (each target's chain then stays the reference's when it is applied step
// by step); the transposed programs never sweep.
void build_solve_programs(const LuStructure& s, TriProgram& fwd, TriProgram& bwd,
                          TriProgram& tfwd, TriProgram& tbwd, int sweep_max = 0);

}  // namespace nsa
