#include "lu_structure.hpp"

#include <algorithm>
#include <stdexcept>
#include <unordered_map>

#include "common.hpp"

namespace nsa {

bool LuStructure::same_as(const LuStructure& o) const {
    return n == o.n && form == o.form && cptr == o.cptr && nU == o.nU && rows == o.rows &&
           a_pos == o.a_pos && row_perm == o.row_perm && col_perm == o.col_perm;
}

namespace {

// Position of pivot-space row `row` inside the segment of column k.
int32_t find_in_col(const LuStructure& s, int k, int32_t row) {
    for (int32_t p = s.cptr[k]; p < s.cptr[k + 1]; ++p)
        if (s.rows[p] == row) return p;
    return -1;
}

}  // namespace

LuStructure structure_from_gp(int n, const int32_t* acp, const int32_t* ari,
                              const std::vector<int32_t>& q, const std::vector<int32_t>& pinv,
                              const std::vector<int32_t>& Lp, const std::vector<int32_t>& Li,
                              const std::vector<int32_t>& Up, const std::vector<int32_t>& Ui,
                              const std::vector<int32_t>& cand_ptr,
                              const std::vector<int32_t>& cand_rows) {
    LuStructure s;
    s.n = n;
    s.nnz_a = acp[n];
    s.form = FORM_GP;
    s.cptr.assign(n + 1, 0);
    s.nU.assign(n, 0);
    for (int k = 0; k < n; ++k) {
        const int nu = Up[k + 1] - Up[k] - 1;          // U diag is last (amd_lu_solver.cpp:158-160)
        const int nl = Lp[k + 1] - Lp[k] - 1;          // unit L diag is first (:163)
        s.nU[k] = nu;
        s.cptr[k + 1] = s.cptr[k] + nu + 1 + nl;
    }
    s.rows.resize(s.cptr[n]);
    for (int k = 0; k < n; ++k) {
        int32_t p = s.cptr[k];
        for (int up = Up[k]; up < Up[k + 1] - 1; ++up) s.rows[p++] = Ui[up];
        s.rows[p++] = k;
        for (int lp = Lp[k] + 1; lp < Lp[k + 1]; ++lp) s.rows[p++] = Li[lp];
    }
    // A(:, q[k]) scatters to pivot rows pinv[row] of column k
    // (scatter_row_, amd_lu_solver.cpp:190-193).
    std::vector<int32_t> qinv(n);
    for (int k = 0; k < n; ++k) qinv[q[k]] = k;
    s.a_pos.assign(s.nnz_a, -1);
    for (int c = 0; c < n; ++c) {
        const int k = qinv[c];
        for (int p = acp[c]; p < acp[c + 1]; ++p) {
            int32_t pos = find_in_col(s, k, pinv[ari[p]]);
            if (pos < 0) throw Error(NSA_ERR_RUNTIME, "GP structure: A entry outside column pattern");
            s.a_pos[p] = pos;
        }
    }
    // diag_gmin lands on x[col] = pivot row pinv[col] (amd_lu_solver.cpp:121, 227-229).
    s.gmin_pos.assign(n, -1);
    for (int k = 0; k < n; ++k) s.gmin_pos[k] = find_in_col(s, k, pinv[q[k]]);
    s.row_perm.assign(n, 0);
    for (int i = 0; i < n; ++i) s.row_perm[pinv[i]] = i;
    s.col_perm = q;
    // pivot candidates in reach order
    s.cand_ptr = cand_ptr;
    s.cand_pos.resize(cand_rows.size());
    s.diag_cand.assign(n, -1);
    for (int k = 0; k < n; ++k) {
        for (int c = cand_ptr[k]; c < cand_ptr[k + 1]; ++c) {
            const int32_t orig = cand_rows[c];
            s.cand_pos[c] = find_in_col(s, k, pinv[orig]);
            if (orig == q[k]) s.diag_cand[k] = c - cand_ptr[k];
        }
    }
    return s;
}

LuStructure structure_from_s13(int n, const int32_t* acp, const int32_t* ari,
                               const std::vector<int32_t>& i2e_row,
                               const std::vector<int32_t>& i2e_col,
                               const std::vector<int32_t>& col_ptr1,
                               const std::vector<int32_t>& rows1) {
    LuStructure s;
    s.n = n;
    s.nnz_a = acp[n];
    s.form = FORM_S13;
    s.cptr.assign(n + 1, 0);
    s.nU.assign(n, 0);
    s.rows.resize(col_ptr1[n]);
    for (int k = 0; k < n; ++k) {
        int nu = 0;
        bool has_diag = false;
        for (int p = col_ptr1[k]; p < col_ptr1[k + 1]; ++p) {
            const int r = rows1[p] - 1;
            s.rows[p] = r;
            if (r < k) ++nu;
            if (r == k) has_diag = true;
        }
        if (!has_diag) throw Error(NSA_ERR_RUNTIME, "S13 structure: missing diagonal");
        s.nU[k] = nu;
        s.cptr[k + 1] = col_ptr1[k + 1];
    }
    std::vector<int32_t> e2i_row(n + 1), e2i_col(n + 1);
    for (int i = 1; i <= n; ++i) {
        e2i_row[i2e_row[i - 1]] = i;
        e2i_col[i2e_col[i - 1]] = i;
    }
    s.a_pos.assign(s.nnz_a, -1);
    for (int c = 0; c < n; ++c)
        for (int p = acp[c]; p < acp[c + 1]; ++p) {
            const int k = e2i_col[c + 1] - 1;
            const int r = e2i_row[ari[p] + 1] - 1;
            int32_t pos = find_in_col(s, k, r);
            if (pos < 0) throw Error(NSA_ERR_RUNTIME, "S13 structure: A entry outside column pattern");
            s.a_pos[p] = pos;
        }
    // add_diag_gmin adds to every internal diagonal (utils.cpp:91-97).
    s.gmin_pos.assign(n, -1);
    for (int k = 0; k < n; ++k) s.gmin_pos[k] = s.dpos(k);
    s.row_perm.assign(n, 0);
    s.col_perm.assign(n, 0);
    for (int k = 0; k < n; ++k) {
        s.row_perm[k] = i2e_row[k] - 1;
        s.col_perm[k] = i2e_col[k] - 1;
    }
    return s;
}

RefactorSchedule build_refactor_schedule(const LuStructure& s) {
    RefactorSchedule rs;
    const int n = s.n;
    std::vector<int32_t> lev(n, 0);
    int32_t maxlev = -1;
    for (int k = 0; k < n; ++k) {
        int32_t l = 0;
        for (int p = s.cptr[k]; p < s.cptr[k] + s.nU[k]; ++p)
            l = std::max(l, lev[s.rows[p]] + 1);
        lev[k] = l;
        maxlev = std::max(maxlev, l);
    }
    const int nl = maxlev + 1;
    rs.level_ptr.assign(nl + 1, 0);
    for (int k = 0; k < n; ++k) rs.level_ptr[lev[k] + 1]++;
    for (int l = 0; l < nl; ++l) rs.level_ptr[l + 1] += rs.level_ptr[l];
    rs.cols.resize(n);
    std::vector<int32_t> fill(rs.level_ptr.begin(), rs.level_ptr.end() - 1);
    rs.level_maxlen.assign(nl, 0);
    for (int k = 0; k < n; ++k) {
        rs.cols[fill[lev[k]]++] = k;
        rs.level_maxlen[lev[k]] = std::max(rs.level_maxlen[lev[k]], s.cptr[k + 1] - s.cptr[k]);
    }
    rs.a_src.assign(s.nlu(), -1);
    for (int p = 0; p < s.nnz_a; ++p) rs.a_src[s.a_pos[p]] = p;
    // The refactor kernel resolves update targets through a row -> slot map
    // of the column, so every row of L(:, r) must be in column k for each U
    // entry r of k (fill closure).  Checked here once per structure.
    std::vector<int32_t> mark(n, -1);
    for (int k = 0; k < n; ++k) {
        for (int p = s.cptr[k]; p < s.cptr[k + 1]; ++p) {
            if (s.rows[p] < 0 || s.rows[p] >= n)
                throw Error(NSA_ERR_RUNTIME, "LU structure: row index out of range");
            mark[s.rows[p]] = k;
        }
        for (int p = s.cptr[k]; p < s.cptr[k] + s.nU[k]; ++p) {
            const int r = s.rows[p];
            if (r >= k) throw Error(NSA_ERR_RUNTIME, "LU structure: U entry not above the diagonal");
            for (int e = s.dpos(r) + 1; e < s.cptr[r + 1]; ++e)
                if (mark[s.rows[e]] != k)
                    throw Error(NSA_ERR_RUNTIME, "LU structure: fill pattern not closed");
        }
    }
    return rs;
}

namespace {

struct Term { int32_t key, pos, src; };

// Levels of a row-gather program: target t waits for every source.
// Targets with skip[t] are not in the program; as sources they are ready
// before it starts (a descending program's sweep runs first).
void level_program(TriProgram& pr, int n, bool ascending, const std::vector<uint8_t>* skip) {
    std::vector<int32_t> lev(n, -1);
    int32_t maxlev = -1;
    for (int it = 0; it < n; ++it) {
        const int t = ascending ? it : n - 1 - it;
        if (skip && (*skip)[t]) continue;
        int32_t l = 0;
        for (int p = pr.term_ptr[t]; p < pr.term_ptr[t + 1]; ++p)
            l = std::max(l, lev[pr.term_src[p]] + 1);
        lev[t] = l;
        maxlev = std::max(maxlev, l);
    }
    const int nl = maxlev + 1;
    pr.level_ptr.assign(nl + 1, 0);
    for (int t = 0; t < n; ++t)
        if (lev[t] >= 0) pr.level_ptr[lev[t] + 1]++;
    for (int l = 0; l < nl; ++l) pr.level_ptr[l + 1] += pr.level_ptr[l];
    pr.order.resize(pr.level_ptr[nl]);
    std::vector<int32_t> fill(pr.level_ptr.begin(), pr.level_ptr.end() - 1);
    for (int t = 0; t < n; ++t)
        if (lev[t] >= 0) pr.order[fill[lev[t]]++] = t;
}

// Plans the column sweep over the last (ascending) or first (descending) D
// targets.  Requires every target's terms to be in processing order, which
// the callers guarantee by their sort; builds the step x target position
// table and cuts the swept terms out of the target lists.
void plan_sweep(TriProgram& pr, std::vector<std::vector<Term>>& terms, int n, bool ascending,
                int sweep_max, std::vector<uint8_t>& skip) {
    const int D = std::min(sweep_max, n);
    skip.assign(n, 0);
    if (D <= 0) return;
    pr.sweep_n = D;
    pr.sweep_first = !ascending;
    pr.sweep_tgt.resize(D);
    std::vector<int32_t> rel(n, -1);
    for (int r = 0; r < D; ++r) {
        const int t = ascending ? n - D + r : n - 1 - r;
        pr.sweep_tgt[r] = t;
        rel[t] = r;
    }
    pr.sweep_pos.assign((size_t)D * D, -1);
    pr.partial.assign(n, 0);
    for (int r = 0; r < D; ++r) {
        const int t = pr.sweep_tgt[r];
        std::vector<Term> keep;
        for (const Term& tm : terms[t]) {
            const int sidx = rel[tm.src];
            if (sidx < 0) {
                keep.push_back(tm);
            } else {
                if (sidx >= r) throw Error(NSA_ERR_LOGIC, "plan_sweep: term from a later step");
                pr.sweep_pos[(size_t)sidx * D + r] = tm.pos;
            }
        }
        // in processing order every pre-segment term precedes the swept ones
        terms[t] = keep;
        if (ascending && !keep.empty()) pr.partial[t] = 1;
        else skip[t] = 1;
    }
}

void pack_terms(TriProgram& pr, int n, std::vector<std::vector<Term>>& terms) {
    pr.term_ptr.assign(n + 1, 0);
    for (int t = 0; t < n; ++t) pr.term_ptr[t + 1] = pr.term_ptr[t] + (int)terms[t].size();
    pr.term_pos.resize(pr.term_ptr[n]);
    pr.term_src.resize(pr.term_ptr[n]);
    for (int t = 0; t < n; ++t)
        for (size_t i = 0; i < terms[t].size(); ++i) {
            pr.term_pos[pr.term_ptr[t] + i] = terms[t][i].pos;
            pr.term_src[pr.term_ptr[t] + i] = terms[t][i].src;
        }
}

}  // namespace

void build_solve_programs(const LuStructure& s, TriProgram& fwd, TriProgram& bwd,
                          TriProgram& tfwd, TriProgram& tbwd, int sweep_max) {
    const int n = s.n;
    // Entries by kind, each (row, col, pos) in pivot space.
    std::vector<std::vector<Term>> lrow(n), urow(n), lcol(n), ucol(n);
    std::vector<int32_t> dp(n);
    for (int k = 0; k < n; ++k) {
        dp[k] = s.dpos(k);
        for (int p = s.cptr[k]; p < s.cptr[k + 1]; ++p) {
            const int r = s.rows[p];
            if (r < k) {                     // U(r, k)
                urow[r].push_back({k, p, k});
                ucol[k].push_back({r, p, r});
            } else if (r > k) {              // L(r, k)
                lrow[r].push_back({k, p, k});
                lcol[k].push_back({r, p, r});
            }
        }
    }
    auto sort_key = [](std::vector<Term>& v, bool asc) {
        std::stable_sort(v.begin(), v.end(), [asc](const Term& a, const Term& b) {
            return asc ? a.key < b.key : a.key > b.key;
        });
    };
    std::vector<std::vector<Term>> tf(n), tb(n), ttf(n), ttb(n);
    if (s.form == FORM_S13) {
        // solve.cpp:21-33: forward column scatter in ascending column i,
        // skipped when inter[i] == 0, pivot row scaled by the stored 1/pivot.
        for (int t = 0; t < n; ++t) { tf[t] = lrow[t]; sort_key(tf[t], true); }
        fwd.op = OP_SKIPZERO | OP_MUL_NZ;
        // solve.cpp:35-43: backward row dot over U(i, c), c ascending.
        for (int t = 0; t < n; ++t) { tb[t] = urow[t]; sort_key(tb[t], true); }
        bwd.op = OP_PLAIN;
        // solve.cpp:117-126: U^T forward, row scatter from rows i ascending,
        // skipped when inter[i] == 0.  Target c gathers U(i, c), i ascending.
        for (int t = 0; t < n; ++t) { ttf[t] = ucol[t]; sort_key(ttf[t], true); }
        tfwd.op = OP_SKIPZERO;
        // solve.cpp:128-137: L^T backward column dot (list order = ascending
        // row), then scaled by the stored 1/pivot.
        for (int t = 0; t < n; ++t) { ttb[t] = lcol[t]; sort_key(ttb[t], true); }
        tbwd.op = OP_MUL_DIAG;
    } else {
        // amd_lu_solver.cpp:320-327: unit-L forward scatter, j ascending.
        for (int t = 0; t < n; ++t) { tf[t] = lrow[t]; sort_key(tf[t], true); }
        fwd.op = OP_PLAIN;
        // amd_lu_solver.cpp:329-336: backward, j descending: x[j] /= U(j,j),
        // then scatter U(:, j).  Target t gathers U(t, c), c descending, and
        // divides last.
        for (int t = 0; t < n; ++t) { tb[t] = urow[t]; sort_key(tb[t], false); }
        bwd.op = OP_DIV_DIAG;
        // Transposed: U^T forward (column k of U is row k of U^T) in U's
        // column storage order, divided by U(k,k); then unit L^T backward.
        for (int t = 0; t < n; ++t) ttf[t] = ucol[t];
        tfwd.op = OP_DIV_DIAG;
        for (int t = 0; t < n; ++t) ttb[t] = lcol[t];
        tbwd.op = OP_PLAIN;
    }
    // Sweeps: the forward programs of both forms gather ascending sources in
    // an ascending program; the AMD-LU backward gathers descending sources in
    // a descending one.  Sparse 1.3's backward row dot runs ascending inside
    // a descending program, so it keeps its levels.
    std::vector<uint8_t> fskip, bskip;
    plan_sweep(fwd, tf, n, true, sweep_max, fskip);
    if (s.form != FORM_S13) plan_sweep(bwd, tb, n, false, sweep_max, bskip);
    pack_terms(fwd, n, tf);
    pack_terms(bwd, n, tb);
    pack_terms(tfwd, n, ttf);
    pack_terms(tbwd, n, ttb);
    fwd.dpos = dp; bwd.dpos = dp; tfwd.dpos = dp; tbwd.dpos = dp;
    level_program(fwd, n, true, fskip.empty() ? nullptr : &fskip);
    level_program(bwd, n, false, bskip.empty() ? nullptr : &bskip);
    level_program(tfwd, n, true, nullptr);
    level_program(tbwd, n, false, nullptr);
}

}  // namespace nsa
