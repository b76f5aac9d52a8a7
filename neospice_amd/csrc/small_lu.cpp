#include "small_lu.hpp"

#include <algorithm>
#include <map>

#include "common.hpp"

namespace nsa {

SmallLuPlans::~SmallLuPlans() {
    if (d_words_) (void)hipFree(d_words_);
    if (d_off_) (void)hipFree(d_off_);
}

namespace {

// Right-looking step tasks of a Sparse 1.3 structure, grouped in levels:
// a step goes one level after the last step that wrote any entry it reads
// or writes.  Reads of a step are final entries of row and column k, which
// no later step writes, so read-after-write is the only hazard besides the
// write order of a target (same rule).
void build_tasks(const LuStructure& s, std::vector<int32_t>& lev_ptr, std::vector<int32_t>& tasks) {
    const int n = s.n;
    std::vector<std::vector<int32_t>> lpos(n), upos(n);  // L(:,k) / U(k,:) positions
    std::map<std::pair<int, int>, int32_t> pos_of;        // (row, col) -> position
    for (int k = 0; k < n; ++k)
        for (int p = s.cptr[k]; p < s.cptr[k + 1]; ++p) {
            const int r = s.rows[p];
            pos_of[{r, k}] = p;
            if (r > k) lpos[k].push_back(p);
            else if (r < k) upos[r].push_back(p);
        }
    std::vector<int> wlev(s.nlu(), -1);
    std::vector<std::vector<int32_t>> by_level;
    for (int k = 0; k < n; ++k) {
        if (lpos[k].empty() || upos[k].empty()) continue;
        const int32_t piv = s.dpos(k);
        int lev = std::max(wlev[piv], -1);
        for (int32_t p : lpos[k]) lev = std::max(lev, wlev[p]);
        for (int32_t p : upos[k]) lev = std::max(lev, wlev[p]);
        std::vector<int32_t> step;
        for (int32_t pu : upos[k]) {
            // column j of U(k, j): the column whose segment holds position pu
            const int j = (int)(std::upper_bound(s.cptr.begin(), s.cptr.end(), pu) - s.cptr.begin()) - 1;
            for (int32_t pl : lpos[k]) {
                const int i = s.rows[pl];
                auto it = pos_of.find({i, j});
                if (it == pos_of.end())
                    throw Error(NSA_ERR_RUNTIME, "small LU plan: update outside the fill pattern");
                lev = std::max(lev, wlev[it->second]);
                step.insert(step.end(), {it->second, pu, pl, piv});
            }
        }
        ++lev;
        for (size_t t = 0; t < step.size(); t += 4) wlev[step[t]] = lev;
        if ((int)by_level.size() <= lev) by_level.resize(lev + 1);
        by_level[lev].insert(by_level[lev].end(), step.begin(), step.end());
    }
    lev_ptr.assign(1, 0);
    tasks.clear();
    for (auto& l : by_level) {
        tasks.insert(tasks.end(), l.begin(), l.end());
        lev_ptr.push_back((int32_t)(tasks.size() / 4));
    }
}

void align4(std::vector<int32_t>& w) {
    while (w.size() % 4) w.push_back(0);
}

// A triangular program as level-ordered row records and terms.
void append_program(const TriProgram& t, std::vector<int32_t>& w, int32_t base, int32_t* hdr3) {
    const int nl = t.nlevels();
    hdr3[0] = nl;
    align4(w);
    hdr3[1] = (int32_t)w.size() - base;
    for (int l = 0; l <= nl; ++l) w.push_back(t.level_ptr[l]);
    // rows (4 words each), then terms (2 words each)
    align4(w);
    const int nrows = (int)t.order.size();
    const int32_t rows_at = (int32_t)w.size();
    hdr3[2] = rows_at - base;
    w.resize(w.size() + 4 * (size_t)nrows);
    int32_t term_base = (int32_t)w.size() - base;
    std::vector<int32_t> terms;
    for (int r = 0; r < nrows; ++r) {
        const int tg = t.order[r];
        const int a = t.term_ptr[tg], b = t.term_ptr[tg + 1];
        int32_t* rec = &w[rows_at + 4 * (size_t)r];
        rec[0] = tg;
        rec[1] = term_base + (int32_t)terms.size();  // relative word offset of the first term
        rec[2] = b - a;
        rec[3] = t.dpos.empty() ? -1 : t.dpos[tg];
        for (int q = a; q < b; ++q) {
            terms.push_back(t.term_pos[q]);
            terms.push_back(t.term_src[q]);
        }
    }
    w.insert(w.end(), terms.begin(), terms.end());
}

}  // namespace

int SmallLuPlans::add(const LuStructure& s) {
    if (!fits(s)) throw Error(NSA_ERR_UNSUPPORTED, "small LU plan: structure does not fit the packed kernel");
    for (size_t p = 0; p < st_.size(); ++p)
        if (st_[p].same_as(s) && st_[p].nnz_a == s.nnz_a && st_[p].gmin_pos == s.gmin_pos)
            return (int)p;
    const int n = s.n;
    align4(words_);
    const int32_t base = (int32_t)words_.size();
    off_.push_back(base);
    words_.resize(words_.size() + SH_WORDS, 0);
    auto H = [&](int f) -> int32_t& { return words_[base + f]; };
    auto put = [&](int f, const std::vector<int32_t>& v) {
        align4(words_);
        H(f) = (int32_t)words_.size() - base;
        words_.insert(words_.end(), v.begin(), v.end());
    };
    H(SH_N) = n;
    H(SH_NLU) = s.nlu();
    H(SH_NNZA) = s.nnz_a;
    std::vector<int32_t> lev_ptr, tasks;
    build_tasks(s, lev_ptr, tasks);
    H(SH_NLEV) = (int32_t)lev_ptr.size() - 1;
    H(SH_NTASK) = (int32_t)(tasks.size() / 4);
    put(SH_APOS, s.a_pos);
    put(SH_GMIN, s.gmin_pos);
    put(SH_LEV, lev_ptr);
    put(SH_TASK, tasks);
    put(SH_CPTR, s.cptr);
    put(SH_NU, s.nU);
    put(SH_ROWS, s.rows);
    std::vector<int32_t> pcol(s.nlu());
    for (int k = 0; k < n; ++k)
        for (int p = s.cptr[k]; p < s.cptr[k + 1]; ++p) pcol[p] = k;
    put(SH_PCOL, pcol);
    put(SH_RPERM, s.row_perm);
    put(SH_CPERM, s.col_perm);
    TriProgram fwd, bwd, tfwd, tbwd;
    build_solve_programs(s, fwd, bwd, tfwd, tbwd, 0);
    int32_t h3[3];
    const TriProgram* progs[4] = {&fwd, &bwd, &tfwd, &tbwd};
    for (int p = 0; p < 4; ++p) {
        append_program(*progs[p], words_, base, h3);
        H(SH_PROG + 3 * p) = h3[0];
        H(SH_PROG + 3 * p + 1) = h3[1];
        H(SH_PROG + 3 * p + 2) = h3[2];
    }
    st_.push_back(s);
    lev_.push_back(H(SH_NLEV));
    slev_.push_back(fwd.nlevels() + bwd.nlevels());
    max_nlu_ = std::max(max_nlu_, s.nlu());
    max_n_ = std::max(max_n_, n);
    return (int)st_.size() - 1;
}

void SmallLuPlans::sync(hipStream_t s) {
    if (synced_plans_ == st_.size()) return;
    const bool grow = words_.size() > d_words_cap_ || off_.size() > d_off_cap_;
    if (grow) {
        // launches in flight may still read the old copy
        NSA_HIP(hipStreamSynchronize(s));
        if (d_words_) NSA_HIP(hipFree(d_words_));
        if (d_off_) NSA_HIP(hipFree(d_off_));
        d_words_cap_ = std::max<size_t>(words_.size() * 2, 4096);
        d_off_cap_ = std::max<size_t>(off_.size() * 2, 64);
        NSA_HIP(hipMalloc(&d_words_, d_words_cap_ * sizeof(int32_t)));
        NSA_HIP(hipMalloc(&d_off_, d_off_cap_ * sizeof(int32_t)));
        synced_words_ = 0;
    }
    // Appended plans never move earlier ones, so only the tail is copied;
    // words a running launch reads are never rewritten.
    const size_t w0 = synced_words_;
    NSA_HIP(hipMemcpyAsync(d_words_ + w0, words_.data() + w0, (words_.size() - w0) * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
    NSA_HIP(hipMemcpyAsync(d_off_, off_.data(), off_.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
    // the host vectors are not modified before the copies complete: callers
    // add plans only after synchronizing the stream
    NSA_HIP(hipStreamSynchronize(s));
    synced_words_ = words_.size();
    synced_plans_ = st_.size();
}

}  // namespace nsa
