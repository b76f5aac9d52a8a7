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
// a task goes one level after the last task that wrote any entry it reads
// or writes.  Reads of a step are final entries of row and column k, which
// no later step writes, so read-after-write is the only hazard besides the
// write order of a target (same rule).
//
// With the forward substitution folded in, the right-hand side is one more
// column: vector entry i sits at position nlu + i, and step k adds, for
// every L(i,k), the task y_i -= (y_k * (1 / x_kk)) * L(i,k) -- solve.cpp:21-33
// applies exactly these terms, in ascending k per target, and skips step k
// when y_k == 0.  Such a task carries its pivot position as -(pos + 1).
void build_tasks(const LuStructure& s, std::vector<int32_t>& lev_ptr, std::vector<int32_t>& tasks,
                 bool with_rhs) {
    const int n = s.n, nlu = s.nlu();
    std::vector<std::vector<int32_t>> lpos(n), upos(n);  // L(:,k) / U(k,:) positions
    std::map<std::pair<int, int>, int32_t> pos_of;        // (row, col) -> position
    for (int k = 0; k < n; ++k)
        for (int p = s.cptr[k]; p < s.cptr[k + 1]; ++p) {
            const int r = s.rows[p];
            pos_of[{r, k}] = p;
            if (r > k) lpos[k].push_back(p);
            else if (r < k) upos[r].push_back(p);
        }
    std::vector<int> wlev(nlu + n, -1);
    std::vector<std::vector<int32_t>> by_level;
    auto place = [&](int lev, const std::vector<int32_t>& step) {
        for (size_t t = 0; t < step.size(); t += 4) wlev[step[t]] = lev;
        if ((int)by_level.size() <= lev) by_level.resize(lev + 1);
        by_level[lev].insert(by_level[lev].end(), step.begin(), step.end());
    };
    for (int k = 0; k < n; ++k) {
        const int32_t piv = s.dpos(k);
        if (!lpos[k].empty() && !upos[k].empty()) {
            int lev = wlev[piv];
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
            place(lev + 1, step);
        }
        if (with_rhs && !lpos[k].empty()) {
            const int32_t yk = nlu + k;
            int lev = std::max(wlev[piv], wlev[yk]);
            for (int32_t p : lpos[k]) lev = std::max({lev, wlev[p], wlev[nlu + s.rows[p]]});
            std::vector<int32_t> step;
            for (int32_t pl : lpos[k]) step.insert(step.end(), {nlu + s.rows[pl], yk, pl, -(piv + 1)});
            place(lev + 1, step);
        }
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

// A triangular program as level-ordered row records and terms, in LDS byte
// addresses of the launch's layout (factor value p at 8p, vector entry i at
// 8 (nlu + i), a zero at 8 (nlu + n)).  A row record is 8 words: target,
// number of terms after the first two, diagonal value (the zero when the
// program has none), plan offset of those terms, then value and source of
// terms 0 and 1 (the zero for both when a row has fewer terms: x - 0 * 0 is
// x, so every row runs two terms without a branch).  Circuit rows are
// mostly that short, so a level needs no term fetch before its operand
// loads.
void append_program(const TriProgram& t, int nlu, int n, std::vector<int32_t>& w, int32_t base,
                    int32_t* hdr4, int min_run) {
    const int32_t zero_b = 8 * (nlu + n);
    auto vb = [&](int pos) { return (int32_t)(8 * pos); };
    auto yb = [&](int i) { return (int32_t)(8 * (nlu + i)); };
    const int nl = t.nlevels();
    hdr4[0] = nl;
    align4(w);
    hdr4[1] = (int32_t)w.size() - base;
    for (int l = 0; l <= nl; ++l) w.push_back(t.level_ptr[l]);
    align4(w);
    // phases: maximal runs of at least min_run one-row levels are
    // serial, what lies between them parallel
    std::vector<int32_t> ph;
    for (int l = 0; l < nl;) {
        int e = l;
        while (e < nl && t.level_ptr[e + 1] - t.level_ptr[e] == 1) ++e;
        if (e - l >= min_run) {
            ph.insert(ph.end(), {PH_SERIAL, l, e, 0});
            l = e;
            continue;
        }
        // a short run or a wider level: part of a parallel phase
        const int q = std::max(e, l + 1);
        if (!ph.empty() && ph[ph.size() - 4] == PH_PARALLEL) ph[ph.size() - 2] = q;
        else ph.insert(ph.end(), {PH_PARALLEL, l, q, 0});
        l = q;
    }
    hdr4[3] = (int32_t)w.size() - base;
    w.insert(w.end(), {(int32_t)(ph.size() / 4), 0, 0, 0});
    w.insert(w.end(), ph.begin(), ph.end());
    const int nrows = (int)t.order.size();
    const int32_t rows_at = (int32_t)w.size();
    hdr4[2] = rows_at - base;
    w.resize(w.size() + 8 * (size_t)nrows);
    const int32_t term_base = (int32_t)w.size() - base;
    std::vector<int32_t> terms;
    for (int r = 0; r < nrows; ++r) {
        const int tg = t.order[r];
        const int a = t.term_ptr[tg], b = t.term_ptr[tg + 1];
        int32_t* rec = &w[rows_at + 8 * (size_t)r];
        rec[0] = yb(tg);
        rec[1] = std::max(0, b - a - 2);
        rec[2] = (t.dpos.empty() || t.dpos[tg] < 0) ? zero_b : vb(t.dpos[tg]);
        rec[3] = term_base + (int32_t)terms.size();
        rec[4] = rec[5] = rec[6] = rec[7] = zero_b;
        for (int q = a; q < b; ++q) {
            if (q - a < 2) {
                rec[4 + 2 * (q - a)] = vb(t.term_pos[q]);
                rec[5 + 2 * (q - a)] = yb(t.term_src[q]);
            } else {
                terms.push_back(vb(t.term_pos[q]));
                terms.push_back(yb(t.term_src[q]));
            }
        }
    }
    w.insert(w.end(), terms.begin(), terms.end());
    // serial phases: for terms 0 and 1 of each row, 1 + the row of the same
    // phase that produces the term's source (0: the source is final before
    // the phase), packed low/high 16 bits; the phases header's second word
    // points here
    std::vector<int32_t> producer(n, -1);  // target -> row
    for (int r = 0; r < nrows; ++r) producer[t.order[r]] = r;
    std::vector<int32_t> srow(nrows, 0);
    for (size_t q = 0; q < ph.size(); q += 4) {
        if (ph[q] != PH_SERIAL) continue;
        const int rb = t.level_ptr[ph[q + 1]], re = t.level_ptr[ph[q + 2]];
        for (int r = rb; r < re; ++r) {
            const int tg = t.order[r];
            const int a = t.term_ptr[tg], b = t.term_ptr[tg + 1];
            int32_t packed = 0;
            for (int k = 0; k < 2 && a + k < b; ++k) {
                const int pr = producer[t.term_src[a + k]];
                if (pr >= rb && pr < r) packed |= (pr + 1) << (16 * k);
            }
            srow[r] = packed;
        }
    }
    align4(w);
    w[base + hdr4[3] + 1] = (int32_t)w.size() - base;
    w.insert(w.end(), srow.begin(), srow.end());
}

}  // namespace

int SmallLuPlans::add(const LuStructure& s) {
    if (!fits(s)) throw Error(NSA_ERR_UNSUPPORTED, "small LU plan: structure does not fit the packed kernel");
    for (size_t p = 0; p < st_.size(); ++p)
        if (st_[p].same_as(s) && st_[p].nnz_a == s.nnz_a && st_[p].gmin_pos == s.gmin_pos)
            return (int)p;
    return add_new(s);
}

int SmallLuPlans::add_new(const LuStructure& s) {
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
    put(SH_APOS, s.a_pos);
    put(SH_GMIN, s.gmin_pos);
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
    int32_t h4[4];
    // the forward program (solve-only launches) and the transposed ones
    const TriProgram* cold[3] = {&fwd, &tfwd, &tbwd};
    const int cold_id[3] = {0, 2, 3};
    for (int c = 0; c < 3; ++c) {
        append_program(*cold[c], s.nlu(), n, words_, base, h4, opts_.min_serial_run);
        for (int q = 0; q < 4; ++q) H(SH_PROG + 4 * cold_id[c] + q) = h4[q];
    }
    // the hot block: what the level loops of a factor-and-solve launch read
    // (the task levels with the forward substitution folded in, and the
    // backward program), contiguous so that a launch can stage it in LDS
    std::vector<int32_t> lev_ptr, tasks;
    build_tasks(s, lev_ptr, tasks, true);
    const int nlev = (int)lev_ptr.size() - 1;
    H(SH_NLEV) = nlev;
    H(SH_NTASK) = (int32_t)(tasks.size() / 4);
    // three copies of the end, so that a level reads the word three levels
    // ahead without a bounds test
    lev_ptr.insert(lev_ptr.end(), 3, lev_ptr.back());
    align4(words_);
    const int32_t hot0 = (int32_t)words_.size() - base;
    put(SH_LEV, lev_ptr);
    // tasks in LDS byte addresses (a forward task's pivot as -(address + 1))
    for (size_t q = 0; q < tasks.size(); q += 4) {
        for (int f = 0; f < 3; ++f) tasks[q + f] *= 8;
        const int32_t pv = tasks[q + 3];
        tasks[q + 3] = pv >= 0 ? 8 * pv : -(8 * (-pv - 1) + 1);
    }
    put(SH_TASK, tasks);
    append_program(bwd, s.nlu(), n, words_, base, h4, opts_.min_serial_run);
    for (int q = 0; q < 4; ++q) H(SH_PROG + 4 + q) = h4[q];
    align4(words_);
    H(SH_HOT) = hot0;
    H(SH_HOTLEN) = (int32_t)words_.size() - base - hot0;
    st_.push_back(s);
    lev_.push_back(H(SH_NLEV));
    slev_.push_back(fwd.nlevels() + bwd.nlevels());
    blev_.push_back(bwd.nlevels());
    max_nlu_ = std::max(max_nlu_, s.nlu());
    max_n_ = std::max(max_n_, n);
    max_hot_ = std::max(max_hot_, (int)words_[base + SH_HOTLEN]);
    return (int)st_.size() - 1;
}

void SmallLuPlans::rebuild(const SluOpts& o) {
    std::vector<LuStructure> st;
    st.swap(st_);
    opts_ = o;
    words_.clear();
    off_.clear();
    lev_.clear();
    slev_.clear();
    blev_.clear();
    max_nlu_ = max_n_ = max_hot_ = 0;
    for (const LuStructure& x : st) {
        // add() reuses an identical earlier plan; keep one plan per entry so
        // that plan ids stay what the instances hold
        const size_t before = st_.size();
        (void)add_new(x);
        if (st_.size() != before + 1) throw Error(NSA_ERR_RUNTIME, "SmallLuPlans::rebuild: plan ids moved");
    }
    synced_words_ = 0;
    synced_plans_ = 0;
    d_words_cap_ = 0;  // a fresh device copy: the old one may be in use
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
