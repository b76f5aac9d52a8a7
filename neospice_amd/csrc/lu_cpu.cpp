// Host-side pivot-order computations (see lu_cpu.hpp).
//
// The Markowitz search below must choose the same pivots as the reference
// on the same values, so its steps (candidate scan order, tie rules,
// thresholds, Markowitz-count bookkeeping and the order in which the linked
// row/column lists are spliced) follow src/solver/factor.cpp and
// src/solver/utils.cpp step for step.  The representation is this file's
// own: elements live in flat arrays and links are element indices (-1 =
// end of list), with a small "link slot" helper standing for the
// reference's pointer-to-pointer splicing.
#include "lu_cpu.hpp"

#include "cplx.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace nsa {

// ---------------------------------------------------------------------------
// Minimum degree (src/core/amd.cpp:14-143)
// ---------------------------------------------------------------------------
std::vector<int32_t> min_degree_order(int n, const int32_t* col_ptr, const int32_t* row_idx) {
    std::vector<int32_t> perm;
    if (n <= 0) return perm;
    std::vector<std::vector<int32_t>> adj(n);
    for (int j = 0; j < n; ++j)
        for (int p = col_ptr[j]; p < col_ptr[j + 1]; ++p) {
            const int i = row_idx[p];
            if (i != j) { adj[j].push_back(i); adj[i].push_back(j); }
        }
    for (auto& a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::vector<int32_t> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = (int32_t)adj[i].size();
    int dense_thresh = (int)(10.0 * std::sqrt((double)n));
    dense_thresh = std::min(std::max(dense_thresh, 16), n);
    std::vector<char> dense(n, 0), gone(n, 0);
    std::vector<int32_t> dense_nodes;
    for (int i = 0; i < n; ++i)
        if (deg[i] > dense_thresh) { dense[i] = 1; dense_nodes.push_back(i); }
    if (!dense_nodes.empty())
        for (int i = 0; i < n; ++i) {
            if (dense[i]) continue;
            auto& a = adj[i];
            a.erase(std::remove_if(a.begin(), a.end(), [&](int32_t v) { return dense[v] != 0; }),
                    a.end());
            deg[i] = (int32_t)a.size();
        }
    for (int32_t d : dense_nodes) gone[d] = 1;
    perm.reserve(n);
    std::vector<int32_t> live;
    const int rounds = n - (int)dense_nodes.size();
    for (int step = 0; step < rounds; ++step) {
        int best = -1, best_deg = n + 1;
        for (int i = 0; i < n; ++i)
            if (!gone[i] && deg[i] < best_deg) { best_deg = deg[i]; best = i; }
        perm.push_back(best);
        gone[best] = 1;
        live.clear();
        for (int32_t v : adj[best]) if (!gone[v]) live.push_back(v);
        for (size_t a = 0; a < live.size(); ++a)
            for (size_t b = a + 1; b < live.size(); ++b) {
                const int32_t u = live[a], v = live[b];
                auto it = std::lower_bound(adj[u].begin(), adj[u].end(), v);
                if (it != adj[u].end() && *it == v) continue;
                adj[u].insert(it, v);
                adj[v].insert(std::lower_bound(adj[v].begin(), adj[v].end(), u), u);
            }
        for (int32_t v : live) {
            adj[v].erase(std::lower_bound(adj[v].begin(), adj[v].end(), best));
            int32_t d = 0;
            for (int32_t x : adj[v]) d += !gone[x];
            deg[v] = d;
        }
    }
    // Dense nodes last by ascending degree, with the reference's call
    // (amd.cpp:134-137): std::sort is not stable, and for more than 16 dense
    // nodes the tie order is the one its introsort leaves.
    std::sort(dense_nodes.begin(), dense_nodes.end(),
              [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
    for (int32_t d : dense_nodes) perm.push_back(d);
    return perm;
}

// ---------------------------------------------------------------------------
// Gilbert-Peierls structure (AmdLuSolver::factor_, amd_lu_solver.cpp:74-197)
// ---------------------------------------------------------------------------
bool gp_factor_structure(int n, const int32_t* cp, const int32_t* ri, const double* val,
                         double gmin, const std::vector<int32_t>& q, LuStructure& out) {
    const double tol = 1e-3;
    std::vector<int32_t> Lp(n + 1, 0), Up(n + 1, 0), Li, Ui, pinv(n, -1);
    std::vector<double> Lx, Ux, x(n, 0.0);
    std::vector<int32_t> xi(n), stack(n), cand_ptr(n + 1, 0), cand_rows;
    std::vector<char> seen(n, 0);
    for (int k = 0; k < n; ++k) {
        const int col = q[k];
        Lp[k] = (int32_t)Li.size();
        Up[k] = (int32_t)Ui.size();
        int top = n;
        std::fill(seen.begin(), seen.end(), 0);
        // reach of A(:, col) in the graph of L, iterative DFS (lu_dfs, :44-72)
        for (int p = cp[col]; p < cp[col + 1]; ++p) {
            int start = ri[p];
            if (seen[start]) continue;
            int head = 0;
            xi[0] = start;
            while (head >= 0) {
                const int j = xi[head];
                const int jn = pinv[j];
                if (!seen[j]) { seen[j] = 1; stack[head] = (jn < 0) ? 0 : Lp[jn]; }
                bool done = true;
                const int pend = (jn < 0) ? 0 : Lp[jn + 1];
                for (int pp = stack[head]; pp < pend; ++pp) {
                    const int i = Li[pp];
                    if (seen[i]) continue;
                    stack[head] = pp;
                    xi[++head] = i;
                    done = false;
                    break;
                }
                if (done) { --head; xi[--top] = j; }
            }
        }
        for (int p = top; p < n; ++p) x[xi[p]] = 0.0;
        for (int p = cp[col]; p < cp[col + 1]; ++p) x[ri[p]] = val[p];
        if (gmin != 0.0) x[col] += gmin;
        for (int px = top; px < n; ++px) {
            const int j = xi[px], J = pinv[j];
            if (J < 0) continue;
            const double xj = x[j];
            for (int p = Lp[J] + 1; p < Lp[J + 1]; ++p) x[Li[p]] -= Lx[p] * xj;
        }
        int ipiv = -1;
        double amax = -1.0;
        cand_ptr[k] = (int32_t)cand_rows.size();
        for (int p = top; p < n; ++p) {
            const int i = xi[p];
            if (pinv[i] < 0) {
                cand_rows.push_back(i);
                const double t = std::fabs(x[i]);
                if (t > amax) { amax = t; ipiv = i; }
            } else {
                Ui.push_back(pinv[i]);
                Ux.push_back(x[i]);
            }
        }
        cand_ptr[k + 1] = (int32_t)cand_rows.size();
        if (ipiv == -1 || amax <= 0.0) return false;
        if (pinv[col] < 0 && std::fabs(x[col]) >= amax * tol) ipiv = col;
        const double pivot = x[ipiv];
        if (pivot == 0.0) return false;
        Ui.push_back(k);
        Ux.push_back(pivot);
        pinv[ipiv] = k;
        Li.push_back(ipiv);
        Lx.push_back(1.0);
        for (int p = top; p < n; ++p) {
            const int i = xi[p];
            if (pinv[i] < 0) { Li.push_back(i); Lx.push_back(x[i] / pivot); }
            x[i] = 0.0;
        }
    }
    Lp[n] = (int32_t)Li.size();
    Up[n] = (int32_t)Ui.size();
    for (auto& r : Li) r = pinv[r];
    out = structure_from_gp(n, cp, ri, q, pinv, Lp, Li, Up, Ui, cand_ptr, cand_rows);
    return true;
}

// ---------------------------------------------------------------------------
// Sparse 1.3 Markowitz (src/solver/{build,utils,factor}.cpp)
// ---------------------------------------------------------------------------
struct Sparse13::Impl {
    int n = 0;
    // element arrays
    std::vector<double> re, im;  // im: zero unless the matrix is complex
    bool cplx = false;
    std::vector<int32_t> erow, ecol, nrow, ncol;
    // list heads (1-based, [0] unused)
    std::vector<int32_t> fc, fr, dg;
    std::vector<int32_t> i2e_r, i2e_c, e2i_r, e2i_c;
    std::vector<int32_t> mr, mc;
    std::vector<long> mp;
    int singletons = 0;
    double rel = 1e-3, absth = 0.0;
    bool needs_ordering = true, rows_linked = false, preordered = false, vectors = false;
    bool singular = false;
    int fillins = 0;
    int current = 0;
    std::vector<int32_t> slot_elem;   // A slot -> element
    std::vector<char> is_fill;

    // ---- link slots: head of a column/row list, or an element's next field
    int32_t& col_link(int pred, int c) { return pred < 0 ? fc[c] : ncol[pred]; }
    int32_t& row_link(int pred, int r) { return pred < 0 ? fr[r] : nrow[pred]; }

    int new_element(int r, int c, bool fill) {
        re.push_back(0.0);
        im.push_back(0.0);
        erow.push_back(r);
        ecol.push_back(c);
        nrow.push_back(-1);
        ncol.push_back(-1);
        is_fill.push_back(fill ? 1 : 0);
        return (int)re.size() - 1;
    }

    // translate (build.cpp:197-249): first touch assigns the next internal
    // index to an external index, for rows and columns alike.
    int touch(int ext) {
        if (e2i_r[ext] == -1) {
            ++current;
            e2i_r[ext] = current;
            e2i_c[ext] = current;
            i2e_r[current] = ext;
            i2e_c[current] = ext;
        }
        return e2i_r[ext];
    }

    // get_element before rows are linked (build.cpp:55-80, 103-181)
    int get_element(int ext_r, int ext_c) {
        const int r = touch(ext_r);
        const int c = e2i_c[ext_c] == -1 ? touch(ext_c) : e2i_c[ext_c];
        if (r == c && dg[r] >= 0) return dg[r];
        int pred = -1, e = fc[c];
        while (e >= 0) {
            if (erow[e] < r) { pred = e; e = ncol[e]; }
            else if (erow[e] == r) return e;
            else break;
        }
        const int ne = new_element(r, c, false);
        if (r == c) dg[r] = ne;
        ncol[ne] = col_link(pred, c);
        col_link(pred, c) = ne;
        return ne;
    }

    void link_rows() {  // build.cpp:183-195
        for (int c = n; c >= 1; --c)
            for (int e = fc[c]; e >= 0; e = ncol[e]) {
                ecol[e] = c;
                nrow[e] = fr[erow[e]];
                fr[erow[e]] = e;
            }
        rows_linked = true;
    }

    // ---- MNA pre-ordering (utils.cpp:8-89)
    int count_twins(int c, int& t1, int& t2) {
        int twins = 0;
        for (int a = fc[c]; a >= 0; a = ncol[a]) {
            if (std::fabs(re[a]) == 1.0) {
                const int r = erow[a];
                int b = fc[r];
                while (b >= 0 && erow[b] != c) b = ncol[b];
                if (b >= 0 && std::fabs(re[b]) == 1.0) {
                    if (++twins >= 2) return twins;
                    t1 = a; ecol[a] = c;
                    t2 = b; ecol[b] = r;
                }
            }
        }
        return twins;
    }
    void swap_cols(int t1, int t2) {
        const int c1 = ecol[t1], c2 = ecol[t2];
        std::swap(fc[c1], fc[c2]);
        std::swap(i2e_c[c1], i2e_c[c2]);
        e2i_c[i2e_c[c2]] = c2;
        e2i_c[i2e_c[c1]] = c1;
        dg[c1] = t2;
        dg[c2] = t1;
    }
    void mna_preorder() {
        if (rows_linked) return;
        int start = 1;
        bool again;
        do {
            again = false;
            bool swapped = false;
            for (int j = start; j <= n; ++j) {
                if (dg[j] >= 0) continue;
                int t1 = -1, t2 = -1;
                const int tw = count_twins(j, t1, t2);
                if (tw == 1) { swap_cols(t1, t2); swapped = true; }
                else if (tw > 1 && !again) { again = true; start = j; }
            }
            if (again)
                for (int j = start; !swapped && j <= n; ++j) {
                    if (dg[j] >= 0) continue;
                    int t1 = -1, t2 = -1;
                    count_twins(j, t1, t2);
                    if (t1 < 0) throw Error(NSA_ERR_RUNTIME, "mna_preorder: no twin to swap");
                    swap_cols(t1, t2);
                    swapped = true;
                }
        } while (again);
    }

    // ---- Markowitz counts (factor.cpp:61-115)
    void count_markowitz(int step) {
        for (int i = step; i <= n; ++i) {
            int cnt = -1, e = fr[i];
            while (e >= 0 && ecol[e] < step) e = nrow[e];
            for (; e >= 0; e = nrow[e]) ++cnt;
            mr[i] = cnt;
        }
        for (int i = step; i <= n; ++i) {
            int cnt = -1, e = fc[i];
            while (e >= 0 && erow[e] < step) e = ncol[e];
            for (; e >= 0; e = ncol[e]) ++cnt;
            mc[i] = cnt;
        }
    }
    static long product_of(int a, int b) {
        if ((a > SHRT_MAX && b != 0) || (b > SHRT_MAX && a != 0)) {
            const double fp = (double)a * (double)b;
            return fp >= (double)LONG_MAX ? LONG_MAX : (long)fp;
        }
        return (long)a * b;
    }
    void markowitz_products(int step) {
        singletons = 0;
        for (int i = step; i <= n; ++i) {
            const int a = mr[i], b = mc[i];
            if ((a > SHRT_MAX && b != 0) || (b > SHRT_MAX && a != 0)) {
                mp[i] = product_of(a, b);
            } else {
                mp[i] = (long)a * b;
                if (mp[i] == 0) ++singletons;
            }
        }
    }

    // element_mag (element.hpp:18-20); im is zero on a real matrix
    double mag(int e) const { return cmag(re[e], im[e]); }

    double biggest_in_col_excluding(int e, int step) {  // factor.cpp:38-55
        const int r = erow[e], c = ecol[e];
        int x = fc[c];
        while (x >= 0 && erow[x] < step) x = ncol[x];
        double largest = (erow[x] != r) ? mag(x) : 0.0;
        while ((x = ncol[x]) >= 0) {
            const double m = mag(x);
            if (m > largest && erow[x] != r) largest = m;
        }
        return largest;
    }
    // ---- pivot searches (factor.cpp:121-396)
    int search_singleton(int step) {
        mp[n + 1] = mp[step];
        mp[step - 1] = 0;
        int count = singletons--;
        int scan = n + 1;
        while (count-- > 0) {
            while (mp[scan] != 0) --scan;
            int i = scan;
            --scan;
            if (i < step) break;
            if (i > n) i = step;
            int chosen = dg[i];
            if (chosen >= 0) {
                const double m = mag(chosen);
                if (m > absth && m > rel * biggest_in_col_excluding(chosen, step)) return chosen;
            } else {
                // A singleton without a diagonal ends the singleton search
                // (the reduced row/column still holds its element).
                if (mc[i] == 0) {
                    chosen = fc[i];
                    while (chosen >= 0 && erow[chosen] < step) chosen = ncol[chosen];
                    if (chosen >= 0) break;
                    break;
                } else {
                    chosen = fr[i];
                    while (chosen >= 0 && ecol[chosen] < step) chosen = nrow[chosen];
                    if (chosen >= 0) break;
                    break;
                }
            }
        }
        ++singletons;
        return -1;
    }

    int quick_diagonal(int step) {
        int chosen = -1;
        long minmp = LONG_MAX;
        mp[n + 1] = mp[step];
        mp[step - 1] = -1;
        int scan = n + 2;
        for (;;) {
            --scan;
            while (mp[scan] >= minmp) --scan;
            int i = scan;
            if (i < step) break;
            if (i > n) i = step;
            const int d = dg[i];
            if (d < 0) continue;
            const double m = mag(d);
            if (m <= absth) continue;
            if (mp[scan] == 1) {
                int oir = nrow[d], oic = ncol[d];
                if (oir < 0 && oic < 0) {
                    oir = fr[i];
                    while (oir >= 0) {
                        if (ecol[oir] >= step && ecol[oir] != i) break;
                        oir = nrow[oir];
                    }
                    oic = fc[i];
                    while (oic >= 0) {
                        if (erow[oic] >= step && erow[oic] != i) break;
                        oic = ncol[oic];
                    }
                }
                if (oir >= 0 && oic >= 0 && ecol[oir] == erow[oic]) {
                    const double off = std::max(mag(oir), mag(oic));
                    if (m >= off) return d;
                }
            }
            minmp = mp[scan];
            chosen = d;
        }
        if (chosen >= 0) {
            const double big = biggest_in_col_excluding(chosen, step);
            if (mag(chosen) <= rel * big) chosen = -1;
        }
        return chosen;
    }

    int search_diagonal(int step) {
        int chosen = -1;
        long minmp = LONG_MAX;
        int ties = 0;
        double ratio_acc = 0.0;
        mp[n + 1] = mp[step];
        for (int j = n + 1; j > step; --j) {
            if (mp[j] > minmp) continue;
            const int i = (j > n) ? step : j;
            const int d = dg[i];
            if (d < 0) continue;
            const double m = mag(d);
            if (m <= absth) continue;
            const double big = biggest_in_col_excluding(d, step);
            if (m <= rel * big) continue;
            if (mp[j] < minmp) {
                chosen = d;
                minmp = mp[j];
                ratio_acc = big / m;
                ties = 0;
            } else {
                ++ties;
                const double ratio = big / m;
                if (ratio < ratio_acc) { chosen = d; ratio_acc = ratio; }
                if (ties >= minmp * 5) return chosen;
            }
        }
        return chosen;
    }

    int search_entire(int step) {
        int chosen = -1, largest_el = -1;
        double largest_mag = 0.0;
        long minmp = LONG_MAX;
        int ties = 0;
        double ratio_acc = 0.0;
        for (int i = step; i <= n; ++i) {
            int e = fc[i];
            while (e >= 0 && erow[e] < step) e = ncol[e];
            double lic = 0.0;
            for (int x = e; x >= 0; x = ncol[x]) if (mag(x) > lic) lic = mag(x);
            if (lic == 0.0) continue;
            for (; e >= 0; e = ncol[e]) {
                const double m = mag(e);
                if (m > largest_mag) { largest_mag = m; largest_el = e; }
                const long prod = (long)mr[erow[e]] * mc[ecol[e]];
                if (prod <= minmp && m > rel * lic && m > absth) {
                    if (prod < minmp) {
                        chosen = e;
                        minmp = prod;
                        ratio_acc = lic / m;
                        ties = 0;
                    } else {
                        ++ties;
                        const double ratio = lic / m;
                        if (ratio < ratio_acc) { chosen = e; ratio_acc = ratio; }
                        if (ties >= minmp * 5) return chosen;
                    }
                }
            }
        }
        if (chosen >= 0) return chosen;
        if (largest_mag == 0.0) { singular = true; return -1; }
        return largest_el;
    }

    int search_pivot(int step) {
        if (singletons) {
            const int c = search_singleton(step);
            if (c >= 0) return c;
        }
        int c = quick_diagonal(step);
        if (c >= 0) return c;
        c = search_diagonal(step);
        if (c >= 0) return c;
        return search_entire(step);
    }

    // ---- list surgery for row/column exchange (factor.cpp:402-530)
    void exchange_col_elements(int row1, int e1, int row2, int e2, int col) {
        int pred1 = -1, scan = fc[col];
        while (erow[scan] < row1) { pred1 = scan; scan = ncol[scan]; }
        if (e1 >= 0) {
            const int below1 = ncol[e1];
            if (e2 < 0) {
                if (below1 >= 0 && erow[below1] < row2) {
                    col_link(pred1, col) = below1;
                    int p = below1, pred2;
                    do { pred2 = p; p = ncol[p]; } while (p >= 0 && erow[p] < row2);
                    ncol[pred2] = e1;
                    ncol[e1] = p;
                }
                erow[e1] = row2;
            } else {
                if (erow[below1] == row2) {
                    ncol[e1] = ncol[e2];
                    ncol[e2] = e1;
                    col_link(pred1, col) = e2;
                } else {
                    int p = below1, pred2;
                    do { pred2 = p; p = ncol[p]; } while (erow[p] < row2);
                    const int below2 = ncol[e2];
                    col_link(pred1, col) = e2;
                    ncol[e2] = below1;
                    ncol[pred2] = e1;
                    ncol[e1] = below2;
                }
                erow[e1] = row2;
                erow[e2] = row1;
            }
        } else {
            const int below1 = scan;
            if (erow[below1] != row2) {
                int p = scan, pred2;
                do { pred2 = p; p = ncol[p]; } while (erow[p] < row2);
                ncol[pred2] = ncol[e2];
                col_link(pred1, col) = e2;
                ncol[e2] = below1;
            }
            erow[e2] = row1;
        }
    }
    void exchange_row_elements(int col1, int e1, int col2, int e2, int row) {
        int pred1 = -1, scan = fr[row];
        while (ecol[scan] < col1) { pred1 = scan; scan = nrow[scan]; }
        if (e1 >= 0) {
            const int right1 = nrow[e1];
            if (e2 < 0) {
                if (right1 >= 0 && ecol[right1] < col2) {
                    row_link(pred1, row) = right1;
                    int p = right1, pred2;
                    do { pred2 = p; p = nrow[p]; } while (p >= 0 && ecol[p] < col2);
                    nrow[pred2] = e1;
                    nrow[e1] = p;
                }
                ecol[e1] = col2;
            } else {
                if (ecol[right1] == col2) {
                    nrow[e1] = nrow[e2];
                    nrow[e2] = e1;
                    row_link(pred1, row) = e2;
                } else {
                    int p = right1, pred2;
                    do { pred2 = p; p = nrow[p]; } while (ecol[p] < col2);
                    const int right2 = nrow[e2];
                    row_link(pred1, row) = e2;
                    nrow[e2] = right1;
                    nrow[pred2] = e1;
                    nrow[e1] = right2;
                }
                ecol[e1] = col2;
                ecol[e2] = col1;
            }
        } else {
            const int right1 = scan;
            if (ecol[right1] != col2) {
                int p = scan, pred2;
                do { pred2 = p; p = nrow[p]; } while (ecol[p] < col2);
                nrow[pred2] = nrow[e2];
                row_link(pred1, row) = e2;
                nrow[e2] = right1;
            }
            ecol[e2] = col1;
        }
    }
    void row_exchange(int r1, int r2) {
        if (r1 > r2) std::swap(r1, r2);
        int a = fr[r1], b = fr[r2];
        while (a >= 0 || b >= 0) {
            int column, e1, e2;
            if (a < 0) { column = ecol[b]; e1 = -1; e2 = b; b = nrow[b]; }
            else if (b < 0) { column = ecol[a]; e1 = a; e2 = -1; a = nrow[a]; }
            else if (ecol[a] < ecol[b]) { column = ecol[a]; e1 = a; e2 = -1; a = nrow[a]; }
            else if (ecol[a] > ecol[b]) { column = ecol[b]; e1 = -1; e2 = b; b = nrow[b]; }
            else { column = ecol[a]; e1 = a; e2 = b; a = nrow[a]; b = nrow[b]; }
            exchange_col_elements(r1, e1, r2, e2, column);
        }
        if (vectors) std::swap(mr[r1], mr[r2]);
        std::swap(fr[r1], fr[r2]);
        std::swap(i2e_r[r1], i2e_r[r2]);
        e2i_r[i2e_r[r1]] = r1;
        e2i_r[i2e_r[r2]] = r2;
    }
    void col_exchange(int c1, int c2) {
        if (c1 > c2) std::swap(c1, c2);
        int a = fc[c1], b = fc[c2];
        while (a >= 0 || b >= 0) {
            int row, e1, e2;
            if (a < 0) { row = erow[b]; e1 = -1; e2 = b; b = ncol[b]; }
            else if (b < 0) { row = erow[a]; e1 = a; e2 = -1; a = ncol[a]; }
            else if (erow[a] < erow[b]) { row = erow[a]; e1 = a; e2 = -1; a = ncol[a]; }
            else if (erow[a] > erow[b]) { row = erow[b]; e1 = -1; e2 = b; b = ncol[b]; }
            else { row = erow[a]; e1 = a; e2 = b; a = ncol[a]; b = ncol[b]; }
            exchange_row_elements(c1, e1, c2, e2, row);
        }
        if (vectors) std::swap(mc[c1], mc[c2]);
        std::swap(fc[c1], fc[c2]);
        std::swap(i2e_c[c1], i2e_c[c2]);
        e2i_c[i2e_c[c1]] = c1;
        e2i_c[i2e_c[c2]] = c2;
    }
    int find_in_col(int c, int r) {
        for (int e = fc[c]; e >= 0; e = ncol[e]) {
            if (erow[e] < r) continue;
            return erow[e] == r ? e : -1;
        }
        return -1;
    }
    void adjust_singletons(long now, long before) {
        if ((now == 0) != (before == 0)) {
            if (before == 0) --singletons; else ++singletons;
        }
    }
    void exchange_rows_and_cols(int piv, int step) {  // factor.cpp:642-700
        const int row = erow[piv], col = ecol[piv];
        if (row == step && col == step) return;
        if (row == col) {
            row_exchange(step, row);
            col_exchange(step, col);
            std::swap(mp[step], mp[row]);
            std::swap(dg[row], dg[step]);
        } else {
            const long old_step = mp[step], old_row = mp[row], old_col = mp[col];
            if (row != step) {
                row_exchange(step, row);
                mp[row] = (long)mr[row] * mc[row];
                adjust_singletons(mp[row], old_row);
            }
            if (col != step) {
                col_exchange(step, col);
                mp[col] = (long)mc[col] * mr[col];
                adjust_singletons(mp[col], old_col);
                dg[col] = find_in_col(col, col);
            }
            if (row != step) dg[row] = find_in_col(row, row);
            dg[step] = find_in_col(step, step);
            mp[step] = (long)mc[step] * mr[step];
            adjust_singletons(mp[step], old_step);
        }
    }

    int create_fillin(int r, int c) {  // factor.cpp:706-729 with build.cpp:124-163
        int pred = -1, e = fc[c];
        while (e >= 0 && erow[e] < r) { pred = e; e = ncol[e]; }
        const int ne = new_element(r, c, true);
        ++fillins;
        if (r == c) dg[r] = ne;
        ncol[ne] = col_link(pred, c);
        col_link(pred, c) = ne;
        int rp = -1, s = fr[r];
        while (s >= 0 && ecol[s] < c) { rp = s; s = nrow[s]; }
        nrow[ne] = row_link(rp, r);
        row_link(rp, r) = ne;
        mp[r] = (long)(++mr[r]) * mc[r];
        if (mr[r] == 1 && mc[r] != 0) --singletons;
        mp[c] = (long)(++mc[c]) * mr[c];
        if (mr[c] != 0 && mc[c] == 1) --singletons;
        return ne;
    }

    // complex_row_col_elimination (factor.cpp:769-804)
    bool eliminate_complex(int piv) {
        if (mag(piv) == 0.0) return false;
        crecip(re[piv], im[piv], re[piv], im[piv]);
        for (int up = nrow[piv]; up >= 0; up = nrow[up]) {
            cmul(re[up], im[up], re[piv], im[piv]);
            int sub = ncol[up];
            for (int lo = ncol[piv]; lo >= 0; lo = ncol[lo]) {
                const int r = erow[lo];
                while (sub >= 0 && erow[sub] < r) sub = ncol[sub];
                if (sub < 0 || erow[sub] > r) sub = create_fillin(r, ecol[up]);
                cmsub(re[sub], im[sub], re[up], im[up], re[lo], im[lo]);
                sub = ncol[sub];
            }
        }
        return true;
    }

    bool eliminate(int piv) {  // real_row_col_elimination, factor.cpp:735-767
        if (cplx) return eliminate_complex(piv);
        if (std::fabs(re[piv]) == 0.0) return false;
        re[piv] = 1.0 / re[piv];
        for (int up = nrow[piv]; up >= 0; up = nrow[up]) {
            re[up] *= re[piv];
            int sub = ncol[up];
            for (int lo = ncol[piv]; lo >= 0; lo = ncol[lo]) {
                const int r = erow[lo];
                while (sub >= 0 && erow[sub] < r) sub = ncol[sub];
                if (sub < 0 || erow[sub] > r) sub = create_fillin(r, ecol[up]);
                re[sub] -= re[up] * re[lo];
                sub = ncol[sub];
            }
        }
        return true;
    }

    void update_markowitz(int piv) {  // factor.cpp:806-845
        for (int e = ncol[piv]; e >= 0; e = ncol[e]) {
            const int r = erow[e];
            --mr[r];
            mp[r] = product_of(mr[r], mc[r]);
            if (mr[r] == 0) ++singletons;
        }
        for (int e = nrow[piv]; e >= 0; e = nrow[e]) {
            const int c = ecol[e];
            --mc[c];
            mp[c] = product_of(mr[c], mc[c]);
            if (mc[c] == 0 && mr[c] != 0) ++singletons;
        }
    }

    int order_and_factor() {  // factor.cpp:911-978 (rhs = nullptr)
        singular = false;
        int step = 1;
        if (!needs_ordering) {
            bool reorder = false;
            for (step = 1; step <= n; ++step) {
                const int piv = dg[step];
                if (piv < 0) { reorder = true; break; }
                double largest = 0.0;
                for (int e = ncol[piv]; e >= 0; e = ncol[e]) if (mag(e) > largest) largest = mag(e);
                if (largest * rel < mag(piv)) {
                    if (!eliminate(piv)) return 1;
                } else { reorder = true; break; }
            }
            if (!reorder) { needs_ordering = false; return 0; }
        } else {
            if (!rows_linked) link_rows();
            if (!vectors) {
                mr.assign(n + 1, 0);
                mc.assign(n + 1, 0);
                mp.assign(n + 2, 0);
                vectors = true;
            }
        }
        count_markowitz(step);
        markowitz_products(step);
        for (; step <= n; ++step) {
            const int piv = search_pivot(step);
            if (piv < 0) return 1;
            exchange_rows_and_cols(piv, step);
            if (!eliminate(piv)) return 1;
            update_markowitz(piv);
        }
        needs_ordering = false;
        return 0;
    }
};

Sparse13::Sparse13(const Sparse13& o)
    : impl_(new Impl(*o.impl_)), n_(o.n_), acp_(o.acp_), ari_(o.ari_) {}

Sparse13::Sparse13(int n, const int32_t* col_ptr, const int32_t* row_idx)
    : impl_(new Impl()), n_(n), acp_(col_ptr, col_ptr + n + 1),
      ari_(row_idx, row_idx + col_ptr[n]) {
    Impl& m = *impl_;
    m.n = n;
    const int alloc = std::max(n, 6);
    m.fc.assign(alloc + 1, -1);
    m.fr.assign(alloc + 1, -1);
    m.dg.assign(alloc + 1, -1);
    m.i2e_r.resize(alloc + 1);
    m.i2e_c.resize(alloc + 1);
    for (int i = 0; i <= alloc; ++i) { m.i2e_r[i] = i; m.i2e_c[i] = i; }
    m.e2i_r.assign(alloc + 1, -1);
    m.e2i_c.assign(alloc + 1, -1);
    m.e2i_r[0] = 0;
    m.e2i_c[0] = 0;
    // NeoSolver::symbolic (neo_solver.cpp:113-146): elements in CSC order.
    m.slot_elem.resize(acp_[n]);
    for (int j = 0; j < n; ++j)
        for (int p = acp_[j]; p < acp_[j + 1]; ++p)
            m.slot_elem[p] = m.get_element(ari_[p] + 1, j + 1);
    if (m.current != n)
        throw Error(NSA_ERR_RUNTIME, "Sparse13: pattern leaves unknowns untouched");
}

Sparse13::~Sparse13() { delete impl_; }

std::unique_ptr<Sparse13> Sparse13::clone() const {
    std::unique_ptr<Sparse13> c(new Sparse13(*this));
    return c;
}

int Sparse13::numeric(const double* values, double gmin) {
    Impl& m = *impl_;
    // load_real: fill-ins cleared, originals reloaded (neo_solver.cpp:45-50)
    for (size_t e = 0; e < m.re.size(); ++e) if (m.is_fill[e]) m.re[e] = 0.0;
    for (size_t p = 0; p < m.slot_elem.size(); ++p) m.re[m.slot_elem[p]] = values[p];
    if (!m.preordered) { m.mna_preorder(); m.preordered = true; }
    if (gmin != 0.0)
        for (int i = m.n; i > 0; --i) if (m.dg[i] >= 0) m.re[m.dg[i]] += gmin;
    m.rel = 1e-3;
    m.absth = 1e-13;
    return m.order_and_factor();
}

int Sparse13::numeric_complex(const double* ax) {
    Impl& m = *impl_;
    // load_complex (neo_solver.cpp:104-111): fill-ins cleared, both parts
    // of the originals reloaded
    m.cplx = true;
    for (size_t e = 0; e < m.re.size(); ++e)
        if (m.is_fill[e]) { m.re[e] = 0.0; m.im[e] = 0.0; }
    for (size_t p = 0; p < m.slot_elem.size(); ++p) {
        m.re[m.slot_elem[p]] = ax[2 * p];
        m.im[m.slot_elem[p]] = ax[2 * p + 1];
    }
    if (!m.preordered) { m.mna_preorder(); m.preordered = true; }
    m.rel = 1e-3;
    m.absth = 1e-13;
    return m.order_and_factor();
}

int Sparse13::fillins() const { return impl_->fillins; }

LuStructure Sparse13::structure() const {
    const Impl& m = *impl_;
    std::vector<int32_t> i2e_row(n_), i2e_col(n_), cp(n_ + 1, 0), rows;
    for (int j = 1; j <= n_; ++j) {
        i2e_row[j - 1] = m.i2e_r[j];
        i2e_col[j - 1] = m.i2e_c[j];
        for (int e = m.fc[j]; e >= 0; e = m.ncol[e]) rows.push_back(m.erow[e]);
        cp[j] = (int32_t)rows.size();
    }
    return structure_from_s13(n_, acp_.data(), ari_.data(), i2e_row, i2e_col, cp, rows);
}

}  // namespace nsa
