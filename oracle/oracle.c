/* CPU restatement of the reference's numeric LU path (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's
 * CPU-baseline leg as the checker; never linked into the product.
 *
 * Parity pins (tests/test_oracle.py):
 *   orc_amd_order     - against the compiled reference amd_ordering on grids,
 *                       random patterns and the 64x64 RC-mesh pattern.  Tie
 *                       order among more than 16 dense nodes follows a
 *                       stable sort, which is what libstdc++ std::sort does
 *                       for 16 or fewer elements only: beyond 16 dense nodes
 *                       the order is parity unpinned (no config reaches it).
 *   orc_gp_*          - against the compiled reference AmdLuSolver (factors,
 *                       pivot order, replay, solve) and x_true recovery.
 *   orc_s13_*         - against the compiled reference NeoSolver factors and
 *                       solves, closed forms from the reference's own
 *                       tests/unit/test_neo_solver.cpp, and the transpose
 *                       identity <A^T y, x> = <y, A x>.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Minimum-degree ordering: src/core/amd.cpp:14-143                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t* v;
    int n, cap;
} ivec;

static void iv_push(ivec* a, int32_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 8;
        a->v = (int32_t*)realloc(a->v, sizeof(int32_t) * a->cap);
    }
    a->v[a->n++] = x;
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* first index with v[i] >= x */
static int iv_lower(const ivec* a, int32_t x) {
    int lo = 0, hi = a->n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a->v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static void iv_insert_at(ivec* a, int pos, int32_t x) {
    iv_push(a, 0);
    memmove(a->v + pos + 1, a->v + pos, sizeof(int32_t) * (a->n - 1 - pos));
    a->v[pos] = x;
}

static void iv_erase_at(ivec* a, int pos) {
    memmove(a->v + pos, a->v + pos + 1, sizeof(int32_t) * (a->n - 1 - pos));
    a->n--;
}

int orc_amd_order(int n, const int32_t* col_ptr, const int32_t* row_idx,
                  int32_t* perm) {
    if (n <= 0) return 0;
    ivec* adj = (ivec*)calloc(n, sizeof(ivec));
    for (int j = 0; j < n; ++j)
        for (int p = col_ptr[j]; p < col_ptr[j + 1]; ++p) {
            int i = row_idx[p];
            if (i != j) { iv_push(&adj[j], i); iv_push(&adj[i], j); }
        }
    for (int i = 0; i < n; ++i) {  /* sort + unique (amd.cpp:30-34) */
        if (adj[i].n == 0) continue;
        qsort(adj[i].v, adj[i].n, sizeof(int32_t), cmp_i32);
        int w = 1;
        for (int k = 1; k < adj[i].n; ++k)
            if (adj[i].v[k] != adj[i].v[w - 1]) adj[i].v[w++] = adj[i].v[k];
        adj[i].n = w;
    }
    int* degree = (int*)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) degree[i] = adj[i].n;

    /* dense threshold (amd.cpp:48-51) */
    int dense_thresh = (int)(10.0 * sqrt((double)n));
    if (dense_thresh < 16) dense_thresh = 16;
    if (dense_thresh > n) dense_thresh = n;
    char* is_dense = (char*)calloc(n, 1);
    int32_t* dense = (int32_t*)malloc(sizeof(int32_t) * n);
    int ndense = 0;
    for (int i = 0; i < n; ++i)
        if (degree[i] > dense_thresh) { is_dense[i] = 1; dense[ndense++] = i; }
    if (ndense) {
        for (int i = 0; i < n; ++i) {
            if (is_dense[i]) continue;
            int w = 0;
            for (int k = 0; k < adj[i].n; ++k)
                if (!is_dense[adj[i].v[k]]) adj[i].v[w++] = adj[i].v[k];
            adj[i].n = w;
            degree[i] = w;
        }
    }
    char* elim = (char*)calloc(n, 1);
    for (int k = 0; k < ndense; ++k) elim[dense[k]] = 1;
    int np = 0;
    int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * n);
    for (int step = 0; step < n - ndense; ++step) {
        int best = -1, best_deg = n + 1;   /* strict <: lowest index wins */
        for (int i = 0; i < n; ++i)
            if (!elim[i] && degree[i] < best_deg) { best_deg = degree[i]; best = i; }
        perm[np++] = best;
        elim[best] = 1;
        int nn = 0;
        for (int k = 0; k < adj[best].n; ++k)
            if (!elim[adj[best].v[k]]) nb[nn++] = adj[best].v[k];
        for (int a = 0; a < nn; ++a)           /* clique fill (amd.cpp:102-113) */
            for (int b = a + 1; b < nn; ++b) {
                int32_t u = nb[a], v = nb[b];
                int pu = iv_lower(&adj[u], v);
                if (pu < adj[u].n && adj[u].v[pu] == v) continue;
                iv_insert_at(&adj[u], pu, v);
                iv_insert_at(&adj[v], iv_lower(&adj[v], u), u);
            }
        for (int a = 0; a < nn; ++a) {         /* drop best, recount (:116-122) */
            int32_t x = nb[a];
            iv_erase_at(&adj[x], iv_lower(&adj[x], best));
            int d = 0;
            for (int k = 0; k < adj[x].n; ++k) d += !elim[adj[x].v[k]];
            degree[x] = d;
        }
    }
    /* dense nodes last, ascending original degree (:130-137) */
    for (int a = 1; a < ndense; ++a) {
        int32_t key = dense[a];
        int b = a - 1;
        while (b >= 0 && degree[dense[b]] > degree[key]) { dense[b + 1] = dense[b]; --b; }
        dense[b + 1] = key;
    }
    for (int k = 0; k < ndense; ++k) perm[np++] = dense[k];

    for (int i = 0; i < n; ++i) free(adj[i].v);
    free(adj); free(degree); free(is_dense); free(dense); free(elim); free(nb);
    return np;
}

/* ------------------------------------------------------------------------ */
/* Gilbert-Peierls LU: src/core/amd_lu_solver.cpp:44-197                      */
/* ------------------------------------------------------------------------ */

/* lu_dfs, amd_lu_solver.cpp:44-72 */
static int lu_dfs(int j, const int32_t* Lp, const int32_t* Li, const int32_t* pinv,
                  int top, int32_t* xi, int32_t* pstack, char* visited) {
    int head = 0;
    xi[0] = j;
    while (head >= 0) {
        j = xi[head];
        int jnew = pinv[j];
        if (!visited[j]) {
            visited[j] = 1;
            pstack[head] = (jnew < 0) ? 0 : Lp[jnew];
        }
        int done = 1;
        int p2 = (jnew < 0) ? 0 : Lp[jnew + 1];
        for (int p = pstack[head]; p < p2; ++p) {
            int i = Li[p];
            if (visited[i]) continue;
            pstack[head] = p;
            xi[++head] = i;
            done = 0;
            break;
        }
        if (done) {
            --head;
            xi[--top] = j;
        }
    }
    return top;
}

int orc_gp_factor(int n, const int32_t* col_ptr, const int32_t* row_idx,
                  const double* values, double diag_gmin, double pivot_tol,
                  const int32_t* q, int32_t* pinv, int32_t* Lp, int32_t* Li,
                  double* Lx, int lcap, int32_t* Up, int32_t* Ui, double* Ux,
                  int ucap) {
    double* x = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    int32_t* xi = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    int32_t* pstack = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    char* visited = (char*)calloc(n > 0 ? n : 1, 1);
    int ln = 0, un = 0, rc = 0;
    for (int i = 0; i < n; ++i) pinv[i] = -1;
    for (int k = 0; k < n; ++k) {
        int col = q[k];
        Lp[k] = ln;
        Up[k] = un;
        int top = n;
        memset(visited, 0, n);
        for (int p = col_ptr[col]; p < col_ptr[col + 1]; ++p) {
            int i = row_idx[p];
            if (!visited[i]) top = lu_dfs(i, Lp, Li, pinv, top, xi, pstack, visited);
        }
        for (int p = top; p < n; ++p) x[xi[p]] = 0.0;
        for (int p = col_ptr[col]; p < col_ptr[col + 1]; ++p) x[row_idx[p]] = values[p];
        if (diag_gmin != 0.0) x[col] += diag_gmin;
        for (int px = top; px < n; ++px) {         /* :134-142 */
            int j = xi[px];
            int J = pinv[j];
            if (J < 0) continue;
            double xj = x[j];
            for (int p = Lp[J] + 1; p < Lp[J + 1]; ++p) x[Li[p]] -= Lx[p] * xj;
        }
        int ipiv = -1;
        double amax = -1.0;
        for (int p = top; p < n; ++p) {            /* :145-157 */
            int i = xi[p];
            if (pinv[i] < 0) {
                double t = fabs(x[i]);
                if (t > amax) { amax = t; ipiv = i; }
            } else {
                if (un >= ucap) { rc = -1; goto out; }
                Ui[un] = pinv[i];
                Ux[un++] = x[i];
            }
        }
        if (ipiv == -1 || amax <= 0.0) { rc = 1; goto out; }
        if (pinv[col] < 0 && fabs(x[col]) >= amax * pivot_tol) ipiv = col;
        double pivot = x[ipiv];
        if (pivot == 0.0) { rc = 1; goto out; }
        if (un >= ucap || ln >= lcap) { rc = -1; goto out; }
        Ui[un] = k;
        Ux[un++] = pivot;
        pinv[ipiv] = k;
        Li[ln] = ipiv;
        Lx[ln++] = 1.0;
        for (int p = top; p < n; ++p) {
            int i = xi[p];
            if (pinv[i] < 0) {
                if (ln >= lcap) { rc = -1; goto out; }
                Li[ln] = i;
                Lx[ln++] = x[i] / pivot;
            }
            x[i] = 0.0;
        }
    }
    Lp[n] = ln;
    Up[n] = un;
    for (int p = 0; p < ln; ++p) Li[p] = pinv[Li[p]];   /* :181 */
out:
    free(x); free(xi); free(pstack); free(visited);
    return rc;
}

int orc_gp_refactor(int n, const int32_t* col_ptr, const int32_t* row_idx,
                    const double* values, double diag_gmin, double pivot_tol,
                    const int32_t* q, const int32_t* pinv, const int32_t* Lp,
                    const int32_t* Li, double* Lx, const int32_t* Up,
                    const int32_t* Ui, double* Ux) {
    double* X = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    int ok = 1;
    for (int k = 0; k < n && ok; ++k) {
        int col = q[k];
        for (int p = col_ptr[col]; p < col_ptr[col + 1]; ++p)
            X[pinv[row_idx[p]]] = values[p];        /* scatter_row_ (:190-193) */
        if (diag_gmin != 0.0) X[pinv[col]] += diag_gmin;
        double colmax = 0.0;
        int ub = Up[k], ue = Up[k + 1];
        for (int up = ub; up < ue - 1; ++up) {
            int j = Ui[up];
            double ujk = X[j];
            X[j] = 0.0;
            Ux[up] = ujk;
            double a = fabs(ujk);
            if (a > colmax) colmax = a;
            for (int p = Lp[j] + 1; p < Lp[j + 1]; ++p) X[Li[p]] -= Lx[p] * ujk;
        }
        double ukk = X[k];
        X[k] = 0.0;
        double aukk = fabs(ukk);
        if (aukk > colmax) colmax = aukk;
        if (ukk == 0.0 || (colmax > 0.0 && aukk < pivot_tol * colmax)) { ok = 0; break; }
        Ux[ue - 1] = ukk;
        for (int p = Lp[k] + 1; p < Lp[k + 1]; ++p) {
            int i = Li[p];
            Lx[p] = X[i] / ukk;
            X[i] = 0.0;
        }
    }
    free(X);
    return ok;
}

void orc_gp_solve(int n, const int32_t* q, const int32_t* pinv,
                  const int32_t* Lp, const int32_t* Li, const double* Lx,
                  const int32_t* Up, const int32_t* Ui, const double* Ux,
                  double* rhs) {
    double* x = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    for (int i = 0; i < n; ++i) x[pinv[i]] = rhs[i];
    for (int j = 0; j < n; ++j) {
        double xj = x[j];
        for (int p = Lp[j] + 1; p < Lp[j + 1]; ++p) x[Li[p]] -= Lx[p] * xj;
    }
    for (int j = n - 1; j >= 0; --j) {
        x[j] /= Ux[Up[j + 1] - 1];
        double xj = x[j];
        for (int p = Up[j]; p < Up[j + 1] - 1; ++p) x[Ui[p]] -= Ux[p] * xj;
    }
    for (int k = 0; k < n; ++k) rhs[q[k]] = x[k];
    free(x);
}

void orc_gp_solve_t(int n, const int32_t* q, const int32_t* pinv,
                    const int32_t* Lp, const int32_t* Li, const double* Lx,
                    const int32_t* Up, const int32_t* Ui, const double* Ux,
                    double* rhs) {
    /* A^T x = b;  (P A Q)^T = Q^T A^T P^T = U^T L^T.
     * y = Q^T b  (y[k] = b[q[k]]);  U^T w = y;  L^T z = w;  x = P^T z,
     * i.e. x[i] = z[pinv[i]]. */
    double* y = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    for (int k = 0; k < n; ++k) y[k] = rhs[q[k]];
    /* U^T forward: column k of U is row k of U^T. */
    for (int k = 0; k < n; ++k) {
        double s = y[k];
        for (int p = Up[k]; p < Up[k + 1] - 1; ++p) s -= Ux[p] * y[Ui[p]];
        y[k] = s / Ux[Up[k + 1] - 1];
    }
    /* unit L^T backward: column k of L is row k of L^T. */
    for (int k = n - 1; k >= 0; --k) {
        double s = y[k];
        for (int p = Lp[k] + 1; p < Lp[k + 1]; ++p) s -= Lx[p] * y[Li[p]];
        y[k] = s;
    }
    for (int i = 0; i < n; ++i) rhs[i] = y[pinv[i]];
    free(y);
}

/* ------------------------------------------------------------------------ */
/* Sparse 1.3 refactor / solves: src/solver/factor.cpp:980-1052,             */
/* src/solver/solve.cpp:6-47, 102-141                                         */
/* ------------------------------------------------------------------------ */

static int s13_find(const orc_s13* f, int j, int row) {  /* element (row, j) */
    for (int p = f->cptr[j - 1]; p < f->cptr[j]; ++p)
        if (f->rows[p] == row) return p;
    return -1;
}

int orc_s13_refactor(orc_s13* f, const int32_t* col_ptr, const int32_t* row_idx,
                     const double* values, double diag_gmin) {
    const int n = f->n;
    int32_t* e2i_row = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    int32_t* e2i_col = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    int32_t* dpos = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    double* dest = (double*)calloc(n + 1, sizeof(double));
    for (int i = 1; i <= n; ++i) {
        e2i_row[f->i2e_row[i - 1]] = i;
        e2i_col[f->i2e_col[i - 1]] = i;
    }
    /* load_real: fill-ins cleared, originals overwritten (build.cpp:36-53,
     * neo_solver.cpp:45-50) */
    for (int p = 0; p < f->cptr[n]; ++p) f->vals[p] = 0.0;
    for (int c = 0; c < n; ++c)
        for (int p = col_ptr[c]; p < col_ptr[c + 1]; ++p) {
            int ir = e2i_row[row_idx[p] + 1], ic = e2i_col[c + 1];
            f->vals[s13_find(f, ic, ir)] = values[p];
        }
    for (int j = 1; j <= n; ++j) dpos[j] = s13_find(f, j, j);
    if (diag_gmin != 0.0)                       /* add_diag_gmin */
        for (int i = n; i > 0; --i) f->vals[dpos[i]] += diag_gmin;
    int rc = 0;
    if (n == 0) goto out;
    if (f->vals[dpos[1]] == 0.0) { rc = 1; goto out; }
    f->vals[dpos[1]] = 1.0 / f->vals[dpos[1]];
    for (int step = 2; step <= n; ++step) {
        int cb = f->cptr[step - 1], ce = f->cptr[step];
        for (int p = cb; p < ce; ++p) dest[f->rows[p]] = f->vals[p];
        for (int p = cb; p < ce && f->rows[p] < step; ++p) {
            int r = f->rows[p];
            int dp = dpos[r];
            f->vals[p] = dest[r] * f->vals[dp];
            for (int e = dp + 1; e < f->cptr[r]; ++e)
                dest[f->rows[e]] -= f->vals[p] * f->vals[e];
        }
        for (int p = dpos[step] + 1; p < ce; ++p) f->vals[p] = dest[f->rows[p]];
        if (dest[step] == 0.0) { rc = step; goto out; }
        f->vals[dpos[step]] = 1.0 / dest[step];
    }
out:
    free(e2i_row); free(e2i_col); free(dpos); free(dest);
    return rc;
}

void orc_s13_solve(const orc_s13* f, const double* rhs, double* x) {
    const int n = f->n;
    double* inter = (double*)calloc(n + 1, sizeof(double));
    int32_t* dpos = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    for (int j = 1; j <= n; ++j) dpos[j] = s13_find(f, j, j);
    for (int i = n; i > 0; --i) inter[i] = rhs[f->i2e_row[i - 1] - 1];
    for (int i = 1; i <= n; ++i) {             /* forward, column scatter */
        double t = inter[i];
        if (t != 0.0) {
            inter[i] = (t *= f->vals[dpos[i]]);
            for (int e = dpos[i] + 1; e < f->cptr[i]; ++e) inter[f->rows[e]] -= t * f->vals[e];
        }
    }
    /* backward, row dot: U(i, c) for c > i lives in column c at row i */
    for (int i = n; i > 0; --i) {
        double t = inter[i];
        for (int c = i + 1; c <= n; ++c) {
            int p = s13_find(f, c, i);
            if (p >= 0) t -= f->vals[p] * inter[c];
        }
        inter[i] = t;
    }
    for (int i = n; i > 0; --i) x[f->i2e_col[i - 1] - 1] = inter[i];
    free(inter); free(dpos);
}

void orc_s13_solve_t(const orc_s13* f, const double* rhs, double* x) {
    const int n = f->n;
    double* inter = (double*)calloc(n + 1, sizeof(double));
    int32_t* dpos = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    for (int j = 1; j <= n; ++j) dpos[j] = s13_find(f, j, j);
    for (int i = n; i > 0; --i) inter[i] = rhs[f->i2e_col[i - 1] - 1];
    for (int i = 1; i <= n; ++i) {             /* U^T forward, row scatter */
        double t = inter[i];
        if (t != 0.0)
            for (int c = i + 1; c <= n; ++c) {
                int p = s13_find(f, c, i);
                if (p >= 0) inter[c] -= t * f->vals[p];
            }
    }
    for (int i = n; i > 0; --i) {              /* L^T backward, column dot */
        double t = inter[i];
        for (int e = dpos[i] + 1; e < f->cptr[i]; ++e) t -= f->vals[e] * inter[f->rows[e]];
        inter[i] = t * f->vals[dpos[i]];
    }
    for (int i = n; i > 0; --i) x[f->i2e_row[i - 1] - 1] = inter[i];
    free(inter); free(dpos);
}
