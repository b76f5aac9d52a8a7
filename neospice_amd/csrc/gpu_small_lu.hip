// Packed small-matrix LU: one wave (one workgroup of 64 lanes) per instance,
// the factor in LDS from the A load to the end of the solve.
//
// Refactor (SparseMatrix::factor, src/solver/factor.cpp:980-1052, applied
// right-looking with the same terms in the same order per entry):
//   1. every position takes its A value (fill: 0) and diag_gmin
//      (NeoSolver::load_real + add_diag_gmin, neo_solver.cpp:45-50, utils.cpp:91-97);
//   2. levels of step tasks: x_ij = x_ij - (x_kj * (1 / x_kk)) * x_ik;
//   3. checks: exact zero pivot (factor.cpp:997, 1045) or, for a numeric()
//      on a reused order, the threshold test of order_and_factor's reuse
//      path (factor.cpp:929-946) -- both report the first failing column;
//   4. U(k,j) = x_kj * (1 / x_kk) and the diagonal holds 1 / x_kk.
// Solve (solve.cpp:6-47) and transposed solve (solve.cpp:102-141) run the
// level-ordered row gathers of the plan on the LDS values.
//
// A level costs one LDS round trip plus its arithmetic, so the level loops
// read their descriptors from LDS too: the plan's hot block (task levels and
// the plain solve's programs) is copied there first when it fits, and each
// level's first descriptor is fetched while the previous level finishes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "small_lu.hpp"
#include "dev_terr.hpp"

namespace nsa {

namespace {

constexpr int kW = 64;
constexpr size_t kLdsLimit = 64 * 1024;

// The end of a level.  A workgroup is one wave, and a wave's LDS accesses
// are performed in the order it issues them, so the next level's loads see
// this level's stores without waiting for the stores to complete: only the
// compiler must keep the accesses on their side (a workgroup barrier would
// also wait for every outstanding LDS access).
__device__ inline void level_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Elements [0, n) of a lane-strided loop, U per lane at a time: all U
// loads are issued before any is used, so a batch costs one memory round
// trip (a plain strided loop waits for each element's load in turn).
template <int U, class Load, class Use>
__device__ inline void batched(int n, int lane, Load ld, Use use) {
    for (int k0 = lane; k0 < n; k0 += U * kW) {
        decltype(ld(0)) val[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u * kW < n) val[u] = ld(k0 + u * kW);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u * kW < n) use(k0 + u * kW, val[u]);
    }
}

struct PosVal {
    int q;
    double w;
};

__device__ inline int wave_max_i32(int v) {
    for (int o = kW / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

__device__ inline int wave_min_i32(int v) {
    for (int o = kW / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

// Plan words: the hot block from LDS when it was staged there (STAGED), the
// rest from global memory.  Staged and unstaged launches are separate
// instantiations, so every hot-block read is a plain LDS (or global) access
// rather than a flat one that would wait on both memory counters.
template <bool STAGED>
struct PlanView {
    const int32_t* __restrict__ P;  // the plan in global memory
    const int32_t* hot;             // the staged hot block (LDS)
    int32_t hot0;                   // plan offset of the hot block
    __device__ const int32_t* at(int32_t off) const {
        if (STAGED) return hot + (off - hot0);
        return P + off;
    }
};

// LDS as bytes: plan records hold byte addresses into the instance's
// values (factor value p at 8p, vector entry i at 8 (nlu + i), a zero at
// 8 (nlu + n)), so an access needs no address arithmetic.
__device__ inline double lds_ld(const char* L, int b) { return *reinterpret_cast<const double*>(L + b); }
__device__ inline void lds_st(char* L, int b, double x) { *reinterpret_cast<double*>(L + b) = x; }

// A program row record (small_lu.hpp): target, number of terms after the
// first two, diagonal value, offset of those terms, then value and source
// of terms 0 and 1 (the zero cell for both when a row has fewer terms).
struct RowRec {
    int t, ntail, dp, toff, p0, s0, p1, s1;
};

__device__ inline RowRec load_row(const int32_t* rows, int r) {
    const int4 a = *reinterpret_cast<const int4*>(rows + 8 * r);
    const int4 b = *reinterpret_cast<const int4*>(rows + 8 * r + 4);
    return RowRec{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

// One term of a row's subtraction chain.
template <int OP>
__device__ inline double term(double acc, double a, double b) {
    if ((OP & OP_SKIPZERO) && b == 0.0) return acc;
    return acc - a * b;
}

// The rest of a row's subtraction chain (terms 2..), operand loads issued
// eight terms ahead.
template <int OP>
__device__ inline double row_tail(double acc, const int32_t* tm, int cnt, const char* L) {
    int q = 0;
    for (; q + 8 <= cnt; q += 8) {
        int pos[8], src[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            pos[u] = tm[2 * (q + u)];
            src[u] = tm[2 * (q + u) + 1];
        }
        double a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a[u] = lds_ld(L, pos[u]);
            b[u] = lds_ld(L, src[u]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = term<OP>(acc, a[u], b[u]);
    }
    for (; q < cnt; ++q) acc = term<OP>(acc, lds_ld(L, tm[2 * q]), lds_ld(L, tm[2 * q + 1]));
    return acc;
}

// The diagonal operation of a program row.
template <int OP>
__device__ inline double row_diag(double acc, double dv) {
    if (OP & OP_MUL_NZ) return (acc != 0.0) ? acc * dv : acc;
    if (OP & OP_MUL_DIAG) return acc * dv;
    return acc;
}

// Row `rec` of a triangular program: acc = y[t] - sum v[pos] * y[src] in
// term order (the reference's subtraction chain for target t), then the
// diagonal operation.  yt, a0, a1, dv are y[t], v[p0], v[p1], v[dp],
// loaded ahead by the caller.
template <int OP, bool STAGED>
__device__ inline void run_row(const PlanView<STAGED>& pv, const RowRec& rec, double yt, double a0,
                               double a1, double dv, char* L) {
    // both inline sources are loaded together, so the row waits for one
    // LDS round trip
    const double b0 = lds_ld(L, rec.s0);
    const double b1 = lds_ld(L, rec.s1);
    double acc = term<OP>(term<OP>(yt, a0, b0), a1, b1);
    if (rec.ntail > 0) acc = row_tail<OP>(acc, pv.at(rec.toff), rec.ntail, L);
    lds_st(L, rec.t, row_diag<OP>(acc, dv));
}

// What a level needs of its first row, loaded ahead: the record and the
// operands that are final before the level starts.
struct RowAhead {
    RowRec rec;
    double yt, a0, a1, dv;
};

__device__ inline void fetch_ahead(RowAhead& a, const char* L) {
    a.yt = lds_ld(L, a.rec.t);
    a.a0 = lds_ld(L, a.rec.p0);
    a.a1 = lds_ld(L, a.rec.p1);
    a.dv = lds_ld(L, a.rec.dp);
}

// Level l of a program: runs the level's rows, then fetches what level l+1
// needs.  `cur` holds this level's ahead data, `nxt` receives the next
// level's; callers alternate two RowAhead variables so that no prefetched
// register is ever copied (a copy would wait for every outstanding load).
template <int OP, bool STAGED>
__device__ inline void program_level(const PlanView<STAGED>& pv, const int32_t* lev, const int32_t* rows,
                                     int nl, int last, int l, int& r0, int& r1, int& r2,
                                     const RowAhead& cur, RowAhead& nxt, char* L, int lane) {
    const int r3 = (l + 3 <= nl) ? lev[l + 3] : r2;
    nxt.rec = load_row(rows, min(r1 + lane, last));
    if (r0 + lane < r1) run_row<OP, STAGED>(pv, cur.rec, cur.yt, cur.a0, cur.a1, cur.dv, L);
    for (int r = r0 + lane + kW; r < r1; r += kW) {
        const RowRec rec = load_row(rows, r);
        run_row<OP, STAGED>(pv, rec, lds_ld(L, rec.t), lds_ld(L, rec.p0), lds_ld(L, rec.p1),
                            lds_ld(L, rec.dp), L);
    }
    level_sync();
    r0 = r1;
    r1 = r2;
    r2 = r3;
    fetch_ahead(nxt, L);
}

// Levels [0, nl) of `lev` run one row per lane, a barrier after each.
template <int OP, bool STAGED>
__device__ inline void run_levels(const PlanView<STAGED>& pv, const int32_t* lev, const int32_t* rows,
                                  int nl, char* L, int lane) {
    if (nl <= 0) return;
    const int last = lev[nl] - 1;  // rows are never empty (one per target)
    // Level bounds [r0, r1) and the next level's end r2.  Each level's
    // first row record is fetched one level ahead at a clamped index (so the
    // fetch does not wait for the bound that decides whether it is used),
    // and right after a level's barrier its operands that are already final
    // -- the factor values and the target's own entry -- are loaded, so a
    // level waits for one LDS round trip, the source entries of its terms.
    int r0 = lev[0], r1 = lev[1];
    int r2 = (nl > 1) ? lev[2] : r1;
    RowAhead A, B;
    A.rec = load_row(rows, min(r0 + lane, last));
    fetch_ahead(A, L);
    int l = 0;
    for (; l + 1 < nl; l += 2) {
        program_level<OP, STAGED>(pv, lev, rows, nl, last, l, r0, r1, r2, A, B, L, lane);
        program_level<OP, STAGED>(pv, lev, rows, nl, last, l + 1, r0, r1, r2, B, A, L, lane);
    }
    if (l < nl) program_level<OP, STAGED>(pv, lev, rows, nl, last, l, r0, r1, r2, A, B, L, lane);
}

__device__ inline int readlane_i32(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ inline double readlane_f64(double v, int lane) {
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v of lane i - 1 in lane i (lane 0: 0), a DPP wave shift.
__device__ inline double shift_up_f64(double v) {
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, 0x138, 0xf, 0xf, false);
    const unsigned hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Rows [rb, re) of a serial phase (a chain of one-row levels), in batches
// of 64 rows, row rb' + j in lane j: every lane loads its row's record and
// operands -- each source read where it is final before the batch (the
// phase's earlier batches stored theirs) -- and a term whose source a row
// of the same batch produces (srow) takes that row's result from
// registers.
//
// When every such source is the row just before or one common row of the
// batch (a hub), and no row has terms beyond the first two, the batch runs
// as a wave-wide recurrence: each step every lane evaluates its row from
// its own operands, the previous lane's value of the step before (a DPP
// shift) and, once the hub's step is done, the hub's value (broadcast
// once), so lane j's value is final from step j on and 64 steps finish the
// batch without a branch or an LDS access.  Otherwise the rows run one after the
// other, each taking earlier rows' results from their lanes and storing its
// own (terms beyond the first two are read from LDS when their row runs).
template <int OP, bool STAGED>
__device__ inline void run_serial(const PlanView<STAGED>& pv, const int32_t* rows, const int32_t* srow, int rb,
                                  int re, char* L, int lane) {
    for (int r0 = rb; r0 < re; r0 += kW) {
        const int m = min(kW, re - r0);
        const int my = min(r0 + lane, re - 1);
        const RowRec rec = load_row(rows, my);
        const int sr = srow[my];
        const double yt = lds_ld(L, rec.t), a0 = lds_ld(L, rec.p0), a1 = lds_ld(L, rec.p1);
        const double dv = lds_ld(L, rec.dp), b0 = lds_ld(L, rec.s0), b1 = lds_ld(L, rec.s1);
        // the batch's loads are complete before its rows run: said here, so
        // that no row waits for the stores of the rows before it
        asm volatile("" ::"v"(yt), "v"(a0), "v"(a1), "v"(dv), "v"(b0), "v"(b1), "v"(sr), "v"(rec.t),
                     "v"(rec.ntail), "v"(rec.toff));
        // producing rows relative to the batch (< 0: final before it)
        const int q0 = (sr & 0xffff) - 1 - r0, q1 = (sr >> 16) - 1 - r0;
        const bool adj0 = q0 >= 0 && q0 == lane - 1, adj1 = q1 >= 0 && q1 == lane - 1;
        // one other row of the batch (a hub) may feed any number of rows
        const bool hub0 = lane < m && q0 >= 0 && !adj0, hub1 = lane < m && q1 >= 0 && !adj1;
        const int hub = wave_max_i32(hub0 ? q0 : hub1 ? q1 : -1);
        const bool other = lane < m && ((hub0 && q0 != hub) || (hub1 && q1 != hub) || rec.ntail > 0);
        if (__ballot(other) == 0) {
            // steps up to the hub's own (its consumers are not final before
            // it), then the rest with the hub's value broadcast once
            double acc = 0.0, up = 0.0, x0 = b0, x1 = b1;
            const int hend = min(hub + 1, m);
            int j = 0;
#pragma unroll 4
            for (; j < hend; ++j) {
                acc = row_diag<OP>(term<OP>(term<OP>(yt, a0, adj0 ? up : x0), a1, adj1 ? up : x1), dv);
                up = shift_up_f64(acc);
            }
            if (hub >= 0) {
                const double hv = readlane_f64(acc, hub);
                x0 = hub0 ? hv : b0;
                x1 = hub1 ? hv : b1;
            }
#pragma unroll 4
            for (; j < m; ++j) {
                acc = row_diag<OP>(term<OP>(term<OP>(yt, a0, adj0 ? up : x0), a1, adj1 ? up : x1), dv);
                up = shift_up_f64(acc);
            }
            if (lane < m) lds_st(L, rec.t, acc);
        } else {
            double res = 0.0, prev = 0.0;
            for (int j = 0; j < m; ++j) {
                const int srj = readlane_i32(sr, j);
                const int Q0 = (srj & 0xffff) - 1 - r0, Q1 = (srj >> 16) - 1 - r0;
                double B0 = b0, B1 = b1;
                if (Q0 >= 0) B0 = (Q0 == j - 1) ? prev : readlane_f64(res, Q0);
                if (Q1 >= 0) B1 = (Q1 == j - 1) ? prev : readlane_f64(res, Q1);
                double acc = term<OP>(term<OP>(yt, a0, B0), a1, B1);
                if (lane == j) {
                    if (rec.ntail > 0) acc = row_tail<OP>(acc, pv.at(rec.toff), rec.ntail, L);
                    acc = row_diag<OP>(acc, dv);
                    lds_st(L, rec.t, acc);
                    res = acc;
                }
                prev = readlane_f64(acc, j);
            }
        }
        level_sync();
    }
}

// Program `prog` phase by phase.
template <int OP, bool STAGED>
__device__ inline void run_program(const PlanView<STAGED>& pv, int prog, char* L, int lane) {
    const int32_t* P = pv.P;
    const int nl = P[SH_PROG + 4 * prog];
    if (nl <= 0) return;
    const int32_t* lev = pv.at(P[SH_PROG + 4 * prog + 1]);
    const int32_t* rows = pv.at(P[SH_PROG + 4 * prog + 2]);
    const int32_t* ph = pv.at(P[SH_PROG + 4 * prog + 3]);
    const int nph = ph[0];
    const int32_t* srow = pv.at(ph[1]);
    for (int q = 0; q < nph; ++q) {
        const int4 h = reinterpret_cast<const int4*>(ph)[1 + q];
        if (h.x == PH_SERIAL) run_serial<OP, STAGED>(pv, rows, srow, lev[h.y], lev[h.z], L, lane);
        else run_levels<OP, STAGED>(pv, lev + h.y, rows, h.z - h.y, L, lane);
    }
}

// One task: x_t = x_t - (x_kj * (1 / x_kk)) * x_ik, the result stored at
// `dst`.  A forward-substitution task (pivot stored as -(address + 1)) is
// skipped when its y_k is zero, as the forward loop skips a zero entry
// (solve.cpp:22-23).
__device__ inline void run_task(const int4 tk, char* L, int dst, const int4& ahead, int far_w) {
    const bool fwd = tk.w < 0;
    const int piv = fwd ? -tk.w - 1 : tk.w;
    const double xkk = lds_ld(L, piv), xkj = lds_ld(L, tk.y), xik = lds_ld(L, tk.z), xt = lds_ld(L, tk.x);
    double r = 1.0 / xkk;
    // what this level fetched for the next ones before its own operands --
    // the next level's task (a level earlier) and the pivot word of the one
    // after -- is in registers by now; saying so (once the reciprocal is
    // done, so the division does not wait behind the claim) keeps the next
    // level from waiting for this task's store to complete before it can
    // use them
    asm volatile("" : "+v"(r) : "v"(ahead.x), "v"(ahead.y), "v"(ahead.z), "v"(ahead.w), "v"(far_w));
    const double nv = xt - (xkj * r) * xik;
    lds_st(L, dst, (fwd && xkj == 0.0) ? xt : nv);
}

// Level l of the refactor: the level's tasks, and the fetch of the first
// task of level l+2 (`far`; `cur` and `nxt` were fetched two levels and one
// level ahead, the callers rotate three variables so that no fetched
// register is ever copied).  Every lane runs a first task, the lanes past
// the level's end the clamped one with its result sent to `junk`, so that
// the common one-pass level has no branch.  t0..t2 are the task offsets of
// levels l..l+2.
__device__ inline void factor_level(const int32_t* lev, const int4* task, int last, int l, int& t0,
                                    int& t1, int& t2, const int4& cur, const int4& nxt, int4& far, char* L,
                                    int lane, int junk) {
    const int t3 = lev[l + 3];
    far = task[min(t2 + lane, last)];
    run_task(cur, L, (t0 + lane < t1) ? cur.x : junk, nxt, far.w);
    for (int t = t0 + lane + kW; t < t1; t += kW) {
        const int4 tk = task[t];
        run_task(tk, L, tk.x, nxt, far.w);
    }
    level_sync();
    t0 = t1;
    t1 = t2;
    t2 = t3;
}

template <bool STAGED>
__global__ __launch_bounds__(kW) void k_small_lu(const int32_t* __restrict__ words,
                                                const int32_t* __restrict__ offs,
                                                const SluInst* __restrict__ insts, int mode,
                                                const double* __restrict__ A, double gmin_all,
                                                const double* __restrict__ gmin_inst,
                                                const int32_t* __restrict__ kind,
                                                int32_t* __restrict__ fail,
                                                int32_t* __restrict__ fail_col, double* LU,
                                                const double* rhs, double* x, int max_nlu, int max_n,
                                                int ninst, SluSide side) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int ii = blockIdx.x;
    if (ii >= ninst) {
        // a side workgroup: truncation proposals that run beside the
        // instances' factor-and-solve chains on otherwise idle SIMDs
        trunc_lane(side.t, side.ctx, side.s0, side.s1, side.s2, side.s3, side.B, side.o, nullptr, side.tmin, 0,
                   side.tmin + side.ni, side.sel, (ii - ninst) * kW + lane);
        return;
    }
    const int kd = kind ? kind[ii] : SLU_REFACTOR;
    if (kd == SLU_SKIP) return;
    const SluInst d = insts[ii];
    if (d.plan < 0) {
        // no pivot order yet: nothing to factor or solve on the device
        if (lane == 0 && fail) {
            fail[ii] = FAIL_NOPLAN;
            if (fail_col) fail_col[ii] = 0;
        }
        return;
    }
    const int32_t* __restrict__ P = words + offs[d.plan];
    const int n = P[SH_N], nlu = P[SH_NLU];
    // LDS: the factor values, the vector right after them (the forward
    // substitution's tasks address entry i as value nlu + i), a zero (plan
    // records name it for absent terms), the reciprocal pivots (at least one
    // cell per lane), then the staged hot block on a 16-byte boundary
    double* v = lds;
    double* y = lds + nlu;
    char* L = reinterpret_cast<char*>(lds);
    double* dr = lds + max_nlu + max_n + 1;
    if (lane == 0) y[n] = 0.0;
    PlanView<STAGED> pv{P, nullptr, P[SH_HOT]};
    if (STAGED) {
        // the hot block, 16 bytes per lane per step (sections start on
        // 4-word boundaries, so the block is whole int4s)
        int32_t* hw = reinterpret_cast<int32_t*>(lds + ((max_nlu + max_n + 2 + max(max_n, kW)) & ~1));
        const int4* src = reinterpret_cast<const int4*>(P + pv.hot0);
        int4* dst = reinterpret_cast<int4*>(hw);
        batched<8>((P[SH_HOTLEN] + 3) / 4, lane, [&](int k) { return src[k]; },
                   [&](int k, int4 w) { dst[k] = w; });
        pv.hot = hw;
    }
    const int32_t* cptr = P + P[SH_CPTR];
    const int32_t* nU = P + P[SH_NU];
    const double gmin = gmin_inst ? gmin_inst[ii] : gmin_all;
    const bool tr = (mode & SLU_SOLVE_T) != 0;
    // a factor-and-solve launch folds the forward substitution into the factor
    const bool fused = (mode & SLU_FACTOR) && (mode & SLU_SOLVE) && !tr;
    const int32_t* in_perm = P + P[tr ? SH_CPERM : SH_RPERM];
    const int32_t* out_perm = P + P[tr ? SH_RPERM : SH_CPERM];
    if (mode & SLU_FACTOR) {
        for (int p = lane; p < nlu; p += kW) v[p] = 0.0;
        // the right-hand side rides along as column n (zeros: no solve asked)
        if (fused) {
            batched<4>(n, lane, [&](int k) { return in_perm[k]; },
                       [&](int k, int i) { y[k] = rhs[d.xoff + i]; });
        } else {
            for (int k = lane; k < n; k += kW) y[k] = 0.0;
        }
        __syncthreads();
        const int32_t* apos = P + P[SH_APOS];
        const int nnz = P[SH_NNZA];
        const double* a = A + d.aoff;
        batched<8>(nnz, lane, [&](int p) { return PosVal{apos[p], a[p]}; },
                   [&](int, PosVal e) { v[e.q] = e.w; });
        __syncthreads();
        if (gmin != 0.0) {
            const int32_t* gp = P + P[SH_GMIN];
            batched<4>(n, lane, [&](int k) { return gp[k]; }, [&](int, int q) {
                if (q >= 0) v[q] += gmin;
            });
            __syncthreads();
        }
        // right-looking steps, level by level; each level's first task is
        // fetched two levels ahead at a clamped index
        const int nlev = P[SH_NLEV];
        const int ntask = P[SH_NTASK];
        if (nlev > 0 && ntask > 0) {
            const int32_t* lev = pv.at(P[SH_LEV]);
            const int4* task = reinterpret_cast<const int4*>(pv.at(P[SH_TASK]));
            const int last = ntask - 1;
            int t0 = lev[0], t1 = lev[1], t2 = lev[2];
            int4 ta = task[min(t0 + lane, last)];
            int4 tb = task[min(t1 + lane, last)], tc;
            // results of the lanes without a task: a reciprocal pivot cell
            // of their own (one address for all would serialize the store),
            // written after the factor
            const int junk = (int)(reinterpret_cast<char*>(dr + lane) - L);
            int l = 0;
            for (; l + 2 < nlev; l += 3) {
                factor_level(lev, task, last, l, t0, t1, t2, ta, tb, tc, L, lane, junk);
                factor_level(lev, task, last, l + 1, t0, t1, t2, tb, tc, ta, L, lane, junk);
                factor_level(lev, task, last, l + 2, t0, t1, t2, tc, ta, tb, L, lane, junk);
            }
            if (l < nlev) factor_level(lev, task, last, l, t0, t1, t2, ta, tb, tc, L, lane, junk);
            if (l + 1 < nlev) factor_level(lev, task, last, l + 1, t0, t1, t2, tb, tc, ta, L, lane, junk);
        }
        // checks (the first failing column, as the reference stops there),
        // reciprocal pivots and the forward substitution's closing scale by
        // the stored 1/pivot (solve.cpp:24, zero entries left alone)
        int bad = 0x7fffffff, code = FAIL_NONE;
        batched<4>(n, lane, [&](int k) { return int2{cptr[k] + nU[k], cptr[k + 1]}; }, [&](int k, int2 c) {
            const double xkk = v[c.x];
            bool f;
            if (kd == SLU_VERIFY) {
                double lmax = 0.0;
                for (int p = c.x + 1; p < c.y; ++p) lmax = fmax(lmax, fabs(v[p]));
                f = !(lmax * 1e-3 < fabs(xkk));
            } else {
                f = (xkk == 0.0);
            }
            if (f && k < bad) bad = k;
            const double r = 1.0 / xkk;
            dr[k] = r;
            const double yk = y[k];
            if (yk != 0.0) y[k] = yk * r;
        });
        bad = wave_min_i32(bad);
        if (bad != 0x7fffffff) code = (kd == SLU_VERIFY) ? FAIL_REORDER : FAIL_ZERO_PIVOT;
        __syncthreads();
        // the stored U values and diagonal
        const int32_t* rows = P + P[SH_ROWS];
        const int32_t* pcol = P + P[SH_PCOL];
        batched<8>(nlu, lane, [&](int p) { return int2{rows[p], pcol[p]}; }, [&](int p, int2 rc) {
            if (rc.x < rc.y) v[p] = v[p] * dr[rc.x];
            else if (rc.x == rc.y) v[p] = dr[rc.x];
        });
        __syncthreads();
        double* lu = LU + d.luoff;
        for (int p = lane; p < nlu; p += kW) lu[p] = v[p];
        if (lane == 0 && fail) {
            fail[ii] = code;
            if (fail_col) fail_col[ii] = (code == FAIL_NONE) ? 0x7fffffff : bad;
        }
        if (code != FAIL_NONE) return;
    } else {
        const double* lu = LU + d.luoff;
        batched<8>(nlu, lane, [&](int p) { return lu[p]; }, [&](int p, double w) { v[p] = w; });
    }
    if (!(mode & (SLU_SOLVE | SLU_SOLVE_T))) return;
    if (fused) {
        // forward substitution done with the factor: the backward one
        run_program<OP_PLAIN, STAGED>(pv, 1, L, lane);
    } else {
        __syncthreads();
        batched<4>(n, lane, [&](int k) { return in_perm[k]; }, [&](int k, int i) { y[k] = rhs[d.xoff + i]; });
        __syncthreads();
        if (!tr) {
            run_program<OP_SKIPZERO | OP_MUL_NZ, false>(PlanView<false>{P, nullptr, 0}, 0, L, lane);
            run_program<OP_PLAIN, STAGED>(pv, 1, L, lane);
        } else {
            // the transposed programs are not in the hot block
            const PlanView<false> g{P, nullptr, 0};
            run_program<OP_SKIPZERO, false>(g, 2, L, lane);
            run_program<OP_MUL_DIAG, false>(g, 3, L, lane);
        }
    }
    batched<4>(n, lane, [&](int k) { return out_perm[k]; }, [&](int k, int i) { x[d.xoff + i] = y[k]; });
}

}  // namespace

void launch_small_lu(const SmallLuPlans& plans, const SluInst* d_inst, int ninst, int mode,
                     const double* A, double diag_gmin, const int32_t* kind, int32_t* fail,
                     int32_t* fail_col, double* LU, const double* rhs, double* x, hipStream_t s,
                     const double* gmin_inst, const SluSide* side) {
    if (ninst <= 0) return;
    SluSide sd{};
    int side_blocks = 0;
    if (side && side->t.ndev > 0) {
        sd = *side;
        side_blocks = (int)(((long)sd.t.ndev * sd.B + kW - 1) / kW);
    }
    const unsigned grid = (unsigned)(ninst + side_blocks);
    const int max_nlu = plans.max_nlu(), max_n = plans.max_n();
    // the hot block starts on a 16-byte boundary after the values, the
    // vector, the zero and the reciprocal pivots (at least one per lane)
    const size_t base =
        sizeof(double) * (((size_t)max_nlu + max_n + 2 + std::max(max_n, kW)) & ~(size_t)1);
    const size_t hot = sizeof(int32_t) * (((size_t)plans.max_hot() + 3) / 4 * 4);
    const int stage = (plans.opts().stage && base + hot <= kLdsLimit) ? 1 : 0;
    const size_t lds = base + (stage ? hot : 0);
    if (stage)
        hipLaunchKernelGGL(k_small_lu<true>, dim3(grid), dim3(kW), lds, s, plans.d_words(),
                           plans.d_offsets(), d_inst, mode, A, diag_gmin, gmin_inst, kind, fail, fail_col,
                           LU, rhs, x, max_nlu, max_n, ninst, sd);
    else
        hipLaunchKernelGGL(k_small_lu<false>, dim3(grid), dim3(kW), lds, s, plans.d_words(),
                           plans.d_offsets(), d_inst, mode, A, diag_gmin, gmin_inst, kind, fail, fail_col,
                           LU, rhs, x, max_nlu, max_n, ninst, sd);
    NSA_HIP(hipGetLastError());
}

}  // namespace nsa
