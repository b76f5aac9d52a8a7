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
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "small_lu.hpp"

namespace nsa {

namespace {

constexpr int kW = 64;

__device__ inline int wave_min_i32(int v) {
    for (int o = kW / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

// One triangular program: per level, lane-strided rows; each row is the
// reference's own subtraction chain for its target.
template <int OP>
__device__ inline void run_program(const int32_t* __restrict__ P, int prog, const double* v, double* y,
                                   int lane) {
    const int nl = P[SH_PROG + 3 * prog];
    const int32_t* lev = P + P[SH_PROG + 3 * prog + 1];
    const int32_t* rows = P + P[SH_PROG + 3 * prog + 2];
    for (int l = 0; l < nl; ++l) {
        const int r0 = lev[l], r1 = lev[l + 1];
        for (int r = r0 + lane; r < r1; r += kW) {
            const int4 rec = *reinterpret_cast<const int4*>(rows + 4 * r);
            const int32_t* tm = P + rec.y;
            double acc = y[rec.x];
            for (int q = 0; q < rec.z; ++q) {
                const int pos = tm[2 * q], src = tm[2 * q + 1];
                const double ys = y[src];
                if ((OP & OP_SKIPZERO) && ys == 0.0) continue;
                acc = acc - v[pos] * ys;
            }
            if (OP & OP_MUL_NZ) {
                if (acc != 0.0) acc = acc * v[rec.w];
            } else if (OP & OP_MUL_DIAG) {
                acc = acc * v[rec.w];
            }
            y[rec.x] = acc;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kW) void k_small_lu(const int32_t* __restrict__ words,
                                                const int32_t* __restrict__ offs,
                                                const SluInst* __restrict__ insts, int mode,
                                                const double* __restrict__ A, double gmin_all,
                                                const double* __restrict__ gmin_inst,
                                                const int32_t* __restrict__ kind,
                                                int32_t* __restrict__ fail,
                                                int32_t* __restrict__ fail_col, double* LU,
                                                const double* rhs, double* x, int max_nlu) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int ii = blockIdx.x;
    const int kd = kind ? kind[ii] : SLU_REFACTOR;
    if (kd == SLU_SKIP) return;
    const SluInst d = insts[ii];
    const int32_t* __restrict__ P = words + offs[d.plan];
    const int n = P[SH_N], nlu = P[SH_NLU];
    double* v = lds;            // nlu factor values
    double* y = lds + max_nlu;  // n: reciprocal pivots, then the solve vector
    const int32_t* cptr = P + P[SH_CPTR];
    const int32_t* nU = P + P[SH_NU];
    const double gmin = gmin_inst ? gmin_inst[ii] : gmin_all;
    if (mode & SLU_FACTOR) {
        for (int p = lane; p < nlu; p += kW) v[p] = 0.0;
        __syncthreads();
        const int32_t* apos = P + P[SH_APOS];
        const int nnz = P[SH_NNZA];
        const double* a = A + d.aoff;
        for (int p = lane; p < nnz; p += kW) v[apos[p]] = a[p];
        __syncthreads();
        if (gmin != 0.0) {
            const int32_t* gp = P + P[SH_GMIN];
            for (int k = lane; k < n; k += kW) {
                const int q = gp[k];
                if (q >= 0) v[q] += gmin;
            }
            __syncthreads();
        }
        // right-looking steps, level by level
        const int nlev = P[SH_NLEV];
        const int32_t* lev = P + P[SH_LEV];
        const int4* task = reinterpret_cast<const int4*>(P + P[SH_TASK]);
        int t_next = lev[0] + lane;
        int4 tk = (t_next < lev[nlev]) ? task[t_next] : make_int4(0, 0, 0, 0);
        for (int l = 0; l < nlev; ++l) {
            const int t1 = lev[l + 1];
            for (int t = lev[l] + lane; t < t1; t += kW) {
                const int4 cur = (t == t_next) ? tk : task[t];
                const double xkk = v[cur.w];
                const double xkj = v[cur.y];
                const double xik = v[cur.z];
                const double xt = v[cur.x];
                const double u = xkj * (1.0 / xkk);
                v[cur.x] = xt - u * xik;
            }
            // the first task of the next level, fetched while this level's
            // stores drain
            t_next = t1 + lane;
            if (t_next < lev[nlev]) tk = task[t_next];
            __syncthreads();
        }
        // checks (the first failing column, as the reference stops there)
        int bad = 0x7fffffff, code = FAIL_NONE;
        for (int k = lane; k < n; k += kW) {
            const int dp = cptr[k] + nU[k];
            const double xkk = v[dp];
            bool f;
            if (kd == SLU_VERIFY) {
                double lmax = 0.0;
                for (int p = dp + 1; p < cptr[k + 1]; ++p) lmax = fmax(lmax, fabs(v[p]));
                f = !(lmax * 1e-3 < fabs(xkk));
            } else {
                f = (xkk == 0.0);
            }
            if (f && k < bad) bad = k;
        }
        bad = wave_min_i32(bad);
        if (bad != 0x7fffffff) code = (kd == SLU_VERIFY) ? FAIL_REORDER : FAIL_ZERO_PIVOT;
        // reciprocal pivots, then the stored U values and diagonal
        for (int k = lane; k < n; k += kW) y[k] = 1.0 / v[cptr[k] + nU[k]];
        __syncthreads();
        const int32_t* rows = P + P[SH_ROWS];
        const int32_t* pcol = P + P[SH_PCOL];
        for (int p = lane; p < nlu; p += kW) {
            const int r = rows[p], c = pcol[p];
            if (r < c) v[p] = v[p] * y[r];
            else if (r == c) v[p] = y[r];
        }
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
        for (int p = lane; p < nlu; p += kW) v[p] = lu[p];
    }
    if (!(mode & (SLU_SOLVE | SLU_SOLVE_T))) return;
    const bool tr = (mode & SLU_SOLVE_T) != 0;
    const int32_t* in_perm = P + P[tr ? SH_CPERM : SH_RPERM];
    const int32_t* out_perm = P + P[tr ? SH_RPERM : SH_CPERM];
    __syncthreads();
    for (int k = lane; k < n; k += kW) y[k] = rhs[d.xoff + in_perm[k]];
    __syncthreads();
    if (!tr) {
        run_program<OP_SKIPZERO | OP_MUL_NZ>(P, 0, v, y, lane);
        run_program<OP_PLAIN>(P, 1, v, y, lane);
    } else {
        run_program<OP_SKIPZERO>(P, 2, v, y, lane);
        run_program<OP_MUL_DIAG>(P, 3, v, y, lane);
    }
    for (int k = lane; k < n; k += kW) x[d.xoff + out_perm[k]] = y[k];
}

}  // namespace

void launch_small_lu(const SmallLuPlans& plans, const SluInst* d_inst, int ninst, int mode,
                     const double* A, double diag_gmin, const int32_t* kind, int32_t* fail,
                     int32_t* fail_col, double* LU, const double* rhs, double* x, hipStream_t s,
                     const double* gmin_inst) {
    if (ninst <= 0) return;
    const int max_nlu = plans.max_nlu();
    const size_t lds = sizeof(double) * ((size_t)max_nlu + plans.max_n());
    hipLaunchKernelGGL(k_small_lu, dim3(ninst), dim3(kW), lds, s, plans.d_words(), plans.d_offsets(),
                       d_inst, mode, A, diag_gmin, gmin_inst, kind, fail, fail_col, LU, rhs, x, max_nlu);
    NSA_HIP(hipGetLastError());
}

}  // namespace nsa
