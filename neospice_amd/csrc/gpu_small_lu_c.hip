// Packed complex LU: one wave per (instance, frequency), the factor in LDS,
// on the plans of the real packed LU (small_lu.hpp): the pivot order and
// fill of a complex Sparse 1.3 factorization are a structure like any other.
//
// Refactor: SparseMatrix::factor_complex (src/solver/factor.cpp:1054-1141),
// applied right-looking with the same terms in the same order per entry:
// step k's reciprocal pivot R_k = 1 / x_kk (Smith's method, reciprocal),
// u = x_kj * R_k, and x_ij -= u * x_ik; an exact zero pivot (|re| + |im|
// == 0) fails at its column.  U(k, j) is stored as x_kj * R_k and the
// diagonal as R_k.  The forward substitution of solve_complex
// (solve.cpp:49-100) rides along as one more column of tasks (a step skips
// when its y_k is zero in both parts), the backward one runs the plan's
// row programs.  Every product and sum is formed as complex_ops forms it
// (cplx.hpp), so results are bit-identical to the reference.
//
// Values are (re, im) pairs of doubles: an LDS byte address of the real plan
// (8 p) is doubled (16 p).  The kernel reads the plan from global memory: a
// frequency sweep runs thousands of these waves at once, which hides the
// latency the real kernel's level-by-level tuning is there to cut.
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"
#include "cplx.hpp"
#include "small_lu.hpp"

namespace nsa {

namespace {

constexpr int kW = 64;

struct C2 {
    double r, i;
};

__device__ inline C2 cld(const char* L, int realb) { return *reinterpret_cast<const C2*>(L + 2 * realb); }
__device__ inline void cst(char* L, int realb, C2 v) { *reinterpret_cast<C2*>(L + 2 * realb) = v; }
__device__ inline bool cnz(C2 v) { return v.r != 0.0 || v.i != 0.0; }

// The end of a level: a workgroup is one wave, whose LDS accesses are
// performed in the order it issues them, so only the compiler must keep
// the accesses on their side.
__device__ inline void level_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int wave_min_i32(int v) {
    for (int o = kW / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

// One task of a level: x_t -= (x_kj * R_k) * x_ik; a forward-substitution
// task (pivot stored as -(address + 1)) is skipped when y_k is zero.
__device__ inline void run_task_c(const int4 tk, char* L) {
    const bool fwd = tk.w < 0;
    const int piv = fwd ? -tk.w - 1 : tk.w;
    const C2 xkj = cld(L, tk.y);
    if (fwd && !cnz(xkj)) return;
    const C2 xkk = cld(L, piv), xik = cld(L, tk.z);
    C2 xt = cld(L, tk.x);
    C2 R;
    crecip(R.r, R.i, xkk.r, xkk.i);
    C2 u = xkj;
    cmul(u.r, u.i, R.r, R.i);
    cmsub(xt.r, xt.i, u.r, u.i, xik.r, xik.i);
    cst(L, tk.x, xt);
}

// Row t of a triangular program: acc = y_t, then acc -= v * y_src per term
// in order (OP_SKIPZERO: a zero y_src skips its term), then the diagonal
// operation (OP_MUL_NZ: a nonzero acc times the stored reciprocal pivot;
// OP_MUL_DIAG: always).
template <int OP>
__device__ inline void run_row_c(const int32_t* P, const int32_t* rec, char* L) {
    C2 acc = cld(L, rec[0]);
    auto term = [&](int pos, int src) {
        const C2 b = cld(L, src);
        if ((OP & OP_SKIPZERO) && !cnz(b)) return;
        const C2 a = cld(L, pos);
        cmsub(acc.r, acc.i, a.r, a.i, b.r, b.i);
    };
    term(rec[4], rec[5]);
    term(rec[6], rec[7]);
    const int32_t* tm = P + rec[3];
    for (int q = 0; q < rec[1]; ++q) term(tm[2 * q], tm[2 * q + 1]);
    if (OP & (OP_MUL_NZ | OP_MUL_DIAG)) {
        if ((OP & OP_MUL_DIAG) || cnz(acc)) {
            const C2 d = cld(L, rec[2]);
            cmul(acc.r, acc.i, d.r, d.i);
        }
    }
    cst(L, rec[0], acc);
}

template <int OP>
__device__ inline void run_program_c(const int32_t* P, int prog, char* L, int lane) {
    const int nl = P[SH_PROG + 4 * prog];
    if (nl <= 0) return;
    const int32_t* lev = P + P[SH_PROG + 4 * prog + 1];
    const int32_t* rows = P + P[SH_PROG + 4 * prog + 2];
    // the level bounds one level ahead (the plan is read from global memory)
    int r0 = lev[0], r1 = lev[1];
    for (int l = 0; l < nl; ++l) {
        const int r2 = (l + 2 <= nl) ? lev[l + 2] : r1;
        for (int r = r0 + lane; r < r1; r += kW) run_row_c<OP>(P, rows + 8 * r, L);
        level_sync();
        r0 = r1;
        r1 = r2;
    }
}

__global__ __launch_bounds__(kW) void k_small_lu_c(const int32_t* __restrict__ words,
                                                  const int32_t* __restrict__ offs,
                                                  const SluInst* __restrict__ insts, int mode,
                                                  const double* __restrict__ ax, int32_t* __restrict__ fail,
                                                  int32_t* __restrict__ fail_col, double* LU, const double* rhs,
                                                  double* x, int max_nlu, int max_n) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const SluInst d = insts[blockIdx.x];
    if (d.plan < 0) {
        if (lane == 0 && fail) {
            fail[blockIdx.x] = FAIL_NOPLAN;
            if (fail_col) fail_col[blockIdx.x] = 0;
        }
        return;
    }
    const int32_t* __restrict__ P = words + offs[d.plan];
    const int n = P[SH_N], nlu = P[SH_NLU];
    // LDS as (re, im) pairs: the factor values, the vector after them, a
    // zero, then the reciprocal pivots
    C2* v = reinterpret_cast<C2*>(lds);
    C2* y = v + nlu;
    C2* dr = v + max_nlu + max_n + 1;
    char* L = reinterpret_cast<char*>(lds);
    if (lane == 0) y[n] = C2{0.0, 0.0};
    const bool fused = (mode & SLU_FACTOR) && (mode & SLU_SOLVE);
    const int32_t* in_perm = P + P[SH_RPERM];
    const int32_t* out_perm = P + P[SH_CPERM];
    const C2* vin = reinterpret_cast<const C2*>(rhs);
    if (mode & SLU_FACTOR) {
        for (int p = lane; p < nlu; p += kW) v[p] = C2{0.0, 0.0};
        for (int k = lane; k < n; k += kW) y[k] = fused ? vin[d.xoff + in_perm[k]] : C2{0.0, 0.0};
        __syncthreads();
        const int32_t* apos = P + P[SH_APOS];
        const C2* a = reinterpret_cast<const C2*>(ax) + d.aoff;
        for (int p = lane; p < P[SH_NNZA]; p += kW) v[apos[p]] = a[p];
        __syncthreads();
        const int nlev = P[SH_NLEV];
        const int32_t* lev = P + P[SH_LEV];
        const int4* task = reinterpret_cast<const int4*>(P + P[SH_TASK]);
        const int last = P[SH_NTASK] - 1;
        if (nlev > 0 && last >= 0) {
            // each level's first task and the next bounds fetched a level
            // ahead (the plan is read from global memory)
            int t0 = lev[0], t1 = lev[1];
            int4 cur = task[min(t0 + lane, last)];
            for (int l = 0; l < nlev; ++l) {
                const int t2 = (l + 2 <= nlev) ? lev[l + 2] : t1;
                const int4 nxt = task[min(t1 + lane, last)];
                if (t0 + lane < t1) run_task_c(cur, L);
                for (int t = t0 + lane + kW; t < t1; t += kW) run_task_c(task[t], L);
                level_sync();
                cur = nxt;
                t0 = t1;
                t1 = t2;
            }
        }
        const int32_t* cptr = P + P[SH_CPTR];
        const int32_t* nU = P + P[SH_NU];
        int bad = 0x7fffffff;
        for (int k = lane; k < n; k += kW) {
            const C2 xkk = v[cptr[k] + nU[k]];
            if (cmag(xkk.r, xkk.i) == 0.0 && k < bad) bad = k;
            C2 R;
            crecip(R.r, R.i, xkk.r, xkk.i);
            dr[k] = R;
            C2 yk = y[k];
            if (cnz(yk)) {
                cmul(yk.r, yk.i, R.r, R.i);
                y[k] = yk;
            }
        }
        bad = wave_min_i32(bad);
        const int code = (bad != 0x7fffffff) ? FAIL_ZERO_PIVOT : FAIL_NONE;
        __syncthreads();
        const int32_t* rows = P + P[SH_ROWS];
        const int32_t* pcol = P + P[SH_PCOL];
        for (int p = lane; p < nlu; p += kW) {
            const int r = rows[p], c = pcol[p];
            if (r < c) {
                C2 u = v[p];
                cmul(u.r, u.i, dr[r].r, dr[r].i);
                v[p] = u;
            } else if (r == c) {
                v[p] = dr[r];
            }
        }
        __syncthreads();
        if (LU) {
            C2* lu = reinterpret_cast<C2*>(LU) + d.luoff;
            for (int p = lane; p < nlu; p += kW) lu[p] = v[p];
        }
        if (lane == 0 && fail) {
            fail[blockIdx.x] = code;
            if (fail_col) fail_col[blockIdx.x] = (code == FAIL_NONE) ? 0x7fffffff : bad;
        }
        if (code != FAIL_NONE) return;
    } else {
        const C2* lu = reinterpret_cast<const C2*>(LU) + d.luoff;
        for (int p = lane; p < nlu; p += kW) v[p] = lu[p];
    }
    if (!(mode & SLU_SOLVE)) return;
    if (!fused) {
        __syncthreads();
        for (int k = lane; k < n; k += kW) y[k] = vin[d.xoff + in_perm[k]];
        __syncthreads();
        run_program_c<OP_SKIPZERO | OP_MUL_NZ>(P, 0, L, lane);
    }
    run_program_c<OP_PLAIN>(P, 1, L, lane);
    C2* xo = reinterpret_cast<C2*>(x);
    for (int k = lane; k < n; k += kW) xo[d.xoff + out_perm[k]] = y[k];
}

}  // namespace

size_t small_lu_complex_lds(const SmallLuPlans& plans) {
    return 2 * sizeof(double) * ((size_t)plans.max_nlu() + 2 * (size_t)plans.max_n() + 1);
}

void launch_small_lu_complex(const SmallLuPlans& plans, const SluInst* d_inst, int ninst, int mode,
                             const double* ax, int32_t* fail, int32_t* fail_col, double* LU, const double* rhs,
                             double* x, hipStream_t s) {
    if (ninst <= 0) return;
    const size_t lds = small_lu_complex_lds(plans);
    // a workgroup may take all of a CU's LDS (160 KB on gfx950); beyond 64
    // KB the kernel has to be told so first
    int dev = 0, limit = 0;
    NSA_HIP(hipGetDevice(&dev));
    NSA_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (lds > (size_t)limit)
        throw Error(NSA_ERR_UNSUPPORTED, "complex packed LU: the factor does not fit in LDS (" +
                                             std::to_string(lds) + " > " + std::to_string(limit) + " bytes)");
    if (lds > kSluComplexLds)
        NSA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small_lu_c),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_small_lu_c, dim3((unsigned)ninst), dim3(kW), lds, s, plans.d_words(), plans.d_offsets(),
                       d_inst, mode, ax, fail, fail_col, LU, rhs, x, plans.max_nlu(), plans.max_n());
    NSA_HIP(hipGetLastError());
}

}  // namespace nsa

namespace nsa {

namespace {

// ax of (instance i, frequency f), slot k: (g_ik, w_f c_ik) -- ac.cpp:200-204.
__global__ void k_ac_build(const double* __restrict__ g, const double* __restrict__ c, const double* __restrict__ w,
                           int B, int nf, int nnz, double* __restrict__ ax) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)B * nf * nnz;
    if (t >= total) return;
    const int k = (int)(t % nnz);
    const long fi = t / nnz;
    const int f = (int)(fi % nf), i = (int)(fi / nf);
    ax[2 * t] = g[(long)i * nnz + k];
    ax[2 * t + 1] = w[f] * c[(long)i * nnz + k];
}

// v of (instance i, frequency f) = rhs of instance i (complex values).
__global__ void k_ac_rhs(const double* __restrict__ rhs, int B, int nf, int n, double* __restrict__ v) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)B * nf * n;
    if (t >= total) return;
    const int j = (int)(t % n);
    const int i = (int)(t / n / nf);
    v[2 * t] = rhs[2 * ((long)i * n + j)];
    v[2 * t + 1] = rhs[2 * ((long)i * n + j) + 1];
}

}  // namespace

void launch_ac_build(const double* g, const double* c, const double* w, int B, int nf, int nnz, double* ax,
                     hipStream_t s) {
    const long total = (long)B * nf * nnz;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_ac_build, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, c, w, B, nf, nnz, ax);
    NSA_HIP(hipGetLastError());
}

void launch_ac_rhs(const double* rhs, int B, int nf, int n, double* v, hipStream_t s) {
    const long total = (long)B * nf * n;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_ac_rhs, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, rhs, B, nf, n, v);
    NSA_HIP(hipGetLastError());
}

}  // namespace nsa
