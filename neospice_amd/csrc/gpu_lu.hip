// Level-scheduled batched LU refactor and triangular solves (gfx950, FP64).
//
// Layout: every per-instance array is position-major with the instance
// index fastest (v[pos * B + inst]).  A block of the refactor kernel owns
// one column and a tile of IT instances; its 256 threads are IT instance
// lanes x TL target lanes, so the 16 instances of one position are one
// 128-byte line.  The column being factored is staged in LDS; a row ->
// slot map in LDS resolves the update targets (the fill pattern guarantees
// every row of L(:, r) is in the column).
//
// The arithmetic per target is the reference's own sequence: for each U
// entry r of the column, in the reference's processing order, x[t] -= u_r *
// L(t, r) with u_r = x_r * (1/pivot_r) (Sparse 1.3, factor.cpp:1010-1016)
// or u_r = x_r (AMD-LU replay, amd_lu_solver.cpp:342-352), then the
// diagonal/L finalisation of the form.  Code is built without FMA
// contraction, so results are bit-identical to the reference CPU path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "gpu_lu.hpp"

namespace nsa {

namespace {

constexpr int IT = 16;   // instances per refactor block
constexpr int TL = 16;   // target lanes per instance
constexpr int RF_THREADS = IT * TL;

__global__ __launch_bounds__(RF_THREADS) void k_refactor_level(
    const int32_t* __restrict__ cols, const int32_t* __restrict__ cptr,
    const int32_t* __restrict__ nU, const int32_t* __restrict__ rows,
    const int32_t* __restrict__ a_src, const int32_t* __restrict__ gmin_pos,
    const int32_t* __restrict__ cand_ptr, const int32_t* __restrict__ cand_pos,
    const int32_t* __restrict__ diag_cand, const double* __restrict__ A,
    double* __restrict__ LU, int B, const int32_t* __restrict__ imap, int Bglobal, int maxlen,
    double gmin, int form, double pivot_tol, const int32_t* __restrict__ verify_mask,
    int32_t* __restrict__ fail, int32_t* __restrict__ fail_col) {
    extern __shared__ double smem[];
    double* ws = smem;                                   // maxlen x IT
    double* red = smem + (size_t)maxlen * IT;            // TL x IT
    int16_t* map = reinterpret_cast<int16_t*>(red + RF_THREADS);

    const int k = cols[blockIdx.x];
    const int il = threadIdx.x % IT;
    const int tl = threadIdx.x / IT;
    const int inst = blockIdx.y * IT + il;          // lane within this plan's instances
    const bool active = inst < B;
    const int ginst = active ? (imap ? imap[inst] : inst) : 0;   // instance in the batch
    const size_t sB = (size_t)B;
    const size_t sG = (size_t)Bglobal;
    const int cb = cptr[k];
    const int len = cptr[k + 1] - cb;
    const int nu = nU[k];
    const int gp = gmin_pos[k];
    // Per-instance call kind: 0 refactorize (replay checks), 1 numeric whose
    // pivot order must be re-checked, 2 numeric on an order the host just
    // computed from these values (nothing to check).
    const bool verify = verify_mask != nullptr;   // block-uniform
    const int kind = (verify && active) ? verify_mask[ginst] : 0;
    const bool vinst = kind == 1;

    // Column load: originals from A, fill positions zero, diag_gmin added
    // after the load (NeoSolver::load_real + add_diag_gmin; GP scatter + add).
    for (int s = tl; s < len; s += TL) {
        const int src = a_src[cb + s];
        double v = 0.0;
        if (src >= 0 && active) v = A[(size_t)src * sG + ginst];
        if (gmin != 0.0 && cb + s == gp) v += gmin;
        ws[s * IT + il] = v;
        if (il == 0) map[rows[cb + s]] = (int16_t)s;
    }
    __syncthreads();

    double colmax = 0.0;
    for (int u = 0; u < nu; ++u) {
        const int r = rows[cb + u];
        const int rdp = cptr[r] + nU[r];
        const int re = cptr[r + 1];
        const double xr = ws[u * IT + il];
        double uval = xr;
        if (form == FORM_S13 && active) uval = xr * LU[(size_t)rdp * sB + inst];
        if (tl == 0 && active) {
            LU[(size_t)(cb + u) * sB + inst] = uval;
            const double a = fabs(uval);
            if (a > colmax) colmax = a;
        }
        for (int e = rdp + 1 + tl; e < re; e += TL) {
            const int t = map[rows[e]];
            const double l = active ? LU[(size_t)e * sB + inst] : 0.0;
            ws[t * IT + il] -= uval * l;
        }
        __syncthreads();
    }

    const int dslot = nu;
    const double xd = ws[dslot * IT + il];
    const size_t dpk = (size_t)(cb + dslot) * sB + inst;
    if (form == FORM_S13) {
        double lmax = 0.0;
        for (int s = dslot + 1 + tl; s < len; s += TL) {
            const double v = ws[s * IT + il];
            if (active) LU[(size_t)(cb + s) * sB + inst] = v;
            const double a = fabs(v);
            if (a > lmax) lmax = a;
        }
        if (verify) {
            red[tl * IT + il] = lmax;
            __syncthreads();
        }
        if (tl == 0 && active) {
            int32_t code = FAIL_NONE;
            if (vinst) {
                // Reuse path of order_and_factor (factor.cpp:929-946): keep
                // the pivot while largest_below * rel_threshold < |pivot|;
                // otherwise the CPU re-pivots from this step.
                double m = 0.0;
                for (int j = 0; j < TL; ++j) m = fmax(m, red[j * IT + il]);
                if (!(m * 1e-3 < fabs(xd))) code = FAIL_REORDER;
            } else if (xd == 0.0) {
                code = FAIL_ZERO_PIVOT;   // refactorize: exact zero pivot
            }
            if (xd != 0.0) LU[dpk] = 1.0 / xd;
            if (code != FAIL_NONE) {
                atomicMax(&fail[ginst], code);
                atomicMin(&fail_col[ginst], k);
            }
        }
    } else {
        const double ukk = xd;
        if (tl == 0 && active) {
            int32_t code = FAIL_NONE;
            const double aukk = fabs(ukk);
            if (aukk > colmax) colmax = aukk;
            if (kind == 0 && (ukk == 0.0 || (colmax > 0.0 && aukk < pivot_tol * colmax)))
                code = FAIL_GROWTH;   // replay rejected: the reference re-pivots
            if (vinst) {
                // Pivot choice of a fresh GP factorization
                // (amd_lu_solver.cpp:145-160): first max over candidates in
                // reach order, then the diagonal if within tolerance.
                const int c0 = cand_ptr[k], c1 = cand_ptr[k + 1];
                int ipiv = -1;
                double amax = -1.0;
                for (int c = c0; c < c1; ++c) {
                    const double t = fabs(ws[(cand_pos[c] - cb) * IT + il]);
                    if (t > amax) { amax = t; ipiv = c - c0; }
                }
                const int dc = diag_cand[k];
                if (ipiv < 0 || amax <= 0.0) code = FAIL_ZERO_PIVOT;
                else {
                    if (dc >= 0 && fabs(ws[(cand_pos[c0 + dc] - cb) * IT + il]) >= amax * pivot_tol)
                        ipiv = dc;
                    if (cand_pos[c0 + ipiv] != cb + dslot) code = FAIL_REORDER;
                }
            }
            LU[dpk] = ukk;
            if (code != FAIL_NONE) {
                atomicMax(&fail[ginst], code);
                atomicMin(&fail_col[ginst], k);
            }
        }
        for (int s = dslot + 1 + tl; s < len; s += TL) {
            if (active) LU[(size_t)(cb + s) * sB + inst] = ws[s * IT + il] / ukk;
        }
    }
}

__global__ void k_tri_level(const int32_t* __restrict__ order, int nrows,
                            const int32_t* __restrict__ term_ptr,
                            const int32_t* __restrict__ term_pos,
                            const int32_t* __restrict__ term_src,
                            const int32_t* __restrict__ dpos, int op,
                            const double* __restrict__ LU, double* __restrict__ y, int B) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ri = (int)(gid / B);
    const int inst = (int)(gid % B);
    if (ri >= nrows) return;
    const size_t sB = (size_t)B;
    const int t = order[ri];
    double acc = y[(size_t)t * sB + inst];
    const int p0 = term_ptr[t], p1 = term_ptr[t + 1];
    if (op & OP_SKIPZERO) {
        for (int p = p0; p < p1; ++p) {
            const double ys = y[(size_t)term_src[p] * sB + inst];
            if (ys != 0.0) acc -= ys * LU[(size_t)term_pos[p] * sB + inst];
        }
    } else {
        for (int p = p0; p < p1; ++p)
            acc -= LU[(size_t)term_pos[p] * sB + inst] * y[(size_t)term_src[p] * sB + inst];
    }
    if (op & OP_MUL_DIAG) acc = acc * LU[(size_t)dpos[t] * sB + inst];
    else if (op & OP_DIV_DIAG) acc = acc / LU[(size_t)dpos[t] * sB + inst];
    else if (op & OP_MUL_NZ) { if (acc != 0.0) acc = acc * LU[(size_t)dpos[t] * sB + inst]; }
    y[(size_t)t * sB + inst] = acc;
}

// Between the batch vector g (n x Bglobal) and this plan's pivot-space
// vector y (n x B, instances imap[0..B)):  y[k] = g[perm[k]] (gather) or
// g[perm[k]] = y[k] (scatter).
__global__ void k_permute(const int32_t* __restrict__ perm, const int32_t* __restrict__ imap,
                          double* __restrict__ g, double* __restrict__ y, int n, int B,
                          int Bglobal, int scatter) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)n * B) return;
    const int k = (int)(gid / B);
    const int inst = (int)(gid % B);
    const int ginst = imap ? imap[inst] : inst;
    const size_t gi = (size_t)perm[k] * Bglobal + ginst;
    const size_t yi = (size_t)k * B + inst;
    if (scatter) g[gi] = y[yi];
    else y[yi] = g[gi];
}

// LDS of one refactor block: the staged column (maxlen x IT doubles), the
// per-instance reduction row (TL x IT doubles) and the row -> slot map.
size_t lds_bytes(int maxlen, int n) {
    return (size_t)maxlen * IT * 8 + RF_THREADS * 8 + (((size_t)n * 2 + 7) & ~(size_t)7);
}

template <class T>
T* upload(std::vector<void*>& allocs, const std::vector<T>& v) {
    T* d = nullptr;
    size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    NSA_HIP(hipMalloc(&d, bytes));
    allocs.push_back(d);
    if (!v.empty()) NSA_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

}  // namespace

GpuLu::GpuLu(const LuStructure& s, int batch, hipStream_t stream,
             const std::vector<int32_t>* instances, int batch_global)
    : s_(s), sched_(build_refactor_schedule(s)), B_(batch), stream_(stream) {
    if (batch <= 0) throw Error(NSA_ERR_ARGUMENT, "GpuLu: batch must be positive");
    Bglobal_ = batch;
    if (instances) {
        if ((int)instances->size() != batch) throw Error(NSA_ERR_ARGUMENT, "GpuLu: instance map size");
        for (int32_t g : *instances)
            if (g < 0 || g >= batch_global) throw Error(NSA_ERR_ARGUMENT, "GpuLu: instance out of range");
        Bglobal_ = batch_global;
        d_imap_ = upload(allocs_, *instances);
    }
    for (int k = 0; k < s_.n; ++k) max_seg_ = std::max(max_seg_, s_.cptr[k + 1] - s_.cptr[k]);
    const size_t lds = lds_bytes(max_seg_, s_.n);
    if (lds > 64 * 1024 || max_seg_ >= 32768)
        throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: column segment too long for the LDS-staged refactor");
    d_cols_ = upload(allocs_, sched_.cols);
    d_cptr_ = upload(allocs_, s_.cptr);
    d_nU_ = upload(allocs_, s_.nU);
    d_rows_ = upload(allocs_, s_.rows);
    d_asrc_ = upload(allocs_, sched_.a_src);
    d_gpos_ = upload(allocs_, s_.gmin_pos);
    std::vector<int32_t> cp = s_.cand_ptr, cpos = s_.cand_pos, dc = s_.diag_cand;
    if (cp.empty()) cp.assign(s_.n + 1, 0);
    if (dc.empty()) dc.assign(s_.n, -1);
    d_cand_ptr_ = upload(allocs_, cp);
    d_cand_pos_ = upload(allocs_, cpos);
    d_diag_cand_ = upload(allocs_, dc);
    d_row_perm_ = upload(allocs_, s_.row_perm);
    d_col_perm_ = upload(allocs_, s_.col_perm);
    NSA_HIP(hipMalloc(&dLU_, std::max<size_t>(1, (size_t)s_.nlu() * B_) * sizeof(double)));
    allocs_.push_back(dLU_);
    NSA_HIP(hipMemsetAsync(dLU_, 0, (size_t)s_.nlu() * B_ * sizeof(double), stream_));
    NSA_HIP(hipMalloc(&dy_, std::max<size_t>(1, (size_t)s_.n * B_) * sizeof(double)));
    allocs_.push_back(dy_);
    TriProgram f, b, tf, tb;
    build_solve_programs(s_, f, b, tf, tb);
    upload_program(f, fwd_);
    upload_program(b, bwd_);
    upload_program(tf, tfwd_);
    upload_program(tb, tbwd_);
    NSA_HIP(hipStreamSynchronize(stream_));
}

GpuLu::~GpuLu() {
    for (void* p : allocs_) (void)hipFree(p);
}

void GpuLu::upload_program(const TriProgram& t, DevProgram& d) {
    d.op = t.op;
    d.level_ptr = t.level_ptr;
    d.order = upload(allocs_, t.order);
    d.term_ptr = upload(allocs_, t.term_ptr);
    d.term_pos = upload(allocs_, t.term_pos);
    d.term_src = upload(allocs_, t.term_src);
    d.dpos = upload(allocs_, t.dpos);
}

void GpuLu::refactor(const double* dA, double gmin, const int32_t* dverify, int32_t* dfail,
                     int32_t* dfail_col) {
    const int nl = sched_.nlevels();
    const dim3 block(RF_THREADS);
    for (int l = 0; l < nl; ++l) {
        const int c0 = sched_.level_ptr[l], c1 = sched_.level_ptr[l + 1];
        const int maxlen = sched_.level_maxlen[l];
        const size_t lds = lds_bytes(maxlen, s_.n);
        const dim3 grid(c1 - c0, (B_ + IT - 1) / IT);
        hipLaunchKernelGGL(k_refactor_level, grid, block, lds, stream_, d_cols_ + c0, d_cptr_,
                           d_nU_, d_rows_, d_asrc_, d_gpos_, d_cand_ptr_, d_cand_pos_,
                           d_diag_cand_, dA, dLU_, B_, d_imap_, Bglobal_, maxlen, gmin,
                           (int)s_.form, 1e-3, dverify, dfail, dfail_col);
    }
    NSA_HIP(hipGetLastError());
}

void GpuLu::run_program(const DevProgram& p, double* dy) {
    const int nl = (int)p.level_ptr.size() - 1;
    for (int l = 0; l < nl; ++l) {
        const int r0 = p.level_ptr[l], r1 = p.level_ptr[l + 1];
        const long work = (long)(r1 - r0) * B_;
        const int threads = 256;
        const int blocks = (int)((work + threads - 1) / threads);
        hipLaunchKernelGGL(k_tri_level, dim3(blocks), dim3(threads), 0, stream_, p.order + r0,
                           r1 - r0, p.term_ptr, p.term_pos, p.term_src, p.dpos, p.op, dLU_, dy,
                           B_);
    }
}

void GpuLu::solve(double* db, bool transposed) {
    const long total = (long)s_.n * B_;
    const int threads = 256;
    const int blocks = (int)((total + threads - 1) / threads);
    // Plain: y[k] = b[row_perm[k]]; L then U; x[col_perm[k]] = y[k].
    // Transposed: y[k] = b[col_perm[k]]; U^T then L^T; x[row_perm[k]] = y[k].
    const int32_t* in_perm = transposed ? d_col_perm_ : d_row_perm_;
    const int32_t* out_perm = transposed ? d_row_perm_ : d_col_perm_;
    hipLaunchKernelGGL(k_permute, dim3(blocks), dim3(threads), 0, stream_, in_perm, d_imap_, db,
                       dy_, s_.n, B_, Bglobal_, 0);
    if (!transposed) {
        run_program(fwd_, dy_);
        run_program(bwd_, dy_);
    } else {
        run_program(tfwd_, dy_);
        run_program(tbwd_, dy_);
    }
    hipLaunchKernelGGL(k_permute, dim3(blocks), dim3(threads), 0, stream_, out_perm, d_imap_, db,
                       dy_, s_.n, B_, Bglobal_, 1);
    NSA_HIP(hipGetLastError());
}

}  // namespace nsa
