// Batched LU refactor and triangular solves (gfx950, FP64).  The phase
// structure and the layout are described in gpu_lu.hpp.
//
// Arithmetic per value follows the reference's own sequence:
//   Sparse 1.3 (SparseMatrix::factor, factor.cpp:1000-1047): for each U entry
//     r of column j in ascending row order, U(r,j) = x_r * (1/pivot_r) and
//     x_t -= U(r,j) * L(t,r); L kept unnormalised; diagonal = 1/x_j.
//   AMD-LU replay (refactor_replay_, amd_lu_solver.cpp:325-373): for each U
//     entry in the recorded reach order, U(r,j) = x_r and x_t -= L(t,r) * U(r,j);
//     pivot U(j,j) = x_j with the growth check; L(t,j) = x_t / U(j,j).
// Everything is compiled without FMA contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <utility>

#include "common.hpp"
#include "gpu_lu.hpp"

namespace nsa {

namespace {

constexpr int kWave = 64;
constexpr int kRfThreads = 1024;              // refactor workgroup: 16 waves
constexpr int kRfWaves = kRfThreads / kWave;
constexpr int kSvThreads = 512;               // solve workgroup
constexpr int kPf = 8;                        // U-entry steps whose operands are in flight

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Workgroups are dealt to the 8 XCDs round-robin; consecutive instances
// share the cache lines of A and b (instance index fastest), so they are
// kept on one XCD.
__device__ __forceinline__ int block_instance(int b, int B) {
    if ((B & 7) == 0) return (b & 7) * (B >> 3) + (b >> 3);
    return b;
}

__device__ __forceinline__ void flag(int32_t* fail, int32_t* fail_col, int ginst, int code, int col) {
    if (code != FAIL_NONE) {
        atomicMax(&fail[ginst], code);
        atomicMin(&fail_col[ginst], col);
    }
}

struct RfArgs {
    int n, T0, t, nlev, wsz, wregion, nws, union_bytes;
    const int32_t *lev_ptr, *lev_cols;
    const int32_t *cptr, *nU, *rows, *a_src, *gmin_pos, *desc_ptr;
    // per step: (U slot | chunk length << 16, first L position of the chunk,
    // diagonal position of r, offset of the chunk's relative slots)
    const int4* desc;
    const int16_t* rel;    // per L entry of a chunk: slot in the column, -1 = not here
    const int32_t *cand_ptr, *cand_pos, *diag_cand;
    const int32_t *tpos_ptr, *tpos, *twidx, *lrow_ptr, *ucol_ptr, *urow_ptr, *tcand_ptr, *tdiag_cand;
    const int16_t *lrow, *ucol, *urow, *tcand;
    const int32_t *dstep_ptr, *dtask_widx, *dtask_ptr, *dterm;
    const int32_t* sv_pos;
    int nsv;
    double* SV;
    const double* A;
    double* LU;
    int nlu, B;
    const int32_t* imap;
    int Bg;
    double gmin;
    int form;
    double pivot_tol;
    const int32_t* kind;
    int32_t *fail, *fail_col;
};

// One column by one wave.  partial: a trailing-block column, only the
// updates from pivots before the block; its values stay raw for the dense
// phase and its running |U| maximum goes to cmx.
__device__ __forceinline__ void rf_column(const RfArgs& a, int k, bool partial, double* ws,
                                          int4* wdesc, double* LU, int ginst, int kind, double* cmx) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool s13 = a.form == FORM_S13;
    const int cb = a.cptr[k];
    const int len = a.cptr[k + 1] - cb;
    const int gp = a.gmin_pos[k];
    // Column load: originals from A, fill zero, diag_gmin after the load
    // (NeoSolver::load_real + add_diag_gmin; GP scatter + x[col] += gmin).
    // Four chunks of loads are in flight at once.
    for (int s0 = 0; s0 < len; s0 += 4 * kWave) {
        int src[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) src[c] = a.a_src[cb + min(s0 + c * kWave + lane, len - 1)];
        double v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = a.A[(size_t)max(src[c], 0) * a.Bg + ginst];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int sl = s0 + c * kWave + lane;
            if (sl < len) {
                double x = src[c] >= 0 ? v[c] : 0.0;
                if (a.gmin != 0.0 && cb + sl == gp) x += a.gmin;
                ws[sl] = x;
            }
        }
    }
    wave_sync();

    const int d0 = a.desc_ptr[k];
    const int nd = a.desc_ptr[k + 1] - d0;
    double colmax = 0.0;
    // The column's step descriptors are staged in the wave's LDS; the
    // descriptor of step u + kPf is read together with step u's operands.
    for (int u = lane; u < nd; u += kWave) wdesc[u] = a.desc[d0 + u];
    wave_sync();
    // Ring of prefetched steps.  Per step: U slot and chunk length packed and
    // r's diagonal value; per lane: its two relative slots and its two L
    // values.  Every load is unconditional (indices clamped) and the main
    // loop is straight-line, so each wait covers exactly the step it feeds.
    int usnl[kPf];
    int16_t q0r[kPf], q1r[kPf];
    double l0[kPf], l1[kPf], dg[kPf];
    auto fetch = [&](const int4 d, int p) {
        usnl[p] = d.x;
        const int nl = d.x >> 16;
        const int last = nl > 0 ? nl - 1 : 0;
        const int e0 = min(lane, last), e1 = min(lane + kWave, last);
        q0r[p] = a.rel[d.w + e0];
        q1r[p] = a.rel[d.w + e1];
        l0[p] = LU[d.y + e0];
        l1[p] = LU[d.y + e1];
        dg[p] = LU[d.z];
    };
    auto step = [&](int p) {
        const int us = usnl[p] & 0xffff;
        const int nl = usnl[p] >> 16;
        const bool m0 = lane < nl && q0r[p] >= 0, m1 = lane + kWave < nl && q1r[p] >= 0;
        // the pivot row's slot is never a target of its own step, so the
        // three reads go out together
        const double xr = ws[us];
        const double t0 = ws[m0 ? q0r[p] : 0], t1 = ws[m1 ? q1r[p] : 0];
        const double uval = s13 ? xr * dg[p] : xr;
        LU[cb + us] = uval;   // every lane stores the same value
        colmax = fmax(colmax, fabs(uval));
        if (m0) ws[q0r[p]] = t0 - uval * l0[p];
        if (m1) ws[q1r[p]] = t1 - uval * l1[p];
        wave_sync();
    };
    if (nd > 0) {
#pragma unroll
        for (int p = 0; p < kPf; ++p) fetch(wdesc[min(p, nd - 1)], p);
        const int nfull = nd / kPf;
        for (int r = 0; r < nfull; ++r) {
#pragma unroll
            for (int p = 0; p < kPf; ++p) {
                const int4 dn = wdesc[min(r * kPf + p + kPf, nd - 1)];
                step(p);
                fetch(dn, p);
            }
        }
#pragma unroll
        for (int p = 0; p < kPf; ++p)
            if (nfull * kPf + p < nd) step(p);
    }

    if (partial) {
        for (int s = lane; s < len; s += kWave)
            if (a.rows[cb + s] >= a.T0) LU[cb + s] = ws[s];
        if (lane == 0) cmx[k - a.T0] = colmax;
        return;
    }
    const int dslot = a.nU[k];
    const double xd = ws[dslot];
    if (s13) {
        double lmax = 0.0;
        for (int s = dslot + 1 + lane; s < len; s += kWave) {
            const double v = ws[s];
            LU[cb + s] = v;
            lmax = fmax(lmax, fabs(v));
        }
        if (kind == 1) lmax = wave_max(lmax);
        if (lane == 0) {
            int code = FAIL_NONE;
            if (kind == 1) {
                // reuse path of order_and_factor (factor.cpp:929-946)
                if (!(lmax * 1e-3 < fabs(xd))) code = FAIL_REORDER;
            } else if (xd == 0.0) {
                code = FAIL_ZERO_PIVOT;
            }
            if (xd != 0.0) LU[cb + dslot] = 1.0 / xd;
            flag(a.fail, a.fail_col, ginst, code, k);
        }
    } else {
        if (lane == 0) {
            int code = FAIL_NONE;
            const double aukk = fabs(xd);
            double cm = colmax;
            if (aukk > cm) cm = aukk;
            if (kind == 0 && (xd == 0.0 || (cm > 0.0 && aukk < a.pivot_tol * cm))) code = FAIL_GROWTH;
            if (kind == 1) {
                // pivot choice of a fresh factor_ (amd_lu_solver.cpp:145-160)
                const int c0 = a.cand_ptr[k], c1 = a.cand_ptr[k + 1];
                int ipiv = -1;
                double amax = -1.0;
                for (int c = c0; c < c1; ++c) {
                    const double v = fabs(ws[a.cand_pos[c] - cb]);
                    if (v > amax) { amax = v; ipiv = c - c0; }
                }
                const int dc = a.diag_cand[k];
                if (ipiv < 0 || amax <= 0.0) {
                    code = FAIL_ZERO_PIVOT;
                } else {
                    if (dc >= 0 && fabs(ws[a.cand_pos[c0 + dc] - cb]) >= amax * a.pivot_tol) ipiv = dc;
                    if (a.cand_pos[c0 + ipiv] != cb + dslot) code = FAIL_REORDER;
                }
            }
            LU[cb + dslot] = xd;
            flag(a.fail, a.fail_col, ginst, code, k);
        }
        for (int s = dslot + 1 + lane; s < len; s += kWave) LU[cb + s] = ws[s] / xd;
    }
}

// The trailing block of one instance, right-looking in LDS by the whole
// workgroup: W[j * t + i] = entry (tail row i, tail column j).
__device__ __forceinline__ void rf_tail(const RfArgs& a, double* W, double* LU, int ginst, int kind,
                                        const double* cmx, double* sh_piv) {
    const int t = a.t;
    const int tid = threadIdx.x;
    const bool s13 = a.form == FORM_S13;
    for (int i = tid; i < t * t; i += kRfThreads) W[i] = 0.0;
    __syncthreads();
    const int npos = a.tpos_ptr[t];
    for (int q = tid; q < npos; q += kRfThreads) W[a.twidx[q]] = LU[a.tpos[q]];
    __syncthreads();
    int code_all = FAIL_NONE, fcol = 0x7fffffff;
    for (int s = 0; s < t; ++s) {
        double* colS = W + (size_t)s * t;
        if (tid < kWave) {
            const int lane = tid;
            const double piv = colS[s];
            int code = FAIL_NONE;
            if (!s13) {
                double cm = 0.0;
                for (int q = a.urow_ptr[s] + lane; q < a.urow_ptr[s + 1]; q += kWave)
                    cm = fmax(cm, fabs(colS[a.urow[q]]));
                cm = wave_max(cm);
                if (lane == 0) {
                    cm = fmax(cm, cmx[s]);
                    const double aukk = fabs(piv);
                    if (aukk > cm) cm = aukk;
                    if (kind == 0 && (piv == 0.0 || (cm > 0.0 && aukk < a.pivot_tol * cm)))
                        code = FAIL_GROWTH;
                    if (kind == 1) {
                        const int c0 = a.tcand_ptr[s], c1 = a.tcand_ptr[s + 1];
                        int ipiv = -1;
                        double amax = -1.0;
                        for (int c = c0; c < c1; ++c) {
                            const double v = fabs(colS[a.tcand[c]]);
                            if (v > amax) { amax = v; ipiv = c - c0; }
                        }
                        const int dc = a.tdiag_cand[s];
                        if (ipiv < 0 || amax <= 0.0) {
                            code = FAIL_ZERO_PIVOT;
                        } else {
                            if (dc >= 0 && fabs(colS[a.tcand[c0 + dc]]) >= amax * a.pivot_tol) ipiv = dc;
                            if (a.tcand[c0 + ipiv] != s) code = FAIL_REORDER;
                        }
                    }
                    *sh_piv = piv;
                }
            } else {
                double lm = 0.0;
                if (kind == 1) {
                    for (int q = a.lrow_ptr[s] + lane; q < a.lrow_ptr[s + 1]; q += kWave)
                        lm = fmax(lm, fabs(colS[a.lrow[q]]));
                    lm = wave_max(lm);
                }
                if (lane == 0) {
                    if (kind == 1) {
                        if (!(lm * 1e-3 < fabs(piv))) code = FAIL_REORDER;
                    } else if (piv == 0.0) {
                        code = FAIL_ZERO_PIVOT;
                    }
                    *sh_piv = (piv != 0.0) ? 1.0 / piv : piv;
                }
            }
            if (lane == 0 && code != FAIL_NONE) {
                if (code > code_all) code_all = code;
                if (a.T0 + s < fcol) fcol = a.T0 + s;
            }
        }
        __syncthreads();
        const double pv = *sh_piv;
        if (!s13) {
            for (int q = a.lrow_ptr[s] + tid; q < a.lrow_ptr[s + 1]; q += kRfThreads) {
                const int i = a.lrow[q];
                colS[i] = colS[i] / pv;
            }
        } else {
            for (int q = a.ucol_ptr[s] + tid; q < a.ucol_ptr[s + 1]; q += kRfThreads) {
                const int j = a.ucol[q];
                W[(size_t)j * t + s] = W[(size_t)j * t + s] * pv;
            }
            if (tid == 0) colS[s] = pv;
        }
        __syncthreads();
        const int l0 = a.lrow_ptr[s], nl = a.lrow_ptr[s + 1] - l0;
        const int u0 = a.ucol_ptr[s], nuc = a.ucol_ptr[s + 1] - u0;
        const int work = nl * nuc;
        for (int q = tid; q < work; q += kRfThreads) {
            const int i = a.lrow[l0 + q % nl];
            const int j = a.ucol[u0 + q / nl];
            W[(size_t)j * t + i] -= colS[i] * W[(size_t)j * t + s];
        }
        __syncthreads();
        // Updates from earlier (sparse) pivots that the reference's reach
        // order places after pivot s in a column: one task per target entry,
        // its terms in that order.
        const int k0 = a.dstep_ptr[s], k1 = a.dstep_ptr[s + 1];
        if (k1 > k0) {
            for (int q = k0 + tid; q < k1; q += kRfThreads) {
                double acc = W[a.dtask_widx[q]];
                for (int e = a.dtask_ptr[q]; e < a.dtask_ptr[q + 1]; ++e)
                    acc -= LU[a.dterm[2 * e]] * LU[a.dterm[2 * e + 1]];
                W[a.dtask_widx[q]] = acc;
            }
            __syncthreads();
        }
    }
    for (int q = tid; q < npos; q += kRfThreads) LU[a.tpos[q]] = W[a.twidx[q]];
    if (tid == 0) flag(a.fail, a.fail_col, ginst, code_all, fcol);
}

__global__ __launch_bounds__(kRfThreads) void k_refactor(RfArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double sh_piv;
    const int inst = block_instance(blockIdx.x, a.B);
    const int ginst = a.imap ? a.imap[inst] : inst;
    const int kind = a.kind ? a.kind[ginst] : 0;
    if (kind == 3) return;  // workgroup-uniform: instance not part of this run
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    double* LU = a.LU + (size_t)inst * a.nlu;
    double* cmx = reinterpret_cast<double*>(smem + a.union_bytes);
    double* ws = reinterpret_cast<double*>(smem + (size_t)w * a.wregion);
    int4* wdesc = reinterpret_cast<int4*>(smem + (size_t)w * a.wregion + (size_t)a.wsz * 8);
    for (int l = 0; l < a.nlev; ++l) {
        if (w < a.nws)
            for (int c = a.lev_ptr[l] + w; c < a.lev_ptr[l + 1]; c += a.nws)
                rf_column(a, a.lev_cols[c], false, ws, wdesc, LU, ginst, kind, cmx);
        __syncthreads();
    }
    if (a.t > 0) {
        if (w < a.nws)
            for (int j = w; j < a.t; j += a.nws) rf_column(a, a.T0 + j, true, ws, wdesc, LU, ginst, kind, cmx);
        __syncthreads();
        rf_tail(a, reinterpret_cast<double*>(smem), LU, ginst, kind, cmx, &sh_piv);
        __syncthreads();
    }
    double* SV = a.SV + (size_t)inst * a.nsv;
    for (int i = threadIdx.x; i < a.nsv; i += kRfThreads) SV[i] = LU[a.sv_pos[i]];
}

struct SolveArgs {
    int n;
    const int32_t *in_perm, *out_perm;
    DevProgram p1, p2;
    const double* LU;
    int nlu;
    const double* SV;
    int nsv;
    double* b;
    const int32_t* imap;
    int B, Bg;
};

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __hiloint2double(hi, lo);
}

// Operands of one thread's (or wave's) first row of a level, loaded one
// level ahead: the row record and its first terms' values and sources.
struct RowPrefetch {
    int4 ri;
    double v[8];
    int s[8];
};

// Runs one program level by level.  Each row's subtraction chain is
// sequential in term order.  Thread mode: one thread per row, term loads
// issued eight at a time.  Wave mode: one wave per row, the lanes form the
// products of 64 terms at once and the chain then consumes them in order.
// The next level's first-row operands are loaded while the current level
// runs (they depend on nothing the current level writes).
template <bool kSV>
__device__ void run_program(const DevProgram& p, double* y, const double* LU, const double* SV) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
    const bool skip = (p.op & OP_SKIPZERO) != 0;
    const bool has_diag = (p.op & (OP_MUL_DIAG | OP_DIV_DIAG | OP_MUL_NZ)) != 0;
    auto val = [&](int q) { return kSV ? SV[p.sv_base + q] : LU[p.pos[q]]; };
    auto finish = [&](double acc, int dix) {
        if (has_diag) {
            const double d = kSV ? SV[dix] : LU[dix];
            if (p.op & OP_MUL_DIAG) acc = acc * d;
            else if (p.op & OP_DIV_DIAG) acc = acc / d;
            else if (acc != 0.0) acc = acc * d;
        }
        return acc;
    };
    // Loads are clamped rather than skipped: term arrays carry padding.
    auto prefetch = [&](int l, RowPrefetch& f) {
        const int r0 = p.level_ptr[l], r1 = p.level_ptr[l + 1];
        const bool tmode = p.level_mode[l] == 0;
        f.ri = p.rows[min((tmode ? r0 + tid : r0 + w), r1 - 1)];
        const int last = max(f.ri.z - 1, 0);
        if (tmode) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int q = f.ri.y + min(j, last);
                f.v[j] = val(q);
                f.s[j] = p.src[q];
            }
        } else {
            const int q = f.ri.y + min(lane, last);
            f.v[0] = val(q);
            f.s[0] = p.src[q];
        }
    };
    if (p.nlevels == 0) return;
    RowPrefetch cur, nxt;
    prefetch(0, cur);
    for (int l = 0; l < p.nlevels; ++l) {
        if (l + 1 < p.nlevels) prefetch(l + 1, nxt);
        const int r0 = p.level_ptr[l], r1 = p.level_ptr[l + 1];
        if (p.level_mode[l] == 0) {
            for (int idx = r0 + tid; idx < r1; idx += kSvThreads) {
                const bool first = idx < r0 + kSvThreads;
                const int4 ri = first ? cur.ri : p.rows[idx];
                double acc = y[ri.x];
                for (int j0 = 0; j0 < ri.z; j0 += 8) {
                    double v[8];
                    int sx[8];
                    if (first && j0 == 0) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) { v[j] = cur.v[j]; sx[j] = cur.s[j]; }
                    } else {
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const int q = ri.y + min(j0 + j, ri.z - 1);
                            v[j] = val(q);
                            sx[j] = p.src[q];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (j0 + j < ri.z) {
                            const double ys = y[sx[j]];
                            if (skip) {
                                if (ys != 0.0) acc -= ys * v[j];
                            } else {
                                acc -= v[j] * ys;
                            }
                        }
                    }
                }
                y[ri.x] = finish(acc, ri.w);
            }
        } else {
            for (int idx = r0 + w; idx < r1; idx += kSvThreads / kWave) {
                const bool first = idx < r0 + kSvThreads / kWave;
                const int4 ri = first ? cur.ri : p.rows[idx];
                double acc = y[ri.x];
                for (int c = 0; c < ri.z; c += kWave) {
                    double v;
                    int sx;
                    if (first && c == 0) {
                        v = cur.v[0];
                        sx = cur.s[0];
                    } else {
                        const int q = ri.y + min(c + lane, ri.z - 1);
                        v = val(q);
                        sx = p.src[q];
                    }
                    const double ys = y[sx];
                    // a skipped term subtracts +0.0, which leaves acc unchanged
                    const double prod = skip ? (ys != 0.0 ? ys * v : 0.0) : v * ys;
                    const int m = min(kWave, ri.z - c);
                    for (int j = 0; j < m; ++j) acc -= readlane_f64(prod, j);
                }
                acc = finish(acc, ri.w);
                if (lane == 0) y[ri.x] = acc;
            }
        }
        __syncthreads();
        cur = nxt;
    }
}

// One workgroup per instance: the whole solve with the vector in LDS.
template <bool kSV>
__global__ __launch_bounds__(kSvThreads) void k_solve(SolveArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* y = reinterpret_cast<double*>(smem);
    const int inst = block_instance(blockIdx.x, a.B);
    const int ginst = a.imap ? a.imap[inst] : inst;
    const double* LU = a.LU + (size_t)inst * a.nlu;
    const double* SV = a.SV + (size_t)inst * a.nsv;
    for (int k = threadIdx.x; k < a.n; k += kSvThreads) y[k] = a.b[(size_t)a.in_perm[k] * a.Bg + ginst];
    __syncthreads();
    run_program<kSV>(a.p1, y, LU, SV);
    run_program<kSV>(a.p2, y, LU, SV);
    for (int k = threadIdx.x; k < a.n; k += kSvThreads) a.b[(size_t)a.out_perm[k] * a.Bg + ginst] = y[k];
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

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

int max_lds_per_block() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0;
        NSA_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        NSA_HIP(hipGetDeviceProperties(&prop, dev));
        cached = (int)prop.sharedMemPerBlock;
    }
    return cached;
}

}  // namespace

GpuLu::GpuLu(const LuStructure& s, int batch, hipStream_t stream,
             const std::vector<int32_t>* instances, int batch_global)
    : s_(s), B_(batch), stream_(stream) {
    if (batch <= 0) throw Error(NSA_ERR_ARGUMENT, "GpuLu: batch must be positive");
    const RefactorSchedule sched = build_refactor_schedule(s_);  // validates the fill closure
    const int n = s_.n;
    Bglobal_ = batch;
    if (instances) {
        if ((int)instances->size() != batch) throw Error(NSA_ERR_ARGUMENT, "GpuLu: instance map size");
        for (int32_t g : *instances)
            if (g < 0 || g >= batch_global) throw Error(NSA_ERR_ARGUMENT, "GpuLu: instance out of range");
        Bglobal_ = batch_global;
        d_imap_ = upload(allocs_, *instances);
    }
    const size_t lds_max = (size_t)max_lds_per_block() - 64;   // static LDS of the kernels
    if ((size_t)n * 8 > lds_max)
        throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: solve vector exceeds LDS (n too large)");
    int tmax = 128;
    while (tmax > 1 && (size_t)tmax * tmax * 8 + (size_t)tmax * 8 > lds_max) --tmax;
    // The dense phase applies updates from trailing pivots in ascending
    // pivot order; the trailing block is kept to sizes where every trailing
    // column lists its trailing U rows in that order (always for Sparse 1.3).
    auto ascending_tail = [&](int T0) {
        for (int k = T0; k < n; ++k) {
            int last = -1;
            for (int p = s_.cptr[k]; p < s_.cptr[k] + s_.nU[k]; ++p) {
                const int r = s_.rows[p];
                if (r < T0) continue;
                if (r < last) return false;
                last = r;
            }
        }
        return true;
    };
    tail_ = std::min(n, tmax);
    while (tail_ > 0 && !ascending_tail(n - tail_)) --tail_;
    T0_ = n - tail_;
    const int t = tail_;
    std::vector<int32_t> dpos(n), nl(n);
    int maxseg = 1;
    for (int k = 0; k < n; ++k) {
        dpos[k] = s_.dpos(k);
        nl[k] = s_.cptr[k + 1] - dpos[k] - 1;
        maxseg = std::max(maxseg, s_.cptr[k + 1] - s_.cptr[k]);
    }
    if (maxseg >= 32768) throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: column segment too long");
    wsz_ = (maxseg + 1) & ~1;
    // ---- sparse phase: levels of the columns before the trailing block
    std::vector<int32_t> lev(n, 0);
    int maxlev = -1;
    for (int k = 0; k < T0_; ++k) {
        int l = 0;
        for (int p = s_.cptr[k]; p < s_.cptr[k] + s_.nU[k]; ++p) l = std::max(l, lev[s_.rows[p]] + 1);
        lev[k] = l;
        maxlev = std::max(maxlev, l);
    }
    nlev_ = maxlev + 1;
    std::vector<int32_t> lev_ptr(nlev_ + 1, 0);
    for (int k = 0; k < T0_; ++k) lev_ptr[lev[k] + 1]++;
    for (int l = 0; l < nlev_; ++l) lev_ptr[l + 1] += lev_ptr[l];
    std::vector<int32_t> lev_cols(std::max(T0_, 1));
    {
        std::vector<int32_t> fill(lev_ptr.begin(), lev_ptr.end() - 1);
        for (int k = 0; k < T0_; ++k) lev_cols[fill[lev[k]]++] = k;
    }
    d_lev_ptr_ = upload(allocs_, lev_ptr);
    d_lev_cols_ = upload(allocs_, lev_cols);
    // Step descriptors and relative slots.  In a trailing column, a pivot r
    // before the block that the U list puts after a trailing pivot s is
    // "deferred": the sparse pass applies it to the rows before the block
    // only (rel = -1 for the block's rows), and the dense phase applies it
    // to the block's rows right after step s, in list order (dtask lists).
    std::vector<int32_t> desc_ptr(n + 1, 0);
    std::vector<int4> desc;
    std::vector<int16_t> rel;
    std::vector<std::vector<std::pair<int32_t, std::vector<int32_t>>>> step_tasks(std::max(t, 1));
    std::vector<int32_t> slot(n, -1);
    std::vector<int32_t> task_of(std::max(t, 1), -1);  // target row -> task in step_tasks[cur]
    for (int k = 0; k < n; ++k) {
        const int cb = s_.cptr[k];
        for (int p = cb; p < s_.cptr[k + 1]; ++p) slot[s_.rows[p]] = p - cb;
        int cur = -1;                        // last trailing pivot seen in this list
        for (int p = cb; p < cb + s_.nU[k]; ++p) {
            const int r = s_.rows[p];
            if (k >= T0_ && r >= T0_) {      // handled by the dense phase
                cur = r - T0_;
                std::fill(task_of.begin(), task_of.end(), -1);
                continue;
            }
            const bool deferred = k >= T0_ && cur >= 0;
            if (rel.size() + (size_t)nl[r] >= (size_t)INT_MAX)
                throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: plan too large");
            // L(:, r) in chunks of 128 (two entries per lane); a continuation
            // chunk recomputes the same U value (x_r is not among its targets)
            for (int c0 = 0; c0 < std::max(nl[r], 1); c0 += 2 * kWave) {
                const int cl = std::min(2 * kWave, nl[r] - c0);
                desc.push_back(make_int4((p - cb) | (std::max(cl, 0) << 16), dpos[r] + 1 + c0, dpos[r],
                                         (int)rel.size()));
                for (int e = dpos[r] + 1 + c0; e < dpos[r] + 1 + c0 + cl; ++e) {
                    const int i = s_.rows[e];
                    if (slot[i] < 0) throw Error(NSA_ERR_LOGIC, "GpuLu: fill pattern not closed");
                    rel.push_back((deferred && i >= T0_) ? (int16_t)-1 : (int16_t)slot[i]);
                }
            }
            if (!deferred) continue;
            auto& tasks = step_tasks[cur];
            for (int e = dpos[r] + 1; e < s_.cptr[r + 1]; ++e) {
                const int i = s_.rows[e];
                if (i < T0_) continue;
                int& ti = task_of[i - T0_];
                if (ti < 0) {
                    ti = (int)tasks.size();
                    tasks.push_back({(k - T0_) * t + (i - T0_), {}});
                }
                tasks[ti].second.push_back(e);
                tasks[ti].second.push_back(p);
            }
        }
        desc_ptr[k + 1] = (int32_t)desc.size();
        for (int p = cb; p < s_.cptr[k + 1]; ++p) slot[s_.rows[p]] = -1;
        if (k >= T0_) std::fill(task_of.begin(), task_of.end(), -1);
    }
    std::vector<int32_t> dstep_ptr(t + 1, 0), dtask_widx, dtask_ptr(1, 0), dterm;
    for (int s0 = 0; s0 < t; ++s0) {
        for (auto& tk : step_tasks[s0]) {
            dtask_widx.push_back(tk.first);
            dterm.insert(dterm.end(), tk.second.begin(), tk.second.end());
            dtask_ptr.push_back((int32_t)(dterm.size() / 2));
        }
        dstep_ptr[s0 + 1] = (int32_t)dtask_widx.size();
    }
    n_deferred_terms_ = (int64_t)dterm.size() / 2;
    d_dstep_ptr_ = upload(allocs_, dstep_ptr);
    d_dtask_widx_ = upload(allocs_, dtask_widx);
    d_dtask_ptr_ = upload(allocs_, dtask_ptr);
    d_dterm_ = upload(allocs_, dterm);
    d_cptr_ = upload(allocs_, s_.cptr);
    d_nU_ = upload(allocs_, s_.nU);
    d_rows_ = upload(allocs_, s_.rows);
    d_asrc_ = upload(allocs_, sched.a_src);
    d_gpos_ = upload(allocs_, s_.gmin_pos);
    d_desc_ptr_ = upload(allocs_, desc_ptr);
    int maxnd = 1;
    for (int k = 0; k < n; ++k) maxnd = std::max(maxnd, desc_ptr[k + 1] - desc_ptr[k]);
    // LDS: per-wave regions (column workspace + staged descriptors),
    // overlaid by the trailing block, then the trailing columns' |U| maxima.
    wregion_ = (int)align16((size_t)wsz_ * 8 + (size_t)maxnd * 16);
    const size_t cmx_bytes = (size_t)std::max(t, 1) * 8;
    nws_ = (int)std::min<size_t>(kRfWaves, (lds_max - cmx_bytes) / (size_t)wregion_);
    if (nws_ < 1) throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: column workspace exceeds LDS");
    const size_t union_bytes = align16(std::max((size_t)nws_ * wregion_, (size_t)t * t * 8));
    rf_lds_ = union_bytes + cmx_bytes;
    if (rf_lds_ > lds_max) throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: refactor LDS budget");
    desc.push_back(make_int4(0, 0, 0, 0));
    d_desc_ = upload(allocs_, desc);
    rel.push_back(-1);   // a clamped read of an empty chunk at the end stays in bounds
    d_rel_ = upload(allocs_, rel);
    std::vector<int32_t> cp = s_.cand_ptr, cpos = s_.cand_pos, dc = s_.diag_cand;
    if (cp.empty()) cp.assign(n + 1, 0);
    if (dc.empty()) dc.assign(n, -1);
    d_cand_ptr_ = upload(allocs_, cp);
    d_cand_pos_ = upload(allocs_, cpos);
    d_diag_cand_ = upload(allocs_, dc);
    d_row_perm_ = upload(allocs_, s_.row_perm);
    d_col_perm_ = upload(allocs_, s_.col_perm);
    // ---- dense trailing block lists (tail-local indices)
    std::vector<int32_t> tpos_ptr(t + 1, 0), tpos, twidx;
    std::vector<int32_t> lrow_ptr(t + 1, 0), urow_ptr(t + 1, 0), ucol_ptr(t + 1, 0);
    std::vector<int16_t> lrow, urow;
    std::vector<std::vector<int16_t>> ucol_lists(t);
    std::vector<int32_t> tcand_ptr(t + 1, 0), tdiag(std::max(t, 1), -1);
    std::vector<int16_t> tcand;
    for (int j = 0; j < t; ++j) {
        const int k = T0_ + j;
        for (int p = s_.cptr[k]; p < s_.cptr[k + 1]; ++p) {
            const int r = s_.rows[p];
            if (r < T0_) continue;
            const int i = r - T0_;
            tpos.push_back(p);
            twidx.push_back(j * t + i);
            if (i > j) lrow.push_back((int16_t)i);
            if (i < j) { urow.push_back((int16_t)i); ucol_lists[i].push_back((int16_t)j); }
        }
        tpos_ptr[j + 1] = (int32_t)tpos.size();
        lrow_ptr[j + 1] = (int32_t)lrow.size();
        urow_ptr[j + 1] = (int32_t)urow.size();
        if (s_.form == FORM_GP && !s_.cand_ptr.empty()) {
            for (int c = s_.cand_ptr[k]; c < s_.cand_ptr[k + 1]; ++c)
                tcand.push_back((int16_t)(s_.rows[s_.cand_pos[c]] - T0_));
            tdiag[j] = s_.diag_cand[k];
        }
        tcand_ptr[j + 1] = (int32_t)tcand.size();
    }
    std::vector<int16_t> ucol;
    for (int i = 0; i < t; ++i) {
        ucol.insert(ucol.end(), ucol_lists[i].begin(), ucol_lists[i].end());
        ucol_ptr[i + 1] = (int32_t)ucol.size();
    }
    d_tpos_ptr_ = upload(allocs_, tpos_ptr);
    d_tpos_ = upload(allocs_, tpos);
    d_twidx_ = upload(allocs_, twidx);
    d_lrow_ptr_ = upload(allocs_, lrow_ptr);
    d_lrow_ = upload(allocs_, lrow);
    d_urow_ptr_ = upload(allocs_, urow_ptr);
    d_urow_ = upload(allocs_, urow);
    d_ucol_ptr_ = upload(allocs_, ucol_ptr);
    d_ucol_ = upload(allocs_, ucol);
    d_tcand_ptr_ = upload(allocs_, tcand_ptr);
    d_tcand_ = upload(allocs_, tcand);
    d_tdiag_cand_ = upload(allocs_, tdiag);
    // ---- solve programs and the solve-ordered value copy of the plain solve
    TriProgram f, b, tf, tb;
    build_solve_programs(s_, f, b, tf, tb);
    std::vector<int32_t> sv_pos;
    upload_program(f, fwd_, &sv_pos);
    upload_program(b, bwd_, &sv_pos);
    upload_program(tf, tfwd_, nullptr);
    upload_program(tb, tbwd_, nullptr);
    nsv_ = (int)sv_pos.size();
    d_sv_pos_ = upload(allocs_, sv_pos);
    NSA_HIP(hipMalloc(&dSV_, std::max<size_t>(1, (size_t)nsv_ * B_) * sizeof(double)));
    allocs_.push_back(dSV_);
    // 256 doubles of padding past the last instance for clamped reads
    NSA_HIP(hipMalloc(&dLU_, ((size_t)s_.nlu() * B_ + 256) * sizeof(double)));
    allocs_.push_back(dLU_);
    NSA_HIP(hipMemsetAsync(dLU_, 0, (size_t)s_.nlu() * B_ * sizeof(double), stream_));
    NSA_HIP(hipMemsetAsync(dSV_, 0, (size_t)nsv_ * B_ * sizeof(double), stream_));
    if (rf_lds_ > 65536)
        NSA_HIP(hipFuncSetAttribute((const void*)k_refactor, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)rf_lds_));
    if ((size_t)n * 8 > 65536) {
        NSA_HIP(hipFuncSetAttribute((const void*)k_solve<true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, n * 8));
        NSA_HIP(hipFuncSetAttribute((const void*)k_solve<false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, n * 8));
    }
    NSA_HIP(hipStreamSynchronize(stream_));
}

GpuLu::~GpuLu() {
    for (void* p : allocs_) (void)hipFree(p);
}

void GpuLu::upload_program(const TriProgram& t, DevProgram& d, std::vector<int32_t>* sv_pos) {
    d.op = t.op;
    d.nlevels = t.nlevels();
    d.sv_base = sv_pos ? (int32_t)sv_pos->size() : 0;
    const bool has_diag = (t.op & (OP_MUL_DIAG | OP_DIV_DIAG | OP_MUL_NZ)) != 0;
    std::vector<int32_t> mode(std::max(d.nlevels, 1), 0);
    std::vector<int4> rows(std::max<size_t>(t.order.size(), 1));
    std::vector<int16_t> src;
    std::vector<int32_t> pos;
    for (int l = 0; l < d.nlevels; ++l) {
        int maxc = 0;
        const int r0 = t.level_ptr[l], r1 = t.level_ptr[l + 1];
        for (int idx = r0; idx < r1; ++idx) {
            const int tg = t.order[idx];
            const int q0 = t.term_ptr[tg], cnt = t.term_ptr[tg + 1] - q0;
            maxc = std::max(maxc, cnt);
            const int first = (int)pos.size();
            for (int q = q0; q < q0 + cnt; ++q) {
                src.push_back((int16_t)t.term_src[q]);
                pos.push_back(t.term_pos[q]);
            }
            rows[idx] = make_int4(tg, first, cnt, has_diag ? t.dpos[tg] : -1);
        }
        // latency model (cycles): a round of dependent loads ~2000, a chained
        // subtraction ~10 (thread) or ~12 (wave, with the lane read)
        const double R = r1 - r0;
        const double thread_cost = std::ceil(R / kSvThreads) * (std::ceil(maxc / 8.0) * 2000 + 10.0 * maxc);
        const double wave_cost = std::ceil(R / (kSvThreads / kWave)) * (std::ceil(maxc / 64.0) * 2000 + 12.0 * maxc);
        mode[l] = wave_cost < thread_cost ? 1 : 0;
    }
    if (sv_pos) {
        sv_pos->insert(sv_pos->end(), pos.begin(), pos.end());
        sv_pos->push_back(0);   // padding for clamped prefetches
        if (has_diag) {
            for (size_t idx = 0; idx < t.order.size(); ++idx) {
                const int tg = rows[idx].x;
                if (t.dpos[tg] < 0) throw Error(NSA_ERR_LOGIC, "GpuLu: solve program without diagonal");
                rows[idx].w = (int)sv_pos->size();
                sv_pos->push_back(t.dpos[tg]);
            }
        }
    }
    if (s_.n > 32767) throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: solve sources exceed int16");
    // padding: clamped prefetches of a row without terms read one past the end
    src.push_back(0);
    pos.push_back(0);
    d.level_ptr = upload(allocs_, t.level_ptr);
    d.level_mode = upload(allocs_, mode);
    d.rows = upload(allocs_, rows);
    d.src = upload(allocs_, src);
    d.pos = upload(allocs_, pos);
}

void GpuLu::refactor(const double* dA, double gmin, const int32_t* dkind, int32_t* dfail,
                     int32_t* dfail_col, hipStream_t on) {
    RfArgs a{};
    a.n = s_.n; a.T0 = T0_; a.t = tail_; a.nlev = nlev_; a.wsz = wsz_; a.wregion = wregion_; a.nws = nws_;
    a.union_bytes = (int)(rf_lds_ - (size_t)std::max(tail_, 1) * 8);
    a.lev_ptr = d_lev_ptr_; a.lev_cols = d_lev_cols_;
    a.cptr = d_cptr_; a.nU = d_nU_; a.rows = d_rows_; a.a_src = d_asrc_; a.gmin_pos = d_gpos_;
    a.desc_ptr = d_desc_ptr_; a.desc = d_desc_; a.rel = d_rel_;
    a.cand_ptr = d_cand_ptr_; a.cand_pos = d_cand_pos_; a.diag_cand = d_diag_cand_;
    a.tpos_ptr = d_tpos_ptr_; a.tpos = d_tpos_; a.twidx = d_twidx_;
    a.lrow_ptr = d_lrow_ptr_; a.ucol_ptr = d_ucol_ptr_; a.urow_ptr = d_urow_ptr_;
    a.tcand_ptr = d_tcand_ptr_; a.tdiag_cand = d_tdiag_cand_;
    a.lrow = d_lrow_; a.ucol = d_ucol_; a.urow = d_urow_; a.tcand = d_tcand_;
    a.dstep_ptr = d_dstep_ptr_; a.dtask_widx = d_dtask_widx_; a.dtask_ptr = d_dtask_ptr_;
    a.dterm = d_dterm_;
    a.sv_pos = d_sv_pos_; a.nsv = nsv_; a.SV = dSV_;
    a.A = dA; a.LU = dLU_; a.nlu = s_.nlu(); a.B = B_; a.imap = d_imap_; a.Bg = Bglobal_;
    a.gmin = gmin; a.form = (int)s_.form; a.pivot_tol = 1e-3; a.kind = dkind;
    a.fail = dfail; a.fail_col = dfail_col;
    hipLaunchKernelGGL(k_refactor, dim3(B_), dim3(kRfThreads), rf_lds_, on ? on : stream_, a);
    NSA_HIP(hipGetLastError());
}

void GpuLu::solve(double* db, bool transposed, hipStream_t on) {
    hipStream_t st = on ? on : stream_;
    SolveArgs a{};
    a.n = s_.n;
    // plain: y[k] = b[row_perm[k]]; L then U; x[col_perm[k]] = y[k]
    // transposed: y[k] = b[col_perm[k]]; U^T then L^T; x[row_perm[k]] = y[k]
    a.in_perm = transposed ? d_col_perm_ : d_row_perm_;
    a.out_perm = transposed ? d_row_perm_ : d_col_perm_;
    a.p1 = transposed ? tfwd_ : fwd_;
    a.p2 = transposed ? tbwd_ : bwd_;
    a.LU = dLU_;
    a.nlu = s_.nlu();
    a.SV = dSV_;
    a.nsv = nsv_;
    a.b = db;
    a.imap = d_imap_;
    a.B = B_;
    a.Bg = Bglobal_;
    if (transposed)
        hipLaunchKernelGGL(k_solve<false>, dim3(B_), dim3(kSvThreads), (size_t)s_.n * 8, st, a);
    else
        hipLaunchKernelGGL(k_solve<true>, dim3(B_), dim3(kSvThreads), (size_t)s_.n * 8, st, a);
    NSA_HIP(hipGetLastError());
}

}  // namespace nsa
