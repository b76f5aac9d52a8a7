// Batched LU refactor and triangular solves (gfx950, FP64).  The phase
// structure and the layout are described in gpu_lu.hpp.
//
// Arithmetic per value follows the reference's own sequence:
//   Sparse 1.3 (SparseMatrix::factor, factor.cpp:1000-1047): for each U entry
//     r of column j in ascending row order, U(r,j) = x_r * (1/pivot_r) and
//     x_t -= U(r,j) * L(t,r); L kept unnormalised; diagonal = 1/x_j.
//   AMD-LU replay (refactor_replay_, amd_lu_solver.cpp:214-271): for each U
//     entry in the recorded reach order, U(r,j) = x_r and x_t -= L(t,r) * U(r,j);
//     pivot U(j,j) = x_j with the growth check; L(t,j) = x_t / U(j,j).
// Everything is compiled without FMA contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>
#include <utility>

#include "common.hpp"
#include "gpu_lu.hpp"

namespace nsa {

namespace {

constexpr int kWave = 64;
constexpr int kRfThreads = 1024;              // refactor workgroup: 16 waves
constexpr int kRfWaves = kRfThreads / kWave;
constexpr int kSvThreads = 768;               // solve workgroup
constexpr int kPf = 16;                       // U-entry steps whose operands are in flight
constexpr int kContinued = 1 << 30;           // step descriptor: continuation chunk of an L column
constexpr int kSpinMax = 1 << 24;              // bounded wait for a column (about 0.5 s)
constexpr int kChunk = kWave - 1;              // L entries per step (lane 63 carries the diagonal)
constexpr int kBufferFlags = 0x00020000;       // raw buffer resource, dword 3 (gfx9 family)
constexpr unsigned kOutOfRange = 0xfffffff0u;  // buffer offset past any buffer here: load returns 0

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kSweepMax = 4 * kWave;          // solve column sweep: up to 4 targets per lane
constexpr int kSweepPf = 8;                   // sweep steps whose operands are in flight
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __hiloint2double(hi, lo);
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
    const int32_t* lev_ptr;   // nlev + 1 into the column positions
    const int4* cmeta;        // per column position: (k, cb, len, nU), (d0, nd, 0, 0)
    const int32_t *cptr, *nU, *rows, *a_src, *gmin_pos, *desc_ptr;
    // per step: (U slot | chunk length << 16 | continuation flag, first L
    // position of the chunk, pivot column r, offset of the chunk's relative
    // slots)
    const int4* desc;
    const int16_t* rel;    // per L entry of a chunk: slot in the column, -1 = not here
    int nrel;              // entries of rel
    const int32_t *cand_ptr, *cand_pos, *diag_cand;
    const int32_t *tposmap, *tcand_ptr, *tdiag_cand;   // tposmap: 16 LU positions per thread (-1: none)
    const int16_t* tcand;
    const uint32_t* tmask;   // per trailing step: L rows, U columns, U rows (3 x 128 bits)
    // deferred updates: per trailing step a table index (-1: none); per table
    // and thread a task range; per task (register, first product, end)
    const int32_t *dstep_ptr, *dtask_ptr, *dterm;
    const int4* dtask;
    int ndterm;
    double* dprod;         // per instance: ndterm deferred products
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

// Completion bits of the columns (LDS, one bit per column).  A column's
// factor values are published by a workgroup-scope release fence followed
// by setting its bit; a reader that sees the bit set issues an acquire
// fence before loading them.
__device__ __forceinline__ bool col_done(uint32_t* done, int r) {
    return (__hip_atomic_load(&done[r >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> (r & 31)) & 1u;
}

__device__ __forceinline__ void col_publish(uint32_t* done, int k) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if ((threadIdx.x & (kWave - 1)) == 0)
        __hip_atomic_fetch_or(&done[k >> 5], 1u << (k & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One column by one wave.  partial: a trailing-block column, only the
// updates from pivots before the block; its values stay raw for the dense
// phase and its running |U| maximum goes to cmx.
//
// Columns run as soon as a wave takes them, not level by level: the step
// of pivot r reads L(:, r) only once column r's bit is set (columns are
// taken in a topological order, so every wait ends; a bounded spin flags
// FAIL_TIMEOUT if one did not).
__device__ __forceinline__ void rf_column(const RfArgs& a, const int4 m0, const int4 m1, bool partial,
                                          double* ws, int4* wdesc, double* LU, int ginst, int kind,
                                          double* cmx, uint32_t* done) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool s13 = a.form == FORM_S13;
    const int k = m0.x, cb = m0.y, len = m0.z;
    const __amdgpu_buffer_rsrc_t lu_rsrc = __builtin_amdgcn_make_buffer_rsrc(LU, 0, a.nlu * 8, kBufferFlags);
    const __amdgpu_buffer_rsrc_t rel_rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.rel, 0, a.nrel * 2, kBufferFlags);
    // Column load: the loaded values (A plus diag_gmin) already sit in the
    // column's own positions; four chunks in flight at once.
    for (int s0 = 0; s0 < len; s0 += 4 * kWave) {
        double v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = LU[cb + min(s0 + c * kWave + lane, len - 1)];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (s0 + c * kWave + lane < len) ws[s0 + c * kWave + lane] = v[c];
    }
    wave_sync();

    const int d0 = m1.x;
    const int nd = m1.y;
    // The column's step descriptors are staged in the wave's LDS; the
    // descriptor of step u + kPf is read together with step u's operands.
    for (int u = lane; u < nd; u += kWave) wdesc[u] = a.desc[d0 + u];
    wave_sync();
    // first step at or after u whose pivot column is not finished (nd: none)
    auto first_open = [&](int u) {
        for (int u0 = u; u0 < nd; u0 += kWave) {
            const bool open = u0 + lane < nd && !col_done(done, wdesc[u0 + lane].z);
            const unsigned long long b = __ballot(open);
            if (b) return u0 + __ffsll(b) - 1;
        }
        return nd;
    };
    auto await = [&](int r) {
        int spins = 0;
        while (!col_done(done, r)) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinMax) {
                if (lane == 0) flag(a.fail, a.fail_col, ginst, FAIL_TIMEOUT, k);
                break;
            }
        }
    };
    // Ring of prefetched steps, D deep.  Per step: U slot and chunk length
    // packed (wave-uniform); per lane: its relative slot and its L value.
    // Every load is unconditional (indices clamped to the segment) and the
    // ring loop is straight-line, so each wait covers exactly the step it
    // feeds.  U values stay in the column workspace (no later step of the
    // column reads or targets their slots) and are written out with the
    // column.  The address unit, which bounds this loop, spends about the
    // same on any vector load, so loads are few: rings of 8 or 16 start at
    // a step group and fetch a lane's relative slots for eight steps with
    // one load; shorter rings (short segments) load one slot per step.
    // The depth follows the segment's length, since a ring fetches up to D
    // steps past the segment's end.
    auto ring = [&](auto depth, const int u0, const int u1) {
        constexpr int D = decltype(depth)::value;
        constexpr bool kGrouped = D % 8 == 0;   // u0 is then a multiple of 8
        int usnl[D];
        int16_t qr[D];
        double lv[D];
        u32x4 gq[kGrouped ? D / 8 : 1];
        auto fetch = [&](int u, int p) {
            const int uc = min(u, u1 - 1);
            // the descriptor is wave-uniform: scalar copies keep the
            // offsets' selects below free of exec-mask branches
            const int4 dv = wdesc[uc];
            const int4 d = make_int4(__builtin_amdgcn_readfirstlane(dv.x), __builtin_amdgcn_readfirstlane(dv.y),
                                     0, __builtin_amdgcn_readfirstlane(dv.w));
            usnl[p] = d.x;
            const int nl = (d.x >> 16) & 0x7f;
            if constexpr (kGrouped) {
                // a group's slots are loaded at its last slot's fetch: the
                // steps of the group before it still read the previous ring
                // cycle's register (a clamped uc stays in the same group)
                if (p % 8 == 7)
                    gq[p / 8] = __builtin_amdgcn_raw_buffer_load_b128(rel_rsrc, (unsigned)(d.w * 512 + lane * 8) * 2u, 0, 0);
            } else {
                qr[p] = (int16_t)__builtin_amdgcn_raw_buffer_load_b16(
                    rel_rsrc, (unsigned)(d.w * 512 + lane * 8 + (uc & 7)) * 2u, 0, 0);
            }
            // Buffer loads: lanes past the chunk take an offset past the
            // buffer's end, which returns 0 without a memory access.
            // Chunks hold at most 63 entries, so lane 63 is always past the
            // chunk: for Sparse 1.3 it loads r's stored diagonal (first
            // chunk: just above L(:, r)) with the chunk.
            // (bitwise selects: a conditional expression here compiles to
            // exec-mask branches)
            const unsigned in_mask = 0u - (unsigned)(lane < nl);
            const unsigned dg_mask = 0u - (unsigned)(lane == kWave - 1 && s13);
            const unsigned past = ((unsigned)(d.y - 1) * 8u & dg_mask) | (kOutOfRange & ~dg_mask);
            const unsigned lo = ((unsigned)(d.y + lane) * 8u & in_mask) | (past & ~in_mask);
            const u32x2 lw = __builtin_amdgcn_raw_buffer_load_b64(lu_rsrc, lo, 0, 0);
            lv[p] = __hiloint2double((int)lw.y, (int)lw.x);
        };
        auto slot_of = [&](int p) -> int {
            if constexpr (kGrouped) {
                const u32x4 g = gq[p / 8];
                const int j = p % 8;
                const unsigned w = j < 2 ? g.x : j < 4 ? g.y : j < 6 ? g.z : g.w;
                return (int)(int16_t)(w >> (16 * (j & 1)));
            } else {
                return qr[p];
            }
        };
        auto step = [&](int p) {
            const int ud = __builtin_amdgcn_readfirstlane(usnl[p]);   // uniform: scalar branches below
            const int us = ud & 0xffff;
            const int nl = (ud >> 16) & 0x7f;
            const bool scale = s13 && !(ud & kContinued);   // a continuation finds U(r,j) in place
            const int q = slot_of(p);
            const bool m = lane < nl && q >= 0;
            // the pivot row's slot is never a target of its own step, so both
            // reads are independent; the empty asm takes both values, which
            // keeps the target read from sinking into the masked update below
            // (it would then wait for the pivot read before it is issued)
            const double xr = ws[us];
            const double tv = ws[m ? q : 0];
            asm volatile("" ::"v"(xr), "v"(tv));
            const double uval = scale ? xr * readlane_f64(lv[p], kWave - 1) : xr;
            if (m) ws[q] = tv - uval * lv[p];
            if (scale) {
                if (lane == 0) ws[us] = uval;
            }
            wave_sync();
        };
        const int cnt = u1 - u0;
        const int ng = (cnt + D - 1) / D;   // groups of D steps
#pragma unroll
        for (int p = 0; p < D; ++p) fetch(u0 + p, p);
        for (int g = 0; g + 1 < ng; ++g) {
#pragma unroll
            for (int p = 0; p < D; ++p) {
                step(p);
                fetch(u0 + (g + 1) * D + p, p);
            }
        }
        const int base = (ng - 1) * D;
#pragma unroll
        for (int p = 0; p < D; ++p)
            if (base + p < cnt) step(p);
        // the ring's last fetches (clamped, unused) are drained here, so the
        // next segment starts with no loads outstanding
        __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0): vmcnt bits 3:0 and 15:14 zero, expcnt 7, lgkmcnt 15
    };
    // steps [s, e) through the ring its length calls for (grouped rings
    // only from a group boundary)
    auto run = [&](int s, int e) {
        const int cnt = e - s;
        if (cnt >= 2 * kPf && (s & 7) == 0) ring(std::integral_constant<int, kPf>{}, s, e);
        else if (cnt >= 12 && (s & 7) == 0) ring(std::integral_constant<int, 8>{}, s, e);
        else if (cnt >= 4) ring(std::integral_constant<int, 4>{}, s, e);
        else if (cnt >= 2) ring(std::integral_constant<int, 2>{}, s, e);
        else ring(std::integral_constant<int, 1>{}, s, e);
    };
    // Segments: the steps from u0 whose pivot columns are finished run
    // through the rings; at a step whose pivot is still open the wave waits
    // for that one column and scans on.  A dependency thus never holds
    // ready steps behind it.  A long segment that starts inside a step
    // group first runs up to the group's end.
    int u0 = 0;
    while (u0 < nd) {
        int u1 = first_open(u0);
        if (u1 == u0) {
            await(wdesc[u0].z);
            u1 = first_open(u0 + 1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        int s = u0;
        if (u1 - s >= 20 && (s & 7) != 0) {
            run(s, (s | 7) + 1);
            s = (s | 7) + 1;
        }
        run(s, u1);
        u0 = u1;
    }
    // U part of the column: write out, and its largest magnitude
    const int nu = m0.w;
    double colmax = 0.0;
    for (int u = lane; u < nu; u += kWave) {
        const double v = ws[u];
        if (!partial || a.rows[cb + u] < a.T0) {
            LU[cb + u] = v;
            colmax = fmax(colmax, fabs(v));
        }
    }
    if (!s13) colmax = wave_max(colmax);

    if (partial) {
        for (int s = lane; s < len; s += kWave)
            if (a.rows[cb + s] >= a.T0) LU[cb + s] = ws[s];
        if (lane == 0) cmx[k - a.T0] = colmax;
        return;
    }
    const int dslot = m0.w;
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

__host__ __device__ constexpr size_t align16_d(size_t x) { return (x + 15) & ~(size_t)15; }

__device__ __forceinline__ bool bit128(const uint32_t* m, int i) { return (m[i >> 5] >> (i & 31)) & 1u; }

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
    return v;
}

// The trailing block of one instance (t <= 128), right-looking, held in
// registers: thread (ti, tj) = (tid % 32, tid / 32) owns entries
// (ti + 32a, tj + 32b), a, b < 4.  Per step s the wave holding column s
// checks the pivot and publishes the column (L, scaled for AMD-LU), the
// owners of row s publish the row, and after one barrier every thread
// applies the rank-1 update to its entries -- exactly where L(i,s) and
// U(s,j) are both in the pattern (128-bit masks per step), as in the
// sparse reference.  Deferred sparse updates of a step are applied by the
// owner of their target entry.
__device__ __forceinline__ void rf_tail(const RfArgs& a, double* lds, double* LU, int ginst, int kind,
                                        double* cmx, uint32_t* msk, const double* dprod) {
    const int t = a.t;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int ti = tid & 31, tj = tid >> 5;
    const int wv = tid / kWave;
    const bool s13 = a.form == FORM_S13;
    double* Lcol = lds;          // 2 x 128 (double-buffered by step parity)
    double* Urow = lds + 256;    // 2 x 128
    double* shpv = lds + 512;    // 2
    double* colbuf = lds + 516;  // 128: raw pivot column for the candidate check
    for (int i = tid; i < t * 12; i += kRfThreads) msk[i] = a.tmask[i];
    double R[4][4];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int pos = a.tposmap[tid * 16 + q];
        R[q >> 2][q & 3] = pos >= 0 ? LU[pos] : 0.0;
    }
    __syncthreads();
    int dt_next = a.dstep_ptr[0];   // deferred-task table of the next step, loaded a step ahead
    // One pivot step.  The register column of pivot s (bs = s / 32) is a
    // constant of each instantiation, so column s's registers are read and
    // written directly instead of through four-way selects.
    auto step = [&](auto bs_tag, const int s) {
        constexpr int bs = decltype(bs_tag)::value;
        const int buf = s & 1;
        const int ts = s & 31;                 // column s: owners tj == ts, register column bs
        const uint32_t* lm = msk + 12 * s;     // L rows of column s
        const uint32_t* um = lm + 4;           // U columns of row s
        const int dt = dt_next;
        dt_next = a.dstep_ptr[min(s + 1, t - 1)];
        if (wv == (ts >> 1)) {
            // the wave holding column s (its half tj == ts owns it)
            const bool own = tj == ts;
            const int src = ts + 32 * (ts & 1);    // lane owning (s, s)
            double cv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) cv[q] = R[q][bs];
            const double mine = cv[bs];
            const double piv = readlane_f64(mine, src);
            int code = FAIL_NONE;
            if (!s13) {
                double cm = fmax(cmx[s], 0.0);
                const double aukk = fabs(piv);
                if (aukk > cm) cm = aukk;
                if (kind == 0 && (piv == 0.0 || (cm > 0.0 && aukk < a.pivot_tol * cm))) code = FAIL_GROWTH;
                if (kind == 1) {
                    // first maximum over the candidates in reach order, then
                    // the diagonal within tolerance (amd_lu_solver.cpp:145-160)
                    if (own)
#pragma unroll
                        for (int q = 0; q < 4; ++q) colbuf[ti + 32 * q] = cv[q];
                    wave_sync();
                    const int c0 = a.tcand_ptr[s], nc = a.tcand_ptr[s + 1] - c0;
                    double v0 = -1.0, v1 = -1.0;
                    if (lane < nc) v0 = fabs(colbuf[a.tcand[c0 + lane]]);
                    if (lane + kWave < nc) v1 = fabs(colbuf[a.tcand[c0 + lane + kWave]]);
                    const double amax = wave_max(fmax(v0, v1));
                    int c = 0x7fffffff;
                    if (lane + kWave < nc && v1 == amax) c = lane + kWave;
                    if (lane < nc && v0 == amax) c = lane;
                    int ipiv = wave_min_i(c);
                    const int dc = a.tdiag_cand[s];
                    if (nc == 0 || ipiv == 0x7fffffff || amax <= 0.0) {
                        code = FAIL_ZERO_PIVOT;
                    } else {
                        if (dc >= 0 && fabs(colbuf[a.tcand[c0 + dc]]) >= amax * a.pivot_tol) ipiv = dc;
                        if (a.tcand[c0 + ipiv] != s) code = FAIL_REORDER;
                    }
                }
                // L(i, s) = x_i / U(s, s); the four quotients are formed
                // unconditionally (independent, so they overlap) and kept
                // where the pattern has the entry
                if (own) {
                    const uint4 lw = *reinterpret_cast<const uint4*>(lm);
                    const uint32_t lwa[4] = {lw.x, lw.y, lw.z, lw.w};
                    double dv[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) dv[q] = cv[q] / piv;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int i = ti + 32 * q;
                        if (i > s && i < t && ((lwa[q] >> ti) & 1u)) {
                            cv[q] = dv[q];
                            Lcol[buf * 128 + i] = dv[q];
                        }
                    }
                }
            } else {
                if (kind == 1) {
                    double m = 0.0;
                    if (own)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int i = ti + 32 * q;
                            if (i > s && i < t && bit128(lm, i)) m = fmax(m, fabs(cv[q]));
                        }
                    m = wave_max(m);
                    if (!(m * 1e-3 < fabs(piv))) code = FAIL_REORDER;
                } else if (piv == 0.0) {
                    code = FAIL_ZERO_PIVOT;
                }
                const double pv = (piv != 0.0) ? 1.0 / piv : piv;   // the diagonal stores 1 / pivot
                if (own) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int i = ti + 32 * q;
                        if (i > s && i < t && bit128(lm, i)) Lcol[buf * 128 + i] = cv[q];
                        if (i == s) cv[q] = pv;
                    }
                }
                if (lane == src) shpv[buf] = pv;
            }
            if (own) {
#pragma unroll
                for (int q = 0; q < 4; ++q) R[q][bs] = cv[q];
            }
            if (lane == src && code != FAIL_NONE) flag(a.fail, a.fail_col, ginst, code, a.T0 + s);
        }
        if (ti == ts) {
            // owners of row s: publish U(s, j) (raw) and fold it into the
            // columns' running |U| maxima (the maxima are read first, all
            // four at once; each owner has its own columns j)
            const uint4 uw = *reinterpret_cast<const uint4*>(um);
            const uint32_t uwa[4] = {uw.x, uw.y, uw.z, uw.w};
            double cm[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) cm[b] = s13 ? 0.0 : cmx[min(tj + 32 * b, t - 1)];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = tj + 32 * b;
                const double raw = R[bs][b];
                if (j > s && j < t && ((uwa[b] >> tj) & 1u)) {
                    Urow[buf * 128 + j] = raw;
                    if (!s13) cmx[j] = fmax(cm[b], fabs(raw));
                }
            }
        }
        __syncthreads();
        const double pv = s13 ? shpv[buf] : 1.0;
        double lv[4], uv[4];
        bool lok[4], uok[4];
        const uint4 lw = *reinterpret_cast<const uint4*>(lm);
        const uint4 uw = *reinterpret_cast<const uint4*>(um);
        const uint32_t lwa[4] = {lw.x, lw.y, lw.z, lw.w}, uwa[4] = {uw.x, uw.y, uw.z, uw.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = ti + 32 * q, j = tj + 32 * q;
            lok[q] = i > s && i < t && ((lwa[q] >> ti) & 1u);
            uok[q] = j > s && j < t && ((uwa[q] >> tj) & 1u);
            // unused operands (absent entries) are read but never applied
            lv[q] = Lcol[buf * 128 + i];
            const double ur = Urow[buf * 128 + j];
            // U(s, j) = x_s(j) * (1 / pivot) for Sparse 1.3, x_s(j) for AMD-LU
            uv[q] = s13 ? ur * pv : ur;
        }
        if (s13 && ti == ts) {
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (uok[b]) R[bs][b] = uv[b];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (lok[q] && uok[b]) R[q][b] = R[q][b] - lv[q] * uv[b];
        // Updates from earlier (sparse) pivots that the reference's reach
        // order places after pivot s in a column, by the target's owner, in
        // that order; the products were formed before the dense phase.
        if (dt >= 0) {
            const int k0 = a.dtask_ptr[dt * (kRfThreads + 1) + tid];
            const int k1 = a.dtask_ptr[dt * (kRfThreads + 1) + tid + 1];
            for (int k = k0; k < k1; ++k) {
                const int4 tk = a.dtask[k];   // (register, first product, end, -)
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    if (tk.x == q) acc = R[q >> 2][q & 3];
                for (int e = tk.y; e < tk.z; e += 16) {
                    double pr[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) pr[u] = dprod[min(e + u, tk.z - 1)];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        if (e + u < tk.z) acc -= pr[u];
                }
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    if (tk.x == q) R[q >> 2][q & 3] = acc;
            }
        }
    };
    for (int s = 0; s < min(t, 32); ++s) step(std::integral_constant<int, 0>{}, s);
    for (int s = 32; s < min(t, 64); ++s) step(std::integral_constant<int, 1>{}, s);
    for (int s = 64; s < min(t, 96); ++s) step(std::integral_constant<int, 2>{}, s);
    for (int s = 96; s < t; ++s) step(std::integral_constant<int, 3>{}, s);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int pos = a.tposmap[tid * 16 + q];
        if (pos >= 0) LU[pos] = R[q >> 2][q & 3];
    }
}

__global__ __launch_bounds__(kRfThreads) void k_refactor(RfArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int inst = block_instance(blockIdx.x, a.B);
    const int ginst = a.imap ? a.imap[inst] : inst;
    const int kind = a.kind ? a.kind[ginst] : 0;
    if (kind == 3) return;  // workgroup-uniform: instance not part of this run
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    double* LU = a.LU + (size_t)inst * a.nlu;
    double* cmx = reinterpret_cast<double*>(smem + a.union_bytes);
    // after the union, the |U| maxima and the trailing masks: the column
    // completion bits, then the counter that deals out the columns
    uint32_t* done = reinterpret_cast<uint32_t*>(smem + a.union_bytes + align16_d((size_t)a.t * 8) + (size_t)a.t * 48);
    const int nwords = (a.n + 31) / 32;
    int* next_col = reinterpret_cast<int*>(done + nwords);
    double* ws = reinterpret_cast<double*>(smem + (size_t)w * a.wregion);
    int4* wdesc = reinterpret_cast<int4*>(smem + (size_t)w * a.wregion + (size_t)a.wsz * 8);
    // Load phase (NeoSolver::load_real + add_diag_gmin; GP scatter + gmin):
    // every factor position takes its A value (fill: zero), eight loads in
    // flight per thread.
    for (int p0 = threadIdx.x; p0 < a.nlu; p0 += 8 * kRfThreads) {
        int src[8];
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) src[u] = a.a_src[min(p0 + u * kRfThreads, a.nlu - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = a.A[(size_t)max(src[u], 0) * a.Bg + ginst];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (p0 + u * kRfThreads < a.nlu) LU[p0 + u * kRfThreads] = src[u] >= 0 ? v[u] : 0.0;
    }
    for (int i = threadIdx.x; i < nwords; i += kRfThreads) done[i] = 0u;
    if (threadIdx.x == 0) *next_col = 0;
    __syncthreads();
    if (a.gmin != 0.0) {
        for (int k = threadIdx.x; k < a.n; k += kRfThreads)
            if (a.gmin_pos[k] >= 0) LU[a.gmin_pos[k]] += a.gmin;
        __syncthreads();
    }
    // Sparse columns, then the trailing columns' partial pass.  Column
    // positions: the sparse columns in level order, then the trailing ones
    // (a topological order); each wave takes the next position from a
    // counter and loads its metadata while it runs the current one.
    auto meta = [&](int c, int4& x0, int4& x1) { x0 = a.cmeta[2 * c]; x1 = a.cmeta[2 * c + 1]; };
    auto take = [&]() {
        int c = 0;
        if ((threadIdx.x & (kWave - 1)) == 0) c = atomicAdd(next_col, 1);
        return __builtin_amdgcn_readfirstlane(c);
    };
    if (w < a.nws) {
        int c = take();
        int4 m0 = make_int4(0, 0, 0, 0), m1 = m0;
        if (c < a.n) meta(c, m0, m1);
        while (c < a.n) {
            const int cn = take();
            int4 n0 = m0, n1 = m1;
            if (cn < a.n) meta(cn, n0, n1);
            rf_column(a, m0, m1, c >= a.T0, ws, wdesc, LU, ginst, kind, cmx, done);
            col_publish(done, m0.x);
            c = cn;
            m0 = n0;
            m1 = n1;
        }
    }
    __syncthreads();
    if (a.t > 0) {
        // products L(i, r) * U(r, j) of the deferred updates, all final now
        double* dprod = a.dprod + (size_t)inst * a.ndterm;
        for (int e0 = threadIdx.x; e0 < a.ndterm; e0 += 4 * kRfThreads) {
            int2 q[4];
            double l[4], u[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) q[j] = reinterpret_cast<const int2*>(a.dterm)[min(e0 + j * kRfThreads, a.ndterm - 1)];
#pragma unroll
            for (int j = 0; j < 4; ++j) { l[j] = LU[q[j].x]; u[j] = LU[q[j].y]; }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (e0 + j * kRfThreads < a.ndterm) dprod[e0 + j * kRfThreads] = l[j] * u[j];
        }
        __syncthreads();
        rf_tail(a, reinterpret_cast<double*>(smem), LU, ginst, kind, cmx,
                reinterpret_cast<uint32_t*>(smem + a.union_bytes + align16_d((size_t)a.t * 8)), dprod);
        __syncthreads();
    }
    double* SV = a.SV + (size_t)inst * a.nsv;
    // solve-ordered copy, eight gathers in flight per thread; entries the
    // sweeps mask (position -1) are not copied
    for (int i0 = threadIdx.x; i0 < a.nsv; i0 += 8 * kRfThreads) {
        int q[8];
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = a.sv_pos[min(i0 + u * kRfThreads, a.nsv - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = q[u] >= 0 ? LU[q[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + u * kRfThreads < a.nsv && q[u] >= 0) SV[i0 + u * kRfThreads] = v[u];
    }
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

// Operands of one thread's (or wave's) first row of a level, loaded one
// level ahead: the row record and its first terms' values and sources
// (thread mode: 8 terms; wave mode: 8 chunks of 64).
struct RowPrefetch {
    int4 ri;     // record of this thread's (wave's) first row of the level
    int4 rn;     // record of the level three ahead, in flight
    double v[8];
    int s[8];
};

// Column sweep over a program's densest end (DevProgram::sweep_*), one
// wave.  Lane l holds the targets l + 64 m (sweep order) in registers; the
// block of 64 targets the current steps finalize is always register 0 (the
// registers rotate after each block).  Step s takes target s's value from
// its lane, applies the diagonal operation, stores it, and subtracts its
// contribution from every later target with a term from it, so each target
// receives its terms in step order -- the order of its row in the reference
// loop.  A row-gather level costs its longest chain of dependent
// subtractions; a sweep step costs one, for all targets at once.
struct SweepArgs {
    int n, sv;
    const int32_t* tgt;
    const int32_t* bits;
};

template <int kOp>
__device__ __forceinline__ void run_sweep(const SweepArgs p, double* y, const double* SV) {
    constexpr bool skip = (kOp & OP_SKIPZERO) != 0;
    const int lane = threadIdx.x & (kWave - 1);
    const int D = p.n;
    const double* V = SV + p.sv;
    double a[4];
    int tg[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        tg[m] = p.tgt[min(lane + m * kWave, D - 1)];
        a[m] = y[tg[m]];
    }
    // Operands of step s (four values per lane, relative to the step's own
    // block of 64 targets, and the lane's presence bits) live in ring slot
    // s % kSweepPf; every load is a vector load issued kSweepPf steps ahead.
    double ring[kSweepPf][4];
    int bring[kSweepPf];
    auto fetch = [&](int st, int slot) {
        st = min(st, D - 1);
        const int base = (st / kWave) * kWave + lane;
#pragma unroll
        for (int m = 0; m < 4; ++m) ring[slot][m] = V[(size_t)st * D + min(base + m * kWave, D - 1)];
        bring[slot] = p.bits[(size_t)st * kWave + lane];
    };
#pragma unroll
    for (int q = 0; q < kSweepPf; ++q) fetch(q, q);
    for (int mb = 0; mb * kWave < D; ++mb) {
#pragma unroll
        for (int tt = 0; tt < kWave; ++tt) {
            const int st = mb * kWave + tt;
            if (st >= D) break;
            const int slot = tt % kSweepPf;
            double v[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) v[m] = ring[slot][m];
            const int bits = bring[slot] >> mb;
            fetch(st + kSweepPf, slot);
            double x = readlane_f64(a[0], tt);
            if constexpr ((kOp & (OP_MUL_DIAG | OP_DIV_DIAG | OP_MUL_NZ)) != 0) {
                const double d = readlane_f64(v[0], tt);   // the block diagonal
                if constexpr ((kOp & OP_MUL_DIAG) != 0) x = x * d;
                else if constexpr ((kOp & OP_DIV_DIAG) != 0) x = x / d;
                else if (x != 0.0) x = x * d;
            }
            if (lane == tt) y[tg[0]] = x;
            if (!(skip && x == 0.0)) {
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    if (m + mb < 4) {
                        const double na = a[m] - v[m] * x;
                        a[m] = ((bits >> m) & 1) ? na : a[m];
                    }
                }
            }
        }
        a[0] = a[1]; a[1] = a[2]; a[2] = a[3];
        tg[0] = tg[1]; tg[1] = tg[2]; tg[2] = tg[3];
    }
}

// Runs one program level by level.  Each row's subtraction chain is
// sequential in term order.  Thread mode: one thread per row, term loads
// issued eight at a time.  Wave mode: one wave per row, the lanes form the
// products of 64 terms at once and the chain then consumes them in order.
//
// Prefetch: the operands of each thread's first row are loaded two levels
// ahead and its row record three levels ahead (none of them depend on
// values a level writes).  The three buffers rotate by name -- the level
// loop is unrolled three times -- so no buffer is ever copied: a copy of a
// register that a load is still filling would make the compiler wait for
// every outstanding load at the end of each level.
template <bool kSV>
__device__ __forceinline__ void run_program(const DevProgram& p, const int* lvl, const int* lmode,
                                            double* y, const double* LU, const double* SV) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
    const bool skip = (p.op & OP_SKIPZERO) != 0;
    const bool has_diag = (p.op & (OP_MUL_DIAG | OP_DIV_DIAG | OP_MUL_NZ)) != 0;
    const int nlev = p.nlevels;
    auto val = [&](int q) { return kSV ? SV[p.sv_base + q] : LU[p.pos[q]]; };
    auto finish = [&](double acc, int dix) {
        if (has_diag && dix >= 0) {   // dix < 0: a partial row the sweep finishes
            const double d = kSV ? SV[dix] : LU[dix];
            if (p.op & OP_MUL_DIAG) acc = acc * d;
            else if (p.op & OP_DIV_DIAG) acc = acc / d;
            else if (acc != 0.0) acc = acc * d;
        }
        return acc;
    };
    // Loads are clamped rather than skipped: the level table is read at a
    // valid level and the term arrays carry padding.
    auto record = [&](int l) {
        l = min(l, nlev - 1);
        const int r0 = lvl[l], r1 = lvl[l + 1];
        return p.rows[min((lmode[l] == 0 ? r0 + tid : r0 + w), r1 - 1)];
    };
    auto operands = [&](int l, RowPrefetch& f) {
        l = min(l, nlev - 1);
        const bool tmode = lmode[l] == 0;
        const int last = max(f.ri.z - 1, 0);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int q = f.ri.y + min(tmode ? j : j * kWave + lane, last);
            f.v[j] = val(q);
            f.s[j] = p.src[q];
        }
    };
    auto level = [&](int l, const RowPrefetch& cur) {
        const int r0 = lvl[l], r1 = lvl[l + 1];
        if (lmode[l] == 0) {
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
                for (int c0 = 0; c0 < ri.z; c0 += 8 * kWave) {
                    // up to 8 chunks of operands in flight at once
                    double v[8];
                    int sx[8];
                    if (first && c0 == 0) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) { v[j] = cur.v[j]; sx[j] = cur.s[j]; }
                    } else {
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const int q = ri.y + min(c0 + j * kWave + lane, ri.z - 1);
                            v[j] = val(q);
                            sx[j] = p.src[q];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int c = c0 + j * kWave;
                        if (c < ri.z) {
                            const double ys = y[sx[j]];
                            // a skipped term subtracts +0.0, which leaves acc unchanged
                            const double prod = skip ? (ys != 0.0 ? ys * v[j] : 0.0) : v[j] * ys;
                            // the row record is wave-uniform: a scalar term count lets
                            // a full chunk run as 64 straight-line subtractions with
                            // immediate lane indices (same order as the tail loop)
                            const int m = __builtin_amdgcn_readfirstlane(min(kWave, ri.z - c));
                            if (m == kWave) {
#pragma unroll
                                for (int k = 0; k < kWave; ++k) acc -= readlane_f64(prod, k);
                            } else {
                                for (int k = 0; k < m; ++k) acc -= readlane_f64(prod, k);
                            }
                        }
                    }
                }
                acc = finish(acc, ri.w);
                if (lane == 0) y[ri.x] = acc;
            }
        }
        __syncthreads();
    };
    // One level: start the operands of level l + 2 (their record arrived a
    // level ago) and the record of level l + 3 (into the slot the current
    // buffer no longer needs), then run level l.
    auto body = [&](int l, RowPrefetch& cur, RowPrefetch& ahead2) {
        ahead2.ri = ahead2.rn;
        operands(l + 2, ahead2);
        cur.rn = record(l + 3);
        level(l, cur);
    };
    if (nlev > 0) {
    RowPrefetch F0, F1, F2;
    F0.ri = record(0);
    operands(0, F0);
    F1.ri = record(1);
    operands(1, F1);
    F2.rn = record(2);
    int l = 0;
    for (; l + 3 <= nlev; l += 3) {
        body(l, F0, F2);
        body(l + 1, F1, F0);
        body(l + 2, F2, F1);
    }
    if (l < nlev) body(l, F0, F2);
    if (l + 1 < nlev) body(l + 1, F1, F0);
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
    // level tables of both programs after the vector
    int* lv1 = reinterpret_cast<int*>(smem + (size_t)a.n * 8);
    int* lm1 = lv1 + a.p1.nlevels + 1;
    int* lv2 = lm1 + a.p1.nlevels;
    int* lm2 = lv2 + a.p2.nlevels + 1;
    for (int k = threadIdx.x; k < a.n; k += kSvThreads) y[k] = a.b[(size_t)a.in_perm[k] * a.Bg + ginst];
    for (int l = threadIdx.x; l <= a.p1.nlevels; l += kSvThreads) lv1[l] = a.p1.level_ptr[l];
    for (int l = threadIdx.x; l < a.p1.nlevels; l += kSvThreads) lm1[l] = a.p1.level_mode[l];
    for (int l = threadIdx.x; l <= a.p2.nlevels; l += kSvThreads) lv2[l] = a.p2.level_ptr[l];
    for (int l = threadIdx.x; l < a.p2.nlevels; l += kSvThreads) lm2[l] = a.p2.level_mode[l];
    __syncthreads();
    run_program<kSV>(a.p1, lv1, lm1, y, LU, SV);
    if constexpr (kSV) {
        // The forward program's sweep ends it, the backward program's starts
        // it; both run here, one after the other, by the first wave.
#pragma nounroll
        for (int k = 0; k < 2; ++k) {
            const bool on = k == 0 ? (a.p1.sweep_n > 0 && !a.p1.sweep_first)
                                   : (a.p2.sweep_n > 0 && a.p2.sweep_first);
            if (!on) continue;
            const SweepArgs sw = k == 0 ? SweepArgs{a.p1.sweep_n, a.p1.sweep_sv, a.p1.sweep_tgt, a.p1.sweep_bits}
                                        : SweepArgs{a.p2.sweep_n, a.p2.sweep_sv, a.p2.sweep_tgt, a.p2.sweep_bits};
            const int op = k == 0 ? a.p1.op : a.p2.op;
            if (threadIdx.x < kWave) {
                if (op == OP_PLAIN) run_sweep<OP_PLAIN>(sw, y, SV);
                else if (op == OP_DIV_DIAG) run_sweep<OP_DIV_DIAG>(sw, y, SV);
                else run_sweep<OP_SKIPZERO | OP_MUL_NZ>(sw, y, SV);
            }
            __syncthreads();
        }
    }
    run_program<kSV>(a.p2, lv2, lm2, y, LU, SV);
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
    // Step descriptors and relative slots.  In a trailing column, a pivot r
    // before the block that the U list puts after a trailing pivot s is
    // "deferred": the sparse pass applies it to the rows before the block
    // only (rel = -1 for the block's rows), and the dense phase applies it
    // to the block's rows right after step s, in list order (dtask lists).
    //
    // Relative slots are grouped by eight consecutive steps of a column:
    // group G holds, for lane l and step j = u % 8 of the group, the slot
    // of chunk entry l at rel[G * 512 + l * 8 + j] (-1: none), so one
    // 16-byte load per lane brings a lane's slots for eight steps.  desc.w
    // is the step's group.
    std::vector<int32_t> desc_ptr(n + 1, 0);
    std::vector<int4> desc;
    std::vector<int16_t> rel;
    int ngroups = 0;
    std::vector<std::vector<std::pair<int32_t, std::vector<int32_t>>>> step_tasks(std::max(t, 1));
    std::vector<int32_t> slot(n, -1);
    std::vector<int32_t> task_of(std::max(t, 1), -1);  // target row -> task in step_tasks[cur]
    for (int k = 0; k < n; ++k) {
        const int cb = s_.cptr[k];
        const int d_begin = (int)desc.size(), g_begin = ngroups;
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
            if ((size_t)(ngroups + 2) * 512 * 2 >= (size_t)INT_MAX)
                throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: plan too large");
            // L(:, r) in chunks of 63 (one entry per lane; lane 63 carries
            // the diagonal for Sparse 1.3); a continuation chunk recomputes
            // the same U value (x_r is not among its targets)
            for (int c0 = 0; c0 < std::max(nl[r], 1); c0 += kChunk) {
                const int cl = std::min(kChunk, nl[r] - c0);
                const int u = (int)desc.size() - d_begin;   // step index in the column
                const int G = g_begin + u / 8;
                if (u % 8 == 0) {
                    ngroups = G + 1;
                    rel.resize((size_t)ngroups * 512, (int16_t)-1);
                }
                desc.push_back(make_int4((p - cb) | (std::max(cl, 0) << 16) | (c0 > 0 ? kContinued : 0),
                                         dpos[r] + 1 + c0, r, G));
                for (int e = 0; e < cl; ++e) {
                    const int i = s_.rows[dpos[r] + 1 + c0 + e];
                    if (slot[i] < 0) throw Error(NSA_ERR_LOGIC, "GpuLu: fill pattern not closed");
                    rel[(size_t)G * 512 + e * 8 + u % 8] = (deferred && i >= T0_) ? (int16_t)-1 : (int16_t)slot[i];
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
    // Deferred tasks by trailing step, regrouped by the thread owning the
    // target entry: tail row i, column j -> thread (i % 32) + 32 (j % 32),
    // register (i / 32) * 4 + j / 32.
    std::vector<int32_t> dstep_ptr(std::max(t, 1), -1), dtask_ptr, dterm;
    std::vector<int4> dtask;
    for (int s0 = 0; s0 < t; ++s0) {
        if (step_tasks[s0].empty()) continue;
        std::vector<std::vector<int4>> by_tid(kRfThreads);
        for (auto& tk : step_tasks[s0]) {
            const int j = tk.first / t, i = tk.first % t;
            const int e0 = (int)(dterm.size() / 2);
            dterm.insert(dterm.end(), tk.second.begin(), tk.second.end());
            const int e1 = (int)(dterm.size() / 2);
            by_tid[(i & 31) + 32 * (j & 31)].push_back(make_int4((i >> 5) * 4 + (j >> 5), e0, e1, 0));
        }
        dstep_ptr[s0] = (int32_t)(dtask_ptr.size() / (kRfThreads + 1));
        for (int th = 0; th < kRfThreads; ++th) {
            dtask_ptr.push_back((int32_t)dtask.size());
            dtask.insert(dtask.end(), by_tid[th].begin(), by_tid[th].end());
        }
        dtask_ptr.push_back((int32_t)dtask.size());
    }
    n_deferred_terms_ = (int64_t)dterm.size() / 2;
    d_dstep_ptr_ = upload(allocs_, dstep_ptr);
    d_dtask_ptr_ = upload(allocs_, dtask_ptr);
    d_dtask_ = upload(allocs_, dtask);
    d_dterm_ = upload(allocs_, dterm);
    NSA_HIP(hipMalloc(&d_dprod_, std::max<size_t>(1, dterm.size() / 2 * (size_t)B_) * sizeof(double)));
    allocs_.push_back(d_dprod_);
    d_cptr_ = upload(allocs_, s_.cptr);
    d_nU_ = upload(allocs_, s_.nU);
    d_rows_ = upload(allocs_, s_.rows);
    d_asrc_ = upload(allocs_, sched.a_src);
    d_gpos_ = upload(allocs_, s_.gmin_pos);
    d_desc_ptr_ = upload(allocs_, desc_ptr);
    int maxnd = 1;
    for (int k = 0; k < n; ++k) maxnd = std::max(maxnd, desc_ptr[k + 1] - desc_ptr[k]);
    // column positions: the sparse columns in level order, then the trailing ones
    std::vector<int4> cmeta(2 * (size_t)std::max(n, 1));
    for (int c = 0; c < n; ++c) {
        const int k = c < T0_ ? lev_cols[c] : c;
        cmeta[2 * c] = make_int4(k, s_.cptr[k], s_.cptr[k + 1] - s_.cptr[k], s_.nU[k]);
        cmeta[2 * c + 1] = make_int4(desc_ptr[k], desc_ptr[k + 1] - desc_ptr[k], 0, 0);
    }
    d_cmeta_ = upload(allocs_, cmeta);
    // LDS: per-wave regions (column workspace + staged descriptors),
    // overlaid by the trailing block, then the trailing columns' |U| maxima.
    wregion_ = (int)align16((size_t)wsz_ * 8 + (size_t)maxnd * 16);
    // after the union: |U| maxima of the trailing columns (t doubles, padded
    // to 16 bytes), then the trailing structure masks (12 words per step,
    // read 4 at a time)
    const size_t cmx_bytes = align16((size_t)std::max(t, 1) * 8) + (size_t)std::max(t, 1) * 48;
    // the column completion bits and the dealing counter
    const size_t flag_bytes = align16((size_t)((n + 31) / 32) * 4 + 4);
    if (cmx_bytes + flag_bytes >= lds_max) throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: completion bits exceed LDS");
    nws_ = (int)std::min<size_t>(kRfWaves, (lds_max - cmx_bytes - flag_bytes) / (size_t)wregion_);
    if (nws_ < 1) throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: column workspace exceeds LDS");
    union_bytes_ = align16(std::max((size_t)nws_ * wregion_, (size_t)(516 + 128) * 8));
    rf_lds_ = union_bytes_ + cmx_bytes + flag_bytes;
    if (rf_lds_ > lds_max) throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: refactor LDS budget");
    desc.push_back(make_int4(0, 0, 0, 0));
    d_desc_ = upload(allocs_, desc);
    if (rel.empty()) rel.assign(512, (int16_t)-1);
    d_rel_ = upload(allocs_, rel);
    nrel_ = (int)rel.size();
    // the refactor reads the factor and rel through buffer resources (32-bit byte ranges)
    if ((int64_t)s_.nlu() * 8 >= INT_MAX || (int64_t)nrel_ * 2 >= INT_MAX)
        throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: factor exceeds a buffer resource's range");
    std::vector<int32_t> cp = s_.cand_ptr, cpos = s_.cand_pos, dc = s_.diag_cand;
    if (cp.empty()) cp.assign(n + 1, 0);
    if (dc.empty()) dc.assign(n, -1);
    d_cand_ptr_ = upload(allocs_, cp);
    d_cand_pos_ = upload(allocs_, cpos);
    d_diag_cand_ = upload(allocs_, dc);
    d_row_perm_ = upload(allocs_, s_.row_perm);
    d_col_perm_ = upload(allocs_, s_.col_perm);
    // ---- dense trailing block: per-thread register positions and structure masks
    if (t > 128) throw Error(NSA_ERR_LOGIC, "GpuLu: trailing block larger than 128");
    std::vector<int32_t> tposmap((size_t)kRfThreads * 16, -1);
    std::vector<uint32_t> tmask((size_t)std::max(t, 1) * 12, 0u);
    auto setbit = [&](int step, int word0, int i) {
        tmask[(size_t)step * 12 + word0 + (i >> 5)] |= 1u << (i & 31);
    };
    std::vector<int32_t> tcand_ptr(t + 1, 0), tdiag(std::max(t, 1), -1);
    std::vector<int16_t> tcand;
    for (int j = 0; j < t; ++j) {
        const int k = T0_ + j;
        for (int p = s_.cptr[k]; p < s_.cptr[k + 1]; ++p) {
            const int r = s_.rows[p];
            if (r < T0_) continue;
            const int i = r - T0_;
            tposmap[(size_t)((i & 31) + 32 * (j & 31)) * 16 + (i >> 5) * 4 + (j >> 5)] = p;
            if (i > j) setbit(j, 0, i);                       // L(i, j): an L row of step j
            if (i < j) { setbit(i, 4, j); setbit(j, 8, i); }  // U(i, j): a U column of step i, a U row of column j
        }
        if (s_.form == FORM_GP && !s_.cand_ptr.empty()) {
            for (int c = s_.cand_ptr[k]; c < s_.cand_ptr[k + 1]; ++c)
                tcand.push_back((int16_t)(s_.rows[s_.cand_pos[c]] - T0_));
            tdiag[j] = s_.diag_cand[k];
        }
        tcand_ptr[j + 1] = (int32_t)tcand.size();
        if (tcand_ptr[j + 1] - tcand_ptr[j] > 2 * kWave)
            throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: trailing step with more than 128 candidates");
    }
    d_tposmap_ = upload(allocs_, tposmap);
    d_tmask_ = upload(allocs_, tmask);
    d_tcand_ptr_ = upload(allocs_, tcand_ptr);
    d_tcand_ = upload(allocs_, tcand);
    d_tdiag_cand_ = upload(allocs_, tdiag);
    // ---- solve programs and the solve-ordered value copy of the plain solve
    TriProgram f, b, tf, tb;
    build_solve_programs(s_, f, b, tf, tb, kSweepMax);
    std::vector<int32_t> sv_pos;
    upload_program(f, fwd_, &sv_pos);
    upload_program(b, bwd_, &sv_pos);
    upload_program(tf, tfwd_, nullptr);
    upload_program(tb, tbwd_, nullptr);
    nsv_ = (int)sv_pos.size();
    d_sv_pos_ = upload(allocs_, sv_pos);
    // solve LDS: the vector, then both programs' level tables
    sv_lds_ = (size_t)n * 8 + (size_t)(2 * fwd_.nlevels + 1 + 2 * bwd_.nlevels + 1) * 4;
    sv_lds_t_ = (size_t)n * 8 + (size_t)(2 * tfwd_.nlevels + 1 + 2 * tbwd_.nlevels + 1) * 4;
    if (std::max(sv_lds_, sv_lds_t_) > lds_max)
        throw Error(NSA_ERR_UNSUPPORTED, "GpuLu: solve vector and level tables exceed LDS");
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
    if (sv_lds_ > 65536)
        NSA_HIP(hipFuncSetAttribute((const void*)k_solve<true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sv_lds_));
    if (sv_lds_t_ > 65536)
        NSA_HIP(hipFuncSetAttribute((const void*)k_solve<false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sv_lds_t_));
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
            const bool part = !t.partial.empty() && t.partial[tg];
            rows[idx] = make_int4(tg, first, cnt, (has_diag && !part) ? t.dpos[tg] : -1);
        }
        // latency model (cycles): a round of dependent loads ~2000 (8 terms per
        // thread, 8 x 64 per wave), a chained subtraction ~10 (thread) or ~12
        // (wave, with the lane read)
        const double R = r1 - r0;
        const double thread_cost = std::ceil(R / kSvThreads) * (std::ceil(maxc / 8.0) * 2000 + 10.0 * maxc);
        const double wave_cost = std::ceil(R / (kSvThreads / kWave)) * (std::ceil(maxc / 512.0) * 2000 + 12.0 * maxc);
        mode[l] = wave_cost < thread_cost ? 1 : 0;
    }
    if (sv_pos) {
        sv_pos->insert(sv_pos->end(), pos.begin(), pos.end());
        sv_pos->push_back(0);   // padding for clamped prefetches
        if (has_diag) {
            for (size_t idx = 0; idx < t.order.size(); ++idx) {
                const int tg = rows[idx].x;
                if (rows[idx].w < 0) continue;   // partial target: the sweep finishes it
                if (t.dpos[tg] < 0) throw Error(NSA_ERR_LOGIC, "GpuLu: solve program without diagonal");
                rows[idx].w = (int)sv_pos->size();
                sv_pos->push_back(t.dpos[tg]);
            }
        }
        if (t.sweep_n > 0) {
            // D x D value block: row st holds step st's terms on later targets
            // and, on the diagonal, target st's own diagonal value (absent
            // terms read position 0 and are masked out)
            const int D = t.sweep_n;
            if (t.op != OP_PLAIN && t.op != OP_DIV_DIAG && t.op != (OP_SKIPZERO | OP_MUL_NZ))
                throw Error(NSA_ERR_LOGIC, "GpuLu: no sweep kernel for this solve operation");
            d.sweep_n = D;
            d.sweep_first = t.sweep_first ? 1 : 0;
            d.sweep_sv = (int)sv_pos->size();
            std::vector<int32_t> bits((size_t)D * kWave, 0);
            for (int st = 0; st < D; ++st) {
                const int tg = t.sweep_tgt[st];
                if (has_diag && t.dpos[tg] < 0)
                    throw Error(NSA_ERR_LOGIC, "GpuLu: swept target without diagonal");
                for (int r = 0; r < D; ++r) {
                    int q = t.sweep_pos[(size_t)st * D + r];
                    if (q >= 0) bits[(size_t)st * kWave + r % kWave] |= 1 << (r / kWave);
                    if (r == st) q = has_diag ? t.dpos[tg] : -1;
                    // -1: an absent term (or the unused diagonal of a plain
                    // program); the sweep masks it, so the copy skips it
                    sv_pos->push_back(q);
                }
            }
            d.sweep_tgt = upload(allocs_, t.sweep_tgt);
            d.sweep_bits = upload(allocs_, bits);
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
    a.union_bytes = (int)union_bytes_;
    a.lev_ptr = d_lev_ptr_; a.cmeta = d_cmeta_;
    a.cptr = d_cptr_; a.nU = d_nU_; a.rows = d_rows_; a.a_src = d_asrc_; a.gmin_pos = d_gpos_;
    a.desc_ptr = d_desc_ptr_; a.desc = d_desc_; a.rel = d_rel_; a.nrel = nrel_;
    a.cand_ptr = d_cand_ptr_; a.cand_pos = d_cand_pos_; a.diag_cand = d_diag_cand_;
    a.tposmap = d_tposmap_;
    a.tcand_ptr = d_tcand_ptr_; a.tdiag_cand = d_tdiag_cand_; a.tcand = d_tcand_; a.tmask = d_tmask_;
    a.dstep_ptr = d_dstep_ptr_; a.dtask = d_dtask_; a.dtask_ptr = d_dtask_ptr_;
    a.dterm = d_dterm_;
    a.ndterm = (int)n_deferred_terms_;
    a.dprod = d_dprod_;
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
    const size_t lds = transposed ? sv_lds_t_ : sv_lds_;
    if (transposed)
        hipLaunchKernelGGL(k_solve<false>, dim3(B_), dim3(kSvThreads), lds, st, a);
    else
        hipLaunchKernelGGL(k_solve<true>, dim3(B_), dim3(kSvThreads), lds, st, a);
    NSA_HIP(hipGetLastError());
}

}  // namespace nsa
