// Device loads (one lane per device instance, instance index fastest) and
// the Newton bookkeeping kernels.  Device kernels write each stamp call's
// value into its own "site"; the gather then sums the sites of every matrix
// slot / rhs row in load order, so no atomics are needed and every slot is
// accumulated in the reference's order (NumericMatrix::add from 0.0).
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "circuit.hpp"
#include "gpu_dev.hpp"
#include "dev_src.hpp"
#include "dev_terr.hpp"

namespace nsa {

namespace {

constexpr int TPB = 256;

inline int blocks_for(long work) { return (int)((work + TPB - 1) / TPB); }

__global__ void k_res_sites(int count, const double* __restrict__ ohms,
                            double* __restrict__ contrib, int site, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    const size_t sB = (size_t)B;
    // Resistor::evaluate (resistor.cpp:33-42)
    const double g = 1.0 / ohms[gid];
    double* c = contrib + (size_t)(site + 4 * d) * sB + inst;
    c[0] = g;
    c[sB] = -g;
    c[2 * sB] = -g;
    c[3 * sB] = g;
}

__global__ void k_const_sites(int count, const double* __restrict__ vals,
                              double* __restrict__ contrib, int site, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    contrib[(size_t)(site + d) * B + inst] = vals[d];
}

// Capacitor::evaluate, state-vector path (capacitor.cpp:28-110).
__global__ void k_cap(CapTable t, const DevCtx* __restrict__ ctx, const double* __restrict__ x,
                      double* __restrict__ s0, double* __restrict__ s1,
                      const double* __restrict__ s2, const int32_t* __restrict__ done,
                      double* __restrict__ cm, double* __restrict__ cr, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)t.count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    const int ci = t.inst_of ? t.inst_of[gid] : inst;
    if (done && done[ci]) return;
    const size_t sB = (size_t)B;
    const DevCtx c = ctx[ci];
    double* m = cm + (size_t)(t.site + 4 * d) * sB + inst;
    double* r = cr + (size_t)(t.rsite + 2 * d) * sB + inst;
    double geq = 0.0, ceq = 0.0;
    if (c.mode & (MODETRAN | MODEAC | MODETRANOP)) {
        const int np = t.nodes[2 * d], nn = t.nodes[2 * d + 1];
        const double cap = t.farads[gid];
        const bool cond1 = ((c.mode & MODEDC) && (c.mode & MODEINITJCT)) ||
                           ((c.mode & MODEUIC) && (c.mode & MODEINITTRAN));
        double vcap;
        if (cond1) {
            vcap = 0.0;  // no IC= on capacitors in this build
        } else {
            const double vp = np >= 0 ? x[(size_t)np * sB + inst] : 0.0;
            const double vn = nn >= 0 ? x[(size_t)nn * sB + inst] : 0.0;
            vcap = vp - vn;
        }
        const size_t q = (size_t)t.state[d] * sB + inst;  // state slot 0: charge
        const size_t iq = q + sB;                          // state slot 1: current
        if (c.mode & (MODETRAN | MODEAC)) {
            if (c.mode & MODEINITPRED) {
                s0[q] = s1[q];
            } else {
                s0[q] = cap * vcap;
                if (c.mode & MODEINITTRAN) s1[q] = s0[q];
            }
            int order = c.order;
            if (order < 1) order = 1;
            if (order > 2) order = 2;
            double deriv;
            if (order == 1) {
                deriv = c.ag0 * s0[q] + c.ag1 * s1[q];
            } else if (c.method == 0) {
                deriv = -s1[iq] * c.xmu_ratio + c.ag0 * s0[q] + c.ag1 * s1[q];
            } else {
                deriv = c.ag0 * s0[q] + c.ag1 * s1[q] + c.ag2 * s2[q];
            }
            s0[iq] = deriv;
            geq = c.ag0 * cap;
            ceq = s0[iq] - c.ag0 * s0[q];
            if (c.mode & MODEINITTRAN) s1[iq] = s0[iq];
        } else {
            s0[q] = cap * vcap;
        }
    }
    // multiplier m = 1 (m * geq == geq exactly)
    m[0] = geq;
    m[sB] = -geq;
    m[2 * sB] = -geq;
    m[3 * sB] = geq;
    r[0] = -ceq;
    r[sB] = ceq;
}

// VSource::evaluate (vsource.cpp:185-213): vsrc_lane per (device, instance).
__global__ void k_vsrc(VsrcTable t, const DevCtx* __restrict__ ctx,
                       const int32_t* __restrict__ done, double* __restrict__ cr, int B) {
    vsrc_lane(t, ctx, done, cr, B, (long)blockIdx.x * TPB + threadIdx.x);
}

// ISource::evaluate (isource.cpp:157-174), DC sources.
__global__ void k_isrc(IsrcTable t, const DevCtx* __restrict__ ctx,
                       const int32_t* __restrict__ done, double* __restrict__ cr, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)t.count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    const int ci = t.inst_of ? t.inst_of[gid] : inst;
    if (done && done[ci]) return;
    const double I = t.dc[gid] * ctx[ci].src_fact;
    double* r = cr + (size_t)(t.rsite + 2 * d) * B + inst;
    r[0] = -I;
    r[(size_t)B] = I;
}

constexpr int kWaveSize = 64;

// The sum of a row's sites, in site order from 0.0 (the slot's load-order
// accumulation).  Rows of more than 4 sites issue their loads a chunk at a
// time ahead of the in-order adds: 8, or 32 for long rows (a supply node
// touched by hundreds of devices), which then cost one memory latency per
// chunk instead of one per site.
template <int kChunk>
__device__ inline double sum_chunks(const int32_t* __restrict__ sites,
                                    const double* __restrict__ contrib, int p0, int p1, int inst,
                                    size_t sB) {
    double acc = 0.0;
    for (int p = p0; p < p1; p += kChunk) {
        int s[kChunk];
        double v[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) s[u] = sites[min(p + u, p1 - 1)];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) v[u] = contrib[(size_t)s[u] * sB + inst];
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
            if (p + u < p1) acc += v[u];
    }
    return acc;
}

__device__ inline double sum_sites(const int32_t* __restrict__ ptr, const int32_t* __restrict__ sites,
                                   const double* __restrict__ contrib, int r, int inst, size_t sB,
                                   bool uniform) {
    int p0 = ptr[r], p1 = ptr[r + 1];
    if (uniform) {
        p0 = __builtin_amdgcn_readfirstlane(p0);
        p1 = __builtin_amdgcn_readfirstlane(p1);
    }
    if (p1 - p0 > 16) return sum_chunks<32>(sites, contrib, p0, p1, inst, sB);
    if (p1 - p0 > 4) return sum_chunks<8>(sites, contrib, p0, p1, inst, sB);
    double acc = 0.0;
    for (int p = p0; p < p1; ++p) acc += contrib[(size_t)sites[p] * sB + inst];
    return acc;
}

__global__ void k_gather(int nrows, const int32_t* __restrict__ ptr,
                         const int32_t* __restrict__ sites, const int32_t* __restrict__ ptr2,
                         const int32_t* __restrict__ sites2, const double* __restrict__ contrib,
                         const double* __restrict__ contrib2, double* __restrict__ out, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)nrows * B) return;
    const int r = (int)(gid / B), inst = (int)(gid % B);
    const size_t sB = (size_t)B;
    // B a multiple of the wave size: the row is wave-uniform, so the site
    // list is read through the scalar cache
    const bool uniform = (B & (kWaveSize - 1)) == 0;
    double acc = sum_sites(ptr, sites, contrib, r, inst, sB, uniform);
    if (ptr2) acc += sum_sites(ptr2, sites2, contrib2, r, inst, sB, uniform);
    out[gid] = acc;
}

// Gather of a packed batch (B = 1): rows of at most kLongRow sites take one
// lane each; longer rows (a supply node's diagonal collects a stamp from
// every device on the rail) take a wave each, whose lanes load the row's
// site values into LDS with one round trip before one lane adds them in
// site order -- the same additions from 0.0 as the lane-per-row form.
constexpr int kLongRow = 32;
constexpr int kLongChunk = 1024;   // values a long-row wave stages per pass

__device__ inline double gsum_lane(const int32_t* __restrict__ ptr, const int32_t* __restrict__ sites,
                                   const double* __restrict__ contrib, int r) {
    const int p0 = ptr[r], p1 = ptr[r + 1];
    if (p1 - p0 > 4) return sum_chunks<8>(sites, contrib, p0, p1, 0, 1);
    double acc = 0.0;
    for (int p = p0; p < p1; ++p) acc += contrib[sites[p]];
    return acc;
}

__device__ inline double gsum_wave(const int32_t* __restrict__ ptr, const int32_t* __restrict__ sites,
                                   const double* __restrict__ contrib, int r, double* buf, int lane) {
    const int p0 = ptr[r], p1 = ptr[r + 1];
    double acc = 0.0;
    constexpr int kPer = kLongChunk / kWaveSize;
    for (int c0 = p0; c0 < p1; c0 += kLongChunk) {
        const int c1 = min(p1, c0 + kLongChunk);
        // every lane issues all its index loads, then all its value loads,
        // so the staging costs two memory latencies per chunk
        int sx[kPer];
        double v[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int p = c0 + lane + u * kWaveSize;
            sx[u] = p < c1 ? sites[p] : -1;
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) v[u] = sx[u] >= 0 ? contrib[sx[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int p = c0 + lane + u * kWaveSize;
            if (p < c1) buf[p - c0] = v[u];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane == 0)
            for (int p = 0; p < c1 - c0; ++p) acc += buf[p];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    return acc;
}

__global__ __launch_bounds__(TPB) void k_gather_packed(GatherPacked m, GatherPacked r) {
    __shared__ double buf[TPB / kWaveSize][kLongChunk];
    const int lane = threadIdx.x & (kWaveSize - 1), wave = threadIdx.x / kWaveSize;
    const long short_blocks_m = ((long)m.nshort + TPB - 1) / TPB;
    const long short_blocks_r = ((long)r.nshort + TPB - 1) / TPB;
    long b = blockIdx.x;
    // blocks: short rows of m, short rows of r, then long rows of m and r
    // (TPB / 64 rows per block)
    if (b < short_blocks_m + short_blocks_r) {
        const GatherPacked& g = (b < short_blocks_m) ? m : r;
        if (b >= short_blocks_m) b -= short_blocks_m;
        const long k = b * TPB + threadIdx.x;
        if (k >= g.nshort) return;
        const int row = g.short_rows[k];
        double acc = gsum_lane(g.ptr, g.sites, g.contrib, row);
        if (g.ptr2) acc += gsum_lane(g.ptr2, g.sites2, g.contrib2, row);
        g.out[row] = acc;
        return;
    }
    b -= short_blocks_m + short_blocks_r;
    const long k = b * (TPB / kWaveSize) + wave;
    const GatherPacked& g = (k < m.nlong) ? m : r;
    const long kk = (k < m.nlong) ? k : k - m.nlong;
    if (kk >= g.nlong) return;
    const int row = g.long_rows[kk];
    // either list of a row may be long; each runs the form its length needs
    double acc = (g.ptr[row + 1] - g.ptr[row] > kLongRow) ? gsum_wave(g.ptr, g.sites, g.contrib, row, buf[wave], lane)
                                                         : gsum_lane(g.ptr, g.sites, g.contrib, row);
    if (g.ptr2) {
        const double a2 = (g.ptr2[row + 1] - g.ptr2[row] > kLongRow)
                              ? gsum_wave(g.ptr2, g.sites2, g.contrib2, row, buf[wave], lane)
                              : gsum_lane(g.ptr2, g.sites2, g.contrib2, row);
        acc += a2;
    }
    if (lane == 0) g.out[row] = acc;
}

// Newton convergence test (newton.cpp:267-292): any failing unknown marks
// its instance non-converged.
__global__ void k_conv(int n, int num_nodes, const double* __restrict__ xn,
                       const double* __restrict__ xo, double reltol, double vntol, double abstol,
                       const int32_t* __restrict__ done, int32_t* __restrict__ nonconv, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)n * B) return;
    const int i = (int)(gid / B), inst = (int)(gid % B);
    if (done && done[inst]) return;
    const double vn = xn[gid], vo = xo[gid];
    bool bad;
    if (isnan(vn) || isinf(vn)) {
        bad = true;
    } else {
        const double diff = fabs(vn - vo);
        const double tol = reltol * fmax(fabs(vn), fabs(vo)) + (i < num_nodes ? vntol : abstol);
        bad = diff > tol;
    }
    if (bad) atomicOr(&nonconv[inst], 1);
}

__global__ void k_commit(int n, const double* __restrict__ src, double* __restrict__ dst,
                         const int32_t* __restrict__ mask, int mask_value, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)n * B) return;
    const int inst = (int)(gid % B);
    if (mask && mask[inst] != mask_value) return;
    dst[gid] = src[gid];
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int count) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid < count) p[gid] = v;
}

__global__ void k_take_instance(const double* __restrict__ A, int B, int inst, int count,
                                double* __restrict__ out) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid < count) out[gid] = A[gid * B + inst];
}

__global__ void k_trunc(TruncTable t, const DevCtx* __restrict__ ctx, const double* __restrict__ s0,
                        const double* __restrict__ s1, const double* __restrict__ s2,
                        const double* __restrict__ s3, int B, TruncOpts o, double* __restrict__ dt_dev,
                        unsigned long long* __restrict__ dt_min, int order_override,
                        unsigned long long* __restrict__ dt_min2, const int32_t* __restrict__ sel) {
    trunc_lane(t, ctx, s0, s1, s2, s3, B, o, dt_dev, dt_min, order_override, dt_min2, sel,
               blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

}  // namespace

void launch_trunc(const TruncTable& t, const DevCtx* ctx, const double* s0, const double* s1,
                  const double* s2, const double* s3, int B, const TruncOpts& o, double* dt_dev,
                  unsigned long long* dt_min, hipStream_t s, int order_override,
                  unsigned long long* dt_min2, const int32_t* sel) {
    const int total = t.ndev * B;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_trunc, dim3((total + 255) / 256), dim3(256), 0, s, t, ctx, s0, s1, s2, s3, B, o,
                       dt_dev, dt_min, order_override, dt_min2, sel);
}

void launch_fill_u64(unsigned long long* p, unsigned long long v, int count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill_u64, dim3((count + 255) / 256), dim3(256), 0, s, p, v, count);
}

void launch_res_sites(int count, const double* ohms, double* contrib, int site, int B,
                      hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_res_sites, dim3(blocks_for((long)count * B)), dim3(TPB), 0, s, count,
                       ohms, contrib, site, B);
}

void launch_const_sites(int count, const double* vals, double* contrib, int site, int B,
                        hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_const_sites, dim3(blocks_for((long)count * B)), dim3(TPB), 0, s, count,
                       vals, contrib, site, B);
}

void launch_cap(const CapTable& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                double* s2, const int32_t* done, double* cm, double* cr, int B, hipStream_t s) {
    if (!t.count) return;
    hipLaunchKernelGGL(k_cap, dim3(blocks_for((long)t.count * B)), dim3(TPB), 0, s, t, ctx, x, s0,
                       s1, s2, done, cm, cr, B);
}

void launch_vsrc(const VsrcTable& t, const DevCtx* ctx, const int32_t* done, double* cr, int B,
                 hipStream_t s) {
    if (!t.count) return;
    hipLaunchKernelGGL(k_vsrc, dim3(blocks_for((long)t.count * B)), dim3(TPB), 0, s, t, ctx, done,
                       cr, B);
}

void launch_isrc(const IsrcTable& t, const DevCtx* ctx, const int32_t* done, double* cr, int B,
                 hipStream_t s) {
    if (!t.count) return;
    hipLaunchKernelGGL(k_isrc, dim3(blocks_for((long)t.count * B)), dim3(TPB), 0, s, t, ctx, done,
                       cr, B);
}

void launch_gather(int nrows, const int32_t* ptr, const int32_t* sites, const int32_t* ptr2,
                   const int32_t* sites2, const double* contrib, const double* contrib2,
                   double* out, int B, hipStream_t s) {
    if (!nrows) return;
    hipLaunchKernelGGL(k_gather, dim3(blocks_for((long)nrows * B)), dim3(TPB), 0, s, nrows, ptr,
                       sites, ptr2, sites2, contrib, contrib2, out, B);
}

void launch_gather_packed(const GatherPacked& m, const GatherPacked& r, hipStream_t s) {
    const long blocks = ((long)m.nshort + TPB - 1) / TPB + ((long)r.nshort + TPB - 1) / TPB +
                        ((long)m.nlong + r.nlong + TPB / kWaveSize - 1) / (TPB / kWaveSize);
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_gather_packed, dim3((unsigned)blocks), dim3(TPB), 0, s, m, r);
}

void launch_conv(int n, int num_nodes, const double* xnew, const double* xold, double reltol,
                 double vntol, double abstol, const int32_t* done, int32_t* nonconv, int B,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_conv, dim3(blocks_for((long)n * B)), dim3(TPB), 0, s, n, num_nodes,
                       xnew, xold, reltol, vntol, abstol, done, nonconv, B);
}

void launch_commit(int n, const double* src, double* dst, const int32_t* mask, int mask_value,
                   int B, hipStream_t s) {
    hipLaunchKernelGGL(k_commit, dim3(blocks_for((long)n * B)), dim3(TPB), 0, s, n, src, dst,
                       mask, mask_value, B);
}

void launch_fill_i32(int32_t* p, int32_t v, int count, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(count)), dim3(TPB), 0, s, p, v, count);
}

void launch_take_instance(const double* A, int B, int inst, int count, double* out, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_take_instance, dim3(blocks_for(count)), dim3(TPB), 0, s, A, B, inst, count,
                       out);
}

}  // namespace nsa
