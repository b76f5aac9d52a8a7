// Adjoint kernels: parameter gradients -lambda^T dF/dp of the device
// families and the capacitor-current recursion of the discrete transient
// adjoint.  One lane per (device, instance), instance fastest; node-side
// accumulations are gathers over a node -> device incidence list in a fixed
// order, so every result is deterministic (no atomics).
#include <hip/hip_runtime.h>

#include "gpu_adjoint.hpp"

namespace nsa {

namespace {

constexpr int TPB = 256;
inline int blocks_for(long work) { return (int)((work + TPB - 1) / TPB); }

__device__ inline double at(const double* v, int node, size_t sB, int inst) {
    return node >= 0 ? v[(size_t)node * sB + inst] : 0.0;
}

// Resistor (resistor.cpp:33-42): F_a += (v_a - v_b)/R, F_b -= (v_a - v_b)/R, so
// dL/dR = -lambda^T dF/dR = (lambda_a - lambda_b)(v_a - v_b) / R^2.
__global__ void k_grad_res(int count, const int32_t* __restrict__ nodes,
                           const double* __restrict__ ohms, const double* __restrict__ x,
                           const double* __restrict__ lam, double* __restrict__ g, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    const size_t sB = (size_t)B;
    const int a = nodes[2 * d], b = nodes[2 * d + 1];
    const double R = ohms[gid];
    const double v = at(x, a, sB, inst) - at(x, b, sB, inst);
    const double l = at(lam, a, sB, inst) - at(lam, b, sB, inst);
    g[gid] += l * v / (R * R);
}

// Voltage source (vsource.cpp:193-212): rhs[branch] = V * src_fact, so
// dL/dV = lambda_branch * src_fact.  Current source (isource.cpp:157-174):
// rhs[p] -= I, rhs[n] += I, so dL/dI = (lambda_n - lambda_p) * src_fact.
__global__ void k_grad_src(int nv, const int32_t* __restrict__ vbranch, int ni,
                           const int32_t* __restrict__ inodes, const DevCtx* __restrict__ ctx,
                           const double* __restrict__ lam, double* __restrict__ gv,
                           double* __restrict__ gi, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    const size_t sB = (size_t)B;
    if (gid < (long)nv * B) {
        const int d = (int)(gid / B), inst = (int)(gid % B);
        gv[gid] += at(lam, vbranch[d], sB, inst) * ctx[inst].src_fact;
        return;
    }
    const long g2 = gid - (long)nv * B;
    if (g2 >= (long)ni * B) return;
    const int d = (int)(g2 / B), inst = (int)(g2 % B);
    gi[g2] += (at(lam, inodes[2 * d + 1], sB, inst) - at(lam, inodes[2 * d], sB, inst)) *
              ctx[inst].src_fact;
}

// Per-capacitor part of one backward step (capacitor.cpp:69-85 companion
// i_k = a0 C v_k + a1 C v_{k-1} - r i_{k-1}, r = xmu_ratio for trapezoidal
// order 2, else 0).  With u = lambda_a - lambda_b and w = a0 v_k + a1 v_{k-1}:
//   dL/dC += (ibar - u) w;   q = C (ibar - u)  (scattered by a1 into xbar_{k-1});
//   ibar_{k-1} = r (u - ibar).
__global__ void k_cap_back(int count, const int32_t* __restrict__ nodes,
                           const double* __restrict__ farads, const DevCtx* __restrict__ ctx,
                           const double* __restrict__ xk, const double* __restrict__ xkm1,
                           const double* __restrict__ lam, double* __restrict__ ibar,
                           double* __restrict__ q, double* __restrict__ g, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    const size_t sB = (size_t)B;
    const DevCtx c = ctx[inst];
    const int a = nodes[2 * d], b = nodes[2 * d + 1];
    const double r = (c.order >= 2 && c.method == 0) ? c.xmu_ratio : 0.0;
    const double u = at(lam, a, sB, inst) - at(lam, b, sB, inst);
    const double w = c.ag0 * (at(xk, a, sB, inst) - at(xk, b, sB, inst)) +
                     c.ag1 * (at(xkm1, a, sB, inst) - at(xkm1, b, sB, inst));
    const double ib = ibar[gid];
    g[gid] += (ib - u) * w;
    q[gid] = farads[gid] * (ib - u);
    ibar[gid] = r * (u - ib);
}

// out[node] (+)= coef * sum over incident capacitors of sign * val[cap],
// coef = ctx.ag0 (which = 0) or ctx.ag1 (which = 1); val optionally scaled
// by C (scale_c).
__global__ void k_cap_gather(int n, const int32_t* __restrict__ ptr,
                             const int32_t* __restrict__ ent, const double* __restrict__ farads,
                             const double* __restrict__ val, const DevCtx* __restrict__ ctx,
                             int which, int scale_c, int accumulate, double* __restrict__ out,
                             int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)n * B) return;
    const int i = (int)(gid / B), inst = (int)(gid % B);
    const size_t sB = (size_t)B;
    double acc = 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
        const int e = ent[p];
        const int cap = e >> 1;
        const double v = val[(size_t)cap * sB + inst] *
                         (scale_c ? farads[(size_t)cap * sB + inst] : 1.0);
        acc += (e & 1) ? -v : v;
    }
    const double coef = which == 0 ? ctx[inst].ag0 : ctx[inst].ag1;
    const double add = coef * acc;
    out[gid] = accumulate ? out[gid] + add : add;
}

// out[node(k)] += src[k] for probe k (dst n x B, src nprobe x B); probes are
// distinct nodes.
__global__ void k_probe_scatter(int nprobe, const int32_t* __restrict__ probes,
                                const double* __restrict__ src, double* __restrict__ dst, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)nprobe * B) return;
    const int k = (int)(gid / B), inst = (int)(gid % B);
    dst[(size_t)probes[k] * B + inst] += src[gid];
}

__global__ void k_probe_gather(int nprobe, const int32_t* __restrict__ probes,
                               const double* __restrict__ x, double* __restrict__ dst, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)nprobe * B) return;
    const int k = (int)(gid / B), inst = (int)(gid % B);
    dst[gid] = x[(size_t)probes[k] * B + inst];
}

}  // namespace

void launch_grad_res(int count, const int32_t* nodes, const double* ohms, const double* x,
                     const double* lam, double* g, int B, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_grad_res, dim3(blocks_for((long)count * B)), dim3(TPB), 0, s, count,
                       nodes, ohms, x, lam, g, B);
}

void launch_grad_src(int nv, const int32_t* vbranch, int ni, const int32_t* inodes,
                     const DevCtx* ctx, const double* lam, double* gv, double* gi, int B,
                     hipStream_t s) {
    if (!nv && !ni) return;
    hipLaunchKernelGGL(k_grad_src, dim3(blocks_for((long)(nv + ni) * B)), dim3(TPB), 0, s, nv,
                       vbranch, ni, inodes, ctx, lam, gv, gi, B);
}

void launch_cap_back(int count, const int32_t* nodes, const double* farads, const DevCtx* ctx,
                     const double* xk, const double* xkm1, const double* lam, double* ibar,
                     double* q, double* g, int B, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_cap_back, dim3(blocks_for((long)count * B)), dim3(TPB), 0, s, count,
                       nodes, farads, ctx, xk, xkm1, lam, ibar, q, g, B);
}

void launch_cap_gather(int n, const int32_t* ptr, const int32_t* ent, const double* farads,
                       const double* val, const DevCtx* ctx, int which, int scale_c,
                       int accumulate, double* out, int B, hipStream_t s) {
    hipLaunchKernelGGL(k_cap_gather, dim3(blocks_for((long)n * B)), dim3(TPB), 0, s, n, ptr, ent,
                       farads, val, ctx, which, scale_c, accumulate, out, B);
}

void launch_probe_scatter(int nprobe, const int32_t* probes, const double* src, double* dst,
                          int B, hipStream_t s) {
    if (!nprobe) return;
    hipLaunchKernelGGL(k_probe_scatter, dim3(blocks_for((long)nprobe * B)), dim3(TPB), 0, s,
                       nprobe, probes, src, dst, B);
}

void launch_probe_gather(int nprobe, const int32_t* probes, const double* x, double* dst, int B,
                         hipStream_t s) {
    if (!nprobe) return;
    hipLaunchKernelGGL(k_probe_gather, dim3(blocks_for((long)nprobe * B)), dim3(TPB), 0, s,
                       nprobe, probes, x, dst, B);
}

}  // namespace nsa
