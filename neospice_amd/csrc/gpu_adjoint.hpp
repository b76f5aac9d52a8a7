// Launch wrappers of the adjoint kernels (gpu_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gpu_dev.hpp"

namespace nsa {

// g[d, inst] += dL/dR of resistor d (count x B), given the solution x and
// the adjoint lambda (n x B).
void launch_grad_res(int count, const int32_t* nodes, const double* ohms, const double* x,
                     const double* lam, double* g, int B, hipStream_t s);
// gv += dL/dV (dc value), gi += dL/dI (dc value).
void launch_grad_src(int nv, const int32_t* vbranch, int ni, const int32_t* inodes,
                     const DevCtx* ctx, const double* lam, double* gv, double* gi, int B,
                     hipStream_t s);
// One backward step of the capacitor-current recursion; see the kernel.
void launch_cap_back(int count, const int32_t* nodes, const double* farads, const DevCtx* ctx,
                     const double* xk, const double* xkm1, const double* lam, double* ibar,
                     double* q, double* g, int B, hipStream_t s);
// out (+)= ag[which] * P (scale_c ? C val : val) over the capacitor incidence
// list: ptr[n+1], ent = 2 * cap + (1 if the node is the capacitor's minus pin).
void launch_cap_gather(int n, const int32_t* ptr, const int32_t* ent, const double* farads,
                       const double* val, const DevCtx* ctx, int which, int scale_c,
                       int accumulate, double* out, int B, hipStream_t s);
void launch_probe_scatter(int nprobe, const int32_t* probes, const double* src, double* dst,
                          int B, hipStream_t s);
void launch_probe_gather(int nprobe, const int32_t* probes, const double* x, double* dst, int B,
                         hipStream_t s);

}  // namespace nsa
