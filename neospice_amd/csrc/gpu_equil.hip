// Row/column equilibration of a batch of matrices (NeoSolver::
// load_real_equilibrated, src/core/neo_solver.cpp:52-102, with pow2_round
// :16-33) and the matching scaling of right-hand sides and solutions
// (NeoSolver::solve, neo_solver.cpp:204-225).  One workgroup per instance;
// batched arrays are slot-major with the instance fastest (a[k * B + inst]).
//
// The scale factors are powers of two, so every scaled value is exact
// (barring overflow and subnormals) and the order of the two products is
// immaterial; it is kept as the reference writes it anyway.  Row and column
// maxima are taken with integer atomics on the bit patterns of non-negative
// doubles, which order like their values.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "gpu_equil.hpp"

namespace nsa {

namespace {

constexpr int kT = 256;

// The power of two nearest to v in log2 (1 for a non-positive or
// non-finite v).
__device__ inline double pow2_round(double v) {
    if (!(v > 0.0) || !isfinite(v)) return 1.0;
    int e = 0;
    const double m = frexp(v, &e);
    return (m < 0.70710678118654752440) ? ldexp(1.0, e - 1) : ldexp(1.0, e);
}

__device__ inline unsigned long long bits(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ inline double from_bits(unsigned long long b) { return __longlong_as_double((long long)b); }

__global__ __launch_bounds__(kT) void k_equilibrate(int n, int nnz, const int32_t* __restrict__ rows,
                                                   const int32_t* __restrict__ cols, const double* __restrict__ A,
                                                   double gmin, int B, double* __restrict__ As,
                                                   double* __restrict__ dr, double* __restrict__ dc) {
    extern __shared__ unsigned long long mx[];  // n maxima (bit patterns)
    const int inst = blockIdx.x, t = threadIdx.x;
    auto a = [&](int k) { return A[(size_t)k * B + inst]; };
    auto R = [&](int i) -> double& { return dr[(size_t)i * B + inst]; };
    auto Cc = [&](int j) -> double& { return dc[(size_t)j * B + inst]; };
    // row infinity norms: D_r[i] = pow2(1 / max_j |A_ij|)
    for (int i = t; i < n; i += kT) mx[i] = 0ull;
    __syncthreads();
    for (int k = t; k < nnz; k += kT) atomicMax(&mx[rows[k]], bits(fabs(a(k))));
    __syncthreads();
    for (int i = t; i < n; i += kT) {
        const double m = from_bits(mx[i]);
        R(i) = pow2_round(m > 0.0 ? 1.0 / m : 1.0);
        mx[i] = 0ull;
    }
    __syncthreads();
    // column norms of the row-scaled matrix: D_c[j] = pow2(1 / max_i |D_r[i] A_ij|)
    for (int k = t; k < nnz; k += kT) atomicMax(&mx[cols[k]], bits(fabs(R(rows[k]) * a(k))));
    __syncthreads();
    for (int j = t; j < n; j += kT) {
        const double m = from_bits(mx[j]);
        Cc(j) = pow2_round(m > 0.0 ? 1.0 / m : 1.0);
        mx[j] = 0ull;
    }
    __syncthreads();
    // one refining row sweep on the fully scaled matrix
    for (int k = t; k < nnz; k += kT) atomicMax(&mx[rows[k]], bits(fabs(R(rows[k]) * a(k) * Cc(cols[k]))));
    __syncthreads();
    for (int i = t; i < n; i += kT) {
        const double m = from_bits(mx[i]);
        if (m > 0.0) R(i) *= pow2_round(1.0 / m);
    }
    __syncthreads();
    // the scaled values, then the scaled regularization on diagonal slots
    for (int k = t; k < nnz; k += kT) {
        const int r = rows[k], c = cols[k];
        double v = R(r) * a(k) * Cc(c);
        if (gmin != 0.0 && r == c) v += R(r) * Cc(r) * gmin;
        As[(size_t)k * B + inst] = v;
    }
}

__global__ void k_scale_vec(int n, int B, const double* __restrict__ s, double* __restrict__ v) {
    const long k = (long)blockIdx.x * kT + threadIdx.x;
    if (k < (long)n * B) v[k] = s[k] * v[k];
}

}  // namespace

void launch_equilibrate(int n, int nnz, const int32_t* rows, const int32_t* cols, const double* A, double gmin,
                        int B, double* As, double* dr, double* dc, hipStream_t s) {
    if (B <= 0 || n <= 0) return;
    const size_t lds = sizeof(unsigned long long) * (size_t)n;
    if (lds > 64 * 1024) throw Error(NSA_ERR_UNSUPPORTED, "equilibration: more than 8192 unknowns");
    hipLaunchKernelGGL(k_equilibrate, dim3(B), dim3(kT), lds, s, n, nnz, rows, cols, A, gmin, B, As, dr, dc);
    NSA_HIP(hipGetLastError());
}

void launch_scale_vec(int n, int B, const double* scale, double* v, hipStream_t s) {
    const long total = (long)n * B;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_scale_vec, dim3((unsigned)((total + kT - 1) / kT)), dim3(kT), 0, s, n, B, scale, v);
}

}  // namespace nsa
