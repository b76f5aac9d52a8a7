// Row/column equilibration of a batch of matrices (NeoSolver's radix-2
// scaling, src/core/neo_solver.cpp:52-102) and the scaling of vectors by
// its factors.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nsa {

// For every instance of A (nnz x B, slot k at row rows[k], column cols[k]):
// the scale factors dr, dc (n x B) and As = D_r A D_c with dr*dc*gmin added
// to the diagonal slots.  Asynchronous.
void launch_equilibrate(int n, int nnz, const int32_t* rows, const int32_t* cols, const double* A, double gmin,
                        int B, double* As, double* dr, double* dc, hipStream_t s);
// v[k] = scale[k] * v[k] for k < n * B.
void launch_scale_vec(int n, int B, const double* scale, double* v, hipStream_t s);

}  // namespace nsa
