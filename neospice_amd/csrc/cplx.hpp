// Complex arithmetic of the reference's Sparse 1.3 complex paths
// (complex_ops in src/solver/element.hpp:22-75), on (real, imaginary)
// pairs of doubles, for the host ordering and the device kernels alike.
// The operation order is the reference's: a complex factor or solve is
// bit-identical only if every product and sum is formed the same way.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace nsa {

// |re| + |im| (element_mag, element.hpp:18-20): the pivot magnitude.
__host__ __device__ inline double cmag(double r, double i) { return fabs(r) + fabs(i); }

// 1 / (dr + j di) by Smith's method with the reference's branch test
// (reciprocal, element.hpp:42-54).
__host__ __device__ inline void crecip(double& tr, double& ti, double dr, double di) {
    if ((dr >= di && dr > -di) || (dr < di && dr <= -di)) {
        const double r = di / dr;
        tr = 1.0 / (dr + r * di);
        ti = -r * tr;
    } else {
        const double r = dr / di;
        ti = -1.0 / (di + r * dr);
        tr = -r * ti;
    }
}

// t = t * f (mult_assign, element.hpp:28-33).
__host__ __device__ inline void cmul(double& tr, double& ti, double fr, double fi) {
    const double a = tr;
    tr = a * fr - ti * fi;
    ti = a * fi + ti * fr;
}

// t -= a * b (mult_subt_assign, element.hpp:35-40).
__host__ __device__ inline void cmsub(double& tr, double& ti, double ar, double ai, double br, double bi) {
    tr -= ar * br - ai * bi;
    ti -= ar * bi + ai * br;
}

}  // namespace nsa
