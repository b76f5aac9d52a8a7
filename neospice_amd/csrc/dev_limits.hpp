// Device-side helpers shared by the UCB-style nonlinear loads (diode,
// MOS1, BSIM4v7): the ngspice limiting functions, NIintegrate and the
// solution read.  DEVfetlim, DEVlimvds, DEVpnjlim and NIintegrate are the
// same in the device shims that restate them (mos1_shim.cpp:44-153,
// bsim4v7_shim.cpp:44-155, dio_shim.cpp:46-73).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

#include "gpu_dev.hpp"

namespace nsa {
namespace devk {

// UCB MIN / MAX macros (ucb_compat.hpp:38-43), operand order kept.
__device__ inline double umax(double a, double b) { return (a > b) ? a : b; }
__device__ inline double umin(double a, double b) { return (a < b) ? a : b; }
// std::max as the device wrappers use it
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }

__device__ inline double sol(const double* x, int node, size_t sB, int inst) {
    return node >= 0 ? x[(size_t)node * sB + inst] : 0.0;
}

// DEVpnjlim (dio_shim.cpp / mos1_shim.cpp:101-121)
__device__ inline double pnjlim(double vnew, double vold, double vt, double vcrit, int* icheck) {
    double arg;
    if ((vnew > vcrit) && (fabs(vnew - vold) > (vt + vt))) {
        if (vold > 0) {
            arg = (vnew - vold) / vt;
            if (arg > 0) vnew = vold + vt * (2 + log(arg - 2));
            else vnew = vold - vt * (2 + log(2 - arg));
        } else {
            vnew = vt * log(vnew / vt);
        }
        *icheck = 1;
    } else {
        if (vnew < 0) {
            if (vold > 0) arg = -1 * vold - 1;
            else arg = 2 * vold - 1;
            if (vnew < arg) { vnew = arg; *icheck = 1; }
            else *icheck = 0;
        } else {
            *icheck = 0;
        }
    }
    return vnew;
}

// DEVlimvds (mos1_shim.cpp:86-99)
__device__ inline double limvds(double vnew, double vold) {
    if (vold >= 3.5) {
        if (vnew > vold) {
            vnew = fmin(vnew, (3 * vold) + 2);
        } else {
            if (vnew < 3.5) vnew = fmax(vnew, 2);
        }
    } else {
        if (vnew > vold) vnew = fmin(vnew, 4);
        else vnew = fmax(vnew, -.5);
    }
    return vnew;
}

// DEVfetlim (mos1_shim.cpp:123-153)
__device__ inline double fetlim(double vnew, double vold, double vto) {
    const double vtsthi = fabs(2 * (vold - vto)) + 2;
    const double vtstlo = fabs(vold - vto) + 1;
    const double vtox = vto + 3.5;
    const double delv = vnew - vold;
    if (vold >= vto) {
        if (vold >= vtox) {
            if (delv <= 0) {
                if (vnew >= vtox) {
                    if (-delv > vtstlo) vnew = vold - vtstlo;
                } else {
                    vnew = fmax(vnew, vto + 2);
                }
            } else {
                if (delv >= vtsthi) vnew = vold + vtsthi;
            }
        } else {
            if (delv <= 0) vnew = fmax(vnew, vto - .5);
            else vnew = fmin(vnew, vto + 4);
        }
    } else {
        if (delv <= 0) {
            if (-delv > vtsthi) vnew = vold - vtsthi;
        } else {
            const double vtemp = vto + .5;
            if (vnew <= vtemp) {
                if (delv > vtstlo) vnew = vold + vtstlo;
            } else {
                vnew = vtemp;
            }
        }
    }
    return vnew;
}

// NIintegrate (dio_shim.cpp:46-73, mos1_shim.cpp:44-73): charge at q,
// current at q + 1 (slot strides of sB).
__device__ inline void niintegrate(const DevCtx& c, double* s0, const double* s1,
                                   const double* s2, size_t q, size_t sB, double cap, double* geq,
                                   double* ceq) {
    int order = c.order;
    if (order < 1) order = 1;
    if (order > 2) order = 2;
    double deriv;
    if (order == 1) {
        deriv = c.ag0 * s0[q] + c.ag1 * s1[q];
    } else if (c.method == 0) {
        deriv = -s1[q + sB] * c.xmu_ratio + c.ag0 * s0[q] + c.ag1 * s1[q];
    } else {
        deriv = c.ag0 * s0[q] + c.ag1 * s1[q] + c.ag2 * s2[q];
    }
    s0[q + sB] = deriv;
    *geq = c.ag0 * cap;
    *ceq = s0[q + sB] - (*geq) * s0[q];
}

}  // namespace devk
}  // namespace nsa
