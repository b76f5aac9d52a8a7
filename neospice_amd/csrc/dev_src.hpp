// Voltage-source evaluation shared by its own kernel (gpu_dev.hip) and the
// BSIM4v7 load's extra workgroups (gpu_bsim4.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "gpu_dev.hpp"

namespace nsa {

// VSource::value_at (vsource.cpp:80-155) for DC, PULSE and PWL.
__device__ inline double wave_value(const VsrcTable& t, int d, double dcv, double time) {
    const int w = t.wave[d];
    if (w == WAVE_PULSE) {
        const double* p = t.pulse + 7 * d;  // v1 v2 td tr tf pw per
        const double v1 = p[0], v2 = p[1], td = p[2], tr = p[3], tf = p[4], pw = p[5], per = p[6];
        if (time < td) return v1;
        const double t_rel = time - td;
        const double t_in = (per > 0.0) ? fmod(t_rel, per) : t_rel;
        if (t_in < tr) return v1 + (v2 - v1) * (t_in / tr);
        if (t_in < tr + pw) return v2;
        if (t_in < tr + pw + tf) {
            const double tfall = t_in - (tr + pw);
            return v2 + (v1 - v2) * (tfall / tf);
        }
        return v1;
    }
    if (w == WAVE_PWL) {
        const int a = t.pwl_ptr[d], b = t.pwl_ptr[d + 1];
        if (a == b) return dcv;
        if (time <= t.pwl_t[a]) return t.pwl_v[a];
        if (time >= t.pwl_t[b - 1]) return t.pwl_v[b - 1];
        int hi = a + 1;
        for (; hi < b; ++hi)
            if (t.pwl_t[hi] >= time) break;
        const int lo = hi - 1;
        const double t0 = t.pwl_t[lo], v0 = t.pwl_v[lo];
        const double t1 = t.pwl_t[hi], v1 = t.pwl_v[hi];
        if (t1 == t0) return v1;
        const double frac = (time - t0) / (t1 - t0);
        return v0 + (v1 - v0) * frac;
    }
    return dcv;
}

// VSource::evaluate rhs part (vsource.cpp:193-212) of (device, instance)
// gid; the +-1 matrix stamps are constant sites.
__device__ inline void vsrc_lane(const VsrcTable& t, const DevCtx* __restrict__ ctx,
                                 const int32_t* __restrict__ done, double* __restrict__ cr, int B, long gid) {
    if (gid >= (long)t.count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    const int ci = t.inst_of ? t.inst_of[gid] : inst;
    if (done && done[ci]) return;
    const DevCtx c = ctx[ci];
    const double dcv = t.dc[gid];
    double val;
    if (t.dc_given[d] && t.wave[d] != WAVE_DC && (c.mode & (MODEDCOP | MODEDCTRANCURVE)))
        val = dcv;
    else
        val = wave_value(t, d, dcv, c.time);
    val *= c.src_fact;
    cr[(size_t)(t.rsite + d) * B + inst] = val;
}

}  // namespace nsa
