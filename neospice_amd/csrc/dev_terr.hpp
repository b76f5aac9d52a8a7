// Truncation-error helpers shared by the truncation kernel and the fused
// BSIM4v7 tick load: ckt_terr per charge and a per-instance minimum over a
// wave.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gpu_dev.hpp"

namespace nsa {

// ckt_terr (src/devices/ckt_terr.hpp:19-64) for the charge at state q of
// one instance, folded into dt the way the reference folds it (min of the
// running value and the proposal; a NaN proposal leaves it).  std::max and
// std::min keep their argument order: (a < b) ? b : a and (b < a) ? b : a.
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

__device__ __forceinline__ double terr(int q, const double* s0, const double* s1, const double* s2,
                                       const double* s3, int B, int inst, const DevCtx& c,
                                       const TruncOpts& o, double dt) {
    const int order = c.order;
    const size_t iq = (size_t)q * B + inst, ic = iq + B;   // ccap = qcap + 1
    const double volttol = o.abstol + o.reltol * smax(fabs(s0[ic]), fabs(s1[ic]));
    double chargetol = smax(fabs(s0[iq]), fabs(s1[iq]));
    chargetol = o.reltol * smax(chargetol, o.chgtol) / c.delta;
    const double tol = smax(volttol, chargetol);
    // divided differences of the charge history (order <= 2: four states)
    const double* st[4] = {s0, s1, s2, s3};
    const double dold[3] = {c.delta, c.delta_old1, c.delta_old2};   // CKTdeltaOld[0] is CKTdelta
    double diff[4], deltmp[3];
    for (int i = order + 1; i >= 0; --i) diff[i] = st[i][iq];
    for (int i = 0; i <= order; ++i) deltmp[i] = dold[i];
    int j = order;
    for (;;) {
        for (int i = 0; i <= j; ++i) diff[i] = (diff[i] - diff[i + 1]) / deltmp[i];
        if (--j < 0) break;
        for (int i = 0; i <= j; ++i) deltmp[i] = deltmp[i + 1] + dold[i];
    }
    // LTE coefficients of the integration methods (the reference's tables)
    const double gear[2] = {0.5, 0.2222222222};
    const double trap[2] = {0.5, 0.08333333333};
    const double factor = c.method == 1 ? gear[order - 1] : trap[order - 1];
    double del = o.trtol * tol / smax(o.abstol, factor * fabs(diff[0]));
    if (order == 2) del = sqrt(del);
    return (del < dt) ? del : dt;
}

// compute_trunc of a BSIM4v7 device at context c: its charges qb, qg, qd
// (state slots 11, 13, 15 from the device's base; bsim4v7_device.cpp:
// 1015-1017); 1e30 where the reference's guard skips it.
__device__ __forceinline__ double b4_trunc(int base, const double* s0, const double* s1, const double* s2,
                                           const double* s3, int B, int inst, const DevCtx& c,
                                           const TruncOpts& o) {
    double dt = 1e30;
    // compute_trunc's guard; orders above 2 would need more than four states
    if (c.order >= 1 && c.order <= 2 && c.delta > 0.0) {
        dt = terr(base + 11, s0, s1, s2, s3, B, inst, c, o, dt);
        dt = terr(base + 13, s0, s1, s2, s3, B, inst, c, o, dt);
        dt = terr(base + 15, s0, s1, s2, s3, B, inst, c, o, dt);
    }
    return dt;
}

// The minimum of v over the lanes of this wave that share ci, folded into
// dt_min[ci] with one atomic per distinct ci (devices of one instance sit
// in consecutive lanes, so a wave usually holds one or two instances).
// Every lane of the wave must call it (all 64 active); lanes with ci < 0
// take part in the reductions and update nothing.
__device__ inline void seg_atomic_min(unsigned long long* dt_min, int ci, unsigned long long v) {
    unsigned long long todo = __ballot(1);
    const int lane = __lane_id();
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int lci = __shfl(ci, leader);
        const bool mine = (ci == lci);
        unsigned long long m = mine ? v : ~0ull;
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned lo = __shfl_xor((unsigned)(m & 0xffffffffu), o);
            const unsigned hi = __shfl_xor((unsigned)(m >> 32), o);
            const unsigned long long w = ((unsigned long long)hi << 32) | lo;
            m = (w < m) ? w : m;
        }
        if (lane == leader && lci >= 0) atomicMin(&dt_min[lci], m);
        todo &= ~__ballot(mine);
    }
}

// One lane of the truncation kernel: device (idx / B) of instance slot
// (idx % B) of table t proposes its step at its context (and at order 2
// into dt_min2), folded into dt_min per instance.  Lanes past the table and
// lanes of instances sel leaves out take part in the wave's reductions with
// no instance, so every lane of the wave must call it.
__device__ inline void trunc_lane(const TruncTable& t, const DevCtx* __restrict__ ctx,
                                  const double* __restrict__ s0, const double* __restrict__ s1,
                                  const double* __restrict__ s2, const double* __restrict__ s3, int B,
                                  const TruncOpts& o, double* __restrict__ dt_dev,
                                  unsigned long long* __restrict__ dt_min, int order_override,
                                  unsigned long long* __restrict__ dt_min2, const int32_t* __restrict__ sel,
                                  int idx) {
    bool valid = idx < t.ndev * B;
    const int inst = valid ? idx % B : 0, dev = valid ? idx / B : 0;
    int ci = !valid ? -1 : (t.inst_of ? t.inst_of[idx] : inst);
    if (valid && sel && !sel[ci]) {
        valid = false;
        ci = -1;
    }
    DevCtx c = ctx[valid ? ci : 0];
    if (order_override > 0) c.order = order_override;
    // the device's proposal at context c (compute_trunc's body)
    auto device_dt = [&](const DevCtx& cx) {
        double dt = 1e30;
        // compute_trunc's guard; orders above 2 would need more than four states
        if (cx.order >= 1 && cx.order <= 2 && cx.delta > 0.0) {
            const int base = t.base[dev];
            if (t.kind[dev] == 0) {
                dt = terr(base, s0, s1, s2, s3, B, inst, cx, o, dt);
            } else if (t.kind[dev] == 2) {
                dt = terr(base + 3, s0, s1, s2, s3, B, inst, cx, o, dt);   // DIOcapCharge
            } else if (t.kind[dev] == 3) {
                dt = b4_trunc(base, s0, s1, s2, s3, B, inst, cx, o);
            } else {
                const int qoff[5] = {5, 8, 11, 13, 15};   // qgs, qgd, qgb, qbd, qbs
                for (int k = 0; k < 5; ++k) dt = terr(base + qoff[k], s0, s1, s2, s3, B, inst, cx, o, dt);
            }
        }
        return dt;
    };
    const double dt = valid ? device_dt(c) : 1e30;
    if (dt_dev && valid) dt_dev[idx] = dt;
    // dt >= 0: the bit patterns of non-negative doubles order like the values
    seg_atomic_min(dt_min, ci, (unsigned long long)__double_as_longlong(dt));
    if (dt_min2) {
        // the same states at order 2 (the order-promotion test)
        c.order = 2;
        const double dt2 = valid ? device_dt(c) : 1e30;
        seg_atomic_min(dt_min2, ci, (unsigned long long)__double_as_longlong(dt2));
    }
}

}  // namespace nsa
