// BSIM4v7load (src/devices/bsim4v7/bsim4v7_load.cpp:69-5104) for the
// default-mode cards of bsim4.hpp, one lane per (device, instance) with the
// instance index fastest: the lanes of a wave share one device's constant
// record (a broadcast read) and touch consecutive instances of every state,
// site and solution entry.
//
// The lane restates the reference's statements in its operation order,
// without FMA contraction.  Terms that are identically zero under the
// covered cards (substrate and GIDL/GISL currents, gate tunnelling, junction
// tunnelling, external S/D and gate resistances, NQS, the poly-depletion
// shift) are left out: each enters the reference's sums as an added or
// subtracted zero, which leaves every sum unchanged.  For the same reason
// the stamps whose value is identically zero (the D/S external-node
// entries, GIDL and GISL) are not written.  Bypass is disabled in the
// reference (CKTbypass = 0, bsim4v7_device.cpp:279).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

#include "bsim4.hpp"
#include "common.hpp"
#include "dev_limits.hpp"
#include "dual.hpp"
#include "gpu_dev.hpp"
#include "dev_src.hpp"

namespace nsa {

namespace {

using namespace devk;

constexpr int TPB = 256;
inline int blocks_for(long work) { return (int)((work + TPB - 1) / TPB); }

// bsim4v7_load.cpp:37-50
constexpr double kMaxExp = 5.834617425e14;
constexpr double kMinExp = 1.713908431e-15;
constexpr double kExpThreshold = 34.0;
constexpr double kEpsSi = 1.03594e-10;
constexpr double kDelta1 = 0.02;
constexpr double kDelta3 = 0.02;
constexpr double kDelta4 = 0.02;
// CONSTvt0 (ucb_compat.hpp:12)
constexpr double kVt0 = 1.38064852e-23 * 300.15 / 1.6021766208e-19;

// field f of the device's constant record (pb, ps set in the kernel)
#define PF(f) P((int)(offsetof(B4Param, f) / sizeof(double)))
static_assert(kB4Fields == (int)(offsetof(B4Param, off) / sizeof(double)), "B4Param layout");

// UCB MAX (ucb_compat.hpp:38-43)
template <class T>
__device__ inline T ucb_max(const T& a, const T& b) { return (a > b) ? a : b; }

// The lane's live state (~140 doubles at the charge model) fits in 280
// VGPRs without spilling when the kernel may run one wave per SIMD; at the
// compiler's default occupancy target it spilled 680 bytes per lane to
// scratch, which doubled its HBM traffic and its duration.
template <class T>
__device__ __forceinline__ void b4_lane(const B4Table& t, const DevCtx* __restrict__ ctx,
                                        const double* __restrict__ x, double* __restrict__ s0,
                                        double* __restrict__ s1, const double* __restrict__ s2,
                                        const int32_t* __restrict__ done, double* __restrict__ cm,
                                        double* __restrict__ cr, int32_t* __restrict__ devnc, DevTol tol,
                                        int B, const double* __restrict__ dpar,
                                        const double* __restrict__ lam, double* __restrict__ gout) {
    // the gradient form: dual numbers seeded by d(record)/dp, DC adjoint only
    constexpr bool kGrad = is_dual<T>::value;
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)t.count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    const int ci = t.inst_of ? t.inst_of[gid] : inst;
    if (done && done[ci]) return;
    const size_t sB = (size_t)B;
    const DevCtx c = ctx[ci];
    // the device's constants: one shared record, or per-instance records
    // stored field-major (a wave reads 64 consecutive instances of a field)
    const double* pb = t.par_inst ? t.par_inst + (size_t)d * kB4Fields * sB + inst
                                  : reinterpret_cast<const double*>(t.par + d);
    const size_t ps = t.par_inst ? sB : 1;
    const double* dpb = kGrad ? dpar + (size_t)d * kB4Fields * sB + inst : nullptr;
    auto P = [&](int q) -> T {
        T v = pb[(size_t)q * ps];
        if constexpr (kGrad) v.d[0] = dpb[(size_t)q * sB];
        return v;
    };
    const int mode = c.mode;
    // stores of the load (none in the gradient form)
    auto put = [&](double* a, size_t k, const T& v) {
        if constexpr (!kGrad) a[k] = value_of(v);
    };
    const size_t st = (size_t)t.state[d] * sB + inst;
    auto S = [&](int k) { return st + (size_t)k * sB; };
    // state slots (bsim4v7_def.hpp:514-545)
    enum { vbdS = 0, vbsS = 1, vgsS = 2, vdsS = 3, qbS = 11, cqbS = 12, qgS = 13, cqgS = 14,
           qdS = 15, cqdS = 16, qbsS = 19, qbdS = 21, qsS = 28 };
    double* aux = t.aux + (size_t)d * kB4Aux * sB + inst;
    const int nD = t.nodes[4 * d], nG = t.nodes[4 * d + 1], nS = t.nodes[4 * d + 2],
              nB = t.nodes[4 * d + 3];
    // terminal k's voltage: node of the device (batch of one circuit) or the
    // lane's own solution index (packed batch of circuits)
    auto V = [&](int node, int k) -> T {
        if (t.pin) {
            const int q = t.pin[(size_t)k * sB + inst];
            return q >= 0 ? x[q] : 0.0;
        }
        return sol(x, node, sB, inst);
    };
    const T type = PF(type);
    const bool charge = !kGrad && ((mode & (MODEDCTRANCURVE | MODEAC | MODETRAN | MODEINITSMSIG)) ||
                         ((mode & MODETRANOP) && (mode & MODEUIC)));

    // instance fields carried from the previous load
    const T h_von = aux[B4_VON * sB];
    const T h_mode = aux[B4_MODE * sB];
    const T h_cd = aux[B4_CD * sB], h_cbs = aux[B4_CBS * sB], h_cbd = aux[B4_CBD * sB];
    const T h_gbs = aux[B4_GBS * sB], h_gbd = aux[B4_GBD * sB];
    const T h_gm = aux[B4_GM * sB], h_gds = aux[B4_GDS * sB], h_gmbs = aux[B4_GMBS * sB];

    // ---- terminal voltages (bsim4v7_load.cpp:237-663)
    int Check = 1;
    T vds, vgs, vbs, vbd, vgd;
    T cdhat = 0.0, cbhat = 0.0;
    if (kGrad || (mode & MODEADJOINT)) {
        // the converged solution itself, unlimited (the adjoint's Jacobian)
        vds = type * (V(nD, 0) - V(nS, 2));
        vgs = type * (V(nG, 1) - V(nS, 2));
        vbs = type * (V(nB, 3) - V(nS, 2));
        Check = 0;
    } else {
    if constexpr (!kGrad) {
    if (mode & MODEINITTRAN) {
        vds = s1[S(vdsS)];
        vgs = s1[S(vgsS)];
        vbs = s1[S(vbsS)];
    } else if (mode & MODEINITJCT) {
        // no IC= on the instance: icVDS = icVGS = icVBS = 0
        vds = type * 0.0;
        vgs = type * 0.0;
        vbs = type * 0.0;
        if ((vds == 0.0) && (vgs == 0.0) && (vbs == 0.0) &&
            ((mode & (MODETRAN | MODEAC | MODEDCOP | MODEDCTRANCURVE)) || (!(mode & MODEUIC)))) {
            vds = 0.1;
            vgs = type * PF(vth0) + 0.1;
            vbs = 0.0;
        }
    } else {
        if (mode & MODEINITPRED) {
            const T xfact = c.delta / c.delta_old1;
            put(s0, S(vdsS), s1[S(vdsS)]);
            vds = (1.0 + xfact) * (s1[S(vdsS)]) - (xfact * (s2[S(vdsS)]));
            put(s0, S(vgsS), s1[S(vgsS)]);
            vgs = (1.0 + xfact) * (s1[S(vgsS)]) - (xfact * (s2[S(vgsS)]));
            put(s0, S(vbsS), s1[S(vbsS)]);
            vbs = (1.0 + xfact) * (s1[S(vbsS)]) - (xfact * (s2[S(vbsS)]));
            put(s0, S(vbdS), s0[S(vbsS)] - s0[S(vdsS)]);
        } else {
            vds = type * (V(nD, 0) - V(nS, 2));
            vgs = type * (V(nG, 1) - V(nS, 2));
            vbs = type * (V(nB, 3) - V(nS, 2));
        }
        const T s0vds = s0[S(vdsS)], s0vgs = s0[S(vgsS)], s0vbs = s0[S(vbsS)],
                     s0vbd = s0[S(vbdS)];
        const T vgdo = s0vgs - s0vds;
        vbd = vbs - vds;
        vgd = vgs - vds;
        const T delvbd = vbd - s0vbd;
        const T delvgd = vgd - vgdo;
        const T delvds = vds - s0vds;
        const T delvgs = vgs - s0vgs;
        const T delvbs = vbs - s0vbs;
        // the linearised currents of the previous load at this iterate, for
        // the current-based convergence test (:421-473)
        if (h_mode >= 0) {
            const T Idtot = h_cd - h_cbd;
            cdhat = Idtot - h_gbd * delvbd + h_gmbs * delvbs + h_gm * delvgs + h_gds * delvds;
            const T Ibtot = h_cbs + h_cbd;
            cbhat = Ibtot + h_gbd * delvbd + h_gbs * delvbs;
        } else {
            const T Idtot = h_cd + h_cbd;
            cdhat = Idtot + h_gbd * delvbd + h_gmbs * delvbd + h_gm * delvgd - h_gds * delvds;
            const T Ibtot = h_cbs + h_cbd;
            cbhat = Ibtot + h_gbs * delvbs + h_gbd * delvbd;
        }
        // limiting (:582-662)
        const T von = h_von;
        if (s0vds >= 0.0) {
            vgs = fetlim(vgs, s0vgs, von);
            vds = vgs - vgd;
            vds = limvds(vds, s0vds);
            vgd = vgs - vds;
        } else {
            vgd = fetlim(vgd, vgdo, von);
            vds = vgs - vgd;
            vds = -limvds(-vds, -(s0vds));
            vgs = vgd + vds;
        }
        if (vds >= 0.0) {
            vbs = pnjlim(vbs, s0vbs, kVt0, PF(vcrit), &Check);
            vbd = vbs - vds;
        } else {
            vbd = pnjlim(vbd, s0vbd, kVt0, PF(vcrit), &Check);
            vbs = vbd + vds;
        }
    }
    }
    }
    vbd = vbs - vds;
    vgd = vgs - vds;
    const T vgb = vgs - vbs;
    const T vbs_jct = vbs, vbd_jct = vbd;

    // ---- junction currents: zero saturation current, gmin only (:682-795)
    const T gbs = dev_gmin(tol, ci), cbs = gbs * vbs_jct;
    const T gbd = dev_gmin(tol, ci), cbd = gbd * vbd_jct;

    int dmode;
    T Vds, Vbs;
    if (vds >= 0.0) {
        dmode = 1;
        Vds = vds;
        Vbs = vbs;
    } else {
        dmode = -1;
        Vds = -vds;
        Vbs = vbd;
    }
    const T toxe = PF(toxe), epssub = kEpsSi;

    // ---- Vbseff, Phis, Xdep (:1013-1038)
    T T0, T1, T2, T3, T4, T5, T6, T7, T8, T9, T10, T11, T12, T13;
    T Vbseff, dVbseff_dVb;
    T0 = Vbs - PF(vbsc) - 0.001;
    T1 = sqrt(T0 * T0 - 0.004 * PF(vbsc));
    if (T0 >= 0.0) {
        Vbseff = PF(vbsc) + 0.5 * (T0 + T1);
        dVbseff_dVb = 0.5 * (1.0 + T0 / T1);
    } else {
        T2 = -0.002 / (T1 - T0);
        Vbseff = PF(vbsc) * (1.0 + T2);
        dVbseff_dVb = T2 * PF(vbsc) / T1;
    }
    T9 = 0.95 * PF(phi);
    T0 = T9 - Vbseff - 0.001;
    T1 = sqrt(T0 * T0 + 0.004 * T9);
    Vbseff = T9 - 0.5 * (T0 + T1);
    dVbseff_dVb *= 0.5 * (1.0 + T0 / T1);
    const T Phis = PF(phi) - Vbseff;
    const T dPhis_dVb = -1.0;
    const T sqrtPhis = sqrt(Phis);
    const T dsqrtPhis_dVb = -0.5 / sqrtPhis;
    const T Xdep = PF(Xdep0) * sqrtPhis / PF(sqrtPhi);
    const T dXdep_dVb = (PF(Xdep0) / PF(sqrtPhi)) * dsqrtPhis_dVb;
    const T Leff = PF(leff);
    const T Vtm = PF(vtm);

    // ---- Vth (:1044-1141)
    T3 = sqrt(Xdep);
    const T V0 = PF(vbi) - PF(phi);
    T0 = PF(dvt2) * Vbseff;
    if (T0 >= -0.5) {
        T1 = 1.0 + T0;
        T2 = PF(dvt2);
    } else {
        T4 = 1.0 / (3.0 + 8.0 * T0);
        T1 = (1.0 + 3.0 * T0) * T4;
        T2 = PF(dvt2) * T4 * T4;
    }
    const T lt1 = PF(factor1) * T3 * T1;
    const T dlt1_dVb = PF(factor1) * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    T0 = PF(dvt2w) * Vbseff;
    if (T0 >= -0.5) {
        T1 = 1.0 + T0;
        T2 = PF(dvt2w);
    } else {
        T4 = 1.0 / (3.0 + 8.0 * T0);
        T1 = (1.0 + 3.0 * T0) * T4;
        T2 = PF(dvt2w) * T4 * T4;
    }
    const T ltw = PF(factor1) * T3 * T1;
    const T dltw_dVb = PF(factor1) * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    T Theta0, dTheta0_dVb;
    T0 = PF(dvt1) * Leff / lt1;
    if (T0 < kExpThreshold) {
        T1 = exp(T0);
        T2 = T1 - 1.0;
        T3 = T2 * T2;
        T4 = T3 + 2.0 * T1 * kMinExp;
        Theta0 = T1 / T4;
        const T dT1_dVb = -T0 * T1 * dlt1_dVb / lt1;
        dTheta0_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + kMinExp)) / T4 / T4;
    } else {
        Theta0 = 1.0 / (kMaxExp - 2.0);
        dTheta0_dVb = 0.0;
    }
    const T thetavth = PF(dvt0) * Theta0;
    const T Delt_vth = thetavth * V0;
    const T dDelt_vth_dVb = PF(dvt0) * dTheta0_dVb * V0;
    T dT5_dVb;
    T0 = PF(dvt1w) * PF(weff) * Leff / ltw;
    if (T0 < kExpThreshold) {
        T1 = exp(T0);
        T2 = T1 - 1.0;
        T3 = T2 * T2;
        T4 = T3 + 2.0 * T1 * kMinExp;
        T5 = T1 / T4;
        const T dT1_dVb = -T0 * T1 * dltw_dVb / ltw;
        dT5_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + kMinExp)) / T4 / T4;
    } else {
        T5 = 1.0 / (kMaxExp - 2.0);
        dT5_dVb = 0.0;
    }
    T0 = PF(dvt0w) * T5;
    T2 = T0 * V0;
    const T dT2_dVb_w = PF(dvt0w) * dT5_dVb * V0;
    const T TempRatio = PF(temp_ratio);
    T0 = sqrt(1.0 + PF(lpe0) / Leff);
    T1 = PF(k1ox) * (T0 - 1.0) * PF(sqrtPhi) + (PF(kt1) + PF(kt1l) / Leff + PF(kt2) * Vbseff) * TempRatio;
    const T Vth_NarrowW = toxe * PF(phi) / (PF(weff) + PF(w0));
    T3 = PF(eta0) + PF(etab) * Vbseff;
    if (T3 < 1.0e-4) {
        T9 = 1.0 / (3.0 - 2.0e4 * T3);
        T3 = (2.0e-4 - T3) * T9;
        T4 = T9 * T9;
    } else {
        T4 = 1.0;
    }
    const T dDIBL_Sft_dVd = T3 * PF(theta0vb0);
    const T DIBL_Sft = dDIBL_Sft_dVd * Vds;
    const T Lpe_Vb = sqrt(1.0 + PF(lpeb) / Leff);
    const T Vth = type * PF(vth0) + (PF(k1ox) * sqrtPhis - PF(k1) * PF(sqrtPhi)) * Lpe_Vb -
                       PF(k2ox) * Vbseff - Delt_vth - T2 + (PF(k3) + PF(k3b) * Vbseff) * Vth_NarrowW +
                       T1 - DIBL_Sft;
    const T dVth_dVb = Lpe_Vb * PF(k1ox) * dsqrtPhis_dVb - PF(k2ox) - dDelt_vth_dVb - dT2_dVb_w +
                            PF(k3b) * Vth_NarrowW - PF(etab) * Vds * PF(theta0vb0) * T4 +
                            PF(kt2) * TempRatio;
    const T dVth_dVd = -dDIBL_Sft_dVd;

    // ---- n (:1145-1165)
    T n, dn_dVb, dn_dVd;
    {
        const T tmp1 = epssub / Xdep;
        const T tmp2 = PF(nfactor) * tmp1;
        const T tmp3 = PF(cdsc) + PF(cdscb) * Vbseff + PF(cdscd) * Vds;
        const T tmp4 = (tmp2 + tmp3 * Theta0 + PF(cit)) / PF(coxe);
        if (tmp4 >= -0.5) {
            n = 1.0 + tmp4;
            dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + PF(cdscb) * Theta0) / PF(coxe);
            dn_dVd = PF(cdscd) * Theta0 / PF(coxe);
        } else {
            T0 = 1.0 / (3.0 + 8.0 * tmp4);
            n = (1.0 + 3.0 * tmp4) * T0;
            T0 *= T0;
            dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + PF(cdscb) * Theta0) / PF(coxe) * T0;
            dn_dVd = PF(cdscd) * Theta0 / PF(coxe) * T0;
        }
    }
    const T von_new = Vth;

    // ---- Vgsteff (:1222-1307); no poly depletion (NGATE = 0)
    const T Vgs_eff = (dmode > 0) ? vgs : vgd;
    const T dVgs_eff_dVg = 1.0;
    const T Vgst = Vgs_eff - Vth;
    T dT10_dVg, dT10_dVd, dT10_dVb, dT9_dVg, dT9_dVd, dT9_dVb;
    T0 = n * Vtm;
    T1 = PF(mstar) * Vgst;
    T2 = T1 / T0;
    if (T2 > kExpThreshold) {
        T10 = T1;
        dT10_dVg = PF(mstar) * dVgs_eff_dVg;
        dT10_dVd = -dVth_dVd * PF(mstar);
        dT10_dVb = -dVth_dVb * PF(mstar);
    } else if (T2 < -kExpThreshold) {
        T10 = Vtm * log(1.0 + kMinExp);
        dT10_dVg = 0.0;
        dT10_dVd = T10 * dn_dVd;
        dT10_dVb = T10 * dn_dVb;
        T10 *= n;
    } else {
        const T ExpVgst = exp(T2);
        T3 = Vtm * log(1.0 + ExpVgst);
        T10 = n * T3;
        dT10_dVg = PF(mstar) * ExpVgst / (1.0 + ExpVgst);
        dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
        dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
        dT10_dVg *= dVgs_eff_dVg;
    }
    T1 = PF(voffcbn) - (1.0 - PF(mstar)) * Vgst;
    T2 = T1 / T0;
    if (T2 < -kExpThreshold) {
        T3 = PF(coxe) * kMinExp / PF(cdep0);
        T9 = PF(mstar) + T3 * n;
        dT9_dVg = 0.0;
        dT9_dVd = dn_dVd * T3;
        dT9_dVb = dn_dVb * T3;
    } else if (T2 > kExpThreshold) {
        T3 = PF(coxe) * kMaxExp / PF(cdep0);
        T9 = PF(mstar) + T3 * n;
        dT9_dVg = 0.0;
        dT9_dVd = dn_dVd * T3;
        dT9_dVb = dn_dVb * T3;
    } else {
        const T ExpVgst = exp(T2);
        T3 = PF(coxe) / PF(cdep0);
        T4 = T3 * ExpVgst;
        T5 = T1 * T4 / T0;
        T9 = PF(mstar) + n * T4;
        dT9_dVg = T3 * (PF(mstar) - 1.0) * ExpVgst / Vtm;
        dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
        dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
        dT9_dVg *= dVgs_eff_dVg;
    }
    T Vgsteff = T10 / T9;
    T11 = T9 * T9;
    T dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
    T dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
    T dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;

    // ---- effective width, Rds (:1310-1347)
    T9 = sqrtPhis - PF(sqrtPhi);
    T Weff = PF(weff) - 2.0 * (PF(dwg) * Vgsteff + PF(dwb) * T9);
    T dWeff_dVg = -2.0 * PF(dwg);
    T dWeff_dVb = -2.0 * PF(dwb) * dsqrtPhis_dVb;
    if (Weff < 2.0e-8) {
        T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
        Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
        T0 *= T0 * 4.0e-16;
        dWeff_dVg *= T0;
        dWeff_dVb *= T0;
    }
    T Rds, dRds_dVg, dRds_dVb;
    {
        T0 = 1.0 + PF(prwg) * Vgsteff;
        const T dT0_dVg = -PF(prwg) / T0 / T0;
        T1 = PF(prwb) * T9;
        const T dT1_dVb = PF(prwb) * dsqrtPhis_dVb;
        T2 = 1.0 / T0 + T1;
        T3 = T2 + sqrt(T2 * T2 + 0.01);
        T dT3_dVg = 1.0 + T2 / (T3 - T2);
        const T dT3_dVb = dT3_dVg * dT1_dVb;
        dT3_dVg *= dT0_dVg;
        T4 = PF(rds0) * 0.5;
        Rds = PF(rdswmin) + T3 * T4;
        dRds_dVg = T4 * dT3_dVg;
        dRds_dVb = T4 * dT3_dVb;
    }

    // ---- Abulk (:1350-1403)
    T Abulk, Abulk0, dAbulk_dVg, dAbulk_dVb, dAbulk0_dVb;
    {
        T9 = 0.5 * PF(k1ox) * Lpe_Vb / sqrtPhis;
        T1 = T9 + PF(k2ox) - PF(k3b) * Vth_NarrowW;
        const T dT1_dVb = -T9 / sqrtPhis * dsqrtPhis_dVb;
        T9 = sqrt(PF(xj) * Xdep);
        const T tmp1 = Leff + 2.0 * T9;
        T5 = Leff / tmp1;
        const T tmp2 = PF(a0) * T5;
        const T tmp3 = PF(weff) + PF(b1);
        const T tmp4 = PF(b0) / tmp3;
        T2 = tmp2 + tmp4;
        const T dT2_dVb = -T9 / tmp1 / Xdep * dXdep_dVb;
        T6 = T5 * T5;
        T7 = T5 * T6;
        Abulk0 = 1.0 + T1 * T2;
        dAbulk0_dVb = T1 * tmp2 * dT2_dVb + T2 * dT1_dVb;
        T8 = PF(ags) * PF(a0) * T7;
        dAbulk_dVg = -T1 * T8;
        Abulk = Abulk0 + dAbulk_dVg * Vgsteff;
        dAbulk_dVb = dAbulk0_dVb - T8 * Vgsteff * (dT1_dVb + 3.0 * T1 * dT2_dVb);
    }
    if (Abulk0 < 0.1) {
        T9 = 1.0 / (3.0 - 20.0 * Abulk0);
        Abulk0 = (0.2 - Abulk0) * T9;
        dAbulk0_dVb *= T9 * T9;
    }
    if (Abulk < 0.1) {
        T9 = 1.0 / (3.0 - 20.0 * Abulk);
        Abulk = (0.2 - Abulk) * T9;
        T10 = T9 * T9;
        dAbulk_dVb *= T10;
        dAbulk_dVg *= T10;
    }
    {
        T dT0_dVb;
        T2 = PF(keta) * Vbseff;
        if (T2 >= -0.9) {
            T0 = 1.0 / (1.0 + T2);
            dT0_dVb = -PF(keta) * T0 * T0;
        } else {
            T1 = 1.0 / (0.8 + T2);
            T0 = (17.0 + 20.0 * T2) * T1;
            dT0_dVb = -PF(keta) * T1 * T1;
        }
        dAbulk_dVg *= T0;
        dAbulk_dVb = dAbulk_dVb * T0 + Abulk * dT0_dVb;
        dAbulk0_dVb = dAbulk0_dVb * T0 + Abulk0 * dT0_dVb;
        Abulk *= T0;
        Abulk0 *= T0;
    }

    // ---- mobility, mobMod = 0 (:1405-1514)
    T ueff, dueff_dVg, dueff_dVd, dueff_dVb;
    {
        const T T14 = 0.0;
        T0 = Vgsteff + Vth + Vth - T14;
        T2 = PF(ua) + PF(uc) * Vbseff;
        T3 = T0 / toxe;
        T12 = sqrt(Vth * Vth + 0.0001);
        T9 = 1.0 / (Vgsteff + 2 * T12);
        T10 = T9 * toxe;
        T8 = PF(ud) * T10 * T10 * Vth;
        T6 = T8 * Vth;
        T5 = T3 * (T2 + PF(ub) * T3) + T6;
        T7 = -2.0 * T6 * T9;
        T11 = T7 * Vth / T12;
        T dDenomi_dVg = (T2 + 2.0 * PF(ub) * T3) / toxe;
        T13 = 2.0 * (dDenomi_dVg + T11 + T8);
        T dDenomi_dVd = T13 * dVth_dVd;
        T dDenomi_dVb = T13 * dVth_dVb + PF(uc) * T3;
        dDenomi_dVg += T7;
        T Denomi;
        if (T5 >= -0.8) {
            Denomi = 1.0 + T5;
        } else {
            T9 = 1.0 / (7.0 + 10.0 * T5);
            Denomi = (0.6 + T5) * T9;
            T9 *= T9;
            dDenomi_dVg *= T9;
            dDenomi_dVd *= T9;
            dDenomi_dVb *= T9;
        }
        ueff = PF(u0temp) / Denomi;
        T9 = -ueff / Denomi;
        dueff_dVg = T9 * dDenomi_dVg;
        dueff_dVd = T9 * dDenomi_dVd;
        dueff_dVb = T9 * dDenomi_dVb;
    }

    // ---- Vdsat (:1517-1616); A1 = 0: Lambda = A2
    const T WVCox = Weff * PF(vsattemp) * PF(coxe);
    const T WVCoxRds = WVCox * Rds;
    const T Esat = 2.0 * PF(vsattemp) / ueff;
    const T EsatL = Esat * Leff;
    T0 = -EsatL / ueff;
    const T dEsatL_dVg = T0 * dueff_dVg;
    const T dEsatL_dVd = T0 * dueff_dVd;
    const T dEsatL_dVb = T0 * dueff_dVb;
    const T Lambda = PF(a2);
    const T dLambda_dVg = 0.0;
    const T Vgst2Vtm = Vgsteff + 2.0 * Vtm;
    T tmp1, tmp2, tmp3;
    if (Rds > 0) {
        tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff;
        tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff;
    } else {
        tmp2 = dWeff_dVg / Weff;
        tmp3 = dWeff_dVb / Weff;
    }
    T Vdsat, dVdsat_dVg, dVdsat_dVd, dVdsat_dVb;
    if ((Rds == 0.0) && (Lambda == 1.0)) {
        T0 = 1.0 / (Abulk * EsatL + Vgst2Vtm);
        tmp1 = 0.0;
        T1 = T0 * T0;
        T2 = Vgst2Vtm * T0;
        T3 = EsatL * Vgst2Vtm;
        Vdsat = T3 * T0;
        const T dT0_dVg = -(Abulk * dEsatL_dVg + EsatL * dAbulk_dVg + 1.0) * T1;
        const T dT0_dVd = -(Abulk * dEsatL_dVd) * T1;
        const T dT0_dVb = -(Abulk * dEsatL_dVb + dAbulk_dVb * EsatL) * T1;
        dVdsat_dVg = T3 * dT0_dVg + T2 * dEsatL_dVg + EsatL * T0;
        dVdsat_dVd = T3 * dT0_dVd + T2 * dEsatL_dVd;
        dVdsat_dVb = T3 * dT0_dVb + T2 * dEsatL_dVb;
    } else {
        tmp1 = dLambda_dVg / (Lambda * Lambda);
        T9 = Abulk * WVCoxRds;
        T8 = Abulk * T9;
        T7 = Vgst2Vtm * T9;
        T6 = Vgst2Vtm * WVCoxRds;
        T0 = 2.0 * Abulk * (T9 - 1.0 + 1.0 / Lambda);
        const T dT0_dVg = 2.0 * (T8 * tmp2 - Abulk * tmp1 +
                                      (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
        const T dT0_dVb = 2.0 * (T8 * (2.0 / Abulk * dAbulk_dVb + tmp3) +
                                      (1.0 / Lambda - 1.0) * dAbulk_dVb);
        const T dT0_dVd = 0.0;
        T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + Abulk * EsatL + 3.0 * T7;
        const T dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + Abulk * dEsatL_dVg +
                               EsatL * dAbulk_dVg + 3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
        const T dT1_dVb =
            Abulk * dEsatL_dVb + EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
        const T dT1_dVd = Abulk * dEsatL_dVd;
        T2 = Vgst2Vtm * (EsatL + 2.0 * T6);
        const T dT2_dVg = EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
        const T dT2_dVb = Vgst2Vtm * (dEsatL_dVb + 2.0 * T6 * tmp3);
        const T dT2_dVd = Vgst2Vtm * dEsatL_dVd;
        T3 = sqrt(T1 * T1 - 2.0 * T0 * T2);
        Vdsat = (T1 - T3) / T0;
        dVdsat_dVg = (dT1_dVg - (T1 * dT1_dVg - dT0_dVg * T2 - T0 * dT2_dVg) / T3 - Vdsat * dT0_dVg) / T0;
        dVdsat_dVb = (dT1_dVb - (T1 * dT1_dVb - dT0_dVb * T2 - T0 * dT2_dVb) / T3 - Vdsat * dT0_dVb) / T0;
        dVdsat_dVd = (dT1_dVd - (T1 * dT1_dVd - T0 * dT2_dVd) / T3) / T0;
        (void)dT0_dVd;
    }

    // ---- Vdseff (:1619-1657)
    T Vdseff, dVdseff_dVg, dVdseff_dVd, dVdseff_dVb;
    {
        T1 = Vdsat - Vds - PF(delta);
        const T dT1_dVg = dVdsat_dVg;
        const T dT1_dVd = dVdsat_dVd - 1.0;
        const T dT1_dVb = dVdsat_dVb;
        T2 = sqrt(T1 * T1 + 4.0 * PF(delta) * Vdsat);
        T0 = T1 / T2;
        T9 = 2.0 * PF(delta);
        T3 = T9 / T2;
        const T dT2_dVg = T0 * dT1_dVg + T3 * dVdsat_dVg;
        const T dT2_dVd = T0 * dT1_dVd + T3 * dVdsat_dVd;
        const T dT2_dVb = T0 * dT1_dVb + T3 * dVdsat_dVb;
        if (T1 >= 0.0) {
            Vdseff = Vdsat - 0.5 * (T1 + T2);
            dVdseff_dVg = dVdsat_dVg - 0.5 * (dT1_dVg + dT2_dVg);
            dVdseff_dVd = dVdsat_dVd - 0.5 * (dT1_dVd + dT2_dVd);
            dVdseff_dVb = dVdsat_dVb - 0.5 * (dT1_dVb + dT2_dVb);
        } else {
            T4 = T9 / (T2 - T1);
            T5 = 1.0 - T4;
            T6 = Vdsat * T4 / (T2 - T1);
            Vdseff = Vdsat * T5;
            dVdseff_dVg = dVdsat_dVg * T5 + T6 * (dT2_dVg - dT1_dVg);
            dVdseff_dVd = dVdsat_dVd * T5 + T6 * (dT2_dVd - dT1_dVd);
            dVdseff_dVb = dVdsat_dVb * T5 + T6 * (dT2_dVb - dT1_dVb);
        }
        if (Vds == 0.0) {
            Vdseff = 0.0;
            dVdseff_dVg = 0.0;
            dVdseff_dVb = 0.0;
        }
        if (Vdseff > Vds) Vdseff = Vds;
    }
    const T diffVds = Vds - Vdseff;

    // ---- Vasat (:1702-1724)
    T Vasat, dVasat_dVg, dVasat_dVb, dVasat_dVd;
    {
        const T tmp4 = 1.0 - 0.5 * Abulk * Vdsat / Vgst2Vtm;
        T9 = WVCoxRds * Vgsteff;
        T8 = T9 / Vgst2Vtm;
        T0 = EsatL + Vdsat + 2.0 * T9 * tmp4;
        T7 = 2.0 * WVCoxRds * tmp4;
        const T dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * Vgsteff) -
                               T8 * (Abulk * dVdsat_dVg - Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
        const T dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * Vgsteff -
                               T8 * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
        const T dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * Abulk * dVdsat_dVd;
        T9 = WVCoxRds * Abulk;
        T1 = 2.0 / Lambda - 1.0 + T9;
        const T dT1_dVg = -2.0 * tmp1 + WVCoxRds * (Abulk * tmp2 + dAbulk_dVg);
        const T dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
        Vasat = T0 / T1;
        dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
        dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
        dVasat_dVd = dT0_dVd / T1;
    }

    // ---- Idl (:1728-1785)
    T Idl, dIdl_dVg, dIdl_dVd, dIdl_dVb;
    {
        tmp1 = PF(vtfbphi2);
        tmp2 = 2.0e8 * PF(toxp);
        T dT0_dVg = 1.0 / tmp2;
        T0 = (Vgsteff + tmp1) * dT0_dVg;
        tmp3 = exp(PF(bdos) * 0.7 * log(T0));
        T1 = 1.0 + tmp3;
        T2 = PF(bdos) * 0.7 * tmp3 / T0;
        const T Tcen = PF(ados) * 1.9e-9 / T1;
        const T dTcen_dVg = -Tcen * T2 * dT0_dVg / T1;
        const T Coxeff = epssub * PF(coxp) / (epssub + PF(coxp) * Tcen);
        const T dCoxeff_dVg = -Coxeff * Coxeff * dTcen_dVg / epssub;
        const T CoxeffWovL = Coxeff * Weff / Leff;
        const T beta = ueff * CoxeffWovL;
        T3 = ueff / Leff;
        const T dbeta_dVg = CoxeffWovL * dueff_dVg + T3 * (Weff * dCoxeff_dVg + Coxeff * dWeff_dVg);
        const T dbeta_dVd = CoxeffWovL * dueff_dVd;
        const T dbeta_dVb = CoxeffWovL * dueff_dVb + T3 * Coxeff * dWeff_dVb;
        const T AbovVgst2Vtm = Abulk / Vgst2Vtm;
        T0 = 1.0 - 0.5 * Vdseff * AbovVgst2Vtm;
        dT0_dVg = -0.5 * (Abulk * dVdseff_dVg - Abulk * Vdseff / Vgst2Vtm + Vdseff * dAbulk_dVg) / Vgst2Vtm;
        const T dT0_dVd = -0.5 * Abulk * dVdseff_dVd / Vgst2Vtm;
        const T dT0_dVb = -0.5 * (Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff) / Vgst2Vtm;
        const T fgche1 = Vgsteff * T0;
        const T dfgche1_dVg = Vgsteff * dT0_dVg + T0;
        const T dfgche1_dVd = Vgsteff * dT0_dVd;
        const T dfgche1_dVb = Vgsteff * dT0_dVb;
        T9 = Vdseff / EsatL;
        const T fgche2 = 1.0 + T9;
        const T dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / EsatL;
        const T dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / EsatL;
        const T dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / EsatL;
        const T gche = beta * fgche1 / fgche2;
        const T dgche_dVg = (beta * dfgche1_dVg + fgche1 * dbeta_dVg - gche * dfgche2_dVg) / fgche2;
        const T dgche_dVd = (beta * dfgche1_dVd + fgche1 * dbeta_dVd - gche * dfgche2_dVd) / fgche2;
        const T dgche_dVb = (beta * dfgche1_dVb + fgche1 * dbeta_dVb - gche * dfgche2_dVb) / fgche2;
        T0 = 1.0 + gche * Rds;
        Idl = gche / T0;
        T1 = (1.0 - Idl * Rds) / T0;
        T2 = Idl * Idl;
        dIdl_dVg = T1 * dgche_dVg - T2 * dRds_dVg;
        dIdl_dVd = T1 * dgche_dVd;
        dIdl_dVb = T1 * dgche_dVb - T2 * dRds_dVb;
    }

    // ---- output resistance: FP = 1 (FPROUT = 0), VACLM, VADIBL (:1789-1897)
    const T FP = 1.0, dFP_dVg = 0.0;
    T PvagTerm, dPvagTerm_dVg, dPvagTerm_dVb, dPvagTerm_dVd;
    T8 = PF(pvag) / EsatL;
    T9 = T8 * Vgsteff;
    if (T9 > -0.9) {
        PvagTerm = 1.0 + T9;
        dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL);
        dPvagTerm_dVb = -T9 * dEsatL_dVb / EsatL;
        dPvagTerm_dVd = -T9 * dEsatL_dVd / EsatL;
    } else {
        T4 = 1.0 / (17.0 + 20.0 * T9);
        PvagTerm = (0.8 + T9) * T4;
        T4 *= T4;
        dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL) * T4;
        T9 *= T4 / EsatL;
        dPvagTerm_dVb = -T9 * dEsatL_dVb;
        dPvagTerm_dVd = -T9 * dEsatL_dVd;
    }
    T Cclm, dCclm_dVg, dCclm_dVd, dCclm_dVb, VACLM, dVACLM_dVg, dVACLM_dVd, dVACLM_dVb;
    if ((PF(pclm) > kMinExp) && (diffVds > 1.0e-10)) {
        T0 = 1.0 + Rds * Idl;
        const T dT0_dVg = dRds_dVg * Idl + Rds * dIdl_dVg;
        const T dT0_dVd = Rds * dIdl_dVd;
        const T dT0_dVb = dRds_dVb * Idl + Rds * dIdl_dVb;
        T2 = Vdsat / Esat;
        T1 = Leff + T2;
        const T dT1_dVg = (dVdsat_dVg - T2 * dEsatL_dVg / Leff) / Esat;
        const T dT1_dVd = (dVdsat_dVd - T2 * dEsatL_dVd / Leff) / Esat;
        const T dT1_dVb = (dVdsat_dVb - T2 * dEsatL_dVb / Leff) / Esat;
        Cclm = FP * PvagTerm * T0 * T1 / (PF(pclm) * PF(litl));
        dCclm_dVg = Cclm * (dFP_dVg / FP + dPvagTerm_dVg / PvagTerm + dT0_dVg / T0 + dT1_dVg / T1);
        dCclm_dVb = Cclm * (dPvagTerm_dVb / PvagTerm + dT0_dVb / T0 + dT1_dVb / T1);
        dCclm_dVd = Cclm * (dPvagTerm_dVd / PvagTerm + dT0_dVd / T0 + dT1_dVd / T1);
        VACLM = Cclm * diffVds;
        dVACLM_dVg = dCclm_dVg * diffVds - dVdseff_dVg * Cclm;
        dVACLM_dVb = dCclm_dVb * diffVds - dVdseff_dVb * Cclm;
        dVACLM_dVd = dCclm_dVd * diffVds + (1.0 - dVdseff_dVd) * Cclm;
    } else {
        VACLM = Cclm = kMaxExp;
        dVACLM_dVd = dVACLM_dVg = dVACLM_dVb = 0.0;
        dCclm_dVd = dCclm_dVg = dCclm_dVb = 0.0;
    }
    T VADIBL, dVADIBL_dVg, dVADIBL_dVb, dVADIBL_dVd;
    if (PF(thetaRout) > kMinExp) {
        T8 = Abulk * Vdsat;
        T0 = Vgst2Vtm * T8;
        const T dT0_dVg = Vgst2Vtm * Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
        const T dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
        const T dT0_dVd = Vgst2Vtm * Abulk * dVdsat_dVd;
        T1 = Vgst2Vtm + T8;
        const T dT1_dVg = 1.0 + Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
        const T dT1_dVb = Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
        const T dT1_dVd = Abulk * dVdsat_dVd;
        T9 = T1 * T1;
        T2 = PF(thetaRout);
        VADIBL = (Vgst2Vtm - T0 / T1) / T2;
        dVADIBL_dVg = (1.0 - dT0_dVg / T1 + T0 * dT1_dVg / T9) / T2;
        dVADIBL_dVb = (-dT0_dVb / T1 + T0 * dT1_dVb / T9) / T2;
        dVADIBL_dVd = (-dT0_dVd / T1 + T0 * dT1_dVd / T9) / T2;
        T7 = PF(pdiblb) * Vbseff;
        if (T7 >= -0.9) {
            T3 = 1.0 / (1.0 + T7);
            VADIBL *= T3;
            dVADIBL_dVg *= T3;
            dVADIBL_dVb = (dVADIBL_dVb - VADIBL * PF(pdiblb)) * T3;
            dVADIBL_dVd *= T3;
        } else {
            T4 = 1.0 / (0.8 + T7);
            T3 = (17.0 + 20.0 * T7) * T4;
            dVADIBL_dVg *= T3;
            dVADIBL_dVb = dVADIBL_dVb * T3 - VADIBL * PF(pdiblb) * T4 * T4;
            dVADIBL_dVd *= T3;
            VADIBL *= T3;
        }
        dVADIBL_dVg = dVADIBL_dVg * PvagTerm + VADIBL * dPvagTerm_dVg;
        dVADIBL_dVb = dVADIBL_dVb * PvagTerm + VADIBL * dPvagTerm_dVb;
        dVADIBL_dVd = dVADIBL_dVd * PvagTerm + VADIBL * dPvagTerm_dVd;
        VADIBL *= PvagTerm;
    } else {
        VADIBL = kMaxExp;
        dVADIBL_dVd = dVADIBL_dVg = dVADIBL_dVb = 0.0;
    }
    const T Va = Vasat + VACLM;
    const T dVa_dVg = dVasat_dVg + dVACLM_dVg;
    const T dVa_dVb = dVasat_dVb + dVACLM_dVb;
    const T dVa_dVd = dVasat_dVd + dVACLM_dVd;
    // VADITS with PDITS = 0 (:1905-1926)
    const T VADITS = kMaxExp, dVADITS_dVg = 0.0, dVADITS_dVd = 0.0;
    // VASCBE (:1929-1947)
    T VASCBE, dVASCBE_dVg, dVASCBE_dVd, dVASCBE_dVb;
    if ((PF(pscbe2) > 0.0) && (PF(pscbe1) >= 0.0)) {
        if (diffVds > PF(pscbe1) * PF(litl) / kExpThreshold) {
            T0 = PF(pscbe1) * PF(litl) / diffVds;
            VASCBE = Leff * exp(T0) / PF(pscbe2);
            T1 = T0 * VASCBE / diffVds;
            dVASCBE_dVg = T1 * dVdseff_dVg;
            dVASCBE_dVd = -T1 * (1.0 - dVdseff_dVd);
            dVASCBE_dVb = T1 * dVdseff_dVb;
        } else {
            VASCBE = kMaxExp * Leff / PF(pscbe2);
            dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
        }
    } else {
        VASCBE = kMaxExp;
        dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
    }

    // ---- Ids (:1950-2047)
    T Idsa, dIdsa_dVg, dIdsa_dVd, dIdsa_dVb;
    T9 = diffVds / VADIBL;
    T0 = 1.0 + T9;
    Idsa = Idl * T0;
    dIdsa_dVg = T0 * dIdl_dVg - Idl * (dVdseff_dVg + T9 * dVADIBL_dVg) / VADIBL;
    dIdsa_dVd = T0 * dIdl_dVd + Idl * (1.0 - dVdseff_dVd - T9 * dVADIBL_dVd) / VADIBL;
    dIdsa_dVb = T0 * dIdl_dVb - Idl * (dVdseff_dVb + T9 * dVADIBL_dVb) / VADIBL;
    T9 = diffVds / VADITS;
    T0 = 1.0 + T9;
    dIdsa_dVg = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVADITS_dVg) / VADITS;
    dIdsa_dVd = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVADITS_dVd) / VADITS;
    dIdsa_dVb = T0 * dIdsa_dVb - Idsa * dVdseff_dVb / VADITS;
    Idsa *= T0;
    {
        T0 = log(Va / Vasat);
        const T dT0_dVg = dVa_dVg / Va - dVasat_dVg / Vasat;
        const T dT0_dVb = dVa_dVb / Va - dVasat_dVb / Vasat;
        const T dT0_dVd = dVa_dVd / Va - dVasat_dVd / Vasat;
        T1 = T0 / Cclm;
        T9 = 1.0 + T1;
        const T dT9_dVg2 = (dT0_dVg - T1 * dCclm_dVg) / Cclm;
        const T dT9_dVb2 = (dT0_dVb - T1 * dCclm_dVb) / Cclm;
        const T dT9_dVd2 = (dT0_dVd - T1 * dCclm_dVd) / Cclm;
        dIdsa_dVg = dIdsa_dVg * T9 + Idsa * dT9_dVg2;
        dIdsa_dVb = dIdsa_dVb * T9 + Idsa * dT9_dVb2;
        dIdsa_dVd = dIdsa_dVd * T9 + Idsa * dT9_dVd2;
        Idsa *= T9;
    }
    T9 = diffVds / VASCBE;
    T0 = 1.0 + T9;
    const T Ids = Idsa * T0;
    T Gm = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVASCBE_dVg) / VASCBE;
    T Gds = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVASCBE_dVd) / VASCBE;
    T Gmb = T0 * dIdsa_dVb - Idsa * (dVdseff_dVb + T9 * dVASCBE_dVb) / VASCBE;
    tmp1 = Gds + Gm * dVgsteff_dVd;
    tmp2 = Gmb + Gm * dVgsteff_dVb;
    tmp3 = Gm;
    Gm = (Ids * dVdseff_dVg + Vdseff * tmp3) * dVgsteff_dVg;
    Gds = Ids * (dVdseff_dVd + dVdseff_dVg * dVgsteff_dVd) + Vdseff * tmp1;
    Gmb = (Ids * (dVdseff_dVb + dVdseff_dVg * dVgsteff_dVb) + Vdseff * tmp2) * dVbseff_dVb;
    const T cdrain = Ids * Vdseff;
    const T h_gds_new = Gds, h_gm_new = Gm, h_gmbs_new = Gmb;

    // ---- C-V, capMod = 2 (charge-thickness model, :2910-3810), XPART < 0.5
    T qgate = 0.0, qbulk = 0.0, qdrn = 0.0;
    T cggb = 0.0, cgsb = 0.0, cgdb = 0.0, cdgb = 0.0, cdsb = 0.0, cddb = 0.0;
    T cbgb = 0.0, cbsb = 0.0, cbdb = 0.0;
    if (charge) {
        T VbseffCV, dVbseffCV_dVb;
        if (Vbseff < 0.0) {
            VbseffCV = Vbseff;
            dVbseffCV_dVb = 1.0;
        } else {
            VbseffCV = PF(phi) - Phis;
            dVbseffCV_dVb = -dPhis_dVb;
        }
        const T CoxWL = PF(coxe) * PF(weffCV) * PF(leffCV) * 1.0;
        {
            const T noff = n * PF(noff);
            const T dnoff_dVd = PF(noff) * dn_dVd;
            const T dnoff_dVb = PF(noff) * dn_dVb;
            T0 = Vtm * noff;
            const T voffcv = PF(voffcv);
            const T VgstNVt = (Vgst - voffcv) / T0;
            if (VgstNVt > kExpThreshold) {
                Vgsteff = Vgst - voffcv;
                dVgsteff_dVg = dVgs_eff_dVg;
                dVgsteff_dVd = -dVth_dVd;
                dVgsteff_dVb = -dVth_dVb;
            } else if (VgstNVt < -kExpThreshold) {
                Vgsteff = T0 * log(1.0 + kMinExp);
                dVgsteff_dVg = 0.0;
                dVgsteff_dVd = Vgsteff / noff;
                dVgsteff_dVb = dVgsteff_dVd * dnoff_dVb;
                dVgsteff_dVd *= dnoff_dVd;
            } else {
                const T ExpVgst = exp(VgstNVt);
                Vgsteff = T0 * log(1.0 + ExpVgst);
                dVgsteff_dVg = ExpVgst / (1.0 + ExpVgst);
                dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd) +
                               Vgsteff / noff * dnoff_dVd;
                dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb) +
                               Vgsteff / noff * dnoff_dVb;
                dVgsteff_dVg *= dVgs_eff_dVg;
            }
        }
        T V3 = PF(vfbzb) - Vgs_eff + VbseffCV - kDelta3;
        if (PF(vfbzb) <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * kDelta3 * PF(vfbzb));
        else T0 = sqrt(V3 * V3 + 4.0 * kDelta3 * PF(vfbzb));
        T1 = 0.5 * (1.0 + V3 / T0);
        const T Vfbeff = PF(vfbzb) - 0.5 * (V3 + T0);
        const T dVfbeff_dVg = T1 * dVgs_eff_dVg;
        const T dVfbeff_dVb = -T1 * dVbseffCV_dVb;
        const T Cox = PF(coxp);
        T Tox = 1.0e8 * PF(toxp);
        T0 = (Vgs_eff - VbseffCV - PF(vfbzb)) / Tox;
        const T dT0_dVg = dVgs_eff_dVg / Tox;
        const T dT0_dVb = -dVbseffCV_dVb / Tox;
        T Tcen, dTcen_dVg, dTcen_dVb, dTcen_dVd;
        T tmp = T0 * PF(acde);
        if ((-kExpThreshold < tmp) && (tmp < kExpThreshold)) {
            Tcen = PF(ldeb) * exp(tmp);
            dTcen_dVg = PF(acde) * Tcen;
            dTcen_dVb = dTcen_dVg * dT0_dVb;
            dTcen_dVg *= dT0_dVg;
        } else if (tmp <= -kExpThreshold) {
            Tcen = PF(ldeb) * kMinExp;
            dTcen_dVg = dTcen_dVb = 0.0;
        } else {
            Tcen = PF(ldeb) * kMaxExp;
            dTcen_dVg = dTcen_dVb = 0.0;
        }
        const T LINK = 1.0e-3 * PF(toxp);
        V3 = PF(ldeb) - Tcen - LINK;
        const T V4 = sqrt(V3 * V3 + 4.0 * LINK * PF(ldeb));
        Tcen = PF(ldeb) - 0.5 * (V3 + V4);
        T1 = 0.5 * (1.0 + V3 / V4);
        dTcen_dVg *= T1;
        dTcen_dVb *= T1;
        T Ccen = epssub / Tcen;
        T2 = Cox / (Cox + Ccen);
        T Coxeff = T2 * Ccen;
        T3 = -Ccen / Tcen;
        T dCoxeff_dVg = T2 * T2 * T3;
        T dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        T CoxWLcen = CoxWL * Coxeff / PF(coxe);
        const T Qac0 = CoxWLcen * (Vfbeff - PF(vfbzb));
        T QovCox = Qac0 / Coxeff;
        const T dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVg;
        const T dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVb;
        T0 = 0.5 * PF(k1ox);
        T3 = Vgs_eff - Vfbeff - VbseffCV - Vgsteff;
        if (PF(k1ox) == 0.0) {
            T1 = 0.0;
            T2 = 0.0;
        } else if (T3 < 0.0) {
            T1 = T0 + T3 / PF(k1ox);
            T2 = CoxWLcen;
        } else {
            T1 = sqrt(T0 * T0 + T3);
            T2 = CoxWLcen * T0 / T1;
        }
        const T Qsub0 = CoxWLcen * PF(k1ox) * (T1 - T0);
        QovCox = Qsub0 / Coxeff;
        const T dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg) + QovCox * dCoxeff_dVg;
        const T dQsub0_dVd = -T2 * dVgsteff_dVd;
        const T dQsub0_dVb =
            -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb) + QovCox * dCoxeff_dVb;
        // gate-bias dependent delta Phis
        T Denomi;
        if (PF(k1ox) <= 0.0) {
            Denomi = 0.25 * PF(moin) * Vtm;
            T0 = 0.5 * PF(sqrtPhi);
        } else {
            Denomi = PF(moin) * Vtm * PF(k1ox) * PF(k1ox);
            T0 = PF(k1ox) * PF(sqrtPhi);
        }
        T1 = 2.0 * T0 + Vgsteff;
        const T DeltaPhi = Vtm * log(1.0 + T1 * Vgsteff / Denomi);
        const T dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (Denomi + T1 * Vgsteff);
        T0 = Vgsteff - DeltaPhi - 0.001;
        const T dT0_dVg2 = 1.0 - dDeltaPhi_dVg;
        T1 = sqrt(T0 * T0 + Vgsteff * 0.004);
        const T VgDP = 0.5 * (T0 + T1);
        const T dVgDP_dVg = 0.5 * (dT0_dVg2 + (T0 * dT0_dVg2 + 0.002) / T1);
        Tox += Tox;
        T0 = (Vgsteff + PF(vtfbphi2)) / Tox;
        tmp = exp(PF(bdos) * 0.7 * log(T0));
        T1 = 1.0 + tmp;
        T2 = PF(bdos) * 0.7 * tmp / (T0 * Tox);
        Tcen = PF(ados) * 1.9e-9 / T1;
        dTcen_dVg = -Tcen * T2 / T1;
        dTcen_dVd = dTcen_dVg * dVgsteff_dVd;
        dTcen_dVb = dTcen_dVg * dVgsteff_dVb;
        dTcen_dVg *= dVgsteff_dVg;
        Ccen = epssub / Tcen;
        T0 = Cox / (Cox + Ccen);
        Coxeff = T0 * Ccen;
        T1 = -Ccen / Tcen;
        dCoxeff_dVg = T0 * T0 * T1;
        const T dCoxeff_dVd = dCoxeff_dVg * dTcen_dVd;
        dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        CoxWLcen = CoxWL * Coxeff / PF(coxe);
        const T AbulkCV = Abulk0 * PF(abulkCVfactor);
        const T dAbulkCV_dVb = PF(abulkCVfactor) * dAbulk0_dVb;
        const T VdsatCV = VgDP / AbulkCV;
        T0 = VdsatCV - Vds - kDelta4;
        const T dT0_dVg3 = dVgDP_dVg / AbulkCV;
        const T dT0_dVb3 = -VdsatCV * dAbulkCV_dVb / AbulkCV;
        T1 = sqrt(T0 * T0 + 4.0 * kDelta4 * VdsatCV);
        T dT1_dVg = (T0 + kDelta4 + kDelta4) / T1;
        const T dT1_dVd = -T0 / T1;
        const T dT1_dVb = dT1_dVg * dT0_dVb3;
        dT1_dVg *= dT0_dVg3;
        T VdseffCV, dVdseffCV_dVg, dVdseffCV_dVd, dVdseffCV_dVb;
        if (T0 >= 0.0) {
            VdseffCV = VdsatCV - 0.5 * (T0 + T1);
            dVdseffCV_dVg = 0.5 * (dT0_dVg3 - dT1_dVg);
            dVdseffCV_dVd = 0.5 * (1.0 - dT1_dVd);
            dVdseffCV_dVb = 0.5 * (dT0_dVb3 - dT1_dVb);
        } else {
            T3 = (kDelta4 + kDelta4) / (T1 - T0);
            T4 = 1.0 - T3;
            T5 = VdsatCV * T3 / (T1 - T0);
            VdseffCV = VdsatCV * T4;
            dVdseffCV_dVg = dT0_dVg3 * T4 + T5 * (dT1_dVg - dT0_dVg3);
            dVdseffCV_dVd = T5 * (dT1_dVd + 1.0);
            dVdseffCV_dVb = dT0_dVb3 * (T4 - T5) + T5 * dT1_dVb;
        }
        if (Vds == 0.0) {
            VdseffCV = 0.0;
            dVdseffCV_dVg = 0.0;
            dVdseffCV_dVb = 0.0;
        }
        T0 = AbulkCV * VdseffCV;
        T1 = VgDP;
        T2 = 12.0 * (T1 - 0.5 * T0 + 1.0e-20);
        T3 = T0 / T2;
        T4 = 1.0 - 12.0 * T3 * T3;
        T5 = AbulkCV * (6.0 * T0 * (4.0 * T1 - T0) / (T2 * T2) - 0.5);
        T6 = T5 * VdseffCV / AbulkCV;
        qgate = CoxWLcen * (T1 - T0 * (0.5 - T3));
        QovCox = qgate / Coxeff;
        T Cgg1 = CoxWLcen * (T4 * dVgDP_dVg + T5 * dVdseffCV_dVg);
        const T Cgd1 = CoxWLcen * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
        const T Cgb1 = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) +
                            Cgg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
        Cgg1 = Cgg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        T7 = 1.0 - AbulkCV;
        T8 = T2 * T2;
        T9 = 12.0 * T7 * T0 * T0 / (T8 * AbulkCV);
        T10 = T9 * dVgDP_dVg;
        T11 = -T7 * T5 / AbulkCV;
        T12 = -(T9 * T1 / AbulkCV + VdseffCV * (0.5 - T0 / T2));
        qbulk = CoxWLcen * T7 * (0.5 * VdseffCV - T0 * VdseffCV / T2);
        QovCox = qbulk / Coxeff;
        T Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
        const T Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
        const T Cbb1 = CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) +
                            Cbg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
        Cbg1 = Cbg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        // 40/60 partition
        T2 = T2 / 12.0;
        T3 = 0.5 * CoxWLcen / (T2 * T2);
        T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
        const T qsrc0 = -T3 * T4;
        QovCox = qsrc0 / Coxeff;
        T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
        T5 = -2.0 * qsrc0 / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
        T6 = AbulkCV * (qsrc0 / T2 + T3 * T8);
        T7 = T6 * VdseffCV / AbulkCV;
        T Csg = T5 * dVgDP_dVg + T6 * dVdseffCV_dVg;
        const T Csd = Csg * dVgsteff_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
        T Csb = Csg * dVgsteff_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVb;
        Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        qgate += Qac0 + Qsub0 - qbulk;
        qbulk -= (Qac0 + Qsub0);
        qdrn = -(qgate + qbulk + qsrc0);
        const T Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
        const T Cbd = Cbd1 - dQsub0_dVd;
        T Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
        const T Cgg = Cgg1 - Cbg;
        const T Cgd = Cgd1 - Cbd;
        T Cgb = Cgb1 - Cbb;
        Cgb *= dVbseff_dVb;
        Cbb *= dVbseff_dVb;
        Csb *= dVbseff_dVb;
        cggb = Cgg;
        cgsb = -(Cgg + Cgd + Cgb);
        cgdb = Cgd;
        cdgb = -(Cgg + Cbg + Csg);
        cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
        cddb = -(Cgd + Cbd + Csd);
        cbgb = Cbg;
        cbsb = -(Cbg + Cbd + Cbb);
        cbdb = Cbd;
    }

    // ---- junction C-V (:3843-3960)
    T qbs = 0.0, qbd = 0.0, capbs = 0.0, capbd = 0.0;
    if constexpr (!kGrad) {
    if (charge) {
        const T czbd = PF(czbd), czbs = PF(czbs), czbdsw = PF(czbdsw), czbdswg = PF(czbdswg);
        const T czbssw = PF(czbssw), czbsswg = PF(czbsswg);
        if (vbs_jct == 0.0) {
            qbs = 0.0;
            capbs = czbs + czbssw + czbsswg;
        } else if (vbs_jct < 0.0) {
            if (czbs > 0.0) {
                const T arg = 1.0 - vbs_jct / PF(PhiBS);
                const T sarg = (PF(MJS) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJS) * log(arg));
                qbs = PF(PhiBS) * czbs * (1.0 - arg * sarg) / (1.0 - PF(MJS));
                capbs = czbs * sarg;
            } else {
                qbs = 0.0;
                capbs = 0.0;
            }
            if (czbssw > 0.0) {
                const T arg = 1.0 - vbs_jct / PF(PhiBSWS);
                const T sarg = (PF(MJSWS) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJSWS) * log(arg));
                qbs += PF(PhiBSWS) * czbssw * (1.0 - arg * sarg) / (1.0 - PF(MJSWS));
                capbs += czbssw * sarg;
            }
            if (czbsswg > 0.0) {
                const T arg = 1.0 - vbs_jct / PF(PhiBSWGS);
                const T sarg = (PF(MJSWGS) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJSWGS) * log(arg));
                qbs += PF(PhiBSWGS) * czbsswg * (1.0 - arg * sarg) / (1.0 - PF(MJSWGS));
                capbs += czbsswg * sarg;
            }
        } else {
            T0 = czbs + czbssw + czbsswg;
            T1 = vbs_jct * (czbs * PF(MJS) / PF(PhiBS) + czbssw * PF(MJSWS) / PF(PhiBSWS) +
                            czbsswg * PF(MJSWGS) / PF(PhiBSWGS));
            qbs = vbs_jct * (T0 + 0.5 * T1);
            capbs = T0 + T1;
        }
        if (vbd_jct == 0.0) {
            qbd = 0.0;
            capbd = czbd + czbdsw + czbdswg;
        } else if (vbd_jct < 0.0) {
            if (czbd > 0.0) {
                const T arg = 1.0 - vbd_jct / PF(PhiBD);
                const T sarg = (PF(MJD) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJD) * log(arg));
                qbd = PF(PhiBD) * czbd * (1.0 - arg * sarg) / (1.0 - PF(MJD));
                capbd = czbd * sarg;
            } else {
                qbd = 0.0;
                capbd = 0.0;
            }
            if (czbdsw > 0.0) {
                const T arg = 1.0 - vbd_jct / PF(PhiBSWD);
                const T sarg = (PF(MJSWD) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJSWD) * log(arg));
                qbd += PF(PhiBSWD) * czbdsw * (1.0 - arg * sarg) / (1.0 - PF(MJSWD));
                capbd += czbdsw * sarg;
            }
            if (czbdswg > 0.0) {
                const T arg = 1.0 - vbd_jct / PF(PhiBSWGD);
                const T sarg = (PF(MJSWGD) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJSWGD) * log(arg));
                qbd += PF(PhiBSWGD) * czbdswg * (1.0 - arg * sarg) / (1.0 - PF(MJSWGD));
                capbd += czbdswg * sarg;
            }
        } else {
            T0 = czbd + czbdsw + czbdswg;
            T1 = vbd_jct * (czbd * PF(MJD) / PF(PhiBD) + czbdsw * PF(MJSWD) / PF(PhiBSWD) +
                            czbdswg * PF(MJSWGD) / PF(PhiBSWGD));
            qbd = vbd_jct * (T0 + 0.5 * T1);
            capbd = T0 + T1;
        }
        put(s0, S(qbsS), qbs);
        put(s0, S(qbdS), qbd);
    }
    }

    // ---- convergence (:3967-4011), with this load's currents
    int noncon = 0;
    if (Check == 1) {
        noncon = 1;
    } else if (!(mode & MODEADJOINT)) {
        const T Idtot = (dmode >= 0) ? cdrain - cbd : cdrain + cbd;
        const T tol0 = tol.reltol * ucb_max(fabs(cdhat), fabs(Idtot)) + tol.abstol;
        if (fabs(cdhat - Idtot) >= tol0) {
            noncon = 1;
        } else {
            const T Ibtot = cbs + cbd;
            const T tol6 = tol.reltol * ucb_max(fabs(cbhat), fabs(Ibtot)) + tol.abstol;
            if (fabs(cbhat - Ibtot) > tol6) noncon = 1;
        }
    }
    put(s0, S(vdsS), vds);
    put(s0, S(vgsS), vgs);
    put(s0, S(vbsS), vbs);
    put(s0, S(vbdS), vbd);

    // ---- charges and their companion conductances (:4026-4546)
    T gcggb = 0.0, gcgdb = 0.0, gcgsb = 0.0, gcgbb = 0.0, gcdgb = 0.0, gcddb = 0.0,
           gcdsb = 0.0, gcdbb = 0.0, gcsgb = 0.0, gcsdb = 0.0, gcssb = 0.0, gcsbb = 0.0,
           gcbdb = 0.0, gcbgb = 0.0, gcbsb = 0.0, gcbbb = 0.0;
    T ceqqg = 0.0, ceqqd = 0.0, ceqqb = 0.0;
    if constexpr (!kGrad) {
    if (charge) {
        const T vgdx = vgd, vgsx = vgs;
        T0 = vgdx + kDelta1;
        T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
        T2 = 0.5 * (T0 - T1);
        T3 = PF(weffCV) * PF(cgdl);
        T4 = sqrt(1.0 - 4.0 * T2 / PF(ckappad));
        const T cgdo = PF(cgdo) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        const T qgdo = (PF(cgdo) + T3) * vgdx - T3 * (T2 + 0.5 * PF(ckappad) * (T4 - 1.0));
        T0 = vgsx + kDelta1;
        T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
        T2 = 0.5 * (T0 - T1);
        T3 = PF(weffCV) * PF(cgsl);
        T4 = sqrt(1.0 - 4.0 * T2 / PF(ckappas));
        const T cgso = PF(cgso) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        const T qgso = (PF(cgso) + T3) * vgsx - T3 * (T2 + 0.5 * PF(ckappas) * (T4 - 1.0));
        const T ag0 = c.ag0;
        T qsrc;
        if (dmode > 0) {
            qdrn -= qgdo;
            gcggb = (cggb + cgdo + cgso + PF(cgbo)) * ag0;
            gcgdb = (cgdb - cgdo) * ag0;
            gcgsb = (cgsb - cgso) * ag0;
            gcgbb = -(gcggb + gcgdb + gcgsb);
            gcdgb = (cdgb - cgdo) * ag0;
            gcsgb = -(cggb + cbgb + cdgb + cgso) * ag0;
            gcbgb = (cbgb - PF(cgbo)) * ag0;
            const T qgb = PF(cgbo) * vgb;
            qgate += qgdo + qgso + qgb;
            qbulk -= qgb;
            qsrc = -(qgate + qbulk + qdrn);
            gcddb = (cddb + capbd + cgdo) * ag0;
            gcdsb = cdsb * ag0;
            gcsdb = -(cgdb + cbdb + cddb) * ag0;
            gcssb = (capbs + cgso - (cgsb + cbsb + cdsb)) * ag0;
            gcdbb = -(gcdgb + gcddb + gcdsb + 0.0);
            gcsbb = -(gcsgb + gcsdb + gcssb + 0.0);
            gcbdb = (cbdb - capbd) * ag0;
            gcbsb = (cbsb - capbs) * ag0;
            gcbbb = -(gcbdb + gcbgb + gcbsb + 0.0);
        } else {
            qsrc = qdrn - qgso;
            gcggb = (cggb + cgdo + cgso + PF(cgbo)) * ag0;
            gcgdb = (cgsb - cgdo) * ag0;
            gcgsb = (cgdb - cgso) * ag0;
            gcgbb = -(gcggb + gcgdb + gcgsb);
            gcdgb = -(cggb + cbgb + cdgb + cgdo) * ag0;
            gcsgb = (cdgb - cgso) * ag0;
            gcbgb = (cbgb - PF(cgbo)) * ag0;
            const T qgb = PF(cgbo) * vgb;
            qgate += qgdo + qgso + qgb;
            qbulk -= qgb;
            qdrn = -(qgate + qbulk + qsrc);
            gcddb = (capbd + cgdo - (cgsb + cbsb + cdsb)) * ag0;
            gcdsb = -(cgdb + cbdb + cddb) * ag0;
            gcsdb = cdsb * ag0;
            gcssb = (cddb + capbs + cgso) * ag0;
            gcdbb = -(gcdgb + gcddb + gcdsb + 0.0);
            gcsbb = -(gcsgb + gcsdb + gcssb + 0.0);
            gcbdb = (cbsb - capbd) * ag0;
            gcbsb = (cbdb - capbs) * ag0;
            gcbbb = -(gcbgb + gcbdb + gcbsb + 0.0);
        }
        put(s0, S(qgS), qgate);
        put(s0, S(qdS), qdrn - qbd);
        put(s0, S(qsS), qsrc - qbs);
        put(s0, S(qbS), qbulk + qbd + qbs);
        if (mode & MODEINITTRAN) {
            put(s1, S(qbS), s0[S(qbS)]);
            put(s1, S(qgS), s0[S(qgS)]);
            put(s1, S(qdS), s0[S(qdS)]);
        }
        T geq, ceq;
        niintegrate(c, s0, s1, s2, S(qbS), sB, 0.0, &geq, &ceq);
        niintegrate(c, s0, s1, s2, S(qgS), sB, 0.0, &geq, &ceq);
        niintegrate(c, s0, s1, s2, S(qdS), sB, 0.0, &geq, &ceq);
        const T cqgate = s0[S(cqgS)], cqbody = s0[S(cqbS)], cqdrn = s0[S(cqdS)];
        ceqqg = cqgate - gcggb * vgb + gcgdb * vbd + gcgsb * vbs;
        ceqqd = cqdrn - gcdgb * vgb + gcddb * vbd + gcdsb * vbs;
        ceqqb = cqbody - gcbgb * vgb + gcbdb * vbd + gcbsb * vbs;
        if (mode & MODEINITTRAN) {
            put(s1, S(cqbS), s0[S(cqbS)]);
            put(s1, S(cqgS), s0[S(cqgS)]);
            put(s1, S(cqdS), s0[S(cqdS)]);
        }
    }
    }

    // ---- currents into the right-hand side (:4643-4914)
    T Gmr, Gmbsr, FwdSum, RevSum, ceqdrn;
    if (dmode >= 0) {
        Gmr = Gm;
        Gmbsr = Gmb;
        FwdSum = Gmr + Gmbsr;
        RevSum = 0.0;
        ceqdrn = type * (cdrain - Gds * vds - Gmr * vgs - Gmbsr * vbs);
    } else {
        Gmr = -Gm;
        Gmbsr = -Gmb;
        FwdSum = 0.0;
        RevSum = -(Gmr + Gmbsr);
        ceqdrn = -type * (cdrain + Gds * vds + Gmr * vgd + Gmbsr * vbd);
    }
    T ceqjs, ceqjd;
    if (type > 0) {
        ceqjs = (cbs - gbs * vbs_jct);
        ceqjd = (cbd - gbd * vbd_jct);
    } else {
        ceqjs = -(cbs - gbs * vbs_jct);
        ceqjd = -(cbd - gbd * vbd_jct);
        ceqqg = -ceqqg;
        ceqqd = -ceqqd;
        ceqqb = -ceqqb;
    }
    T rv[4], mv[16];
    rv[0] = (ceqjd - ceqdrn - ceqqd);                              // dNodePrime +=
    rv[1] = -(ceqqg);                                             // gNodePrime -=
    rv[2] = (-ceqjd - ceqjs - ceqqb);                         // bNodePrime +=
    rv[3] = (ceqdrn + ceqjs + ceqqg + ceqqb + ceqqd);         // sNodePrime +=

    // ---- matrix (:4980-5024), stamp order of the reference
    const T gjbd = gbd, gjbs = gbs;
    mv[0] = gcggb;                                                  // GPgp
    mv[1] = gcgdb;                                                 // GPdp
    mv[2] = gcgsb;                                             // GPsp
    mv[3] = gcgbb;                                             // GPbp
    mv[4] = (Gds + gbd + RevSum + gcddb);                      // DPdp
    mv[5] = (Gmr + gcdgb);                                     // DPgp
    mv[6] = -(Gds + FwdSum - gcdsb);                           // DPsp
    mv[7] = -(gjbd - Gmbsr - gcdbb);                           // DPbp
    mv[8] = -(Gds + RevSum - gcsdb);                           // SPdp
    mv[9] = (gcsgb - Gmr);                                     // SPgp
    mv[10] = (Gds + gbs + FwdSum + gcssb);                     // SPsp
    mv[11] = -(gjbs + Gmbsr - gcsbb);                          // SPbp
    mv[12] = (gcbdb - gjbd);                                   // BPdp
    mv[13] = (gcbgb);                                          // BPgp
    mv[14] = (gcbsb - gjbs);                                   // BPsp
    mv[15] = (gjbd + gjbs + gcbbb);                            // BPbp

    if constexpr (!kGrad) {
        double* r = cr + (size_t)(t.rsite + 4 * d) * sB + inst;
        for (int k = 0; k < 4; ++k) r[(size_t)k * sB] = rv[k];
        double* m = cm + (size_t)(t.site + 16 * d) * sB + inst;
        for (int k = 0; k < 16; ++k) m[(size_t)k * sB] = mv[k];
    } else {
        // dL/dp = -lambda^T dF/dp with F_row = sum_col A(row, col) x_col - rhs_row
        // over the device's rows; stamp order: matrix rows G, D, S, B with
        // columns (g d s b), (d g s b), (d g s b), (d g s b); rhs rows D, G, B, S
        const double vd = value_of(V(nD, 0)), vg = value_of(V(nG, 1)), vs = value_of(V(nS, 2)),
                     vb = value_of(V(nB, 3));
        auto Lm = [&](int node, int k) -> double {
            if (t.pin) {
                const int q = t.pin[(size_t)k * sB + inst];
                return q >= 0 ? lam[q] : 0.0;
            }
            return sol(lam, node, sB, inst);
        };
        const double ld = Lm(nD, 0), lg = Lm(nG, 1), ls = Lm(nS, 2), lb = Lm(nB, 3);
        const double dFg = mv[0].d[0] * vg + mv[1].d[0] * vd + mv[2].d[0] * vs + mv[3].d[0] * vb - rv[1].d[0];
        const double dFd = mv[4].d[0] * vd + mv[5].d[0] * vg + mv[6].d[0] * vs + mv[7].d[0] * vb - rv[0].d[0];
        const double dFs = mv[8].d[0] * vd + mv[9].d[0] * vg + mv[10].d[0] * vs + mv[11].d[0] * vb - rv[3].d[0];
        const double dFb = mv[12].d[0] * vd + mv[13].d[0] * vg + mv[14].d[0] * vs + mv[15].d[0] * vb - rv[2].d[0];
        gout[gid] = -(lg * dFg + ld * dFd + ls * dFs + lb * dFb);
    }

    put(aux, B4_VON * sB, von_new);
    put(aux, B4_MODE * sB, T(dmode));
    put(aux, B4_CD * sB, cdrain);
    put(aux, B4_CBS * sB, cbs);
    put(aux, B4_CBD * sB, cbd);
    put(aux, B4_GBS * sB, gbs);
    put(aux, B4_GBD * sB, gbd);
    put(aux, B4_GM * sB, h_gm_new);
    put(aux, B4_GDS * sB, h_gds_new);
    put(aux, B4_GMBS * sB, h_gmbs_new);
    if constexpr (!kGrad) {
        if (noncon) atomicOr(&devnc[ci], 1);
    }
}


__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_bsim4(
    B4Table t, const DevCtx* __restrict__ ctx, const double* __restrict__ x, double* __restrict__ s0,
    double* __restrict__ s1, const double* __restrict__ s2, const int32_t* __restrict__ done,
    double* __restrict__ cm, double* __restrict__ cr, int32_t* __restrict__ devnc, DevTol tol, int B, int nb4,
    VsrcSide vs) {
    if ((int)blockIdx.x >= nb4) {
        // an extra workgroup: voltage sources (their own sites)
        vsrc_lane(vs.t, vs.ctx, vs.done, vs.cr, vs.B, (long)(blockIdx.x - nb4) * TPB + threadIdx.x);
        return;
    }
    b4_lane<double>(t, ctx, x, s0, s1, s2, done, cm, cr, devnc, tol, B, nullptr, nullptr, nullptr);
}

// dL/dp of every (device, instance) for one direction p of the record:
// dpar holds d(record)/dp in the per-instance layout of par_inst.
__global__ __launch_bounds__(TPB) void k_bsim4_grad(B4Table t, const DevCtx* __restrict__ ctx,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ dpar,
                                                    const double* __restrict__ lam,
                                                    double* __restrict__ g, DevTol tol, int B) {
    b4_lane<Dual<1>>(t, ctx, x, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, tol, B, dpar,
                     lam, g);
}

}  // namespace

void launch_bsim4_grad(const B4Table& t, const DevCtx* ctx, const double* x, const double* dpar,
                       const double* lam, double* g, DevTol tol, int B, hipStream_t s) {
    if (!t.count) return;
    if (!t.par_inst) throw Error(NSA_ERR_LOGIC, "BSIM4v7 gradient: needs per-instance records");
    hipLaunchKernelGGL(k_bsim4_grad, dim3(blocks_for((long)t.count * B)), dim3(TPB), 0, s, t, ctx, x, dpar,
                       lam, g, tol, B);
}

void launch_bsim4(const B4Table& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                  const double* s2, const int32_t* done, double* cm, double* cr, int32_t* devnc,
                  DevTol tol, int B, hipStream_t s, const VsrcSide* vs) {
    if (!t.count) return;
    const int nb4 = (int)blocks_for((long)t.count * B);
    VsrcSide v{};
    int nv = 0;
    if (vs && vs->t.count) {
        v = *vs;
        nv = (int)blocks_for((long)v.t.count * v.B);
    }
    hipLaunchKernelGGL(k_bsim4, dim3(nb4 + nv), dim3(TPB), 0, s, t, ctx, x, s0, s1, s2, done, cm, cr, devnc, tol,
                       B, nb4, v);
}

}  // namespace nsa
