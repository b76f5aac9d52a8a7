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
#include "dev_limits.hpp"
#include "gpu_dev.hpp"

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
#define PF(f) (pb[(offsetof(B4Param, f) / sizeof(double)) * ps])
static_assert(kB4Fields == (int)(offsetof(B4Param, off) / sizeof(double)), "B4Param layout");

// UCB MAX (ucb_compat.hpp:38-43)
__device__ inline double ucb_max(double a, double b) { return (a > b) ? a : b; }

// The lane's live state (~140 doubles at the charge model) fits in 280
// VGPRs without spilling when the kernel may run one wave per SIMD; at the
// compiler's default occupancy target it spilled 680 bytes per lane to
// scratch, which doubled its HBM traffic and its duration.
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_bsim4(B4Table t, const DevCtx* __restrict__ ctx, const double* __restrict__ x,
                        double* __restrict__ s0, double* __restrict__ s1,
                        const double* __restrict__ s2, const int32_t* __restrict__ done,
                        double* __restrict__ cm, double* __restrict__ cr,
                        int32_t* __restrict__ devnc, DevTol tol, int B) {
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
    const int mode = c.mode;
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
    auto V = [&](int node, int k) -> double {
        if (t.pin) {
            const int q = t.pin[(size_t)k * sB + inst];
            return q >= 0 ? x[q] : 0.0;
        }
        return sol(x, node, sB, inst);
    };
    const double type = PF(type);
    const bool charge = ((mode & (MODEDCTRANCURVE | MODEAC | MODETRAN | MODEINITSMSIG)) ||
                         ((mode & MODETRANOP) && (mode & MODEUIC)));

    // instance fields carried from the previous load
    const double h_von = aux[B4_VON * sB];
    const double h_mode = aux[B4_MODE * sB];
    const double h_cd = aux[B4_CD * sB], h_cbs = aux[B4_CBS * sB], h_cbd = aux[B4_CBD * sB];
    const double h_gbs = aux[B4_GBS * sB], h_gbd = aux[B4_GBD * sB];
    const double h_gm = aux[B4_GM * sB], h_gds = aux[B4_GDS * sB], h_gmbs = aux[B4_GMBS * sB];

    // ---- terminal voltages (bsim4v7_load.cpp:237-663)
    int Check = 1;
    double vds, vgs, vbs, vbd, vgd;
    double cdhat = 0.0, cbhat = 0.0;
    if (mode & MODEADJOINT) {
        // the converged solution itself, unlimited (the adjoint's Jacobian)
        vds = type * (V(nD, 0) - V(nS, 2));
        vgs = type * (V(nG, 1) - V(nS, 2));
        vbs = type * (V(nB, 3) - V(nS, 2));
        Check = 0;
    } else if (mode & MODEINITTRAN) {
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
            const double xfact = c.delta / c.delta_old1;
            s0[S(vdsS)] = s1[S(vdsS)];
            vds = (1.0 + xfact) * (s1[S(vdsS)]) - (xfact * (s2[S(vdsS)]));
            s0[S(vgsS)] = s1[S(vgsS)];
            vgs = (1.0 + xfact) * (s1[S(vgsS)]) - (xfact * (s2[S(vgsS)]));
            s0[S(vbsS)] = s1[S(vbsS)];
            vbs = (1.0 + xfact) * (s1[S(vbsS)]) - (xfact * (s2[S(vbsS)]));
            s0[S(vbdS)] = s0[S(vbsS)] - s0[S(vdsS)];
        } else {
            vds = type * (V(nD, 0) - V(nS, 2));
            vgs = type * (V(nG, 1) - V(nS, 2));
            vbs = type * (V(nB, 3) - V(nS, 2));
        }
        const double s0vds = s0[S(vdsS)], s0vgs = s0[S(vgsS)], s0vbs = s0[S(vbsS)],
                     s0vbd = s0[S(vbdS)];
        const double vgdo = s0vgs - s0vds;
        vbd = vbs - vds;
        vgd = vgs - vds;
        const double delvbd = vbd - s0vbd;
        const double delvgd = vgd - vgdo;
        const double delvds = vds - s0vds;
        const double delvgs = vgs - s0vgs;
        const double delvbs = vbs - s0vbs;
        // the linearised currents of the previous load at this iterate, for
        // the current-based convergence test (:421-473)
        if (h_mode >= 0) {
            const double Idtot = h_cd - h_cbd;
            cdhat = Idtot - h_gbd * delvbd + h_gmbs * delvbs + h_gm * delvgs + h_gds * delvds;
            const double Ibtot = h_cbs + h_cbd;
            cbhat = Ibtot + h_gbd * delvbd + h_gbs * delvbs;
        } else {
            const double Idtot = h_cd + h_cbd;
            cdhat = Idtot + h_gbd * delvbd + h_gmbs * delvbd + h_gm * delvgd - h_gds * delvds;
            const double Ibtot = h_cbs + h_cbd;
            cbhat = Ibtot + h_gbs * delvbs + h_gbd * delvbd;
        }
        // limiting (:582-662)
        const double von = h_von;
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
    vbd = vbs - vds;
    vgd = vgs - vds;
    const double vgb = vgs - vbs;
    const double vbs_jct = vbs, vbd_jct = vbd;

    // ---- junction currents: zero saturation current, gmin only (:682-795)
    const double gbs = tol.gmin, cbs = gbs * vbs_jct;
    const double gbd = tol.gmin, cbd = gbd * vbd_jct;

    int dmode;
    double Vds, Vbs;
    if (vds >= 0.0) {
        dmode = 1;
        Vds = vds;
        Vbs = vbs;
    } else {
        dmode = -1;
        Vds = -vds;
        Vbs = vbd;
    }
    const double toxe = PF(toxe), epssub = kEpsSi;

    // ---- Vbseff, Phis, Xdep (:1013-1038)
    double T0, T1, T2, T3, T4, T5, T6, T7, T8, T9, T10, T11, T12, T13;
    double Vbseff, dVbseff_dVb;
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
    const double Phis = PF(phi) - Vbseff;
    const double dPhis_dVb = -1.0;
    const double sqrtPhis = sqrt(Phis);
    const double dsqrtPhis_dVb = -0.5 / sqrtPhis;
    const double Xdep = PF(Xdep0) * sqrtPhis / PF(sqrtPhi);
    const double dXdep_dVb = (PF(Xdep0) / PF(sqrtPhi)) * dsqrtPhis_dVb;
    const double Leff = PF(leff);
    const double Vtm = PF(vtm);

    // ---- Vth (:1044-1141)
    T3 = sqrt(Xdep);
    const double V0 = PF(vbi) - PF(phi);
    T0 = PF(dvt2) * Vbseff;
    if (T0 >= -0.5) {
        T1 = 1.0 + T0;
        T2 = PF(dvt2);
    } else {
        T4 = 1.0 / (3.0 + 8.0 * T0);
        T1 = (1.0 + 3.0 * T0) * T4;
        T2 = PF(dvt2) * T4 * T4;
    }
    const double lt1 = PF(factor1) * T3 * T1;
    const double dlt1_dVb = PF(factor1) * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    T0 = PF(dvt2w) * Vbseff;
    if (T0 >= -0.5) {
        T1 = 1.0 + T0;
        T2 = PF(dvt2w);
    } else {
        T4 = 1.0 / (3.0 + 8.0 * T0);
        T1 = (1.0 + 3.0 * T0) * T4;
        T2 = PF(dvt2w) * T4 * T4;
    }
    const double ltw = PF(factor1) * T3 * T1;
    const double dltw_dVb = PF(factor1) * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    double Theta0, dTheta0_dVb;
    T0 = PF(dvt1) * Leff / lt1;
    if (T0 < kExpThreshold) {
        T1 = exp(T0);
        T2 = T1 - 1.0;
        T3 = T2 * T2;
        T4 = T3 + 2.0 * T1 * kMinExp;
        Theta0 = T1 / T4;
        const double dT1_dVb = -T0 * T1 * dlt1_dVb / lt1;
        dTheta0_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + kMinExp)) / T4 / T4;
    } else {
        Theta0 = 1.0 / (kMaxExp - 2.0);
        dTheta0_dVb = 0.0;
    }
    const double thetavth = PF(dvt0) * Theta0;
    const double Delt_vth = thetavth * V0;
    const double dDelt_vth_dVb = PF(dvt0) * dTheta0_dVb * V0;
    double dT5_dVb;
    T0 = PF(dvt1w) * PF(weff) * Leff / ltw;
    if (T0 < kExpThreshold) {
        T1 = exp(T0);
        T2 = T1 - 1.0;
        T3 = T2 * T2;
        T4 = T3 + 2.0 * T1 * kMinExp;
        T5 = T1 / T4;
        const double dT1_dVb = -T0 * T1 * dltw_dVb / ltw;
        dT5_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + kMinExp)) / T4 / T4;
    } else {
        T5 = 1.0 / (kMaxExp - 2.0);
        dT5_dVb = 0.0;
    }
    T0 = PF(dvt0w) * T5;
    T2 = T0 * V0;
    const double dT2_dVb_w = PF(dvt0w) * dT5_dVb * V0;
    const double TempRatio = PF(temp_ratio);
    T0 = sqrt(1.0 + PF(lpe0) / Leff);
    T1 = PF(k1ox) * (T0 - 1.0) * PF(sqrtPhi) + (PF(kt1) + PF(kt1l) / Leff + PF(kt2) * Vbseff) * TempRatio;
    const double Vth_NarrowW = toxe * PF(phi) / (PF(weff) + PF(w0));
    T3 = PF(eta0) + PF(etab) * Vbseff;
    if (T3 < 1.0e-4) {
        T9 = 1.0 / (3.0 - 2.0e4 * T3);
        T3 = (2.0e-4 - T3) * T9;
        T4 = T9 * T9;
    } else {
        T4 = 1.0;
    }
    const double dDIBL_Sft_dVd = T3 * PF(theta0vb0);
    const double DIBL_Sft = dDIBL_Sft_dVd * Vds;
    const double Lpe_Vb = sqrt(1.0 + PF(lpeb) / Leff);
    const double Vth = type * PF(vth0) + (PF(k1ox) * sqrtPhis - PF(k1) * PF(sqrtPhi)) * Lpe_Vb -
                       PF(k2ox) * Vbseff - Delt_vth - T2 + (PF(k3) + PF(k3b) * Vbseff) * Vth_NarrowW +
                       T1 - DIBL_Sft;
    const double dVth_dVb = Lpe_Vb * PF(k1ox) * dsqrtPhis_dVb - PF(k2ox) - dDelt_vth_dVb - dT2_dVb_w +
                            PF(k3b) * Vth_NarrowW - PF(etab) * Vds * PF(theta0vb0) * T4 +
                            PF(kt2) * TempRatio;
    const double dVth_dVd = -dDIBL_Sft_dVd;

    // ---- n (:1145-1165)
    double n, dn_dVb, dn_dVd;
    {
        const double tmp1 = epssub / Xdep;
        const double tmp2 = PF(nfactor) * tmp1;
        const double tmp3 = PF(cdsc) + PF(cdscb) * Vbseff + PF(cdscd) * Vds;
        const double tmp4 = (tmp2 + tmp3 * Theta0 + PF(cit)) / PF(coxe);
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
    const double von_new = Vth;

    // ---- Vgsteff (:1222-1307); no poly depletion (NGATE = 0)
    const double Vgs_eff = (dmode > 0) ? vgs : vgd;
    const double dVgs_eff_dVg = 1.0;
    const double Vgst = Vgs_eff - Vth;
    double dT10_dVg, dT10_dVd, dT10_dVb, dT9_dVg, dT9_dVd, dT9_dVb;
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
        const double ExpVgst = exp(T2);
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
        const double ExpVgst = exp(T2);
        T3 = PF(coxe) / PF(cdep0);
        T4 = T3 * ExpVgst;
        T5 = T1 * T4 / T0;
        T9 = PF(mstar) + n * T4;
        dT9_dVg = T3 * (PF(mstar) - 1.0) * ExpVgst / Vtm;
        dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
        dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
        dT9_dVg *= dVgs_eff_dVg;
    }
    double Vgsteff = T10 / T9;
    T11 = T9 * T9;
    double dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
    double dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
    double dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;

    // ---- effective width, Rds (:1310-1347)
    T9 = sqrtPhis - PF(sqrtPhi);
    double Weff = PF(weff) - 2.0 * (PF(dwg) * Vgsteff + PF(dwb) * T9);
    double dWeff_dVg = -2.0 * PF(dwg);
    double dWeff_dVb = -2.0 * PF(dwb) * dsqrtPhis_dVb;
    if (Weff < 2.0e-8) {
        T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
        Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
        T0 *= T0 * 4.0e-16;
        dWeff_dVg *= T0;
        dWeff_dVb *= T0;
    }
    double Rds, dRds_dVg, dRds_dVb;
    {
        T0 = 1.0 + PF(prwg) * Vgsteff;
        const double dT0_dVg = -PF(prwg) / T0 / T0;
        T1 = PF(prwb) * T9;
        const double dT1_dVb = PF(prwb) * dsqrtPhis_dVb;
        T2 = 1.0 / T0 + T1;
        T3 = T2 + sqrt(T2 * T2 + 0.01);
        double dT3_dVg = 1.0 + T2 / (T3 - T2);
        const double dT3_dVb = dT3_dVg * dT1_dVb;
        dT3_dVg *= dT0_dVg;
        T4 = PF(rds0) * 0.5;
        Rds = PF(rdswmin) + T3 * T4;
        dRds_dVg = T4 * dT3_dVg;
        dRds_dVb = T4 * dT3_dVb;
    }

    // ---- Abulk (:1350-1403)
    double Abulk, Abulk0, dAbulk_dVg, dAbulk_dVb, dAbulk0_dVb;
    {
        T9 = 0.5 * PF(k1ox) * Lpe_Vb / sqrtPhis;
        T1 = T9 + PF(k2ox) - PF(k3b) * Vth_NarrowW;
        const double dT1_dVb = -T9 / sqrtPhis * dsqrtPhis_dVb;
        T9 = sqrt(PF(xj) * Xdep);
        const double tmp1 = Leff + 2.0 * T9;
        T5 = Leff / tmp1;
        const double tmp2 = PF(a0) * T5;
        const double tmp3 = PF(weff) + PF(b1);
        const double tmp4 = PF(b0) / tmp3;
        T2 = tmp2 + tmp4;
        const double dT2_dVb = -T9 / tmp1 / Xdep * dXdep_dVb;
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
        double dT0_dVb;
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
    double ueff, dueff_dVg, dueff_dVd, dueff_dVb;
    {
        const double T14 = 0.0;
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
        double dDenomi_dVg = (T2 + 2.0 * PF(ub) * T3) / toxe;
        T13 = 2.0 * (dDenomi_dVg + T11 + T8);
        double dDenomi_dVd = T13 * dVth_dVd;
        double dDenomi_dVb = T13 * dVth_dVb + PF(uc) * T3;
        dDenomi_dVg += T7;
        double Denomi;
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
    const double WVCox = Weff * PF(vsattemp) * PF(coxe);
    const double WVCoxRds = WVCox * Rds;
    const double Esat = 2.0 * PF(vsattemp) / ueff;
    const double EsatL = Esat * Leff;
    T0 = -EsatL / ueff;
    const double dEsatL_dVg = T0 * dueff_dVg;
    const double dEsatL_dVd = T0 * dueff_dVd;
    const double dEsatL_dVb = T0 * dueff_dVb;
    const double Lambda = PF(a2);
    const double dLambda_dVg = 0.0;
    const double Vgst2Vtm = Vgsteff + 2.0 * Vtm;
    double tmp1, tmp2, tmp3;
    if (Rds > 0) {
        tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff;
        tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff;
    } else {
        tmp2 = dWeff_dVg / Weff;
        tmp3 = dWeff_dVb / Weff;
    }
    double Vdsat, dVdsat_dVg, dVdsat_dVd, dVdsat_dVb;
    if ((Rds == 0.0) && (Lambda == 1.0)) {
        T0 = 1.0 / (Abulk * EsatL + Vgst2Vtm);
        tmp1 = 0.0;
        T1 = T0 * T0;
        T2 = Vgst2Vtm * T0;
        T3 = EsatL * Vgst2Vtm;
        Vdsat = T3 * T0;
        const double dT0_dVg = -(Abulk * dEsatL_dVg + EsatL * dAbulk_dVg + 1.0) * T1;
        const double dT0_dVd = -(Abulk * dEsatL_dVd) * T1;
        const double dT0_dVb = -(Abulk * dEsatL_dVb + dAbulk_dVb * EsatL) * T1;
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
        const double dT0_dVg = 2.0 * (T8 * tmp2 - Abulk * tmp1 +
                                      (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
        const double dT0_dVb = 2.0 * (T8 * (2.0 / Abulk * dAbulk_dVb + tmp3) +
                                      (1.0 / Lambda - 1.0) * dAbulk_dVb);
        const double dT0_dVd = 0.0;
        T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + Abulk * EsatL + 3.0 * T7;
        const double dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + Abulk * dEsatL_dVg +
                               EsatL * dAbulk_dVg + 3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
        const double dT1_dVb =
            Abulk * dEsatL_dVb + EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
        const double dT1_dVd = Abulk * dEsatL_dVd;
        T2 = Vgst2Vtm * (EsatL + 2.0 * T6);
        const double dT2_dVg = EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
        const double dT2_dVb = Vgst2Vtm * (dEsatL_dVb + 2.0 * T6 * tmp3);
        const double dT2_dVd = Vgst2Vtm * dEsatL_dVd;
        T3 = sqrt(T1 * T1 - 2.0 * T0 * T2);
        Vdsat = (T1 - T3) / T0;
        dVdsat_dVg = (dT1_dVg - (T1 * dT1_dVg - dT0_dVg * T2 - T0 * dT2_dVg) / T3 - Vdsat * dT0_dVg) / T0;
        dVdsat_dVb = (dT1_dVb - (T1 * dT1_dVb - dT0_dVb * T2 - T0 * dT2_dVb) / T3 - Vdsat * dT0_dVb) / T0;
        dVdsat_dVd = (dT1_dVd - (T1 * dT1_dVd - T0 * dT2_dVd) / T3) / T0;
        (void)dT0_dVd;
    }

    // ---- Vdseff (:1619-1657)
    double Vdseff, dVdseff_dVg, dVdseff_dVd, dVdseff_dVb;
    {
        T1 = Vdsat - Vds - PF(delta);
        const double dT1_dVg = dVdsat_dVg;
        const double dT1_dVd = dVdsat_dVd - 1.0;
        const double dT1_dVb = dVdsat_dVb;
        T2 = sqrt(T1 * T1 + 4.0 * PF(delta) * Vdsat);
        T0 = T1 / T2;
        T9 = 2.0 * PF(delta);
        T3 = T9 / T2;
        const double dT2_dVg = T0 * dT1_dVg + T3 * dVdsat_dVg;
        const double dT2_dVd = T0 * dT1_dVd + T3 * dVdsat_dVd;
        const double dT2_dVb = T0 * dT1_dVb + T3 * dVdsat_dVb;
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
    const double diffVds = Vds - Vdseff;

    // ---- Vasat (:1702-1724)
    double Vasat, dVasat_dVg, dVasat_dVb, dVasat_dVd;
    {
        const double tmp4 = 1.0 - 0.5 * Abulk * Vdsat / Vgst2Vtm;
        T9 = WVCoxRds * Vgsteff;
        T8 = T9 / Vgst2Vtm;
        T0 = EsatL + Vdsat + 2.0 * T9 * tmp4;
        T7 = 2.0 * WVCoxRds * tmp4;
        const double dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * Vgsteff) -
                               T8 * (Abulk * dVdsat_dVg - Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
        const double dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * Vgsteff -
                               T8 * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
        const double dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * Abulk * dVdsat_dVd;
        T9 = WVCoxRds * Abulk;
        T1 = 2.0 / Lambda - 1.0 + T9;
        const double dT1_dVg = -2.0 * tmp1 + WVCoxRds * (Abulk * tmp2 + dAbulk_dVg);
        const double dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
        Vasat = T0 / T1;
        dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
        dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
        dVasat_dVd = dT0_dVd / T1;
    }

    // ---- Idl (:1728-1785)
    double Idl, dIdl_dVg, dIdl_dVd, dIdl_dVb;
    {
        tmp1 = PF(vtfbphi2);
        tmp2 = 2.0e8 * PF(toxp);
        double dT0_dVg = 1.0 / tmp2;
        T0 = (Vgsteff + tmp1) * dT0_dVg;
        tmp3 = exp(PF(bdos) * 0.7 * log(T0));
        T1 = 1.0 + tmp3;
        T2 = PF(bdos) * 0.7 * tmp3 / T0;
        const double Tcen = PF(ados) * 1.9e-9 / T1;
        const double dTcen_dVg = -Tcen * T2 * dT0_dVg / T1;
        const double Coxeff = epssub * PF(coxp) / (epssub + PF(coxp) * Tcen);
        const double dCoxeff_dVg = -Coxeff * Coxeff * dTcen_dVg / epssub;
        const double CoxeffWovL = Coxeff * Weff / Leff;
        const double beta = ueff * CoxeffWovL;
        T3 = ueff / Leff;
        const double dbeta_dVg = CoxeffWovL * dueff_dVg + T3 * (Weff * dCoxeff_dVg + Coxeff * dWeff_dVg);
        const double dbeta_dVd = CoxeffWovL * dueff_dVd;
        const double dbeta_dVb = CoxeffWovL * dueff_dVb + T3 * Coxeff * dWeff_dVb;
        const double AbovVgst2Vtm = Abulk / Vgst2Vtm;
        T0 = 1.0 - 0.5 * Vdseff * AbovVgst2Vtm;
        dT0_dVg = -0.5 * (Abulk * dVdseff_dVg - Abulk * Vdseff / Vgst2Vtm + Vdseff * dAbulk_dVg) / Vgst2Vtm;
        const double dT0_dVd = -0.5 * Abulk * dVdseff_dVd / Vgst2Vtm;
        const double dT0_dVb = -0.5 * (Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff) / Vgst2Vtm;
        const double fgche1 = Vgsteff * T0;
        const double dfgche1_dVg = Vgsteff * dT0_dVg + T0;
        const double dfgche1_dVd = Vgsteff * dT0_dVd;
        const double dfgche1_dVb = Vgsteff * dT0_dVb;
        T9 = Vdseff / EsatL;
        const double fgche2 = 1.0 + T9;
        const double dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / EsatL;
        const double dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / EsatL;
        const double dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / EsatL;
        const double gche = beta * fgche1 / fgche2;
        const double dgche_dVg = (beta * dfgche1_dVg + fgche1 * dbeta_dVg - gche * dfgche2_dVg) / fgche2;
        const double dgche_dVd = (beta * dfgche1_dVd + fgche1 * dbeta_dVd - gche * dfgche2_dVd) / fgche2;
        const double dgche_dVb = (beta * dfgche1_dVb + fgche1 * dbeta_dVb - gche * dfgche2_dVb) / fgche2;
        T0 = 1.0 + gche * Rds;
        Idl = gche / T0;
        T1 = (1.0 - Idl * Rds) / T0;
        T2 = Idl * Idl;
        dIdl_dVg = T1 * dgche_dVg - T2 * dRds_dVg;
        dIdl_dVd = T1 * dgche_dVd;
        dIdl_dVb = T1 * dgche_dVb - T2 * dRds_dVb;
    }

    // ---- output resistance: FP = 1 (FPROUT = 0), VACLM, VADIBL (:1789-1897)
    const double FP = 1.0, dFP_dVg = 0.0;
    double PvagTerm, dPvagTerm_dVg, dPvagTerm_dVb, dPvagTerm_dVd;
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
    double Cclm, dCclm_dVg, dCclm_dVd, dCclm_dVb, VACLM, dVACLM_dVg, dVACLM_dVd, dVACLM_dVb;
    if ((PF(pclm) > kMinExp) && (diffVds > 1.0e-10)) {
        T0 = 1.0 + Rds * Idl;
        const double dT0_dVg = dRds_dVg * Idl + Rds * dIdl_dVg;
        const double dT0_dVd = Rds * dIdl_dVd;
        const double dT0_dVb = dRds_dVb * Idl + Rds * dIdl_dVb;
        T2 = Vdsat / Esat;
        T1 = Leff + T2;
        const double dT1_dVg = (dVdsat_dVg - T2 * dEsatL_dVg / Leff) / Esat;
        const double dT1_dVd = (dVdsat_dVd - T2 * dEsatL_dVd / Leff) / Esat;
        const double dT1_dVb = (dVdsat_dVb - T2 * dEsatL_dVb / Leff) / Esat;
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
    double VADIBL, dVADIBL_dVg, dVADIBL_dVb, dVADIBL_dVd;
    if (PF(thetaRout) > kMinExp) {
        T8 = Abulk * Vdsat;
        T0 = Vgst2Vtm * T8;
        const double dT0_dVg = Vgst2Vtm * Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
        const double dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
        const double dT0_dVd = Vgst2Vtm * Abulk * dVdsat_dVd;
        T1 = Vgst2Vtm + T8;
        const double dT1_dVg = 1.0 + Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
        const double dT1_dVb = Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
        const double dT1_dVd = Abulk * dVdsat_dVd;
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
    const double Va = Vasat + VACLM;
    const double dVa_dVg = dVasat_dVg + dVACLM_dVg;
    const double dVa_dVb = dVasat_dVb + dVACLM_dVb;
    const double dVa_dVd = dVasat_dVd + dVACLM_dVd;
    // VADITS with PDITS = 0 (:1905-1926)
    const double VADITS = kMaxExp, dVADITS_dVg = 0.0, dVADITS_dVd = 0.0;
    // VASCBE (:1929-1947)
    double VASCBE, dVASCBE_dVg, dVASCBE_dVd, dVASCBE_dVb;
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
    double Idsa, dIdsa_dVg, dIdsa_dVd, dIdsa_dVb;
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
        const double dT0_dVg = dVa_dVg / Va - dVasat_dVg / Vasat;
        const double dT0_dVb = dVa_dVb / Va - dVasat_dVb / Vasat;
        const double dT0_dVd = dVa_dVd / Va - dVasat_dVd / Vasat;
        T1 = T0 / Cclm;
        T9 = 1.0 + T1;
        const double dT9_dVg2 = (dT0_dVg - T1 * dCclm_dVg) / Cclm;
        const double dT9_dVb2 = (dT0_dVb - T1 * dCclm_dVb) / Cclm;
        const double dT9_dVd2 = (dT0_dVd - T1 * dCclm_dVd) / Cclm;
        dIdsa_dVg = dIdsa_dVg * T9 + Idsa * dT9_dVg2;
        dIdsa_dVb = dIdsa_dVb * T9 + Idsa * dT9_dVb2;
        dIdsa_dVd = dIdsa_dVd * T9 + Idsa * dT9_dVd2;
        Idsa *= T9;
    }
    T9 = diffVds / VASCBE;
    T0 = 1.0 + T9;
    const double Ids = Idsa * T0;
    double Gm = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVASCBE_dVg) / VASCBE;
    double Gds = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVASCBE_dVd) / VASCBE;
    double Gmb = T0 * dIdsa_dVb - Idsa * (dVdseff_dVb + T9 * dVASCBE_dVb) / VASCBE;
    tmp1 = Gds + Gm * dVgsteff_dVd;
    tmp2 = Gmb + Gm * dVgsteff_dVb;
    tmp3 = Gm;
    Gm = (Ids * dVdseff_dVg + Vdseff * tmp3) * dVgsteff_dVg;
    Gds = Ids * (dVdseff_dVd + dVdseff_dVg * dVgsteff_dVd) + Vdseff * tmp1;
    Gmb = (Ids * (dVdseff_dVb + dVdseff_dVg * dVgsteff_dVb) + Vdseff * tmp2) * dVbseff_dVb;
    const double cdrain = Ids * Vdseff;
    const double h_gds_new = Gds, h_gm_new = Gm, h_gmbs_new = Gmb;

    // ---- C-V, capMod = 2 (charge-thickness model, :2910-3810), XPART < 0.5
    double qgate = 0.0, qbulk = 0.0, qdrn = 0.0;
    double cggb = 0.0, cgsb = 0.0, cgdb = 0.0, cdgb = 0.0, cdsb = 0.0, cddb = 0.0;
    double cbgb = 0.0, cbsb = 0.0, cbdb = 0.0;
    if (charge) {
        double VbseffCV, dVbseffCV_dVb;
        if (Vbseff < 0.0) {
            VbseffCV = Vbseff;
            dVbseffCV_dVb = 1.0;
        } else {
            VbseffCV = PF(phi) - Phis;
            dVbseffCV_dVb = -dPhis_dVb;
        }
        const double CoxWL = PF(coxe) * PF(weffCV) * PF(leffCV) * 1.0;
        {
            const double noff = n * PF(noff);
            const double dnoff_dVd = PF(noff) * dn_dVd;
            const double dnoff_dVb = PF(noff) * dn_dVb;
            T0 = Vtm * noff;
            const double voffcv = PF(voffcv);
            const double VgstNVt = (Vgst - voffcv) / T0;
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
                const double ExpVgst = exp(VgstNVt);
                Vgsteff = T0 * log(1.0 + ExpVgst);
                dVgsteff_dVg = ExpVgst / (1.0 + ExpVgst);
                dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd) +
                               Vgsteff / noff * dnoff_dVd;
                dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb) +
                               Vgsteff / noff * dnoff_dVb;
                dVgsteff_dVg *= dVgs_eff_dVg;
            }
        }
        double V3 = PF(vfbzb) - Vgs_eff + VbseffCV - kDelta3;
        if (PF(vfbzb) <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * kDelta3 * PF(vfbzb));
        else T0 = sqrt(V3 * V3 + 4.0 * kDelta3 * PF(vfbzb));
        T1 = 0.5 * (1.0 + V3 / T0);
        const double Vfbeff = PF(vfbzb) - 0.5 * (V3 + T0);
        const double dVfbeff_dVg = T1 * dVgs_eff_dVg;
        const double dVfbeff_dVb = -T1 * dVbseffCV_dVb;
        const double Cox = PF(coxp);
        double Tox = 1.0e8 * PF(toxp);
        T0 = (Vgs_eff - VbseffCV - PF(vfbzb)) / Tox;
        const double dT0_dVg = dVgs_eff_dVg / Tox;
        const double dT0_dVb = -dVbseffCV_dVb / Tox;
        double Tcen, dTcen_dVg, dTcen_dVb, dTcen_dVd;
        double tmp = T0 * PF(acde);
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
        const double LINK = 1.0e-3 * PF(toxp);
        V3 = PF(ldeb) - Tcen - LINK;
        const double V4 = sqrt(V3 * V3 + 4.0 * LINK * PF(ldeb));
        Tcen = PF(ldeb) - 0.5 * (V3 + V4);
        T1 = 0.5 * (1.0 + V3 / V4);
        dTcen_dVg *= T1;
        dTcen_dVb *= T1;
        double Ccen = epssub / Tcen;
        T2 = Cox / (Cox + Ccen);
        double Coxeff = T2 * Ccen;
        T3 = -Ccen / Tcen;
        double dCoxeff_dVg = T2 * T2 * T3;
        double dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        double CoxWLcen = CoxWL * Coxeff / PF(coxe);
        const double Qac0 = CoxWLcen * (Vfbeff - PF(vfbzb));
        double QovCox = Qac0 / Coxeff;
        const double dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVg;
        const double dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVb;
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
        const double Qsub0 = CoxWLcen * PF(k1ox) * (T1 - T0);
        QovCox = Qsub0 / Coxeff;
        const double dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg) + QovCox * dCoxeff_dVg;
        const double dQsub0_dVd = -T2 * dVgsteff_dVd;
        const double dQsub0_dVb =
            -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb) + QovCox * dCoxeff_dVb;
        // gate-bias dependent delta Phis
        double Denomi;
        if (PF(k1ox) <= 0.0) {
            Denomi = 0.25 * PF(moin) * Vtm;
            T0 = 0.5 * PF(sqrtPhi);
        } else {
            Denomi = PF(moin) * Vtm * PF(k1ox) * PF(k1ox);
            T0 = PF(k1ox) * PF(sqrtPhi);
        }
        T1 = 2.0 * T0 + Vgsteff;
        const double DeltaPhi = Vtm * log(1.0 + T1 * Vgsteff / Denomi);
        const double dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (Denomi + T1 * Vgsteff);
        T0 = Vgsteff - DeltaPhi - 0.001;
        const double dT0_dVg2 = 1.0 - dDeltaPhi_dVg;
        T1 = sqrt(T0 * T0 + Vgsteff * 0.004);
        const double VgDP = 0.5 * (T0 + T1);
        const double dVgDP_dVg = 0.5 * (dT0_dVg2 + (T0 * dT0_dVg2 + 0.002) / T1);
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
        const double dCoxeff_dVd = dCoxeff_dVg * dTcen_dVd;
        dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        CoxWLcen = CoxWL * Coxeff / PF(coxe);
        const double AbulkCV = Abulk0 * PF(abulkCVfactor);
        const double dAbulkCV_dVb = PF(abulkCVfactor) * dAbulk0_dVb;
        const double VdsatCV = VgDP / AbulkCV;
        T0 = VdsatCV - Vds - kDelta4;
        const double dT0_dVg3 = dVgDP_dVg / AbulkCV;
        const double dT0_dVb3 = -VdsatCV * dAbulkCV_dVb / AbulkCV;
        T1 = sqrt(T0 * T0 + 4.0 * kDelta4 * VdsatCV);
        double dT1_dVg = (T0 + kDelta4 + kDelta4) / T1;
        const double dT1_dVd = -T0 / T1;
        const double dT1_dVb = dT1_dVg * dT0_dVb3;
        dT1_dVg *= dT0_dVg3;
        double VdseffCV, dVdseffCV_dVg, dVdseffCV_dVd, dVdseffCV_dVb;
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
        double Cgg1 = CoxWLcen * (T4 * dVgDP_dVg + T5 * dVdseffCV_dVg);
        const double Cgd1 = CoxWLcen * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
        const double Cgb1 = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) +
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
        double Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
        const double Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
        const double Cbb1 = CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) +
                            Cbg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
        Cbg1 = Cbg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        // 40/60 partition
        T2 = T2 / 12.0;
        T3 = 0.5 * CoxWLcen / (T2 * T2);
        T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
        const double qsrc0 = -T3 * T4;
        QovCox = qsrc0 / Coxeff;
        T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
        T5 = -2.0 * qsrc0 / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
        T6 = AbulkCV * (qsrc0 / T2 + T3 * T8);
        T7 = T6 * VdseffCV / AbulkCV;
        double Csg = T5 * dVgDP_dVg + T6 * dVdseffCV_dVg;
        const double Csd = Csg * dVgsteff_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
        double Csb = Csg * dVgsteff_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVb;
        Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        qgate += Qac0 + Qsub0 - qbulk;
        qbulk -= (Qac0 + Qsub0);
        qdrn = -(qgate + qbulk + qsrc0);
        const double Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
        const double Cbd = Cbd1 - dQsub0_dVd;
        double Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
        const double Cgg = Cgg1 - Cbg;
        const double Cgd = Cgd1 - Cbd;
        double Cgb = Cgb1 - Cbb;
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
    double qbs = 0.0, qbd = 0.0, capbs = 0.0, capbd = 0.0;
    if (charge) {
        const double czbd = PF(czbd), czbs = PF(czbs), czbdsw = PF(czbdsw), czbdswg = PF(czbdswg);
        const double czbssw = PF(czbssw), czbsswg = PF(czbsswg);
        if (vbs_jct == 0.0) {
            qbs = 0.0;
            capbs = czbs + czbssw + czbsswg;
        } else if (vbs_jct < 0.0) {
            if (czbs > 0.0) {
                const double arg = 1.0 - vbs_jct / PF(PhiBS);
                const double sarg = (PF(MJS) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJS) * log(arg));
                qbs = PF(PhiBS) * czbs * (1.0 - arg * sarg) / (1.0 - PF(MJS));
                capbs = czbs * sarg;
            } else {
                qbs = 0.0;
                capbs = 0.0;
            }
            if (czbssw > 0.0) {
                const double arg = 1.0 - vbs_jct / PF(PhiBSWS);
                const double sarg = (PF(MJSWS) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJSWS) * log(arg));
                qbs += PF(PhiBSWS) * czbssw * (1.0 - arg * sarg) / (1.0 - PF(MJSWS));
                capbs += czbssw * sarg;
            }
            if (czbsswg > 0.0) {
                const double arg = 1.0 - vbs_jct / PF(PhiBSWGS);
                const double sarg = (PF(MJSWGS) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJSWGS) * log(arg));
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
                const double arg = 1.0 - vbd_jct / PF(PhiBD);
                const double sarg = (PF(MJD) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJD) * log(arg));
                qbd = PF(PhiBD) * czbd * (1.0 - arg * sarg) / (1.0 - PF(MJD));
                capbd = czbd * sarg;
            } else {
                qbd = 0.0;
                capbd = 0.0;
            }
            if (czbdsw > 0.0) {
                const double arg = 1.0 - vbd_jct / PF(PhiBSWD);
                const double sarg = (PF(MJSWD) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJSWD) * log(arg));
                qbd += PF(PhiBSWD) * czbdsw * (1.0 - arg * sarg) / (1.0 - PF(MJSWD));
                capbd += czbdsw * sarg;
            }
            if (czbdswg > 0.0) {
                const double arg = 1.0 - vbd_jct / PF(PhiBSWGD);
                const double sarg = (PF(MJSWGD) == 0.5) ? 1.0 / sqrt(arg) : exp(-PF(MJSWGD) * log(arg));
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
        s0[S(qbsS)] = qbs;
        s0[S(qbdS)] = qbd;
    }

    // ---- convergence (:3967-4011), with this load's currents
    int noncon = 0;
    if (Check == 1) {
        noncon = 1;
    } else if (!(mode & MODEADJOINT)) {
        const double Idtot = (dmode >= 0) ? cdrain - cbd : cdrain + cbd;
        const double tol0 = tol.reltol * ucb_max(fabs(cdhat), fabs(Idtot)) + tol.abstol;
        if (fabs(cdhat - Idtot) >= tol0) {
            noncon = 1;
        } else {
            const double Ibtot = cbs + cbd;
            const double tol6 = tol.reltol * ucb_max(fabs(cbhat), fabs(Ibtot)) + tol.abstol;
            if (fabs(cbhat - Ibtot) > tol6) noncon = 1;
        }
    }
    s0[S(vdsS)] = vds;
    s0[S(vgsS)] = vgs;
    s0[S(vbsS)] = vbs;
    s0[S(vbdS)] = vbd;

    // ---- charges and their companion conductances (:4026-4546)
    double gcggb = 0.0, gcgdb = 0.0, gcgsb = 0.0, gcgbb = 0.0, gcdgb = 0.0, gcddb = 0.0,
           gcdsb = 0.0, gcdbb = 0.0, gcsgb = 0.0, gcsdb = 0.0, gcssb = 0.0, gcsbb = 0.0,
           gcbdb = 0.0, gcbgb = 0.0, gcbsb = 0.0, gcbbb = 0.0;
    double ceqqg = 0.0, ceqqd = 0.0, ceqqb = 0.0;
    if (charge) {
        const double vgdx = vgd, vgsx = vgs;
        T0 = vgdx + kDelta1;
        T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
        T2 = 0.5 * (T0 - T1);
        T3 = PF(weffCV) * PF(cgdl);
        T4 = sqrt(1.0 - 4.0 * T2 / PF(ckappad));
        const double cgdo = PF(cgdo) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        const double qgdo = (PF(cgdo) + T3) * vgdx - T3 * (T2 + 0.5 * PF(ckappad) * (T4 - 1.0));
        T0 = vgsx + kDelta1;
        T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
        T2 = 0.5 * (T0 - T1);
        T3 = PF(weffCV) * PF(cgsl);
        T4 = sqrt(1.0 - 4.0 * T2 / PF(ckappas));
        const double cgso = PF(cgso) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        const double qgso = (PF(cgso) + T3) * vgsx - T3 * (T2 + 0.5 * PF(ckappas) * (T4 - 1.0));
        const double ag0 = c.ag0;
        double qsrc;
        if (dmode > 0) {
            qdrn -= qgdo;
            gcggb = (cggb + cgdo + cgso + PF(cgbo)) * ag0;
            gcgdb = (cgdb - cgdo) * ag0;
            gcgsb = (cgsb - cgso) * ag0;
            gcgbb = -(gcggb + gcgdb + gcgsb);
            gcdgb = (cdgb - cgdo) * ag0;
            gcsgb = -(cggb + cbgb + cdgb + cgso) * ag0;
            gcbgb = (cbgb - PF(cgbo)) * ag0;
            const double qgb = PF(cgbo) * vgb;
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
            const double qgb = PF(cgbo) * vgb;
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
        s0[S(qgS)] = qgate;
        s0[S(qdS)] = qdrn - qbd;
        s0[S(qsS)] = qsrc - qbs;
        s0[S(qbS)] = qbulk + qbd + qbs;
        if (mode & MODEINITTRAN) {
            s1[S(qbS)] = s0[S(qbS)];
            s1[S(qgS)] = s0[S(qgS)];
            s1[S(qdS)] = s0[S(qdS)];
        }
        double geq, ceq;
        niintegrate(c, s0, s1, s2, S(qbS), sB, 0.0, &geq, &ceq);
        niintegrate(c, s0, s1, s2, S(qgS), sB, 0.0, &geq, &ceq);
        niintegrate(c, s0, s1, s2, S(qdS), sB, 0.0, &geq, &ceq);
        const double cqgate = s0[S(cqgS)], cqbody = s0[S(cqbS)], cqdrn = s0[S(cqdS)];
        ceqqg = cqgate - gcggb * vgb + gcgdb * vbd + gcgsb * vbs;
        ceqqd = cqdrn - gcdgb * vgb + gcddb * vbd + gcdsb * vbs;
        ceqqb = cqbody - gcbgb * vgb + gcbdb * vbd + gcbsb * vbs;
        if (mode & MODEINITTRAN) {
            s1[S(cqbS)] = s0[S(cqbS)];
            s1[S(cqgS)] = s0[S(cqgS)];
            s1[S(cqdS)] = s0[S(cqdS)];
        }
    }

    // ---- currents into the right-hand side (:4643-4914)
    double Gmr, Gmbsr, FwdSum, RevSum, ceqdrn;
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
    double ceqjs, ceqjd;
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
    double* r = cr + (size_t)(t.rsite + 4 * d) * sB + inst;
    r[0] = (ceqjd - ceqdrn - ceqqd);                              // dNodePrime +=
    r[sB] = -(ceqqg);                                             // gNodePrime -=
    r[2 * sB] = (-ceqjd - ceqjs - ceqqb);                         // bNodePrime +=
    r[3 * sB] = (ceqdrn + ceqjs + ceqqg + ceqqb + ceqqd);         // sNodePrime +=

    // ---- matrix (:4980-5024), stamp order of the reference
    const double gjbd = gbd, gjbs = gbs;
    double* m = cm + (size_t)(t.site + 16 * d) * sB + inst;
    m[0] = gcggb;                                                  // GPgp
    m[sB] = gcgdb;                                                 // GPdp
    m[2 * sB] = gcgsb;                                             // GPsp
    m[3 * sB] = gcgbb;                                             // GPbp
    m[4 * sB] = (Gds + gbd + RevSum + gcddb);                      // DPdp
    m[5 * sB] = (Gmr + gcdgb);                                     // DPgp
    m[6 * sB] = -(Gds + FwdSum - gcdsb);                           // DPsp
    m[7 * sB] = -(gjbd - Gmbsr - gcdbb);                           // DPbp
    m[8 * sB] = -(Gds + RevSum - gcsdb);                           // SPdp
    m[9 * sB] = (gcsgb - Gmr);                                     // SPgp
    m[10 * sB] = (Gds + gbs + FwdSum + gcssb);                     // SPsp
    m[11 * sB] = -(gjbs + Gmbsr - gcsbb);                          // SPbp
    m[12 * sB] = (gcbdb - gjbd);                                   // BPdp
    m[13 * sB] = (gcbgb);                                          // BPgp
    m[14 * sB] = (gcbsb - gjbs);                                   // BPsp
    m[15 * sB] = (gjbd + gjbs + gcbbb);                            // BPbp

    aux[B4_VON * sB] = von_new;
    aux[B4_MODE * sB] = (double)dmode;
    aux[B4_CD * sB] = cdrain;
    aux[B4_CBS * sB] = cbs;
    aux[B4_CBD * sB] = cbd;
    aux[B4_GBS * sB] = gbs;
    aux[B4_GBD * sB] = gbd;
    aux[B4_GM * sB] = h_gm_new;
    aux[B4_GDS * sB] = h_gds_new;
    aux[B4_GMBS * sB] = h_gmbs_new;
    if (noncon) atomicOr(&devnc[ci], 1);
}

}  // namespace

void launch_bsim4(const B4Table& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                  const double* s2, const int32_t* done, double* cm, double* cr, int32_t* devnc,
                  DevTol tol, int B, hipStream_t s) {
    if (!t.count) return;
    hipLaunchKernelGGL(k_bsim4, dim3(blocks_for((long)t.count * B)), dim3(TPB), 0, s, t, ctx, x,
                       s0, s1, s2, done, cm, cr, devnc, tol, B);
}

}  // namespace nsa
