// Nonlinear device loads: level-1 diode (DIOload, src/devices/dio/dio_load.cpp)
// and MOS1 (MOS1load, src/devices/mos1/mos1_load.cpp), one lane per
// (device, instance) with the instance index fastest, so the lanes of a
// wave share one device's parameter record (a broadcast read) and touch
// consecutive instances of every state, site and solution entry.
//
// Each lane restates its device's load statement by statement, in the
// reference's operation order and without FMA contraction, and writes one
// value per stamp call into the device's sites; the gather sums the sites of
// a slot in load order.  Bypass is disabled in the reference
// (CKTbypass = 0, dio_device.cpp:207, mos1_device.cpp:207), so it is not
// restated; neither is the device-gain homotopy (device_gain_fact = 1).
// Transcendentals come from the device math library, so results agree with
// the reference's glibc values to the last few ulps rather than bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

#include "dev_limits.hpp"
#include "gpu_dev.hpp"

namespace nsa {

namespace {

using namespace devk;

constexpr int TPB = 256;
inline int blocks_for(long work) { return (int)((work + TPB - 1) / TPB); }

constexpr double kE = 2.7182818284590452354;  // CONSTe (dio_shim.hpp:26)
constexpr double kMaxExpArg = 709.0;          // MAX_EXP_ARG (mos1_shim.hpp:213-216)

// DEVqmeyer (mos1_shim.cpp:155-206)
__device__ void qmeyer(double vgs, double vgd, double von, double vdsat, double* capgs,
                       double* capgd, double* capgb, double phi, double cox) {
    const double magic_vds = 0.025;
    const double vgst = vgs - von;
    vdsat = fmax(vdsat, magic_vds);
    if (vgst <= -phi) {
        *capgb = cox / 2;
        *capgs = 0;
        *capgd = 0;
    } else if (vgst <= -phi / 2) {
        *capgb = -vgst * cox / (2 * phi);
        *capgs = 0;
        *capgd = 0;
    } else if (vgst <= 0) {
        *capgb = -vgst * cox / (2 * phi);
        *capgs = vgst * cox / (1.5 * phi) + cox / 3;
        const double vds = vgs - vgd;
        if (vds >= vdsat) {
            *capgd = 0;
        } else {
            const double vddif = 2.0 * vdsat - vds;
            const double vddif1 = vdsat - vds;
            const double vddif2 = vddif * vddif;
            *capgd = *capgs * (1.0 - vdsat * vdsat / vddif2);
            *capgs = *capgs * (1.0 - vddif1 * vddif1 / vddif2);
        }
    } else {
        const double vds = vgs - vgd;
        vdsat = fmax(vdsat, magic_vds);
        if (vdsat <= vds) {
            *capgs = cox / 3;
            *capgd = 0;
            *capgb = 0;
        } else {
            const double vddif = 2.0 * vdsat - vds;
            const double vddif1 = vdsat - vds;
            const double vddif2 = vddif * vddif;
            *capgd = cox * (1.0 - vdsat * vdsat / vddif2) / 3;
            *capgs = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3;
            *capgb = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// DIOload (dio_load.cpp:21-409) for one instance of one diode.
__global__ void k_diode(DioTable t, const DevCtx* __restrict__ ctx, const double* __restrict__ x,
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
    const DioParam& p = t.par[d];
    const int mode = c.mode;
    const size_t sv = (size_t)t.state[d] * sB + inst;  // DIOvoltage
    const size_t sc = sv + sB, sg = sv + 2 * sB, sq = sv + 3 * sB, si = sv + 4 * sB;
    const double v_pos = sol(x, t.nodes[2 * d], sB, inst);
    const double v_neg = sol(x, t.nodes[2 * d + 1], sB, inst);

    double cdb = 0.0, cdsw = 0.0, gdb = 0.0, gdsw = 0.0;
    double csat = p.csat;
    const double csatsw = p.csatsw;
    const double gspr = p.gspr;
    const double vte = p.vte, vtebrk = p.vtebrk, vterec = p.vterec;
    const int fl = p.flags;
    int Check = 1;
    double vd;
    if (mode & MODEADJOINT) {
        vd = v_pos - v_neg;
        Check = 0;
    } else if (mode & MODEINITSMSIG) {
        vd = s0[sv];
    } else if (mode & MODEINITTRAN) {
        vd = s1[sv];
    } else if ((mode & MODEINITJCT) && (mode & MODETRANOP) && (mode & MODEUIC)) {
        vd = 0.0;  // no IC= on diodes in this build
    } else if (mode & MODEINITJCT) {
        vd = p.tVcrit;
    } else {
        if (mode & MODEINITPRED) {
            s0[sv] = s1[sv];
            const double xfact = c.delta / c.delta_old1;  // DEVpred (dio_shim.hpp:201-205)
            vd = ((1.0 + xfact) * s1[sv]) - (xfact * s2[sv]);
            s0[sc] = s1[sc];
            s0[sg] = s1[sg];
        } else {
            vd = v_pos - v_neg;
        }
        if ((fl & DIO_BV) && (vd < umin(0, -p.tBrkdwnV + 10 * vtebrk))) {
            double vdtemp = -(vd + p.tBrkdwnV);
            vdtemp = pnjlim(vdtemp, -(s0[sv] + p.tBrkdwnV), vtebrk, p.tVcrit, &Check);
            vd = -(vdtemp + p.tBrkdwnV);
        } else {
            vd = pnjlim(vd, s0[sv], vte, p.tVcrit, &Check);
        }
    }
    // dc current and derivatives
    if (fl & DIO_SW) {
        if (fl & DIO_SWN) {
            const double vtesw = p.vtesw;
            if (vd >= -3 * vtesw) {
                const double evd = exp(vd / vtesw);
                cdsw = csatsw * (evd - 1);
                gdsw = csatsw * evd / vtesw;
            } else if ((!(fl & DIO_BV)) || vd >= -p.tBrkdwnV) {
                double argsw = 3 * vtesw / (vd * kE);
                argsw = argsw * argsw * argsw;
                cdsw = -csatsw * (1 + argsw);
                gdsw = csatsw * 3 * argsw / vd;
            } else {
                const double evrev = exp(-(p.tBrkdwnV + vd) / vtebrk);
                cdsw = -csatsw * evrev;
                gdsw = csatsw * evrev / vtebrk;
            }
        } else {
            csat = csat + csatsw;
        }
    }
    if (vd >= -3 * vte) {
        const double evd = exp(vd / vte);
        cdb = csat * (evd - 1);
        gdb = csat * evd / vte;
        if (fl & DIO_REC) {
            const double evd_rec = exp(vd / vterec);
            double cdb_rec = p.tRecSatCur * (evd_rec - 1);
            double gdb_rec = p.tRecSatCur * evd_rec / vterec;
            const double t1 = pow((1 - vd / p.tJctPot), 2) + 0.005;
            const double gen_fac = pow(t1, p.tGradingCoeff / 2);
            const double gen_fac_vd = -p.tGradingCoeff * (1 - vd / p.tJctPot) / p.tJctPot *
                                      pow(t1, (p.tGradingCoeff / 2 - 1));
            cdb_rec = cdb_rec * gen_fac;
            gdb_rec = gdb_rec * gen_fac + cdb_rec * gen_fac_vd;
            cdb = cdb + cdb_rec;
            gdb = gdb + gdb_rec;
        }
    } else if ((!(fl & DIO_BV)) || vd >= -p.tBrkdwnV) {
        double arg = 3 * vte / (vd * kE);
        arg = arg * arg * arg;
        cdb = -csat * (1 + arg);
        gdb = csat * 3 * arg / vd;
        if (fl & DIO_REC) {
            const double evd_rec = exp((-3 * vte) / vterec);
            const double cdb_rec = p.tRecSatCur * (evd_rec - 1);
            const double t1 = pow((1 - (-3 * vte) / p.tJctPot), 2) + 0.005;
            const double gen_fac = pow(t1, p.tGradingCoeff / 2);
            cdb = cdb + gen_fac * cdb_rec;
        }
    } else {
        const double evrev = exp(-(p.tBrkdwnV + vd) / vtebrk);
        if (fl & DIO_REC) {
            const double evd_rec = exp((-3 * vte) / vterec);
            const double cdb_rec = p.tRecSatCur * (evd_rec - 1);
            const double t1 = pow((1 - (-3 * vte) / p.tJctPot), 2) + 0.005;
            const double gen_fac = pow(t1, p.tGradingCoeff / 2);
            cdb = -csat * evrev + gen_fac * cdb_rec;
        } else {
            cdb = -csat * evrev;
        }
        gdb = csat * evrev / vtebrk;
    }
    if (fl & DIO_TUNSW) {
        const double evd = exp(-vd / p.vtetun);
        cdsw = cdsw - p.tTunSatSWCur * (evd - 1);
        gdsw = gdsw + p.tTunSatSWCur * evd / p.vtetun;
    }
    if (fl & DIO_TUN) {
        const double evd = exp(-vd / p.vtetun);
        cdb = cdb - p.tTunSatCur * (evd - 1);
        gdb = gdb + p.tTunSatCur * evd / p.vtetun;
    }
    double cd = cdb + cdsw;
    double gd = gdb + gdsw;
    const double gmin = dev_gmin(tol, ci);
    if (vd >= -3 * vte) {
        if ((fl & DIO_IKF) && (cd > 1.0e-18)) {
            const double ik = p.ikf;
            const double sq = sqrt(cd / ik);
            gd = ((1 + sq) * gd - cd * gd / (2 * sq * ik)) / (1 + 2 * sq + cd / ik) + gmin;
            cd = cd / (1 + sq) + gmin * vd;
        } else {
            gd = gd + gmin;
            cd = cd + gmin * vd;
        }
    } else {
        if ((fl & DIO_IKR) && (cd < -1.0e-18)) {
            const double ik = p.ikr;
            const double sq = sqrt(cd / (-ik));
            gd = ((1 + sq) * gd + cd * gd / (2 * sq * ik)) / (1 + 2 * sq - cd / ik) + gmin;
            cd = cd / (1 + sq) + gmin * vd;
        } else {
            gd = gd + gmin;
            cd = cd + gmin * vd;
        }
    }
    double* m = cm + (size_t)(t.site + 7 * d) * sB + inst;
    double* r = cr + (size_t)(t.rsite + 2 * d) * sB + inst;
    if ((mode & (MODETRAN | MODEAC | MODEINITSMSIG)) ||
        ((mode & MODETRANOP) && (mode & MODEUIC))) {
        // charge storage elements
        const double czero = p.tJctCap;
        double deplcharge, deplcap;
        if (vd < p.tDepCap) {
            const double arg = 1 - vd / p.tJctPot;
            const double sarg = exp(-p.tGradingCoeff * log(arg));
            deplcharge = p.tJctPot * czero * (1 - arg * sarg) / (1 - p.tGradingCoeff);
            deplcap = czero * sarg;
        } else {
            const double czof2 = czero / p.tF2;
            deplcharge = czero * p.tF1 +
                         czof2 * (p.tF3 * (vd - p.tDepCap) +
                                  (p.tGradingCoeff / (p.tJctPot + p.tJctPot)) *
                                      (vd * vd - p.tDepCap * p.tDepCap));
            deplcap = czof2 * (p.tF3 + p.tGradingCoeff * vd / p.tJctPot);
        }
        const double czeroSW = p.tJctSWCap;
        double deplchargeSW, deplcapSW;
        if (vd < p.tDepSWCap) {
            const double argSW = 1 - vd / p.tJctSWPot;
            const double sargSW = exp(-p.gradSW * log(argSW));
            deplchargeSW = p.tJctSWPot * czeroSW * (1 - argSW * sargSW) / (1 - p.gradSW);
            deplcapSW = czeroSW * sargSW;
        } else {
            const double czof2SW = czeroSW / p.tF2SW;
            deplchargeSW = czeroSW * p.tF1 +
                           czof2SW * (p.tF3SW * (vd - p.tDepSWCap) +
                                      (p.gradSW / (p.tJctSWPot + p.tJctSWPot)) *
                                          (vd * vd - p.tDepSWCap * p.tDepSWCap));
            deplcapSW = czof2SW * (p.tF3SW + p.gradSW * vd / p.tJctSWPot);
        }
        const double diffcharge = p.tTransitTime * cdb;
        const double diffchargeSW = p.tTransitTime * cdsw;
        s0[sq] = diffcharge + diffchargeSW + deplcharge + deplchargeSW;
        const double diffcap = p.tTransitTime * gdb;
        const double diffcapSW = p.tTransitTime * gdsw;
        const double capd = diffcap + diffcapSW + deplcap + deplcapSW;
        if ((!(mode & MODETRANOP)) || (!(mode & MODEUIC))) {
            if (mode & MODEINITSMSIG) {
                // small-signal setup stores the capacitance and stamps nothing
                s0[si] = capd;
                for (int k = 0; k < 7; ++k) m[k * sB] = 0.0;
                r[0] = 0.0;
                r[sB] = 0.0;
                return;
            }
            if (mode & MODEINITTRAN) s1[sq] = s0[sq];
            double geq, ceq;
            niintegrate(c, s0, s1, s2, sq, sB, capd, &geq, &ceq);
            gd = gd + geq;
            cd = cd + s0[si];
            if (mode & MODEINITTRAN) s1[si] = s0[si];
        }
    }
    // convergence: the device counts in CKTnoncon unless it is OFF in INITFIX
    int noncon = (Check == 1) ? 1 : 0;
    s0[sv] = vd;
    s0[sc] = cd;
    s0[sg] = gd;
    // DIODevice::conv_test on the iterate this load saw (dio_device.cpp:350-361)
    {
        const double vdo = v_pos - v_neg;
        const double delvd = vdo - s0[sv];
        const double cdhat = s0[sc] + s0[sg] * delvd;
        const double cdv = s0[sc];
        const double tl = tol.reltol * smax(fabs(cdhat), fabs(cdv)) + tol.abstol;
        if (!(fabs(cdhat - cdv) <= tl)) noncon = 1;
    }
    if (noncon) atomicOr(&devnc[ci], 1);
    const double cdeq = cd - gd * vd;
    r[0] = cdeq;        // CKTrhs[neg] += cdeq
    r[sB] = -cdeq;      // CKTrhs[posPrime] -= cdeq
    m[0] = gspr;                 // posPos
    m[sB] = gd;                  // negNeg
    m[2 * sB] = (gd + gspr);     // posPrimePosPrime
    m[3 * sB] = -(gspr);         // posPosPrime
    m[4 * sB] = -(gd);           // negPosPrime
    m[5 * sB] = -(gspr);         // posPrimePos
    m[6 * sB] = -(gd);           // posPrimeNeg
}

// ---------------------------------------------------------------------------
// MOS1load (mos1_load.cpp:21-896) for one instance of one MOS1 device.
__global__ void k_mos1(Mos1Table t, const DevCtx* __restrict__ ctx, const double* __restrict__ x,
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
    const Mos1Param& p = t.par[d];
    const int mode = c.mode;
    const size_t st = (size_t)t.state[d] * sB + inst;
    auto S = [&](int k) { return st + (size_t)k * sB; };
    enum { vbdS = 0, vbsS, vgsS, vdsS, capgsS, qgsS, cqgsS, capgdS, qgdS, cqgdS, capgbS, qgbS,
           cqgbS, qbdS, cqbdS, qbsS, cqbsS };
    double* aux = t.aux + (size_t)d * kMos1Aux * sB + inst;
    const int nD = t.nodes[4 * d], nG = t.nodes[4 * d + 1], nS = t.nodes[4 * d + 2],
              nB = t.nodes[4 * d + 3];
    const double type = p.type;
    const double vt = p.vt;
    const double gmin = dev_gmin(tol, ci);
    int Check = 1;
    double vbs, vgs, vds, vbd, vgd, vgb, von, vdsat;
    double xfact = 0.0;
    double capgs = 0.0, capgd = 0.0, capgb = 0.0;

    if ((mode & (MODEINITFLOAT | MODEINITPRED | MODEINITSMSIG | MODEINITTRAN)) ||
        ((mode & MODEINITFIX) && (!p.off))) {
        if (mode & (MODEINITPRED | MODEINITTRAN)) {
            // predictor step
            xfact = c.delta / c.delta_old1;
            s0[S(vbsS)] = s1[S(vbsS)];
            vbs = (1 + xfact) * (s1[S(vbsS)]) - (xfact * (s2[S(vbsS)]));
            s0[S(vgsS)] = s1[S(vgsS)];
            vgs = (1 + xfact) * (s1[S(vgsS)]) - (xfact * (s2[S(vgsS)]));
            s0[S(vdsS)] = s1[S(vdsS)];
            vds = (1 + xfact) * (s1[S(vdsS)]) - (xfact * (s2[S(vdsS)]));
            s0[S(vbdS)] = s0[S(vbsS)] - s0[S(vdsS)];
        } else {
            // general iteration (dNodePrime = d, sNodePrime = s)
            const double xs = sol(x, nS, sB, inst);
            vbs = type * (sol(x, nB, sB, inst) - xs);
            vgs = type * (sol(x, nG, sB, inst) - xs);
            vds = type * (sol(x, nD, sB, inst) - xs);
        }
        vbd = vbs - vds;
        vgd = vgs - vds;
        const double vgdo = s0[S(vgsS)] - s0[S(vdsS)];
        von = type * aux[M1_VON * sB];
        // limiting
        if (mode & MODEADJOINT) {
            Check = 0;
        } else if (s0[S(vdsS)] >= 0) {
            vgs = fetlim(vgs, s0[S(vgsS)], von);
            vds = vgs - vgd;
            vds = limvds(vds, s0[S(vdsS)]);
            vgd = vgs - vds;
        } else {
            vgd = fetlim(vgd, vgdo, von);
            vds = vgs - vgd;
            vds = -limvds(-vds, -(s0[S(vdsS)]));  // CKTfixLimit = 0
            vgs = vgd + vds;
        }
        if (mode & MODEADJOINT) {
        } else if (vds >= 0) {
            vbs = pnjlim(vbs, s0[S(vbsS)], vt, p.sourceVcrit, &Check);
            vbd = vbs - vds;
        } else {
            vbd = pnjlim(vbd, s0[S(vbdS)], vt, p.drainVcrit, &Check);
            vbs = vbd + vds;
        }
    } else {
        if ((mode & MODEINITJCT) && !p.off) {
            // no IC= on MOS1 in this build: icVDS = icVGS = icVBS = 0
            vds = type * 0.0;
            vgs = type * 0.0;
            vbs = type * 0.0;
            if ((vds == 0) && (vgs == 0) && (vbs == 0) &&
                ((mode & (MODETRAN | MODEDCOP | MODEDCTRANCURVE)) || (!(mode & MODEUIC)))) {
                vbs = -1;
                vgs = type * p.tVto;
                vds = 0;
            }
        } else {
            vbs = vgs = vds = 0;
        }
    }
    vbd = vbs - vds;
    vgd = vgs - vds;
    vgb = vgs - vbs;

    // bulk-source and bulk-drain diodes
    double gbs, cbs, gbd, cbd;
    if (vbs <= -3 * vt) {
        gbs = gmin;
        cbs = gbs * vbs - p.SourceSatCur;
    } else {
        const double evbs = exp(umin(kMaxExpArg, vbs / vt));
        gbs = p.SourceSatCur * evbs / vt + gmin;
        cbs = p.SourceSatCur * (evbs - 1) + gmin * vbs;
    }
    if (vbd <= -3 * vt) {
        gbd = gmin;
        cbd = gbd * vbd - p.DrainSatCur;
    } else {
        const double evbd = exp(umin(kMaxExpArg, vbd / vt));
        gbd = p.DrainSatCur * evbd / vt + gmin;
        cbd = p.DrainSatCur * (evbd - 1) + gmin * vbd;
    }
    const int dmode = (vds >= 0) ? 1 : -1;

    // Shichman-Hodges drain current
    double cdrain, gm, gds, gmbs;
    {
        const double vbx = (dmode == 1 ? vbs : vbd);
        double sarg;
        if (vbx <= 0) {
            sarg = sqrt(p.tPhi - vbx);
        } else {
            sarg = sqrt(p.tPhi);
            sarg = sarg - vbx / (sarg + sarg);
            sarg = umax(0, sarg);
        }
        von = (p.tVbi * type) + p.gamma * sarg;
        const double vgst = (dmode == 1 ? vgs : vgd) - von;
        vdsat = umax(vgst, 0);
        const double arg = (sarg <= 0) ? 0 : p.gamma / (sarg + sarg);
        if (vgst <= 0) {
            cdrain = 0;
            gm = 0;
            gds = 0;
            gmbs = 0;
        } else {
            const double vdsm = vds * dmode;
            const double betap = p.Beta * (1 + p.lambda * vdsm);
            if (vgst <= vdsm) {
                cdrain = betap * vgst * vgst * .5;
                gm = betap * vgst;
                gds = p.lambda * p.Beta * vgst * vgst * .5;
                gmbs = gm * arg;
            } else {
                cdrain = betap * vdsm * (vgst - .5 * vdsm);
                gm = betap * vdsm;
                gds = betap * (vgst - vdsm) + p.lambda * p.Beta * vdsm * (vgst - .5 * vdsm);
                gmbs = gm * arg;
            }
        }
    }
    aux[M1_VON * sB] = type * von;
    double cd = dmode * cdrain - cbd;

    if (mode & (MODETRAN | MODETRANOP | MODEINITSMSIG)) {
        // bulk junction depletion charges
        double capbs, capbd;
        if (p.Cbs != 0 || p.Cbssw != 0) {
            if (vbs < p.tDepCap) {
                const double arg = 1 - vbs / p.tBulkPot;
                double sarg, sargsw;
                if (p.mj == p.mjsw) {
                    if (p.mj == .5) {
                        sarg = sargsw = 1 / sqrt(arg);
                    } else {
                        sarg = sargsw = exp(-p.mj * log(arg));
                    }
                } else {
                    sarg = (p.mj == .5) ? 1 / sqrt(arg) : exp(-p.mj * log(arg));
                    sargsw = (p.mjsw == .5) ? 1 / sqrt(arg) : exp(-p.mjsw * log(arg));
                }
                s0[S(qbsS)] = p.tBulkPot * (p.Cbs * (1 - arg * sarg) / (1 - p.mj) +
                                            p.Cbssw * (1 - arg * sargsw) / (1 - p.mjsw));
                capbs = p.Cbs * sarg + p.Cbssw * sargsw;
            } else {
                s0[S(qbsS)] = p.f4s + vbs * (p.f2s + vbs * (p.f3s / 2));
                capbs = p.f2s + p.f3s * vbs;
            }
        } else {
            s0[S(qbsS)] = 0;
            capbs = 0;
        }
        if (p.Cbd != 0 || p.Cbdsw != 0) {
            if (vbd < p.tDepCap) {
                const double arg = 1 - vbd / p.tBulkPot;
                double sarg, sargsw;
                if (p.mj == .5 && p.mjsw == .5) {
                    sarg = sargsw = 1 / sqrt(arg);
                } else {
                    sarg = (p.mj == .5) ? 1 / sqrt(arg) : exp(-p.mj * log(arg));
                    sargsw = (p.mjsw == .5) ? 1 / sqrt(arg) : exp(-p.mjsw * log(arg));
                }
                s0[S(qbdS)] = p.tBulkPot * (p.Cbd * (1 - arg * sarg) / (1 - p.mj) +
                                            p.Cbdsw * (1 - arg * sargsw) / (1 - p.mjsw));
                capbd = p.Cbd * sarg + p.Cbdsw * sargsw;
            } else {
                s0[S(qbdS)] = p.f4d + vbd * (p.f2d + vbd * p.f3d / 2);
                capbd = p.f2d + vbd * p.f3d;
            }
        } else {
            s0[S(qbdS)] = 0;
            capbd = 0;
        }
        if ((mode & MODETRAN) || ((mode & MODEINITTRAN) && !(mode & MODEUIC))) {
            double geq, ceq;
            niintegrate(c, s0, s1, s2, S(qbdS), sB, capbd, &geq, &ceq);
            gbd += geq;
            cbd += s0[S(cqbdS)];
            cd -= s0[S(cqbdS)];
            niintegrate(c, s0, s1, s2, S(qbsS), sB, capbs, &geq, &ceq);
            gbs += geq;
            cbs += s0[S(cqbsS)];
        }
    }
    // convergence: OFF devices are exempt only in INITFIX / INITSMSIG
    const bool noncon = (p.off == 0 || !(mode & (MODEINITFIX | MODEINITSMSIG))) && Check == 1;
    if (noncon) atomicOr(&devnc[ci], 1);
    s0[S(vbsS)] = vbs;
    s0[S(vbdS)] = vbd;
    s0[S(vgsS)] = vgs;
    s0[S(vdsS)] = vds;

    // Meyer's gate capacitances
    if (mode & (MODETRAN | MODETRANOP | MODEINITSMSIG)) {
        if (dmode > 0) {
            qmeyer(vgs, vgd, von, vdsat, &s0[S(capgsS)], &s0[S(capgdS)], &s0[S(capgbS)], p.tPhi,
                   p.OxideCap);
        } else {
            qmeyer(vgd, vgs, von, vdsat, &s0[S(capgdS)], &s0[S(capgsS)], &s0[S(capgbS)], p.tPhi,
                   p.OxideCap);
        }
        const double vgs1 = s1[S(vgsS)];
        const double vgd1 = vgs1 - s1[S(vdsS)];
        const double vgb1 = vgs1 - s1[S(vbsS)];
        if (mode & (MODETRANOP | MODEINITSMSIG)) {
            capgs = 2 * s0[S(capgsS)] + p.GateSourceOverlapCap;
            capgd = 2 * s0[S(capgdS)] + p.GateDrainOverlapCap;
            capgb = 2 * s0[S(capgbS)] + p.GateBulkOverlapCap;
        } else {
            capgs = (s0[S(capgsS)] + s1[S(capgsS)] + p.GateSourceOverlapCap);
            capgd = (s0[S(capgdS)] + s1[S(capgdS)] + p.GateDrainOverlapCap);
            capgb = (s0[S(capgbS)] + s1[S(capgbS)] + p.GateBulkOverlapCap);
        }
        if (mode & (MODEINITPRED | MODEINITTRAN)) {
            s0[S(qgsS)] = (1 + xfact) * s1[S(qgsS)] - xfact * s2[S(qgsS)];
            s0[S(qgdS)] = (1 + xfact) * s1[S(qgdS)] - xfact * s2[S(qgdS)];
            s0[S(qgbS)] = (1 + xfact) * s1[S(qgbS)] - xfact * s2[S(qgbS)];
        } else if (mode & MODETRAN) {
            s0[S(qgsS)] = (vgs - vgs1) * capgs + s1[S(qgsS)];
            s0[S(qgdS)] = (vgd - vgd1) * capgd + s1[S(qgdS)];
            s0[S(qgbS)] = (vgb - vgb1) * capgb + s1[S(qgbS)];
        } else {
            s0[S(qgsS)] = vgs * capgs;
            s0[S(qgdS)] = vgd * capgd;
            s0[S(qgbS)] = vgb * capgb;
        }
    }
    double gcgs, ceqgs, gcgd, ceqgd, gcgb, ceqgb;
    if ((mode & MODEINITTRAN) || (!(mode & MODETRAN))) {
        gcgs = 0;
        ceqgs = 0;
        gcgd = 0;
        ceqgd = 0;
        gcgb = 0;
        ceqgb = 0;
    } else {
        if (capgs == 0) s0[S(cqgsS)] = 0;
        if (capgd == 0) s0[S(cqgdS)] = 0;
        if (capgb == 0) s0[S(cqgbS)] = 0;
        niintegrate(c, s0, s1, s2, S(qgsS), sB, capgs, &gcgs, &ceqgs);
        niintegrate(c, s0, s1, s2, S(qgdS), sB, capgd, &gcgd, &ceqgd);
        niintegrate(c, s0, s1, s2, S(qgbS), sB, capgb, &gcgb, &ceqgb);
        ceqgs = ceqgs - gcgs * vgs + c.ag0 * s0[S(qgsS)];
        ceqgd = ceqgd - gcgd * vgd + c.ag0 * s0[S(qgdS)];
        ceqgb = ceqgb - gcgb * vgb + c.ag0 * s0[S(qgbS)];
    }
    // fields the next load and the convergence test read
    aux[M1_MODE * sB] = (double)dmode;
    aux[M1_CD * sB] = cd;
    aux[M1_CBS * sB] = cbs;
    aux[M1_CBD * sB] = cbd;
    aux[M1_GBS * sB] = gbs;
    aux[M1_GBD * sB] = gbd;
    aux[M1_GM * sB] = gm;
    aux[M1_GDS * sB] = gds;
    aux[M1_GMBS * sB] = gmbs;

    // current vector
    const double ceqbs = type * (cbs - (gbs)*vbs);
    const double ceqbd = type * (cbd - (gbd)*vbd);
    int xnrm, xrev;
    double cdreq;
    if (dmode >= 0) {
        xnrm = 1;
        xrev = 0;
        cdreq = type * (cdrain - gds * vds - gm * vgs - gmbs * vbs);
    } else {
        xnrm = 0;
        xrev = 1;
        cdreq = -(type) * (cdrain - gds * (-vds) - gm * vgd - gmbs * vbd);
    }
    double* r = cr + (size_t)(t.rsite + 4 * d) * sB + inst;
    r[0] = -((type * (ceqgs + ceqgb + ceqgd)));          // G   -=
    r[sB] = -((ceqbs + ceqbd - type * ceqgb));           // B   -=
    r[2 * sB] = (ceqbd - cdreq + type * ceqgd);          // DP  +=
    r[3 * sB] = cdreq + ceqbs + type * ceqgs;            // SP  +=
    // Y matrix (drain/source conductances are 0 with RD = RS = 0)
    const double dcond = p.drainConductance, scond = p.sourceConductance;
    const double xd = (double)(xnrm - xrev);
    double* m = cm + (size_t)(t.site + 22 * d) * sB + inst;
    m[0] = (dcond);                                                       // Dd
    m[1 * sB] = ((gcgd + gcgs + gcgb));                                   // Gg
    m[2 * sB] = (scond);                                                  // Ss
    m[3 * sB] = (gbd + gbs + gcgb);                                       // Bb
    m[4 * sB] = (dcond + gds + gbd + xrev * (gm + gmbs) + gcgd);          // DPdp
    m[5 * sB] = (scond + gds + gbs + xnrm * (gm + gmbs) + gcgs);          // SPsp
    m[6 * sB] = (-dcond);                                                 // Ddp
    m[7 * sB] = -(gcgb);                                                  // Gb
    m[8 * sB] = -(gcgd);                                                  // Gdp
    m[9 * sB] = -(gcgs);                                                  // Gsp
    m[10 * sB] = (-scond);                                                // Ssp
    m[11 * sB] = -(gcgb);                                                 // Bg
    m[12 * sB] = -(gbd);                                                  // Bdp
    m[13 * sB] = -(gbs);                                                  // Bsp
    m[14 * sB] = (-dcond);                                                // DPd
    m[15 * sB] = (xd * gm - gcgd);                                        // DPg
    m[16 * sB] = (-gbd + xd * gmbs);                                      // DPb
    m[17 * sB] = (-gds - xnrm * (gm + gmbs));                             // DPsp
    m[18 * sB] = ((double)(-(xnrm - xrev)) * gm - gcgs);                  // SPg
    m[19 * sB] = (-scond);                                                // SPs
    m[20 * sB] = (-gbs - xd * gmbs);                                      // SPb
    m[21 * sB] = (-gds - xrev * (gm + gmbs));                             // SPdp
}

// MOS1Device::conv_test (mos1_device.cpp:376-416) on the proposed solution.
__global__ void k_mos1_conv(Mos1Table t, const double* __restrict__ x,
                            const double* __restrict__ s0, const int32_t* __restrict__ skip,
                            int32_t* __restrict__ devnc, DevTol tol, int B) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid >= (long)t.count * B) return;
    const int d = (int)(gid / B), inst = (int)(gid % B);
    const int ci = t.inst_of ? t.inst_of[gid] : inst;
    if (skip && skip[ci]) return;
    const size_t sB = (size_t)B;
    const Mos1Param& p = t.par[d];
    const size_t st = (size_t)t.state[d] * sB + inst;
    const double* aux = t.aux + (size_t)d * kMos1Aux * sB + inst;
    const double type = p.type;
    const double xs = sol(x, t.nodes[4 * d + 2], sB, inst);
    const double vbs = type * (sol(x, t.nodes[4 * d + 3], sB, inst) - xs);
    const double vgs = type * (sol(x, t.nodes[4 * d + 1], sB, inst) - xs);
    const double vds = type * (sol(x, t.nodes[4 * d], sB, inst) - xs);
    const double vbd = vbs - vds;
    const double vgd = vgs - vds;
    const double vgdo = s0[st + 2 * sB] - s0[st + 3 * sB];
    const double delvbs = vbs - s0[st + 1 * sB];
    const double delvbd = vbd - s0[st];
    const double delvgs = vgs - s0[st + 2 * sB];
    const double delvds = vds - s0[st + 3 * sB];
    const double delvgd = vgd - vgdo;
    const double cdv = aux[M1_CD * sB], gbd = aux[M1_GBD * sB], gmbs = aux[M1_GMBS * sB];
    const double gm = aux[M1_GM * sB], gds = aux[M1_GDS * sB], gbs = aux[M1_GBS * sB];
    double cdhat;
    if (aux[M1_MODE * sB] >= 0) {
        cdhat = cdv - gbd * delvbd + gmbs * delvbs + gm * delvgs + gds * delvds;
    } else {
        cdhat = cdv - (gbd - gmbs) * delvbd - gm * delvgd + gds * delvds;
    }
    const double cb = aux[M1_CBS * sB] + aux[M1_CBD * sB];
    const double cbhat = cb + gbd * delvbd + gbs * delvbs;
    double tl = tol.reltol * smax(fabs(cdhat), fabs(cdv)) + tol.abstol;
    bool ok = !(fabs(cdhat - cdv) >= tl);
    if (ok) {
        tl = tol.reltol * smax(fabs(cbhat), fabs(cb)) + tol.abstol;
        ok = fabs(cbhat - cb) <= tl;
    }
    if (!ok) atomicOr(&devnc[ci], 1);
}

__global__ void k_or_i32(int32_t* __restrict__ dst, const int32_t* __restrict__ src, int count) {
    const long gid = (long)blockIdx.x * TPB + threadIdx.x;
    if (gid < count) dst[gid] |= src[gid];
}

}  // namespace

void launch_diode(const DioTable& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                  const double* s2, const int32_t* done, double* cm, double* cr, int32_t* devnc,
                  DevTol tol, int B, hipStream_t s) {
    if (!t.count) return;
    hipLaunchKernelGGL(k_diode, dim3(blocks_for((long)t.count * B)), dim3(TPB), 0, s, t, ctx, x,
                       s0, s1, s2, done, cm, cr, devnc, tol, B);
}

void launch_mos1(const Mos1Table& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                 const double* s2, const int32_t* done, double* cm, double* cr, int32_t* devnc,
                 DevTol tol, int B, hipStream_t s) {
    if (!t.count) return;
    hipLaunchKernelGGL(k_mos1, dim3(blocks_for((long)t.count * B)), dim3(TPB), 0, s, t, ctx, x, s0,
                       s1, s2, done, cm, cr, devnc, tol, B);
}

void launch_mos1_conv(const Mos1Table& t, const double* x, const double* s0,
                      const int32_t* skip, int32_t* devnc, DevTol tol, int B, hipStream_t s) {
    if (!t.count) return;
    hipLaunchKernelGGL(k_mos1_conv, dim3(blocks_for((long)t.count * B)), dim3(TPB), 0, s, t, x,
                       s0, skip, devnc, tol, B);
}

void launch_or_i32(int32_t* dst, const int32_t* src, int count, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_or_i32, dim3(blocks_for(count)), dim3(TPB), 0, s, dst, src, count);
}

}  // namespace nsa
