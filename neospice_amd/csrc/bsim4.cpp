// BSIM4v7 host constants: setup defaulting, BSIM4v7temp for one size and
// instance, BSIM4v7PAeffGeo for geoMod 0 -- restated for the default-mode
// cards described in bsim4.hpp.
#include "bsim4.hpp"

#include <cmath>
#include <set>
#include <string>

#include "common.hpp"
#include "dual.hpp"

namespace nsa {

namespace {

// Constants of bsim4v7_temp.cpp:37-46 and bsim4v7_setup.cpp:37-43 (PI is
// the truncated value both files define).
constexpr double kKboQ = 8.617087e-5;
constexpr double kEps0 = 8.85418e-12;
constexpr double kEpsSi = 1.03594e-10;
constexpr double kPi = 3.141592654;
constexpr double kMaxExp = 5.834617425e14;
constexpr double kMinExp = 1.713908431e-15;
constexpr double kExpThreshold = 34.0;
constexpr double kChargeQ = 1.60219e-19;
// ucb_compat.hpp:12-15
constexpr double kVt0 = 1.38064852e-23 * 300.15 / 1.6021766208e-19;
constexpr double kRoot2 = 1.4142135623730950488;

// Card parameters that only change values: each is read below with its
// BSIM4v7setup default (bsim4v7_setup.cpp, "if (!...Given)" lines).
const std::set<std::string>& value_params() {
    static const std::set<std::string> s = {
        "vth0", "u0", "toxe", "toxm", "vsat", "rdsw", "ua", "ub", "uc", "nfactor", "voff",
        "eta0", "etab", "dsub", "drout", "pclm", "pdiblc1", "pdiblc2", "pscbe1", "pscbe2",
        "delta", "keta", "a0", "cdsc", "ndep", "nsd", "xj", "k3", "dvt0", "dvt1", "dvt2",
        "w0", "lpe0", "minv", "cgso", "cgdo", "cgbo"};
    return s;
}

template <class T>
T smooth_exp_ratio(T T0) {
    // T1 / (T2^2 + 2 T1 MIN_EXP) with T1 = exp(T0), T2 = T1 - 1, and the
    // large-argument limit 1 / (MAX_EXP - 2) (bsim4v7_temp.cpp:1452-1460)
    if (T0 < kExpThreshold) {
        const T T1 = exp(T0);
        const T T2 = T1 - 1.0;
        const T T3 = T2 * T2;
        const T T4 = T3 + 2.0 * T1 * kMinExp;
        return T1 / T4;
    }
    return T(1.0 / (kMaxExp - 2.0));
}

template <class T>
B4ParamT<T> param_t(const ModelCard& card, const Device& d, const DeviceEnv& env, T Wd, T Ld) {
    for (const auto& kv : card.p) {
        if (kv.first == "level" || kv.first == "version") continue;
        if (!value_params().count(kv.first))
            throw Error(NSA_ERR_UNSUPPORTED,
                        "BSIM4v7 model " + card.name + ": parameter '" + kv.first +
                            "' is outside the covered default-mode subset");
    }
    if (!card.has("vth0"))
        throw Error(NSA_ERR_UNSUPPORTED, "BSIM4v7 model " + card.name + ": VTH0 must be given");
    if (env.temp != env.tnom)
        throw Error(NSA_ERR_UNSUPPORTED, "BSIM4v7: only TEMP == TNOM is covered");
    auto g = [&](const char* k, T dflt) -> T { return card.has(k) ? T(card.get(k, 0.0)) : dflt; };
    const bool nmos = card.type == "nmos";
    const T type = nmos ? 1.0 : -1.0;

    // ---- BSIM4v7setup defaults (model)
    const T toxe = g("toxe", 30.0e-10);
    const T toxm = g("toxm", toxe);
    const T epsrox = 3.9;
    const T xj = g("xj", .15e-6);
    const T ndep = g("ndep", 1.7e17);
    const T nsd = g("nsd", 1.0e20);
    const T nsub = 6.0e16, xt = 1.55e-7, vbm0 = -3.0, phin = 0.0;
    const T vth0 = g("vth0", nmos ? 0.7 : -0.7);
    const T u0 = g("u0", nmos ? 0.067 : 0.025);
    const T vsat = g("vsat", 8.0e4), at = 3.3e4, ute = -1.5, up = 0.0, lp = 1.0e-8;
    const T rdsw = g("rdsw", 200.0), rdswmin_m = 0.0, prt = 0.0, wr = 1.0;
    const T cdsc = g("cdsc", 2.4e-4), nfactor = g("nfactor", 1.0);
    const T voff = g("voff", -0.08), voffl = 0.0, minv = g("minv", 0.0);
    const T eta0 = g("eta0", 0.08), etab = g("etab", -0.07);
    const T drout = g("drout", 0.56), dsub = g("dsub", drout);
    const T pdibl1 = g("pdiblc1", 0.39), pdibl2 = g("pdiblc2", 0.0086);
    const T dvt0 = g("dvt0", 2.2), dvt1 = g("dvt1", 0.53), dvt2 = g("dvt2", -0.032);
    const T dvt0w = 0.0, dvt1w = 5.3e6, dvt2w = -0.032;
    const T k3 = g("k3", 80.0), w0 = g("w0", 2.5e-6), lpe0 = g("lpe0", 1.74e-7);
    const T kt1 = -0.11, kt1l = 0.0, kt2 = 0.022;
    const T clc = 0.1e-6, cle = 0.6, acde = 1.0;
    // cf (bsim4v7_setup.cpp): fringing capacitance from the oxide thickness
    const T cf = 2.0 * epsrox * kEps0 / kPi * log(1.0 + 0.4e-6 / toxe);
    const T Temp = env.temp, Tnom = env.tnom;

    B4ParamT<T> p{};
    p.type = type;
    p.ua = g("ua", 1.0e-9);
    p.ub = g("ub", 1.0e-19);
    p.uc = g("uc", -0.0465e-9);
    p.ud = 0.0;
    p.a0 = g("a0", 1.0);
    p.ags = 0.0;
    p.b0 = 0.0;
    p.b1 = 0.0;
    p.keta = g("keta", -0.047);
    p.a2 = 1.0;
    p.delta = g("delta", 0.01);
    p.pvag = 0.0;
    p.pclm = g("pclm", 1.3);
    p.pdiblb = 0.0;
    p.pditsd = 0.0;
    p.pscbe1 = g("pscbe1", 4.24e8);
    p.pscbe2 = g("pscbe2", 1.0e-5);
    p.noff = 1.0;
    p.cgdl = 0.0;
    p.cgsl = 0.0;
    p.ckappas = 0.6;
    p.ckappad = p.ckappas;
    p.moin = 15.0;
    p.cdscb = 0.0;
    p.cdscd = 0.0;
    p.cit = 0.0;
    p.k3b = 0.0;
    p.lpeb = 0.0;
    p.kt1 = kt1;
    p.kt1l = kt1l;
    p.kt2 = kt2;
    p.dvt0 = dvt0;
    p.dvt1 = dvt1;
    p.dvt2 = dvt2;
    p.dvt0w = dvt0w;
    p.dvt1w = dvt1w;
    p.dvt2w = dvt2w;
    p.k3 = k3;
    p.w0 = w0;
    p.lpe0 = lpe0;
    p.etab = etab;
    p.cdsc = cdsc;
    p.dwg = 0.0;
    p.dwb = 0.0;
    p.prwg = 1.0;
    p.prwb = 0.0;
    p.xj = xj;
    p.bdos = 1.0;
    p.ados = 1.0;
    p.off = 0;

    // ---- BSIM4v7temp, model part (bsim4v7_temp.cpp:100-409)
    const T epssub = kEpsSi;
    const T toxp = toxe - 0.0;  // TOXE given, TOXP not: toxp = toxe - dtox
    p.toxe = toxe;
    p.toxp = toxp;
    p.coxe = epsrox * kEps0 / toxe;
    p.coxp = epsrox * kEps0 / toxp;
    const T cgdo_m = card.has("cgdo") ? g("cgdo", 0.0) : 0.6 * xj * p.coxe;
    const T cgso_m = card.has("cgso") ? g("cgso", 0.0) : 0.6 * xj * p.coxe;
    const T dwc = 0.0;
    const T cgbo_m = card.has("cgbo") ? g("cgbo", 0.0) : 2.0 * dwc * p.coxe;
    const T TRatio = Temp / Tnom;
    p.vcrit = kVt0 * log(kVt0 / (kRoot2 * 1.0e-14));
    p.factor1 = sqrt(epssub / (epsrox * kEps0) * toxe);
    const T Vtm0 = kKboQ * Tnom;
    const T Eg0 = 1.16 - 7.02e-4 * Tnom * Tnom / (Tnom + 1108.0);
    const T ni = 1.45e10 * (Tnom / 300.15) * sqrt(Tnom / 300.15) *
                      exp(21.5565981 - Eg0 / (2.0 * Vtm0));
    p.vtm = kKboQ * Temp;
    const T delTemp = Temp - Tnom;
    // junctions: default potentials and grading, temperature factors at T = TNOM
    const T cj = 5.0e-4, cjsw = 5.0e-10, cjswg = cjsw;
    const T unitArea = cj * (1.0 + 0.0 * delTemp);
    const T unitSw = cjsw * (1.0 + 0.0 * delTemp);
    const T unitSwg = cjswg * (1.0 + 0.0 * delTemp);
    p.PhiBS = p.PhiBD = 1.0 - 0.0 * delTemp;
    p.PhiBSWS = p.PhiBSWD = 1.0 - 0.0 * delTemp;
    p.PhiBSWGS = p.PhiBSWGD = 1.0 - 0.0 * delTemp;
    p.MJS = p.MJD = 0.5;
    p.MJSWS = p.MJSWD = 0.33;
    p.MJSWGS = p.MJSWGD = 0.33;
    p.temp_ratio = Temp / Tnom - 1.0;  // TempRatio of bsim4v7_load.cpp:1110

    // ---- size-dependent part (bsim4v7_temp.cpp:433-1549), NF = 1
    const T nf = 1.0;
    if (!d.l_given || !d.w_given || !(Ld > 0.0) || !(Wd > 0.0))
        throw Error(NSA_ERR_ARGUMENT, "BSIM4v7 device " + d.name + ": W and L must be positive");
    const T Lnew = Ld + 0.0;
    const T Wnew = Wd / nf + 0.0;
    // Lint = Wint = 0 and no L/W dependence of the offsets: dl = dw = dlc = dwc = dwj = 0
    const T dl = 0.0 + (0.0 / Lnew + 0.0 / Wnew + 0.0 / (Lnew * Wnew));
    const T dw = dl;
    p.leff = Lnew - 2.0 * dl;
    p.weff = Wnew - 2.0 * dw;
    p.leffCV = Lnew - 2.0 * dl;
    p.weffCV = Wnew - 2.0 * dw;
    const T weffCJ = Wnew - 2.0 * dw;
    if (!(p.leff > 0.0) || !(p.weff > 0.0))
        throw Error(NSA_ERR_ARGUMENT, "BSIM4v7 device " + d.name + ": effective size <= 0");
    p.abulkCVfactor = 1.0 + pow((clc / p.leffCV), cle);
    const T T0t = TRatio - 1.0;
    const T PowWeffWr = pow(weffCJ * 1.0e6, wr) * nf;
    const T T10 = prt * T0t;
    p.rds0 = (rdsw + T10) * nf / PowWeffWr;
    p.rdswmin = (rdswmin_m + T10) * nf / PowWeffWr;
    {
        const T T5 = 1.0 - up * exp(-p.leff / lp);
        p.u0temp = u0 * T5 * pow(TRatio, ute);
    }
    p.vsattemp = vsat - at * T0t;
    const T voff_t = voff * (1.0 + 0.0 * delTemp);
    p.nfactor = nfactor + 0.0 * delTemp / Tnom;
    p.voffcv = 0.0 * (1.0 + 0.0 * delTemp);
    const T eta0_t = eta0 + 0.0 * delTemp / Tnom;
    p.cgdo = (cgdo_m + cf) * p.weffCV;
    p.cgso = (cgso_m + cf) * p.weffCV;
    p.cgbo = cgbo_m * p.leffCV * nf;
    p.phi = Vtm0 * log(ndep / ni) + phin + 0.4;
    p.sqrtPhi = sqrt(p.phi);
    p.Xdep0 = sqrt(2.0 * epssub / (kChargeQ * ndep * 1.0e6)) * p.sqrtPhi;
    p.litl = sqrt(3.0 * 3.9 / epsrox * xj * toxe);
    p.vbi = Vtm0 * log(nsd * ndep / (ni * ni));
    p.cdep0 = sqrt(kChargeQ * epssub * ndep * 1.0e6 / 2.0 / p.phi);
    p.mstar = 0.5 + atan(minv) / kPi;
    p.voffcbn = voff_t + voffl / p.leff;
    p.ldeb = sqrt(epssub * Vtm0 / (kChargeQ * ndep * 1.0e6)) / 3.0;
    p.acde = acde * pow((ndep / 2.0e16), -0.25);
    // K1, K2 from the doping profile (bsim4v7_temp.cpp:1387-1413)
    T vbx = p.phi - 7.7348e-4 * ndep * xt * xt;
    if (vbx > 0.0) vbx = -vbx;
    T vbm = vbm0;
    if (vbm > 0.0) vbm = -vbm;
    const T gamma1 = 5.753e-12 * sqrt(ndep) / p.coxe;
    const T gamma2 = 5.753e-12 * sqrt(nsub) / p.coxe;
    T k2;
    {
        const T T0 = gamma1 - gamma2;
        const T T1 = sqrt(p.phi - vbx) - p.sqrtPhi;
        const T T2 = sqrt(p.phi * (p.phi - vbm)) - p.phi;
        k2 = T0 * T1 / (2.0 * T2 + vbm);
        p.k1 = gamma2 - 2.0 * k2 * sqrt(p.phi - vbm);
    }
    const T vfb = type * vth0 - p.phi - p.k1 * p.sqrtPhi;
    p.k1ox = p.k1 * toxe / toxm;
    {
        const T tmp = sqrt(epssub / (epsrox * kEps0) * toxe * p.Xdep0);
        p.theta0vb0 = smooth_exp_ratio(dsub * p.leff / tmp);
        const T T5 = smooth_exp_ratio(drout * p.leff / tmp);
        p.thetaRout = pdibl1 * T5 + pdibl2;
    }
    T vfbzbfactor;
    {
        const T tmp = sqrt(p.Xdep0);
        const T tmp1 = p.vbi - p.phi;
        const T tmp2 = p.factor1 * tmp;
        T T8 = smooth_exp_ratio(dvt1w * p.weff * p.leff / tmp2);
        T T0 = dvt0w * T8;
        T8 = T0 * tmp1;
        T T9 = smooth_exp_ratio(dvt1 * p.leff / tmp2);
        T9 = dvt0 * T9 * tmp1;
        const T T4 = toxe * p.phi / (p.weff + w0);
        T0 = sqrt(1.0 + lpe0 / p.leff);
        const T T3 = (kt1 + kt1l / p.leff) * (TRatio - 1.0);
        const T T5 = p.k1ox * (T0 - 1.0) * p.sqrtPhi + T3;
        vfbzbfactor = -T8 - T9 + k3 * T4 + T5 - p.phi - p.k1 * p.sqrtPhi;
    }

    // ---- instance part (bsim4v7_temp.cpp:1591-1686), no stress, DELVTO = 0
    p.vth0 = vth0 + 0.0;
    p.vsattemp = p.vsattemp;
    p.eta0 = eta0_t;
    const T vfb_i = vfb + type * 0.0;
    {
        const T T3 = type * p.vth0 - vfb_i - p.phi;
        p.vtfbphi2 = 4.0 * T3;
        if (p.vtfbphi2 < 0.0) p.vtfbphi2 = 0.0;
    }
    if (k2 < 0.0) {
        const T T0 = 0.5 * p.k1 / k2;
        p.vbsc = 0.9 * (p.phi - T0 * T0);
        if (p.vbsc > -3.0) p.vbsc = -3.0;
        else if (p.vbsc < -30.0) p.vbsc = -30.0;
    } else {
        p.vbsc = -30.0;
    }
    if (p.vbsc > vbm) p.vbsc = vbm;
    p.k2ox = k2 * toxe / toxm;
    p.vfbzb = vfbzbfactor + type * p.vth0;

    // ---- geometry (BSIM4v7PAeffGeo, geoMod 0, one finger, DMCG = DMCI = 0):
    // perimeter = weffCJ, area = 0 on both sides
    const T T0g = 0.0 + 0.0;
    const T Pseff = 1.0 * (T0g + T0g + weffCJ) + 0.0 * (0.0 + 0.0);
    const T Aseff = 1.0 * (T0g * weffCJ) + 0.0 * (0.0 * weffCJ);
    // zero saturation current (JSS applies to area 0, JSSWS = JSSWGS = 0):
    // the junction currents reduce to gmin (bsim4v7_load.cpp:682-695)
    p.czbd = unitArea * Aseff;
    p.czbs = unitArea * Aseff;
    p.czbdsw = unitSw * Pseff;
    p.czbdswg = unitSwg * weffCJ * nf;
    p.czbssw = unitSw * Pseff;
    p.czbsswg = unitSwg * weffCJ * nf;
    return p;
}

}  // namespace

B4Param bsim4_param(const ModelCard& card, const Device& d, const DeviceEnv& env) {
    return param_t<double>(card, d, env, d.w, d.l);
}

void bsim4_param_derivs(const ModelCard& card, const Device& d, const DeviceEnv& env, double* val,
                        double* dw, double* dl) {
    using D2 = Dual<2>;
    D2 W(d.w), L(d.l);
    W.d[0] = 1.0;
    L.d[1] = 1.0;
    const B4ParamT<D2> p = param_t<D2>(card, d, env, W, L);
    const D2* f = reinterpret_cast<const D2*>(&p);
    for (int q = 0; q < kB4Fields; ++q) {
        val[q] = f[q].v;
        dw[q] = f[q].d[0];
        dl[q] = f[q].d[1];
    }
}

}  // namespace nsa
