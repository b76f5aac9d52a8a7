#include "devices.hpp"

#include <algorithm>
#include <cmath>
#include <map>

#include "common.hpp"

namespace nsa {

namespace {

// Physical constants as the reference's device shims define them
// (src/devices/dio/dio_shim.hpp:22-33, src/devices/ucb_compat.hpp:11-22).
constexpr double kKoverQ = 1.38064852e-23 / 1.6021766208e-19;
constexpr double kBoltz = 1.38064852e-23;
constexpr double kCharge = 1.6021766208e-19;
constexpr double kRefTemp = 300.15;
constexpr double kRoot2 = 1.4142135623730950488;
constexpr double kCtoK = 273.15;

// One model field: value and "given" flag, defaulted like the UCB setup code.
struct Field {
    double v = 0.0;
    bool given = false;
    void def(double d) { if (!given) v = d; }
};

// Card names -> canonical field key (aliases as in the reference's tables).
using Aliases = std::map<std::string, std::string>;

std::map<std::string, Field> collect(const ModelCard& card, const Aliases& names,
                                     const char* family) {
    std::map<std::string, Field> f;
    for (const auto& kv : card.p) {
        if (kv.first == "level") continue;
        auto it = names.find(kv.first);
        if (it == names.end()) continue;  // unknown names are ignored (model_card_utils.hpp:82-86)
        if (it->second == "!bad")
            throw Error(NSA_ERR_ARGUMENT, std::string("model '") + card.name + "': " + family +
                                              " parameter '" + kv.first + "' is not settable");
        Field& x = f[it->second];
        x.v = kv.second;
        x.given = true;
    }
    return f;
}

const Aliases& dio_names() {
    static const Aliases a = {
        {"is", "is"}, {"js", "is"}, {"jsw", "jsw"}, {"tnom", "tnom"}, {"tref", "tnom"},
        {"rs", "rs"}, {"trs", "trs1"}, {"trs1", "trs1"}, {"trs2", "trs2"}, {"n", "n"},
        {"ns", "ns"}, {"tt", "tt"}, {"ttt1", "ttt1"}, {"ttt2", "ttt2"}, {"cjo", "cjo"},
        {"cj0", "cjo"}, {"cj", "cjo"}, {"vj", "vj"}, {"pb", "vj"}, {"m", "m"}, {"mj", "m"},
        {"tm1", "tm1"}, {"tm2", "tm2"}, {"cjp", "cjsw"}, {"cjsw", "cjsw"}, {"php", "php"},
        {"mjsw", "mjsw"}, {"ikf", "ikf"}, {"ik", "ikf"}, {"ikr", "ikr"}, {"nbv", "nbv"},
        {"area", "area"}, {"pj", "pj"}, {"tlev", "tlev"}, {"tlevc", "tlevc"}, {"eg", "eg"},
        {"xti", "xti"}, {"cta", "cta"}, {"ctc", "cta"}, {"ctp", "ctp"}, {"tpb", "tpb"},
        {"tvj", "tpb"}, {"tphp", "tphp"}, {"jtun", "jtun"}, {"jtunsw", "jtunsw"},
        {"ntun", "ntun"}, {"xtitun", "xtitun"}, {"keg", "keg"}, {"kf", "kf"}, {"af", "af"},
        {"ef", "ef"}, {"fc", "fc"}, {"fcs", "fcs"}, {"bv", "bv"}, {"ibv", "ibv"}, {"ib", "ibv"},
        {"tcv", "tcv"}, {"fv_max", "fv_max"}, {"bv_max", "bv_max"}, {"isr", "isr"},
        {"nr", "nr"}, {"tikf", "tikf"}, {"d", "d"}, {"cond", "!bad"},
    };
    return a;
}

const Aliases& mos1_names() {
    static const Aliases a = {
        {"vto", "vto"}, {"vt0", "vto"}, {"kp", "kp"}, {"gamma", "gamma"}, {"phi", "phi"},
        {"lambda", "lambda"}, {"rd", "rd"}, {"rs", "rs"}, {"cbd", "cbd"}, {"cbs", "cbs"},
        {"is", "is"}, {"pb", "pb"}, {"cgso", "cgso"}, {"cgdo", "cgdo"}, {"cgbo", "cgbo"},
        {"rsh", "rsh"}, {"cj", "cj"}, {"mj", "mj"}, {"cjsw", "cjsw"}, {"mjsw", "mjsw"},
        {"js", "js"}, {"tox", "tox"}, {"ld", "ld"}, {"u0", "u0"}, {"uo", "u0"}, {"fc", "fc"},
        {"nsub", "nsub"}, {"tpg", "tpg"}, {"nss", "nss"}, {"tnom", "tnom"}, {"kf", "kf"},
        {"af", "af"},
    };
    return a;
}

}  // namespace

bool card_needs_internal_nodes(const ModelCard& card) {
    if (card.type == "d") return card.get("rs", 0.0) != 0.0;
    // mos1_setup.cpp:127-150: RD, RS, or RSH with the default one square
    return card.get("rd", 0.0) != 0.0 || card.get("rs", 0.0) != 0.0 ||
           card.get("rsh", 0.0) != 0.0;
}

DioParam diode_param(const ModelCard& card, const Device& d, const DeviceEnv& env) {
    if (card_needs_internal_nodes(card))
        throw Error(NSA_ERR_UNSUPPORTED, "diode '" + d.name + "': RS != 0 (internal node) not covered");
    auto f = collect(card, dio_names(), "DIO");
    // ---- DIOsetup model defaults (dio_setup.cpp:35-188)
    f["level"].v = card.level; f["level"].given = card.has("level");
    f["level"].def(1);
    f["n"].def(1); f["is"].def(1e-14); f["jsw"].def(0.0); f["ns"].def(1); f["ibv"].def(1e-3);
    f["vj"].def(1); f["m"].def(.5); f["tm1"].def(0.0); f["tm2"].def(0.0); f["fc"].def(.5);
    f["fcs"].def(.5); f["tt"].def(0); f["ttt1"].def(0.0); f["ttt2"].def(0.0); f["cjo"].def(0);
    f["cjsw"].def(0); f["php"].def(1); f["mjsw"].def(.33);
    f["ikf"].def(0.0);
    if (f["ikf"].given && f["ikf"].v < 1e-28) { f["ikf"].given = false; f["ikf"].v = 0.0; }
    f["ikr"].def(0.0);
    if (f["ikr"].given && f["ikr"].v < 1e-28) { f["ikr"].given = false; f["ikr"].v = 0.0; }
    f["nbv"].def(f["n"].v);
    f["tlev"].def(0); f["tlevc"].def(0); f["eg"].def(1.11); f["xti"].def(3); f["cta"].def(0.0);
    f["ctp"].def(0.0); f["tpb"].def(0.0); f["tphp"].def(0.0); f["trs1"].def(0.0);
    f["trs2"].def(0.0); f["tcv"].def(0.0); f["area"].def(1.0); f["pj"].def(0.0);
    f["jtun"].def(0.0); f["jtunsw"].def(0.0); f["ntun"].def(30.0); f["xtitun"].def(3.0);
    f["keg"].def(1.0); f["isr"].def(1e-14); f["nr"].def(2); f["tikf"].def(0.0);
    // ---- instance defaults (dio_setup.cpp:193-223): typed-API geometry is
    // area 1 (not given), pj 0, no w/l, m 1
    const double m = 1.0;
    double area = f["area"].v * m;
    double pj = f["pj"].v * m;
    double inst_ikf = f["ikf"].v * area;
    double inst_ikr = f["ikr"].v * area;
    const double junctionCap = f["cjo"].v * area;
    const double junctionSWCap = f["cjsw"].v * pj;
    // ---- DIOtemp model part (dio_temp.cpp:42-94)
    double nomTemp = f["tnom"].given ? f["tnom"].v + kCtoK : env.tnom;
    const double vtnom = kKoverQ * nomTemp;
    double grading = f["m"].v;
    if (grading > .9) grading = .9;
    double eg = f["eg"].v;
    if (eg < .1) eg = .1;
    double fc = f["fc"].v;
    if (fc > .95) fc = .95;
    double fcs = f["fcs"].v;
    if (fcs > .95) fcs = .95;
    const double conductance = 0.0;  // RS absent or zero, no PSpice compat mode
    const double xfc = std::log(1 - fc);
    const double xfcs = std::log(1 - fcs);
    // ---- DIOtemp instance part (dio_temp.cpp:96-297)
    const double temp = env.temp;  // no per-instance temp / dtemp / model T_ABS
    const double dt = temp - nomTemp;
    double factor = 1.0 + (f["tm1"].v * dt) + (f["tm2"].v * dt * dt);
    double tGrading = grading * factor;
    if (tGrading > .9) tGrading = .9;
    const double vt = kKoverQ * temp;
    const double fact2 = temp / kRefTemp;
    const double egfet = 1.16 - (7.02e-4 * temp * temp) / (temp + 1108);
    const double arg = -egfet / (2 * kBoltz * temp) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
    const double pbfact = -2 * vt * (1.5 * std::log(fact2) + kCharge * arg);
    const double egfet1 = 1.16 - (7.02e-4 * nomTemp * nomTemp) / (nomTemp + 1108);
    const double arg1 = -egfet1 / (kBoltz * 2 * nomTemp) + 1.1150877 / (2 * kBoltz * kRefTemp);
    const double fact1 = nomTemp / kRefTemp;
    const double pbfact1 = -2 * vtnom * (1.5 * std::log(fact1) + kCharge * arg1);
    double tJctCap = 0, tJctPot = 0, tJctSWCap = 0, tJctSWPot = 0;
    const double vj = f["vj"].v, php = f["php"].v, mjsw = f["mjsw"].v;
    const int tlevc = (int)f["tlevc"].v;
    if (tlevc == 0) {
        const double pbo = (vj - pbfact1) / fact1;
        const double gmaold = (vj - pbo) / pbo;
        tJctCap = junctionCap / (1 + tGrading * (400e-6 * (nomTemp - kRefTemp) - gmaold));
        tJctPot = pbfact + fact2 * pbo;
        const double gmanew = (tJctPot - pbo) / pbo;
        tJctCap *= 1 + tGrading * (400e-6 * (temp - kRefTemp) - gmanew);
    } else if (tlevc == 1) {
        tJctPot = vj - f["tpb"].v * (temp - kRefTemp);
        tJctCap = junctionCap * (1 + f["cta"].v * (temp - kRefTemp));
    }
    if (tlevc == 0) {
        const double pboSW = (php - pbfact1) / fact1;
        const double gmaSWold = (php - pboSW) / pboSW;
        tJctSWCap = junctionSWCap / (1 + mjsw * (400e-6 * (nomTemp - kRefTemp) - gmaSWold));
        tJctSWPot = pbfact + fact2 * pboSW;
        const double gmaSWnew = (tJctSWPot - pboSW) / pboSW;
        tJctSWCap *= 1 + mjsw * (400e-6 * (temp - kRefTemp) - gmaSWnew);
    } else if (tlevc == 1) {
        tJctSWPot = php - f["tphp"].v * (temp - kRefTemp);
        tJctSWCap = junctionSWCap * (1 + f["ctp"].v * (temp - kRefTemp));
    }
    const double n = f["n"].v, ns = f["ns"].v, nr = f["nr"].v, xti = f["xti"].v;
    const double tratio = temp / nomTemp;
    const double tSatCur = f["is"].v * area *
        std::exp(((tratio) - 1) * eg / (n * vt) + xti / n * std::log(tratio));
    const double tSatSWCur = f["jsw"].v * pj *
        std::exp(((tratio) - 1) * eg / (ns * vt) + xti / ns * std::log(tratio));
    const double tTunSatCur = f["jtun"].v * area *
        std::exp(((tratio) - 1) * f["keg"].v * eg / vt + f["xtitun"].v * std::log(tratio));
    const double tTunSatSWCur = f["jtunsw"].v * pj *
        std::exp(((tratio) - 1) * f["keg"].v * eg / vt + f["xtitun"].v * std::log(tratio));
    const double tRecSatCur = f["isr"].v * area * m *
        std::exp(((tratio) - 1) * eg / (nr * vt) + xti / nr * std::log(tratio));
    const double tF1 = tJctPot * (1 - std::exp((1 - tGrading) * xfc)) / (1 - tGrading);
    double tDepCap = fc * tJctPot;
    double tDepSWCap = fcs * tJctSWPot;
    const double vte = n * vt;
    const double tVcrit = vte * std::log(vte / (kRoot2 * tSatCur));
    if (tDepCap > 1.0) {
        tJctPot = 1.0 / fc;
        tDepCap = fc * tJctPot;
    }
    if (tDepSWCap > 1.0) {
        tJctSWPot = 1.0 / fcs;
        tDepSWCap = fcs * tJctSWPot;
    }
    double tBrkdwnV = 0.0;
    const double nbv = f["nbv"].v;
    if (f["bv"].given) {
        const double bv = f["bv"].v, tcv = f["tcv"].v;
        const double tBV = ((int)f["tlev"].v == 0) ? bv - tcv * dt : bv * (1 - tcv * dt);
        double cbv = ((int)f["level"].v == 1) ? f["ibv"].v : f["ibv"].v * area;
        double xbv;
        if (cbv < tSatCur * tBV / vt) {
            cbv = tSatCur * tBV / vt;
            xbv = tBV;
        } else {
            const double tol = env.reltol * cbv;
            xbv = tBV - nbv * vt * std::log(1 + cbv / (tSatCur));
            for (int iter = 0; iter < 25; iter++) {
                xbv = tBV - nbv * vt * std::log(cbv / (tSatCur) + 1 - xbv / vt);
                const double xcbv = tSatCur * (std::exp((tBV - xbv) / (nbv * vt)) - 1 + xbv / vt);
                if (std::fabs(xcbv - cbv) <= tol) break;
            }
        }
        tBrkdwnV = xbv;
    }
    factor = 1.0 + (f["ttt1"].v * dt) + (f["ttt2"].v * dt * dt);
    const double tTransitTime = f["tt"].v * factor;
    const double tConductance = conductance;
    if (f["tikf"].given && inst_ikf > 0.0) inst_ikf *= (1.0 + f["tikf"].v * dt);
    const double tF2 = std::exp((1 + tGrading) * xfc);
    const double tF3 = 1 - fc * (1 + tGrading);
    const double tF2SW = std::exp((1 + mjsw) * xfcs);
    const double tF3SW = 1 - fcs * (1 + mjsw);

    DioParam p{};
    p.csat = tSatCur;
    p.csatsw = tSatSWCur;
    p.gspr = tConductance * area;
    p.vt = vt;
    p.vte = n * vt;
    p.vtebrk = nbv * vt;
    p.vterec = nr * vt;
    p.vtesw = ns * vt;
    p.vtetun = f["ntun"].v * vt;
    p.tVcrit = tVcrit;
    p.tBrkdwnV = tBrkdwnV;
    p.tRecSatCur = tRecSatCur;
    p.tJctPot = tJctPot;
    p.tGradingCoeff = tGrading;
    p.tTunSatSWCur = tTunSatSWCur;
    p.tTunSatCur = tTunSatCur;
    p.ikf = inst_ikf;
    p.ikr = inst_ikr;
    p.tJctCap = tJctCap;
    p.tDepCap = tDepCap;
    p.gradSW = mjsw;
    p.tJctSWCap = tJctSWCap;
    p.tDepSWCap = tDepSWCap;
    p.tJctSWPot = tJctSWPot;
    p.tF1 = tF1;
    p.tF2 = tF2;
    p.tF3 = tF3;
    p.tF2SW = tF2SW;
    p.tF3SW = tF3SW;
    p.tTransitTime = tTransitTime;
    p.flags = (f["bv"].given ? DIO_BV : 0) | (f["jsw"].given ? DIO_SW : 0) |
              (f["ns"].given ? DIO_SWN : 0) | (f["isr"].given ? DIO_REC : 0) |
              (f["jtunsw"].given ? DIO_TUNSW : 0) | (f["jtun"].given ? DIO_TUN : 0) |
              (f["ikf"].v > 0.0 ? DIO_IKF : 0) | (f["ikr"].v > 0.0 ? DIO_IKR : 0);
    return p;
}

Mos1Param mos1_param(const ModelCard& card, const Device& d, const DeviceEnv& env) {
    if (card_needs_internal_nodes(card))
        throw Error(NSA_ERR_UNSUPPORTED,
                    "MOS1 '" + d.name + "': RD/RS/RSH != 0 (internal nodes) not covered");
    auto f = collect(card, mos1_names(), "MOS1");
    const double type = (card.type == "pmos") ? -1.0 : 1.0;
    // ---- MOS1setup model defaults (mos1_setup.cpp:33-96)
    f["ld"].def(0); f["js"].def(0); f["is"].def(1e-14); f["kp"].def(2e-5); f["cgso"].def(0);
    f["cgdo"].def(0); f["cgbo"].def(0); f["vto"].def(0); f["cj"].def(0); f["cjsw"].def(0);
    f["pb"].def(.8); f["mj"].def(.5); f["mjsw"].def(.5); f["fc"].def(.5); f["phi"].def(.6);
    f["lambda"].def(0); f["gamma"].def(0);
    // ---- MOS1temp model part (mos1_temp.cpp:42-117)
    const double tnom = f["tnom"].given ? f["tnom"].v + kCtoK : env.tnom;
    const double fact1 = tnom / kRefTemp;
    const double vtnom = tnom * kKoverQ;
    const double kt1 = kBoltz * tnom;
    const double egfet1 = 1.16 - (7.02e-4 * tnom * tnom) / (tnom + 1108);
    const double arg1 = -egfet1 / (kt1 + kt1) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
    const double pbfact1 = -2 * vtnom * (1.5 * std::log(fact1) + kCharge * arg1);
    if (f["phi"].v <= 0.0)
        throw Error(NSA_ERR_ARGUMENT, "MOS1 model '" + card.name + "': Phi is not positive");
    double oxideCapFactor = 0;
    if (f["tox"].given && f["tox"].v != 0) {
        oxideCapFactor = 3.9 * 8.854214871e-12 / f["tox"].v;
        if (!f["kp"].given) {
            if (!f["u0"].given) f["u0"].v = 600;
            f["kp"].v = f["u0"].v * oxideCapFactor * 1e-4;
        }
        if (f["nsub"].given) {
            const double nsub = f["nsub"].v;
            if (nsub * 1e6 > 1.45e16) {
                if (!f["phi"].given) {
                    f["phi"].v = 2 * vtnom * std::log(nsub * 1e6 / 1.45e16);
                    f["phi"].v = std::max(.1, f["phi"].v);
                }
                const double fermis = type * .5 * f["phi"].v;
                double wkfng = 3.2;
                if (!f["tpg"].given) f["tpg"].v = 1;
                if ((int)f["tpg"].v != 0) {
                    const double fermig = type * (int)f["tpg"].v * .5 * egfet1;
                    wkfng = 3.25 + .5 * egfet1 - fermig;
                }
                const double wkfngs = wkfng - (3.25 + .5 * egfet1 + fermis);
                if (!f["gamma"].given)
                    f["gamma"].v = std::sqrt(2 * 11.70 * 8.854214871e-12 * kCharge * nsub * 1e6) /
                                   oxideCapFactor;
                if (!f["vto"].given) {
                    if (!f["nss"].given) f["nss"].v = 0;
                    const double vfb = wkfngs - f["nss"].v * 1e4 * kCharge / oxideCapFactor;
                    f["vto"].v = vfb + type * (f["gamma"].v * std::sqrt(f["phi"].v) + f["phi"].v);
                }
            } else {
                throw Error(NSA_ERR_ARGUMENT, "MOS1 model '" + card.name + "': Nsub < Ni");
            }
        }
    }
    // ---- instance defaults (mos1_setup.cpp:98-126, mos1_temp.cpp:150-165):
    // typed-API devices give W and L; AD/AS/PD/PS 0, NRD/NRS 1, M 1, no IC
    const double m = 1.0;
    const double w = d.w_given ? d.w : 1e-4;
    const double l = d.l_given ? d.l : 1e-4;
    const double drainArea = 0.0, sourceArea = 0.0, drainPerim = 0.0, sourcePerim = 0.0;
    // ---- MOS1temp instance part (mos1_temp.cpp:131-333)
    const double temp = env.temp;
    const double vt = temp * kKoverQ;
    const double ratio = temp / tnom;
    const double fact2 = temp / kRefTemp;
    const double kt = temp * kBoltz;
    const double egfet = 1.16 - (7.02e-4 * temp * temp) / (temp + 1108);
    const double arg = -egfet / (kt + kt) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
    const double pbfact = -2 * vt * (1.5 * std::log(fact2) + kCharge * arg);
    const double ratio4 = ratio * std::sqrt(ratio);
    const double tTransconductance = f["kp"].v / ratio4;
    const double phi = f["phi"].v, gamma = f["gamma"].v;
    const double phio = (phi - pbfact1) / fact1;
    const double tPhi = fact2 * phio + pbfact;
    const double tVbi = f["vto"].v - type * (gamma * std::sqrt(phi)) + .5 * (egfet1 - egfet) +
                        type * .5 * (tPhi - phi);
    const double tVto = tVbi + type * gamma * std::sqrt(tPhi);
    const double tSatCur = f["is"].v * std::exp(-egfet / vt + egfet1 / vtnom);
    const double tSatCurDens = f["js"].v * std::exp(-egfet / vt + egfet1 / vtnom);
    const double pb = f["pb"].v, mj = f["mj"].v, mjsw = f["mjsw"].v;
    const double pbo = (pb - pbfact1) / fact1;
    const double gmaold = (pb - pbo) / pbo;
    double capfact = 1 / (1 + mj * (4e-4 * (tnom - kRefTemp) - gmaold));
    double tCbd = f["cbd"].v * capfact;
    double tCbs = f["cbs"].v * capfact;
    double tCj = f["cj"].v * capfact;
    capfact = 1 / (1 + mjsw * (4e-4 * (tnom - kRefTemp) - gmaold));
    double tCjsw = f["cjsw"].v * capfact;
    const double tBulkPot = fact2 * pbo + pbfact;
    const double gmanew = (tBulkPot - pbo) / pbo;
    capfact = (1 + mj * (4e-4 * (temp - kRefTemp) - gmanew));
    tCbd *= capfact;
    tCbs *= capfact;
    tCj *= capfact;
    capfact = (1 + mjsw * (4e-4 * (temp - kRefTemp) - gmanew));
    tCjsw *= capfact;
    const double fcoef = f["fc"].v;
    const double tDepCap = fcoef * tBulkPot;
    double sourceVcrit, drainVcrit;
    if (tSatCurDens == 0 || drainArea == 0 || sourceArea == 0) {
        sourceVcrit = drainVcrit = vt * std::log(vt / (kRoot2 * m * tSatCur));
    } else {
        drainVcrit = vt * std::log(vt / (kRoot2 * m * tSatCurDens * drainArea));
        sourceVcrit = vt * std::log(vt / (kRoot2 * m * tSatCurDens * sourceArea));
    }
    double czbd, czbdsw, czbs, czbssw;
    if (f["cbd"].given) czbd = tCbd * m;
    else czbd = f["cj"].given ? tCj * m * drainArea : 0;
    czbdsw = f["cjsw"].given ? tCjsw * drainPerim * m : 0;
    double a = 1 - fcoef;
    double sarg = std::exp((-mj) * std::log(a));
    double sargsw = std::exp((-mjsw) * std::log(a));
    const double f2d = czbd * (1 - fcoef * (1 + mj)) * sarg / a +
                       czbdsw * (1 - fcoef * (1 + mjsw)) * sargsw / a;
    const double f3d = czbd * mj * sarg / a / tBulkPot + czbdsw * mjsw * sargsw / a / tBulkPot;
    const double f4d = czbd * tBulkPot * (1 - a * sarg) / (1 - mj) +
                       czbdsw * tBulkPot * (1 - a * sargsw) / (1 - mjsw) -
                       f3d / 2 * (tDepCap * tDepCap) - tDepCap * f2d;
    if (f["cbs"].given) czbs = tCbs * m;
    else czbs = f["cj"].given ? tCj * sourceArea * m : 0;
    czbssw = f["cjsw"].given ? tCjsw * sourcePerim * m : 0;
    a = 1 - fcoef;
    sarg = std::exp((-mj) * std::log(a));
    sargsw = std::exp((-mjsw) * std::log(a));
    const double f2s = czbs * (1 - fcoef * (1 + mj)) * sarg / a +
                       czbssw * (1 - fcoef * (1 + mjsw)) * sargsw / a;
    const double f3s = czbs * mj * sarg / a / tBulkPot + czbssw * mjsw * sargsw / a / tBulkPot;
    const double f4s = czbs * tBulkPot * (1 - a * sarg) / (1 - mj) +
                       czbssw * tBulkPot * (1 - a * sargsw) / (1 - mjsw) -
                       f3s / 2 * (tDepCap * tDepCap) - tDepCap * f2s;
    // ---- per-load constants (mos1_load.cpp:110-131)
    const double EffectiveLength = l - 2 * f["ld"].v;
    Mos1Param p{};
    p.type = type;
    p.vt = kKoverQ * temp;
    if (tSatCurDens == 0 || drainArea == 0 || sourceArea == 0) {
        p.DrainSatCur = m * tSatCur;
        p.SourceSatCur = m * tSatCur;
    } else {
        p.DrainSatCur = tSatCurDens * m * drainArea;
        p.SourceSatCur = tSatCurDens * m * sourceArea;
    }
    p.GateSourceOverlapCap = f["cgso"].v * m * w;
    p.GateDrainOverlapCap = f["cgdo"].v * m * w;
    p.GateBulkOverlapCap = f["cgbo"].v * m * EffectiveLength;
    p.Beta = tTransconductance * m * w / EffectiveLength;
    p.OxideCap = oxideCapFactor * EffectiveLength * m * w;
    p.tPhi = tPhi;
    p.tVbi = tVbi;
    p.tVto = tVto;
    p.gamma = gamma;
    p.lambda = f["lambda"].v;
    p.sourceVcrit = sourceVcrit;
    p.drainVcrit = drainVcrit;
    p.Cbs = czbs;
    p.Cbssw = czbssw;
    p.Cbd = czbd;
    p.Cbdsw = czbdsw;
    p.tDepCap = tDepCap;
    p.tBulkPot = tBulkPot;
    p.mj = mj;
    p.mjsw = mjsw;
    p.f2s = f2s; p.f3s = f3s; p.f4s = f4s;
    p.f2d = f2d; p.f3d = f3d; p.f4d = f4d;
    p.drainConductance = 0.0;
    p.sourceConductance = 0.0;
    p.off = 0;
    return p;
}

}  // namespace nsa
