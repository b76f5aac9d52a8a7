// BSIM4v7 (LEVEL=14) per-device constants for the GPU load, computed once
// on the host: the model defaulting of BSIM4v7setup
// (src/devices/bsim4v7/bsim4v7_setup.cpp:60-1500), the size-dependent and
// instance part of BSIM4v7temp (bsim4v7_temp.cpp:98-2317) and the geometry
// of BSIM4v7PAeffGeo (bsim4v7_geo.cpp:63-161).
//
// Scope: the model's default operating modes.  Cards are accepted when
// every parameter they give is in a list of value-only parameters (the
// threshold, mobility, velocity, output-resistance, short-channel and
// overlap parameters below), which change numbers but not which branches of
// BSIM4v7temp and BSIM4v7load run.  Everything else -- mode selectors
// (capMod, mobMod, rdsMod, rgateMod, rbodyMod, trnqsMod, igcMod, igbMod,
// gidlMod, tempMod, mtrlMod ...), binning (L/W/P prefixed), junction,
// gate-current, GIDL, impact-ionisation, stress and temperature parameters
// -- is refused with NSA_ERR_UNSUPPORTED, as are instances other than W, L
// with NF = 1 and simulation temperatures other than TNOM.  Under these
// cards the following terms of BSIM4v7load are identically zero and are not
// restated: substrate current, GIDL/GISL, gate tunnelling currents, the
// trap-assisted junction tunnelling currents, source/drain junction
// currents beyond gmin (zero area, zero sidewall saturation current),
// external source/drain and gate resistances, NQS.
//
// The arithmetic is the reference's, in its order, with glibc
// transcendentals, so every entry is bit-equal to the field the reference
// holds in its model, size-dependent (pParam) or instance record.
#pragma once
#include <cstdint>

#include "circuit.hpp"
#include "devices.hpp"

namespace nsa {

// The record with its numeric fields of type T (double for the load; dual
// numbers for the derivatives with respect to W and L).
template <class T>
struct B4ParamT {
    // model
    T type, vtm, coxe, coxp, toxe, toxp, factor1, vcrit, bdos, ados, temp_ratio;
    T PhiBS, PhiBSWS, PhiBSWGS, PhiBD, PhiBSWD, PhiBSWGD;
    T MJS, MJSWS, MJSWGS, MJD, MJSWD, MJSWGD;
    // junction capacitances of this instance (zero-bias, area/perimeter applied)
    T czbs, czbssw, czbsswg, czbd, czbdsw, czbdswg;
    // size-dependent (pParam)
    T leff, weff, leffCV, weffCV, phi, sqrtPhi, Xdep0, vbi, k1, k1ox, k3, k3b;
    T dvt0, dvt1, dvt2, dvt0w, dvt1w, dvt2w, lpe0, lpeb, kt1, kt1l, kt2, w0;
    T theta0vb0, etab, nfactor, cdsc, cdscb, cdscd, cit, mstar, voffcbn, cdep0;
    T dwg, dwb, prwg, prwb, rds0, rdswmin, xj, a0, ags, b0, b1, keta;
    T ua, ub, uc, ud, a2, delta, pvag, pclm, litl, thetaRout, pdiblb, pditsd;
    T pscbe1, pscbe2, noff, voffcv, abulkCVfactor, acde, ldeb, moin;
    T cgdo, cgso, cgbo, cgdl, cgsl, ckappad, ckappas;
    // instance
    T vth0, k2ox, eta0, vbsc, u0temp, vsattemp, vtfbphi2, vfbzb;
    int32_t off, pad;
};
using B4Param = B4ParamT<double>;

// The doubles before `off`: the record's numeric fields.
constexpr int kB4Fields = 110;

// Throws Error(NSA_ERR_UNSUPPORTED) for a card, instance or temperature
// outside the scope above and Error(NSA_ERR_ARGUMENT) for an invalid one.
B4Param bsim4_param(const ModelCard& card, const Device& d, const DeviceEnv& env);
// The record (val) and its derivatives with respect to the instance's W
// (dw) and L (dl), each kB4Fields doubles in field order, by forward-mode
// dual numbers through the same computation (val equals bsim4_param's).
void bsim4_param_derivs(const ModelCard& card, const Device& d, const DeviceEnv& env, double* val,
                        double* dw, double* dl);

// Whether a model card selects BSIM4v7: LEVEL=14, as the reference's
// registry has it (mosfet_common.cpp:479-481, bsim4v7_factory.cpp:8-10), or
// ngspice's LEVEL=54 for the same model.
inline bool is_bsim4_level(int level) { return level == 14 || level == 54; }

}  // namespace nsa
