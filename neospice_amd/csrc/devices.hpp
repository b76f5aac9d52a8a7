// Per-instance parameters of the nonlinear device families, computed once on
// the host and uploaded as read-only tables for the GPU loads.
//
//   * diode:  DIOsetup's defaulting (src/devices/dio/dio_setup.cpp:33-223)
//             and DIOtemp (dio_temp.cpp:24-300), with the card names of
//             DIOmPTable (dio_devsup.cpp:50-123) and DIOmParam
//             (dio_mpar.cpp:24-240);
//   * MOS1:   MOS1setup's defaulting (mos1_setup.cpp:33-126), MOS1temp
//             (mos1_temp.cpp:24-335) and the per-load constants at the top
//             of MOS1load (mos1_load.cpp:110-131), with the card names of
//             MOS1mPTable (mos1_devsup.cpp:125-158).
//
// The arithmetic is the reference's, in its order, on the host (glibc
// transcendentals, no FMA contraction), so every table entry is bit-equal
// to the instance field the reference computes.
#pragma once
#include <cstdint>
#include <string>

#include "circuit.hpp"

namespace nsa {

// Simulator settings a device load reads (SimOptions, src/core/types.hpp:30-60).
struct DeviceEnv {
    double temp = 300.15, tnom = 300.15;
    double gmin = 1e-12, reltol = 1e-3, abstol = 1e-12, vntol = 1e-6;
};

enum DioFlags : int32_t {
    DIO_BV = 1,       // model BV given (breakdown branch, limiting)
    DIO_SW = 2,       // JSW given (sidewall current)
    DIO_SWN = 4,      // NS given (own sidewall characteristic)
    DIO_REC = 8,      // ISR given (recombination current)
    DIO_TUNSW = 16,   // JTUNSW given
    DIO_TUN = 32,     // JTUN given
    DIO_IKF = 64,     // model IKF > 0
    DIO_IKR = 128,    // model IKR > 0
};

struct DioParam {
    double csat, csatsw, gspr, vt, vte, vtebrk, vterec, vtesw, vtetun;
    double tVcrit, tBrkdwnV, tRecSatCur, tJctPot, tGradingCoeff;
    double tTunSatSWCur, tTunSatCur, ikf, ikr;
    double tJctCap, tDepCap, gradSW, tJctSWCap, tDepSWCap, tJctSWPot;
    double tF1, tF2, tF3, tF2SW, tF3SW, tTransitTime;
    int32_t flags, pad;
};

struct Mos1Param {
    double type, vt, DrainSatCur, SourceSatCur;
    double GateSourceOverlapCap, GateDrainOverlapCap, GateBulkOverlapCap, Beta, OxideCap;
    double tPhi, tVbi, tVto, gamma, lambda, sourceVcrit, drainVcrit;
    double Cbs, Cbssw, Cbd, Cbdsw, tDepCap, tBulkPot, mj, mjsw;
    double f2s, f3s, f4s, f2d, f3d, f4d;
    double drainConductance, sourceConductance;
    int32_t off, pad;
};

// Throws Error(NSA_ERR_ARGUMENT) for an unknown card parameter and
// Error(NSA_ERR_UNSUPPORTED) for cards that would need internal nodes.
DioParam diode_param(const ModelCard& card, const Device& d, const DeviceEnv& env);
Mos1Param mos1_param(const ModelCard& card, const Device& d, const DeviceEnv& env);

// Whether the card creates internal nodes (RS for diodes; RD/RS/RSH for
// MOS1), which this build's circuit numbering does not model.
bool card_needs_internal_nodes(const ModelCard& card);

}  // namespace nsa
