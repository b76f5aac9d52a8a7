// Device-evaluation and Newton helper kernels (launch wrappers).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "devices.hpp"

namespace nsa {

// Per-instance integrator context (the fields of the reference's
// IntegratorCtx that device loads read, src/core/types.hpp:73-94).
struct DevCtx {
    int32_t mode;
    int32_t order;
    int32_t method;
    int32_t iter;     // Newton iteration of the current load (0: the solve's first)
    double ag0, ag1, ag2;
    double xmu_ratio;
    double time;      // current_time of the step being solved
    double src_fact;  // SimOptions::src_fact
    double delta;       // CKTdelta: the step being solved
    double delta_old1;  // CKTdeltaOld[1]: the previous step (predictor ratio)
    double delta_old2;  // CKTdeltaOld[2] (truncation error of order 2)
};

// CKTmode bits (ngspice cktdefs.h, used by src/core/newton.cpp:109-115).
enum ModeBits : int32_t {
    MODETRAN = 0x1,
    MODEAC = 0x2,
    MODEDCOP = 0x10,
    MODETRANOP = 0x20,
    MODEDCTRANCURVE = 0x40,
    MODEDC = 0x70,
    MODEINITFLOAT = 0x100,
    MODEINITJCT = 0x200,
    MODEINITFIX = 0x400,
    MODEINITSMSIG = 0x800,
    MODEINITTRAN = 0x1000,
    MODEINITPRED = 0x2000,
    MODEUIC = 0x10000,
    INITF_MASK = 0x3F00,
    // Not an ngspice bit: evaluate the devices at the given solution with no
    // junction limiting or prediction, for the adjoint's Jacobian at a
    // converged point (where the forward's last load did not limit either).
    MODEADJOINT = 0x40000000,
};

struct CapTable {
    int32_t count;
    const int32_t* nodes;      // 2 per device
    const int32_t* state;      // state base per device
    const double* farads;      // count x B
    int32_t site;              // first matrix site (4 per device)
    int32_t rsite;             // first rhs site (2 per device)
    const int32_t* inst_of;    // null, or the context instance of every lane (see below)
};

struct VsrcTable {
    int32_t count;
    const int32_t* wave;       // WAVE_* per device
    const double* dc;          // count x B
    const int32_t* dc_given;   // per device
    const double* pulse;       // 7 per device
    const int32_t* pwl_ptr;    // count+1
    const double* pwl_t;
    const double* pwl_v;
    int32_t rsite;             // 1 per device
    const int32_t* inst_of;
};

struct IsrcTable {
    int32_t count;
    const double* dc;          // count x B
    int32_t rsite;             // 2 per device
    const int32_t* inst_of;
};

// Level-1 diodes (RS = 0: posPrime is the anode).  7 matrix sites and 2 rhs
// sites per device, in DIOload's stamp order.
struct DioTable {
    int32_t count;
    const int32_t* nodes;   // 2 per device: anode, cathode (-1 = ground)
    const int32_t* state;   // state base per device (5 slots)
    const DioParam* par;    // per device
    int32_t site, rsite;
    const int32_t* inst_of;
};

// MOS1 (RD = RS = 0: dNodePrime = d, sNodePrime = s).  22 matrix sites and
// 4 rhs sites per device, in MOS1load's stamp order.  `aux` holds the
// instance fields that outlive one load (count x kMos1Aux x B).
enum Mos1Aux : int32_t {
    M1_VON = 0, M1_MODE, M1_CD, M1_CBS, M1_CBD, M1_GBS, M1_GBD, M1_GM, M1_GDS, M1_GMBS, kMos1Aux
};
struct Mos1Table {
    int32_t count;
    const int32_t* nodes;   // 4 per device: d, g, s, b
    const int32_t* state;   // state base per device (17 slots)
    const Mos1Param* par;   // per device
    double* aux;
    int32_t site, rsite;
    const int32_t* inst_of;
};

// BSIM4v7: the instance fields BSIM4v7load reads back on the next load
// (the convergence test's linearisation and the limiting's von).
enum B4Aux : int32_t {
    B4_VON = 0, B4_MODE, B4_CD, B4_CBS, B4_CBD, B4_GBS, B4_GBD, B4_GM, B4_GDS, B4_GMBS, kB4Aux
};
template <class T>
struct B4ParamT;
using B4Param = B4ParamT<double>;
struct B4Table {
    int32_t count;
    const int32_t* nodes;   // 4 per device: d, g, s, b
    const int32_t* state;   // state base per device (29 slots)
    const B4Param* par;     // per device, shared by the instances
    // per-instance constants instead (null: use par): for device d, field f
    // and instance i at par_inst[(d * fields + f) * B + i]
    const double* par_inst;
    double* aux;            // count x kB4Aux x B
    int32_t site, rsite;    // 16 matrix sites, 4 rhs sites per device
    const int32_t* inst_of;
    // null, or the terminals' solution indices per lane (4 x B, terminal
    // major, -1 = ground) in place of nodes[]: a packed batch of circuits
    // (one device per lane) reads its instance's own unknowns
    const int32_t* pin;
};

// `inst_of` (every table): null for a batch of one circuit, where lane
// (d, inst) reads ctx[inst], done[inst] and sets devnc[inst].  A packed
// batch of different circuits lays each family out as one batch whose lanes
// are the devices of every instance; inst_of[lane] (lane = d * B + inst)
// then names the instance whose context, completion flag and
// non-convergence flag the lane uses.

// Tolerances a device load reads from SimOptions.
struct DevTol {
    double gmin, reltol, abstol;
    // per-instance device gmin (SimOptions::gmin as true gmin stepping
    // publishes it, convergence.cpp:383-386), or null: gmin for all
    const double* gmin_inst = nullptr;
};

// The device gmin an instance's loads use.
__host__ __device__ inline double dev_gmin(const DevTol& t, int inst) {
    return t.gmin_inst ? t.gmin_inst[inst] : t.gmin;
}

void launch_res_sites(int count, const double* ohms, double* contrib, int site, int B,
                      hipStream_t s);
void launch_const_sites(int count, const double* vals, double* contrib, int site, int B,
                        hipStream_t s);
void launch_cap(const CapTable& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                double* s2, const int32_t* done, double* cm, double* cr, int B, hipStream_t s);
void launch_vsrc(const VsrcTable& t, const DevCtx* ctx, const int32_t* done, double* cr, int B,
                 hipStream_t s);
void launch_isrc(const IsrcTable& t, const DevCtx* ctx, const int32_t* done, double* cr, int B,
                 hipStream_t s);
// DIOload / MOS1load; a device that limits its junction voltages or fails
// its own convergence test sets devnc[inst] (CKTnoncon, device_converged).
void launch_diode(const DioTable& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                  const double* s2, const int32_t* done, double* cm, double* cr, int32_t* devnc,
                  DevTol tol, int B, hipStream_t s);
void launch_mos1(const Mos1Table& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                 const double* s2, const int32_t* done, double* cm, double* cr, int32_t* devnc,
                 DevTol tol, int B, hipStream_t s);
// Voltage sources a launch may evaluate in extra workgroups (vsrc_lane):
// the BSIM4v7 load carries them, so that a batch of rings needs no launch
// of its own for its supplies.
struct VsrcSide {
    VsrcTable t;
    const DevCtx* ctx;
    const int32_t* done;
    double* cr;
    int B;
};
// BSIM4v7load for the default-mode cards (bsim4v7_load.cpp:69-5104); the
// load's own current test sets devnc (device_converged).  vs: voltage
// sources evaluated in extra workgroups of the same launch.
void launch_bsim4(const B4Table& t, const DevCtx* ctx, const double* x, double* s0, double* s1,
                  const double* s2, const int32_t* done, double* cm, double* cr, int32_t* devnc,
                  DevTol tol, int B, hipStream_t s, const VsrcSide* vs = nullptr);
// dL/dp of every BSIM4v7 (device, instance) at a converged DC point for one
// direction p of the constant record: the load's stamps in MODEADJOINT
// differentiated by dual numbers seeded with dpar = d(record)/dp (the
// per-instance layout of par_inst), contracted with the adjoint solution
// lam: g = -lam^T (dA/dp x - drhs/dp) over the device's rows.  ctx must
// carry MODEADJOINT.
void launch_bsim4_grad(const B4Table& t, const DevCtx* ctx, const double* x, const double* dpar,
                       const double* lam, double* g, DevTol tol, int B, hipStream_t s);
// MOS1Device::conv_test on the proposed solution x (mos1_device.cpp:376-416).
void launch_mos1_conv(const Mos1Table& t, const double* x, const double* s0,
                      const int32_t* skip, int32_t* devnc, DevTol tol, int B, hipStream_t s);
// dst[i] |= src[i], i < count
void launch_or_i32(int32_t* dst, const int32_t* src, int count, hipStream_t s);
void launch_gather(int nrows, const int32_t* ptr, const int32_t* sites, const int32_t* ptr2,
                   const int32_t* sites2, const double* contrib, const double* contrib2,
                   double* out, int B, hipStream_t s);
// The gather of a packed batch of circuits (B = 1): out[row] = the sum of
// the row's sites in ptr/sites (plus, after it, those in ptr2/sites2) from
// 0.0 in site order.  Rows are split into short ones (a lane each) and long
// ones (a wave each); both lists hold row indices.
struct GatherPacked {
    int nshort, nlong;
    const int32_t* short_rows;
    const int32_t* long_rows;
    const int32_t* ptr;
    const int32_t* sites;
    const double* contrib;
    const int32_t* ptr2;
    const int32_t* sites2;
    const double* contrib2;
    double* out;
};
// One launch for two gathers (the matrix and the rhs).
void launch_gather_packed(const GatherPacked& m, const GatherPacked& r, hipStream_t s);
void launch_conv(int n, int num_nodes, const double* xnew, const double* xold, double reltol,
                 double vntol, double abstol, const int32_t* done, int32_t* nonconv, int B,
                 hipStream_t s);
void launch_commit(int n, const double* src, double* dst, const int32_t* mask, int mask_value,
                   int B, hipStream_t s);
void launch_fill_i32(int32_t* p, int32_t v, int count, hipStream_t s);

// Local truncation error (ckt_terr, src/devices/ckt_terr.hpp:19-64) of the
// devices that propose a step: capacitors (one charge at their state base,
// capacitor.cpp:242-250), diodes (the depletion charge at base + 3,
// dio_device.cpp:278-291) and MOS1 (qgs, qgd, qgb, qbd, qbs at base + 5, 8,
// 11, 13, 15; mos1_device.cpp:352-365).
struct TruncTable {
    int ndev;
    const int32_t* base;   // state base per device
    const int32_t* kind;   // 0 capacitor, 1 MOS1, 2 diode, 3 BSIM4v7
    const int32_t* inst_of;
};
struct TruncOpts {
    double abstol, reltol, chgtol, trtol;
};
// dt_dev (ndev x B, optional) gets each device's compute_trunc; dt_min (B,
// the bits of non-negative doubles, filled with those of 1e30 first) their
// minimum per instance.
// order_override > 0 evaluates the devices at that integration order instead
// of the context's; dt_min2 (optional) receives the per-instance minimum at
// order 2 from the same states (the order-promotion test,
// transient.cpp:946-958).  sel (optional, per instance): only instances with
// sel[i] != 0 are evaluated; the others' dt_min entries are left alone.
void launch_trunc(const TruncTable& t, const DevCtx* ctx, const double* s0, const double* s1,
                  const double* s2, const double* s3, int B, const TruncOpts& o, double* dt_dev,
                  unsigned long long* dt_min, hipStream_t s, int order_override = 0,
                  unsigned long long* dt_min2 = nullptr, const int32_t* sel = nullptr);
void launch_fill_u64(unsigned long long* p, unsigned long long v, int count, hipStream_t s);
// out[k] = A[k * B + inst], k < count (one instance of a slot-major batch)
void launch_take_instance(const double* A, int B, int inst, int count, double* out, hipStream_t s);

}  // namespace nsa
