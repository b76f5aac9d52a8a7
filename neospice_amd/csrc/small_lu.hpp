// Packed LU for batches of small circuits: one wave per instance, the
// instance's factor held in LDS for the whole refactor and solve.
//
// Every instance may have its own matrix (its own circuit topology and its
// own pivot order): instances point at a plan, and instances whose Sparse 1.3
// structures are identical share one.  Values are packed instance after
// instance (A at aoff, factors at luoff, vectors at xoff) instead of being
// interleaved, so batches of different sizes sit side by side.
//
// The refactor restates SparseMatrix::factor (src/solver/factor.cpp:980-1052)
// right-looking: step k takes d_k = 1/x_kk and applies, to every (i, j) with
// L(i,k) and U(k,j) in the pattern, x_ij -= (x_kj * d_k) * L(i,k).  Entry
// (i, j) receives exactly the terms of its left-looking column loop, in the
// same ascending-k order, with the same two roundings per term, so the
// factor is bit-identical; steps whose entries do not overlap share a level.
// Sparse 1.3 stores U(k,j) = x_kj * d_k and the reciprocal pivot, which a
// closing pass writes.  A factor-and-solve launch folds the forward
// substitution (solve.cpp:21-33) into those levels as one more column; the
// backward substitution and the solve-only launches run the level-ordered
// row gathers of build_solve_programs (solve.cpp:6-47, 102-141).
//
// Only FORM_S13 structures are planned here: make_solver picks AMD-LU only
// for linear circuits of at least 256 unknowns (make_solver.cpp:156-176),
// which run on the instance-interleaved kernels of gpu_lu.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "gpu_dev.hpp"
#include "lu_structure.hpp"

namespace nsa {

// Per-instance work descriptor of a packed launch.
struct SluInst {
    int32_t plan;   // plan id
    int32_t aoff;   // first matrix value of the instance (slot order)
    int32_t luoff;  // first factor value
    int32_t xoff;   // first unknown
};

// Header of one plan inside the flattened plan data (int32 words; every
// offset is relative to the plan's first word, and every section starts on
// a 4-word boundary so that 16-byte records load as one vector).
enum SluHdr : int32_t {
    SH_N = 0, SH_NLU, SH_NNZA, SH_NLEV, SH_NTASK,
    SH_APOS,    // nnz_a: position receiving A slot p
    SH_GMIN,    // n: position receiving diag_gmin, or -1
    SH_LEV,     // nlev+1 task offsets (relative to SH_TASK, in tasks) and three
                // more copies of the last
    SH_TASK,    // tasks, 4 words each, LDS byte addresses (value p at 8p, vector entry
                // i at 8 (nlu + i)): target, U(k,j), L(i,k), pivot (k,k); forward-
                // substitution tasks: y_i, y_k, L(i,k), -(pivot address + 1)
    SH_CPTR,    // n+1 column segments
    SH_NU,      // n: U entries per column
    SH_ROWS,    // nlu: pivot-space row of each position
    SH_PCOL,    // nlu: pivot-space column of each position
    SH_RPERM, SH_CPERM,
    SH_PROG,    // 4 programs (fwd, bwd, tfwd, tbwd), SH_PROG + 4*p: nlev, levptr off,
                // rows off, phases off
    SH_HOT = SH_PROG + 16,  // first word and length of the block a factor-and-solve
    SH_HOTLEN,              // launch reads in its level loops (LEV, TASK, bwd)
    SH_WORDS,
};
// Program rows: 8 words each, LDS byte addresses (value p at 8p, vector
// entry i at 8 (nlu + i), a zero at 8 (nlu + n)): target, number of terms
// after the first two, diagonal value (the zero when none), offset of those
// terms, then value and source of terms 0 and 1 (both the zero when
// absent); terms: 2 words each (value, source).  The program op is fixed by the form:
// fwd SKIPZERO|MUL_NZ, bwd PLAIN, tfwd SKIPZERO, tbwd MUL_DIAG.
// Phases: 4 words of header (count, offset of the serial source rows, 0,
// 0), then 4 words each (kind, first level, end level, 0).  A parallel
// phase runs its levels one row per lane; a serial phase is a run of
// one-row levels (a dependency chain) that the wave runs 64 rows at a time
// from registers: every lane loads one row's operands, then the rows run
// in order, each taking the results of earlier rows of its batch from
// registers.  Serial source rows: one word per row, 1 + the row of the same
// serial phase producing the source of term 0 (low 16 bits) and term 1
// (high 16 bits), 0 when the source is final before the phase.
enum SluPhase : int32_t { PH_PARALLEL = 0, PH_SERIAL = 1 };
constexpr int kSluMinSerialRun = 4;

// Kinds per instance (the same codes the interleaved kernels use).
enum SluKind : int32_t { SLU_REFACTOR = 0, SLU_VERIFY = 1, SLU_TRUSTED = 2, SLU_SKIP = 3 };
// Launch modes.
enum SluMode : int32_t { SLU_FACTOR = 1, SLU_SOLVE = 2, SLU_SOLVE_T = 4 };

constexpr int kSluMaxLu = 6144;   // factor values per instance held in LDS
constexpr int kSluMaxN = 1024;

// Schedule choices of the packed LU, the same results under every choice
// (per entry the same operations in the same order): the shortest run of
// one-row levels that a solve program runs as a serial phase (a wave
// recurrence) instead of level by level, and whether a launch stages the
// plan's hot block in LDS when it fits.
struct SluOpts {
    int min_serial_run = kSluMinSerialRun;
    bool stage = true;
};

// The plans of a packed batch, built on the host and mirrored on the device.
class SmallLuPlans {
public:
    const SluOpts& opts() const { return opts_; }
    // Rebuilds every plan under new schedule choices (plan ids unchanged);
    // the next sync() uploads them afresh.
    void rebuild(const SluOpts& o);
    ~SmallLuPlans();
    // Whether a structure fits the packed kernel.
    static bool fits(const LuStructure& s) {
        return s.form == FORM_S13 && s.nlu() <= kSluMaxLu && s.n <= kSluMaxN;
    }
    // Plan id of `s` (an existing identical plan is reused).
    int add(const LuStructure& s);
    int add_new(const LuStructure& s);
    const LuStructure& structure(int plan) const { return st_[plan]; }
    int count() const { return (int)st_.size(); }
    int max_nlu() const { return max_nlu_; }
    int max_n() const { return max_n_; }
    int max_hot() const { return max_hot_; }
    // Refactor levels and the solve programs' levels of a plan (statistics).
    int levels(int plan) const { return lev_[plan]; }
    int solve_levels(int plan) const { return slev_[plan]; }
    // Backward-substitution levels: a factor-and-solve launch runs
    // levels() (the forward substitution folded in) and then these.
    int bwd_levels(int plan) const { return blev_[plan]; }
    // Copies new plans to the device (stream-ordered); returns the device
    // words and offsets valid until the next sync() that adds plans.
    void sync(hipStream_t s);
    const int32_t* d_words() const { return d_words_; }
    const int32_t* d_offsets() const { return d_off_; }
    // host copies of the flattened plans (what sync() uploads)
    const std::vector<int32_t>& host_words() const { return words_; }
    const std::vector<int32_t>& host_offsets() const { return off_; }

private:
    SluOpts opts_;
    std::vector<LuStructure> st_;
    std::vector<int32_t> words_, off_;
    std::vector<int> lev_, slev_, blev_;
    int max_nlu_ = 0, max_n_ = 0, max_hot_ = 0;
    int32_t* d_words_ = nullptr;
    int32_t* d_off_ = nullptr;
    size_t d_words_cap_ = 0, d_off_cap_ = 0, synced_words_ = 0, synced_plans_ = 0;
};

// Independent work a packed launch may carry in extra workgroups: the
// truncation proposals of a device table (as launch_trunc with dt_min =
// tmin, dt_min2 = tmin + ni and the mask sel), which then run beside the
// instances' factor-and-solve chains, whose one wave each leaves most
// SIMDs idle.
struct SluSide {
    TruncTable t;
    const DevCtx* ctx;
    const double *s0, *s1, *s2, *s3;
    int B;
    TruncOpts o;
    unsigned long long* tmin;
    int ni;
    const int32_t* sel;
};

// One packed launch: factor (per-instance kinds, fail codes and first
// failing column), solve (rhs -> x, may alias) or transposed solve.  A, LU,
// rhs, x are packed device arrays addressed through `inst`.  diag_gmin is
// added to every pivot's diagonal, or gmin_inst[i] per instance when given.
// side (optional): extra work carried by the same launch.  Asynchronous.
void launch_small_lu(const SmallLuPlans& plans, const SluInst* d_inst, int ninst, int mode,
                     const double* A, double diag_gmin, const int32_t* kind, int32_t* fail,
                     int32_t* fail_col, double* LU, const double* rhs, double* x, hipStream_t s,
                     const double* gmin_inst = nullptr, const SluSide* side = nullptr);

// The complex packed LU (gpu_small_lu_c.hip): factor_complex and
// solve_complex on the same plans, one wave per SluInst, every array of
// (re, im) pairs with offsets counted in complex values (ax at aoff, LU at
// luoff -- LU may be null when only the solution is wanted -- rhs and x at
// xoff; x may alias rhs).  Modes SLU_FACTOR and SLU_SOLVE (both: the
// forward substitution folded into the factor).  fail[i] = FAIL_ZERO_PIVOT
// with the first failing column in fail_col.
constexpr size_t kSluComplexLds = 64 * 1024;  // beyond it a launch raises the kernel's LDS limit
size_t small_lu_complex_lds(const SmallLuPlans& plans);
void launch_small_lu_complex(const SmallLuPlans& plans, const SluInst* d_inst, int ninst, int mode,
                             const double* ax, int32_t* fail, int32_t* fail_col, double* LU, const double* rhs,
                             double* x, hipStream_t s);
// The AC sweep's matrices and right-hand sides for B instances x nf
// frequencies, (instance, frequency) blocks, instance-major: ax (complex,
// nnz per block) = (g_i, w_f c_i) per slot; v (complex, n per block) =
// rhs_i (complex, n per instance).
void launch_ac_build(const double* g, const double* c, const double* w, int B, int nf, int nnz, double* ax,
                     hipStream_t s);
void launch_ac_rhs(const double* rhs, int B, int nf, int n, double* v, hipStream_t s);

}  // namespace nsa
