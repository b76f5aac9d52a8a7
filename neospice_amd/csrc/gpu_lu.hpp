// Batched LU refactorization and triangular solves on the GPU for a fixed
// LuStructure shared by a group of instances.
//
// Factor values are stored instance-major (LU[inst * nlu + pos]): the
// refactor and the solves are latency-bound dependency chains inside one
// instance, so each instance is owned by one workgroup that walks its own
// columns; A and b stay slot-major with the instance index fastest (the
// device loads produce them that way).
//
// Refactor = one launch, one workgroup (16 waves) per instance:
//   sparse levels  columns before the trailing block, level by level
//                  (workgroup barrier between levels), one wave per column,
//                  left-looking exactly as the reference, updates in the
//                  order of the column's U list;
//   partial pass   trailing-block columns: A load plus the updates from
//                  pivots before the block;
//   dense tail     the trailing block (<= 128 columns) factored right-looking
//                  by the whole workgroup, held in registers;
//   solve gather   the factor values the plain solve reads, copied into the
//                  solve's own term order so that it streams them.
// Every value is bit-identical to the reference: each entry receives its
// updates in the order of its column's U list.  The dense phase applies
// trailing pivots in ascending order, so the block is sized such that every
// trailing column lists its trailing pivots ascending (always true for
// Sparse 1.3); earlier pivots that an AMD-LU reach order lists after a
// trailing pivot are applied to the block right after that pivot's step.
//
// Solve = one launch, one workgroup per instance, the vector in LDS.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lu_structure.hpp"

namespace nsa {

// A triangular program on the device, rows in level order.
struct DevProgram {
    int32_t op = 0;
    int32_t nlevels = 0;
    int32_t sv_base = 0;             // first term value in the solve-ordered copy
    int32_t* level_ptr = nullptr;    // nlevels+1 into row positions
    int32_t* level_mode = nullptr;   // 0: one thread per row, 1: one wave per row
    int4* rows = nullptr;            // (target, first term, term count, diagonal value index)
    int16_t* src = nullptr;          // per term: source index
    int32_t* pos = nullptr;          // per term: LU position (programs that read LU directly)
    // column sweep (TriProgram::sweep_*): one wave, before (descending) or
    // after (ascending) the levels; values in a sweep_n x sweep_n block of the
    // solve-ordered copy at sweep_sv (row = step, column = target; the step's
    // own diagonal on the block diagonal), presence bits per step and lane
    // (bit m: target lane + 64 m; a full word each, so that the kernel's
    // prefetch of them needs no conversion that would wait for the load)
    int32_t sweep_n = 0;
    int32_t sweep_first = 0;
    int32_t sweep_sv = 0;
    int32_t* sweep_tgt = nullptr;
    int32_t* sweep_bits = nullptr;
};

class GpuLu {
public:
    // One plan over `batch` instances.  With `instances`, the plan serves the
    // listed instances of a larger batch (batch_global): A, b and the
    // failure flags are addressed by batch instance, the factors are
    // stored for the plan's own instances.
    GpuLu(const LuStructure& s, int batch, hipStream_t stream,
          const std::vector<int32_t>* instances = nullptr, int batch_global = 0);
    ~GpuLu();
    GpuLu(const GpuLu&) = delete;
    GpuLu& operator=(const GpuLu&) = delete;

    // Refactor from A values (device, nnz_a x Bglobal, slot-major).  kind
    // (device, per batch instance): 0 refactorize (replay checks), 1
    // numeric whose pivot order must be re-checked, 2 numeric on an order
    // just computed from these values, 3 skip.  Failures go to fail[inst]
    // (FailCode) and fail_col[inst].  Asynchronous on the stream.
    // `on` overrides the plan's stream.
    void refactor(const double* dA, double diag_gmin, const int32_t* dkind, int32_t* dfail,
                  int32_t* dfail_col, hipStream_t on = nullptr);

    // In-place solve of b (device, n x Bglobal) for this plan's instances;
    // transposed = A^T x = b.
    void solve(double* db, bool transposed, hipStream_t on = nullptr);

    const LuStructure& structure() const { return s_; }
    int batch() const { return B_; }
    double* lu_values() { return dLU_; }
    int32_t nlevels_refactor() const { return nlev_ + (tail_ ? 1 : 0); }
    int32_t nlevels_solve() const { return fwd_.nlevels + bwd_.nlevels; }
    int32_t tail_size() const { return tail_; }
    int64_t deferred_terms() const { return n_deferred_terms_; }
    hipStream_t stream() const { return stream_; }

private:
    // Uploads a program; with sv_pos, its term and diagonal values are read
    // from the solve-ordered copy and their LU positions appended to sv_pos.
    void upload_program(const TriProgram& t, DevProgram& d, std::vector<int32_t>* sv_pos);

    LuStructure s_;
    int B_;
    int Bglobal_ = 0;
    int32_t* d_imap_ = nullptr;
    hipStream_t stream_;
    std::vector<void*> allocs_;
    int tail_ = 0, T0_ = 0, nlev_ = 0, wsz_ = 0, wregion_ = 0, nws_ = 0;
    size_t rf_lds_ = 0, union_bytes_ = 0, sv_lds_ = 0, sv_lds_t_ = 0;
    int64_t n_deferred_terms_ = 0;
    // sparse phase
    int32_t* d_lev_ptr_ = nullptr;
    int4* d_cmeta_ = nullptr;
    int32_t *d_cptr_ = nullptr, *d_nU_ = nullptr, *d_rows_ = nullptr, *d_asrc_ = nullptr,
            *d_gpos_ = nullptr, *d_desc_ptr_ = nullptr;
    int4* d_desc_ = nullptr;
    int16_t* d_rel_ = nullptr;
    int nrel_ = 0;
    int32_t *d_cand_ptr_ = nullptr, *d_cand_pos_ = nullptr, *d_diag_cand_ = nullptr,
            *d_row_perm_ = nullptr, *d_col_perm_ = nullptr;
    // dense tail
    int32_t* d_tposmap_ = nullptr;
    uint32_t* d_tmask_ = nullptr;
    int32_t* d_tcand_ptr_ = nullptr;
    int16_t* d_tcand_ = nullptr;
    int32_t* d_tdiag_cand_ = nullptr;
    int32_t *d_dstep_ptr_ = nullptr, *d_dtask_ptr_ = nullptr, *d_dterm_ = nullptr;
    int4* d_dtask_ = nullptr;
    double* d_dprod_ = nullptr;   // deferred products, ndterm per instance
    // solve-ordered copy of the values the plain solve reads
    int32_t* d_sv_pos_ = nullptr;
    int nsv_ = 0;
    double* dSV_ = nullptr;
    double* dLU_ = nullptr;
    DevProgram fwd_, bwd_, tfwd_, tbwd_;
};

}  // namespace nsa
