// Batched, instance-interleaved LU refactorization and triangular solves on
// the GPU for a fixed LuStructure (one structure shared by B instances).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lu_structure.hpp"

namespace nsa {

struct DevProgram {
    int32_t op = 0;
    std::vector<int32_t> level_ptr;   // host copy (launch bounds)
    int32_t* order = nullptr;
    int32_t* term_ptr = nullptr;
    int32_t* term_pos = nullptr;
    int32_t* term_src = nullptr;
    int32_t* dpos = nullptr;
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

    // Refactor from A values (device, nnz_a x B, slot-major).  For the
    // instances whose dverify[inst] != 0 (dverify may be null) it also checks a fresh numeric factorization
    // of these values would keep every pivot of this structure (the
    // "numeric()" contract); failures are written to fail[inst] (FailCode)
    // and the failing column to fail_col[inst].  Asynchronous on the stream.
    void refactor(const double* dA, double diag_gmin, const int32_t* dverify, int32_t* dfail,
                  int32_t* dfail_col);

    // In-place solve of b (device, n x B, row-major by unknown) with the
    // current factors; transposed = A^T x = b.
    void solve(double* db, bool transposed);

    const LuStructure& structure() const { return s_; }
    int batch() const { return B_; }
    double* lu_values() { return dLU_; }
    int32_t nlevels_refactor() const { return sched_.nlevels(); }
    int32_t nlevels_solve() const { return fwd_.level_ptr.size() + bwd_.level_ptr.size() - 2; }
    size_t lu_bytes() const { return (size_t)s_.nlu() * B_ * sizeof(double); }
    hipStream_t stream() const { return stream_; }

private:
    void run_program(const DevProgram& p, double* dy);
    void upload_program(const TriProgram& t, DevProgram& d);

    LuStructure s_;
    RefactorSchedule sched_;
    int B_;
    int Bglobal_ = 0;
    int32_t* d_imap_ = nullptr;
    hipStream_t stream_;
    std::vector<void*> allocs_;
    int32_t *d_cols_ = nullptr, *d_cptr_ = nullptr, *d_nU_ = nullptr, *d_rows_ = nullptr,
            *d_asrc_ = nullptr, *d_gpos_ = nullptr, *d_cand_ptr_ = nullptr,
            *d_cand_pos_ = nullptr, *d_diag_cand_ = nullptr, *d_row_perm_ = nullptr,
            *d_col_perm_ = nullptr;
    double* dLU_ = nullptr;
    double* dy_ = nullptr;
    DevProgram fwd_, bwd_, tfwd_, tbwd_;
    int max_seg_ = 0;
};

}  // namespace nsa
