// Shared helpers for the host side of neospice_amd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace nsa {

// Error raised by the host library; the C-ABI maps it to a status code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes of the C-ABI (include/neospice_amd.h).
enum Status : int {
    NSA_OK = 0,
    NSA_SINGULAR = 1,          // maps to ISolver's `true` (singular)
    NSA_ERR_LOGIC = -1,        // misuse ordering (std::logic_error in the reference)
    NSA_ERR_ARGUMENT = -2,     // bad size / argument (std::invalid_argument)
    NSA_ERR_RUNTIME = -3,      // runtime failure (std::runtime_error)
    NSA_ERR_HIP = -4,          // HIP runtime error
    NSA_ERR_UNSUPPORTED = -5,  // a reference feature this build does not cover
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw Error(NSA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define NSA_HIP(x) ::nsa::hip_check((x), #x)

// Per-instance failure codes written by the refactor kernels.
enum FailCode : int32_t {
    FAIL_NONE = 0,
    FAIL_ZERO_PIVOT = 1,   // exact zero pivot (refactorize) or singular (numeric)
    FAIL_GROWTH = 2,       // GP replay pivot-growth check failed
    FAIL_REORDER = 3,      // pivot no longer threshold-acceptable: CPU re-pivot
    FAIL_TIMEOUT = 4,      // a kernel's bounded wait for a dependency ran out (a bug)
    FAIL_NOPLAN = 5,       // no factorization yet (a fresh solver): the host orders it
};

}  // namespace nsa
