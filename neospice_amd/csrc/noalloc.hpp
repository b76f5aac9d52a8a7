// The library's static device scratch for entry points that may not
// allocate (noalloc.hip).
#pragma once
#include <cstddef>

namespace nsa {

// Size of the array on each GPU.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;

// The array's device address on the current device (looked up on the
// first call on each device).  Every call on a GPU shares that GPU's array.
double* noalloc_scratch();

}  // namespace nsa
