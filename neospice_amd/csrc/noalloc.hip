// A device array owned by the library for entry points that may not
// allocate: a file-scope __device__ array, resident on every GPU the library
// runs on until the process ends.  Host code never passes the array's name
// to a kernel or a copy (on the host it is only a stand-in symbol); it uses
// the device address hipGetSymbolAddress gives on each device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "common.hpp"
#include "noalloc.hpp"

namespace nsa {

namespace {

constexpr size_t kScratchDoubles = kNoAllocScratchBytes / sizeof(double);

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
std::vector<double*> g_ptr;  // per device: the array's device address, once looked up

}  // namespace

double* noalloc_scratch() {
    int dev = 0;
    NSA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_ptr.size() <= dev) g_ptr.resize(dev + 1, nullptr);
    if (!g_ptr[dev]) {
        // once per device, with that device current
        void* p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !p)
            throw Error(NSA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
        g_ptr[dev] = static_cast<double*>(p);
    }
    return g_ptr[dev];
}

}  // namespace nsa
