// noalloc.hip -- the device scratch of the allocation-free entry point.
//
// The array below is a host-side stand-in in host code: its name must never
// reach a kernel or a copy.  Its device address on each GPU is resolved once
// per device, with that device current, and only that address is used.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>

#include "common.hpp"
#include "noalloc.hpp"

namespace v2h {

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
std::mutex g_mu;
double* g_addr[kNoAllocMaxDevices] = {};
}  // namespace

double* noalloc_scratch(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[device]) {
        int cur = -1;
        V2H_HIP(hipGetDevice(&cur));
        if (cur != device) throw std::logic_error("noalloc_scratch: the device must be current");
        void* p = nullptr;
        V2H_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipError("hipGetSymbolAddress returned no address for the scratch");
        g_addr[device] = static_cast<double*>(p);
    }
    return g_addr[device];
}

}  // namespace v2h
