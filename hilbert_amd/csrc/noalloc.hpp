// noalloc.hpp -- the device scratch of the allocation-free entry point
// (v2h_batch_Au_noalloc): a file-scope device array in noalloc.hip, one copy
// per GPU, held until the process ends.
#pragma once

#include <cstdint>

namespace v2h {

// Doubles of the scratch (4 MiB): the input vector of A u, then its output,
// so n_primal + n_dual must fit.
constexpr int64_t kNoAllocDoubles = int64_t(1) << 19;
constexpr int kNoAllocMaxDevices = 64;

// Device address of the scratch on `device`, which must be the current
// device.  The first call on a device resolves the address with
// hipGetSymbolAddress and keeps it; later calls return the kept address.
double* noalloc_scratch(int device);

}  // namespace v2h
