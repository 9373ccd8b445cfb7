// dbuf.hpp -- owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

namespace v2h {

// Device buffer owned by the solver.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t count) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count) V2H_HIP(hipMalloc(&p, sizeof(T) * count));
    }
    void upload(const std::vector<T>& v) {
        alloc(v.size());
        if (!v.empty()) V2H_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    }
};

}  // namespace v2h
