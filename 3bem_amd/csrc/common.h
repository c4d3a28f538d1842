// Shared host-side plumbing: errors, HIP checks, device buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tbem_fmm.h"

namespace tbem {

// Error carried from deep inside the library to the C boundary, where it becomes
// an int status plus tbem_last_error().
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define TBEM_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::tbem::Error(e_ == hipErrorOutOfMemory ? TBEM_E_NOMEM : TBEM_E_HIP,       \
                                std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                                    __FILE__ + ":" + std::to_string(__LINE__));              \
    } while (0)

#define TBEM_REQUIRE(cond, code, msg)                       \
    do {                                                    \
        if (!(cond)) throw ::tbem::Error((code), (msg));    \
    } while (0)

#define TBEM_LAUNCH_CHECK() TBEM_HIP(hipGetLastError())

// Owning device allocation (hipMalloc'd, freed on destruction).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) TBEM_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s = 0) {
        if (count > n) alloc(count);
        if (count) TBEM_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s = 0) { upload(v.data(), v.size(), s); }
    void download(T* h, size_t count, hipStream_t s = 0) const {
        if (count) TBEM_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    std::vector<T> to_host(hipStream_t s = 0) const {
        std::vector<T> v(n);
        download(v.data(), n, s);
        TBEM_HIP(hipStreamSynchronize(s));
        return v;
    }
    void zero(hipStream_t s = 0) {
        if (n) TBEM_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
};

inline int div_up(size_t a, size_t b) { return (int)((a + b - 1) / b); }

// Marks a value that every lane of the wave holds alike as wave-uniform
// (readfirstlane puts it in an SGPR). Any loop or branch that contains a
// __syncthreads() and depends on a value loaded from memory (global or LDS)
// must go through this: the compiler cannot prove such a loaded value uniform,
// and it lowers a barrier inside what it takes for divergent control flow with
// per-lane exec masking, which has dropped the later LDS stores of the lanes
// that were active in a loop's last partial iteration.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

}  // namespace tbem
