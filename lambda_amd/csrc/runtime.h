// Host-side helpers shared by the library's translation units: the error
// type every entry point turns into a code and message, the HIP status check
// and a typed device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "lambda_hip.h"

namespace lam {

// Threads per block of the path kernels.
constexpr int kBlock = 256;

// ------------------------------------------------------------------ errors

struct Failure {
    int code;
    std::string msg;
};

inline void set_err(char *err, size_t len, const std::string &msg) {
    if (!err || len == 0) return;
    size_t n = std::min(msg.size(), len - 1);
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
}

#define HIP_OK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw Failure{e_ == hipErrorOutOfMemory ? LAMBDA_ENOMEM : LAMBDA_EHIP,            \
                          std::string(#call) + ": " + hipGetErrorString(e_)};                 \
    } while (0)

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return;
        HIP_OK(hipMalloc(&p, count * sizeof(T)));
    }
    // alloc unless it already holds exactly count elements
    void ensure(size_t count) {
        if (!p || n != count) alloc(count);
    }
    void upload(const T *src, size_t count) {
        alloc(count);
        if (count) HIP_OK(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    }
    void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
};

}  // namespace lam
