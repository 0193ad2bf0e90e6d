// Shared helpers for the gfx950 kernels of libpamo_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace pamo {

// Records the first failing HIP call in a thread-local string and returns a
// nonzero status from the enclosing function.
void set_error(const std::string& msg);
const char* get_error();

#define PAMO_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::pamo::set_error(std::string(#call) + ": " + hipGetErrorString(e_) + " (" +      \
                              __FILE__ + ":" + std::to_string(__LINE__) + ")");               \
            return 1;                                                                         \
        }                                                                                     \
    } while (0)

#define PAMO_CHECK(cond, msg)                                                                 \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            ::pamo::set_error(msg);                                                           \
            return 2;                                                                         \
        }                                                                                     \
    } while (0)

#define PAMO_TRY(call)                \
    do {                              \
        int rc_ = (call);             \
        if (rc_ != 0) return rc_;     \
    } while (0)

constexpr int kWave = 64;

inline unsigned grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

// Number of set bits of `mask` in lanes strictly below the calling lane.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

}  // namespace pamo
