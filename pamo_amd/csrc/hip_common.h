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
void clear_error();
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

// Kernel ids of the launch profiler (HIP events around selected launches).
enum KernelId {
    K_RING_COUNT, K_RING_FILL, K_RING_SORT, K_EDGE_COUNT, K_EDGE_FILL, K_QUADRIC, K_EDGE_COST, K_PROPAGATE,
    K_SELECT_COLLAPSE, K_LINE_FLAG, K_MARK_CHANGED, K_FACE_BOX, K_MORTON, K_RADIX_HIST, K_RADIX_SCATTER,
    K_KARRAS, K_BVH_BOXES, K_GUARD_QUERY, K_UNDO_CAND, K_COMPACT_MESH, K_SCAN_TILES, K_SCAN_SUMS, K_OTHER,
    K_COUNT
};

// Called around a launch; no-ops unless the calling thread's active context
// profiles this kernel id.
void prof_begin(int id, double bytes, hipStream_t st);
void prof_end(int id, hipStream_t st);

// Tree walks are bounded: over a valid tree of n leaves (2n - 1 nodes) an
// escape-link walk visits each node at most once, a parent-link walk at most
// three times, and a shared frontier pops at most (queries x nodes) entries.
// A walk that exceeds its bound (a corrupted tree) stops and sets *err, which
// the host reports as an error status instead of the GPU spinning.
__device__ __forceinline__ void walk_overrun(int* err) { atomicOr(err, 1); }

inline unsigned grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

// Number of set bits of `mask` in lanes strictly below the calling lane.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

}  // namespace pamo
