// Device-wide primitives written for wave64 / LDS on gfx950:
//   * exclusive prefix sum of int32 (tile sums, then each 2048-item tile
//     scanned from the sum of the tiles before it);
//   * stable LSD radix sort of (uint32 key, uint32 value) pairs, 8 bits per
//     pass; within a tile, ranks of equal digits come from wave64 ballots
//     (match of the 8 digit bits) and an LDS prefix across the 4 waves, so
//     the order of equal keys is preserved exactly.
#pragma once
#include <algorithm>

#include "hip_common.h"

namespace pamo {
// Internal linkage: every translation unit that includes this header gets
// its own copy of the kernels.
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanIpt = 8;
constexpr int kScanTile = kScanBlock * kScanIpt;

// Block-wide exclusive scan of one value per thread (256 threads, 4 waves).
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds_wave /*[4]*/, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds_wave[w] = x;
    __syncthreads();
    int wbase = 0;
    int tot = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int s = lds_wave[i];
        if (i < w) wbase += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wbase + x - v;
}

__global__ void __launch_bounds__(kScanBlock) scan_tiles_k(const int* __restrict__ in, int* __restrict__ out,
                                                           int* __restrict__ tile_sums, long long n) {
    __shared__ int lds[4];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanIpt;
    int v[kScanIpt];
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanIpt; ++i) {
        long long idx = base + i;
        v[i] = idx < n ? in[idx] : 0;
        s += v[i];
    }
    int total;
    int ex = block_exclusive_scan(s, lds, total);
#pragma unroll
    for (int i = 0; i < kScanIpt; ++i) {
        long long idx = base + i;
        if (idx < n) out[idx] = ex;
        ex += v[i];
    }
    if (threadIdx.x == 0 && tile_sums) tile_sums[blockIdx.x] = total;
}

// Tiles of kScanTile items: pass 1 writes each tile's sum; pass 2 gives
// every tile the sum of the tile sums before it (read by the whole block)
// and scans the tile from there.  Two launches for any n, no waiting between
// blocks; a single tile is scanned in one launch.
__global__ void __launch_bounds__(kScanBlock) tile_sum_k(const int* __restrict__ in, int* __restrict__ sums,
                                                         long long n) {
    __shared__ int lds[4];
    const long long base = (long long)blockIdx.x * kScanTile + threadIdx.x;
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanIpt; ++i) {
        const long long idx = base + (long long)i * kScanBlock;
        s += idx < n ? in[idx] : 0;
    }
    int total;
    (void)block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanBlock) scan_tiles_prefix_k(const int* in, int* out,
                                                                  const int* __restrict__ sums, long long n) {
    __shared__ int lds[4];
    int p = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += kScanBlock) p += sums[t];
    int before;
    (void)block_exclusive_scan(p, lds, before);
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanIpt;
    int v[kScanIpt];
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanIpt; ++i) {
        const long long idx = base + i;
        v[i] = idx < n ? in[idx] : 0;
        s += v[i];
    }
    int total;
    int ex = before + block_exclusive_scan(s, lds, total);
#pragma unroll
    for (int i = 0; i < kScanIpt; ++i) {
        const long long idx = base + i;
        if (idx < n) out[idx] = ex;
        ex += v[i];
    }
}

// One workgroup of 1024 threads scans up to kBlockScanMax items in one
// launch: each thread sums a contiguous run of `per` items (four at a time
// with 16-byte loads when both arrays are 16-byte aligned), the 16 waves scan
// the run sums, and each thread writes its run from its offset.  Replaces the
// two launches of the tiled scan for mid-size arrays, whose cost is mostly
// the second launch and the wait between the two.  Measured on the 2M-face
// workload: up to 16k items it is slightly faster than two launches; at 64k
// the serial runs make it slower.
constexpr int kBlockScan = 1024;
constexpr long long kBlockScanMax = 1 << 14;

__global__ void __launch_bounds__(kBlockScan) scan_block_k(const int* in, int* out, int n, int per, int vec) {
    __shared__ int lds[kBlockScan / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b = threadIdx.x * per;
    const int e = b + per < n ? b + per : n;
    int s = 0;
    if (vec) {
        for (int i = b; i < e; i += 4) {
            if (i + 4 <= e) {
                const int4 q = *reinterpret_cast<const int4*>(in + i);
                s += q.x + q.y + q.z + q.w;
            } else {
                for (int k = i; k < e; ++k) s += in[k];
            }
        }
    } else {
        for (int i = b; i < e; ++i) s += in[i];
    }
    int x = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    int ex = x - s;
    for (int i = 0; i < w; ++i) ex += lds[i];
    if (vec) {
        for (int i = b; i < e; i += 4) {
            if (i + 4 <= e) {
                const int4 q = *reinterpret_cast<const int4*>(in + i);
                int4 o;
                o.x = ex;
                o.y = o.x + q.x;
                o.z = o.y + q.y;
                o.w = o.z + q.z;
                ex = o.w + q.w;
                *reinterpret_cast<int4*>(out + i) = o;
            } else {
                for (int k = i; k < e; ++k) {
                    const int v = in[k];
                    out[k] = ex;
                    ex += v;
                }
            }
        }
    } else {
        for (int i = b; i < e; ++i) {
            const int v = in[i];
            out[i] = ex;
            ex += v;
        }
    }
}

// Single-launch scans by decoupled look-back.  Each workgroup takes a tile
// number from a ticket counter (so a tile only waits on tiles whose
// workgroups have already started), scans its items, publishes its total,
// and wave 0 then reads the states of up to 64 preceding tiles at a time:
// it adds their totals back to the nearest tile that has published its
// inclusive prefix, and publishes its own inclusive prefix.  A state word
// packs the call's epoch (bits 34-63), a status (bit 32: total, bit 33:
// inclusive prefix) and the 32-bit value, so words left by earlier calls
// never match and the states need no clearing between calls.  Every wait is
// bounded: after kLbSpinCap reads the tile reports bit 2 in *err (the host
// turns it into an error status) and goes on.
//
// The item values and what is done with their prefixes come from an
// operation `Op`: Op::count(i) is item i's value (i < n), Op::emit(i, v,
// pos) receives its exclusive prefix, Op::done(total) runs once, in the last
// tile, with the sum of all values, and Op::kIpt is the number of
// consecutive items per thread.  So a flag computation, the scan and the
// compaction it feeds are one launch.
constexpr unsigned long long kLbTotal = 1ull << 32, kLbInclusive = 2ull << 32;
constexpr int kLbSpinCap = 1 << 20;

struct LookbackScan {
    unsigned long long* state = nullptr;  // one word per tile
    unsigned* ticket = nullptr;           // tiles handed out so far (mod 2^32)
    int* err = nullptr;
    long long cap = 0;                    // tiles
    unsigned epoch = 0, base = 0;         // this call's epoch; ticket value at its start
};

__device__ __forceinline__ unsigned long long lb_word(unsigned epoch, unsigned long long status, int v) {
    return ((unsigned long long)epoch << 34) | status | (unsigned long long)(unsigned)v;
}

// Exclusive prefix of this tile's total among the tiles before it (wave 0
// of the workgroup; lane 0's return value is the one used).
__device__ __forceinline__ int lb_prefix(int tile, int total, unsigned long long* __restrict__ state,
                                         int* __restrict__ err, unsigned epoch) {
    const int lane = threadIdx.x;
    int excl = 0;
    if (tile == 0) {
        if (lane == 0)
            __hip_atomic_store(state, lb_word(epoch, kLbInclusive, total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    if (lane == 0)
        __hip_atomic_store(state + tile, lb_word(epoch, kLbTotal, total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    for (int end = tile - 1;; end -= kWave) {
        const int j = end - lane;
        unsigned long long w = lb_word(epoch, kLbInclusive, 0);  // before tile 0: prefix 0
        if (j >= 0) {
            for (int spins = 0;; ++spins) {
                w = __hip_atomic_load(state + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((unsigned)(w >> 34) == epoch && (w & (kLbTotal | kLbInclusive)) != 0) break;
                if (spins >= kLbSpinCap) {
                    atomicOr(err, 2);
                    w = lb_word(epoch, kLbInclusive, 0);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        const unsigned long long inc = __ballot((w & kLbInclusive) != 0);
        const int stop = inc ? __ffsll((long long)inc) - 1 : kWave - 1;
        int part = lane <= stop ? (int)(unsigned)w : 0;
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) part += __shfl_xor(part, d, kWave);
        excl += part;
        if (inc) break;
    }
    if (lane == 0)
        __hip_atomic_store(state + tile, lb_word(epoch, kLbInclusive, excl + total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}

template <class Op>
__global__ void __launch_bounds__(kScanBlock) scan_op_k(Op op, long long n, unsigned long long* __restrict__ state,
                                                        unsigned* __restrict__ ticket, int* __restrict__ err,
                                                        unsigned epoch, unsigned base) {
    constexpr int kIpt = Op::kIpt;
    __shared__ int lds[4];
    __shared__ int sh[2];
    if (threadIdx.x == 0) sh[0] = (int)(atomicAdd(ticket, 1u) - base);
    __syncthreads();
    const int tile = sh[0];
    if ((unsigned)tile >= gridDim.x) {  // ticket out of step with the host's count
        if (threadIdx.x == 0) atomicOr(err, 2);
        return;
    }
    const long long first = ((long long)tile * kScanBlock + threadIdx.x) * kIpt;
    int v[kIpt];
    int s = 0;
#pragma unroll
    for (int i = 0; i < kIpt; ++i) {
        const long long idx = first + i;
        v[i] = idx < n ? op.count(idx) : 0;
        s += v[i];
    }
    int total;
    int ex = block_exclusive_scan(s, lds, total);
    if (threadIdx.x < kWave) {
        const int excl = lb_prefix(tile, total, state, err, epoch);
        if (threadIdx.x == 0) {
            sh[1] = excl;
            if (tile == (int)gridDim.x - 1) op.done(excl + total);
        }
    }
    __syncthreads();
    ex += sh[1];
#pragma unroll
    for (int i = 0; i < kIpt; ++i) {
        const long long idx = first + i;
        if (idx < n) op.emit(idx, v[i], ex);
        ex += v[i];
    }
}

// out[i] = sum_{j<i} in[j] (in may alias out: a tile reads all its items
// before writing any).
struct ScanOp {
    static constexpr int kIpt = kScanIpt;
    const int* in;
    int* out;
    __device__ int count(long long i) const { return in[i]; }
    __device__ void emit(long long i, int, int pos) const { out[i] = pos; }
    __device__ void done(int) const {}
};

// Workspace (ints) needed by scan_exclusive for n elements.
inline long long scan_workspace(long long n) { return (n + kScanTile - 1) / kScanTile + 16; }

// out[i] = sum_{j<i} in[j]; in may alias out (each tile is read and written
// by one block only).
inline int scan_exclusive(const int* in, int* out, long long n, int* ws, hipStream_t st) {
    if (n <= 0) return 0;
    const long long nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        prof_begin(K_SCAN_TILES, 8.0 * n, st);
        hipLaunchKernelGGL(scan_tiles_k, dim3(1), dim3(kScanBlock), 0, st, in, out, (int*)nullptr, n);
        prof_end(K_SCAN_TILES, st);
        PAMO_HIP(hipGetLastError());
        return 0;
    }
    if (n <= kBlockScanMax) {
        // runs of a multiple of 4 items, so that aligned arrays stay aligned per thread
        const int per = (int)(((n + kBlockScan - 1) / kBlockScan + 3) & ~3LL);
        const int vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0 ? 1 : 0;
        prof_begin(K_SCAN_TILES, 8.0 * n, st);
        hipLaunchKernelGGL(scan_block_k, dim3(1), dim3(kBlockScan), 0, st, in, out, (int)n, per, vec);
        prof_end(K_SCAN_TILES, st);
        PAMO_HIP(hipGetLastError());
        return 0;
    }
    int* sums = ws;
    prof_begin(K_SCAN_SUMS, 4.0 * n, st);
    hipLaunchKernelGGL(tile_sum_k, dim3((unsigned)nb), dim3(kScanBlock), 0, st, in, sums, n);
    prof_end(K_SCAN_SUMS, st);
    PAMO_HIP(hipGetLastError());
    prof_begin(K_SCAN_TILES, 8.0 * n, st);
    hipLaunchKernelGGL(scan_tiles_prefix_k, dim3((unsigned)nb), dim3(kScanBlock), 0, st, in, out,
                       (const int*)sums, n);
    prof_end(K_SCAN_TILES, st);
    PAMO_HIP(hipGetLastError());
    return 0;
}

// Room for scan_op launches of up to n items at any items-per-thread count
// (tile states zeroed when they grow; the ticket counter and its host copy
// stay in step).
inline int lookback_reserve(LookbackScan& lb, long long n, int* err, hipStream_t st) {
    lb.err = err;
    if (!lb.ticket) {
        PAMO_HIP(hipMalloc((void**)&lb.ticket, sizeof(unsigned)));
        PAMO_HIP(hipMemsetAsync(lb.ticket, 0, sizeof(unsigned), st));
        lb.base = 0;
    }
    const long long tiles = (n + kScanBlock - 1) / kScanBlock + 16;  // room for one item per thread
    if (tiles > lb.cap) {
        if (lb.state) PAMO_HIP(hipFree(lb.state));
        lb.state = nullptr;
        const long long cap = tiles + tiles / 4;
        PAMO_HIP(hipMalloc((void**)&lb.state, sizeof(unsigned long long) * cap));
        PAMO_HIP(hipMemsetAsync(lb.state, 0, sizeof(unsigned long long) * cap, st));
        lb.cap = cap;
        lb.epoch = 0;
    }
    return 0;
}

inline void lookback_free(LookbackScan& lb) {
    if (lb.state) (void)hipFree(lb.state);
    if (lb.ticket) (void)hipFree(lb.ticket);
    lb = LookbackScan{};
}

// One launch of scan_op_k over n >= 0 items (a single tile needs no waiting).
template <class Op>
inline int scan_op(const Op& op, long long n, LookbackScan& lb, hipStream_t st, int prof_id = K_SCAN_TILES,
                   double bytes = 0.0) {
    if (n < 0) return 0;
    constexpr long long kTile = (long long)kScanBlock * Op::kIpt;
    const long long nb = std::max(1LL, (n + kTile - 1) / kTile);  // n = 0 still runs Op::done(0)
    PAMO_CHECK(lb.state && nb <= lb.cap, "look-back scan: tile states not reserved for this size");
    if (lb.epoch + 1 >= (1u << 30)) {  // epochs exhausted: clear the states and start again
        PAMO_HIP(hipMemsetAsync(lb.state, 0, sizeof(unsigned long long) * lb.cap, st));
        lb.epoch = 0;
    }
    ++lb.epoch;
    prof_begin(prof_id, bytes, st);
    hipLaunchKernelGGL(scan_op_k<Op>, dim3((unsigned)nb), dim3(kScanBlock), 0, st, op, n, lb.state, lb.ticket,
                       lb.err, lb.epoch, lb.base);
    prof_end(prof_id, st);
    PAMO_HIP(hipGetLastError());
    lb.base += (unsigned)nb;
    return 0;
}

// scan_exclusive in one launch for any size.  in may alias out.
inline int scan_exclusive_lb(const int* in, int* out, long long n, LookbackScan& lb, hipStream_t st) {
    return scan_op(ScanOp{in, out}, n, lb, st, K_SCAN_TILES, 8.0 * n);
}

// ---------------------------------------------------------------- radix sort
constexpr int kSortBlock = 256;
constexpr int kSortIpt = 8;
constexpr int kSortTile = kSortBlock * kSortIpt;

__global__ void __launch_bounds__(kSortBlock) radix_hist_k(const uint32_t* __restrict__ keys, int n, int shift,
                                                           int* __restrict__ counts, int nb) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kSortTile;
#pragma unroll
    for (int c = 0; c < kSortIpt; ++c) {
        long long idx = base + c * kSortBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1);
    }
    __syncthreads();
    counts[(long long)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// raw = 0: offs holds the exclusive scan of the per-block digit counts
// (digit-major); raw = 1: offs holds the counts themselves and each block
// forms its own offsets -- all blocks' totals of smaller digits plus the
// counts of earlier blocks for its digit (used for small grids, where it
// replaces two scan launches).
__global__ void __launch_bounds__(kSortBlock) radix_scatter_k(const uint32_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout,
                                                              uint32_t* __restrict__ vout, int n, int shift,
                                                              const int* __restrict__ offs, int nb, int raw) {
    __shared__ int base[256];
    __shared__ int wcnt[4][256];
    __shared__ int lds4[4];
    const int t = threadIdx.x, w = t >> 6;
    if (raw) {
        int before = 0, total = 0;
        for (int b = 0; b < nb; ++b) {
            const int c = offs[(long long)t * nb + b];
            if (b < (int)blockIdx.x) before += c;
            total += c;
        }
        int all;
        const int smaller = block_exclusive_scan(total, lds4, all);
        base[t] = smaller + before;
    } else {
        base[t] = offs[(long long)t * nb + blockIdx.x];
    }
    const long long tile = (long long)blockIdx.x * kSortTile;
    for (int c = 0; c < kSortIpt; ++c) {
        const long long idx = tile + c * kSortBlock + t;
        const bool valid = idx < n;
        const uint32_t key = valid ? kin[idx] : 0u;
        const uint32_t d = (key >> shift) & 255u;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) wcnt[ww][t] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = lanes_below(peers);
        if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        int run = 0;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) {
            int x = wcnt[ww][t];
            wcnt[ww][t] = run;
            run += x;
        }
        __syncthreads();
        if (valid) {
            const int pos = base[d] + wcnt[w][d] + rank;
            kout[pos] = key;
            vout[pos] = vin[idx];
        }
        __syncthreads();
        base[t] += run;
        __syncthreads();
    }
}

// grids of at most this many tiles form their digit offsets inside the
// scatter kernel (each block reads 256 x nb counts)
constexpr int kRawOffsetBlocks = 64;

inline long long radix_workspace(long long n) {
    long long nb = (n + kSortTile - 1) / kSortTile;
    return 256 * nb + scan_workspace(256 * nb) + 16;
}

// Stable sort of (keys, vals) on the low `bits` bits (multiple of 8).
// Results end in (keys, vals); k2/v2 are n-element scratch.
inline int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* k2, uint32_t* v2, int n, int bits, int* ws,
                            hipStream_t st) {
    if (n <= 1) return 0;
    const int nb = (n + kSortTile - 1) / kSortTile;
    int* counts = ws;
    int* sws = ws + 256LL * nb;
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    for (int shift = 0; shift < bits; shift += 8) {
        prof_begin(K_RADIX_HIST, 4.0 * n + 1024.0 * nb, st);
        hipLaunchKernelGGL(radix_hist_k, dim3(nb), dim3(kSortBlock), 0, st, ka, n, shift, counts, nb);
        prof_end(K_RADIX_HIST, st);
        PAMO_HIP(hipGetLastError());
        const int raw = nb <= kRawOffsetBlocks ? 1 : 0;
        if (!raw) PAMO_TRY(scan_exclusive(counts, counts, 256LL * nb, sws, st));
        prof_begin(K_RADIX_SCATTER, 16.0 * n + 1024.0 * nb, st);
        hipLaunchKernelGGL(radix_scatter_k, dim3(nb), dim3(kSortBlock), 0, st, ka, va, kb, vb, n, shift,
                           (const int*)counts, nb, raw);
        prof_end(K_RADIX_SCATTER, st);
        PAMO_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        PAMO_HIP(hipMemcpyAsync(keys, ka, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        PAMO_HIP(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    }
    return 0;
}

template <typename T>
__global__ void fill_k(T* __restrict__ p, T v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T>
inline int fill(T* p, T v, long long n, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fill_k<T>, dim3(grid_for(n, 256)), dim3(256), 0, st, p, v, n);
    PAMO_HIP(hipGetLastError());
    return 0;
}

}  // namespace
}  // namespace pamo
