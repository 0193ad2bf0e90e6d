// Device kernels of the stage-2 round (see stage2.hip for the mapping to
// the reference).  Layout in HBM:
//   positions   float4 (x, y, z, 0) per vertex: every consumer gathers whole
//               vertices, so one 16-byte load replaces three 4-byte loads;
//   faces       SoA int arrays i[], j[], k[] (coalesced per-face passes);
//   rings       ring_off[V+1] + ring[3F] (face ids, ascending per vertex) and
//               rv[3F] = int4 (i, j, k, face) so one-ring passes read each
//               ring face with one 16-byte load;
//   quadrics    3 float4 per vertex (10 symmetric entries + 2 pad).
#pragma once
#include "exact_guard.h"
#include "guard_ref.h"
#include "hip_common.h"

namespace pamo {
// Internal linkage: every translation unit that includes this header gets
// its own copy of the kernels.
namespace {

constexpr int kBlock = 256;

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ F3 f3sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 f3add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ float f3dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 f3cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float f3norm(F3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ F3 f3divs(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// ternary max-then-min clamp of cusimp_free.cu:29-34 (NaN -> upper bound)
__device__ __forceinline__ float tclamp(float x, float a, float b) {
    float m = a > x ? a : x;
    return m < b ? m : b;
}
// float -> uint32 with saturation (CUDA's conversion rule, cusimp_free.cu:584)
__device__ __forceinline__ uint32_t sat_u32(float q) {
    if (!(q < 4294967296.0f)) return 0xFFFFFFFFu;
    if (q <= 0.0f) return 0u;
    return (uint32_t)q;
}

struct Mesh {
    float4* p;
    int *i, *j, *k;
};

__device__ __forceinline__ F3 ld(const Mesh& m, int v) { return f3(m.p[v]); }
__device__ __forceinline__ void st(const Mesh& m, int v, F3 q) { m.p[v] = make_float4(q.x, q.y, q.z, 0.0f); }
__device__ __forceinline__ bool face_dead(int a, int b, int c) { return a == -1 && b == -1 && c == -1; }

struct Feistel {
    uint32_t lbits, rbits, lmask, rmask;
    uint32_t key[24];
};

enum Counter {
    C_LINE = 0,     // line-collapse count
    C_NREG = 1,     // regular collapse records
    C_NLIVE = 2,    // guard tree leaves (faces live at the start of the round)
    C_PUSHED = 3,   // undo candidates
    C_NV = 4,       // next-round vertices
    C_NF = 5,       // next-round faces
    C_NEDGE = 6,
    C_NQUERY = 7,   // guard queries
    C_UNDECIDED = 8,  // exact guard: pairs the filtered predicates left undecided
    C_WALK_ERR = 9,   // a guard tree walk exceeded its bound (walk_overrun)
    C_BAD_UNDO = 10,  // an incoming undo entry outside the context's points
    C_LIVE_LEAVES = 11,  // guard tree leaves still live after a compaction
    C_COUNT = 16
};

// ---------------------------------------------------------------- load / shuffle (A3)
__global__ void load_points_k(const float* __restrict__ aos, float4* __restrict__ p, float4* __restrict__ op,
                              int* __restrict__ occ, int n) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    if (v == n) {
        occ[n] = 0;
        return;
    }
    float4 q = make_float4(aos[3 * (long long)v], aos[3 * (long long)v + 1], aos[3 * (long long)v + 2], 0.0f);
    p[v] = q;
    op[v] = q;
    occ[v] = 1;
}

__global__ void load_tris_k(const int* __restrict__ aos, int* ti, int* tj, int* tk, int n) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    ti[f] = aos[3 * (long long)f];
    tj[f] = aos[3 * (long long)f + 1];
    tk[f] = aos[3 * (long long)f + 2];
}

__device__ __forceinline__ uint32_t feistel_apply(const Feistel& fs, uint32_t v) {
    uint32_t L = v >> fs.rbits, R = v & fs.rmask;
    for (int r = 0; r < 24; ++r) {
        uint64_t prod = 0xD2B74407B1CE6E93ull * (uint64_t)L;
        uint32_t hi = (uint32_t)(prod >> 32), lo = (uint32_t)prod;
        lo = (lo << (fs.rbits - fs.lbits)) | (R >> fs.lbits);
        L = ((hi ^ fs.key[r]) ^ R) & fs.lmask;
        R = lo & fs.rmask;
    }
    return (L << fs.rbits) | R;
}

// Feistel permutation of [0, 2^bits); images below m are kept in order.
__global__ void shuffle_flags_k(Feistel fs, uint32_t domain, uint32_t m, int* __restrict__ flag,
                                uint32_t* __restrict__ img) {
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx > domain) return;
    if (idx == domain) {
        flag[idx] = 0;
        return;
    }
    uint32_t k = feistel_apply(fs, idx);
    img[idx] = k;
    flag[idx] = k < m ? 1 : 0;
}

__global__ void shuffle_scatter_k(uint32_t domain, uint32_t m, const int* __restrict__ pos,
                                  const uint32_t* __restrict__ img, const int* __restrict__ si,
                                  const int* __restrict__ sj, const int* __restrict__ sk, int* di, int* dj, int* dk) {
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= domain) return;
    uint32_t k = img[idx];
    if (k >= m) return;
    int q = pos[idx];
    di[q] = si[k];
    dj[q] = sj[k];
    dk[q] = sk[k];
}

// ---------------------------------------------------------------- undo-list prologue
// pts_map holds cap + 1 entries; an entry outside them cannot come from an
// earlier call of the context: it is reported and maps to 0.
__global__ void stuck_prologue_k(const int* __restrict__ undo_in, int n, int* tmp_undo, int* undo_list,
                                 const int* __restrict__ pts_map, int cap, int* err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = undo_in[i];
    tmp_undo[i] = v;
    if (v < 0 || v > cap) {
        atomicOr(err, 1);
        undo_list[i] = 0;
        return;
    }
    undo_list[i] = pts_map[v];
}

// The invalid table holds epochs: vertex v is invalid when invalid[v] equals
// the current epoch.  Starting a new epoch clears the table (the
// reference's memset, cusimp_free.cu:905-908) without touching it.
__global__ void invalid_set_k(const int* __restrict__ undo_in, int n, uint32_t* invalid, int cap, uint32_t epoch) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = undo_in[i];
    if (v >= 0 && v < cap) invalid[v] = epoch;
}

// ---------------------------------------------------------------- rings and edges (A4, A5)

__global__ void ring_fill_k(Mesh m, int nT, const int* __restrict__ off, int* __restrict__ fill,
                            int* __restrict__ raw, int* __restrict__ owner) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nT) return;
    int a = m.i[f], b = m.j[f], c = m.k[f];
    if (face_dead(a, b, c)) return;
    int q = off[a] + atomicAdd(&fill[a], 1);
    raw[q] = f;
    owner[q] = a;
    q = off[b] + atomicAdd(&fill[b], 1);
    raw[q] = f;
    owner[q] = b;
    q = off[c] + atomicAdd(&fill[c], 1);
    raw[q] = f;
    owner[q] = c;
}

// Canonical ring order (ascending face index) with the packed int4 entries,
// one lane per ring slot: the slot's place is its rank among the slots of its
// vertex, ties (a face holding the vertex twice) broken by slot.  Independent
// loads only, so a high-valence ring costs no serial chain.
__global__ void __launch_bounds__(kBlock) ring_rank_k(Mesh m, const int* __restrict__ off,
                                                      const int* __restrict__ raw, const int* __restrict__ owner,
                                                      int n_slots_cap, const int* __restrict__ n_slots,
                                                      int* __restrict__ ring, int4* __restrict__ rv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots_cap || s >= *n_slots) return;
    const int v = owner[s], b = off[v], e = off[v + 1];
    const int x = raw[s];
    int r = 0;
    for (int q = b; q < e; ++q) {
        const int y = raw[q];
        r += (y < x || (y == x && q < s)) ? 1 : 0;
    }
    ring[b + r] = x;
    rv[b + r] = make_int4(m.i[x], m.j[x], m.k[x], x);
}

// Per-face counts in one pass: each live face adds 1 to its three vertices'
// ring counters (ring_count_k) and stores its emitted-edge count (edge_count_k).
__global__ void topo_count_k(Mesh m, int nT, int* __restrict__ ring_cnt, int* __restrict__ edge_cnt) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > nT) return;
    if (f == nT) {
        edge_cnt[nT] = 0;
        return;
    }
    int a = m.i[f], b = m.j[f], c = m.k[f];
    if (face_dead(a, b, c)) {
        edge_cnt[f] = 0;
        return;
    }
    atomicAdd(&ring_cnt[a], 1);
    atomicAdd(&ring_cnt[b], 1);
    atomicAdd(&ring_cnt[c], 1);
    edge_cnt[f] = (int)(a > b) + (int)(b > c) + (int)(c > a);
}

// Also stores the edge count off[nT] to cnt[C_NEDGE].
// Also counts the edges per smaller endpoint into order_cnt (zeroed by
// round_init_k), the histogram of the edge-cost launch order.
// also counts the edges per smaller endpoint (order_cnt) and keeps each
// edge's rank among them (rank), for the edge-cost launch order
__global__ void edge_fill_k(Mesh m, int nT, const int* __restrict__ off, int* __restrict__ eu, int* __restrict__ ev,
                            int* __restrict__ cnt, int* __restrict__ order_cnt, int* __restrict__ rank) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f == 0) cnt[C_NEDGE] = off[nT];
    if (f >= nT) return;
    int a = m.i[f], b = m.j[f], c = m.k[f];
    if (face_dead(a, b, c)) return;
    int q = off[f];
    if (a > b) { eu[q] = b; ev[q] = a; rank[q] = atomicAdd(&order_cnt[b], 1); ++q; }
    if (b > c) { eu[q] = c; ev[q] = b; rank[q] = atomicAdd(&order_cnt[c], 1); ++q; }
    if (c > a) { eu[q] = a; ev[q] = c; rank[q] = atomicAdd(&order_cnt[a], 1); ++q; }
}

// ---------------------------------------------------------------- quadrics and costs (A6, A7)
__global__ void __launch_bounds__(kBlock) quadric_k(Mesh m, const int* __restrict__ off,
                                                    const int4* __restrict__ rv, int nP, float4* __restrict__ Q) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nP) return;
    float q[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) q[c] = 0.0f;
    const int e = off[v + 1];
    for (int r = off[v]; r < e; ++r) {
        const int4 t = rv[r];
        F3 a = ld(m, t.x), b = ld(m, t.y), d = ld(m, t.z);
        F3 n = f3cross(f3sub(b, a), f3sub(d, a));
        n = f3divs(n, f3norm(n));
        float w = -f3dot(n, a);
        float p[4] = {n.x, n.y, n.z, w};
        int c = 0;
#pragma unroll
        for (int ii = 0; ii < 4; ++ii)
#pragma unroll
            for (int jj = ii; jj < 4; ++jj) q[c++] += p[ii] * p[jj];
    }
    Q[3 * (long long)v] = make_float4(q[0], q[1], q[2], q[3]);
    Q[3 * (long long)v + 1] = make_float4(q[4], q[5], q[6], q[7]);
    Q[3 * (long long)v + 2] = make_float4(q[8], q[9], 0.0f, 0.0f);
}

// Makes the stores of a group's lanes visible to the other lanes of the
// workgroup (the group shares one wave) before they read the same faces.
__device__ __forceinline__ void group_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Group mask of the calling lane's kG-lane group within its wave.
template <int kG>
__device__ __forceinline__ unsigned long long group_mask() {
    const int lane = threadIdx.x & (kWave - 1);
    return ((kG == 64) ? ~0ull : ((1ull << kG) - 1)) << (lane & ~(kG - 1));
}

// quadric_k with kG lanes per vertex: each lane forms the plane of every
// kG-th ring face and all lanes add the ten products face by face in the
// single-lane order (the planes come from their lanes by shuffles), so the
// sums are bit-identical to quadric_k's.
template <int kG>
__global__ void __launch_bounds__(kBlock) quadric_group_k(Mesh m, const int* __restrict__ off,
                                                          const int4* __restrict__ rv, int nP,
                                                          float4* __restrict__ Q) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = (int)(threadIdx.x & (kG - 1));
    const long long vv = gt / kG;
    if (vv >= nP) return;  // the whole group
    const int v = (int)vv;
    float q[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) q[c] = 0.0f;
    const int b = off[v], nf = off[v + 1] - b;
    for (int r0 = 0; r0 < nf; r0 += kG) {  // the same trip count in every lane of the group
        float p[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (r0 + g < nf) {
            const int4 t = rv[b + r0 + g];
            F3 a = ld(m, t.x), bb = ld(m, t.y), d = ld(m, t.z);
            F3 n = f3cross(f3sub(bb, a), f3sub(d, a));
            n = f3divs(n, f3norm(n));
            p[0] = n.x;
            p[1] = n.y;
            p[2] = n.z;
            p[3] = -f3dot(n, a);
        }
        const int nj = nf - r0 < kG ? nf - r0 : kG;
        for (int j = 0; j < kG; ++j) {
            float pj[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) pj[k] = __shfl(p[k], j, kG);
            if (j < nj) {
                int c = 0;
#pragma unroll
                for (int ii = 0; ii < 4; ++ii)
#pragma unroll
                    for (int jj = ii; jj < 4; ++jj) q[c++] += pj[ii] * pj[jj];
            }
        }
    }
    if (g != 0) return;
    Q[3 * (long long)v] = make_float4(q[0], q[1], q[2], q[3]);
    Q[3 * (long long)v + 1] = make_float4(q[4], q[5], q[6], q[7]);
    Q[3 * (long long)v + 2] = make_float4(q[8], q[9], 0.0f, 0.0f);
}

__device__ __forceinline__ int succ_in(int4 t, int x) {
    if (t.x == x) return t.y;
    if (t.y == x) return t.z;
    if (t.z == x) return t.x;
    return -2;
}

__device__ __forceinline__ float heron(F3 a, F3 b, F3 d) {
    float x = f3norm(f3sub(a, b)), y = f3norm(f3sub(b, d)), z = f3norm(f3sub(d, a));
    float s = (x + y + z) / 2;
    return sqrtf(s * (s - x) * (s - y) * (s - z));
}

constexpr int kLinkRing = 24;

// Link condition of edge (u, v) (:405-444): dup = #{(a, b) : succ_u(a) ==
// succ_v(b)} over the faces a of ring(u) and b of ring(v), where succ_x(t) is
// the vertex after x in face t; more than 2 rejects the collapse, fewer keeps
// the stale cost.  dup is symmetric in u and v and only its total matters, so
// the smaller ring sits in registers, kLinkRing successors per pass, and the
// larger one streams past it four faces per load batch.  Returns 3 as soon
// as the count exceeds 2.
__device__ __forceinline__ int link_dup(const int4* __restrict__ rv, int u, int ub, int ue, int v, int vb, int ve) {
    const bool swap = (ve - vb) > (ue - ub);
    const int ib = swap ? ub : vb, ie = swap ? ue : ve, iw = swap ? u : v;
    const int ob = swap ? vb : ub, oe = swap ? ve : ue, ow = swap ? v : u;
    int dup = 0;
    for (int c0 = ib; c0 < ie; c0 += kLinkRing) {
        int sv[kLinkRing];
#pragma unroll
        for (int b = 0; b < kLinkRing; ++b) sv[b] = c0 + b < ie ? succ_in(rv[c0 + b], iw) : -2;
        int a = ob;
        for (; a + 4 <= oe; a += 4) {
            const int4 t0 = rv[a], t1 = rv[a + 1], t2 = rv[a + 2], t3 = rv[a + 3];
            const int s0 = succ_in(t0, ow), s1 = succ_in(t1, ow), s2 = succ_in(t2, ow), s3 = succ_in(t3, ow);
#pragma unroll
            for (int b = 0; b < kLinkRing; ++b)
                dup += (s0 == sv[b] ? 1 : 0) + (s1 == sv[b] ? 1 : 0) + (s2 == sv[b] ? 1 : 0) + (s3 == sv[b] ? 1 : 0);
            if (dup > 2) return 3;
        }
        for (; a < oe; ++a) {
            const int su = succ_in(rv[a], ow);
#pragma unroll
            for (int b = 0; b < kLinkRing; ++b) dup += (su == sv[b]) ? 1 : 0;
        }
        if (dup > 2) return 3;
    }
    return dup;
}

struct CostParams {
    float edge_s, tres;
    int is_stuck;
    uint32_t inv_epoch;  // the invalid table's current epoch
};

// Opt-in optimal placement (the reference always collapses to the midpoint,
// cusimp_free.cu:491,687): the minimiser of x^T M x over x = (p, 1) for
// M = Q_u + Q_v, from the 3x3 system A p = -b solved in FP64 by cofactors in
// registers.  The midpoint is kept when A is singular
// (|det| <= 1e-9 max|a_ij|^3, NaN included) or when the minimiser lies
// farther from the midpoint than the edge is long.  M holds the 10 float
// sums 00 01 02 03 11 12 13 22 23 33.
__device__ __forceinline__ F3 place_optimal(const float M[10], F3 p0, F3 p1) {
    const F3 mid = f3divs(f3add(p0, p1), 2.0f);
    const double a00 = M[0], a01 = M[1], a02 = M[2], a11 = M[4], a12 = M[5], a22 = M[7];
    const double b0 = -(double)M[3], b1 = -(double)M[6], b2 = -(double)M[8];
    const double c00 = a11 * a22 - a12 * a12;
    const double c01 = a02 * a12 - a01 * a22;
    const double c02 = a01 * a12 - a02 * a11;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    double amax = fabs(a00);
    amax = fmax(amax, fabs(a01));
    amax = fmax(amax, fabs(a02));
    amax = fmax(amax, fabs(a11));
    amax = fmax(amax, fabs(a12));
    amax = fmax(amax, fabs(a22));
    if (!(fabs(det) > 1e-9 * amax * amax * amax)) return mid;
    const double c11 = a00 * a22 - a02 * a02;
    const double c12 = a01 * a02 - a00 * a12;
    const double c22 = a00 * a11 - a01 * a01;
    const double x = (c00 * b0 + c01 * b1 + c02 * b2) / det;
    const double y = (c01 * b0 + c11 * b1 + c12 * b2) / det;
    const double z = (c02 * b0 + c12 * b1 + c22 * b2) / det;
    const double ex = (double)p1.x - (double)p0.x, ey = (double)p1.y - (double)p0.y, ez = (double)p1.z - (double)p0.z;
    const double dx = x - (double)mid.x, dy = y - (double)mid.y, dz = z - (double)mid.z;
    if (!(dx * dx + dy * dy + dz * dz <= ex * ex + ey * ey + ez * ez)) return mid;
    return F3{(float)x, (float)y, (float)z};
}

// M = Q_u + Q_v from the per-vertex records (3 float4: 00 01 02 03 | 11 12 13 22 | 23 33 - -).
__device__ __forceinline__ void edge_quadric(const float4* __restrict__ Q, int u, int v, float M[10]) {
    const float4 ua = Q[3 * (long long)u], ub = Q[3 * (long long)u + 1], uc = Q[3 * (long long)u + 2];
    const float4 va = Q[3 * (long long)v], vb = Q[3 * (long long)v + 1], vc = Q[3 * (long long)v + 2];
    M[0] = ua.x + va.x; M[1] = ua.y + va.y; M[2] = ua.z + va.z; M[3] = ua.w + va.w;
    M[4] = ub.x + vb.x; M[5] = ub.y + vb.y; M[6] = ub.z + vb.z; M[7] = ub.w + vb.w;
    M[8] = uc.x + vc.x; M[9] = uc.y + vc.y;
}

// Launch order of the edge-cost lanes: edge ids grouped by their smaller
// endpoint (counting sort: histogram with ranks, exclusive scan, scatter).  Vertex ids of
// a mesh are spatially coherent while edge ids follow the shuffled face
// order, so in this order the lanes of a wave read neighbouring rings,
// positions and quadrics.  The order inside one group (the ranks edge_fill_k
// took) depends on atomic timing; it changes only which lane computes an
// edge, never a result.
__global__ void edge_order_scatter_k(const int* __restrict__ eu, const int* __restrict__ rank, int nE,
                                     const int* __restrict__ start, int* __restrict__ order) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nE) order[start[eu[e]] + rank[e]] = e;
}

__device__ __forceinline__ uint64_t edge_key(uint32_t c, int e) { return ((uint64_t)c << 32) | (uint32_t)e; }

// propagate_edge_cost_kernel (:587-610) fused into the cost kernels: the
// edge's key (cost << 32 | edge) lowers the per-vertex minimum key of both
// endpoints (see propagate_k).  An edge whose cost the round leaves as it was
// (stuck round, link count below 2) propagates the stored value, as the
// separate pass read it.
__device__ __forceinline__ void prop_key(unsigned long long* __restrict__ vmin, int u, int v, uint32_t c, int e) {
    const unsigned long long k = edge_key(c, e);
    atomicMin(&vmin[u], k);
    atomicMin(&vmin[v], k);
}

// compute_edge_cost_kernel (cusimp_free.cu:382-585), one lane per edge; lane
// t takes edge order[t] and writes cost[order[t]], then propagates its key.

template <bool kOptimal>
__global__ void __launch_bounds__(kBlock) edge_cost_k(Mesh m, const int* __restrict__ off,
                                                      const int4* __restrict__ rv, const int* __restrict__ eu,
                                                      const int* __restrict__ ev, const int* __restrict__ order,
                                                      int nE, const uint32_t* __restrict__ invalid,
                                                      const float4* __restrict__ Q, CostParams cp,
                                                      uint32_t* __restrict__ cost,
                                                      unsigned long long* __restrict__ vmin) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nE) return;
    const int e = order[t];
    const int u = eu[e], v = ev[e];
    if (invalid[u] == cp.inv_epoch && invalid[v] == cp.inv_epoch) {
        cost[e] = 0xFFFFFFFFu;
        prop_key(vmin, u, v, 0xFFFFFFFFu, e);
        return;
    }
    if (cp.is_stuck) {
        prop_key(vmin, u, v, cost[e], e);
        return;
    }
    const int ub = off[u], ue = off[u + 1], vb = off[v], ve = off[v + 1];
    const int dup = link_dup(rv, u, ub, ue, v, vb, ve);
    if (dup > 2) {
        cost[e] = 0xFFFFFFFFu;
        prop_key(vmin, u, v, 0xFFFFFFFFu, e);
        return;
    }
    if (dup != 2) {
        prop_key(vmin, u, v, cost[e], e);
        return;
    }
    const F3 p0 = ld(m, u), p1 = ld(m, v);
    float M[10];
    edge_quadric(Q, u, v, M);
    const F3 vm = kOptimal ? place_optimal(M, p0, p1) : f3divs(f3add(p0, p1), 2.0f);
    const float k4s3 = 4.0f * sqrtf(3.0f);
    // one pass over both rings: the edge-length sum and the skinny / flip
    // terms read the same three vertices; each sum keeps its own order
    float elen = 0.0f, qa = 0.0f;
    int nedge = 0, ntri = 0;
    for (int side = 0; side < 2; ++side) {
        const int w = side ? v : u, other = side ? u : v;
        const int b0 = side ? vb : ub, b1 = side ? ve : ue;
        for (int a = b0; a < b1; ++a) {
            const int4 t = rv[a];
            const F3 pi = ld(m, t.x), pj = ld(m, t.y), pk = ld(m, t.z);
            if (t.x > t.y) { elen += f3norm(f3sub(pi, pj)); ++nedge; }
            if (t.y > t.z) { elen += f3norm(f3sub(pj, pk)); ++nedge; }
            if (t.z > t.x) { elen += f3norm(f3sub(pk, pi)); ++nedge; }
            if (t.x == other || t.y == other || t.z == other) continue;
            const F3 n0 = t.x == w ? vm : pi, n1 = t.y == w ? vm : pj, n2 = t.z == w ? vm : pk;
            F3 on = f3cross(f3sub(pj, pi), f3sub(pk, pi));
            on = f3divs(on, f3norm(on));
            F3 nn = f3cross(f3sub(n1, n0), f3sub(n2, n0));
            nn = f3divs(nn, f3norm(nn));
            if (f3dot(on, nn) < 0) {
                cost[e] = 0xFFFFFFFFu;
                prop_key(vmin, u, v, 0xFFFFFFFFu, e);
                return;
            }
            float l01 = f3norm(f3sub(n0, n1)), l12 = f3norm(f3sub(n1, n2)), l20 = f3norm(f3sub(n2, n0));
            float qv = k4s3 * heron(n0, n1, n2) / (l01 * l01) + l12 * l12 + l20 * l20 + 0.0000001f;
            qa += 1.0f - tclamp(qv, 0.0f, 1.0f);
            ++ntri;
        }
    }
    elen = elen / (float)nedge / cp.edge_s * cp.tres;
    elen = elen + f3norm(f3sub(p0, p1)) / cp.edge_s * cp.tres;
    qa = 5.0f * qa;
    // M = Q_u + Q_v; storage 0:00 1:01 2:02 3:03 4:11 5:12 6:13 7:22 8:23 9:33
    const float M0 = M[0], M1 = M[1], M2 = M[2], M3 = M[3], M4 = M[4], M5 = M[5], M6 = M[6], M7 = M[7];
    const float M8 = M[8], M9 = M[9];
    const float x0 = vm.x, x1 = vm.y, x2 = vm.z, x3 = 1.0f;
    float r0 = M0 * x0 + M1 * x1 + M2 * x2 + M3 * x3;
    float r1 = M1 * x0 + M4 * x1 + M5 * x2 + M6 * x3;
    float r2 = M2 * x0 + M5 * x1 + M7 * x2 + M8 * x3;
    float r3 = M3 * x0 + M6 * x1 + M8 * x2 + M9 * x3;
    float qc = (r0 * x0 + r1 * x1 + r2 * x2 + r3 * x3) / (cp.edge_s * cp.edge_s);
    float tot = qc + elen + qa / (float)ntri * cp.tres;
    const uint32_t ce = sat_u32(tclamp(tot, 0.0f, 10.0f) / 10.0f * 4294967296.0f);
    cost[e] = ce;
    prop_key(vmin, u, v, ce, e);
}

// link_dup's count over the outer-ring faces g, g + kG, ... (one lane's
// share of a kG-lane group); the shares add up to link_dup's count, and a
// share above 2 is reported as 3.
template <int kG>
__device__ __forceinline__ int link_dup_share(const int4* __restrict__ rv, int u, int ub, int ue, int v, int vb,
                                              int ve, int g) {
    const bool swap = (ve - vb) > (ue - ub);
    const int ib = swap ? ub : vb, ie = swap ? ue : ve, iw = swap ? u : v;
    const int ob = swap ? vb : ub, oe = swap ? ve : ue, ow = swap ? v : u;
    int dup = 0;
    for (int c0 = ib; c0 < ie; c0 += kLinkRing) {
        int sv[kLinkRing];
#pragma unroll
        for (int b = 0; b < kLinkRing; ++b) sv[b] = c0 + b < ie ? succ_in(rv[c0 + b], iw) : -2;
        for (int a = ob + g; a < oe; a += kG) {
            const int su = succ_in(rv[a], ow);
#pragma unroll
            for (int b = 0; b < kLinkRing; ++b) dup += (su == sv[b]) ? 1 : 0;
        }
        if (dup > 2) return 3;
    }
    return dup;
}

// edge_cost_k with kG lanes per edge (kG divides the wave).  The link count
// is split over the lanes (an integer sum); each lane then evaluates every
// kG-th face of the two rings (positions, lengths, normals, flip test,
// quality term) and all lanes of the group add the terms up face by face
// in the single-lane order, taking each face's terms from its lane by
// shuffles, so every float sum is formed in the same order as in
// edge_cost_k and the costs are equal bit for bit.  A flip anywhere sets
// the maximum cost, as the single-lane loop's first flip does.  The ring
// loads of one edge then take about 1/kG of the dependent round trips.
template <bool kOptimal, int kG>
__global__ void __launch_bounds__(kBlock) edge_cost_group_k(Mesh m, const int* __restrict__ off,
                                                            const int4* __restrict__ rv, const int* __restrict__ eu,
                                                            const int* __restrict__ ev, const int* __restrict__ order,
                                                            int nE, const uint32_t* __restrict__ invalid,
                                                            const float4* __restrict__ Q, CostParams cp,
                                                            uint32_t* __restrict__ cost,
                                                            unsigned long long* __restrict__ vmin) {
    static_assert(kWave % kG == 0, "a group must not straddle waves");
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = (int)(threadIdx.x & (kG - 1));
    const long long t = gt / kG;
    if (t >= nE) return;  // the whole group
    const int e = order[t];
    const int u = eu[e], v = ev[e];
    if (invalid[u] == cp.inv_epoch && invalid[v] == cp.inv_epoch) {
        if (g == 0) {
            cost[e] = 0xFFFFFFFFu;
            prop_key(vmin, u, v, 0xFFFFFFFFu, e);
        }
        return;
    }
    if (cp.is_stuck) {
        if (g == 0) prop_key(vmin, u, v, cost[e], e);
        return;
    }
    const int ub = off[u], ue = off[u + 1], vb = off[v], ve = off[v + 1];
    int dup = link_dup_share<kG>(rv, u, ub, ue, v, vb, ve, g);
#pragma unroll
    for (int d = kG / 2; d > 0; d >>= 1) dup += __shfl_xor(dup, d, kG);
    if (dup > 2) {
        if (g == 0) {
            cost[e] = 0xFFFFFFFFu;
            prop_key(vmin, u, v, 0xFFFFFFFFu, e);
        }
        return;
    }
    if (dup != 2) {
        if (g == 0) prop_key(vmin, u, v, cost[e], e);
        return;
    }
    const F3 p0 = ld(m, u), p1 = ld(m, v);
    float M[10];
    edge_quadric(Q, u, v, M);
    const F3 vm = kOptimal ? place_optimal(M, p0, p1) : f3divs(f3add(p0, p1), 2.0f);
    const float k4s3 = 4.0f * sqrtf(3.0f);
    const int nu = ue - ub, nf = nu + (ve - vb);
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned long long gmask = ((kG == 64) ? ~0ull : ((1ull << kG) - 1)) << (lane & ~(kG - 1));
    float elen = 0.0f, qa = 0.0f;
    int nedge = 0, ntri = 0;
    for (int r0 = 0; r0 < nf; r0 += kG) {  // the same trip count in every lane of the group
        const int i = r0 + g;
        float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f, qt = 0.0f;
        int bits = 0;  // 1, 2, 4: edge-length terms present; 8: quality term present
        bool flip = false;
        if (i < nf) {
            const bool side = i >= nu;
            const int w = side ? v : u, other = side ? u : v;
            const int4 tt = rv[side ? vb + (i - nu) : ub + i];
            const F3 pi = ld(m, tt.x), pj = ld(m, tt.y), pk = ld(m, tt.z);
            if (tt.x > tt.y) { l0 = f3norm(f3sub(pi, pj)); bits |= 1; }
            if (tt.y > tt.z) { l1 = f3norm(f3sub(pj, pk)); bits |= 2; }
            if (tt.z > tt.x) { l2 = f3norm(f3sub(pk, pi)); bits |= 4; }
            if (!(tt.x == other || tt.y == other || tt.z == other)) {
                const F3 n0 = tt.x == w ? vm : pi, n1 = tt.y == w ? vm : pj, n2 = tt.z == w ? vm : pk;
                F3 on = f3cross(f3sub(pj, pi), f3sub(pk, pi));
                on = f3divs(on, f3norm(on));
                F3 nn = f3cross(f3sub(n1, n0), f3sub(n2, n0));
                nn = f3divs(nn, f3norm(nn));
                if (f3dot(on, nn) < 0) {
                    flip = true;
                } else {
                    float l01 = f3norm(f3sub(n0, n1)), l12 = f3norm(f3sub(n1, n2)), l20 = f3norm(f3sub(n2, n0));
                    float qv = k4s3 * heron(n0, n1, n2) / (l01 * l01) + l12 * l12 + l20 * l20 + 0.0000001f;
                    qt = 1.0f - tclamp(qv, 0.0f, 1.0f);
                    bits |= 8;
                }
            }
        }
        if (__ballot(flip) & gmask) {
            if (g == 0) {
                cost[e] = 0xFFFFFFFFu;
                prop_key(vmin, u, v, 0xFFFFFFFFu, e);
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < kG; ++j) {
            const int bj = __shfl(bits, j, kG);
            const float a0 = __shfl(l0, j, kG), a1 = __shfl(l1, j, kG), a2 = __shfl(l2, j, kG);
            const float qj = __shfl(qt, j, kG);
            if (bj & 1) { elen += a0; ++nedge; }
            if (bj & 2) { elen += a1; ++nedge; }
            if (bj & 4) { elen += a2; ++nedge; }
            if (bj & 8) { qa += qj; ++ntri; }
        }
    }
    if (g != 0) return;
    elen = elen / (float)nedge / cp.edge_s * cp.tres;
    elen = elen + f3norm(f3sub(p0, p1)) / cp.edge_s * cp.tres;
    qa = 5.0f * qa;
    const float M0 = M[0], M1 = M[1], M2 = M[2], M3 = M[3], M4 = M[4], M5 = M[5], M6 = M[6], M7 = M[7];
    const float M8 = M[8], M9 = M[9];
    const float x0 = vm.x, x1 = vm.y, x2 = vm.z, x3 = 1.0f;
    float r0 = M0 * x0 + M1 * x1 + M2 * x2 + M3 * x3;
    float r1 = M1 * x0 + M4 * x1 + M5 * x2 + M6 * x3;
    float r2 = M2 * x0 + M5 * x1 + M7 * x2 + M8 * x3;
    float r3 = M3 * x0 + M6 * x1 + M8 * x2 + M9 * x3;
    float qc = (r0 * x0 + r1 * x1 + r2 * x2 + r3 * x3) / (cp.edge_s * cp.edge_s);
    float tot = qc + elen + qa / (float)ntri * cp.tres;
    const uint32_t ce = sat_u32(tclamp(tot, 0.0f, 10.0f) / 10.0f * 4294967296.0f);
    cost[e] = ce;
    prop_key(vmin, u, v, ce, e);
}

// ---------------------------------------------------------------- independent set + collapse (A8, A9)

// propagate_edge_cost_kernel (:587-610) keeps, per face t, the minimum key
// over edges whose endpoint one-rings contain t.  A face lies in ring(x)
// exactly when x is one of its vertices, so tri_min[t] = min over t's three
// vertices of vmin[x], vmin[x] = min key over edges incident to x: two
// atomics per edge instead of one per ring face.
__global__ void __launch_bounds__(kBlock) propagate_k(const int* __restrict__ eu, const int* __restrict__ ev, int nE,
                                                      const uint32_t* __restrict__ cost,
                                                      unsigned long long* __restrict__ vmin) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const unsigned long long k = edge_key(cost[e], e);
    atomicMin(&vmin[eu[e]], k);
    atomicMin(&vmin[ev[e]], k);
}

__device__ __forceinline__ bool tri_has(int a, int b, int c, int x) { return a == x || b == x || c == x; }

// Face rewrite shared by the regular and the line collapse (:692-715, :249-273).
__device__ __forceinline__ void rewrite_rings(Mesh m, const int* __restrict__ off, const int* __restrict__ ring,
                                              int u, int v) {
    for (int a = off[u]; a < off[u + 1]; ++a) {
        int f = ring[a];
        if (tri_has(m.i[f], m.j[f], m.k[f], v)) { m.i[f] = -1; m.j[f] = -1; m.k[f] = -1; }
    }
    for (int a = off[v]; a < off[v + 1]; ++a) {
        int f = ring[a];
        int ii = m.i[f], jj = m.j[f], kk = m.k[f];
        if (tri_has(ii, jj, kk, u)) { m.i[f] = -1; m.j[f] = -1; m.k[f] = -1; }
        else if (ii == v) m.i[f] = u;
        else if (jj == v) m.j[f] = u;
        else if (kk == v) m.k[f] = u;
    }
}

// is_skinny_triangle (:617-637): area below 1e-6 * (longest edge)^2, the
// threshold formed in FP64 as written.
__device__ __forceinline__ bool skinny(F3 p, F3 q, F3 r) {
    F3 pq = f3sub(q, p), pr = f3sub(r, p), qr = f3sub(r, q);
    float lpq = sqrtf(f3dot(pq, pq)), lpr = sqrtf(f3dot(pr, pr)), lqr = sqrtf(f3dot(qr, qr));
    float ml = fmaxf(lpq, fmaxf(lpr, lqr));
    F3 cr = {pq.y * pr.z - pq.z * pr.y, pq.z * pr.x - pq.x * pr.z, pq.x * pr.y - pq.y * pr.x};
    float thr = (float)(1e-6 * (double)ml * (double)ml);
    float area = 0.5f * sqrtf(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
    return area < thr;
}

struct Orig {
    const float4* p;
    const int *i, *j, *k;
};

__device__ __forceinline__ void restore_edge(Mesh m, Orig o, int* occ, const int* __restrict__ off,
                                             const int* __restrict__ ring, int u, int v) {
    occ[v] = 1;
    m.p[u] = o.p[u];
    m.p[v] = o.p[v];
    for (int side = 0; side < 2; ++side) {
        int w = side ? v : u;
        for (int a = off[w]; a < off[w + 1]; ++a) {
            int f = ring[a];
            m.i[f] = o.i[f]; m.j[f] = o.j[f]; m.k[f] = o.k[f];
        }
    }
}

// Selection (:639-681) fused with the collapse (:683-773).  Selected edges
// have pairwise disjoint one-ring face sets, and no selected edge moves a
// vertex of another selected edge's rings, so the fused pass is race free
// and its result does not depend on execution order.
// Selected <=> cost <= collapse_t and every face t of ring(u) u ring(v) has
// tri_min[t] == key(e), i.e. min(vmin of t's vertices) == key(e).
template <bool kOptimal>
__global__ void __launch_bounds__(kBlock) select_collapse_k(Mesh m, Orig o, int* occ, const int* __restrict__ off,
                                                            const int* __restrict__ ring, const int4* __restrict__ rv,
                                                            const int* __restrict__ eu, const int* __restrict__ ev,
                                                            int nE, const uint32_t* __restrict__ cost,
                                                            const unsigned long long* __restrict__ vmin,
                                                            uint32_t collapse_t, int* __restrict__ sel,
                                                            const float4* __restrict__ Q,
                                                            uint8_t* __restrict__ changed) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > nE) return;
    if (e == nE) {
        sel[nE] = 0;
        return;
    }
    const uint32_t c = cost[e];
    int s = 0;
    if (c <= collapse_t) {
        const unsigned long long key = edge_key(c, e);
        const int u = eu[e], v = ev[e];
        bool ok = true;
        for (int side = 0; side < 2 && ok; ++side) {
            const int w = side ? v : u;
            for (int a = off[w]; a < off[w + 1] && ok; ++a) {
                const int4 t = rv[a];
                unsigned long long mn = vmin[t.x];
                unsigned long long y = vmin[t.y], z = vmin[t.z];
                mn = y < mn ? y : mn;
                mn = z < mn ? z : mn;
                ok = mn == key;
            }
        }
        if (ok) {
            s = 1;
            // the record's one-ring faces are the guard's queries
            for (int a = off[u]; a < off[u + 1]; ++a) changed[ring[a]] = 1;
            for (int a = off[v]; a < off[v + 1]; ++a) changed[ring[a]] = 1;
            F3 vm;
            if (kOptimal) {
                float M[10];
                edge_quadric(Q, u, v, M);
                vm = place_optimal(M, ld(m, u), ld(m, v));
            } else {
                vm = f3divs(f3add(ld(m, u), ld(m, v)), 2.0f);
            }
            st(m, u, vm);
            st(m, v, F3{0.0f, 0.0f, 0.0f});
            occ[v] = 0;
            rewrite_rings(m, off, ring, u, v);
            bool sk = false;
            for (int side = 0; side < 2 && !sk; ++side) {
                int w = side ? v : u;
                for (int a = off[w]; a < off[w + 1]; ++a) {
                    int f = ring[a];
                    int ii = m.i[f];
                    if (ii == -1) continue;
                    if (skinny(ld(m, ii), ld(m, m.j[f]), ld(m, m.k[f]))) { sk = true; break; }
                }
            }
            if (sk) restore_edge(m, o, occ, off, ring, u, v);
        }
    }
    sel[e] = s;
}

// select_collapse_k with kG lanes per edge.  The lanes split the faces of
// the two rings: the selection test and the skinny test are
// order-independent (every face must hold the edge's key; any face
// skinny), and each face's rewrite and restore touch that face only.  The
// rewrite keeps the single-lane order between the rings: every face of
// ring(u) is rewritten before any face of ring(v) is read, since a face in
// both rings (it holds u and v) must be removed by the first pass.
template <bool kOptimal, int kG>
__global__ void __launch_bounds__(kBlock) select_collapse_group_k(Mesh m, Orig o, int* occ,
                                                                  const int* __restrict__ off,
                                                                  const int* __restrict__ ring,
                                                                  const int4* __restrict__ rv,
                                                                  const int* __restrict__ eu,
                                                                  const int* __restrict__ ev, int nE,
                                                                  const uint32_t* __restrict__ cost,
                                                                  const unsigned long long* __restrict__ vmin,
                                                                  uint32_t collapse_t, int* __restrict__ sel,
                                                                  const float4* __restrict__ Q,
                                                                  uint8_t* __restrict__ changed) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = (int)(threadIdx.x & (kG - 1));
    const long long ee = gt / kG;
    if (ee > nE) return;  // the whole group
    const int e = (int)ee;
    if (e == nE) {
        if (g == 0) sel[nE] = 0;
        return;
    }
    const uint32_t c = cost[e];
    if (c > collapse_t) {
        if (g == 0) sel[e] = 0;
        return;
    }
    const unsigned long long key = edge_key(c, e);
    const int u = eu[e], v = ev[e];
    const int ub = off[u], vb = off[v];
    const int nu = off[u + 1] - ub, nv = off[v + 1] - vb, nf = nu + nv;
    const unsigned long long gm = group_mask<kG>();
    bool bad = false;
    for (int i = g; i < nf && !bad; i += kG) {
        const int4 t = rv[i < nu ? ub + i : vb + (i - nu)];
        unsigned long long mn = vmin[t.x];
        const unsigned long long y = vmin[t.y], z = vmin[t.z];
        mn = y < mn ? y : mn;
        mn = z < mn ? z : mn;
        bad = mn != key;
    }
    if (__ballot(bad) & gm) {
        if (g == 0) sel[e] = 0;
        return;
    }
    // the record's one-ring faces are the guard's queries
    for (int i = g; i < nf; i += kG) changed[ring[i < nu ? ub + i : vb + (i - nu)]] = 1;
    F3 vm;
    if (kOptimal) {
        float M[10];
        edge_quadric(Q, u, v, M);
        vm = place_optimal(M, ld(m, u), ld(m, v));
    } else {
        vm = f3divs(f3add(ld(m, u), ld(m, v)), 2.0f);
    }
    if (g == 0) {
        st(m, u, vm);
        st(m, v, F3{0.0f, 0.0f, 0.0f});
        occ[v] = 0;
    }
    for (int i = g; i < nu; i += kG) {
        const int f = ring[ub + i];
        if (tri_has(m.i[f], m.j[f], m.k[f], v)) { m.i[f] = -1; m.j[f] = -1; m.k[f] = -1; }
    }
    group_sync();
    for (int i = g; i < nv; i += kG) {
        const int f = ring[vb + i];
        const int ii = m.i[f], jj = m.j[f], kk = m.k[f];
        if (tri_has(ii, jj, kk, u)) { m.i[f] = -1; m.j[f] = -1; m.k[f] = -1; }
        else if (ii == v) m.i[f] = u;
        else if (jj == v) m.j[f] = u;
        else if (kk == v) m.k[f] = u;
    }
    group_sync();
    bool sk = false;
    for (int i = g; i < nf && !sk; i += kG) {
        const int f = ring[i < nu ? ub + i : vb + (i - nu)];
        const int ii = m.i[f];
        if (ii == -1) continue;
        sk = skinny(ld(m, ii), ld(m, m.j[f]), ld(m, m.k[f]));
    }
    if (__ballot(sk) & gm) {
        // restore_edge across the group (every face gets its snapshot corners)
        if (g == 0) {
            occ[v] = 1;
            m.p[u] = o.p[u];
            m.p[v] = o.p[v];
        }
        for (int i = g; i < nf; i += kG) {
            const int f = ring[i < nu ? ub + i : vb + (i - nu)];
            m.i[f] = o.i[f];
            m.j[f] = o.j[f];
            m.k[f] = o.k[f];
        }
    }
    if (g == 0) sel[e] = 1;
}

// ---------------------------------------------------------------- legacy PaSP round (cusimp.cu)
// compute_edge_cost_kernel of the legacy simplifier (cusimp.cu:283-482):
// the link test, edge-length, skinny and flip terms and midpoint quadric cost
// of edge_cost_k, with the legacy arithmetic: the skinny ratio is formed in
// FP64 (`4.0f * sqrt(3)` takes the double root, so the expression is double
// until its float cast), it is not scaled by the skinny penalty, and the
// averaged edge length is counted twice (`edge_len += edge_len + |v0 - v1|
// / s * t`, cusimp.cu:378).  No invalid-vertex or stuck rules.
__global__ void __launch_bounds__(kBlock) pasp_cost_k(Mesh m, const int* __restrict__ off, const int4* __restrict__ rv,
                                                      const int* __restrict__ eu, const int* __restrict__ ev,
                                                      const int* __restrict__ order, int nE,
                                                      const float4* __restrict__ Q, float edge_s, float tres,
                                                      uint32_t* __restrict__ cost) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nE) return;
    const int e = order[t];
    const int u = eu[e], v = ev[e];
    const int ub = off[u], ue = off[u + 1], vb = off[v], ve = off[v + 1];
    const int dup = link_dup(rv, u, ub, ue, v, vb, ve);
    if (dup > 2) {
        cost[e] = 0xFFFFFFFFu;
        return;
    }
    if (dup != 2) return;
    const F3 p0 = ld(m, u), p1 = ld(m, v);
    const F3 vm = f3divs(f3add(p0, p1), 2.0f);
    const double k4s3 = 4.0 * sqrt(3.0);
    float elen = 0.0f, qa = 0.0f;
    int nedge = 0, ntri = 0;
    for (int side = 0; side < 2; ++side) {
        const int w = side ? v : u, other = side ? u : v;
        const int b0 = side ? vb : ub, b1 = side ? ve : ue;
        for (int a = b0; a < b1; ++a) {
            const int4 tt = rv[a];
            const F3 pi = ld(m, tt.x), pj = ld(m, tt.y), pk = ld(m, tt.z);
            if (tt.x > tt.y) { elen += f3norm(f3sub(pi, pj)); ++nedge; }
            if (tt.y > tt.z) { elen += f3norm(f3sub(pj, pk)); ++nedge; }
            if (tt.z > tt.x) { elen += f3norm(f3sub(pk, pi)); ++nedge; }
            if (tt.x == other || tt.y == other || tt.z == other) continue;
            const F3 n0 = tt.x == w ? vm : pi, n1 = tt.y == w ? vm : pj, n2 = tt.z == w ? vm : pk;
            F3 on = f3cross(f3sub(pj, pi), f3sub(pk, pi));
            on = f3divs(on, f3norm(on));
            F3 nn = f3cross(f3sub(n1, n0), f3sub(n2, n0));
            nn = f3divs(nn, f3norm(nn));
            if (f3dot(on, nn) < 0) {
                cost[e] = 0xFFFFFFFFu;
                return;
            }
            const float l01 = f3norm(f3sub(n0, n1)), l12 = f3norm(f3sub(n1, n2)), l20 = f3norm(f3sub(n2, n0));
            const double qd = k4s3 * (double)heron(n0, n1, n2) / (double)(l01 * l01) + (double)(l12 * l12) +
                              (double)(l20 * l20) + (double)0.0000001f;
            qa += 1.0f - tclamp((float)qd, 0.0f, 1.0f);
            ++ntri;
        }
    }
    elen = elen / (float)nedge / edge_s * tres;
    elen = elen + (elen + f3norm(f3sub(p0, p1)) / edge_s * tres);
    const float4 ua = Q[3 * (long long)u], ub4 = Q[3 * (long long)u + 1], uc = Q[3 * (long long)u + 2];
    const float4 va = Q[3 * (long long)v], vb4 = Q[3 * (long long)v + 1], vc = Q[3 * (long long)v + 2];
    const float M0 = ua.x + va.x, M1 = ua.y + va.y, M2 = ua.z + va.z, M3 = ua.w + va.w;
    const float M4 = ub4.x + vb4.x, M5 = ub4.y + vb4.y, M6 = ub4.z + vb4.z, M7 = ub4.w + vb4.w;
    const float M8 = uc.x + vc.x, M9 = uc.y + vc.y;
    const float x0 = vm.x, x1 = vm.y, x2 = vm.z, x3 = 1.0f;
    const float r0 = M0 * x0 + M1 * x1 + M2 * x2 + M3 * x3;
    const float r1 = M1 * x0 + M4 * x1 + M5 * x2 + M6 * x3;
    const float r2 = M2 * x0 + M5 * x1 + M7 * x2 + M8 * x3;
    const float r3 = M3 * x0 + M6 * x1 + M8 * x2 + M9 * x3;
    const float qc = (r0 * x0 + r1 * x1 + r2 * x2 + r3 * x3) / (edge_s * edge_s);
    const float tot = qc + elen + qa / (float)ntri * tres;
    cost[e] = sat_u32(tclamp(tot, 0.0f, 10.0f) / 10.0f * 4294967296.0f);
}

// collapse_edge_kernel (cusimp.cu:511-585): an edge collapses when its cost
// is at most collapse_t and its key is the minimum of every face of
// ring(u) u ring(v) (propagate_k's per-vertex minima, as in
// select_collapse_k); u moves to the midpoint, v is removed, faces holding
// both are deleted and v's other faces take u.  Selected edges have disjoint
// rings, so the pass is race free and order independent.
__global__ void __launch_bounds__(kBlock) pasp_collapse_k(Mesh m, int* occ, const int* __restrict__ off,
                                                          const int* __restrict__ ring, const int4* __restrict__ rv,
                                                          const int* __restrict__ eu, const int* __restrict__ ev,
                                                          int nE, const uint32_t* __restrict__ cost,
                                                          const unsigned long long* __restrict__ vmin,
                                                          uint32_t collapse_t) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const uint32_t c = cost[e];
    if (c > collapse_t) return;
    const unsigned long long key = edge_key(c, e);
    const int u = eu[e], v = ev[e];
    for (int side = 0; side < 2; ++side) {
        const int w = side ? v : u;
        for (int a = off[w]; a < off[w + 1]; ++a) {
            const int4 t = rv[a];
            unsigned long long mn = vmin[t.x];
            const unsigned long long y = vmin[t.y], z = vmin[t.z];
            mn = y < mn ? y : mn;
            mn = z < mn ? z : mn;
            if (mn != key) return;
        }
    }
    const F3 vm = f3divs(f3add(ld(m, u), ld(m, v)), 2.0f);
    st(m, u, vm);
    st(m, v, F3{0.0f, 0.0f, 0.0f});
    occ[v] = 0;
    rewrite_rings(m, off, ring, u, v);
}

// ---------------------------------------------------------------- cleanup (A10)
// Line collapses are rare; applying them in edge order in one lane keeps
// the result deterministic where the reference races on shared faces.  The
// lane walks the list CleanupOp built, not the whole edge array.
__global__ void line_apply_k(Mesh m, int* occ, const int* __restrict__ off, const int* __restrict__ ring,
                             const int* __restrict__ eu, const int* __restrict__ ev, const int* __restrict__ recs,
                             const int* cnt, uint8_t* __restrict__ changed) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int base = cnt[C_NREG];
    const int nl = cnt[C_LINE];
    for (int w = 0; w < nl; ++w) {
        const int e = recs[base + w];
        const int u = eu[e], v = ev[e];
        // the record's one-ring faces are the guard's queries (the regular
        // records' are marked by the selection kernel)
        for (int a = off[u]; a < off[u + 1]; ++a) changed[ring[a]] = 1;
        for (int a = off[v]; a < off[v + 1]; ++a) changed[ring[a]] = 1;
        st(m, v, F3{0.0f, 0.0f, 0.0f});
        occ[v] = 0;
        rewrite_rings(m, off, ring, u, v);
    }
}

// ---------------------------------------------------------------- LBVH (A11)
__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// 30-bit Morton code at resolution 1024 (morton_code.cuh:19-37)
__device__ __forceinline__ uint32_t morton3(float x, float y, float z) {
    bool sx = x < 0.0f, sy = y < 0.0f, sz = z < 0.0f;
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    uint32_t xx = expand_bits((uint32_t)x), yy = expand_bits((uint32_t)y), zz = expand_bits((uint32_t)z);
    return (sx ? 0x80000000u : 0u) | (sy ? 0x40000000u : 0u) | (sz ? 0x20000000u : 0u) | (xx * 4 + yy * 2 + zz);
}

__device__ __forceinline__ void block_minmax6(float lo[3], float hi[3], float (*red)[kBlock]) {
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fminf(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmaxf(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
}

// Face boxes (an empty box, lo = +inf and hi = -inf, for a removed face),
// per-block partial bounding boxes of the live faces, per-block counts of
// the guard tree's leaves, and the guard's flags cleared (flagged[f] = 0).
// The leaves are the faces live at the start of the round (snap_i[f] != -1):
// every face live after the collapse, and every face an undo can restore.
__global__ void __launch_bounds__(kBlock) face_box_k(Mesh m, int nT, float4* __restrict__ fbox,
                                                     float* __restrict__ part, int* __restrict__ leaf_part,
                                                     uint8_t* __restrict__ flagged, const int* __restrict__ snap_i) {
    __shared__ float red[6][kBlock];
    __shared__ int nleaf;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const float inf = __builtin_huge_valf();
    float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    bool leaf = false;
    if (threadIdx.x == 0) nleaf = 0;
    __syncthreads();
    if (f == nT) {
        // the box of a carried tree's dead leaves (face id nT): empty
        fbox[2 * (long long)f] = make_float4(inf, inf, inf, 0.0f);
        fbox[2 * (long long)f + 1] = make_float4(-inf, -inf, -inf, 0.0f);
    }
    if (f < nT) {
        flagged[f] = 0;
        int a = m.i[f];
        if (a != -1) {
            F3 p = ld(m, a), q = ld(m, m.j[f]), r = ld(m, m.k[f]);
            lo[0] = fminf(p.x, fminf(q.x, r.x));
            lo[1] = fminf(p.y, fminf(q.y, r.y));
            lo[2] = fminf(p.z, fminf(q.z, r.z));
            hi[0] = fmaxf(p.x, fmaxf(q.x, r.x));
            hi[1] = fmaxf(p.y, fmaxf(q.y, r.y));
            hi[2] = fmaxf(p.z, fmaxf(q.z, r.z));
        }
        leaf = a != -1 || snap_i[f] != -1;
        fbox[2 * (long long)f] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        fbox[2 * (long long)f + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    uint64_t lm = __ballot(leaf);
    if ((threadIdx.x & 63) == 0 && lm) atomicAdd(&nleaf, (int)__popcll(lm));
    block_minmax6(lo, hi, red);
    __syncthreads();
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
    if (threadIdx.x == 0) leaf_part[blockIdx.x] = nleaf;
}

// Per-block partial bounding box of the vertices.
__global__ void __launch_bounds__(kBlock) vert_box_k(const float4* __restrict__ p, int nP, float* __restrict__ part) {
    __shared__ float red[6][kBlock];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const float inf = __builtin_huge_valf();
    float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    if (v < nP) {
        float4 q = p[v];
        lo[0] = hi[0] = q.x;
        lo[1] = hi[1] = q.y;
        lo[2] = hi[2] = q.z;
    }
    block_minmax6(lo, hi, red);
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Whole box from the partials (written after them); with live_part, also
// cnt[C_NLIVE] = the sum of the per-block counts (the guard tree's leaves).
__global__ void __launch_bounds__(kBlock) whole_box_k(float* __restrict__ part, int nparts,
                                                      const int* __restrict__ live_part = nullptr,
                                                      int* __restrict__ cnt = nullptr) {
    __shared__ float red[6][kBlock];
    __shared__ int nlive;
    const float inf = __builtin_huge_valf();
    float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    if (threadIdx.x == 0) nlive = 0;
    __syncthreads();
    int my = 0;
    for (int q = threadIdx.x; q < nparts; q += kBlock) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = fminf(lo[d], part[6 * q + d]);
            hi[d] = fmaxf(hi[d], part[6 * q + 3 + d]);
        }
        if (live_part) my += live_part[q];
    }
    if (live_part && my) atomicAdd(&nlive, my);
    block_minmax6(lo, hi, red);
    __syncthreads();
    if (threadIdx.x < 6) part[6 * nparts + threadIdx.x] = red[threadIdx.x][0];  // whole box after the partials
    if (live_part && threadIdx.x == 0) cnt[C_NLIVE] = nlive;
}

// Start-of-round state in one pass: the undo snapshot of points and faces
// (when snap.p is set), occupancy 1 (0 at nP), zeroed ring counters and
// changed flags (when given), per-vertex minimum keys at the maximum.
__global__ void round_init_k(Mesh m, Mesh snap, int nP, int nT, int* __restrict__ occ, int* __restrict__ ring_off,
                             int* __restrict__ ring_fill, int* __restrict__ order_cnt,
                             unsigned long long* __restrict__ vmin, uint8_t* __restrict__ changed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nP) {
        if (snap.p) snap.p[i] = m.p[i];
        occ[i] = 1;
        vmin[i] = ~0ull;
    }
    if (i == nP) occ[nP] = 0;
    if (i <= nP) {
        ring_off[i] = 0;
        ring_fill[i] = 0;
        order_cnt[i] = 0;
    }
    if (i < nT) {
        if (snap.p) {
            snap.i[i] = m.i[i];
            snap.j[i] = m.j[i];
            snap.k[i] = m.k[i];
        }
        if (changed) changed[i] = 0;
    }
}

// Morton code of each face's box centre in the whole box of the live faces.
// A face removed by this round's collapse but live in the snapshot is still
// a leaf (an undo may restore it): its code is that of its snapshot box,
// with the centre clamped into the whole box, so a restored face sits among
// its neighbours. Faces removed before the round sort behind every leaf.
__global__ void morton_k(Mesh m, Orig snap, int nT, const float4* __restrict__ fbox, const float* __restrict__ whole,
                         uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nT) return;
    val[f] = (uint32_t)f;
    float c[3];
    if (m.i[f] != -1) {
        float4 lo = fbox[2 * (long long)f], hi = fbox[2 * (long long)f + 1];
        c[0] = (hi.x + lo.x) * 0.5f;
        c[1] = (hi.y + lo.y) * 0.5f;
        c[2] = (hi.z + lo.z) * 0.5f;
    } else if (snap.i[f] != -1) {
        F3 p = f3(snap.p[snap.i[f]]), q = f3(snap.p[snap.j[f]]), r = f3(snap.p[snap.k[f]]);
        c[0] = (fmaxf(p.x, fmaxf(q.x, r.x)) + fminf(p.x, fminf(q.x, r.x))) * 0.5f;
        c[1] = (fmaxf(p.y, fmaxf(q.y, r.y)) + fminf(p.y, fminf(q.y, r.y))) * 0.5f;
        c[2] = (fmaxf(p.z, fmaxf(q.z, r.z)) + fminf(p.z, fminf(q.z, r.z))) * 0.5f;
#pragma unroll
        for (int d = 0; d < 3; ++d) c[d] = fminf(fmaxf(c[d], whole[d]), whole[3 + d]);
    } else {
        key[f] = 0xFFFFFFFFu;  // above every code of a leaf (at most 0xE0000000 | 2^30 - 1)
        return;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        c[d] -= whole[d];
        c[d] /= (whole[3 + d] - whole[d]);
    }
    key[f] = morton3(c[0], c[1], c[2]);
}

__device__ __forceinline__ uint64_t node_key(const uint32_t* key, const uint32_t* val, int s) {
    return ((uint64_t)key[s] << 32) | (uint64_t)val[s];
}

__device__ __forceinline__ int delta_at(const uint32_t* key, const uint32_t* val, int n, uint64_t ki, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll(ki ^ node_key(key, val, j));
}

// Karras 2012 radix tree over the sorted keys (code << 32 | face); also
// records each internal node's leaf range for the box pass.
__global__ void karras_k(const uint32_t* __restrict__ key, const uint32_t* __restrict__ val,
                         const int* __restrict__ cnt, int* __restrict__ left, int* __restrict__ right,
                         int* __restrict__ parent, int2* __restrict__ range) {
    const int n = cnt[C_NLIVE];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const uint64_t ki = node_key(key, val, i);
    int first, last;
    if (i == 0) {
        first = 0;
        last = n - 1;
    } else {
        int dl = delta_at(key, val, n, ki, i - 1), dr = delta_at(key, val, n, ki, i + 1);
        int d = dr > dl ? 1 : -1;
        int dmin = min(dl, dr);
        int lmax = 2;
        while (delta_at(key, val, n, ki, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t > 0; t >>= 1)
            if (delta_at(key, val, n, ki, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        first = min(i, j);
        last = max(i, j);
    }
    const uint64_t kf = node_key(key, val, first), kl = node_key(key, val, last);
    int split;
    if (kf == kl) {
        split = (first + last) >> 1;
    } else {
        int dn = __clzll(kf ^ kl);
        split = first;
        int stride = last - first;
        do {
            stride = (stride + 1) >> 1;
            int mid = split + stride;
            if (mid < last && __clzll(kf ^ node_key(key, val, mid)) > dn) split = mid;
        } while (stride > 1);
    }
    const int l = (first == split) ? split + (n - 1) : split;
    const int r = (last == split + 1) ? split + 1 + (n - 1) : split + 1;
    left[i] = l;
    right[i] = r;
    parent[l] = i;
    parent[r] = i;
    range[i] = make_int2(first, last);
}

// Box unions inside aligned blocks of 256 leaves, from a sparse table in
// LDS: level k holds the union of entries [t, t + 2^k), so any range of the
// block is the union of two table entries (min / max are exact and
// idempotent, so overlapping halves give the same box).  One launch writes
// the sorted leaf boxes (level 0: node n-1+s of `node`, lo.w = item id,
// hi.w left as it is), the in-block prefix and suffix boxes, each block's
// box, and (given `range`) the box of every internal node whose leaf range
// lies inside the block: internal node i covers [i, j] or [j, i], so those
// nodes are exactly the ones numbered inside the block.  Level 1 runs the
// same kernel over the block boxes and also stores its table (`tab`, level
// k at tab + 2 k stride) for the nodes that span blocks.
constexpr int kTabLevels = 8;  // 2^0 .. 2^7 entries

struct BoxTab {
    float v[kTabLevels][6][kBlock];
};

__device__ __forceinline__ int tab_level(int len) { return min(31 - __clz(len), kTabLevels - 1); }

// Union of table entries over [l, r] (0 <= l <= r < kBlock) into lo / hi.
__device__ __forceinline__ void tab_union(const BoxTab& T, int l, int r, float* lo, float* hi) {
    const int k = tab_level(r - l + 1), r0 = r - (1 << k) + 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        lo[c] = fminf(T.v[k][c][l], T.v[k][c][r0]);
        hi[c] = fmaxf(T.v[k][3 + c][l], T.v[k][3 + c][r0]);
    }
}

__device__ __forceinline__ void store_xyz(float4* q, const float* v) {
    float* f = reinterpret_cast<float*>(q);
    f[0] = v[0];
    f[1] = v[1];
    f[2] = v[2];
}

__global__ void __launch_bounds__(kBlock) box_tab_k(const float4* __restrict__ src, const uint32_t* __restrict__ gather,
                                                    const int* __restrict__ cnt, int level,
                                                    float4* __restrict__ node, const int2* __restrict__ range,
                                                    float4* __restrict__ pre, float4* __restrict__ suf,
                                                    float4* __restrict__ blk, float4* __restrict__ tab,
                                                    long long tab_stride) {
    __shared__ BoxTab T;
    const int nl = cnt[C_NLIVE];
    const int n = level == 0 ? nl : (nl + kBlock - 1) / kBlock;
    const int base = blockIdx.x * kBlock;
    if (base >= n) return;
    const int t = threadIdx.x, s = base + t;
    const float inf = __builtin_huge_valf();
    float4 a = make_float4(inf, inf, inf, 0.0f), b = make_float4(-inf, -inf, -inf, 0.0f);
    if (s < n) {
        const long long g = gather ? (long long)gather[s] : s;
        a = src[2 * g];
        b = src[2 * g + 1];
        if (node) {
            float4* q = node + 2 * ((long long)nl - 1 + s);
            q[0] = make_float4(a.x, a.y, a.z, __int_as_float((int)g));
            const float hv[3] = {b.x, b.y, b.z};
            store_xyz(q + 1, hv);
        }
    }
    T.v[0][0][t] = a.x; T.v[0][1][t] = a.y; T.v[0][2][t] = a.z;
    T.v[0][3][t] = b.x; T.v[0][4][t] = b.y; T.v[0][5][t] = b.z;
    for (int k = 1; k < kTabLevels; ++k) {
        __syncthreads();
        const int h = 1 << (k - 1);
        const int u = t + h < kBlock ? t + h : t;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T.v[k][c][t] = fminf(T.v[k - 1][c][t], T.v[k - 1][c][u]);
            T.v[k][3 + c][t] = fmaxf(T.v[k - 1][3 + c][t], T.v[k - 1][3 + c][u]);
        }
    }
    __syncthreads();
    float lo[3], hi[3];
    if (s < n) {
        tab_union(T, 0, t, lo, hi);
        pre[2 * (long long)s] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        pre[2 * (long long)s + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        tab_union(T, t, kBlock - 1, lo, hi);
        suf[2 * (long long)s] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        suf[2 * (long long)s + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        if (t == 0) {
            blk[2 * (long long)blockIdx.x] = make_float4(lo[0], lo[1], lo[2], 0.0f);
            blk[2 * (long long)blockIdx.x + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        }
    }
    if (range && s < nl - 1) {
        const int2 r = range[s];
        if ((r.x >> 8) == (r.y >> 8)) {
            tab_union(T, r.x - base, r.y - base, lo, hi);
            store_xyz(node + 2 * (long long)s, lo);
            store_xyz(node + 2 * (long long)s + 1, hi);
        }
    }
    if (tab && s < n) {
        for (int k = 0; k < kTabLevels; ++k) {
            float4* q = tab + 2 * (k * tab_stride + s);
            q[0] = make_float4(T.v[k][0][t], T.v[k][1][t], T.v[k][2][t], 0.0f);
            q[1] = make_float4(T.v[k][3][t], T.v[k][4][t], T.v[k][5][t], 0.0f);
        }
    }
}

__device__ __forceinline__ void grow(float4& lo, float4& hi, const float4* bx, long long q) {
    float4 a = bx[2 * q], b = bx[2 * q + 1];
    lo.x = fminf(lo.x, a.x); lo.y = fminf(lo.y, a.y); lo.z = fminf(lo.z, a.z);
    hi.x = fmaxf(hi.x, b.x); hi.y = fmaxf(hi.y, b.y); hi.z = fmaxf(hi.z, b.z);
}

// Union of boxes bx[base + q] for q in [q0, q1]: eight independent loads per
// step, so a long range costs range/8 memory round trips instead of range
// (min / max are exact and order-free, so the result is the same).
__device__ __forceinline__ void grow_range(float4& lo, float4& hi, const float4* bx, long long base, int q0, int q1) {
    int q = q0;
    for (; q + 8 <= q1 + 1; q += 8) {
        float4 a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a[k] = bx[2 * (base + q + k)];
            b[k] = bx[2 * (base + q + k) + 1];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            lo.x = fminf(lo.x, a[k].x); lo.y = fminf(lo.y, a[k].y); lo.z = fminf(lo.z, a[k].z);
            hi.x = fmaxf(hi.x, b[k].x); hi.y = fmaxf(hi.y, b[k].y); hi.z = fmaxf(hi.z, b[k].z);
        }
    }
    for (; q <= q1; ++q) grow(lo, hi, bx, base + q);
}

// Box of every internal node whose leaf range [a, b] spans blocks (box_tab_k
// has written the others): suffix(a) + prefix(b) + the full blocks between,
// which inside one superblock are two entries of the level-1 table, and
// otherwise block-level suffix / prefix and the full superblocks.  Only the
// xyz fields are written; the .w fields hold the hierarchy's links.
__global__ void __launch_bounds__(kBlock) node_box_k(const int2* __restrict__ range, const int* __restrict__ cnt,
                                                     const float4* __restrict__ pre, const float4* __restrict__ suf,
                                                     const float4* __restrict__ tab, long long tab_stride,
                                                     const float4* __restrict__ bpre,
                                                     const float4* __restrict__ bsuf,
                                                     const float4* __restrict__ sblk, float4* __restrict__ box) {
    const int n = cnt[C_NLIVE];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int2 r = range[i];
    const int A = r.x >> 8, B = r.y >> 8;
    if (A == B) return;
    const float inf = __builtin_huge_valf();
    float4 lo = make_float4(inf, inf, inf, 0.0f), hi = make_float4(-inf, -inf, -inf, 0.0f);
    grow(lo, hi, suf, r.x);
    grow(lo, hi, pre, r.y);
    const int A1 = A + 1, B1 = B - 1;
    if (A1 <= B1) {
        const int SA = A1 >> 8, SB = B1 >> 8;
        if (SA == SB) {
            const int k = tab_level(B1 - A1 + 1);
            grow(lo, hi, tab + 2 * k * tab_stride, A1);
            grow(lo, hi, tab + 2 * k * tab_stride, B1 - (1 << k) + 1);
        } else {
            grow(lo, hi, bsuf, A1);
            grow(lo, hi, bpre, B1);
            grow_range(lo, hi, sblk, 0, SA + 1, SB - 1);
        }
    }
    const float l3[3] = {lo.x, lo.y, lo.z}, h3[3] = {hi.x, hi.y, hi.z};
    store_xyz(box + 2 * (long long)i, l3);
    store_xyz(box + 2 * (long long)i + 1, h3);
}

// ---------------------------------------------------------------- guard (A12, A13)
// Escape ("rope") links: node records become two float4 (lo.xyz, left child
// or, for a leaf, its item id in lo.w; hi.xyz, escape node in hi.w).  The
// escape of x is the next node in depth-first order once x's subtree is
// done: the right sibling of the first ancestor-or-self that is a left
// child, -1 past the end.  Traversal then needs one 32-byte record per step
// and no stack or parent walk.
__global__ void __launch_bounds__(kBlock) rope_k(const int* __restrict__ left, const int2* __restrict__ range,
                                                 const int* __restrict__ cnt, const uint32_t* __restrict__ val,
                                                 float4* __restrict__ node) {
    const int n = cnt[C_NLIVE];
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (n <= 0 || x >= 2 * n - 1) return;
    // x's subtree ends at leaf b; the next node in depth-first order is the
    // right child of the node that splits between leaves b and b+1.  That
    // child is internal node b+1 when that node's range starts at b+1 (an
    // internal node is numbered by the end of its range next to its
    // parent's split, so it is a right child exactly when it starts its
    // range), and otherwise the leaf b+1.
    const bool leaf = x >= n - 1;
    const int b = leaf ? x - (n - 1) : range[x].y;
    int esc = -1;
    if (b < n - 1) {
        const int c = b + 1;
        esc = (c <= n - 2 && range[c].x == c) ? c : c + (n - 1);
    }
    const int lw = leaf ? (int)val[x - (n - 1)] : left[x];
    float* r = reinterpret_cast<float*>(node + 2 * (long long)x);
    r[3] = __int_as_float(lw);
    r[7] = __int_as_float(esc);
}

// Splits a rope-linked tree into the 2^kLog subtrees under the root (fewer
// where a leaf is reached early): lane k of a query walks subtree k, from the
// returned node until *stop (that node's escape).  The right child of a node
// is its left child's escape.  Returns -1 for lanes with no subtree.
template <int kLog>
__device__ __forceinline__ int subtree_start(const float4* __restrict__ node, int n_leaves, int k, int* stop) {
    int cur = 0;
    for (int d = kLog - 1; d >= 0; --d) {
        if (cur >= n_leaves - 1) {
            if ((k & ((2 << d) - 1)) != 0) return -1;
            break;
        }
        const int left = __float_as_int(node[2 * (long long)cur].w);
        cur = ((k >> d) & 1) ? __float_as_int(node[2 * (long long)left + 1].w) : left;
    }
    *stop = __float_as_int(node[2 * (long long)cur + 1].w);
    return cur;
}

// subtree_start with a run-time depth: the root of subtree k of 2^log (k's
// bits choose left / right from the root down), -1 when that subtree does not
// exist because a leaf is reached above depth log (the leaf then belongs to
// the subtree whose remaining bits are zero).  The subtrees partition the
// leaves.
__device__ __forceinline__ int subtree_root(const float4* __restrict__ node, int n_leaves, int k, int log) {
    int cur = 0;
    for (int d = log - 1; d >= 0; --d) {
        if (cur >= n_leaves - 1) return (k & ((2 << d) - 1)) != 0 ? -1 : cur;
        const int left = __float_as_int(node[2 * (long long)cur].w);
        cur = ((k >> d) & 1) ? __float_as_int(node[2 * (long long)left + 1].w) : left;
    }
    return cur;
}

__device__ __forceinline__ bool overlap(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    if (ahi.x < blo.x || bhi.x < alo.x) return false;
    if (ahi.y < blo.y || bhi.y < alo.y) return false;
    if (ahi.z < blo.z || bhi.z < alo.z) return false;
    return true;
}

__device__ __forceinline__ void load_tri9(Mesh m, int f, float T[3][3], int idx[3]) {
    idx[0] = m.i[f]; idx[1] = m.j[f]; idx[2] = m.k[f];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        float4 q = m.p[idx[v]];
        T[v][0] = q.x;
        T[v][1] = q.y;
        T[v][2] = q.z;
    }
}

// Carries the guard tree across a compaction: leaf s keeps its place in the
// hierarchy and takes its face's new index, or nT_new (the empty sentinel
// box) if the face was removed; counts the leaves still live.  Every face
// live at the next round's start is live now, hence already a leaf.
__global__ void carry_leaves_k(uint32_t* __restrict__ val, const int* __restrict__ ti, const int* __restrict__ fpos,
                               int nT_old, int* __restrict__ cnt) {
    __shared__ int nlive;
    const int n = cnt[C_NLIVE], n_new = cnt[C_NF];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) nlive = 0;
    __syncthreads();
    bool live = false;
    if (s < n) {
        const int f = (int)val[s];  // nT_old: already dead
        live = f >= 0 && f < nT_old && ti[f] != -1 && fpos[f] < n_new;
        val[s] = live ? (uint32_t)fpos[f] : (uint32_t)n_new;
    }
    const uint64_t lm = __ballot(live);
    if ((threadIdx.x & 63) == 0 && lm) atomicAdd(&nlive, (int)__popcll(lm));
    __syncthreads();
    if (threadIdx.x == 0 && nlive) atomicAdd(&cnt[C_LIVE_LEAVES], nlive);
}

// Wave-level ordering point for LDS traffic between lanes of one wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Both ordered reference predicates for one candidate pair; flags both faces
// on a hit.
__device__ __forceinline__ void eval_pair(Mesh m, int a, int b, float epsilon, uint8_t* __restrict__ flagged) {
    float A[3][3], B[3][3];
    int ia[3], ib[3];
    load_tri9(m, a, A, ia);
    load_tri9(m, b, B, ib);
    if (ref_guard_d(B, ib, A, ia, epsilon) || ref_guard_d(A, ia, B, ib, epsilon)) {
        flagged[a] = 1;
        flagged[b] = 1;
    }
}

constexpr int kQueue = 128;   // per-wave pair queue (< 64 pending + one append of <= 64)
constexpr int kGuardQ = 8;    // guard queries per wave
constexpr int kGuardStack = 512;       // per-wave frontier entries
constexpr int kGuardStackSmall = 32;   // test variant: forces the overflow walk

// Candidate filter for queried face a (vertex ids ia0..ia2) and a live face b
// whose box overlaps a's (closed intervals): b != a; a pair of queried faces
// is taken from the larger index only; with the reference predicate, pairs
// sharing two vertex indices are dropped before any arithmetic (the
// reference rejects them after its coplanarity test in every outcome).
template <bool kExact = false>
__device__ __forceinline__ bool guard_candidate(Mesh m, const uint8_t* __restrict__ changed, int a, int ia0,
                                                int ia1, int ia2, int b) {
    if (b == a || (changed[b] && b < a)) return false;
    if (kExact) return true;  // the exact test also judges pairs sharing an edge (flat folds)
    const int b0 = m.i[b], b1 = m.j[b], b2 = m.k[b];
    const int sh = (ia0 == b0) + (ia0 == b1) + (ia0 == b2) + (ia1 == b0) + (ia1 == b1) + (ia1 == b2) +
                   (ia2 == b0) + (ia2 == b1) + (ia2 == b2);
    return sh != 2;
}

// Exact mode: where the undecided pairs go.
struct Undecided {
    int2* list;
    int* count;
    int cap;
    int force;  // test setting: every pair goes to the exact pass
};

__device__ __forceinline__ void load_xtri(Mesh m, int f, XP3* p, int* idx) {
    idx[0] = m.i[f]; idx[1] = m.j[f]; idx[2] = m.k[f];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const float4 q = m.p[idx[v]];
        p[v] = XP3{q.x, q.y, q.z};
    }
}

// Exact mode, first pass: the decision with filtered predicates; a pair they
// cannot decide is listed for exact_pairs_k.
__device__ __noinline__ void eval_pair_filtered(Mesh m, int a, int b, uint8_t* __restrict__ flagged, Undecided u) {
    if (!u.force) {
        XP3 pa[3], pb[3];
        int ia[3], ib[3];
        load_xtri(m, a, pa, ia);
        load_xtri(m, b, pb, ib);
        XPred<false> pr;
        const bool hit = x_faces_intersect(pr, ia, pa, ib, pb);
        if (!pr.undecided) {
            if (hit) {
                flagged[a] = 1;
                flagged[b] = 1;
            }
            return;
        }
    }
    const int k = atomicAdd(u.count, 1);
    if (k < u.cap) u.list[k] = make_int2(a, b);
}

template <bool kExact>
__device__ __forceinline__ void guard_eval(Mesh m, int a, int b, float epsilon, uint8_t* __restrict__ flagged,
                                           Undecided u) {
    if (kExact) eval_pair_filtered(m, a, b, flagged, u);
    else eval_pair(m, a, b, epsilon, flagged);
}

// Exact mode, second pass: the listed pairs with exact predicates.
__global__ void __launch_bounds__(kBlock) exact_pairs_k(Mesh m, const int2* __restrict__ list,
                                                        const int* __restrict__ count, int cap,
                                                        uint8_t* __restrict__ flagged) {
    const int n = min(*count, cap);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int2 e = list[k];
        XP3 pa[3], pb[3];
        int ia[3], ib[3];
        load_xtri(m, e.x, pa, ia);
        load_xtri(m, e.y, pb, ib);
        XPred<true> pr;
        if (x_faces_intersect(pr, ia, pa, ib, pb)) {
            flagged[e.x] = 1;
            flagged[e.y] = 1;
        }
    }
}

// Stackless walk of the subtree below internal node x: from its left child
// until x's escape, evaluating candidates in this lane.  Inlined: as a
// called function its register needs add to the caller's (242 registers,
// 2 waves per SIMD, against 129 inlined).
template <bool kExact>
__device__ __forceinline__ void guard_walk_subtree(Mesh m, const uint8_t* __restrict__ changed,
                                                const float4* __restrict__ node, int n, int x, int stop,
                                                float4 qlo, float4 qhi, int a, int ia0, int ia1, int ia2,
                                                float epsilon, uint8_t* __restrict__ flagged, Undecided u,
                                                int* err, long long max_visits) {
    for (long long visits = 0; x != stop; ++visits) {
        if (visits >= max_visits || x < 0) {
            walk_overrun(err);
            return;
        }
        const float4 lo = node[2 * (long long)x], hi = node[2 * (long long)x + 1];
        const int lw = __float_as_int(lo.w), esc = __float_as_int(hi.w);
        if (overlap(qlo, qhi, lo, hi)) {
            if (x >= n - 1) {
                if (guard_candidate<kExact>(m, changed, a, ia0, ia1, ia2, lw))
                    guard_eval<kExact>(m, a, lw, epsilon, flagged, u);
                x = esc;
            } else {
                x = lw;
            }
        } else {
            x = esc;
        }
    }
}

// Guard queries (self_intersect.cuh:151-392 restricted to changed faces):
// every queried face a against every live face b whose box overlaps a's.
//
// A wave serves kGuardQ queries with one shared LIFO frontier of
// (query, node) entries in LDS: each step pops up to 64 entries (one per
// lane), tests their boxes, pushes both children of overlapping internal
// nodes and appends overlapping leaves to a pair queue, which the wave
// evaluates 64 pairs at a time.  The critical path is then the wave's total
// node count / 64 plus the tree depth, instead of the longest single walk.
// When the frontier is full, a lane walks the rest of its subtree by escape
// links (exact, only slower).  The flags are idempotent, so the visiting
// order does not change the result.
template <int kStackCap, bool kExact>
__device__ __forceinline__ void guard_query_body(Mesh m, const int* __restrict__ list, const int* __restrict__ cnt,
                                                 const uint8_t* __restrict__ changed,
                                                 const float4* __restrict__ fbox, const float4* __restrict__ node,
                                                 const int* __restrict__ right, float epsilon,
                                                 uint8_t* __restrict__ flagged, Undecided u, int part_log, int* err,
                                                 long long walk_cap) {
    constexpr int kW = kBlock / kWave;
    __shared__ int2 stk[kW][kStackCap];
    __shared__ int2 pq[kW][kQueue];
    __shared__ float4 qbox[kW][kGuardQ][2];
    __shared__ int4 qid[kW][kGuardQ];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    // wave -> (query group, subtree): the 2^part_log waves of a group each
    // walk one subtree, so one heavy query spreads over several waves
    const long long wid = (long long)blockIdx.x * kW + w;
    const long long q0 = (wid >> part_log) * kGuardQ;
    const int part = (int)(wid & ((1 << part_log) - 1));
    const int nq = cnt[C_NQUERY];
    const int n = cnt[C_NLIVE];
    if (q0 >= nq || n <= 1) return;  // uniform per wave
    const int root = subtree_root(node, n, part, part_log);
    if (root < 0) return;
    const int nl = nq - q0 < kGuardQ ? (int)(nq - q0) : kGuardQ;
    if (lane < nl) {
        const int a = list[q0 + lane];
        qid[w][lane] = make_int4(a, m.i[a], m.j[a], m.k[a]);
        qbox[w][lane][0] = fbox[2 * (long long)a];
        qbox[w][lane][1] = fbox[2 * (long long)a + 1];
        stk[w][lane] = make_int2(lane, root);
    }
    wave_sync();
    int size = nl, qn = 0;  // wave-uniform
    // bounds of a valid tree; walk_cap > 0 lowers them (tests of the error path)
    long long max_steps = (long long)nl * (2LL * n - 1), max_visits = 2LL * n;
    if (walk_cap > 0) {
        max_steps = min(max_steps, walk_cap);
        max_visits = min(max_visits, walk_cap);
    }
    for (long long steps = 0; size > 0; ++steps) {
        if (steps >= max_steps) {  // wave-uniform
            if (lane == 0) walk_overrun(err);
            return;
        }
        const int take = size < kWave ? size : kWave;
        int l = 0, x = -1;
        if (lane < take) {
            const int2 e = stk[w][size - take + lane];
            l = e.x;
            x = e.y;
        }
        size -= take;
        int c0 = -1, c1 = -1, esc = -1, leaf = -1;
        float4 qlo = make_float4(0.f, 0.f, 0.f, 0.f), qhi = qlo;
        if (x >= 0) {
            const float4 lo = node[2 * (long long)x], hi = node[2 * (long long)x + 1];
            const int r = x < n - 1 ? right[x] : -1;
            qlo = qbox[w][l][0];
            qhi = qbox[w][l][1];
            if (overlap(qlo, qhi, lo, hi)) {
                if (x >= n - 1) {
                    leaf = __float_as_int(lo.w);
                } else {
                    c0 = __float_as_int(lo.w);
                    c1 = r;
                    esc = __float_as_int(hi.w);
                }
            }
        }
        wave_sync();  // every pop has read its entry before pushes reuse the slots
        const uint64_t lm = __ballot(leaf >= 0);
        if (leaf >= 0) pq[w][qn + lanes_below(lm)] = make_int2(l, leaf);
        qn += (int)__popcll(lm);
        const uint64_t pm = __ballot(c0 >= 0);
        const int below = lanes_below(pm);
        const bool fits = size + 2 * (below + 1) <= kStackCap;  // pushes that fit form a prefix
        if (c0 >= 0) {
            if (fits) {
                stk[w][size + 2 * below] = make_int2(l, c0);
                stk[w][size + 2 * below + 1] = make_int2(l, c1);
            } else {
                const int4 id = qid[w][l];
                guard_walk_subtree<kExact>(m, changed, node, n, c0, esc, qlo, qhi, id.x, id.y, id.z, id.w, epsilon,
                                           flagged, u, err, max_visits);
            }
        }
        size += 2 * (int)__popcll(__ballot(c0 >= 0 && fits));
        wave_sync();
        if (qn >= kWave) {
            const int2 e = pq[w][qn - kWave + lane];
            const int4 id = qid[w][e.x];
            if (guard_candidate<kExact>(m, changed, id.x, id.y, id.z, id.w, e.y))
                guard_eval<kExact>(m, id.x, e.y, epsilon, flagged, u);
            qn -= kWave;
            wave_sync();
        }
    }
    if (lane < qn) {
        const int2 e = pq[w][lane];
        const int4 id = qid[w][e.x];
        if (guard_candidate<kExact>(m, changed, id.x, id.y, id.z, id.w, e.y))
            guard_eval<kExact>(m, id.x, e.y, epsilon, flagged, u);
    }
}

#define PAMO_GUARD_PARAMS                                                                                  \
    Mesh m, const int* __restrict__ list, const int* __restrict__ cnt, const uint8_t* __restrict__ changed,   \
        const float4* __restrict__ fbox, const float4* __restrict__ node, const int* __restrict__ right,       \
        float epsilon, uint8_t* __restrict__ flagged, Undecided u, int part_log, int* err, long long walk_cap
#define PAMO_GUARD_ARGS m, list, cnt, changed, fbox, node, right, epsilon, flagged, u, part_log, err, walk_cap

template <int kStackCap, bool kExact>
__global__ void __launch_bounds__(kBlock) guard_query_k(PAMO_GUARD_PARAMS) {
    guard_query_body<kStackCap, kExact>(PAMO_GUARD_ARGS);
}

// The same guard with its registers capped for 4 waves per SIMD (the
// compiler's own allocation is one register over); the results do not
// change.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) guard_query_w4_k(PAMO_GUARD_PARAMS) {
    guard_query_body<kGuardStack, false>(PAMO_GUARD_ARGS);
}
#undef PAMO_GUARD_PARAMS
#undef PAMO_GUARD_ARGS

// ---------------------------------------------------------------- undo (A14)
// Restores each pushed collapse and appends its two end vertices, through
// the point map of the previous call (rearrange_undo_list_kernel,
// cusimp_free.cu:869-878), to the undo list at `first`.
__global__ void undo_apply_k(Mesh m, Orig o, int* occ, const int* __restrict__ off, const int* __restrict__ ring,
                             const int* __restrict__ eu, const int* __restrict__ ev, const int* __restrict__ pushed,
                             int npushed, const int* __restrict__ pts_map, int* __restrict__ undo_list, int first) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npushed) return;
    int e = pushed[r];
    int u = eu[e], v = ev[e];
    undo_list[first + 2 * r] = pts_map[u];
    undo_list[first + 2 * r + 1] = pts_map[v];
    restore_edge(m, o, occ, off, ring, u, v);
}

// undo_apply_k with kG lanes per record, splitting the ring faces to
// restore (pushed collapses have disjoint rings).
template <int kG>
__global__ void undo_apply_group_k(Mesh m, Orig o, int* occ, const int* __restrict__ off,
                                   const int* __restrict__ ring, const int* __restrict__ eu,
                                   const int* __restrict__ ev, const int* __restrict__ pushed, int npushed,
                                   const int* __restrict__ pts_map, int* __restrict__ undo_list, int first) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = (int)(threadIdx.x & (kG - 1));
    const long long rr = gt / kG;
    if (rr >= npushed) return;
    const int r = (int)rr;
    const int e = pushed[r];
    const int u = eu[e], v = ev[e];
    if (g == 0) {
        undo_list[first + 2 * r] = pts_map[u];
        undo_list[first + 2 * r + 1] = pts_map[v];
        occ[v] = 1;
        m.p[u] = o.p[u];
        m.p[v] = o.p[v];
    }
    const int ub = off[u], vb = off[v];
    const int nu = off[u + 1] - ub, nf = nu + off[v + 1] - vb;
    for (int i = g; i < nf; i += kG) {
        const int f = ring[i < nu ? ub + i : vb + (i - nu)];
        m.i[f] = o.i[f];
        m.j[f] = o.j[f];
        m.k[f] = o.k[f];
    }
}

// ---------------------------------------------------------------- outputs / compaction (A15)
__global__ void store_points_k(const float4* __restrict__ p, int n, float* __restrict__ aos) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    float4 q = p[v];
    aos[3 * (long long)v] = q.x;
    aos[3 * (long long)v + 1] = q.y;
    aos[3 * (long long)v + 2] = q.z;
}

__global__ void store_tris_k(Mesh m, int n, int* __restrict__ aos) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    aos[3 * (long long)f] = m.i[f];
    aos[3 * (long long)f + 1] = m.j[f];
    aos[3 * (long long)f + 2] = m.k[f];
}

// Centring offset of PaMO.run (pamo/__init__.py:60-72,144): numpy's
// tris.mean(axis=1).mean(axis=0) on the float32 face soup.  The per-face mean
// is ((a + b) + c) / 3 (numpy divides the float32 sum by the count through a
// float64 loop; a float64 quotient rounded to float32 is the correctly
// rounded float32 quotient, since 53 >= 2 * 24 + 2); the mean over faces is
// a sequential float32 running sum from the first face's value, in face
// order, divided the same way.  The face means come from a launch over the
// whole GPU (face_mean_k, rows of `pitch` floats per coordinate).  A float32
// running sum cannot be split, so one workgroup adds them up: waves 1..15
// copy the next chunk of means to LDS while lanes 0..2 of wave 0 add up the
// current chunk, one coordinate per lane, 64 values per LDS batch.
constexpr int kCentThreads = 1024;
constexpr int kCentChunk = 2048;            // multiple of 32
constexpr int kCentPitch = kCentChunk + 4;  // rows start in different LDS banks

__global__ void face_mean_k(const float* __restrict__ pts, const int* __restrict__ tris, int nF, long long pitch,
                            float* __restrict__ means) {
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const int i0 = tris[3 * f], i1 = tris[3 * f + 1], i2 = tris[3 * f + 2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float s = (pts[3 * (long long)i0 + d] + pts[3 * (long long)i1 + d]) + pts[3 * (long long)i2 + d];
        means[d * pitch + f] = s / 3.0f;
    }
}

// means[d * pitch + base ...] -> dst rows; base and pitch are multiples of 4
__device__ inline void centroid_fill(const float* __restrict__ means, long long pitch, long long base, int cnt,
                                     float* __restrict__ dst, int t, int nt) {
    const int n4 = cnt >> 2;
    for (int j = t; j < 3 * n4; j += nt) {
        const int d = j / n4, q = j - d * n4;
        const float4 v = *reinterpret_cast<const float4*>(means + d * pitch + base + 4 * q);
        *reinterpret_cast<float4*>(dst + d * kCentPitch + 4 * q) = v;
    }
    const int tail = cnt - 4 * n4;
    if (t < 3 * tail) {
        const int d = t / tail, q = 4 * n4 + t - d * tail;
        dst[d * kCentPitch + q] = means[d * pitch + base + q];
    }
}

// acc + row[0] + ... + row[n - 1], left to right; row is 16-byte aligned.
// Each batch of 64 values is loaded whole, then added: the adds wait for the
// batch's first load only (a prefetch of the next batch between them made
// the compiler wait on the prefetch before the adds).
__device__ __forceinline__ float centroid_chain(float acc, const float* __restrict__ row, int n) {
    const float4* r4 = reinterpret_cast<const float4*>(row);
    const int full = n >> 6;  // batches of 64
    for (int b = 0; b < full; ++b) {
        float4 v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = r4[16 * b + j];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            acc = acc + v[j].x;
            acc = acc + v[j].y;
            acc = acc + v[j].z;
            acc = acc + v[j].w;
        }
    }
    for (int i = full << 6; i < n; ++i) acc = acc + row[i];
    return acc;
}

__global__ void __launch_bounds__(kCentThreads) centroid_seq_k(const float* __restrict__ means, long long pitch,
                                                               int nF, float* __restrict__ out3) {
    __shared__ __align__(16) float buf[2][3 * kCentPitch];
    const int t = threadIdx.x;
    const int nchunks = (nF + kCentChunk - 1) / kCentChunk;
    centroid_fill(means, pitch, 0, nF < kCentChunk ? nF : kCentChunk, buf[0], t, kCentThreads);
    __syncthreads();
    float acc = 0.0f;
    for (int k = 0; k < nchunks; ++k) {
        const long long base = (long long)k * kCentChunk;
        const int cnt = (int)((nF - base) < kCentChunk ? (nF - base) : kCentChunk);
        if (t >= kWave) {
            const long long nb = base + kCentChunk;
            if (nb < nF) {
                const int ncnt = (int)((nF - nb) < kCentChunk ? (nF - nb) : kCentChunk);
                centroid_fill(means, pitch, nb, ncnt, buf[(k + 1) & 1], t - kWave, kCentThreads - kWave);
            }
        } else if (t < 3) {
            const float* row = buf[k & 1] + t * kCentPitch;
            if (k == 0) {
                // numpy's reduction starts from the first element; the rest
                // of the first chunk from index 1 (unaligned) one at a time
                // up to 32, then in batches
                acc = row[0];
                const int head = cnt < 32 ? cnt : 32;
                for (int i = 1; i < head; ++i) acc = acc + row[i];
                acc = centroid_chain(acc, row + head, cnt - head);
            } else {
                acc = centroid_chain(acc, row, cnt);
            }
        }
        __syncthreads();
    }
    if (t < 3) out3[t] = (float)((double)acc / (double)nF);
}

// ---------------------------------------------------------------- fused compactions
// Operations for scan_op (prim.h): each computes an item's flag, receives
// the item's position among the flagged ones from the same launch's
// look-back scan, and writes the compacted list and its count, where the
// round used a flag kernel, a scan and a compaction kernel.

// Guard queries: the changed live faces among the tree's leaves, in leaf
// order (a carried tree's dead leaves hold the sentinel nT).
struct QueryListOp {
    static constexpr int kIpt = 4;
    const uint32_t* val;
    const uint8_t* changed;
    const int* ti;
    int nT;
    int* list;
    int* cnt;
    __device__ int count(long long s) const {
        const int f = s < cnt[C_NLIVE] ? (int)val[s] : nT;
        return (f < nT && changed[f] && ti[f] != -1) ? 1 : 0;
    }
    __device__ void emit(long long s, int v, int pos) const {
        if (v) list[pos] = (int)val[s];
    }
    __device__ void done(int total) const { cnt[C_NQUERY] = total; }
};

// Regular collapse records in edge order (select_collapse_k's flags).
struct CollapseListOp {
    static constexpr int kIpt = 8;
    const int* flag;
    int* recs;
    int* cnt;
    __device__ int count(long long e) const { return flag[e]; }
    __device__ void emit(long long e, int v, int pos) const {
        if (v) recs[pos] = (int)e;
    }
    __device__ void done(int total) const { cnt[C_NREG] = total; }
};

// Cleanup (cusimp_free.cu:218-274): item t removes face t if two of its
// corners coincide (t < nT) and flags edge t if its ends coincide (t < nE);
// the line collapses go, in edge order, after the regular collapses in the
// records.  Items cover max(nE, nT); the two parts touch disjoint data.
struct CleanupOp {
    static constexpr int kIpt = 2;
    Mesh m;
    const int* eu;
    const int* ev;
    int nE, nT;
    int* recs;
    int* cnt;
    __device__ int count(long long t) const {
        if (t < nT) {
            const int a = m.i[t], b = m.j[t], c = m.k[t];
            if (!face_dead(a, b, c) && (a == b || b == c || c == a)) {
                m.i[t] = -1;
                m.j[t] = -1;
                m.k[t] = -1;
            }
        }
        if (t >= nE) return 0;
        const F3 p = ld(m, eu[t]), q = ld(m, ev[t]);
        return (p.x == q.x && p.y == q.y && p.z == q.z) ? 1 : 0;
    }
    __device__ void emit(long long e, int v, int pos) const {
        if (v) recs[cnt[C_NREG] + pos] = (int)e;
    }
    __device__ void done(int total) const { cnt[C_LINE] = total; }
};

// Undo candidates (get_undo_candidate_kernel, cusimp_free.cu:776-827): the
// records one of whose end vertices touches a flagged face, among the
// first min(records, regular collapses rounded up to 512) records, a count
// taken from the device counters.  One record per thread: each walks two
// vertex rings.
struct UndoListOp {
    static constexpr int kIpt = 1;
    const int* recs;
    const int* off;
    const int* ring;
    const int* eu;
    const int* ev;
    const uint8_t* flagged;
    int* pushed;
    int* cnt;
    __device__ int count(long long r) const {
        // candidates: the first min(records, regular rounded up to 512)
        // records (the reference's grid); the launch covers a bound
        const int nreg = cnt[C_NREG], nrec = nreg + cnt[C_LINE];
        if (r >= min(nrec, (nreg + 511) / 512 * 512)) return 0;
        const int e = recs[r];
        const int u = eu[e], v = ev[e];
        // four ring entries, then their four flags, per step
        for (int side = 0; side < 2; ++side) {
            const int w = side ? v : u;
            const int b = off[w], end = off[w + 1];
            for (int a = b; a < end; a += 4) {
                const int f0 = ring[a];
                const int f1 = a + 1 < end ? ring[a + 1] : f0;
                const int f2 = a + 2 < end ? ring[a + 2] : f0;
                const int f3 = a + 3 < end ? ring[a + 3] : f0;
                if (flagged[f0] | flagged[f1] | flagged[f2] | flagged[f3]) return 1;
            }
        }
        return 0;
    }
    __device__ void emit(long long r, int v, int pos) const {
        if (v) pushed[pos] = recs[r];
    }
    __device__ void done(int total) const { cnt[C_PUSHED] = total; }
};

// Compaction of the mesh (pamo/__init__.py:163-167): item t moves point t
// to pts_map[t] if it is occupied, and face t to its position among the
// live faces with renumbered corners; fpos[t] receives that position for
// t <= nT (fpos[nT] = live faces), which the guard-tree carry reads.  Items
// cover max(nP, nT) + 1.
struct CompactMeshOp {
    static constexpr int kIpt = 2;
    Mesh m;
    int nP, nT;
    const int* occ;
    const int* pts_map;
    int* fpos;
    float4* np;
    int* ni;
    int* nj;
    int* nk;
    int* cnt;
    __device__ int count(long long t) const { return (t < nT && m.i[t] >= 0) ? 1 : 0; }
    __device__ void emit(long long t, int v, int pos) const {
        if (t == 0) cnt[C_LIVE_LEAVES] = 0;  // counted by carry_leaves_k next
        if (t < nP && occ[t]) np[pts_map[t]] = m.p[t];
        if (t == nP) cnt[C_NV] = pts_map[nP];
        if (t <= nT) fpos[t] = pos;
        if (v) {
            ni[pos] = pts_map[m.i[t]];
            nj[pos] = pts_map[m.j[t]];
            nk[pos] = pts_map[m.k[t]];
        }
    }
    __device__ void done(int total) const { cnt[C_NF] = total; }
};

}  // namespace
}  // namespace pamo
