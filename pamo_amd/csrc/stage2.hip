// Stage 2 of PaMO on gfx950: one QEM collapse round with the intersection
// guard and its undo loop, plus the fused multi-round driver.
//
// Reference behaviour followed (file:line in the reference repository):
//   round ............. simp_cuda/src/cusimp_free.cu:891-1048
//   rings / edges ..... :276-335 (ring order made canonical: ascending face)
//   quadrics .......... :337-366
//   edge cost ......... :368-585
//   independent set ... :587-610, :639-681
//   collapse/skinny ... :612-774
//   cleanup ........... :218-274
//   undo .............. :776-889, :1014-1040
//   LBVH + guard ...... simp_cuda/src/bvh/{bvh,morton_code,query,self_intersect}.cuh
//   round loop ........ simp_cuda/pamo/__init__.py:147-185
//
// Layout in HBM: positions and faces are structure-of-arrays (x[], y[], z[];
// i[], j[], k[]) so per-vertex / per-face passes issue coalesced 4-byte
// loads; quadrics are 10 floats per vertex stored component-major; the
// vertex->face CSR is (ring_off[V+1], ring[3F]).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "guard_ref.h"
#include "prim.h"

namespace pamo {

static thread_local std::string g_err;
void set_error(const std::string& msg) {
    if (g_err.empty()) g_err = msg;
}
const char* get_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

constexpr int kBlock = 256;

// ---------------------------------------------------------------- small math
struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 f3add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ float f3dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 f3cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float f3norm(F3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ F3 f3divs(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// ternary min/max (NaN -> upper bound), cusimp_free.cu:29-34
__device__ __forceinline__ float tclamp(float x, float a, float b) {
    float m = a > x ? a : x;
    return m < b ? m : b;
}
__device__ __forceinline__ uint32_t sat_u32(float q) {
    if (!(q < 4294967296.0f)) return 0xFFFFFFFFu;
    if (q <= 0.0f) return 0u;
    return (uint32_t)q;
}

struct Mesh {  // device view of the working mesh (SoA)
    float *x, *y, *z;
    int *i, *j, *k;
};

__device__ __forceinline__ F3 ld(const Mesh& m, int v) { return {m.x[v], m.y[v], m.z[v]}; }
__device__ __forceinline__ void st(const Mesh& m, int v, F3 p) {
    m.x[v] = p.x;
    m.y[v] = p.y;
    m.z[v] = p.z;
}
__device__ __forceinline__ bool face_dead(int a, int b, int c) { return a == -1 && b == -1 && c == -1; }

// ---------------------------------------------------------------- context
struct Feistel {
    uint32_t lbits, rbits, lmask, rmask;
    uint32_t key[24];
};

struct Ctx {
    hipStream_t stream = nullptr;
    int device = 0;
    // logical capacities, grown like the reference (n + n/5)
    size_t cap_pts = 0, cap_tris = 0, cap_edges = 0;
    // physical allocations
    size_t alloc_pts = 0, alloc_tris = 0, alloc_edges = 0, alloc_ws = 0, alloc_parts = 0;
    // point sized (alloc_pts + 1)
    float *px = nullptr, *py = nullptr, *pz = nullptr, *ox = nullptr, *oy = nullptr, *oz = nullptr;
    float *nx = nullptr, *ny = nullptr, *nz = nullptr;  // compacted next-round positions
    int *occ = nullptr, *pts_map = nullptr, *ring_off = nullptr, *ring_fill = nullptr;
    int *undo_list = nullptr, *tmp_undo = nullptr, *undo_prev = nullptr;  // 4 * (alloc_pts + 1)
    uint8_t* invalid = nullptr;
    float* Q = nullptr;  // 10 * alloc_pts
    // face sized (alloc_tris + 1)
    int *ti = nullptr, *tj = nullptr, *tk = nullptr, *oi = nullptr, *oj = nullptr, *ok = nullptr;
    int *ni = nullptr, *nj = nullptr, *nk = nullptr;  // compacted next-round faces / shuffle scratch
    int* ring = nullptr;                              // 3 * alloc_tris
    int* edge_off = nullptr;
    uint64_t* tri_min = nullptr;
    uint8_t *changed = nullptr, *flagged = nullptr;
    int* face_scan = nullptr;
    // BVH (alloc_tris)
    uint32_t *bkey = nullptr, *bval = nullptr, *bkey2 = nullptr, *bval2 = nullptr;
    int *bleft = nullptr, *bright = nullptr, *bparent = nullptr, *bflag = nullptr;  // parent: 2n
    float4* bbox = nullptr;   // 2 float4 per node, 2n nodes
    float4* fbox = nullptr;   // 2 float4 per face
    float* box_part = nullptr;  // per-block partial whole boxes
    // edge sized (alloc_edges + 1)
    int *eu = nullptr, *ev = nullptr;
    uint32_t* edge_cost = nullptr;
    int *esel = nullptr, *recs = nullptr, *pflag = nullptr, *pushed = nullptr;
    // shuffle scratch (2 * alloc_tris + 1)
    int* shuf = nullptr;
    // workspace for scans / sorts
    int* ws = nullptr;
    // counters on device and a pinned mirror
    int* dcnt = nullptr;
    int* hcnt = nullptr;
    // round results visible to the host
    int n_pts = 0, n_tris = 0, n_edges = 0, n_rec = 0, n_regular = 0, n_vertices_undo = 0;
    int undo_iters = 0;
    long long guard_queries = 0;
};

enum Counter {
    C_LINE = 0,      // line-collapse count (scan total)
    C_NREG = 1,      // regular records
    C_NLIVE = 2,     // live faces (BVH leaves)
    C_PUSHED = 3,    // undo candidates
    C_NV = 4,        // next-round vertices
    C_NF = 5,        // next-round faces
    C_NEDGE = 6,
    C_QUERIES = 7,
    C_COUNT = 16
};

template <typename T>
static int dalloc(T** p, size_t n) {
    if (*p) PAMO_HIP(hipFree(*p));
    *p = nullptr;
    PAMO_HIP(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)));
    PAMO_HIP(hipMemset(*p, 0, sizeof(T) * std::max<size_t>(n, 1)));
    return 0;
}

static int ensure_parts(Ctx& c, size_t n_items) {
    size_t need = 6 * (n_items / 256 + 4);
    if (need > c.alloc_parts) {
        c.alloc_parts = need;
        PAMO_TRY(dalloc(&c.box_part, need));
    }
    return 0;
}

static int ensure_ws(Ctx& c, long long n_items) {
    long long need = std::max(radix_workspace(n_items), scan_workspace(n_items + 1)) + 256;
    if ((size_t)need > c.alloc_ws) {
        c.alloc_ws = (size_t)need + need / 4;
        PAMO_TRY(dalloc(&c.ws, c.alloc_ws));
    }
    return 0;
}

// Point-sized buffers.  The logical capacity follows the reference's growth
// rule; persistent buffers (pts_map, undo lists, invalid table) restart at
// zero whenever it grows, as a fresh allocation would.
static int grow_pts(Ctx& c, size_t n) {
    if (n <= c.cap_pts) return 0;
    c.cap_pts = n + n / 5;
    size_t a = c.cap_pts + 1;
    c.alloc_pts = a;
    PAMO_TRY(dalloc(&c.px, a)); PAMO_TRY(dalloc(&c.py, a)); PAMO_TRY(dalloc(&c.pz, a));
    PAMO_TRY(dalloc(&c.ox, a)); PAMO_TRY(dalloc(&c.oy, a)); PAMO_TRY(dalloc(&c.oz, a));
    PAMO_TRY(dalloc(&c.nx, a)); PAMO_TRY(dalloc(&c.ny, a)); PAMO_TRY(dalloc(&c.nz, a));
    PAMO_TRY(dalloc(&c.occ, a + 1));
    PAMO_TRY(dalloc(&c.pts_map, a + 1));
    PAMO_TRY(dalloc(&c.ring_off, a + 2));
    PAMO_TRY(dalloc(&c.ring_fill, a + 1));
    PAMO_TRY(dalloc(&c.undo_list, 4 * a));
    PAMO_TRY(dalloc(&c.tmp_undo, 4 * a));
    PAMO_TRY(dalloc(&c.undo_prev, 4 * a));
    PAMO_TRY(dalloc(&c.invalid, a));
    PAMO_TRY(ensure_parts(c, a));
    PAMO_TRY(dalloc(&c.Q, 10 * a));
    return ensure_ws(c, (long long)a + 1);
}

static int grow_tris(Ctx& c, size_t n) {
    if (n <= c.cap_tris) return 0;
    c.cap_tris = n + n / 5;
    size_t a = c.cap_tris + 1;
    c.alloc_tris = a;
    PAMO_TRY(dalloc(&c.ti, a)); PAMO_TRY(dalloc(&c.tj, a)); PAMO_TRY(dalloc(&c.tk, a));
    PAMO_TRY(dalloc(&c.oi, a)); PAMO_TRY(dalloc(&c.oj, a)); PAMO_TRY(dalloc(&c.ok, a));
    PAMO_TRY(dalloc(&c.ni, a)); PAMO_TRY(dalloc(&c.nj, a)); PAMO_TRY(dalloc(&c.nk, a));
    PAMO_TRY(dalloc(&c.ring, 3 * a));
    PAMO_TRY(dalloc(&c.edge_off, a + 1));
    PAMO_TRY(dalloc(&c.tri_min, a));
    PAMO_TRY(dalloc(&c.changed, a));
    PAMO_TRY(dalloc(&c.flagged, a));
    PAMO_TRY(dalloc(&c.face_scan, a + 1));
    PAMO_TRY(dalloc(&c.bkey, a)); PAMO_TRY(dalloc(&c.bval, a));
    PAMO_TRY(dalloc(&c.bkey2, a)); PAMO_TRY(dalloc(&c.bval2, a));
    PAMO_TRY(dalloc(&c.bleft, a)); PAMO_TRY(dalloc(&c.bright, a));
    PAMO_TRY(dalloc(&c.bparent, 2 * a)); PAMO_TRY(dalloc(&c.bflag, a));
    PAMO_TRY(dalloc(&c.bbox, 4 * a));
    PAMO_TRY(dalloc(&c.fbox, 2 * a));
    PAMO_TRY(ensure_parts(c, a));
    PAMO_TRY(dalloc(&c.shuf, 4 * a + 64));
    // edges: at most 2 per face
    size_t ae = 2 * a + 1;
    if (ae > c.alloc_edges) {
        c.alloc_edges = ae;
        PAMO_TRY(dalloc(&c.eu, ae)); PAMO_TRY(dalloc(&c.ev, ae));
        PAMO_TRY(dalloc(&c.esel, ae + 1)); PAMO_TRY(dalloc(&c.recs, ae + 1));
        PAMO_TRY(dalloc(&c.pflag, ae + 1)); PAMO_TRY(dalloc(&c.pushed, ae + 1));
        PAMO_TRY(dalloc(&c.edge_cost, ae + 1));
    }
    return ensure_ws(c, (long long)(4 * a + 64));
}

// edge_cost keeps stale values between rounds; it restarts at zero when the
// logical edge capacity grows (reference: cusimp_free.cu:180-194).
static int grow_edges(Ctx& c, size_t n) {
    if (n <= c.cap_edges) return 0;
    c.cap_edges = n + n / 5;
    if (c.cap_edges + 1 > c.alloc_edges) {
        c.alloc_edges = c.cap_edges + 1;
        PAMO_TRY(dalloc(&c.eu, c.alloc_edges)); PAMO_TRY(dalloc(&c.ev, c.alloc_edges));
        PAMO_TRY(dalloc(&c.esel, c.alloc_edges + 1)); PAMO_TRY(dalloc(&c.recs, c.alloc_edges + 1));
        PAMO_TRY(dalloc(&c.pflag, c.alloc_edges + 1)); PAMO_TRY(dalloc(&c.pushed, c.alloc_edges + 1));
        PAMO_TRY(dalloc(&c.edge_cost, c.alloc_edges));
    } else {
        PAMO_HIP(hipMemsetAsync(c.edge_cost, 0, sizeof(uint32_t) * c.alloc_edges, c.stream));
    }
    return 0;
}

static void free_all(Ctx& c) {
    void* ptrs[] = {c.px, c.py, c.pz, c.ox, c.oy, c.oz, c.nx, c.ny, c.nz, c.occ, c.pts_map, c.ring_off,
                    c.ring_fill, c.undo_list, c.tmp_undo, c.undo_prev, c.invalid, c.Q, c.ti, c.tj, c.tk, c.oi, c.oj, c.ok,
                    c.ni, c.nj, c.nk, c.ring, c.edge_off, c.tri_min, c.changed, c.flagged, c.face_scan, c.bkey,
                    c.bval, c.bkey2, c.bval2, c.bleft, c.bright, c.bparent, c.bflag, c.bbox, c.fbox,
                    c.box_part, c.shuf, c.eu, c.ev, c.edge_cost, c.esel, c.recs, c.pflag, c.pushed, c.ws,
                    c.dcnt};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c.hcnt) (void)hipHostFree(c.hcnt);
}

// ---------------------------------------------------------------- kernels: load / shuffle
__global__ void load_points_k(const float* __restrict__ aos, Mesh m, float* ox, float* oy, float* oz, int* occ,
                              int n) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    if (v == n) {
        occ[n] = 0;
        return;
    }
    float x = aos[3 * (long long)v], y = aos[3 * (long long)v + 1], z = aos[3 * (long long)v + 2];
    m.x[v] = x; m.y[v] = y; m.z[v] = z;
    ox[v] = x; oy[v] = y; oz[v] = z;
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

// Feistel permutation of [0, 2^bits), flagging images below m (A3).
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
    int p = pos[idx];
    di[p] = si[k];
    dj[p] = sj[k];
    dk[p] = sk[k];
}

static Feistel make_feistel(uint64_t m) {
    Feistel f;
    uint32_t total = 4;
    if (m > 16) {
        total = 0;
        uint64_t t = m - 1;
        while (t) { ++total; t >>= 1; }
    }
    f.lbits = total / 2;
    f.rbits = total - f.lbits;
    f.lmask = (1u << f.lbits) - 1;
    f.rmask = (1u << f.rbits) - 1;
    uint64_t s = 1;
    for (int r = 0; r < 24; ++r) {
        s = (s * 48271ull) % 2147483647ull;
        f.key[r] = (uint32_t)s;
    }
    return f;
}

// ---------------------------------------------------------------- kernels: undo bookkeeping prologue
__global__ void stuck_prologue_k(const int* __restrict__ undo_in, int n, int* tmp_undo, int* undo_list,
                                 const int* __restrict__ pts_map) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = undo_in[i];
    tmp_undo[i] = v;
    undo_list[i] = pts_map[v];
}

__global__ void invalid_set_k(const int* __restrict__ undo_in, int n, uint8_t* invalid, int cap) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = undo_in[i];
    if (v >= 0 && v < cap) invalid[v] = 1;
}

// ---------------------------------------------------------------- kernels: rings and edges (A4, A5)
__global__ void ring_count_k(Mesh m, int nT, int* __restrict__ cnt) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nT) return;
    int a = m.i[f], b = m.j[f], c = m.k[f];
    if (face_dead(a, b, c)) return;
    atomicAdd(&cnt[a], 1);
    atomicAdd(&cnt[b], 1);
    atomicAdd(&cnt[c], 1);
}

__global__ void ring_fill_k(Mesh m, int nT, const int* __restrict__ off, int* __restrict__ fill,
                            int* __restrict__ ring) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nT) return;
    int a = m.i[f], b = m.j[f], c = m.k[f];
    if (face_dead(a, b, c)) return;
    ring[off[a] + atomicAdd(&fill[a], 1)] = f;
    ring[off[b] + atomicAdd(&fill[b], 1)] = f;
    ring[off[c] + atomicAdd(&fill[c], 1)] = f;
}

// Canonical ring order: ascending face index.
__global__ void ring_sort_k(const int* __restrict__ off, int* __restrict__ ring, int nP) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nP) return;
    int b = off[v], e = off[v + 1];
    for (int p = b + 1; p < e; ++p) {
        int x = ring[p];
        int q = p - 1;
        while (q >= b && ring[q] > x) {
            ring[q + 1] = ring[q];
            --q;
        }
        ring[q + 1] = x;
    }
}

__global__ void edge_count_k(Mesh m, int nT, int* __restrict__ cnt) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > nT) return;
    if (f == nT) {
        cnt[nT] = 0;
        return;
    }
    int a = m.i[f], b = m.j[f], c = m.k[f];
    cnt[f] = face_dead(a, b, c) ? 0 : (int)(a > b) + (int)(b > c) + (int)(c > a);
}

__global__ void edge_fill_k(Mesh m, int nT, const int* __restrict__ off, int* __restrict__ eu, int* __restrict__ ev) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nT) return;
    int a = m.i[f], b = m.j[f], c = m.k[f];
    if (face_dead(a, b, c)) return;
    int p = off[f];
    if (a > b) { eu[p] = b; ev[p] = a; ++p; }
    if (b > c) { eu[p] = c; ev[p] = b; ++p; }
    if (c > a) { eu[p] = a; ev[p] = c; ++p; }
}

// ---------------------------------------------------------------- kernels: quadrics and costs (A6, A7)
__global__ void __launch_bounds__(kBlock) quadric_k(Mesh m, const int* __restrict__ off,
                                                    const int* __restrict__ ring, int nP, float* __restrict__ Q,
                                                    long long qs) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nP) return;
    float q[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) q[c] = 0.0f;
    for (int r = off[v]; r < off[v + 1]; ++r) {
        int f = ring[r];
        F3 a = ld(m, m.i[f]), b = ld(m, m.j[f]), d = ld(m, m.k[f]);
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
#pragma unroll
    for (int c = 0; c < 10; ++c) Q[c * qs + v] = q[c];
}

__device__ __forceinline__ int succ_in(int a, int b, int c, int x) {
    if (a == x) return b;
    if (b == x) return c;
    if (c == x) return a;
    return -2;
}

__device__ __forceinline__ float heron_d(F3 a, F3 b, F3 d) {
    float x = f3norm(f3sub(a, b)), y = f3norm(f3sub(b, d)), z = f3norm(f3sub(d, a));
    float s = (x + y + z) / 2;
    return sqrtf(s * (s - x) * (s - y) * (s - z));
}

struct CostParams {
    float edge_s, tres;
    int is_stuck;
};

__global__ void __launch_bounds__(kBlock) edge_cost_k(Mesh m, const int* __restrict__ off,
                                                      const int* __restrict__ ring, const int* __restrict__ eu,
                                                      const int* __restrict__ ev, int nE,
                                                      const uint8_t* __restrict__ invalid,
                                                      const float* __restrict__ Q, long long qs, CostParams cp,
                                                      uint32_t* __restrict__ cost) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int u = eu[e], v = ev[e];
    if (invalid[u] && invalid[v]) {
        cost[e] = 0xFFFFFFFFu;
        return;
    }
    if (cp.is_stuck) return;
    const int ub = off[u], ue = off[u + 1], vb = off[v], ve = off[v + 1];
    // link condition (|N(u) ∩ N(v)| via successors), :405-444
    int dup = 0;
    for (int a = ub; a < ue; ++a) {
        int fa = ring[a];
        int su = succ_in(m.i[fa], m.j[fa], m.k[fa], u);
        for (int b = vb; b < ve; ++b) {
            int fb = ring[b];
            int sv = succ_in(m.i[fb], m.j[fb], m.k[fb], v);
            dup += (su == sv) ? 1 : 0;
            if (dup > 2) {
                cost[e] = 0xFFFFFFFFu;
                return;
            }
        }
    }
    if (dup != 2) return;
    const F3 p0 = ld(m, u), p1 = ld(m, v);
    float elen = 0.0f;
    int nedge = 0;
    for (int side = 0; side < 2; ++side) {
        int b0 = side ? vb : ub, b1 = side ? ve : ue;
        for (int a = b0; a < b1; ++a) {
            int f = ring[a];
            int ii = m.i[f], jj = m.j[f], kk = m.k[f];
            if (ii > jj) { elen += f3norm(f3sub(ld(m, ii), ld(m, jj))); ++nedge; }
            if (jj > kk) { elen += f3norm(f3sub(ld(m, jj), ld(m, kk))); ++nedge; }
            if (kk > ii) { elen += f3norm(f3sub(ld(m, kk), ld(m, ii))); ++nedge; }
        }
    }
    elen = elen / (float)nedge / cp.edge_s * cp.tres;
    elen = elen + f3norm(f3sub(p0, p1)) / cp.edge_s * cp.tres;
    const F3 vm = f3divs(f3add(p0, p1), 2.0f);
    const float k4s3 = 4.0f * sqrtf(3.0f);
    float qa = 0.0f;
    int ntri = 0;
    for (int side = 0; side < 2; ++side) {
        const int w = side ? v : u, other = side ? u : v;
        const int b0 = side ? vb : ub, b1 = side ? ve : ue;
        for (int a = b0; a < b1; ++a) {
            int f = ring[a];
            int ii = m.i[f], jj = m.j[f], kk = m.k[f];
            if (ii == other || jj == other || kk == other) continue;
            F3 o0 = ld(m, ii), o1 = ld(m, jj), o2 = ld(m, kk);
            F3 n0 = ii == w ? vm : o0, n1 = jj == w ? vm : o1, n2 = kk == w ? vm : o2;
            F3 on = f3cross(f3sub(o1, o0), f3sub(o2, o0));
            on = f3divs(on, f3norm(on));
            F3 nn = f3cross(f3sub(n1, n0), f3sub(n2, n0));
            nn = f3divs(nn, f3norm(nn));
            if (f3dot(on, nn) < 0) {
                cost[e] = 0xFFFFFFFFu;
                return;
            }
            float l01 = f3norm(f3sub(n0, n1)), l12 = f3norm(f3sub(n1, n2)), l20 = f3norm(f3sub(n2, n0));
            float qv = k4s3 * heron_d(n0, n1, n2) / (l01 * l01) + l12 * l12 + l20 * l20 + 0.0000001f;
            qa += 1.0f - tclamp(qv, 0.0f, 1.0f);
            ++ntri;
        }
    }
    qa = 5.0f * qa;
    // M = Q_u + Q_v (symmetric, 10 stored entries); r_j = sum_i M[i][j] x_i
    float M[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) M[c] = Q[c * qs + u] + Q[c * qs + v];
    const float x0 = vm.x, x1 = vm.y, x2 = vm.z, x3 = 1.0f;
    // storage: 0:00 1:01 2:02 3:03 4:11 5:12 6:13 7:22 8:23 9:33
    float r0 = M[0] * x0 + M[1] * x1 + M[2] * x2 + M[3] * x3;
    float r1 = M[1] * x0 + M[4] * x1 + M[5] * x2 + M[6] * x3;
    float r2 = M[2] * x0 + M[5] * x1 + M[7] * x2 + M[8] * x3;
    float r3 = M[3] * x0 + M[6] * x1 + M[8] * x2 + M[9] * x3;
    float qc = (r0 * x0 + r1 * x1 + r2 * x2 + r3 * x3) / (cp.edge_s * cp.edge_s);
    float tot = qc + elen + qa / (float)ntri * cp.tres;
    cost[e] = sat_u32(tclamp(tot, 0.0f, 10.0f) / 10.0f * 4294967296.0f);
}

// ---------------------------------------------------------------- kernels: independent set + collapse (A8, A9)
__device__ __forceinline__ uint64_t edge_key(uint32_t c, int e) { return ((uint64_t)c << 32) | (uint32_t)e; }

__global__ void __launch_bounds__(kBlock) propagate_k(const int* __restrict__ off, const int* __restrict__ ring,
                                                      const int* __restrict__ eu, const int* __restrict__ ev, int nE,
                                                      const uint32_t* __restrict__ cost,
                                                      unsigned long long* __restrict__ tri_min) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const unsigned long long k = edge_key(cost[e], e);
    const int u = eu[e], v = ev[e];
    for (int a = off[u]; a < off[u + 1]; ++a) atomicMin(&tri_min[ring[a]], k);
    for (int a = off[v]; a < off[v + 1]; ++a) atomicMin(&tri_min[ring[a]], k);
}

__device__ __forceinline__ bool tri_has(int a, int b, int c, int x) { return a == x || b == x || c == x; }

// Face rewrite shared by the regular and the line collapse.
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

__device__ __forceinline__ bool skinny(F3 p, F3 q, F3 r) {
    F3 pq = f3sub(q, p), pr = f3sub(r, p), qr = f3sub(r, q);
    float lpq = sqrtf(f3dot(pq, pq)), lpr = sqrtf(f3dot(pr, pr)), lqr = sqrtf(f3dot(qr, qr));
    float ml = fmaxf(lpq, fmaxf(lpr, lqr));
    F3 cr = {pq.y * pr.z - pq.z * pr.y, pq.z * pr.x - pq.x * pr.z, pq.x * pr.y - pq.y * pr.x};
    float thr = (float)(1e-6 * (double)ml * (double)ml);
    float area = 0.5f * sqrtf(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
    return area < thr;
}

__device__ __forceinline__ void restore_edge(Mesh m, const float* ox, const float* oy, const float* oz,
                                             const int* oi, const int* oj, const int* ok, int* occ,
                                             const int* __restrict__ off, const int* __restrict__ ring, int u,
                                             int v) {
    occ[v] = 1;
    m.x[u] = ox[u]; m.y[u] = oy[u]; m.z[u] = oz[u];
    m.x[v] = ox[v]; m.y[v] = oy[v]; m.z[v] = oz[v];
    for (int side = 0; side < 2; ++side) {
        int w = side ? v : u;
        for (int a = off[w]; a < off[w + 1]; ++a) {
            int f = ring[a];
            m.i[f] = oi[f]; m.j[f] = oj[f]; m.k[f] = ok[f];
        }
    }
}

// Selection (:639-681) fused with the collapse (:683-773): selected edges
// have pairwise disjoint one-ring face sets, and no selected edge moves a
// vertex of another selected edge's rings, so collapsing in the same pass
// is race free and order independent.
__global__ void __launch_bounds__(kBlock) select_collapse_k(Mesh m, const float* ox, const float* oy,
                                                            const float* oz, const int* oi, const int* oj,
                                                            const int* ok, int* occ, const int* __restrict__ off,
                                                            const int* __restrict__ ring,
                                                            const int* __restrict__ eu, const int* __restrict__ ev,
                                                            int nE, const uint32_t* __restrict__ cost,
                                                            const unsigned long long* __restrict__ tri_min,
                                                            uint32_t collapse_t, int* __restrict__ sel) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > nE) return;
    if (e == nE) {
        sel[nE] = 0;
        return;
    }
    const uint32_t c = cost[e];
    int s = 0;
    if (c <= collapse_t) {
        const unsigned long long k = edge_key(c, e);
        const int u = eu[e], v = ev[e];
        bool ok_ = true;
        for (int a = off[u]; a < off[u + 1] && ok_; ++a) ok_ = tri_min[ring[a]] == k;
        for (int a = off[v]; a < off[v + 1] && ok_; ++a) ok_ = tri_min[ring[a]] == k;
        if (ok_) {
            s = 1;
            F3 vm = f3divs(f3add(ld(m, u), ld(m, v)), 2.0f);
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
            if (sk) restore_edge(m, ox, oy, oz, oi, oj, ok, occ, off, ring, u, v);
        }
    }
    sel[e] = s;
}

__global__ void compact_k(const int* __restrict__ flag, const int* __restrict__ pos, int n, int* __restrict__ out,
                          int base) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    if (flag[e]) out[base + pos[e]] = e;
}

// ---------------------------------------------------------------- kernels: cleanup (A10)
__global__ void remove_invalid_k(Mesh m, int nT) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nT) return;
    int a = m.i[f], b = m.j[f], c = m.k[f];
    if (face_dead(a, b, c)) return;
    if (a == b || b == c || c == a) { m.i[f] = -1; m.j[f] = -1; m.k[f] = -1; }
}

__global__ void line_flag_k(Mesh m, const int* __restrict__ eu, const int* __restrict__ ev, int nE,
                            int* __restrict__ flag) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > nE) return;
    if (e == nE) {
        flag[nE] = 0;
        return;
    }
    F3 p = ld(m, eu[e]), q = ld(m, ev[e]);
    flag[e] = (p.x == q.x && p.y == q.y && p.z == q.z) ? 1 : 0;
}

// Line collapses are rare; applying them in edge order in one lane keeps
// the result deterministic where the reference races on shared faces.
__global__ void line_apply_k(Mesh m, int* occ, const int* __restrict__ off, const int* __restrict__ ring,
                             const int* __restrict__ eu, const int* __restrict__ ev, const int* __restrict__ flag,
                             const int* __restrict__ pos, int nE, int* __restrict__ recs, int* cnt) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int base = cnt[C_NREG];
    const int nl = pos[nE];
    cnt[C_LINE] = nl;
    if (nl == 0) return;
    int w = 0;
    for (int e = 0; e < nE && w < nl; ++e) {
        if (!flag[e]) continue;
        int u = eu[e], v = ev[e];
        recs[base + w++] = e;
        st(m, v, F3{0.0f, 0.0f, 0.0f});
        occ[v] = 0;
        rewrite_rings(m, off, ring, u, v);
    }
}

__global__ void store_count_k(const int* __restrict__ pos, int n, int* cnt, int slot) { cnt[slot] = pos[n]; }

__global__ void mark_changed_k(const int* __restrict__ recs, const int* __restrict__ cnt,
                               const int* __restrict__ off, const int* __restrict__ ring,
                               const int* __restrict__ eu, const int* __restrict__ ev,
                               uint8_t* __restrict__ changed) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= cnt[C_NREG] + cnt[C_LINE]) return;
    int e = recs[r];
    int u = eu[e], v = ev[e];
    for (int a = off[u]; a < off[u + 1]; ++a) changed[ring[a]] = 1;
    for (int a = off[v]; a < off[v + 1]; ++a) changed[ring[a]] = 1;
}

// ---------------------------------------------------------------- kernels: LBVH (A11)
__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton3(float x, float y, float z) {
    bool sx = x < 0.0f, sy = y < 0.0f, sz = z < 0.0f;
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    uint32_t xx = expand_bits((uint32_t)x), yy = expand_bits((uint32_t)y), zz = expand_bits((uint32_t)z);
    return (sx ? 0x80000000u : 0u) | (sy ? 0x40000000u : 0u) | (sz ? 0x20000000u : 0u) | (xx * 4 + yy * 2 + zz);
}

// Face boxes + per-block partial whole box + live count.
__global__ void __launch_bounds__(kBlock) face_box_k(Mesh m, int nT, float4* __restrict__ fbox,
                                                     float* __restrict__ part, int* cnt) {
    __shared__ float red[6][kBlock];
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const float inf = __builtin_huge_valf();
    float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    bool live = false;
    if (f < nT) {
        int a = m.i[f];
        if (a != -1) {
            live = true;
            F3 p = ld(m, a), q = ld(m, m.j[f]), r = ld(m, m.k[f]);
            lo[0] = fminf(p.x, fminf(q.x, r.x));
            lo[1] = fminf(p.y, fminf(q.y, r.y));
            lo[2] = fminf(p.z, fminf(q.z, r.z));
            hi[0] = fmaxf(p.x, fmaxf(q.x, r.x));
            hi[1] = fmaxf(p.y, fmaxf(q.y, r.y));
            hi[2] = fmaxf(p.z, fmaxf(q.z, r.z));
            fbox[2 * f] = make_float4(lo[0], lo[1], lo[2], 0.0f);
            fbox[2 * f + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        }
    }
    uint64_t lm = __ballot(live);
    if ((threadIdx.x & 63) == 0 && lm) atomicAdd(&cnt[C_NLIVE], (int)__popcll(lm));
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fminf(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmaxf(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Per-block partial bounding box of the vertices.
__global__ void __launch_bounds__(kBlock) vert_box_k(Mesh m, int nP, float* __restrict__ part) {
    __shared__ float red[6][kBlock];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const float inf = __builtin_huge_valf();
    float a[6] = {inf, inf, inf, -inf, -inf, -inf};
    if (v < nP) {
        a[0] = a[3] = m.x[v];
        a[1] = a[4] = m.y[v];
        a[2] = a[5] = m.z[v];
    }
    for (int d = 0; d < 6; ++d) red[d][threadIdx.x] = a[d];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fminf(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmaxf(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(kBlock) whole_box_k(float* __restrict__ part, int nparts) {
    __shared__ float red[6][kBlock];
    const float inf = __builtin_huge_valf();
    float acc[6] = {inf, inf, inf, -inf, -inf, -inf};
    for (int p = threadIdx.x; p < nparts; p += kBlock)
        for (int d = 0; d < 3; ++d) {
            acc[d] = fminf(acc[d], part[6 * p + d]);
            acc[3 + d] = fmaxf(acc[3 + d], part[6 * p + 3 + d]);
        }
    for (int d = 0; d < 6; ++d) red[d][threadIdx.x] = acc[d];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fminf(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmaxf(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    // the whole box lands after all partials
    if (threadIdx.x < 6) part[6 * nparts + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void morton_k(Mesh m, int nT, const float4* __restrict__ fbox, const float* __restrict__ whole,
                         uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nT) return;
    val[f] = (uint32_t)f;
    if (m.i[f] == -1) {
        key[f] = 0xFFFFFFFFu;  // removed faces sort behind every live code (< 2^30)
        return;
    }
    float4 lo = fbox[2 * f], hi = fbox[2 * f + 1];
    float c[3] = {(hi.x + lo.x) * 0.5f, (hi.y + lo.y) * 0.5f, (hi.z + lo.z) * 0.5f};
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

// Karras 2012 radix tree over the sorted 64-bit keys (code << 32 | face).
__global__ void karras_k(const uint32_t* __restrict__ key, const uint32_t* __restrict__ val,
                         const int* __restrict__ cnt, int* __restrict__ left, int* __restrict__ right,
                         int* __restrict__ parent, int* __restrict__ flag) {
    const int n = cnt[C_NLIVE];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && n > 0) parent[0] = -1;
    if (i >= n - 1) return;
    flag[i] = 0;
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
    int l = (first == split) ? split + (n - 1) : split;
    int r = (last == split + 1) ? split + 1 + (n - 1) : split + 1;
    left[i] = l;
    right[i] = r;
    parent[l] = i;
    parent[r] = i;
}

// Leaf boxes, then bottom-up merge: the second child to arrive at a node
// merges (first-arrival flag by atomic), as in bvh.cuh:427-459.
__global__ void bvh_boxes_k(const uint32_t* __restrict__ val, const float4* __restrict__ fbox,
                            const int* __restrict__ cnt, const int* __restrict__ left,
                            const int* __restrict__ right, const int* __restrict__ parent, int* __restrict__ flag,
                            float4* __restrict__ box) {
    const int n = cnt[C_NLIVE];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int node = n - 1 + s;
    const int f = (int)val[s];
    float4 lo = fbox[2 * f], hi = fbox[2 * f + 1];
    box[2 * node] = lo;
    box[2 * node + 1] = hi;
    if (n == 1) return;
    int p = parent[node];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&flag[p], 1) == 0) return;
        __threadfence();
        int l = left[p], r = right[p];
        // the agent-scope fences above order these loads after the sibling's stores
        float4 al = box[2 * l], ah = box[2 * l + 1];
        float4 bl = box[2 * r], bh = box[2 * r + 1];
        float4 ml = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
        float4 mh = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
        box[2 * p] = ml;
        box[2 * p + 1] = mh;
        p = parent[p];
    }
}

// ---------------------------------------------------------------- kernels: guard query (A12, A13)
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
        T[v][0] = m.x[idx[v]];
        T[v][1] = m.y[idx[v]];
        T[v][2] = m.z[idx[v]];
    }
}

// Every changed live face queries the tree; a pair of changed faces is
// evaluated from the larger index only.  Both ordered predicates are
// evaluated (the reference visits each unordered pair from both sides).
__global__ void __launch_bounds__(kBlock) guard_query_k(Mesh m, int nT, const uint8_t* __restrict__ changed,
                                                        const float4* __restrict__ fbox,
                                                        const float4* __restrict__ box,
                                                        const int* __restrict__ left, const int* __restrict__ right,
                                                        const uint32_t* __restrict__ val, const int* __restrict__ cnt,
                                                        float epsilon, uint8_t* __restrict__ flagged) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nT) return;
    if (!changed[a] || m.i[a] == -1) return;
    const int n = cnt[C_NLIVE];
    if (n <= 1) return;
    float A[3][3];
    int ia[3];
    load_tri9(m, a, A, ia);
    const float4 qlo = fbox[2 * a], qhi = fbox[2 * a + 1];
    bool hit_any = false;
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int node = stack[--sp];
        const int ch[2] = {left[node], right[node]};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int x = ch[c];
            if (!overlap(qlo, qhi, box[2 * x], box[2 * x + 1])) continue;
            if (x >= n - 1) {
                const int b = (int)val[x - (n - 1)];
                if (b == a) continue;
                if (changed[b] && b < a) continue;
                float B[3][3];
                int ib[3];
                load_tri9(m, b, B, ib);
                bool hit = ref_guard_d(B, ib, A, ia, epsilon) || ref_guard_d(A, ia, B, ib, epsilon);
                if (hit) {
                    hit_any = true;
                    flagged[b] = 1;
                }
            } else if (sp < 64) {
                stack[sp++] = x;
            }
        }
    }
    if (hit_any) flagged[a] = 1;
}

// ---------------------------------------------------------------- kernels: undo (A14)
__global__ void undo_cand_k(const int* __restrict__ recs, int n_rec, const int* __restrict__ off,
                            const int* __restrict__ ring, const int* __restrict__ eu, const int* __restrict__ ev,
                            const uint8_t* __restrict__ flagged, int* __restrict__ pflag) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > n_rec) return;
    if (r == n_rec) {
        pflag[n_rec] = 0;
        return;
    }
    int e = recs[r];
    int u = eu[e], v = ev[e];
    int hit = 0;
    for (int a = off[u]; a < off[u + 1] && !hit; ++a) hit = flagged[ring[a]];
    for (int a = off[v]; a < off[v + 1] && !hit; ++a) hit = flagged[ring[a]];
    pflag[r] = hit ? 1 : 0;
}

__global__ void undo_compact_k(const int* __restrict__ recs, int n_rec, const int* __restrict__ pflag,
                               const int* __restrict__ pos, int* __restrict__ pushed, int* cnt) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) cnt[C_PUSHED] = pos[n_rec];
    if (r >= n_rec) return;
    if (pflag[r]) pushed[pos[r]] = recs[r];
}

__global__ void undo_apply_k(Mesh m, const float* ox, const float* oy, const float* oz, const int* oi,
                             const int* oj, const int* ok, int* occ, const int* __restrict__ off,
                             const int* __restrict__ ring, const int* __restrict__ eu, const int* __restrict__ ev,
                             const int* __restrict__ pushed, int npushed, int* __restrict__ tmp_undo) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npushed) return;
    int e = pushed[r];
    int u = eu[e], v = ev[e];
    tmp_undo[2 * r] = u;
    tmp_undo[2 * r + 1] = v;
    restore_edge(m, ox, oy, oz, oi, oj, ok, occ, off, ring, u, v);
}

__global__ void rearrange_k(const int* __restrict__ tmp_undo, int n, const int* __restrict__ pts_map,
                            int* __restrict__ undo_list, int first) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    undo_list[first + i] = pts_map[tmp_undo[i]];
}

// ---------------------------------------------------------------- kernels: outputs / compaction (A15)
__global__ void store_points_k(Mesh m, int n, float* __restrict__ aos) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    aos[3 * (long long)v] = m.x[v];
    aos[3 * (long long)v + 1] = m.y[v];
    aos[3 * (long long)v + 2] = m.z[v];
}

__global__ void store_tris_k(Mesh m, int n, int* __restrict__ aos) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    aos[3 * (long long)f] = m.i[f];
    aos[3 * (long long)f + 1] = m.j[f];
    aos[3 * (long long)f + 2] = m.k[f];
}

__global__ void live_flag_k(Mesh m, int nT, int* __restrict__ flag) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > nT) return;
    flag[f] = (f < nT && m.i[f] >= 0) ? 1 : 0;
}

__global__ void compact_mesh_k(Mesh m, int nP, int nT, const int* __restrict__ occ,
                               const int* __restrict__ pts_map, const int* __restrict__ fpos, float* nx, float* ny,
                               float* nz, int* ni, int* nj, int* nk, int* cnt) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        cnt[C_NV] = pts_map[nP];
        cnt[C_NF] = fpos[nT];
    }
    if (t < nP && occ[t]) {
        int d = pts_map[t];
        nx[d] = m.x[t];
        ny[d] = m.y[t];
        nz[d] = m.z[t];
    }
    if (t < nT && m.i[t] >= 0) {
        int d = fpos[t];
        ni[d] = pts_map[m.i[t]];
        nj[d] = pts_map[m.j[t]];
        nk[d] = pts_map[m.k[t]];
    }
}

// Mean over faces of the float32 face centroid ((a + b) + c) / 3, summed in
// FP64 per block in a fixed tree (pamo/__init__.py:60-72 centring).
__global__ void __launch_bounds__(kBlock) centroid_part_k(const float* __restrict__ pts, const int* __restrict__ tris,
                                                          int nF, double* __restrict__ part) {
    __shared__ double red[3][kBlock];
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    double a[3] = {0.0, 0.0, 0.0};
    if (f < nF) {
        const int i0 = tris[3 * (long long)f], i1 = tris[3 * (long long)f + 1], i2 = tris[3 * (long long)f + 2];
        for (int d = 0; d < 3; ++d) {
            float s = (pts[3 * (long long)i0 + d] + pts[3 * (long long)i1 + d]) + pts[3 * (long long)i2 + d];
            a[d] = (double)(s / 3.0f);
        }
    }
    for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = a[d];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------- host orchestration
static Mesh cur_mesh(Ctx& c) { return Mesh{c.px, c.py, c.pz, c.ti, c.tj, c.tk}; }

#define LAUNCH(kern, n, ...)                                                                              \
    do {                                                                                                  \
        hipLaunchKernelGGL(kern, dim3(grid_for((long long)(n), kBlock)), dim3(kBlock), 0, c.stream,        \
                           __VA_ARGS__);                                                                  \
        PAMO_HIP(hipGetLastError());                                                                      \
    } while (0)

static int sync_counters(Ctx& c) {
    PAMO_HIP(hipMemcpyAsync(c.hcnt, c.dcnt, sizeof(int) * C_COUNT, hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

// Build the LBVH over live faces and flag faces of intersecting pairs that
// involve a changed face (self_intersect.cuh:151-392 restricted to the
// faces that can change an undo decision).
static int run_guard(Ctx& c, int nT) {
    Mesh m = cur_mesh(c);
    PAMO_HIP(hipMemsetAsync(c.flagged, 0, nT, c.stream));
    PAMO_HIP(hipMemsetAsync(c.dcnt + C_NLIVE, 0, sizeof(int), c.stream));
    const int nparts = (int)grid_for(nT, kBlock);
    LAUNCH(face_box_k, nT, m, nT, c.fbox, c.box_part, c.dcnt);
    hipLaunchKernelGGL(whole_box_k, dim3(1), dim3(kBlock), 0, c.stream, c.box_part, nparts);
    PAMO_HIP(hipGetLastError());
    LAUNCH(morton_k, nT, m, nT, (const float4*)c.fbox, (const float*)(c.box_part + 6 * nparts), c.bkey, c.bval);
    PAMO_TRY(radix_sort_pairs(c.bkey, c.bval, c.bkey2, c.bval2, nT, 32, c.ws, c.stream));
    LAUNCH(karras_k, nT, c.bkey, c.bval, c.dcnt, c.bleft, c.bright, c.bparent, c.bflag);
    LAUNCH(bvh_boxes_k, nT, c.bval, c.fbox, c.dcnt, c.bleft, c.bright, c.bparent, c.bflag, c.bbox);
    LAUNCH(guard_query_k, nT, m, nT, c.changed, c.fbox, c.bbox, c.bleft, c.bright, c.bval, c.dcnt, 1e-3f,
           c.flagged);
    return 0;
}

static int undo_candidates(Ctx& c, int& npushed) {
    LAUNCH(undo_cand_k, c.n_rec + 1, c.recs, c.n_rec, c.ring_off, c.ring, c.eu, c.ev, c.flagged, c.pflag);
    PAMO_TRY(scan_exclusive(c.pflag, c.esel, c.n_rec + 1, c.ws, c.stream));
    LAUNCH(undo_compact_k, std::max(c.n_rec, 1), c.recs, c.n_rec, c.pflag, c.esel, c.pushed, c.dcnt);
    PAMO_TRY(sync_counters(c));
    npushed = c.hcnt[C_PUSHED];
    return 0;
}

// One round on the mesh held in (px.., ti..) with nP points and nT faces
// (cusimp_free.cu:891-1048).  undo_in is a device list of n_undo_in ints.
static int round_device(Ctx& c, int nP, int nT, const int* undo_in, int n_undo_in, float scale, float thr,
                        bool is_stuck, bool init_shuffle_done) {
    (void)init_shuffle_done;
    const float tres = thr, edge_s = scale;
    const uint32_t collapse_t = [&] {
        float x = thr > 0.0f ? thr : 0.0f;  // tclamp(thr, 0, 10)
        if (!(x < 10.0f)) x = 10.0f;
        float q = x / 10.0f * 4294967296.0f;
        if (!(q < 4294967296.0f)) return 0xFFFFFFFFu;
        return (uint32_t)q;
    }();
    Mesh m = cur_mesh(c);
    c.n_pts = nP;
    c.n_tris = nT;
    // snapshot for undo
    PAMO_HIP(hipMemcpyAsync(c.ox, c.px, sizeof(float) * nP, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.oy, c.py, sizeof(float) * nP, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.oz, c.pz, sizeof(float) * nP, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.oi, c.ti, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.oj, c.tj, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.ok, c.tk, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    PAMO_TRY(fill(c.occ, 1, nP, c.stream));
    PAMO_TRY(fill(c.occ + nP, 0, 1, c.stream));
    // vertex -> face rings
    PAMO_HIP(hipMemsetAsync(c.ring_off, 0, sizeof(int) * (nP + 1), c.stream));
    PAMO_HIP(hipMemsetAsync(c.ring_fill, 0, sizeof(int) * (nP + 1), c.stream));
    LAUNCH(ring_count_k, nT, m, nT, c.ring_off);
    PAMO_TRY(scan_exclusive(c.ring_off, c.ring_off, nP + 1, c.ws, c.stream));
    LAUNCH(ring_fill_k, nT, m, nT, c.ring_off, c.ring_fill, c.ring);
    LAUNCH(ring_sort_k, nP, c.ring_off, c.ring, nP);
    // edges
    LAUNCH(edge_count_k, nT + 1, m, nT, c.edge_off);
    PAMO_TRY(scan_exclusive(c.edge_off, c.edge_off, nT + 1, c.ws, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.dcnt + C_NEDGE, c.edge_off + nT, sizeof(int), hipMemcpyDeviceToDevice, c.stream));
    PAMO_TRY(sync_counters(c));
    const int nE = c.hcnt[C_NEDGE];
    c.n_edges = nE;
    PAMO_TRY(grow_edges(c, (size_t)nE));
    LAUNCH(edge_fill_k, nT, m, nT, c.edge_off, c.eu, c.ev);
    // quadrics + cost
    const long long qs = (long long)c.alloc_pts;
    LAUNCH(quadric_k, nP, m, c.ring_off, c.ring, nP, c.Q, qs);
    CostParams cp{edge_s, tres, is_stuck ? 1 : 0};
    LAUNCH(edge_cost_k, nE, m, c.ring_off, c.ring, c.eu, c.ev, nE, c.invalid, c.Q, qs, cp, c.edge_cost);
    // independent set + collapse
    PAMO_TRY(fill<unsigned long long>((unsigned long long*)c.tri_min, ~0ull, nT, c.stream));
    LAUNCH(propagate_k, nE, c.ring_off, c.ring, c.eu, c.ev, nE, c.edge_cost, (unsigned long long*)c.tri_min);
    LAUNCH(select_collapse_k, nE + 1, m, c.ox, c.oy, c.oz, c.oi, c.oj, c.ok, c.occ, c.ring_off, c.ring, c.eu,
           c.ev, nE, c.edge_cost, (const unsigned long long*)c.tri_min, collapse_t, c.esel);
    PAMO_TRY(scan_exclusive(c.esel, c.pflag, nE + 1, c.ws, c.stream));
    LAUNCH(compact_k, nE, c.esel, c.pflag, nE, c.recs, 0);
    hipLaunchKernelGGL(store_count_k, dim3(1), dim3(1), 0, c.stream, c.pflag, nE, c.dcnt, (int)C_NREG);
    PAMO_HIP(hipGetLastError());
    // cleanup
    LAUNCH(remove_invalid_k, nT, m, nT);
    LAUNCH(line_flag_k, nE + 1, m, c.eu, c.ev, nE, c.esel);
    PAMO_TRY(scan_exclusive(c.esel, c.pflag, nE + 1, c.ws, c.stream));
    hipLaunchKernelGGL(line_apply_k, dim3(1), dim3(1), 0, c.stream, m, c.occ, c.ring_off, c.ring, c.eu, c.ev,
                       c.esel, c.pflag, nE, c.recs, c.dcnt);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipMemsetAsync(c.changed, 0, nT, c.stream));
    LAUNCH(mark_changed_k, nE, c.recs, c.dcnt, c.ring_off, c.ring, c.eu, c.ev, c.changed);
    PAMO_TRY(sync_counters(c));
    c.n_regular = c.hcnt[C_NREG];
    c.n_rec = c.n_regular + c.hcnt[C_LINE];
    // guard + undo loop
    int first = c.n_vertices_undo;  // entries already placed by the prologue
    int npushed = 0;
    c.undo_iters = 0;
    if (c.n_rec > 0) {
        PAMO_TRY(run_guard(c, nT));
        ++c.guard_queries;
        PAMO_TRY(undo_candidates(c, npushed));
    }
    c.n_vertices_undo += 2 * npushed;
    int it = 0;
    while (npushed != 0) {
        ++it;
        ++c.undo_iters;
        LAUNCH(undo_apply_k, npushed, m, c.ox, c.oy, c.oz, c.oi, c.oj, c.ok, c.occ, c.ring_off, c.ring, c.eu, c.ev,
               c.pushed, npushed, c.tmp_undo);
        LAUNCH(rearrange_k, c.n_vertices_undo - first, c.tmp_undo, c.n_vertices_undo - first, c.pts_map,
               c.undo_list, first);
        first += 2 * npushed;
        PAMO_TRY(run_guard(c, nT));
        ++c.guard_queries;
        PAMO_TRY(undo_candidates(c, npushed));
        c.n_vertices_undo += 2 * npushed;
        if (it == 5) break;
    }
    // new vertex ids
    PAMO_TRY(scan_exclusive(c.occ, c.pts_map, nP + 1, c.ws, c.stream));
    return 0;
}

// Undo-list prologue (cusimp_free.cu:895-916).
static int prologue(Ctx& c, const int* undo_in, int n_undo_in, bool is_stuck) {
    c.n_vertices_undo = 0;
    if (is_stuck) {
        if (n_undo_in > 0) LAUNCH(stuck_prologue_k, n_undo_in, undo_in, n_undo_in, c.tmp_undo, c.undo_list, c.pts_map);
        c.n_vertices_undo += n_undo_in;
    } else if (n_undo_in != 0) {
        PAMO_HIP(hipMemsetAsync(c.invalid, 0, c.alloc_pts, c.stream));
    }
    if (n_undo_in != 0) LAUNCH(invalid_set_k, n_undo_in, undo_in, n_undo_in, c.invalid, (int)c.alloc_pts);
    return 0;
}

// Drop removed vertices and faces and renumber (pamo/__init__.py:163-167):
// stable compaction by pts_map and by a scan of live faces.
static int compact_round(Ctx& c, int nP, int nT) {
    Mesh m = cur_mesh(c);
    LAUNCH(live_flag_k, nT + 1, m, nT, c.face_scan);
    PAMO_TRY(scan_exclusive(c.face_scan, c.face_scan, nT + 1, c.ws, c.stream));
    LAUNCH(compact_mesh_k, std::max(nP, nT), m, nP, nT, c.occ, c.pts_map, c.face_scan, c.nx, c.ny, c.nz, c.ni, c.nj,
           c.nk, c.dcnt);
    PAMO_TRY(sync_counters(c));
    std::swap(c.px, c.nx); std::swap(c.py, c.ny); std::swap(c.pz, c.nz);
    std::swap(c.ti, c.ni); std::swap(c.tj, c.nj); std::swap(c.tk, c.nk);
    return 0;
}

static int shuffle_faces(Ctx& c, int nT) {
    Feistel fs = make_feistel((uint64_t)nT);
    const uint32_t domain = 1u << (fs.lbits + fs.rbits);
    int* flag = c.shuf;
    uint32_t* img = (uint32_t*)(c.shuf + domain + 1);
    PAMO_CHECK((size_t)(2 * domain + 2) <= 4 * c.alloc_tris + 64, "shuffle scratch too small");
    LAUNCH(shuffle_flags_k, domain + 1, fs, domain, (uint32_t)nT, flag, img);
    PAMO_TRY(scan_exclusive(flag, flag, domain + 1, c.ws, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.ni, c.ti, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.nj, c.tj, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.nk, c.tk, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    LAUNCH(shuffle_scatter_k, domain, domain, (uint32_t)nT, flag, img, c.ni, c.nj, c.nk, c.ti, c.tj, c.tk);
    return 0;
}

}  // namespace pamo

// ================================================================ C ABI
#include "../../include/pamo_hip.h"

using pamo::Ctx;

extern "C" {

const char* pamo_last_error(void) { return pamo::get_error(); }

int pamo_ctx_create(pamo_ctx** out, int device) {
    pamo::clear_error();
    PAMO_CHECK(out != nullptr, "pamo_ctx_create: out is null");
    Ctx* c = new Ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete c;
        pamo::set_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
        return 1;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&c->dcnt, sizeof(int) * pamo::C_COUNT);
    if (e == hipSuccess) e = hipMemset(c->dcnt, 0, sizeof(int) * pamo::C_COUNT);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->hcnt, sizeof(int) * pamo::C_COUNT, 0);
    if (e != hipSuccess) {
        pamo::set_error(std::string("pamo_ctx_create: ") + hipGetErrorString(e));
        pamo::free_all(*c);
        delete c;
        return 1;
    }
    *out = reinterpret_cast<pamo_ctx*>(c);
    return 0;
}

int pamo_ctx_destroy(pamo_ctx* ctx) {
    if (!ctx) return 0;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pamo::free_all(*c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int pamo_round(pamo_ctx* ctx, const float* d_pts, int n_pts, const int* d_tris, int n_tris, const int* d_undo,
               int n_undo, float scale, float threshold, int is_stuck, int init, float* d_out_pts, int* d_out_tris,
               int* d_out_occ, int* d_out_map, int* d_out_undo, int undo_cap, int* n_undo_out) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_round: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    PAMO_CHECK(n_pts > 0 && n_tris > 0, "pamo_round: empty mesh");
    PAMO_CHECK(n_undo >= 0 && (n_undo == 0 || d_undo != nullptr), "pamo_round: bad undo list");
    PAMO_HIP(hipSetDevice(c.device));
    // the prologue reads pts_map / invalid from the previous call before any growth
    if (c.alloc_pts > 0) {
        PAMO_CHECK(n_undo <= (int)(4 * c.alloc_pts), "pamo_round: undo list longer than the state can hold");
        PAMO_TRY(pamo::prologue(c, d_undo, n_undo, is_stuck != 0));
    } else {
        PAMO_CHECK(n_undo == 0, "pamo_round: first call must not pass an undo list");
        c.n_vertices_undo = 0;
    }
    PAMO_TRY(pamo::grow_pts(c, (size_t)n_pts));
    PAMO_TRY(pamo::grow_tris(c, (size_t)n_tris));
    pamo::Mesh m = pamo::cur_mesh(c);
    hipLaunchKernelGGL(pamo::load_points_k, dim3(pamo::grid_for(n_pts + 1, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, d_pts, m, c.ox, c.oy, c.oz, c.occ, n_pts);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(pamo::load_tris_k, dim3(pamo::grid_for(n_tris, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, d_tris, c.ti, c.tj, c.tk, n_tris);
    PAMO_HIP(hipGetLastError());
    if (init) PAMO_TRY(pamo::shuffle_faces(c, n_tris));
    PAMO_TRY(pamo::round_device(c, n_pts, n_tris, d_undo, n_undo, scale, threshold, is_stuck != 0, init != 0));
    PAMO_CHECK(c.n_vertices_undo <= undo_cap, "pamo_round: undo_cap too small");
    hipLaunchKernelGGL(pamo::store_points_k, dim3(pamo::grid_for(n_pts, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, m, n_pts, d_out_pts);
    hipLaunchKernelGGL(pamo::store_tris_k, dim3(pamo::grid_for(n_tris, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, m, n_tris, d_out_tris);
    PAMO_HIP(hipMemcpyAsync(d_out_occ, c.occ, sizeof(int) * n_pts, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(d_out_map, c.pts_map, sizeof(int) * n_pts, hipMemcpyDeviceToDevice, c.stream));
    if (c.n_vertices_undo > 0)
        PAMO_HIP(hipMemcpyAsync(d_out_undo, c.undo_list, sizeof(int) * c.n_vertices_undo, hipMemcpyDeviceToDevice,
                                c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    if (n_undo_out) *n_undo_out = c.n_vertices_undo;
    return 0;
}

int pamo_centroid(pamo_ctx* ctx, const float* d_pts, int n_pts, const int* d_tris, int n_tris, float* out3) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr && out3 != nullptr, "pamo_centroid: null argument");
    PAMO_CHECK(n_pts > 0 && n_tris > 0, "pamo_centroid: empty mesh");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    PAMO_HIP(hipSetDevice(c.device));
    const int nb = (int)pamo::grid_for(n_tris, pamo::kBlock);
    double* part = nullptr;
    PAMO_HIP(hipMallocAsync((void**)&part, sizeof(double) * 3 * nb, c.stream));
    hipLaunchKernelGGL(pamo::centroid_part_k, dim3(nb), dim3(pamo::kBlock), 0, c.stream, d_pts, d_tris, n_tris, part);
    PAMO_HIP(hipGetLastError());
    std::vector<double> h(3 * (size_t)nb);
    PAMO_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipFreeAsync(part, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    for (int d = 0; d < 3; ++d) {
        double acc = 0.0;
        for (int b = 0; b < nb; ++b) acc += h[3 * (size_t)b + d];
        out3[d] = (float)(acc / (double)n_tris);
    }
    return 0;
}

int pamo_round_debug(pamo_ctx* ctx, int what, void* d_out, long long cap, long long* n_out) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_round_debug: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    long long n = 0;
    const void* src = nullptr;
    size_t elt = 4;
    switch (what) {
        case PAMO_DBG_EDGES: {  // 2 ints per edge (u, v): interleave on host side
            n = c.n_edges;
            PAMO_CHECK(2 * n <= cap, "pamo_round_debug: cap too small");
            PAMO_HIP(hipMemcpy2DAsync(d_out, 8, c.eu, 4, 4, (size_t)n, hipMemcpyDeviceToDevice, c.stream));
            PAMO_HIP(hipMemcpy2DAsync((char*)d_out + 4, 8, c.ev, 4, 4, (size_t)n, hipMemcpyDeviceToDevice, c.stream));
            PAMO_HIP(hipStreamSynchronize(c.stream));
            if (n_out) *n_out = n;
            return 0;
        }
        case PAMO_DBG_EDGE_COST: n = c.n_edges; src = c.edge_cost; break;
        case PAMO_DBG_RECORDS: n = c.n_rec; src = c.recs; break;
        case PAMO_DBG_RING_OFF: n = c.n_pts + 1; src = c.ring_off; break;
        case PAMO_DBG_RING: {
            int tot = 0;
            PAMO_HIP(hipMemcpy(&tot, c.ring_off + c.n_pts, sizeof(int), hipMemcpyDeviceToHost));
            n = tot;
            src = c.ring;
            break;
        }
        case PAMO_DBG_QUADRICS: {
            n = 10LL * c.n_pts;
            PAMO_CHECK(n <= cap, "pamo_round_debug: cap too small");
            for (int q = 0; q < 10; ++q)
                PAMO_HIP(hipMemcpyAsync((float*)d_out + (long long)q * c.n_pts, c.Q + (long long)q * c.alloc_pts,
                                        sizeof(float) * c.n_pts, hipMemcpyDeviceToDevice, c.stream));
            PAMO_HIP(hipStreamSynchronize(c.stream));
            if (n_out) *n_out = n;
            return 0;
        }
        case PAMO_DBG_UNDO_ITERS: if (n_out) *n_out = c.undo_iters; return 0;
        case PAMO_DBG_N_REGULAR: if (n_out) *n_out = c.n_regular; return 0;
        default: PAMO_CHECK(false, "pamo_round_debug: unknown selector");
    }
    PAMO_CHECK(n <= cap, "pamo_round_debug: cap too small");
    if (n > 0) PAMO_HIP(hipMemcpyAsync(d_out, src, elt * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    if (n_out) *n_out = n;
    return 0;
}

int pamo_simplify(pamo_ctx* ctx, const float* d_verts, int n_verts, const int* d_faces, int n_faces, double ratio,
                  int tolerance, float threshold, long long iters, long long min_verts, float* d_out_verts,
                  int* d_out_faces, int* n_verts_out, int* n_faces_out, int* face_hist, int hist_cap,
                  int* rounds_out) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_simplify: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    PAMO_CHECK(n_verts > 0 && n_faces > 0, "pamo_simplify: empty mesh");
    PAMO_CHECK(d_verts && d_faces && d_out_verts && d_out_faces, "pamo_simplify: null buffer");
    PAMO_HIP(hipSetDevice(c.device));
    long long target = (long long)(ratio * (double)n_faces);
    if (min_verts > target) target = min_verts;
    PAMO_TRY(pamo::grow_pts(c, (size_t)n_verts));
    PAMO_TRY(pamo::grow_tris(c, (size_t)n_faces));
    pamo::Mesh m = pamo::cur_mesh(c);
    hipLaunchKernelGGL(pamo::load_points_k, dim3(pamo::grid_for(n_verts + 1, pamo::kBlock)), dim3(pamo::kBlock),
                       0, c.stream, d_verts, m, c.ox, c.oy, c.oz, c.occ, n_verts);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(pamo::load_tris_k, dim3(pamo::grid_for(n_faces, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, d_faces, c.ti, c.tj, c.tk, n_faces);
    PAMO_HIP(hipGetLastError());
    // scale = largest bounding-box extent of the input (pamo/__init__.py:153)
    float scale;
    {
        const int nparts = (int)pamo::grid_for(n_verts, pamo::kBlock);
        hipLaunchKernelGGL(pamo::vert_box_k, dim3(nparts), dim3(pamo::kBlock), 0, c.stream, m, n_verts, c.box_part);
        PAMO_HIP(hipGetLastError());
        hipLaunchKernelGGL(pamo::whole_box_k, dim3(1), dim3(pamo::kBlock), 0, c.stream, c.box_part, nparts);
        PAMO_HIP(hipGetLastError());
        float hb[6];
        PAMO_HIP(hipMemcpyAsync(hb, c.box_part + 6 * nparts, sizeof(hb), hipMemcpyDeviceToHost, c.stream));
        PAMO_HIP(hipStreamSynchronize(c.stream));
        float ex = hb[3] - hb[0], ey = hb[4] - hb[1], ez = hb[5] - hb[2];
        scale = ex;
        if (ey > scale) scale = ey;
        if (ez > scale) scale = ez;
    }
    int nV = n_verts, nF = n_faces;
    int n_undo = 0, count = 0, rounds = 0;
    bool is_stuck = false;
    for (long long it = 0; it < iters; ++it) {
        const int prev = nF;
        if (rounds == 0) {
            c.n_vertices_undo = 0;
            PAMO_TRY(pamo::shuffle_faces(c, nF));
        } else {
            PAMO_TRY(pamo::prologue(c, c.undo_prev, n_undo, is_stuck));
        }
        PAMO_TRY(pamo::round_device(c, nV, nF, c.undo_prev, n_undo, scale, threshold, is_stuck, rounds == 0));
        ++rounds;
        n_undo = c.n_vertices_undo;
        if (n_undo > 0)
            PAMO_HIP(hipMemcpyAsync(c.undo_prev, c.undo_list, sizeof(int) * n_undo, hipMemcpyDeviceToDevice,
                                    c.stream));
        PAMO_TRY(pamo::compact_round(c, nV, nF));
        nV = c.hcnt[pamo::C_NV];
        nF = c.hcnt[pamo::C_NF];
        if (face_hist && rounds <= hist_cap) face_hist[rounds - 1] = nF;
        if (nF <= target || nF <= 10) break;
        if (nF == prev) ++count;
        else { count = 0; is_stuck = false; }
        if (count >= 2) is_stuck = true;
        if (count == tolerance) break;
    }
    m = pamo::cur_mesh(c);
    hipLaunchKernelGGL(pamo::store_points_k, dim3(pamo::grid_for(nV, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, m, nV, d_out_verts);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(pamo::store_tris_k, dim3(pamo::grid_for(nF, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, m, nF, d_out_faces);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipStreamSynchronize(c.stream));
    if (n_verts_out) *n_verts_out = nV;
    if (n_faces_out) *n_faces_out = nF;
    if (rounds_out) *rounds_out = rounds;
    return 0;
}

}  // extern "C"
