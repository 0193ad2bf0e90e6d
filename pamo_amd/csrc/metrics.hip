// Surface metrics on gfx950 behind pamo_metrics_* of include/pamo_hip.h:
// point-to-mesh squared distances and the two-sided Hausdorff distance
// between triangle meshes, with the semantics of the libigl calls the
// reference makes (simp_cuda/safe_project/src/pamo_safe_project/metrics.py:31-51):
//   igl.point_mesh_squared_distance(P, V, F) -> (sqrD, I, C)
//   igl.hausdorff(VA, FA, VB, FB) = sqrt(max(max_{a in VA} d(a, B)^2,
//                                            max_{b in VB} d(b, A)^2))
// (libigl 2.5.1 hausdorff: vertex-to-surface both ways).
//
// Distances are FP64.  A point's squared distance to a triangle is the one of
// Ericson's closest point (Real-Time Collision Detection 5.1.5, the region
// test libigl's point_simplex_squared_distance uses), evaluated in double
// without contraction.  The nearest triangle is found by branch and bound
// over an LBVH of float boxes rounded outward from the double corners: a
// subtree is skipped only when its box is farther than the best distance so
// far, so the result is the minimum over all triangles up to rounding at
// exact ties, with ties broken by the smaller face index.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "pamo_hip.h"
#include "lbvh_tree.h"

namespace pamo {
namespace metrics {

using s3::kB;
using s3::Tree;
using s3::TreeView;

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return D3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 load3(const double* __restrict__ p, long long i) {
    return D3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}

// Ericson 5.1.5: closest point of triangle abc to p by Voronoi regions.
__device__ inline D3 closest_on_tri_d(D3 p, D3 a, D3 b, D3 c) {
    const D3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const D3 bp = p - b;
    const double d3 = ddot(ab, bp), d4 = ddot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return a + (d1 / (d1 - d3)) * ab;
    const D3 cp = p - c;
    const double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + (d2 / (d2 - d6)) * ac;
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
        return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    const double den = 1.0 / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    return a + v * ab + w * ac;
}

// Largest float <= x and smallest float >= x.
__device__ __forceinline__ float f_down(double x) {
    float f = (float)x;
    return (double)f > x ? nextafterf(f, -__builtin_huge_valf()) : f;
}
__device__ __forceinline__ float f_up(double x) {
    float f = (float)x;
    return (double)f < x ? nextafterf(f, __builtin_huge_valf()) : f;
}

// Number of faces with a corner outside [0, nV).
__global__ void __launch_bounds__(kB) check_faces_k(const int* __restrict__ F, int nF, int nV, int* bad) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const int a = F[3 * (long long)f], b = F[3 * (long long)f + 1], c = F[3 * (long long)f + 2];
    if (a < 0 || a >= nV || b < 0 || b >= nV || c < 0 || c >= nV) atomicAdd(bad, 1);
}

// Float box of each face, rounded outward from its double corners.
__global__ void __launch_bounds__(kB) face_box_d_k(const double* __restrict__ V, const int* __restrict__ F, int nF,
                                                  float4* __restrict__ item) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const D3 a = load3(V, F[3 * (long long)f]), b = load3(V, F[3 * (long long)f + 1]),
             c = load3(V, F[3 * (long long)f + 2]);
    item[2 * (long long)f] = make_float4(f_down(fmin(fmin(a.x, b.x), c.x)), f_down(fmin(fmin(a.y, b.y), c.y)),
                                         f_down(fmin(fmin(a.z, b.z), c.z)), 0.0f);
    item[2 * (long long)f + 1] = make_float4(f_up(fmax(fmax(a.x, b.x), c.x)), f_up(fmax(fmax(a.y, b.y), c.y)),
                                             f_up(fmax(fmax(a.z, b.z), c.z)), 0.0f);
}

__device__ __forceinline__ double box_d2(D3 p, float4 lo, float4 hi) {
    const double dx = fmax(fmax((double)lo.x - p.x, 0.0), p.x - (double)hi.x);
    const double dy = fmax(fmax((double)lo.y - p.y, 0.0), p.y - (double)hi.y);
    const double dz = fmax(fmax((double)lo.z - p.z, 0.0), p.z - (double)hi.z);
    return dx * dx + dy * dy + dz * dz;
}

// Nearest face of each point: ordered stackless descent (nearer child
// first, parent links to climb), subtrees farther than the best skipped.
// Writes sqd[i] and, when given, I[i] and the closest point C[i].
__global__ void __launch_bounds__(kB) point_mesh_k(const double* __restrict__ P, long long nP, TreeView T,
                                                  const double* __restrict__ V, const int* __restrict__ F,
                                                  double* __restrict__ sqd, int* __restrict__ I,
                                                  double* __restrict__ C, int* err) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nP) return;
    const D3 p = load3(P, i);
    double best = __builtin_huge_val();
    int best_f = -1;
    D3 best_c = D3{0.0, 0.0, 0.0};
    const int n = T.n;
    auto leaf = [&](int f) {
        const D3 c = closest_on_tri_d(p, load3(V, F[3 * (long long)f]), load3(V, F[3 * (long long)f + 1]),
                                      load3(V, F[3 * (long long)f + 2]));
        const D3 d = c - p;
        const double d2 = ddot(d, d);
        if (d2 < best || (d2 == best && f < best_f)) {
            best = d2;
            best_f = f;
            best_c = c;
        }
    };
    auto bd = [&](int node) { return box_d2(p, T.box[2 * (long long)node], T.box[2 * (long long)node + 1]); };
    if (n == 1) {
        leaf((int)T.val[0]);
    } else if (n > 1) {
        int cur = 0, prev = -1;
        const long long max_steps = 3LL * (2LL * n - 1) + 1;
        for (long long steps = 0; cur >= 0; ++steps) {
            if (steps >= max_steps || cur > 2 * n - 2) {
                walk_overrun(err);
                break;
            }
            const int par = cur == 0 ? -1 : T.parent[cur];
            int next;
            if (prev == par) {
                if (bd(cur) > best) {
                    next = par;
                } else if (cur >= n - 1) {
                    leaf((int)T.val[cur - (n - 1)]);
                    next = par;
                } else {
                    const int l = T.left[cur], r = T.right[cur];
                    next = bd(l) <= bd(r) ? l : r;
                }
            } else {
                const int l = T.left[cur], r = T.right[cur];
                const bool left_first = bd(l) <= bd(r);
                const int first = left_first ? l : r, second = left_first ? r : l;
                next = prev == first ? second : par;
            }
            prev = cur;
            cur = next;
        }
    }
    sqd[i] = best;
    if (I) I[i] = best_f;
    if (C) {
        C[3 * i] = best_c.x;
        C[3 * i + 1] = best_c.y;
        C[3 * i + 2] = best_c.z;
    }
}

// Maximum of non-negative doubles (their bit patterns order like the values;
// a NaN orders above every number and propagates).
__global__ void __launch_bounds__(kB) max_bits_k(const double* __restrict__ x, long long n,
                                                unsigned long long* out) {
    __shared__ unsigned long long red[kB];
    unsigned long long m = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = max(m, (unsigned long long)__double_as_longlong(x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

// Device memory of one call: the tree over the mesh's faces, the sort and
// scan workspace, and two flags (faces indexing outside the vertices; a tree
// walk that exceeded its bound).  A metrics context owns one that grows with
// the mesh; pamo_point_mesh_sqdist_noalloc carves one out of a static array.
struct Work {
    Tree tree;
    int* ws = nullptr;
    int* flags = nullptr;  // [2] device
    int* h_flags = nullptr;  // [2] host
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Work w;
    long long cap_ws = 0;
    double* sqd = nullptr;  // Hausdorff: per-vertex squared distances
    long long cap_sqd = 0;
    unsigned long long* mx = nullptr;   // [1]
    unsigned long long* h_mx = nullptr; // pinned
};

static void ctx_free(Ctx& c) {
    s3::tree_free(c.w.tree);
    void* ptrs[] = {c.w.ws, c.sqd, c.w.flags, c.mx};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (c.w.h_flags) (void)hipHostFree(c.w.h_flags);
    if (c.h_mx) (void)hipHostFree(c.h_mx);
    if (c.stream) (void)hipStreamDestroy(c.stream);
}

// Checks every face index, then builds the tree over the faces of (V, F);
// w's tree and workspace must hold nF faces.
static int build_tree(Work& w, hipStream_t st, const double* V, int nV, const int* F, int nF, const char* who) {
    PAMO_CHECK(V && F && nV > 0 && nF > 0, std::string(who) + ": empty or null mesh");
    PAMO_TRY(fill(w.flags, 0, 1, st));
    hipLaunchKernelGGL(check_faces_k, dim3(grid_for(nF, kB)), dim3(kB), 0, st, F, nF, nV, w.flags);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipMemcpyAsync(w.h_flags, w.flags, sizeof(int), hipMemcpyDeviceToHost, st));
    PAMO_HIP(hipStreamSynchronize(st));
    PAMO_CHECK(w.h_flags[0] == 0,
               std::string(who) + ": " + std::to_string(w.h_flags[0]) + " faces index outside the vertices");
    hipLaunchKernelGGL(face_box_d_k, dim3(grid_for(nF, kB)), dim3(kB), 0, st, V, F, nF, w.tree.item);
    PAMO_HIP(hipGetLastError());
    return s3::tree_build(w.tree, nF, w.ws, st);
}

static int query(Work& w, hipStream_t st, const double* P, long long nP, const double* V, const int* F, double* sqd,
                 int* I, double* C) {
    if (nP <= 0) return 0;
    hipLaunchKernelGGL(point_mesh_k, dim3(grid_for(nP, kB)), dim3(kB), 0, st, P, nP, s3::view(w.tree), V, F, sqd, I,
                       C, w.flags + 1);
    PAMO_HIP(hipGetLastError());
    return 0;
}

// Clears the walk flag at the start of a call.
static int begin(Work& w, hipStream_t st) { return fill(w.flags + 1, 0, 1, st); }

// Reads the walk flag after the stream's work and reports an overrun.
static int finish(Work& w, hipStream_t st, const char* who) {
    PAMO_HIP(hipMemcpyAsync(w.h_flags + 1, w.flags + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    PAMO_HIP(hipStreamSynchronize(st));
    PAMO_CHECK(w.h_flags[1] == 0, std::string(who) + ": a tree walk exceeded its bound (corrupted LBVH)");
    return 0;
}

// Sizes the context's tree and workspace for nF faces.
static int reserve(Ctx& c, int nF) {
    const long long need = std::max(radix_workspace(nF + 64LL), scan_workspace(nF + 65LL)) + 256;
    if (need > c.cap_ws) {
        c.cap_ws = need + need / 4;
        PAMO_TRY(s3::dalloc(&c.w.ws, (size_t)c.cap_ws, c.stream));
    }
    return s3::tree_alloc(c.w.tree, nF, c.stream);
}

// max over sqd of the one-sided distances of (VA -> B), written to c.mx
static int one_side(Ctx& c, const double* VA, int nVA, const double* VB, int nVB, const int* FB, int nFB) {
    PAMO_CHECK(nFB > 0, "pamo_hausdorff: empty or null mesh");
    PAMO_TRY(reserve(c, nFB));
    PAMO_TRY(build_tree(c.w, c.stream, VB, nVB, FB, nFB, "pamo_hausdorff"));
    if (nVA > c.cap_sqd) {
        c.cap_sqd = nVA + nVA / 4;
        PAMO_TRY(s3::dalloc(&c.sqd, (size_t)c.cap_sqd, c.stream));
    }
    PAMO_TRY(query(c.w, c.stream, VA, nVA, VB, FB, c.sqd, nullptr, nullptr));
    const unsigned nb = std::min<unsigned>(grid_for(nVA, kB), 1024u);
    hipLaunchKernelGGL(max_bits_k, dim3(nb), dim3(kB), 0, c.stream, (const double*)c.sqd, (long long)nVA, c.mx);
    PAMO_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------ allocation-free variant
// pamo_point_mesh_sqdist_noalloc keeps everything a call needs beyond its
// inputs and outputs (the tree over at most kNoAllocMaxFaces faces, the sort
// and scan workspace and the two flags) in one static device array, so that
// it calls no allocator.  Host code uses the array only through the address
// hipGetSymbolAddress returns for the current device: in host code the
// symbol's name is a host-side stand-in, not a device address.
constexpr int kNoAllocMaxFaces = 1 << 16;
constexpr size_t kNoAllocBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
struct NoAllocDevice {
    std::mutex mu;          // serialises calls on one device
    unsigned char* base = nullptr;  // device address of this device's scratch
    Work w;                 // views into it
};
static NoAllocDevice g_noalloc[kNoAllocMaxDevices];

// Carves the work buffers for up to max_faces faces out of [base, base +
// bytes) (base may be null to only measure).  Returns the bytes needed.
static size_t noalloc_layout(unsigned char* base, int max_faces, Work& w) {
    size_t off = 0;
    auto take = [&](auto*& ptr, size_t count) {
        using T = std::remove_reference_t<decltype(*ptr)>;
        off = (off + 255) & ~(size_t)255;
        ptr = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += sizeof(T) * std::max<size_t>(count, 1);
    };
    // the sizes of s3::tree_alloc for a capacity of max_faces + 16 items
    const size_t a = (size_t)max_faces + 16;
    Tree& t = w.tree;
    take(t.item, 2 * a);
    take(t.key, a); take(t.val, a); take(t.key2, a); take(t.val2, a);
    take(t.left, a); take(t.right, a); take(t.parent, 2 * a);
    take(t.range, a);
    take(t.box, 4 * a);
    take(t.pre, 2 * a); take(t.suf, 2 * a);
    const size_t nb = a / kBlock + 2, ns = nb / kBlock + 2;
    take(t.blk, 2 * nb); take(t.blkpre, 2 * nb); take(t.blksuf, 2 * nb);
    take(t.sblk, 2 * ns);
    take(t.blktab, 2 * kTabLevels * nb);
    t.tab_stride = (long long)nb;
    take(t.part, 6 * (a / kBlock + 4));
    take(t.cnt, (size_t)C_COUNT);
    t.cap = (int)a;
    const long long ws = std::max(radix_workspace(max_faces + 64LL), scan_workspace(max_faces + 65LL)) + 256;
    take(w.ws, (size_t)ws);
    take(w.flags, 2);
    return off;
}

// Resolves this device's scratch address once (with the device current) and
// lays the work buffers out in it.
static int noalloc_setup(int device, NoAllocDevice*& out) {
    PAMO_CHECK(device >= 0 && device < kNoAllocMaxDevices, "pamo_point_mesh_sqdist_noalloc: device id out of range");
    NoAllocDevice& d = g_noalloc[device];
    out = &d;
    if (d.base) return 0;
    void* ptr = nullptr;
    PAMO_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    PAMO_CHECK(ptr != nullptr, "pamo_point_mesh_sqdist_noalloc: no device address for the scratch array");
    Work probe;
    PAMO_CHECK(noalloc_layout(nullptr, kNoAllocMaxFaces, probe) <= kNoAllocBytes,
               "pamo_point_mesh_sqdist_noalloc: the scratch array is smaller than its layout");
    noalloc_layout(static_cast<unsigned char*>(ptr), kNoAllocMaxFaces, d.w);
    d.base = static_cast<unsigned char*>(ptr);
    return 0;
}

}  // namespace metrics
}  // namespace pamo

using MCtx = pamo::metrics::Ctx;

extern "C" {

int pamo_metrics_create(pamo_metrics** out, int device) {
    pamo::clear_error();
    PAMO_CHECK(out != nullptr, "pamo_metrics_create: null argument");
    MCtx* c = new MCtx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&c->w.flags, 2 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&c->mx, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->w.h_flags, 2 * sizeof(int), 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_mx, sizeof(unsigned long long), 0);
    if (e != hipSuccess) {
        pamo::set_error(std::string("pamo_metrics_create: ") + hipGetErrorString(e));
        pamo::metrics::ctx_free(*c);
        delete c;
        return 1;
    }
    *out = reinterpret_cast<pamo_metrics*>(c);
    return 0;
}

int pamo_metrics_destroy(pamo_metrics* m) {
    pamo::clear_error();
    if (!m) return 0;
    MCtx* c = reinterpret_cast<MCtx*>(m);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    pamo::metrics::ctx_free(*c);
    delete c;
    return 0;
}

int pamo_point_mesh_sqdist(pamo_metrics* m, const double* P, long long n_points, const double* V, int n_v,
                           const int* F, int n_f, double* sqr_d, int* face_index, double* closest) {
    pamo::clear_error();
    PAMO_CHECK(m != nullptr, "pamo_point_mesh_sqdist: null context");
    PAMO_CHECK(n_points >= 0 && (n_points == 0 || (P && sqr_d)), "pamo_point_mesh_sqdist: null argument");
    MCtx& c = *reinterpret_cast<MCtx*>(m);
    PAMO_HIP(hipSetDevice(c.device));
    PAMO_CHECK(V && F && n_v > 0 && n_f > 0, "pamo_point_mesh_sqdist: empty or null mesh");
    PAMO_TRY(pamo::metrics::reserve(c, n_f));
    PAMO_TRY(pamo::metrics::begin(c.w, c.stream));
    PAMO_TRY(pamo::metrics::build_tree(c.w, c.stream, V, n_v, F, n_f, "pamo_point_mesh_sqdist"));
    PAMO_TRY(pamo::metrics::query(c.w, c.stream, P, n_points, V, F, sqr_d, face_index, closest));
    return pamo::metrics::finish(c.w, c.stream, "pamo_point_mesh_sqdist");
}

int pamo_hausdorff(pamo_metrics* m, const double* VA, int n_va, const int* FA, int n_fa, const double* VB, int n_vb,
                   const int* FB, int n_fb, double* d_out) {
    pamo::clear_error();
    PAMO_CHECK(m != nullptr && d_out != nullptr, "pamo_hausdorff: null argument");
    MCtx& c = *reinterpret_cast<MCtx*>(m);
    PAMO_HIP(hipSetDevice(c.device));
    PAMO_HIP(hipMemsetAsync(c.mx, 0, sizeof(unsigned long long), c.stream));
    PAMO_TRY(pamo::metrics::begin(c.w, c.stream));
    // vertices of B against the surface of A, then vertices of A against B
    PAMO_TRY(pamo::metrics::one_side(c, VB, n_vb, VA, n_va, FA, n_fa));
    PAMO_TRY(pamo::metrics::one_side(c, VA, n_va, VB, n_vb, FB, n_fb));
    PAMO_HIP(hipMemcpyAsync(c.h_mx, c.mx, sizeof(unsigned long long), hipMemcpyDeviceToHost, c.stream));
    PAMO_TRY(pamo::metrics::finish(c.w, c.stream, "pamo_hausdorff"));
    double d2;
    std::memcpy(&d2, c.h_mx, sizeof(double));
    *d_out = std::sqrt(d2);
    return 0;
}

int pamo_point_mesh_sqdist_noalloc(const double* P, long long n_points, const double* V, int n_v, const int* F,
                                   int n_f, double* sqr_d, int* face_index, double* closest, void* stream_handle) {
    using namespace pamo::metrics;
    hipStream_t stream = static_cast<hipStream_t>(stream_handle);
    pamo::clear_error();
    PAMO_CHECK(n_points >= 0 && (n_points == 0 || (P && sqr_d)), "pamo_point_mesh_sqdist_noalloc: null argument");
    PAMO_CHECK(V && F && n_v > 0 && n_f > 0, "pamo_point_mesh_sqdist_noalloc: empty or null mesh");
    PAMO_CHECK(n_f <= kNoAllocMaxFaces, "pamo_point_mesh_sqdist_noalloc: " + std::to_string(n_f) +
                                           " faces, more than the " + std::to_string(kNoAllocMaxFaces) +
                                           " its static scratch holds (use pamo_point_mesh_sqdist)");
    int device = 0;
    PAMO_HIP(hipGetDevice(&device));
    if (stream) {
        hipDevice_t sdev = 0;
        PAMO_HIP(hipStreamGetDevice(stream, &sdev));
        PAMO_CHECK((int)sdev == device, "pamo_point_mesh_sqdist_noalloc: the stream belongs to another device");
    }
    // the setup and the call hold the device's lock: calls on one device
    // share its scratch and run one at a time, each finishing before the next
    NoAllocDevice* d = nullptr;
    std::unique_lock<std::mutex> lock(g_noalloc[std::max(0, std::min(device, kNoAllocMaxDevices - 1))].mu);
    PAMO_TRY(noalloc_setup(device, d));
    // the flags come back to this call's own host words
    int h_flags[2] = {0, 0};
    Work w = d->w;
    w.h_flags = h_flags;
    PAMO_TRY(begin(w, stream));
    PAMO_TRY(build_tree(w, stream, V, n_v, F, n_f, "pamo_point_mesh_sqdist_noalloc"));
    PAMO_TRY(query(w, stream, P, n_points, V, F, sqr_d, face_index, closest));
    return finish(w, stream, "pamo_point_mesh_sqdist_noalloc");
}

}  // extern "C"
