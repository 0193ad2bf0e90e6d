// PaMO stage-2 CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU and in FP32 with no FMA contraction
// (build with -ffp-contract=off, as the reference builds with --fmad=false,
// simp_cuda/setup.py:35), the stage-2 algorithm of the reference:
//   * simp_cuda/src/cusimp_free.cu:208-1048   (one collapse round, "forward")
//   * simp_cuda/src/bvh/bvh.cuh:89-461         (LBVH: Morton, stable sort, Karras tree)
//   * simp_cuda/src/bvh/morton_code.cuh:9-37   (30-bit Morton codes)
//   * simp_cuda/src/bvh/self_intersect.cuh     (candidate pairs + triangle guard)
//   * simp_cuda/src/bvh/tri_tri_3d.cuh:99-276  (coplanarity + triangle/triangle test)
//   * simp_cuda/pamo/__init__.py:122-190       (the PaMO.run round loop)
//
// It is only ever linked by the test-suite and by bench.py's CPU-baseline
// leg, as the checker of the HIP product path.  Nothing in the product
// imports it.
//
// Canonical semantics where the reference is nondeterministic (each is
// shared with the HIP path so the two are bit-comparable):
//   * vertex -> face rings are in ascending face index (reference: atomicAdd
//     order, cusimp_free.cu:291-304);
//   * collapsed edges are recorded in ascending edge index (reference:
//     atomicAdd order, :684-685), followed by "line" collapses in ascending
//     edge index (:245-246);
//   * "line" collapses (:234-274) are flagged from the positions left by the
//     regular collapse step and then applied one by one in edge order;
//   * undo candidates are pushed in record order (reference: atomicAdd,
//     :795-798);
//   * buffers that the reference reads stale (edge_cost when a cost is not
//     written, pts_map, the undo list tail, the invalid-vertex table) are
//     persistent per context and zero-initialised when (re)allocated.
//   * removed faces never enter the LBVH (reference reads points[-1] for
//     them, self_intersect.cuh:164-169).
//   * the intersection query is issued only for faces that belong to the
//     one-ring of a recorded collapse ("changed" faces): only those faces can
//     change an undo decision (cusimp_free.cu:776-827), so the decisions are
//     identical to querying every face.
//
// Parity of each function: see tests/test_oracle_*.py for the closed forms,
// brute-force and invariant checks that pin it.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "exact_pred.h"

namespace orc {

struct V3 {
    float x, y, z;
};
static inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline float vnorm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline V3 vdivs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }

// Ternary min/max/clamp of cusimp_free.cu:29-34 (NaN -> upper bound).
static inline float tmin(float a, float b) { return a < b ? a : b; }
static inline float tmax(float a, float b) { return a > b ? a : b; }
static inline float tclamp(float x, float a, float b) { return tmin(tmax(a, x), b); }

// float -> uint32 with the saturation CUDA applies (cusimp_free.cu:584,923).
static inline uint32_t sat_u32(float q) {
    if (!(q < 4294967296.0f)) return 0xFFFFFFFFu;  // also NaN
    if (q <= 0.0f) return 0u;
    return (uint32_t)q;
}

struct Tri {
    int i, j, k;
};

// ---------------------------------------------------------------- shuffle
// thrust::shuffle(..., thrust::default_random_engine) (cusimp_free.cu:937-941)
// Method: 24-round Feistel bijection over the next power of two, keyed by 24
// draws of minstd_rand (a=48271, m=2^31-1, seed 1), followed by a stable
// compaction of the indices below m (Mitchell et al., "Bandwidth-optimal
// random shuffling for GPUs", ACM TOPC 2022).
struct Feistel {
    uint64_t lbits, rbits, lmask, rmask;
    uint32_t key[24];
    explicit Feistel(uint64_t m) {
        uint64_t total = 4;
        if (m > 16) {
            total = 0;
            uint64_t t = m - 1;
            while (t) { ++total; t >>= 1; }
        }
        lbits = total / 2;
        rbits = total - lbits;
        lmask = (1ull << lbits) - 1;
        rmask = (1ull << rbits) - 1;
        uint64_t s = 1;
        for (int r = 0; r < 24; ++r) {
            s = (s * 48271ull) % 2147483647ull;
            key[r] = (uint32_t)s;
        }
    }
    uint64_t domain() const { return 1ull << (lbits + rbits); }
    uint64_t operator()(uint64_t v) const {
        uint32_t L = (uint32_t)(v >> rbits), R = (uint32_t)(v & rmask);
        for (int r = 0; r < 24; ++r) {
            uint64_t prod = 0xD2B74407B1CE6E93ull * (uint64_t)L;
            uint32_t hi = (uint32_t)(prod >> 32), lo = (uint32_t)prod;
            lo = (lo << (rbits - lbits)) | (R >> lbits);
            L = ((hi ^ key[r]) ^ R) & (uint32_t)lmask;
            R = lo & (uint32_t)rmask;
        }
        return ((uint64_t)L << rbits) | (uint64_t)R;
    }
};

static void shuffle_perm(uint64_t m, std::vector<int64_t>& perm) {
    perm.clear();
    perm.reserve(m);
    Feistel f(m);
    uint64_t n = f.domain();
    for (uint64_t idx = 0; idx < n; ++idx) {
        uint64_t k = f(idx);
        if (k < m) perm.push_back((int64_t)k);
    }
}

// ---------------------------------------------------------------- LBVH
struct Box {
    float lo[3], hi[3];
};

static inline uint32_t expand_bits(uint32_t v) {
    // spread the low 10 bits of v so that bit b lands at bit 3b
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// morton_code.cuh:19-37 with resolution 1024
static inline uint32_t morton3(float x, float y, float z) {
    bool nx = x < 0.0f, ny = y < 0.0f, nz = z < 0.0f;
    x = std::fmin(std::fmax(x * 1024.0f, 0.0f), 1023.0f);
    y = std::fmin(std::fmax(y * 1024.0f, 0.0f), 1023.0f);
    z = std::fmin(std::fmax(z * 1024.0f, 0.0f), 1023.0f);
    uint32_t xx = expand_bits((uint32_t)x), yy = expand_bits((uint32_t)y), zz = expand_bits((uint32_t)z);
    return (nx ? 0x80000000u : 0u) | (ny ? 0x40000000u : 0u) | (nz ? 0x20000000u : 0u) | (xx * 4 + yy * 2 + zz);
}

static inline Box tri_box(V3 a, V3 b, V3 c) {
    Box bx;
    bx.lo[0] = std::fmin(a.x, std::fmin(b.x, c.x));
    bx.lo[1] = std::fmin(a.y, std::fmin(b.y, c.y));
    bx.lo[2] = std::fmin(a.z, std::fmin(b.z, c.z));
    bx.hi[0] = std::fmax(a.x, std::fmax(b.x, c.x));
    bx.hi[1] = std::fmax(a.y, std::fmax(b.y, c.y));
    bx.hi[2] = std::fmax(a.z, std::fmax(b.z, c.z));
    return bx;
}

static inline bool box_overlap(const Box& a, const Box& b) {
    for (int d = 0; d < 3; ++d)
        if (a.hi[d] < b.lo[d] || b.hi[d] < a.lo[d]) return false;
    return true;
}

struct Lbvh {
    int n = 0;                        // leaves (live faces)
    std::vector<int> leaf_face;       // sorted position -> face index
    std::vector<uint32_t> code;       // Morton code per sorted leaf
    std::vector<uint64_t> key;        // (code << 32) | face
    std::vector<int> left, right, parent;  // internal nodes 0..n-2; children index into 0..2n-2
    std::vector<Box> box;             // 2n-1 node boxes (internal first, then leaves)
};

static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees,
// and k-d trees": node i covers a key range found by the sign of the
// common-prefix difference, and splits at the highest differing bit.
static void karras_node(const std::vector<uint64_t>& k, int n, int i, int& lchild, int& rchild) {
    auto delta = [&](int a, int b) -> int {
        if (b < 0 || b >= n) return -1;
        return clz64(k[a] ^ k[b]);
    };
    int first, last;
    if (i == 0) {
        first = 0;
        last = n - 1;
    } else {
        int dl = delta(i, i - 1), dr = delta(i, i + 1);
        int d = dr > dl ? 1 : -1;
        int dmin = std::min(dl, dr);
        int lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t > 0; t >>= 1)
            if (delta(i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        first = std::min(i, j);
        last = std::max(i, j);
    }
    int split;
    if (k[first] == k[last]) {
        split = (first + last) >> 1;
    } else {
        int dnode = clz64(k[first] ^ k[last]);
        split = first;
        int stride = last - first;
        do {
            stride = (stride + 1) >> 1;
            int mid = split + stride;
            if (mid < last && clz64(k[first] ^ k[mid]) > dnode) split = mid;
        } while (stride > 1);
    }
    lchild = (first == split) ? split + (n - 1) : split;
    rchild = (last == split + 1) ? split + 1 + (n - 1) : split + 1;
}

static void build_lbvh(const std::vector<V3>& P, const std::vector<Tri>& T, Lbvh& b) {
    std::vector<int> live;
    live.reserve(T.size());
    for (int f = 0; f < (int)T.size(); ++f)
        if (T[f].i != -1) live.push_back(f);
    int n = (int)live.size();
    b.n = n;
    b.leaf_face.assign(n, 0);
    b.code.assign(n, 0);
    b.key.assign(n, 0);
    b.left.assign(n > 0 ? n - 1 : 0, -1);
    b.right.assign(n > 0 ? n - 1 : 0, -1);
    b.parent.assign(n > 0 ? 2 * n - 1 : 0, -1);
    b.box.assign(n > 0 ? 2 * n - 1 : 0, Box{});
    if (n == 0) return;
    std::vector<Box> lb(n);
    const float inf = std::numeric_limits<float>::infinity();
    Box whole{{inf, inf, inf}, {-inf, -inf, -inf}};
    for (int s = 0; s < n; ++s) {
        const Tri t = T[live[s]];
        lb[s] = tri_box(P[t.i], P[t.j], P[t.k]);
        for (int d = 0; d < 3; ++d) {
            whole.lo[d] = std::fmin(whole.lo[d], lb[s].lo[d]);
            whole.hi[d] = std::fmax(whole.hi[d], lb[s].hi[d]);
        }
    }
    std::vector<uint32_t> codes(n);
    for (int s = 0; s < n; ++s) {
        float c[3];
        for (int d = 0; d < 3; ++d) {
            c[d] = (lb[s].hi[d] + lb[s].lo[d]) * 0.5f;
            c[d] -= whole.lo[d];
            c[d] /= (whole.hi[d] - whole.lo[d]);
        }
        codes[s] = morton3(c[0], c[1], c[2]);
    }
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return codes[a] < codes[c]; });
    for (int s = 0; s < n; ++s) {
        b.leaf_face[s] = live[ord[s]];
        b.code[s] = codes[ord[s]];
        b.key[s] = ((uint64_t)codes[ord[s]] << 32) | (uint64_t)(uint32_t)live[ord[s]];
        b.box[n - 1 + s] = lb[ord[s]];
    }
    for (int i = 0; i < n - 1; ++i) {
        int l, r;
        karras_node(b.key, n, i, l, r);
        b.left[i] = l;
        b.right[i] = r;
        b.parent[l] = i;
        b.parent[r] = i;
    }
    // bottom-up boxes: children always have larger subtree index structure;
    // a simple recursive post-order avoids relying on any ordering.
    if (n > 1) {
        std::vector<int> st;
        std::vector<char> done(n - 1, 0);
        st.push_back(0);
        while (!st.empty()) {
            int i = st.back();
            int l = b.left[i], r = b.right[i];
            bool lready = (l >= n - 1) || done[l];
            bool rready = (r >= n - 1) || done[r];
            if (lready && rready) {
                for (int d = 0; d < 3; ++d) {
                    b.box[i].lo[d] = std::fmin(b.box[l].lo[d], b.box[r].lo[d]);
                    b.box[i].hi[d] = std::fmax(b.box[l].hi[d], b.box[r].hi[d]);
                }
                done[i] = 1;
                st.pop_back();
            } else {
                if (!rready) st.push_back(r);
                if (!lready) st.push_back(l);
            }
        }
    }
}

// All live faces whose box overlaps q (closed intervals), excluding `self`.
static void lbvh_query(const Lbvh& b, const Box& q, int self, std::vector<int>& out) {
    out.clear();
    int n = b.n;
    if (n == 0) return;
    if (n == 1) {
        if (b.leaf_face[0] != self && box_overlap(q, b.box[0])) out.push_back(b.leaf_face[0]);
        return;
    }
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        int node = stack[--sp];
        int ch[2] = {b.left[node], b.right[node]};
        for (int c : ch) {
            if (!box_overlap(q, b.box[c])) continue;
            if (c >= n - 1) {
                int f = b.leaf_face[c - (n - 1)];
                if (f != self) out.push_back(f);
            } else {
                stack[sp++] = c;
            }
        }
    }
}

// ---------------------------------------------------------------- predicate
// Reference triangle guard, self_intersect.cuh:229-378 together with
// tri_tri_3d.cuh:99-276 and math.cuh:43-52.  "cur" is the current face, "qry"
// the query face; the test is evaluated for one ordered pair.
static inline int floor_mean6(float a, float b, float c, float d, float e, float f) {
    float m = (a + b + c + d + e + f) / 6.0f;
    return (int)std::floor(m);
}

static bool ref_coplanar(const float A[3][3], const float B[3][3]) {
    float sc = 0.0f;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sc = (float)std::fmax(sc, std::fabs(A[i][j]));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sc = (float)std::fmax(sc, std::fabs(B[i][j]));
    double ba[3], bc[3], pl[4];
    for (int d = 0; d < 3; ++d) {
        ba[d] = (double)A[0][d] - (double)A[1][d];
        bc[d] = (double)A[2][d] - (double)A[1][d];
    }
    pl[0] = ba[1] * bc[2] - ba[2] * bc[1];
    pl[1] = ba[2] * bc[0] - ba[0] * bc[2];
    pl[2] = ba[0] * bc[1] - ba[1] * bc[0];
    pl[3] = -(pl[0] * (double)A[1][0] + pl[1] * (double)A[1][1] + pl[2] * (double)A[1][2]);
    double eps = (double)sc * 1e-6;
    for (int v = 0; v < 3; ++v) {
        double s = (pl[0] * (double)B[v][0] + pl[1] * (double)B[v][1] + pl[2] * (double)B[v][2]) + pl[3];
        if (!(std::fabs(s) <= eps)) return false;
    }
    return true;
}

// Blender-derived segment-overlap test along the plane-plane line
// (tri_tri_3d.cuh:130-276).  Returns whether the triangles cross, with the
// overlap endpoints in r1/r2.
static bool ref_isect(const float A[3][3], const float B[3][3], float r1[3], float r2[3]) {
    float side[2][3];
    double ba[3], bc[3], pa[4], pb[4];
    auto plane = [&](const float T[3][3], double p[4]) {
        for (int d = 0; d < 3; ++d) {
            ba[d] = (double)T[0][d] - (double)T[1][d];
            bc[d] = (double)T[2][d] - (double)T[1][d];
        }
        p[0] = ba[1] * bc[2] - ba[2] * bc[1];
        p[1] = ba[2] * bc[0] - ba[0] * bc[2];
        p[2] = ba[0] * bc[1] - ba[1] * bc[0];
        p[3] = -(p[0] * (double)T[1][0] + p[1] * (double)T[1][1] + p[2] * (double)T[1][2]);
    };
    auto sdist = [&](const double p[4], const float v[3]) -> float {
        return (float)((p[0] * (double)v[0] + p[1] * (double)v[1] + p[2] * (double)v[2]) + p[3]);
    };
    plane(A, pa);
    for (int v = 0; v < 3; ++v) side[1][v] = sdist(pa, B[v]);
    if (!side[1][0] && !side[1][1] && !side[1][2]) return false;
    if ((side[1][0] && side[1][1] && side[1][2]) && (side[1][0] < 0.0f) == (side[1][1] < 0.0f) &&
        (side[1][0] < 0.0f) == (side[1][2] < 0.0f))
        return false;
    plane(B, pb);
    for (int v = 0; v < 3; ++v) side[0][v] = sdist(pb, A[v]);
    if ((side[0][0] && side[0][1] && side[0][2]) && (side[0][0] < 0.0f) == (side[0][1] < 0.0f) &&
        (side[0][0] < 0.0f) == (side[0][2] < 0.0f))
        return false;
    double dir[3];
    dir[0] = pa[1] * pb[2] - pa[2] * pb[1];
    dir[1] = pa[2] * pb[0] - pa[0] * pb[2];
    dir[2] = pa[0] * pb[1] - pa[1] * pb[0];
    float rmin[2], rmax[2], loc[2][2][3];
    for (int t = 0; t < 2; ++t) {
        const float(*T)[3] = t == 0 ? A : B;
        const float* sd = side[t];
        int o0, o1, o2;  // o1 is the vertex alone on its side
        if ((sd[0] && sd[1]) && (sd[0] < 0.0f) == (sd[1] < 0.0f)) {
            o0 = 1; o1 = 2; o2 = 0;
        } else if ((sd[1] && sd[2]) && (sd[1] < 0.0f) == (sd[2] < 0.0f)) {
            o0 = 2; o1 = 0; o2 = 1;
        } else if (!sd[1] && (sd[2] < 0.0f) != (sd[0] < 0.0f)) {
            o0 = 1; o1 = 2; o2 = 0;
        } else {
            o0 = 0; o1 = 1; o2 = 2;
        }
        double db = dir[0] * (double)T[o1][0] + dir[1] * (double)T[o1][1] + dir[2] * (double)T[o1][2];
        float sc = sd[o1];
        if (sc) {
            double da = dir[0] * (double)T[o0][0] + dir[1] * (double)T[o0][1] + dir[2] * (double)T[o0][2];
            double dc = dir[0] * (double)T[o2][0] + dir[1] * (double)T[o2][1] + dir[2] * (double)T[o2][2];
            float f0 = sc / (sc - sd[o0]);
            float f1 = sc / (sc - sd[o2]);
            double off0 = f0 * (da - db);
            double off1 = f1 * (dc - db);
            if (off0 > off1) {
                std::swap(off0, off1);
                std::swap(f0, f1);
                std::swap(o0, o2);
            }
            rmin[t] = (float)(db + off0);
            rmax[t] = (float)(db + off1);
            float s0 = 1.0f - f0, s1 = 1.0f - f1;
            for (int d = 0; d < 3; ++d) {
                loc[t][0][d] = s0 * T[o1][d] + f0 * T[o0][d];
                loc[t][1][d] = s1 * T[o1][d] + f1 * T[o2][d];
            }
        } else {
            rmin[t] = rmax[t] = (float)db;
            for (int d = 0; d < 3; ++d) loc[t][0][d] = loc[t][1][d] = T[o1][d];
        }
    }
    if ((rmax[0] > rmin[1]) && (rmin[0] < rmax[1])) {
        if (rmin[0] >= rmin[1]) {
            std::memcpy(r1, loc[0][0], 12);
            if (rmax[0] <= rmax[1]) std::memcpy(r2, loc[0][1], 12);
            else std::memcpy(r2, loc[1][1], 12);
        } else {
            if (rmax[0] <= rmax[1]) {
                std::memcpy(r1, loc[0][1], 12);
                std::memcpy(r2, loc[1][0], 12);
            } else {
                std::memcpy(r1, loc[1][0], 12);
                std::memcpy(r2, loc[1][1], 12);
            }
        }
        return true;
    }
    return false;
}

// One ordered pair as the reference's guard lambda evaluates it.
// cur/qry: 9 floats (3 vertices) and 3 vertex indices each.
static bool ref_guard(const float* cur, const int* ci, const float* qry, const int* qi, float epsilon) {
    float p[6][3];
    for (int v = 0; v < 3; ++v)
        for (int d = 0; d < 3; ++d) {
            p[v][d] = cur[3 * v + d];
            p[3 + v][d] = qry[3 * v + d];
        }
    for (int d = 0; d < 3; ++d) {
        int m = floor_mean6(p[0][d], p[1][d], p[2][d], p[3][d], p[4][d], p[5][d]);
        for (int v = 0; v < 6; ++v) p[v][d] -= (float)m;
    }
    int shared = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            if (ci[a] == qi[b]) ++shared;
    float A[3][3], B[3][3];
    std::memcpy(A, p[0], sizeof(A));
    std::memcpy(B, p[3], sizeof(B));
    if (ref_coplanar(A, B)) return false;
    if (shared == 2) return false;
    // two coordinates-coincident vertex pairs within 1e-6 per axis (self_intersect.cuh:37-48)
    int same = 0;
    for (int a = 0; a < 3; ++a) {
        bool any = false;
        for (int b = 0; b < 3; ++b) {
            if (std::fabs(A[a][0] - B[b][0]) < 1e-6f && std::fabs(A[a][1] - B[b][1]) < 1e-6f &&
                std::fabs(A[a][2] - B[b][2]) < 1e-6f)
                any = true;
        }
        same += any ? 1 : 0;
    }
    if (same >= 2) return false;
    float r1[3], r2[3];
    if (!ref_isect(A, B, r1, r2)) return false;
    float dx = std::fabs(r2[0] - r1[0]), dy = std::fabs(r2[1] - r1[1]), dz = std::fabs(r2[2] - r1[2]);
    float big = -std::numeric_limits<float>::infinity();
    if (dx > big) big = dx;
    if (dy > big) big = dy;
    if (dz > big) big = dz;
    if (big < epsilon && shared == 1) return false;
    return true;
}

// ---------------------------------------------------------------- round state
struct Ctx {
    // persistent, capacity-tracked (zeroed on growth)
    size_t cap_pts = 0, cap_edges = 0;
    std::vector<int> pts_map;        // cap_pts + 1
    std::vector<int> undo_list;      // 4 * (cap_pts + 1)
    std::vector<int> tmp_undo;       // 4 * (cap_pts + 1)
    std::vector<uint8_t> invalid;    // cap_pts + 1
    std::vector<uint32_t> edge_cost; // cap_edges + 1
    // per round
    int n_pts = 0, n_tris = 0, n_edges = 0;
    std::vector<V3> P, P0;
    std::vector<Tri> T, T0;
    std::vector<int> occ;
    std::vector<int> ring_off, ring;
    std::vector<int> eu, ev;
    std::vector<float> Q;            // 10 per vertex: 00 01 02 03 11 12 13 22 23 33
    std::vector<uint64_t> tri_min;
    std::vector<int> collapsed;      // records (edge indices)
    int n_regular = 0;               // records from the regular collapse step
    std::vector<uint8_t> flagged;    // faces flagged by the guard
    std::vector<int> undo_iters_edges;  // debug: undone edge count per iteration
    int n_vertices_undo = 0;
    int pred_mode = 0;
    int placement = 0;               // 0 midpoint (the reference), 1 optimal (opt-in)
    long long guard_pairs = 0;       // ordered pairs evaluated (debug)
};

static void grow_pts(Ctx& c, size_t n) {
    if (n > c.cap_pts) {
        c.cap_pts = n + n / 5;
        c.pts_map.assign(c.cap_pts + 1, 0);
        c.undo_list.assign(4 * (c.cap_pts + 1), 0);
        c.tmp_undo.assign(4 * (c.cap_pts + 1), 0);
        c.invalid.assign(c.cap_pts + 1, 0);
    }
}
static void grow_edges(Ctx& c, size_t n) {
    if (n > c.cap_edges) {
        c.cap_edges = n + n / 5;
        c.edge_cost.assign(c.cap_edges + 1, 0u);
    }
}

static inline int succ_of(const Tri& t, int a) {
    if (t.i == a) return t.j;
    if (t.j == a) return t.k;
    if (t.k == a) return t.i;
    return -2;  // not in face (cannot happen for a valid ring)
}

static void build_rings(Ctx& c) {
    int nP = c.n_pts, nT = c.n_tris;
    c.ring_off.assign(nP + 1, 0);
    for (int f = 0; f < nT; ++f) {
        const Tri t = c.T[f];
        if (t.i == -1 && t.j == -1 && t.k == -1) continue;
        c.ring_off[t.i]++;
        c.ring_off[t.j]++;
        c.ring_off[t.k]++;
    }
    int acc = 0;
    for (int v = 0; v <= nP; ++v) {
        int x = c.ring_off[v];
        c.ring_off[v] = acc;
        acc += x;
    }
    c.ring.assign(acc, 0);
    std::vector<int> fill(nP, 0);
    for (int f = 0; f < nT; ++f) {  // ascending face order -> canonical ring order
        const Tri t = c.T[f];
        if (t.i == -1 && t.j == -1 && t.k == -1) continue;
        c.ring[c.ring_off[t.i] + fill[t.i]++] = f;
        c.ring[c.ring_off[t.j] + fill[t.j]++] = f;
        c.ring[c.ring_off[t.k] + fill[t.k]++] = f;
    }
}

static void build_edges(Ctx& c) {
    c.eu.clear();
    c.ev.clear();
    for (int f = 0; f < c.n_tris; ++f) {
        const Tri t = c.T[f];
        if (t.i == -1 && t.j == -1 && t.k == -1) continue;
        if (t.i > t.j) { c.eu.push_back(t.j); c.ev.push_back(t.i); }
        if (t.j > t.k) { c.eu.push_back(t.k); c.ev.push_back(t.j); }
        if (t.k > t.i) { c.eu.push_back(t.i); c.ev.push_back(t.k); }
    }
    c.n_edges = (int)c.eu.size();
}

static void vertex_quadrics(Ctx& c) {
    c.Q.assign((size_t)c.n_pts * 10, 0.0f);
#pragma omp parallel for schedule(static)
    for (int v = 0; v < c.n_pts; ++v) {
        float q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int r = c.ring_off[v]; r < c.ring_off[v + 1]; ++r) {
            const Tri t = c.T[c.ring[r]];
            V3 a = c.P[t.i], b = c.P[t.j], d = c.P[t.k];
            V3 n = vcross(vsub(b, a), vsub(d, a));
            n = vdivs(n, vnorm(n));
            float w = -vdot(n, a);
            float p[4] = {n.x, n.y, n.z, w};
            int k = 0;
            for (int i = 0; i < 4; ++i)
                for (int j = i; j < 4; ++j) q[k++] += p[i] * p[j];
        }
        std::memcpy(&c.Q[(size_t)v * 10], q, sizeof(q));
    }
}

// index into the 10-float symmetric storage
static inline int qidx(int i, int j) {
    static const int tab[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
    return tab[i][j];
}

static inline float heron(V3 a, V3 b, V3 d) {
    float x = vnorm(vsub(a, b)), y = vnorm(vsub(b, d)), z = vnorm(vsub(d, a));
    float s = (x + y + z) / 2;
    return std::sqrt(s * (s - x) * (s - y) * (s - z));
}

// compute_edge_cost_kernel, cusimp_free.cu:382-585
// Opt-in optimal placement (no counterpart in the reference, which always
// collapses to the midpoint, cusimp_free.cu:491,687): the minimiser of
// x^T M x over x = (p, 1) for M = Q_u + Q_v (10 floats 00 01 02 03 11 12 13
// 22 23 33), from the 3x3 system A p = -b solved in FP64 by cofactors.  The
// midpoint is kept when A is singular (|det| <= 1e-9 max|a_ij|^3, NaN
// included) or when the minimiser lies farther from the midpoint than the
// edge is long.
static V3 place_optimal(const float* M, V3 p0, V3 p1) {
    const V3 mid = vdivs(vadd(p0, p1), 2.0f);
    const double a00 = M[0], a01 = M[1], a02 = M[2], a11 = M[4], a12 = M[5], a22 = M[7];
    const double b0 = -(double)M[3], b1 = -(double)M[6], b2 = -(double)M[8];
    const double c00 = a11 * a22 - a12 * a12;
    const double c01 = a02 * a12 - a01 * a22;
    const double c02 = a01 * a12 - a02 * a11;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    double amax = std::fabs(a00);
    for (double a : {a01, a02, a11, a12, a22}) amax = std::fmax(amax, std::fabs(a));
    if (!(std::fabs(det) > 1e-9 * amax * amax * amax)) return mid;
    const double c11 = a00 * a22 - a02 * a02;
    const double c12 = a01 * a02 - a00 * a12;
    const double c22 = a00 * a11 - a01 * a01;
    const double x = (c00 * b0 + c01 * b1 + c02 * b2) / det;
    const double y = (c01 * b0 + c11 * b1 + c12 * b2) / det;
    const double z = (c02 * b0 + c12 * b1 + c22 * b2) / det;
    const double ex = (double)p1.x - (double)p0.x, ey = (double)p1.y - (double)p0.y, ez = (double)p1.z - (double)p0.z;
    const double dx = x - (double)mid.x, dy = y - (double)mid.y, dz = z - (double)mid.z;
    if (!(dx * dx + dy * dy + dz * dz <= ex * ex + ey * ey + ez * ez)) return mid;
    return V3{(float)x, (float)y, (float)z};
}

// M = Q_u + Q_v in the 10-float storage order.
static void edge_quadric(const Ctx& c, int u, int v, float* M) {
    const float* qu = &c.Q[(size_t)u * 10];
    const float* qv = &c.Q[(size_t)v * 10];
    for (int i = 0; i < 10; ++i) M[i] = qu[i] + qv[i];
}

// The collapse position of edge (u, v).
static V3 collapse_point(const Ctx& c, int u, int v) {
    if (c.placement == 1) {
        float M[10];
        edge_quadric(c, u, v, M);
        return place_optimal(M, c.P[u], c.P[v]);
    }
    return vdivs(vadd(c.P[u], c.P[v]), 2.0f);
}

static void edge_costs(Ctx& c, bool is_stuck, float edge_s, float tres) {
    const float k4s3 = 4.0f * std::sqrt(3.0f);
#pragma omp parallel for schedule(dynamic, 256)
    for (int e = 0; e < c.n_edges; ++e) {
        int u = c.eu[e], v = c.ev[e];
        V3 p0 = c.P[u], p1 = c.P[v];
        if (c.invalid[u] && c.invalid[v]) {
            c.edge_cost[e] = 0xFFFFFFFFu;
            continue;
        }
        if (is_stuck) continue;  // stale
        int dup = 0;
        bool over = false;
        for (int a = c.ring_off[u]; a < c.ring_off[u + 1] && !over; ++a) {
            int su = succ_of(c.T[c.ring[a]], u);
            for (int b = c.ring_off[v]; b < c.ring_off[v + 1]; ++b) {
                int sv = succ_of(c.T[c.ring[b]], v);
                if (su == sv) ++dup;
                if (dup > 2) { over = true; break; }
            }
        }
        if (over) { c.edge_cost[e] = 0xFFFFFFFFu; continue; }
        if (dup != 2) continue;  // stale
        float elen = 0.0f;
        int nedge = 0;
        for (int side = 0; side < 2; ++side) {
            int w = side ? v : u;
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) {
                const Tri t = c.T[c.ring[a]];
                if (t.i > t.j) { elen += vnorm(vsub(c.P[t.i], c.P[t.j])); ++nedge; }
                if (t.j > t.k) { elen += vnorm(vsub(c.P[t.j], c.P[t.k])); ++nedge; }
                if (t.k > t.i) { elen += vnorm(vsub(c.P[t.k], c.P[t.i])); ++nedge; }
            }
        }
        elen = elen / (float)nedge / edge_s * tres;
        elen = elen + vnorm(vsub(p0, p1)) / edge_s * tres;
        V3 vm = collapse_point(c, u, v);
        float qa = 0.0f;
        int ntri = 0;
        bool flip = false;
        for (int side = 0; side < 2 && !flip; ++side) {
            int w = side ? v : u, other = side ? u : v;
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) {
                const Tri t = c.T[c.ring[a]];
                if (t.i == other || t.j == other || t.k == other) continue;
                V3 o0 = c.P[t.i], o1 = c.P[t.j], o2 = c.P[t.k];
                V3 n0 = o0, n1 = o1, n2 = o2;
                if (t.i == w) n0 = vm;
                if (t.j == w) n1 = vm;
                if (t.k == w) n2 = vm;
                V3 on = vcross(vsub(o1, o0), vsub(o2, o0));
                on = vdivs(on, vnorm(on));
                V3 nn = vcross(vsub(n1, n0), vsub(n2, n0));
                nn = vdivs(nn, vnorm(nn));
                if (vdot(on, nn) < 0) { flip = true; break; }
                float l01 = vnorm(vsub(n0, n1)), l12 = vnorm(vsub(n1, n2)), l20 = vnorm(vsub(n2, n0));
                float qv = k4s3 * heron(n0, n1, n2) / (l01 * l01) + l12 * l12 + l20 * l20 + 0.0000001f;
                qa += 1.0f - tclamp(qv, 0.0f, 1.0f);
                ++ntri;
            }
        }
        if (flip) { c.edge_cost[e] = 0xFFFFFFFFu; continue; }
        qa = 5.0f * qa;
        float M[16];
        const float* qu = &c.Q[(size_t)u * 10];
        const float* qv = &c.Q[(size_t)v * 10];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) M[i * 4 + j] = qu[qidx(i, j)] + qv[qidx(i, j)];
        float x[4] = {vm.x, vm.y, vm.z, 1.0f};
        float r[4];
        for (int j = 0; j < 4; ++j) r[j] = M[0 * 4 + j] * x[0] + M[1 * 4 + j] * x[1] + M[2 * 4 + j] * x[2] + M[3 * 4 + j] * x[3];
        float cost = (r[0] * x[0] + r[1] * x[1] + r[2] * x[2] + r[3] * x[3]) / (edge_s * edge_s);
        float tot = cost + elen + qa / (float)ntri * tres;
        c.edge_cost[e] = sat_u32(tclamp(tot, 0.0f, 10.0f) / 10.0f * 4294967296.0f);
    }
}

static inline uint64_t edge_key(uint32_t cost, int e) { return ((uint64_t)cost << 32) | (uint64_t)(uint32_t)e; }

static bool is_skinny(V3 p, V3 q, V3 r) {
    V3 pq = vsub(q, p), pr = vsub(r, p), qr = vsub(r, q);
    float lpq = std::sqrt(vdot(pq, pq)), lpr = std::sqrt(vdot(pr, pr)), lqr = std::sqrt(vdot(qr, qr));
    float ml = std::fmax(lpq, std::fmax(lpr, lqr));
    V3 cr = {pq.y * pr.z - pq.z * pr.y, pq.z * pr.x - pq.x * pr.z, pq.x * pr.y - pq.y * pr.x};
    float thr = (float)(1e-6 * (double)ml * (double)ml);
    float area = 0.5f * std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
    return area < thr;
}

static inline bool tri_has(const Tri& t, int a) { return t.i == a || t.j == a || t.k == a; }

// Rewrite of collapse_edge_kernel (:683-715) / remove_line_edge_collapse (:244-273).
static void rewrite_rings(Ctx& c, int u, int v) {
    for (int a = c.ring_off[u]; a < c.ring_off[u + 1]; ++a) {
        Tri& t = c.T[c.ring[a]];
        if (tri_has(t, v)) t = Tri{-1, -1, -1};
    }
    for (int a = c.ring_off[v]; a < c.ring_off[v + 1]; ++a) {
        Tri& t = c.T[c.ring[a]];
        if (tri_has(t, u)) t = Tri{-1, -1, -1};
        else if (t.i == v) t.i = u;
        else if (t.j == v) t.j = u;
        else if (t.k == v) t.k = u;
    }
}

static void restore_edge(Ctx& c, int u, int v) {
    c.occ[v] = 1;
    c.P[u] = c.P0[u];
    c.P[v] = c.P0[v];
    for (int w : {u, v})
        for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) c.T[c.ring[a]] = c.T0[c.ring[a]];
}

// self_intersect + the flag table: flagged[f] = 1 for every face of an
// intersecting (live, live) pair that involves a changed face.
static void guard(Ctx& c, const std::vector<uint8_t>& changed, float epsilon) {
    c.flagged.assign(c.n_tris, 0);
    Lbvh b;
    build_lbvh(c.P, c.T, b);
    std::vector<int> qlist;
    for (int f = 0; f < c.n_tris; ++f)
        if (changed[f] && c.T[f].i != -1) qlist.push_back(f);
    std::vector<std::vector<int>> hits(qlist.size());
    long long pairs = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : pairs)
    for (size_t qi = 0; qi < qlist.size(); ++qi) {
        int a = qlist[qi];
        const Tri ta = c.T[a];
        Box q = tri_box(c.P[ta.i], c.P[ta.j], c.P[ta.k]);
        std::vector<int> cand;
        lbvh_query(b, q, a, cand);
        float fa[9] = {c.P[ta.i].x, c.P[ta.i].y, c.P[ta.i].z, c.P[ta.j].x, c.P[ta.j].y, c.P[ta.j].z,
                       c.P[ta.k].x, c.P[ta.k].y, c.P[ta.k].z};
        int ia[3] = {ta.i, ta.j, ta.k};
        for (int bf : cand) {
            if (changed[bf] && bf < a) continue;  // evaluated from the other side
            const Tri tb = c.T[bf];
            float fb[9] = {c.P[tb.i].x, c.P[tb.i].y, c.P[tb.i].z, c.P[tb.j].x, c.P[tb.j].y, c.P[tb.j].z,
                           c.P[tb.k].x, c.P[tb.k].y, c.P[tb.k].z};
            int ib[3] = {tb.i, tb.j, tb.k};
            pairs += 2;
            bool hit;
            if (c.pred_mode == 1) {  // exact mode: the exact, adjacency-aware intersection test
                exact::P3 pa[3], pb[3];
                for (int k = 0; k < 3; ++k) {
                    pa[k] = exact::P3{fa[3 * k], fa[3 * k + 1], fa[3 * k + 2]};
                    pb[k] = exact::P3{fb[3 * k], fb[3 * k + 1], fb[3 * k + 2]};
                }
                hit = exact::faces_intersect(ia, pa, ib, pb);
            } else {
                hit = ref_guard(fb, ib, fa, ia, epsilon) || ref_guard(fa, ia, fb, ib, epsilon);
            }
            if (hit) hits[qi].push_back(bf);
        }
    }
    c.guard_pairs += pairs;
    for (size_t qi = 0; qi < qlist.size(); ++qi) {
        if (!hits[qi].empty()) c.flagged[qlist[qi]] = 1;
        for (int bf : hits[qi]) c.flagged[bf] = 1;
    }
}

// get_undo_candidate_kernel (:776-827): returns the pushed edges in record order
// get_undo_candidate_kernel (cusimp_free.cu:776-827) over the first n_cand
// records: its grid is sized from the regular-collapse count read before the
// line collapses (:1004,1016: h_n_collapsed threads rounded up to blocks of
// BLOCK_SIZE = 512), so a line-collapse record past that grid is never a
// candidate, and none is when the round has no regular record.
static int undo_candidate_count(const Ctx& c) {
    const int grid = (c.n_regular + 511) / 512 * 512;
    return std::min((int)c.collapsed.size(), grid);
}

static void undo_candidates(Ctx& c, std::vector<int>& pushed) {
    pushed.clear();
    const int n_cand = undo_candidate_count(c);
    for (int r = 0; r < n_cand; ++r) {
        const int e = c.collapsed[r];
        int u = c.eu[e], v = c.ev[e];
        bool hit = false;
        for (int w : {u, v}) {
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1] && !hit; ++a)
                if (c.flagged[c.ring[a]]) hit = true;
            if (hit) break;
        }
        if (hit) pushed.push_back(e);
    }
}

// One call of CUSimp_Free::forward (cusimp_free.cu:891-1048).
static void forward(Ctx& c, const float* pts, int nP, const int* tris, int nT, const int* undo_in, int n_undo_in,
                    float scale, float thr, bool is_stuck, bool init) {
    const float epsilon = 1e-3f;
    c.n_vertices_undo = 0;
    int first = 0;
    if (is_stuck) {
        // tmp <- incoming list, then mapped through the previous round's pts_map
        for (int i = 0; i < n_undo_in; ++i) c.tmp_undo[i] = undo_in[i];
        c.n_vertices_undo += n_undo_in;
        for (int i = 0; i < c.n_vertices_undo - first; ++i) c.undo_list[first + i] = c.pts_map[c.tmp_undo[i]];
        first += n_undo_in;
    } else if (n_undo_in != 0) {
        std::fill(c.invalid.begin(), c.invalid.end(), 0);
    }
    if (n_undo_in != 0)
        for (int i = 0; i < n_undo_in; ++i)
            if (undo_in[i] >= 0 && (size_t)undo_in[i] < c.invalid.size()) c.invalid[undo_in[i]] = 1;

    const float tres = thr, edge_s = scale;
    const uint32_t collapse_t = sat_u32(tclamp(thr, 0.0f, 10.0f) / 10.0f * 4294967296.0f);
    c.n_pts = nP;
    c.n_tris = nT;
    grow_pts(c, nP);
    c.P.resize(nP);
    for (int v = 0; v < nP; ++v) c.P[v] = V3{pts[3 * v], pts[3 * v + 1], pts[3 * v + 2]};
    c.occ.assign(nP + 1, 1);
    c.occ[nP] = 0;
    c.T.resize(nT);
    if (init) {
        std::vector<int64_t> perm;
        shuffle_perm((uint64_t)nT, perm);
        for (int f = 0; f < nT; ++f) {
            int64_t s = perm[f];
            c.T[f] = Tri{tris[3 * s], tris[3 * s + 1], tris[3 * s + 2]};
        }
    } else {
        for (int f = 0; f < nT; ++f) c.T[f] = Tri{tris[3 * f], tris[3 * f + 1], tris[3 * f + 2]};
    }
    c.P0 = c.P;
    c.T0 = c.T;

    build_rings(c);
    build_edges(c);
    vertex_quadrics(c);
    grow_edges(c, c.n_edges);
    edge_costs(c, is_stuck, edge_s, tres);

    // independent set: min key over the union of both one-rings
    c.tri_min.assign(nT, ~0ull);
    for (int e = 0; e < c.n_edges; ++e) {
        uint64_t k = edge_key(c.edge_cost[e], e);
        for (int w : {c.eu[e], c.ev[e]})
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) {
                uint64_t& m = c.tri_min[c.ring[a]];
                if (k < m) m = k;
            }
    }
    c.collapsed.clear();
    std::vector<char> sel(c.n_edges, 0);
    for (int e = 0; e < c.n_edges; ++e) {
        if (c.edge_cost[e] > collapse_t) continue;
        uint64_t k = edge_key(c.edge_cost[e], e);
        bool ok = true;
        for (int w : {c.eu[e], c.ev[e]}) {
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a)
                if (c.tri_min[c.ring[a]] != k) { ok = false; break; }
            if (!ok) break;
        }
        sel[e] = ok;
    }
    // collapse + skinny undo (selected edges have disjoint ring sets, so the
    // order of application does not change the result)
    for (int e = 0; e < c.n_edges; ++e) {
        if (!sel[e]) continue;
        int u = c.eu[e], v = c.ev[e];
        c.collapsed.push_back(e);
        V3 vm = collapse_point(c, u, v);
        c.P[u] = vm;
        c.P[v] = V3{0, 0, 0};
        c.occ[v] = 0;
        rewrite_rings(c, u, v);
        bool skinny = false;
        for (int w : {u, v}) {
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) {
                const Tri t = c.T[c.ring[a]];
                if (t.i == -1) continue;
                if (is_skinny(c.P[t.i], c.P[t.j], c.P[t.k])) { skinny = true; break; }
            }
            if (skinny) break;
        }
        if (skinny) restore_edge(c, u, v);
    }
    c.n_regular = (int)c.collapsed.size();

    // remove_invalid_faces (:218-232)
    for (int f = 0; f < nT; ++f) {
        Tri& t = c.T[f];
        if (t.i == -1 && t.j == -1 && t.k == -1) continue;
        if (t.i == t.j || t.j == t.k || t.k == t.i) t = Tri{-1, -1, -1};
    }
    // remove_line_edge_collapse (:234-274)
    {
        std::vector<int> line;
        for (int e = 0; e < c.n_edges; ++e) {
            V3 p = c.P[c.eu[e]], q = c.P[c.ev[e]];
            if (p.x == q.x && p.y == q.y && p.z == q.z) line.push_back(e);
        }
        for (int e : line) {
            int u = c.eu[e], v = c.ev[e];
            c.collapsed.push_back(e);
            c.P[v] = V3{0, 0, 0};
            c.occ[v] = 0;
            rewrite_rings(c, u, v);
        }
    }

    // changed faces: the pre-collapse one-rings of every recorded collapse
    std::vector<uint8_t> changed(nT, 0);
    for (int e : c.collapsed)
        for (int w : {c.eu[e], c.ev[e]})
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) changed[c.ring[a]] = 1;

    c.undo_iters_edges.clear();
    std::vector<int> pushed;
    if (undo_candidate_count(c) > 0) {
        guard(c, changed, epsilon);
        undo_candidates(c, pushed);
    }
    c.n_vertices_undo += 2 * (int)pushed.size();
    int it = 0;
    while (!pushed.empty()) {
        ++it;
        c.undo_iters_edges.push_back((int)pushed.size());
        for (size_t i = 0; i < pushed.size(); ++i) {
            c.tmp_undo[2 * i] = c.eu[pushed[i]];
            c.tmp_undo[2 * i + 1] = c.ev[pushed[i]];
        }
        for (int e : pushed) restore_edge(c, c.eu[e], c.ev[e]);
        for (int i = 0; i < c.n_vertices_undo - first; ++i) c.undo_list[first + i] = c.pts_map[c.tmp_undo[i]];
        first += 2 * (int)pushed.size();
        guard(c, changed, epsilon);
        undo_candidates(c, pushed);
        c.n_vertices_undo += 2 * (int)pushed.size();
        if (it == 5) break;
    }
    // pts_map = exclusive scan of occupancy over n_pts + 1 entries
    int acc = 0;
    for (int v = 0; v <= nP; ++v) {
        int o = c.occ[v];
        c.pts_map[v] = acc;
        acc += o;
    }
}

// ---------------------------------------------------------------- legacy PaSP (cusimp.cu)
// compute_edge_cost_kernel of the legacy simplifier, cusimp.cu:283-482: the
// link test, edge-length, skinny and flip terms of edge_costs without the
// invalid / stuck rules, with three differences: the skinny ratio is a
// double expression (`4.0f * sqrt(3)` takes the double root; the squared
// lengths are float squares) cast to float, it is not scaled by 5, and the
// averaged edge length is counted twice (cusimp.cu:378,
// `edge_len += edge_len + |v0 - v1| / s * t`).
static void edge_costs_legacy(Ctx& c, float edge_s, float tres) {
    const double k4s3 = 4.0 * std::sqrt(3.0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int e = 0; e < c.n_edges; ++e) {
        int u = c.eu[e], v = c.ev[e];
        V3 p0 = c.P[u], p1 = c.P[v];
        int dup = 0;
        bool over = false;
        for (int a = c.ring_off[u]; a < c.ring_off[u + 1] && !over; ++a) {
            int su = succ_of(c.T[c.ring[a]], u);
            for (int b = c.ring_off[v]; b < c.ring_off[v + 1]; ++b) {
                int sv = succ_of(c.T[c.ring[b]], v);
                if (su == sv) ++dup;
                if (dup > 2) { over = true; break; }
            }
        }
        if (over) { c.edge_cost[e] = 0xFFFFFFFFu; continue; }
        if (dup != 2) continue;  // stale
        float elen = 0.0f;
        int nedge = 0;
        for (int side = 0; side < 2; ++side) {
            int w = side ? v : u;
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) {
                const Tri t = c.T[c.ring[a]];
                if (t.i > t.j) { elen += vnorm(vsub(c.P[t.i], c.P[t.j])); ++nedge; }
                if (t.j > t.k) { elen += vnorm(vsub(c.P[t.j], c.P[t.k])); ++nedge; }
                if (t.k > t.i) { elen += vnorm(vsub(c.P[t.k], c.P[t.i])); ++nedge; }
            }
        }
        elen = elen / (float)nedge / edge_s * tres;
        const float elen_edge = vnorm(vsub(p0, p1)) / edge_s * tres;
        elen = elen + (elen + elen_edge);
        V3 vm = vdivs(vadd(p0, p1), 2.0f);
        float qa = 0.0f;
        int ntri = 0;
        bool flip = false;
        for (int side = 0; side < 2 && !flip; ++side) {
            int w = side ? v : u, other = side ? u : v;
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) {
                const Tri t = c.T[c.ring[a]];
                if (t.i == other || t.j == other || t.k == other) continue;
                V3 o0 = c.P[t.i], o1 = c.P[t.j], o2 = c.P[t.k];
                V3 n0 = o0, n1 = o1, n2 = o2;
                if (t.i == w) n0 = vm;
                if (t.j == w) n1 = vm;
                if (t.k == w) n2 = vm;
                V3 on = vcross(vsub(o1, o0), vsub(o2, o0));
                on = vdivs(on, vnorm(on));
                V3 nn = vcross(vsub(n1, n0), vsub(n2, n0));
                nn = vdivs(nn, vnorm(nn));
                if (vdot(on, nn) < 0) { flip = true; break; }
                float l01 = vnorm(vsub(n0, n1)), l12 = vnorm(vsub(n1, n2)), l20 = vnorm(vsub(n2, n0));
                float sq01 = l01 * l01, sq12 = l12 * l12, sq20 = l20 * l20;
                double qd = k4s3 * (double)heron(n0, n1, n2) / (double)sq01 + (double)sq12 + (double)sq20 +
                            (double)0.0000001f;
                qa += 1.0f - tclamp((float)qd, 0.0f, 1.0f);
                ++ntri;
            }
        }
        if (flip) { c.edge_cost[e] = 0xFFFFFFFFu; continue; }
        float M[16];
        const float* qu = &c.Q[(size_t)u * 10];
        const float* qv = &c.Q[(size_t)v * 10];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) M[i * 4 + j] = qu[qidx(i, j)] + qv[qidx(i, j)];
        float x[4] = {vm.x, vm.y, vm.z, 1.0f};
        float r[4];
        for (int j = 0; j < 4; ++j) r[j] = M[0 * 4 + j] * x[0] + M[1 * 4 + j] * x[1] + M[2 * 4 + j] * x[2] + M[3 * 4 + j] * x[3];
        float cost = (r[0] * x[0] + r[1] * x[1] + r[2] * x[2] + r[3] * x[3]) / (edge_s * edge_s);
        float tot = cost + elen + qa / (float)ntri * tres;
        c.edge_cost[e] = sat_u32(tclamp(tot, 0.0f, 10.0f) / 10.0f * 4294967296.0f);
    }
}

// CUSimp::forward, cusimp.cu:587-720: load (faces shuffled when init), rings
// in ascending face order, edges, quadrics, legacy costs, per-face minimum
// keys (propagate_edge_cost_kernel :484-509) and the midpoint collapse of
// every edge at or below the threshold holding the minimum of all faces of
// both rings (collapse_edge_kernel :511-585); pts_map = exclusive scan of
// occupancy.
static void pasp_forward(Ctx& c, const float* pts, int nP, const int* tris, int nT, float scale, float thr,
                         bool init) {
    const uint32_t collapse_t = sat_u32(tclamp(thr, 0.0f, 10.0f) / 10.0f * 4294967296.0f);
    c.n_pts = nP;
    c.n_tris = nT;
    grow_pts(c, nP);
    c.P.resize(nP);
    for (int v = 0; v < nP; ++v) c.P[v] = V3{pts[3 * v], pts[3 * v + 1], pts[3 * v + 2]};
    c.occ.assign(nP + 1, 1);
    c.occ[nP] = 0;
    c.T.resize(nT);
    if (init) {
        std::vector<int64_t> perm;
        shuffle_perm((uint64_t)nT, perm);
        for (int f = 0; f < nT; ++f) {
            int64_t s = perm[f];
            c.T[f] = Tri{tris[3 * s], tris[3 * s + 1], tris[3 * s + 2]};
        }
    } else {
        for (int f = 0; f < nT; ++f) c.T[f] = Tri{tris[3 * f], tris[3 * f + 1], tris[3 * f + 2]};
    }
    build_rings(c);
    build_edges(c);
    vertex_quadrics(c);
    grow_edges(c, c.n_edges);
    edge_costs_legacy(c, scale, thr);
    c.tri_min.assign(nT, ~0ull);
    for (int e = 0; e < c.n_edges; ++e) {
        uint64_t k = edge_key(c.edge_cost[e], e);
        for (int w : {c.eu[e], c.ev[e]})
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a) {
                uint64_t& m = c.tri_min[c.ring[a]];
                if (k < m) m = k;
            }
    }
    std::vector<char> sel(c.n_edges, 0);
    for (int e = 0; e < c.n_edges; ++e) {
        if (c.edge_cost[e] > collapse_t) continue;
        uint64_t k = edge_key(c.edge_cost[e], e);
        bool ok = true;
        for (int w : {c.eu[e], c.ev[e]}) {
            for (int a = c.ring_off[w]; a < c.ring_off[w + 1]; ++a)
                if (c.tri_min[c.ring[a]] != k) { ok = false; break; }
            if (!ok) break;
        }
        sel[e] = ok;
    }
    for (int e = 0; e < c.n_edges; ++e) {
        if (!sel[e]) continue;
        int u = c.eu[e], v = c.ev[e];
        c.P[u] = vdivs(vadd(c.P[u], c.P[v]), 2.0f);
        c.P[v] = V3{0, 0, 0};
        c.occ[v] = 0;
        rewrite_rings(c, u, v);
    }
    int acc = 0;
    for (int v = 0; v <= nP; ++v) {
        int o = c.occ[v];
        c.pts_map[v] = acc;
        acc += o;
    }
}

}  // namespace orc

// ---------------------------------------------------------------- C ABI
extern "C" {

void* orc_ctx_create() { return new orc::Ctx(); }

// 0 midpoint (default), 1 optimal placement; kept across calls.
void orc_set_placement(void* ctx, int mode) { ((orc::Ctx*)ctx)->placement = mode; }

// place_optimal on a caller's M (10 floats) and edge end points.
void orc_place_optimal(const float* M10, const float* p0, const float* p1, float* out3) {
    orc::V3 r = orc::place_optimal(M10, orc::V3{p0[0], p0[1], p0[2]}, orc::V3{p1[0], p1[1], p1[2]});
    out3[0] = r.x;
    out3[1] = r.y;
    out3[2] = r.z;
}
void orc_ctx_destroy(void* p) { delete (orc::Ctx*)p; }

int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_shuffle_perm(long long m, long long* perm) {
    std::vector<int64_t> p;
    orc::shuffle_perm((uint64_t)m, p);
    for (long long i = 0; i < m; ++i) perm[i] = p[i];
}

// n-th output (1-based) of minstd_rand seeded with 1 (the Feistel key stream).
unsigned int orc_minstd(long long n) {
    unsigned long long s = 1;
    for (long long i = 0; i < n; ++i) s = (s * 48271ull) % 2147483647ull;
    return (unsigned int)s;
}

unsigned int orc_morton3(float x, float y, float z) { return orc::morton3(x, y, z); }

unsigned long long orc_feistel(long long m, unsigned long long idx) {
    orc::Feistel f((uint64_t)m);
    return f(idx);
}

// One round, mirroring CUDSP_Free.forward (pybind.cpp:52-103).  Outputs:
// out_pts (nP*3), out_tris (nT*3), out_occ (nP), out_map (nP), out_undo (cap).
// Returns the number of undo-list entries, or -1 if undo_cap is too small.
int orc_forward(void* ctx, const float* pts, int nP, const int* tris, int nT, const int* undo_in, int n_undo_in,
                float scale, float thr, int is_stuck, int init, int pred_mode, float* out_pts, int* out_tris,
                int* out_occ, int* out_map, int* out_undo, int undo_cap) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    c.pred_mode = pred_mode;
    orc::forward(c, pts, nP, tris, nT, undo_in, n_undo_in, scale, thr, is_stuck != 0, init != 0);
    for (int v = 0; v < nP; ++v) {
        out_pts[3 * v] = c.P[v].x;
        out_pts[3 * v + 1] = c.P[v].y;
        out_pts[3 * v + 2] = c.P[v].z;
        out_occ[v] = c.occ[v];
        out_map[v] = c.pts_map[v];
    }
    for (int f = 0; f < nT; ++f) {
        out_tris[3 * f] = c.T[f].i;
        out_tris[3 * f + 1] = c.T[f].j;
        out_tris[3 * f + 2] = c.T[f].k;
    }
    if (c.n_vertices_undo > undo_cap) return -1;
    for (int i = 0; i < c.n_vertices_undo; ++i) out_undo[i] = c.undo_list[i];
    return c.n_vertices_undo;
}

// Debug accessors of the last round.
int orc_n_edges(void* ctx) { return ((orc::Ctx*)ctx)->n_edges; }
void orc_get_edges(void* ctx, int* uv) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    for (int e = 0; e < c.n_edges; ++e) { uv[2 * e] = c.eu[e]; uv[2 * e + 1] = c.ev[e]; }
}
void orc_get_edge_cost(void* ctx, unsigned int* out) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    for (int e = 0; e < c.n_edges; ++e) out[e] = c.edge_cost[e];
}
void orc_get_quadrics(void* ctx, float* out) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    std::memcpy(out, c.Q.data(), sizeof(float) * c.Q.size());
}
int orc_get_ring(void* ctx, int* off, int* ring) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    if (off) std::memcpy(off, c.ring_off.data(), sizeof(int) * (c.n_pts + 1));
    if (ring) std::memcpy(ring, c.ring.data(), sizeof(int) * c.ring.size());
    return (int)c.ring.size();
}
int orc_get_collapsed(void* ctx, int* out, int* n_regular) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    if (out) std::memcpy(out, c.collapsed.data(), sizeof(int) * c.collapsed.size());
    if (n_regular) *n_regular = c.n_regular;
    return (int)c.collapsed.size();
}
int orc_get_undo_iters(void* ctx, int* out) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    if (out) std::memcpy(out, c.undo_iters_edges.data(), sizeof(int) * c.undo_iters_edges.size());
    return (int)c.undo_iters_edges.size();
}
long long orc_guard_pairs(void* ctx) { return ((orc::Ctx*)ctx)->guard_pairs; }

// LBVH of the live faces of (pts, tris): codes[s], faces[s] in sorted order,
// internal children left/right (n-1 each).  Returns n (live faces).
int orc_lbvh(const float* pts, int nP, const int* tris, int nT, unsigned int* codes, int* faces, int* left,
             int* right) {
    std::vector<orc::V3> P(nP);
    for (int v = 0; v < nP; ++v) P[v] = orc::V3{pts[3 * v], pts[3 * v + 1], pts[3 * v + 2]};
    std::vector<orc::Tri> T(nT);
    for (int f = 0; f < nT; ++f) T[f] = orc::Tri{tris[3 * f], tris[3 * f + 1], tris[3 * f + 2]};
    orc::Lbvh b;
    orc::build_lbvh(P, T, b);
    for (int s = 0; s < b.n; ++s) {
        if (codes) codes[s] = b.code[s];
        if (faces) faces[s] = b.leaf_face[s];
    }
    for (int i = 0; i + 1 < b.n; ++i) {
        if (left) left[i] = b.left[i];
        if (right) right[i] = b.right[i];
    }
    return b.n;
}

// Karras children of the n leaves with sorted keys (codes[s] << 32 | faces[s])
// (the hierarchy step of build_lbvh, on keys given by the caller).
void orc_karras(const unsigned int* codes, const int* faces, int n, int* left, int* right) {
    std::vector<uint64_t> k(n);
    for (int s = 0; s < n; ++s) k[s] = ((uint64_t)codes[s] << 32) | (uint64_t)(uint32_t)faces[s];
    for (int i = 0; i + 1 < n; ++i) orc::karras_node(k, n, i, left[i], right[i]);
}

// Candidate pairs a<b (live, closed-AABB overlap) by LBVH traversal.  Returns
// the count; writes up to cap pairs.
long long orc_candidates(const float* pts, int nP, const int* tris, int nT, int* pairs, long long cap) {
    std::vector<orc::V3> P(nP);
    for (int v = 0; v < nP; ++v) P[v] = orc::V3{pts[3 * v], pts[3 * v + 1], pts[3 * v + 2]};
    std::vector<orc::Tri> T(nT);
    for (int f = 0; f < nT; ++f) T[f] = orc::Tri{tris[3 * f], tris[3 * f + 1], tris[3 * f + 2]};
    orc::Lbvh b;
    orc::build_lbvh(P, T, b);
    long long n = 0;
    std::vector<int> out;
    for (int f = 0; f < nT; ++f) {
        if (T[f].i == -1) continue;
        orc::Box q = orc::tri_box(P[T[f].i], P[T[f].j], P[T[f].k]);
        orc::lbvh_query(b, q, f, out);
        std::sort(out.begin(), out.end());
        for (int g : out) {
            if (g <= f) continue;
            if (n < cap) { pairs[2 * n] = f; pairs[2 * n + 1] = g; }
            ++n;
        }
    }
    return n;
}

// The reference guard for one ordered pair (current face first).
int orc_ref_guard(const float* cur9, const int* ci, const float* qry9, const int* qi, float epsilon) {
    return orc::ref_guard(cur9, ci, qry9, qi, epsilon) ? 1 : 0;
}

// The PaMO.run stage-2 loop (pamo/__init__.py:147-185) over already-centred
// vertices.  Outputs are written into caller buffers of input size.
// Returns the number of rounds; hist (cap entries) receives the face count
// after each round.
int orc_run(void* ctx, const float* verts, int nV, const int* faces, int nF, double ratio, int tolerance, float thr,
            long long iters, long long min_verts, int pred_mode, float* out_v, int* out_nv, int* out_f, int* out_nf,
            int* hist, int hist_cap) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    c.pred_mode = pred_mode;
    long long target = (long long)(ratio * (double)nF);
    if (min_verts > target) target = min_verts;
    std::vector<float> V(verts, verts + 3 * (size_t)nV);
    std::vector<int> F(faces, faces + 3 * (size_t)nF);
    std::vector<int> undo;
    int count = 0;
    bool is_stuck = false;
    float ext[3];
    for (int d = 0; d < 3; ++d) {
        float mn = std::numeric_limits<float>::infinity(), mx = -mn;
        for (int v = 0; v < nV; ++v) { mn = std::fmin(mn, V[3 * v + d]); mx = std::fmax(mx, V[3 * v + d]); }
        ext[d] = mx - mn;
    }
    float scale = ext[0];
    if (ext[1] > scale) scale = ext[1];
    if (ext[2] > scale) scale = ext[2];
    bool init = true;
    int rounds = 0;
    std::vector<float> op;
    std::vector<int> ot, oo, om, ou;
    for (long long it = 0; it < iters; ++it) {
        int nv = (int)(V.size() / 3), nf = (int)(F.size() / 3);
        int prev = nf;
        orc::forward(c, V.data(), nv, F.data(), nf, undo.data(), (int)undo.size(), scale, thr, is_stuck, init);
        init = false;
        ++rounds;
        undo.assign(c.undo_list.begin(), c.undo_list.begin() + c.n_vertices_undo);
        std::vector<float> V2;
        V2.reserve(V.size());
        for (int v = 0; v < nv; ++v)
            if (c.occ[v]) { V2.push_back(c.P[v].x); V2.push_back(c.P[v].y); V2.push_back(c.P[v].z); }
        std::vector<int> F2;
        F2.reserve(F.size());
        for (int f = 0; f < nf; ++f) {
            if (c.T[f].i < 0) continue;
            F2.push_back(c.pts_map[c.T[f].i]);
            F2.push_back(c.pts_map[c.T[f].j]);
            F2.push_back(c.pts_map[c.T[f].k]);
        }
        V.swap(V2);
        F.swap(F2);
        int cur = (int)(F.size() / 3);
        if (rounds <= hist_cap && hist) hist[rounds - 1] = cur;
        if (cur <= target || cur <= 10) break;
        if (cur == prev) ++count;
        else { count = 0; is_stuck = false; }
        if (count >= 2) is_stuck = true;
        if (count == tolerance) break;
    }
    std::memcpy(out_v, V.data(), sizeof(float) * V.size());
    std::memcpy(out_f, F.data(), sizeof(int) * F.size());
    *out_nv = (int)(V.size() / 3);
    *out_nf = (int)(F.size() / 3);
    return rounds;
}

// cusimp::CUDSP.forward (pybind.cpp:140-183) on the oracle.
void orc_pasp_forward(void* ctx, const float* pts, int nP, const int* tris, int nT, float scale, float thr, int init,
                      float* out_pts, int* out_tris, int* out_occ, int* out_map) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    orc::pasp_forward(c, pts, nP, tris, nT, scale, thr, init != 0);
    for (int v = 0; v < nP; ++v) {
        out_pts[3 * v] = c.P[v].x;
        out_pts[3 * v + 1] = c.P[v].y;
        out_pts[3 * v + 2] = c.P[v].z;
        out_occ[v] = c.occ[v];
        out_map[v] = c.pts_map[v];
    }
    for (int f = 0; f < nT; ++f) {
        out_tris[3 * f] = c.T[f].i;
        out_tris[3 * f + 1] = c.T[f].j;
        out_tris[3 * f + 2] = c.T[f].k;
    }
}

// PaSP.run (pamo/__init__.py:224-253): scale = largest extent, rounds with
// compaction until one removes no face.  Returns the number of rounds.
int orc_pasp_run(void* ctx, const float* verts, int nV, const int* faces, int nF, float thr, long long iters,
                 float* out_v, int* out_nv, int* out_f, int* out_nf) {
    orc::Ctx& c = *(orc::Ctx*)ctx;
    std::vector<float> V(verts, verts + 3 * (size_t)nV);
    std::vector<int> F(faces, faces + 3 * (size_t)nF);
    float ext[3];
    for (int d = 0; d < 3; ++d) {
        float mn = std::numeric_limits<float>::infinity(), mx = -mn;
        for (int v = 0; v < nV; ++v) { mn = std::fmin(mn, V[3 * v + d]); mx = std::fmax(mx, V[3 * v + d]); }
        ext[d] = mx - mn;
    }
    float scale = ext[0];
    if (ext[1] > scale) scale = ext[1];
    if (ext[2] > scale) scale = ext[2];
    int rounds = 0;
    for (long long it = 0; it < iters; ++it) {
        int nv = (int)(V.size() / 3), nf = (int)(F.size() / 3);
        orc::pasp_forward(c, V.data(), nv, F.data(), nf, scale, thr, it == 0);
        ++rounds;
        std::vector<float> V2;
        V2.reserve(V.size());
        for (int v = 0; v < nv; ++v)
            if (c.occ[v]) { V2.push_back(c.P[v].x); V2.push_back(c.P[v].y); V2.push_back(c.P[v].z); }
        std::vector<int> F2;
        F2.reserve(F.size());
        for (int f = 0; f < nf; ++f) {
            if (c.T[f].i < 0) continue;
            F2.push_back(c.pts_map[c.T[f].i]);
            F2.push_back(c.pts_map[c.T[f].j]);
            F2.push_back(c.pts_map[c.T[f].k]);
        }
        V.swap(V2);
        F.swap(F2);
        if ((int)(F.size() / 3) == nf) break;
    }
    std::memcpy(out_v, V.data(), sizeof(float) * V.size());
    std::memcpy(out_f, F.data(), sizeof(int) * F.size());
    *out_nv = (int)(V.size() / 3);
    *out_nf = (int)(F.size() / 3);
    return rounds;
}

}  // extern "C"
