// Stage-1 remeshing on gfx950: the input triangles become the iso-surface
// of an offset distance field, as PaMO.remesh does (simp_cuda/pamo/
// __init__.py:60-120: normalise the triangles into the unit cube, a
// narrow-band distance field on an R^3 grid, the level 0.9 / R, dual
// marching cubes and quads split into triangles).  The reference takes the
// field from torchcumesh2sdf.get_sdf and the extraction from pdmc.DMC,
// neither of which ships with it; the field, its sign and the extraction
// here are restated from their published methods:
//   - unsigned distance to the triangles (closest point on a triangle,
//     Ericson, Real-Time Collision Detection 5.1.5) at every grid sample
//     within the band, by rasterising each triangle column by column along
//     its dominant normal axis;
//   - sign by a flood fill from the grid border through samples farther
//     than half a cell from the surface (a 6-connected path that crosses
//     the surface passes a sample within half a cell of it), so enclosed
//     regions are inside and open sheets get a closed shell;
//   - marching cubes with one face rule for all 256 corner-sign cases (on a
//     face with alternating corner signs the inside corners are cut off
//     separately): the surface patches of a cell are the cycles of its
//     crossing edges joined face by face, each polygon is fanned into
//     triangles, and the vertices are the crossings on the grid edges,
//     shared by the cells around an edge.  Two cells sharing a face always
//     cut it the same way, so the surface is closed and manifold; the
//     reference's dual extraction (one vertex per patch, quads split by an
//     angle rule) is not manifold where the patches of both cells cross the
//     shared face twice, and is not used.
// Grid sample k of an axis sits at (k + 0.5) / R of the normalised cube.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hip_common.h"
#include "pamo_hip.h"
#include "prim.h"
#include "stage3_math.h"

namespace pamo {
namespace s1 {
namespace {

using s3::V3;
using s3::mk;
using s3::v3;

constexpr int kB = 256;

// ------------------------------------------------------------ dual marching cubes case table
// Cube corner c = x + 2 y + 4 z; edges 0-3 along x, 4-7 along y, 8-11 along z.
constexpr int kEdgeCorner[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                    {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// The six faces, corners in cyclic order.
constexpr int kFaceCorner[6][4] = {{0, 2, 6, 4}, {1, 3, 7, 5}, {0, 1, 5, 4},
                                   {2, 3, 7, 6}, {0, 1, 3, 2}, {4, 5, 7, 6}};

struct CaseTable {
    int8_t cyc[256][12];  // the crossing edges, patch after patch, each patch in cycle order
    int8_t plen[256][4];  // edges of each patch
    int8_t npatch[256];
    int8_t ntri[256];     // fan triangles: sum over the patches of (edges - 2)
};

static int edge_of(int a, int b) {
    for (int e = 0; e < 12; ++e)
        if ((kEdgeCorner[e][0] == a && kEdgeCorner[e][1] == b) || (kEdgeCorner[e][0] == b && kEdgeCorner[e][1] == a))
            return e;
    return -1;
}

// Outward normal of face f (x = 0, x = 1, y = 0, y = 1, z = 0, z = 1).
constexpr int kFaceNormal[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

struct Seg {
    int a, b, face;
    double in[3];  // the inside corner(s) the segment cuts off on its face (centroid)
};

// Bit c of `cs` set: corner c is inside (below the level).  The patches are
// the components of the crossing edges linked by the face segments; each is
// ordered so that, seen from outside the cube across every face it crosses,
// the inside corners lie to the right of its segments, which makes the
// fan triangles' normals point from inside to outside.
static bool build_case(int cs, CaseTable& t) {
    auto in = [&](int c) { return (cs >> c) & 1; };
    auto cpos = [](int c, double p[3]) {
        p[0] = c & 1;
        p[1] = (c >> 1) & 1;
        p[2] = c >> 2;
    };
    std::vector<Seg> segs;
    for (int f = 0; f < 6; ++f) {
        const int* q = kFaceCorner[f];
        int fe[4];
        bool cross[4];
        int ncross = 0;
        for (int i = 0; i < 4; ++i) {
            fe[i] = edge_of(q[i], q[(i + 1) & 3]);
            cross[i] = in(q[i]) != in(q[(i + 1) & 3]);
            ncross += cross[i] ? 1 : 0;
        }
        if (ncross == 2) {
            Seg s{-1, -1, f, {0, 0, 0}};
            for (int i = 0; i < 4; ++i)
                if (cross[i]) (s.a < 0 ? s.a : s.b) = fe[i];
            int nin = 0;
            for (int i = 0; i < 4; ++i)
                if (in(q[i])) {
                    double p[3];
                    cpos(q[i], p);
                    for (int d = 0; d < 3; ++d) s.in[d] += p[d];
                    ++nin;
                }
            for (int d = 0; d < 3; ++d) s.in[d] /= nin;
            segs.push_back(s);
        } else if (ncross == 4) {
            // alternating signs: cut each inside corner off on its own
            for (int i = 0; i < 4; ++i)
                if (in(q[i])) {
                    Seg s{fe[(i + 3) & 3], fe[i], f, {0, 0, 0}};
                    cpos(q[i], s.in);
                    segs.push_back(s);
                }
        }
    }
    // each crossing edge lies on two faces: two segments
    std::vector<std::vector<int>> at(12);
    for (int k = 0; k < (int)segs.size(); ++k) {
        at[segs[k].a].push_back(k);
        at[segs[k].b].push_back(k);
    }
    auto mid = [&](int e, double p[3]) {
        double a[3], b[3];
        cpos(kEdgeCorner[e][0], a);
        cpos(kEdgeCorner[e][1], b);
        for (int d = 0; d < 3; ++d) p[d] = 0.5 * (a[d] + b[d]);
    };
    // does going from edge u to edge v along segment k keep the inside on the right?
    auto forward = [&](int k, int u, int v) {
        double a[3], b[3];
        mid(u, a);
        mid(v, b);
        const int* n = kFaceNormal[segs[k].face];
        const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double c[3] = {d[1] * n[2] - d[2] * n[1], d[2] * n[0] - d[0] * n[2], d[0] * n[1] - d[1] * n[0]};
        double w = 0;
        for (int j = 0; j < 3; ++j) w += c[j] * (segs[k].in[j] - 0.5 * (a[j] + b[j]));
        return w > 0;
    };
    bool used[12] = {false};
    int pos = 0, np = 0, ntri = 0;
    for (int e = 0; e < 12; ++e) {
        const bool crossing = in(kEdgeCorner[e][0]) != in(kEdgeCorner[e][1]);
        if (!crossing || used[e]) continue;
        if (at[e].size() != 2) return false;
        // the direction out of e that keeps the inside on the right
        int k0 = at[e][0];
        int nxt = segs[k0].a == e ? segs[k0].b : segs[k0].a;
        if (!forward(k0, e, nxt)) {
            k0 = at[e][1];
            nxt = segs[k0].a == e ? segs[k0].b : segs[k0].a;
            if (!forward(k0, e, nxt)) return false;
        }
        int len = 0, cur = e, k = k0;
        for (;;) {
            t.cyc[cs][pos + len++] = (int8_t)cur;
            used[cur] = true;
            const int n2 = segs[k].a == cur ? segs[k].b : segs[k].a;
            if (!forward(k, cur, n2)) return false;
            if (n2 == e) break;
            if (used[n2] || at[n2].size() != 2) return false;
            k = at[n2][0] == k ? at[n2][1] : at[n2][0];
            cur = n2;
        }
        if (np >= 4 || len < 3) return false;
        t.plen[cs][np++] = (int8_t)len;
        pos += len;
        ntri += len - 2;
    }
    for (int j = pos; j < 12; ++j) t.cyc[cs][j] = -1;
    for (int j = np; j < 4; ++j) t.plen[cs][j] = 0;
    t.npatch[cs] = (int8_t)np;
    t.ntri[cs] = (int8_t)ntri;
    return true;
}

static const CaseTable* case_table() {
    static CaseTable t;
    static const bool ok = [] {
        for (int cs = 0; cs < 256; ++cs)
            if (!build_case(cs, t)) return false;
        return true;
    }();
    return ok ? &t : nullptr;
}

// ------------------------------------------------------------ normalisation
// preprocess_mesh (pamo/__init__.py:60-72) on the device, in the order of
// the reference's float32 numpy expressions: t1 = p - mean; the per-axis
// minimum of t1 over the corners; the maximum of t1 - min over all; then
// ((t1 - min) / max + band) / margin.
__device__ __forceinline__ unsigned f2ord(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void __launch_bounds__(kB) bounds_k(const float* __restrict__ p, const int* __restrict__ f, long long n3,
                                               float mx, float my, float mz, unsigned* __restrict__ out6) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const int v = f[i];
    const float t[3] = {p[3 * (long long)v] - mx, p[3 * (long long)v + 1] - my, p[3 * (long long)v + 2] - mz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        atomicMin(&out6[d], f2ord(t[d]));
        atomicMax(&out6[3 + d], f2ord(t[d]));
    }
}

__global__ void __launch_bounds__(kB) normalize_k(const float* __restrict__ p, const int* __restrict__ f, long long n3,
                                                  float mx, float my, float mz, float lx, float ly, float lz,
                                                  float tmax, float band, float margin, float4* __restrict__ tri) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const int v = f[i];
    const float t0 = p[3 * (long long)v] - mx, t1 = p[3 * (long long)v + 1] - my, t2 = p[3 * (long long)v + 2] - mz;
    tri[i] = make_float4(((t0 - lx) / tmax + band) / margin, ((t1 - ly) / tmax + band) / margin,
                         ((t2 - lz) / tmax + band) / margin, 0.0f);
}

// ------------------------------------------------------------ unsigned distance in the band
struct Raster {
    int axis;        // dominant normal axis (the column direction)
    int lo[3], hi[3];  // sample index box of the triangle grown by the band
};

__device__ __forceinline__ float comp(V3 a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }

__device__ __forceinline__ Raster raster_of(V3 a, V3 b, V3 c, int R, float band) {
    Raster r;
    const V3 n = s3::cross(b - a, c - a);
    const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    r.axis = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    for (int d = 0; d < 3; ++d) {
        const float mn = fminf(comp(a, d), fminf(comp(b, d), comp(c, d))) - band;
        const float mxv = fmaxf(comp(a, d), fmaxf(comp(b, d), comp(c, d))) + band;
        r.lo[d] = max(0, (int)ceilf(mn * (float)R - 0.5f));
        r.hi[d] = min(R - 1, (int)floorf(mxv * (float)R - 0.5f));
    }
    return r;
}

__device__ __forceinline__ V3 tri_corner(const float4* __restrict__ tri, long long f, int k) {
    return v3(tri[3 * f + k]);
}

// Columns (sample lines along the dominant axis) of each triangle's box.
__global__ void __launch_bounds__(kB) columns_k(const float4* __restrict__ tri, int nF, int R, float band,
                                                long long* __restrict__ cols) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > nF) return;
    if (f == nF) {
        cols[nF] = 0;
        return;
    }
    const Raster r = raster_of(tri_corner(tri, f, 0), tri_corner(tri, f, 1), tri_corner(tri, f, 2), R, band);
    long long n = 1;
    for (int d = 0; d < 3; ++d) {
        if (d == r.axis) continue;
        const int w = r.hi[d] - r.lo[d] + 1;
        n *= w > 0 ? w : 0;
    }
    cols[f] = (r.hi[r.axis] >= r.lo[r.axis]) ? n : 0;
}

// Exclusive scan of 64-bit column counts in one workgroup per 2048 items
// and a host-side add of the tile sums (the counts are read once per
// remesh, not per round).
__global__ void __launch_bounds__(kB) scan64_tiles_k(long long* __restrict__ v, long long n,
                                                     long long* __restrict__ sums) {
    __shared__ long long lds[kB];
    const long long base = (long long)blockIdx.x * 2048 + (long long)threadIdx.x * 8;
    long long x[8], s = 0;
    for (int i = 0; i < 8; ++i) {
        x[i] = base + i < n ? v[base + i] : 0;
        s += x[i];
    }
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < kB; d <<= 1) {
        const long long y = threadIdx.x >= (unsigned)d ? lds[threadIdx.x - d] : 0;
        __syncthreads();
        lds[threadIdx.x] += y;
        __syncthreads();
    }
    long long ex = lds[threadIdx.x] - s;
    for (int i = 0; i < 8; ++i) {
        if (base + i < n) v[base + i] = ex;
        ex += x[i];
    }
    if (threadIdx.x == kB - 1) sums[blockIdx.x] = lds[kB - 1];
}

__global__ void __launch_bounds__(kB) scan64_add_k(long long* __restrict__ v, long long n,
                                                   const long long* __restrict__ offs) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] += offs[i / 2048];
}

// One work item per column: the samples of the column within the band of
// the triangle's plane, each given min(current, distance) (the distance's
// bits order like the non-negative floats they encode).
__global__ void __launch_bounds__(kB) udf_k(const float4* __restrict__ tri, int nF, const long long* __restrict__ off,
                                            long long total, int R, float band, unsigned* __restrict__ grid) {
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= total) return;
    int lo = 0, hi = nF - 1;  // the triangle whose column range holds w
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    const int f = lo;
    const V3 a = tri_corner(tri, f, 0), b = tri_corner(tri, f, 1), c = tri_corner(tri, f, 2);
    const Raster r = raster_of(a, b, c, R, band);
    const int ax = r.axis, d1 = ax == 0 ? 1 : 0, d2 = ax == 2 ? 1 : 2;
    const long long k = w - off[f];
    const int w1 = r.hi[d1] - r.lo[d1] + 1;
    const int i1 = r.lo[d1] + (int)(k % w1), i2 = r.lo[d2] + (int)(k / w1);
    const float inv = 1.0f / (float)R;
    const float x1 = ((float)i1 + 0.5f) * inv, x2 = ((float)i2 + 0.5f) * inv;
    // the samples of this column within `band` of the plane
    int k0 = r.lo[ax], k1 = r.hi[ax];
    const V3 n = s3::cross(b - a, c - a);
    const float na = comp(n, ax), nl = s3::len(n);
    if (fabsf(na) > 1e-30f && nl > 0.0f) {
        const float ap = comp(a, ax) - (comp(n, d1) * (x1 - comp(a, d1)) + comp(n, d2) * (x2 - comp(a, d2))) / na;
        const float half = band * nl / fabsf(na) * 1.0001f + 1e-6f;
        k0 = max(k0, (int)ceilf((ap - half) * (float)R - 0.5f));
        k1 = min(k1, (int)floorf((ap + half) * (float)R - 0.5f));
    }
    for (int ka = k0; ka <= k1; ++ka) {
        const float xa = ((float)ka + 0.5f) * inv;
        V3 p;
        if (ax == 0) p = mk(xa, x1, x2);
        else if (ax == 1) p = mk(x1, xa, x2);
        else p = mk(x1, x2, xa);
        const V3 q = s3::closest_on_tri(p, a, b, c);
        const float dist = s3::len(p - q);
        if (!(dist <= band)) continue;
        int idx[3];
        idx[ax] = ka;
        idx[d1] = i1;
        idx[d2] = i2;
        unsigned* cell = grid + ((long long)idx[2] * R + idx[1]) * R + idx[0];
        const unsigned bits = __float_as_uint(dist);
        if (bits < *cell) atomicMin(cell, bits);
    }
}

// ------------------------------------------------------------ sign by flood fill
// A sample is passable when it lies farther than `wall` from the surface;
// the outside is the passable region connected (6-neighbours) to the grid
// border.
__global__ void __launch_bounds__(kB) fill_init_k(const unsigned* __restrict__ grid, int R, unsigned wall_bits,
                                                  uint8_t* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)R * R * R;
    if (i >= n) return;
    const int x = (int)(i % R), y = (int)((i / R) % R), z = (int)(i / ((long long)R * R));
    const bool border = x == 0 || y == 0 || z == 0 || x == R - 1 || y == R - 1 || z == R - 1;
    out[i] = (border && grid[i] > wall_bits) ? 1 : 0;
}

// One workgroup per 8x8x8 brick: the brick's flags and the outside flags of
// its face neighbours are brought to LDS, the fill runs inside the brick
// until it stops changing, and *changed records a new outside sample.
constexpr int kBrick = 8;

__global__ void __launch_bounds__(512) fill_brick_k(const unsigned* __restrict__ grid, int R, unsigned wall_bits,
                                                    uint8_t* __restrict__ out, int* __restrict__ changed) {
    __shared__ uint8_t o[kBrick + 2][kBrick + 2][kBrick + 2];
    __shared__ int any, grew;
    const int nb = (R + kBrick - 1) / kBrick;
    const int bx = blockIdx.x % nb, by = (blockIdx.x / nb) % nb, bz = blockIdx.x / (nb * nb);
    const int tx = threadIdx.x & 7, ty = (threadIdx.x >> 3) & 7, tz = threadIdx.x >> 6;
    const int x = bx * kBrick + tx, y = by * kBrick + ty, z = bz * kBrick + tz;
    const bool valid = x < R && y < R && z < R;
    const long long i = valid ? ((long long)z * R + y) * R + x : 0;
    const bool pass = valid && grid[i] > wall_bits;
    // halo: the outside flags around the brick (zero outside the grid)
    for (int h = threadIdx.x; h < (kBrick + 2) * (kBrick + 2) * (kBrick + 2); h += 512) {
        const int hx = h % (kBrick + 2), hy = (h / (kBrick + 2)) % (kBrick + 2), hz = h / ((kBrick + 2) * (kBrick + 2));
        const int gx = bx * kBrick + hx - 1, gy = by * kBrick + hy - 1, gz = bz * kBrick + hz - 1;
        uint8_t v = 0;
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < R && gy < R && gz < R) v = out[((long long)gz * R + gy) * R + gx];
        o[hz][hy][hx] = v;
    }
    if (threadIdx.x == 0) grew = 0;
    __syncthreads();
    const uint8_t start = valid ? o[tz + 1][ty + 1][tx + 1] : 1;
    for (int it = 0; it < kBrick * kBrick * kBrick; ++it) {
        if (threadIdx.x == 0) any = 0;
        __syncthreads();
        if (pass && !o[tz + 1][ty + 1][tx + 1] &&
            (o[tz + 1][ty + 1][tx] | o[tz + 1][ty + 1][tx + 2] | o[tz + 1][ty][tx + 1] | o[tz + 1][ty + 2][tx + 1] |
             o[tz][ty + 1][tx + 1] | o[tz + 2][ty + 1][tx + 1])) {
            o[tz + 1][ty + 1][tx + 1] = 1;
            any = 1;
        }
        __syncthreads();
        if (!any) break;
    }
    if (valid && o[tz + 1][ty + 1][tx + 1] && !start) {
        out[i] = 1;
        grew = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0 && grew) atomicOr(changed, 1);
}

// ------------------------------------------------------------ marching cubes
__device__ __forceinline__ float sdf_at(const unsigned* __restrict__ grid, const uint8_t* __restrict__ out,
                                        long long i) {
    const float u = __uint_as_float(grid[i]);
    return out[i] ? u : -u;
}

// Crossing grid edges, 3 per sample (+x, +y, +z, edge dir * R^3 + sample):
// one surface vertex each.
__global__ void __launch_bounds__(kB) edge_count_k(const unsigned* __restrict__ grid, const uint8_t* __restrict__ out,
                                                   int R, float iso, int* __restrict__ count) {
    const long long n = (long long)R * R * R;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > 3 * n) return;
    if (i == 3 * n) {
        count[i] = 0;
        return;
    }
    const int dir = (int)(i / n);
    const long long s = i % n;
    const int q[3] = {(int)(s % R), (int)((s / R) % R), (int)(s / ((long long)R * R))};
    int c = 0;
    if (q[dir] + 1 < R) {
        const long long step = dir == 0 ? 1 : (dir == 1 ? R : (long long)R * R);
        c = (sdf_at(grid, out, s) < iso) != (sdf_at(grid, out, s + step) < iso) ? 1 : 0;
    }
    count[i] = c;
}

// The crossing of each crossing edge, by linear interpolation, in sample
// index units.
__global__ void __launch_bounds__(kB) edge_verts_k(const unsigned* __restrict__ grid, const uint8_t* __restrict__ out,
                                                   int R, float iso, const int* __restrict__ vpos,
                                                   float4* __restrict__ verts) {
    const long long n = (long long)R * R * R;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n || vpos[i + 1] == vpos[i]) return;
    const int dir = (int)(i / n);
    const long long s = i % n;
    const long long step = dir == 0 ? 1 : (dir == 1 ? R : (long long)R * R);
    const float d0 = sdf_at(grid, out, s), d1 = sdf_at(grid, out, s + step);
    const float t = (iso - d0) / (d1 - d0);
    float p[3] = {(float)(s % R), (float)((s / R) % R), (float)(s / ((long long)R * R))};
    p[dir] += t;
    verts[vpos[i]] = make_float4(p[0], p[1], p[2], 0.0f);
}

struct McTable {
    int8_t cyc[256][12];
    int8_t plen[256][4];
    int8_t npatch[256];
    int8_t ntri[256];
};

__device__ __forceinline__ int cell_case(const unsigned* __restrict__ grid, const uint8_t* __restrict__ out, int R,
                                         int x, int y, int z, float iso) {
    int cs = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const long long i = ((long long)(z + (c >> 2)) * R + (y + ((c >> 1) & 1))) * R + (x + (c & 1));
        cs |= (sdf_at(grid, out, i) < iso ? 1 : 0) << c;
    }
    return cs;
}

__global__ void __launch_bounds__(kB) cell_count_k(const unsigned* __restrict__ grid, const uint8_t* __restrict__ out,
                                                   int R, float iso, const McTable* __restrict__ tab,
                                                   int* __restrict__ count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long m = (long long)(R - 1);
    if (i > m * m * m) return;
    if (i == m * m * m) {
        count[i] = 0;
        return;
    }
    const int x = (int)(i % m), y = (int)((i / m) % m), z = (int)(i / (m * m));
    count[i] = tab->ntri[cell_case(grid, out, R, x, y, z, iso)];
}

// The fan triangles of each patch polygon, vertices by crossing edge.
__global__ void __launch_bounds__(kB) cell_tris_k(const unsigned* __restrict__ grid, const uint8_t* __restrict__ out,
                                                  int R, float iso, const McTable* __restrict__ tab,
                                                  const int* __restrict__ tpos, const int* __restrict__ vpos,
                                                  int* __restrict__ tris) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long m = (long long)(R - 1);
    if (i >= m * m * m || tpos[i + 1] == tpos[i]) return;
    const int x = (int)(i % m), y = (int)((i / m) % m), z = (int)(i / (m * m));
    const int cs = cell_case(grid, out, R, x, y, z, iso);
    const long long n = (long long)R * R * R;
    auto vid = [&](int e) {
        const int c0 = kEdgeCorner[e][0];
        const long long s = ((long long)(z + (c0 >> 2)) * R + (y + ((c0 >> 1) & 1))) * R + (x + (c0 & 1));
        return vpos[(long long)(e >> 2) * n + s];
    };
    int* o = tris + 3LL * tpos[i];
    int base = 0;
    for (int p = 0; p < tab->npatch[cs]; ++p) {
        const int len = tab->plen[cs][p];
        const int v0 = vid(tab->cyc[cs][base]);
        for (int j = 1; j + 1 < len; ++j) {
            o[0] = v0;
            o[1] = vid(tab->cyc[cs][base + j]);
            o[2] = vid(tab->cyc[cs][base + j + 1]);
            o += 3;
        }
        base += len;
    }
}

// Sample index units to the centred frame: ((i + 0.5) / R * margin - band)
// * tmax + min (the inverse of the normalisation).
__global__ void __launch_bounds__(kB) to_frame_k(const float4* __restrict__ v, int n, int R, float band, float margin,
                                                 float tmax, float lx, float ly, float lz, float* __restrict__ outp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = v[i];
    const float inv = 1.0f / (float)R;
    outp[3 * (long long)i] = ((q.x + 0.5f) * inv * margin - band) * tmax + lx;
    outp[3 * (long long)i + 1] = ((q.y + 0.5f) * inv * margin - band) * tmax + ly;
    outp[3 * (long long)i + 2] = ((q.z + 0.5f) * inv * margin - band) * tmax + lz;
}

template <typename T>
__global__ void fill_val_k(T* p, T v, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------ context
struct Ctx {
    hipStream_t stream = nullptr;
    int device = 0;
    int R = 0;
    float band = 0, margin = 0, iso = 0;
    float mean[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, tmax = 0;
    int n_in = 0, n_verts = 0, n_faces = 0, fill_iters = 0;
    float4* tri = nullptr;
    size_t cap_tri = 0;
    long long* cols = nullptr;
    size_t cap_cols = 0;
    long long* sums = nullptr;
    size_t cap_sums = 0;
    unsigned* grid = nullptr;
    uint8_t* outside = nullptr;
    size_t cap_grid = 0;
    int* cnt = nullptr;       // cells (+1), then crossing edges (+1)
    size_t cap_cnt = 0;
    int* cnt2 = nullptr;
    size_t cap_cnt2 = 0;
    float4* verts = nullptr;
    size_t cap_verts = 0;
    int* tris = nullptr;
    size_t cap_tris = 0;
    unsigned* b6 = nullptr;
    int* flag = nullptr;
    size_t cap_out = 0;
    int* err = nullptr;    // look-back scan error bits
    LookbackScan lb;
    McTable* tab = nullptr;
    int* h_int = nullptr;  // pinned host words
};

template <typename T>
static int ensure(T** p, size_t& cap, size_t n, hipStream_t st) {
    if (n <= cap) return 0;
    if (*p) PAMO_HIP(hipFree(*p));
    *p = nullptr;
    const size_t a = n + n / 8 + 64;
    PAMO_HIP(hipMalloc((void**)p, sizeof(T) * a));
    cap = a;
    (void)st;
    return 0;
}

static void ctx_free(Ctx& c) {
    void* ptrs[] = {c.tri, c.cols, c.sums, c.grid, c.outside, c.cnt, c.cnt2, c.verts, c.tris, c.b6, c.flag, c.err,
                    c.tab};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    lookback_free(c.lb);
    if (c.h_int) (void)hipHostFree(c.h_int);
}

static int read_int(Ctx& c, const int* d, int& v) {
    PAMO_HIP(hipMemcpyAsync(c.h_int, d, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    v = c.h_int[0];
    return 0;
}

static int scan_int(Ctx& c, int* v, long long n) {
    PAMO_TRY(lookback_reserve(c.lb, n, c.err, c.stream));
    return scan_exclusive_lb(v, v, n, c.lb, c.stream);
}

// `mean`: the centring offset of PaMO.run (numpy's float32
// tris.mean(axis=1).mean(axis=0), pamo_centroid).
static int remesh(Ctx& c, const float* d_pts, const int* d_tris, int n_tris, const float mean[3], int R) {
    hipStream_t st = c.stream;
    const long long n3 = 3LL * n_tris;
    for (int d = 0; d < 3; ++d) c.mean[d] = mean[d];
    c.n_in = n_tris;
    c.R = R;
    c.band = 3.0f / (float)R;
    c.margin = c.band * 2.0f + 1.0f;
    c.iso = (float)(0.9 / (double)R);
    // per-axis minimum and maximum of t1 = p - mean over the corners
    if (!c.b6) PAMO_HIP(hipMalloc((void**)&c.b6, 6 * sizeof(unsigned)));
    {
        unsigned init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
        PAMO_HIP(hipMemcpyAsync(c.b6, init, sizeof(init), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(bounds_k, dim3(grid_for(n3, kB)), dim3(kB), 0, st, d_pts, d_tris, n3, c.mean[0], c.mean[1],
                           c.mean[2], c.b6);
        PAMO_HIP(hipGetLastError());
        unsigned h6[6];
        PAMO_HIP(hipMemcpyAsync(h6, c.b6, sizeof(h6), hipMemcpyDeviceToHost, st));
        PAMO_HIP(hipStreamSynchronize(st));
        auto dec = [](unsigned u) {
            const unsigned b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
            float f;
            std::memcpy(&f, &b, 4);
            return f;
        };
        float mx = -INFINITY;
        for (int d = 0; d < 3; ++d) {
            c.lo[d] = dec(h6[d]);
            // max over the corners of (t1 - min): the subtraction is monotonic
            const float diff = dec(h6[3 + d]) - c.lo[d];
            mx = std::max(mx, diff);
        }
        c.tmax = mx;
    }
    PAMO_CHECK(c.tmax > 0.0f, "pamo_s1_remesh: the triangles span no extent");
    PAMO_TRY(ensure(&c.tri, c.cap_tri, (size_t)n3, st));
    hipLaunchKernelGGL(normalize_k, dim3(grid_for(n3, kB)), dim3(kB), 0, st, d_pts, d_tris, n3, c.mean[0], c.mean[1],
                       c.mean[2], c.lo[0], c.lo[1], c.lo[2], c.tmax, c.band, c.margin, c.tri);
    PAMO_HIP(hipGetLastError());
    // distance field in the band
    const long long ng = (long long)R * R * R;
    PAMO_TRY(ensure(&c.grid, c.cap_grid, (size_t)ng, st));
    PAMO_TRY(ensure(&c.outside, c.cap_out, (size_t)ng, st));
    hipLaunchKernelGGL(fill_val_k<unsigned>, dim3(grid_for(ng, kB)), dim3(kB), 0, st, c.grid, 0x7F800000u, ng);
    PAMO_HIP(hipGetLastError());
    PAMO_TRY(ensure(&c.cols, c.cap_cols, (size_t)n_tris + 1, st));
    hipLaunchKernelGGL(columns_k, dim3(grid_for(n_tris + 1, kB)), dim3(kB), 0, st, (const float4*)c.tri, n_tris, R,
                       c.band, c.cols);
    PAMO_HIP(hipGetLastError());
    const long long nc = n_tris + 1, ntile = (nc + 2047) / 2048;
    PAMO_TRY(ensure(&c.sums, c.cap_sums, (size_t)ntile, st));
    hipLaunchKernelGGL(scan64_tiles_k, dim3((unsigned)ntile), dim3(kB), 0, st, c.cols, nc, c.sums);
    PAMO_HIP(hipGetLastError());
    std::vector<long long> hs(ntile), ho(ntile);
    PAMO_HIP(hipMemcpyAsync(hs.data(), c.sums, sizeof(long long) * ntile, hipMemcpyDeviceToHost, st));
    PAMO_HIP(hipStreamSynchronize(st));
    long long run = 0;
    for (long long t = 0; t < ntile; ++t) {
        ho[t] = run;
        run += hs[t];
    }
    const long long total = run;
    PAMO_HIP(hipMemcpyAsync(c.sums, ho.data(), sizeof(long long) * ntile, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(scan64_add_k, dim3(grid_for(nc, kB)), dim3(kB), 0, st, c.cols, nc, (const long long*)c.sums);
    PAMO_HIP(hipGetLastError());
    if (total > 0)
        hipLaunchKernelGGL(udf_k, dim3(grid_for(total, kB)), dim3(kB), 0, st, (const float4*)c.tri, n_tris,
                           (const long long*)c.cols, total, R, c.band, c.grid);
    PAMO_HIP(hipGetLastError());
    // sign: flood fill of the outside from the border
    const float wall = 0.55f / (float)R;
    unsigned wall_bits;
    std::memcpy(&wall_bits, &wall, 4);
    hipLaunchKernelGGL(fill_init_k, dim3(grid_for(ng, kB)), dim3(kB), 0, st, (const unsigned*)c.grid, R, wall_bits,
                       c.outside);
    PAMO_HIP(hipGetLastError());
    if (!c.flag) PAMO_HIP(hipMalloc((void**)&c.flag, sizeof(int)));
    const int nb = (R + kBrick - 1) / kBrick;
    const int max_iters = 8 * R + 64;  // the front advances at least one brick per launch
    c.fill_iters = 0;
    for (;;) {
        PAMO_CHECK(c.fill_iters < max_iters, "pamo_s1_remesh: flood fill did not converge");
        hipLaunchKernelGGL(fill_val_k<int>, dim3(1), dim3(64), 0, st, c.flag, 0, 1LL);
        hipLaunchKernelGGL(fill_brick_k, dim3((unsigned)nb * nb * nb), dim3(512), 0, st, (const unsigned*)c.grid, R,
                           wall_bits, c.outside, c.flag);
        PAMO_HIP(hipGetLastError());
        ++c.fill_iters;
        int ch = 0;
        PAMO_TRY(read_int(c, c.flag, ch));
        if (!ch) break;
    }
    // marching cubes at the level iso
    if (!c.tab) {
        const CaseTable* t = case_table();
        PAMO_CHECK(t != nullptr, "pamo_s1_remesh: marching-cubes case table inconsistent");
        PAMO_HIP(hipMalloc((void**)&c.tab, sizeof(McTable)));
        McTable h;
        std::memcpy(h.cyc, t->cyc, sizeof(h.cyc));
        std::memcpy(h.plen, t->plen, sizeof(h.plen));
        std::memcpy(h.npatch, t->npatch, sizeof(h.npatch));
        std::memcpy(h.ntri, t->ntri, sizeof(h.ntri));
        PAMO_HIP(hipMemcpyAsync(c.tab, &h, sizeof(McTable), hipMemcpyHostToDevice, st));
        PAMO_HIP(hipStreamSynchronize(st));
    }
    // vertices: one per crossing grid edge
    PAMO_TRY(ensure(&c.cnt2, c.cap_cnt2, (size_t)3 * ng + 1, st));
    hipLaunchKernelGGL(edge_count_k, dim3(grid_for(3 * ng + 1, kB)), dim3(kB), 0, st, (const unsigned*)c.grid,
                       (const uint8_t*)c.outside, R, c.iso, c.cnt2);
    PAMO_HIP(hipGetLastError());
    PAMO_TRY(scan_int(c, c.cnt2, 3 * ng + 1));
    int nv = 0;
    PAMO_TRY(read_int(c, c.cnt2 + 3 * ng, nv));
    PAMO_TRY(ensure(&c.verts, c.cap_verts, (size_t)std::max(nv, 1), st));
    if (nv > 0)
        hipLaunchKernelGGL(edge_verts_k, dim3(grid_for(3 * ng, kB)), dim3(kB), 0, st, (const unsigned*)c.grid,
                           (const uint8_t*)c.outside, R, c.iso, (const int*)c.cnt2, c.verts);
    PAMO_HIP(hipGetLastError());
    // triangles: the fans of every cell's patches
    const long long m3 = (long long)(R - 1) * (R - 1) * (R - 1);
    PAMO_TRY(ensure(&c.cnt, c.cap_cnt, (size_t)m3 + 1, st));
    hipLaunchKernelGGL(cell_count_k, dim3(grid_for(m3 + 1, kB)), dim3(kB), 0, st, (const unsigned*)c.grid,
                       (const uint8_t*)c.outside, R, c.iso, (const McTable*)c.tab, c.cnt);
    PAMO_HIP(hipGetLastError());
    PAMO_TRY(scan_int(c, c.cnt, m3 + 1));
    int nt = 0;
    PAMO_TRY(read_int(c, c.cnt + m3, nt));
    PAMO_TRY(ensure(&c.tris, c.cap_tris, (size_t)3 * std::max(nt, 1), st));
    if (nt > 0)
        hipLaunchKernelGGL(cell_tris_k, dim3(grid_for(m3, kB)), dim3(kB), 0, st, (const unsigned*)c.grid,
                           (const uint8_t*)c.outside, R, c.iso, (const McTable*)c.tab, (const int*)c.cnt,
                           (const int*)c.cnt2, c.tris);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipMemcpyAsync(c.h_int, c.err, sizeof(int), hipMemcpyDeviceToHost, st));
    PAMO_HIP(hipStreamSynchronize(st));
    PAMO_CHECK(c.h_int[0] == 0, "pamo_s1_remesh: a look-back scan wait exceeded its bound");
    c.n_verts = nv;
    c.n_faces = nt;
    return 0;
}

}  // namespace
}  // namespace s1
}  // namespace pamo

// ------------------------------------------------------------ C ABI
struct pamo_s1 {};

extern "C" {

int pamo_s1_create(pamo_s1** out, int device) {
    pamo::clear_error();
    PAMO_CHECK(out != nullptr, "pamo_s1_create: out is null");
    auto* c = new pamo::s1::Ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_int, 4 * sizeof(int), 0);
    if (e == hipSuccess) e = hipMalloc((void**)&c->err, sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(c->err, 0, sizeof(int), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        pamo::set_error(std::string("pamo_s1_create: ") + hipGetErrorString(e));
        pamo::s1::ctx_free(*c);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return 1;
    }
    *out = reinterpret_cast<pamo_s1*>(c);
    return 0;
}

int pamo_s1_destroy(pamo_s1* s) {
    if (!s) return 0;
    auto* c = reinterpret_cast<pamo::s1::Ctx*>(s);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pamo::s1::ctx_free(*c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int pamo_s1_remesh(pamo_s1* s, const float* d_points, int n_points, const int* d_faces, int n_faces,
                   const float* mean3, int resolution, int* n_verts_out, int* n_faces_out) {
    pamo::clear_error();
    PAMO_CHECK(s && d_points && d_faces && mean3, "pamo_s1_remesh: null argument");
    PAMO_CHECK(n_points > 0 && n_faces > 0, "pamo_s1_remesh: empty mesh");
    PAMO_CHECK(resolution >= 8 && resolution <= 1024, "pamo_s1_remesh: resolution must be in [8, 1024]");
    auto& c = *reinterpret_cast<pamo::s1::Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    PAMO_TRY(pamo::s1::remesh(c, d_points, d_faces, n_faces, mean3, resolution));
    if (n_verts_out) *n_verts_out = c.n_verts;
    if (n_faces_out) *n_faces_out = c.n_faces;
    return 0;
}

int pamo_s1_get(pamo_s1* s, float* d_verts, int* d_faces) {
    pamo::clear_error();
    PAMO_CHECK(s && d_verts && d_faces, "pamo_s1_get: null argument");
    auto& c = *reinterpret_cast<pamo::s1::Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    using namespace pamo::s1;
    if (c.n_verts > 0)
        hipLaunchKernelGGL(to_frame_k, dim3(pamo::grid_for(c.n_verts, kB)), dim3(kB), 0, c.stream,
                           (const float4*)c.verts, c.n_verts, c.R, c.band, c.margin, c.tmax, c.lo[0], c.lo[1], c.lo[2],
                           d_verts);
    PAMO_HIP(hipGetLastError());
    if (c.n_faces > 0)
        PAMO_HIP(hipMemcpyAsync(d_faces, c.tris, sizeof(int) * 3 * (size_t)c.n_faces, hipMemcpyDeviceToDevice,
                                c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

int pamo_s1_debug(pamo_s1* s, int what, void* h_out, long long cap, long long* n_out) {
    pamo::clear_error();
    PAMO_CHECK(s && n_out, "pamo_s1_debug: null argument");
    auto& c = *reinterpret_cast<pamo::s1::Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    using namespace pamo::s1;
    const long long ng = (long long)c.R * c.R * c.R;
    long long n = 0;
    const void* src = nullptr;
    size_t elem = 0;
    std::vector<unsigned char> host;
    switch (what) {
    case PAMO_S1_DBG_FRAME: {
        n = 8;
        PAMO_CHECK(h_out && cap >= n, "pamo_s1_debug: output too small");
        float* o = static_cast<float*>(h_out);
        o[0] = c.mean[0]; o[1] = c.mean[1]; o[2] = c.mean[2];
        o[3] = c.lo[0]; o[4] = c.lo[1]; o[5] = c.lo[2];
        o[6] = c.tmax; o[7] = c.iso;
        *n_out = n;
        return 0;
    }
    case PAMO_S1_DBG_CASES: {
        n = 256 * 18;
        PAMO_CHECK(h_out && cap >= n, "pamo_s1_debug: output too small");
        const CaseTable* t = case_table();
        PAMO_CHECK(t != nullptr, "pamo_s1_debug: marching-cubes case table inconsistent");
        auto* o = static_cast<int8_t*>(h_out);
        for (int cs = 0; cs < 256; ++cs) {
            for (int e = 0; e < 12; ++e) o[18 * cs + e] = t->cyc[cs][e];
            for (int p = 0; p < 4; ++p) o[18 * cs + 12 + p] = t->plen[cs][p];
            o[18 * cs + 16] = t->npatch[cs];
            o[18 * cs + 17] = t->ntri[cs];
        }
        *n_out = n;
        return 0;
    }
    case PAMO_S1_DBG_FIELD: n = ng; src = c.grid; elem = 4; break;
    case PAMO_S1_DBG_OUTSIDE: n = ng; src = c.outside; elem = 1; break;
    case PAMO_S1_DBG_NORMALIZED: n = 12LL * c.n_in; src = c.tri; elem = 4; break;
    case PAMO_S1_DBG_VERTS: n = 4LL * c.n_verts; src = c.verts; elem = 4; break;
    default: PAMO_CHECK(false, "pamo_s1_debug: unknown item");
    }
    PAMO_CHECK(src != nullptr || n == 0, "pamo_s1_debug: nothing computed yet");
    PAMO_CHECK(h_out && cap >= n, "pamo_s1_debug: output too small");
    if (n > 0) {
        PAMO_HIP(hipMemcpyAsync(h_out, src, elem * (size_t)n, hipMemcpyDeviceToHost, c.stream));
        PAMO_HIP(hipStreamSynchronize(c.stream));
    }
    *n_out = n;
    return 0;
}

}  // extern "C"
