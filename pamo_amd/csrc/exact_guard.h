// Exact triangle-intersection decisions for the guard's exact mode (device
// code for gfx950).  No counterpart in the reference, whose guard uses a
// mixed FP32/FP64 predicate with skip rules (self_intersect.cuh:229-378,
// tri_tri_3d.cuh:99-276); this mode flags exactly the pairs that intersect.
//
// Predicates on float32 coordinates:
//   * filtered: differences in FP64 (exact when no rounding tail appears),
//     products and sums checked against the error bounds of J. R. Shewchuk,
//     "Adaptive Precision Floating-Point Arithmetic and Fast Robust Geometric
//     Predicates" (Discrete Comput. Geom. 18, 1997); a sign that the bound
//     cannot certify is reported as undecided;
//   * exact: orient3d as the 4x4 determinant [p 1] of the raw coordinates,
//     24 signed triple products; a product of two float32 values is exact in
//     FP64 and two-product makes the third factor exact, so each term is a
//     two-double expansion and their sum, grown term by term with two-sum,
//     is exact.  orient2d likewise from 6 exact products.
// Decisions (mesh semantics): faces with no common vertex index must not
// meet; faces sharing one vertex may meet only there; faces sharing an edge
// may meet only along it (a flat fold, both opposite corners on one side of
// the edge in a common plane, is an intersection); repeated faces intersect;
// a degenerate face (collinear corners) is not tested.
#pragma once
#include "hip_common.h"

namespace pamo {
namespace {

struct XP3 {
    float x, y, z;
};
struct XP2 {
    float x, y;
};

__device__ __forceinline__ void x_two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    const double bv = x - a, av = x - bv, br = b - bv, ar = a - av;
    y = ar + br;
}
__device__ __forceinline__ void x_two_diff(double a, double b, double& x, double& y) {
    x = a - b;
    const double bv = a - x, av = x + bv, br = bv - b, ar = a - av;
    y = ar + br;
}
__device__ __forceinline__ void x_split(double a, double& hi, double& lo) {
    const double c = 134217729.0 * a;  // 2^27 + 1
    const double abig = c - a;
    hi = c - abig;
    lo = a - hi;
}
__device__ __forceinline__ void x_two_product(double a, double b, double& x, double& y) {
    x = a * b;
    double ah, al, bh, bl;
    x_split(a, ah, al);
    x_split(b, bh, bl);
    const double e1 = x - ah * bh, e2 = e1 - al * bh, e3 = e2 - ah * bl;
    y = al * bl - e3;
}

constexpr double kXEps = 1.1102230246251565e-16;  // 2^-53
constexpr double kXO3d = (7.0 + 56.0 * kXEps) * kXEps;
constexpr double kXCcw = (3.0 + 16.0 * kXEps) * kXEps;
constexpr int kUndecided = 2;

__device__ int o3_filter(const XP3& a, const XP3& b, const XP3& c, const XP3& d) {
    double adx, ady, adz, bdx, bdy, bdz, cdx, cdy, cdz, t, tails = 0.0;
    x_two_diff(a.x, d.x, adx, t); tails += fabs(t);
    x_two_diff(a.y, d.y, ady, t); tails += fabs(t);
    x_two_diff(a.z, d.z, adz, t); tails += fabs(t);
    x_two_diff(b.x, d.x, bdx, t); tails += fabs(t);
    x_two_diff(b.y, d.y, bdy, t); tails += fabs(t);
    x_two_diff(b.z, d.z, bdz, t); tails += fabs(t);
    x_two_diff(c.x, d.x, cdx, t); tails += fabs(t);
    x_two_diff(c.y, d.y, cdy, t); tails += fabs(t);
    x_two_diff(c.z, d.z, cdz, t); tails += fabs(t);
    if (tails != 0.0) return kUndecided;
    const double bc = bdx * cdy, cb = cdx * bdy, ca = cdx * ady, ac = adx * cdy, ab = adx * bdy, ba = bdx * ady;
    const double det = adz * (bc - cb) + bdz * (ca - ac) + cdz * (ab - ba);
    const double perm = (fabs(bc) + fabs(cb)) * fabs(adz) + (fabs(ca) + fabs(ac)) * fabs(bdz) +
                        (fabs(ab) + fabs(ba)) * fabs(cdz);
    const double bound = kXO3d * perm;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    if (perm == 0.0) return 0;
    return kUndecided;
}

__device__ int o2_filter(const XP2& a, const XP2& b, const XP2& c) {
    double acx, acy, bcx, bcy, t, tails = 0.0;
    x_two_diff(a.x, c.x, acx, t); tails += fabs(t);
    x_two_diff(a.y, c.y, acy, t); tails += fabs(t);
    x_two_diff(b.x, c.x, bcx, t); tails += fabs(t);
    x_two_diff(b.y, c.y, bcy, t); tails += fabs(t);
    if (tails != 0.0) return kUndecided;
    const double l = acx * bcy, r = acy * bcx;
    const double det = l - r, bound = kXCcw * (fabs(l) + fabs(r));
    if (det > bound) return 1;
    if (-det > bound) return -1;
    if (l == 0.0 && r == 0.0) return 0;
    return kUndecided;
}

// Nonoverlapping expansion in increasing magnitude, zero components removed.
struct XExp {
    double h[50];
    int n = 0;
};

__device__ void x_grow(XExp& e, double b) {
    double q = b;
    int k = 0;
    for (int i = 0; i < e.n; ++i) {
        double s, hh;
        x_two_sum(q, e.h[i], s, hh);
        q = s;
        if (hh != 0.0) e.h[k++] = hh;
    }
    if (q != 0.0 || k == 0) e.h[k++] = q;
    e.n = k;
}

__device__ __forceinline__ int x_sign(const XExp& e) {
    const double top = e.h[e.n - 1];
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

// s * u * v * w added exactly (u * v is exact in FP64 for float32 factors)
__device__ __forceinline__ void x_add_triple(XExp& e, float s, float u, float v, float w) {
    double hi, lo;
    x_two_product((double)s * ((double)u * (double)v), (double)w, hi, lo);
    x_grow(e, lo);
    x_grow(e, hi);
}

__device__ void x_add_det3(XExp& e, float s, const XP3& p, const XP3& q, const XP3& r) {
    x_add_triple(e, s, p.x, q.y, r.z);
    x_add_triple(e, -s, p.x, q.z, r.y);
    x_add_triple(e, -s, p.y, q.x, r.z);
    x_add_triple(e, s, p.y, q.z, r.x);
    x_add_triple(e, s, p.z, q.x, r.y);
    x_add_triple(e, -s, p.z, q.y, r.x);
}

// det[a-d; b-d; c-d] = det3(a,b,c) - det3(a,b,d) + det3(a,c,d) - det3(b,c,d)
__device__ int o3_exact(const XP3& a, const XP3& b, const XP3& c, const XP3& d) {
    XExp e;
    x_add_det3(e, 1.0f, a, b, c);
    x_add_det3(e, -1.0f, a, b, d);
    x_add_det3(e, 1.0f, a, c, d);
    x_add_det3(e, -1.0f, b, c, d);
    return x_sign(e);
}

// (a - c) x (b - c) = a.x b.y - a.y b.x - a.x c.y + a.y c.x + b.x c.y - b.y c.x
__device__ int o2_exact(const XP2& a, const XP2& b, const XP2& c) {
    XExp e;
    const float s[6] = {1.f, -1.f, -1.f, 1.f, 1.f, -1.f};
    const float u[6] = {a.x, a.y, a.x, a.y, b.x, b.y}, v[6] = {b.y, b.x, c.y, c.x, c.y, c.x};
#pragma unroll
    for (int k = 0; k < 6; ++k) x_grow(e, (double)s[k] * ((double)u[k] * (double)v[k]));
    return x_sign(e);
}

// Predicate provider: filtered only (undecided signs raise a flag), or
// filtered with the exact evaluation behind it.
template <bool kExact>
struct XPred {
    bool undecided = false;
    __device__ int o3(const XP3& a, const XP3& b, const XP3& c, const XP3& d) {
        const int s = o3_filter(a, b, c, d);
        if (s != kUndecided) return s;
        if (kExact) return o3_exact(a, b, c, d);
        undecided = true;
        return 0;
    }
    __device__ int o2(const XP2& a, const XP2& b, const XP2& c) {
        const int s = o2_filter(a, b, c);
        if (s != kUndecided) return s;
        if (kExact) return o2_exact(a, b, c);
        undecided = true;
        return 0;
    }
};

__device__ __forceinline__ XP2 x_proj(const XP3& p, int drop) {
    return drop == 0 ? XP2{p.y, p.z} : (drop == 1 ? XP2{p.z, p.x} : XP2{p.x, p.y});
}

template <class P>
__device__ int x_proj_axis(P& pr, const XP3& a, const XP3& b, const XP3& c) {
    for (int d = 0; d < 3; ++d)
        if (pr.o2(x_proj(a, d), x_proj(b, d), x_proj(c, d)) != 0) return d;
    return -1;
}

__device__ __forceinline__ bool x_on_segment(const XP2& p, const XP2& q, const XP2& r) {
    return fminf(p.x, q.x) <= r.x && r.x <= fmaxf(p.x, q.x) && fminf(p.y, q.y) <= r.y && r.y <= fmaxf(p.y, q.y);
}

template <class P>
__device__ bool x_seg_seg_2d(P& pr, const XP2& p1, const XP2& p2, const XP2& q1, const XP2& q2) {
    const int o1 = pr.o2(p1, p2, q1), o2 = pr.o2(p1, p2, q2);
    const int o3 = pr.o2(q1, q2, p1), o4 = pr.o2(q1, q2, p2);
    if (o1 * o2 < 0 && o3 * o4 < 0) return true;
    if (o1 == 0 && x_on_segment(p1, p2, q1)) return true;
    if (o2 == 0 && x_on_segment(p1, p2, q2)) return true;
    if (o3 == 0 && x_on_segment(q1, q2, p1)) return true;
    if (o4 == 0 && x_on_segment(q1, q2, p2)) return true;
    return false;
}

template <class P>
__device__ bool x_point_in_tri_2d(P& pr, const XP2& p, const XP2& a, const XP2& b, const XP2& c) {
    const int s1 = pr.o2(a, b, p), s2 = pr.o2(b, c, p), s3 = pr.o2(c, a, p);
    const bool has_neg = s1 < 0 || s2 < 0 || s3 < 0, has_pos = s1 > 0 || s2 > 0 || s3 > 0;
    return !(has_neg && has_pos);
}

// Closed segment pq against closed non-degenerate triangle abc.
template <class P>
__device__ bool x_seg_tri(P& pr, const XP3& p, const XP3& q, const XP3& a, const XP3& b, const XP3& c) {
    const int o1 = pr.o3(a, b, c, p), o2 = pr.o3(a, b, c, q);
    if ((o1 > 0 && o2 > 0) || (o1 < 0 && o2 < 0)) return false;
    if (o1 == 0 && o2 == 0) {
        const int d = x_proj_axis(pr, a, b, c);
        if (d < 0) return false;
        const XP2 P2 = x_proj(p, d), Q2 = x_proj(q, d), A = x_proj(a, d), B = x_proj(b, d), C = x_proj(c, d);
        return x_point_in_tri_2d(pr, P2, A, B, C) || x_point_in_tri_2d(pr, Q2, A, B, C) ||
               x_seg_seg_2d(pr, P2, Q2, A, B) || x_seg_seg_2d(pr, P2, Q2, B, C) || x_seg_seg_2d(pr, P2, Q2, C, A);
    }
    const int s1 = pr.o3(p, q, a, b), s2 = pr.o3(p, q, b, c), s3 = pr.o3(p, q, c, a);
    return (s1 >= 0 && s2 >= 0 && s3 >= 0) || (s1 <= 0 && s2 <= 0 && s3 <= 0);
}

template <class P>
__device__ bool x_tri_tri(P& pr, const XP3* t, const XP3* u) {
    const int s0 = pr.o3(t[0], t[1], t[2], u[0]), s1 = pr.o3(t[0], t[1], t[2], u[1]),
              s2 = pr.o3(t[0], t[1], t[2], u[2]);
    if ((s0 > 0 && s1 > 0 && s2 > 0) || (s0 < 0 && s1 < 0 && s2 < 0)) return false;
    if (s0 == 0 && s1 == 0 && s2 == 0) {
        const int d = x_proj_axis(pr, t[0], t[1], t[2]);
        if (d < 0) return false;
        XP2 a[3], b[3];
        for (int i = 0; i < 3; ++i) {
            a[i] = x_proj(t[i], d);
            b[i] = x_proj(u[i], d);
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (x_seg_seg_2d(pr, a[i], a[(i + 1) % 3], b[j], b[(j + 1) % 3])) return true;
        return x_point_in_tri_2d(pr, a[0], b[0], b[1], b[2]) || x_point_in_tri_2d(pr, b[0], a[0], a[1], a[2]);
    }
    for (int i = 0; i < 3; ++i)
        if (x_seg_tri(pr, t[i], t[(i + 1) % 3], u[0], u[1], u[2]) ||
            x_seg_tri(pr, u[i], u[(i + 1) % 3], t[0], t[1], t[2]))
            return true;
    return false;
}

// x coplanar with non-degenerate (v, a, b) and v -> x entering the closed
// angle of the triangle at v (x != v).
template <class P>
__device__ bool x_in_sector(P& pr, const XP3& v, const XP3& a, const XP3& b, const XP3& x) {
    if (pr.o3(v, a, b, x) != 0) return false;
    if (x.x == v.x && x.y == v.y && x.z == v.z) return false;
    const int d = x_proj_axis(pr, v, a, b);
    if (d < 0) return false;
    const XP2 V = x_proj(v, d), A = x_proj(a, d), B = x_proj(b, d), X = x_proj(x, d);
    const int o = pr.o2(V, A, B);
    const int s1 = pr.o2(V, A, X), s2 = pr.o2(V, X, B);
    if (s1 * o < 0 || s2 * o < 0) return false;
    auto sg = [](float p, float q) { return (p > q) - (p < q); };
    if (s1 == 0 && !(sg(A.x, V.x) == sg(X.x, V.x) && sg(A.y, V.y) == sg(X.y, V.y))) return false;
    if (s2 == 0 && !(sg(B.x, V.x) == sg(X.x, V.x) && sg(B.y, V.y) == sg(X.y, V.y))) return false;
    return true;
}

// Faces (ids ia, corners pa) and (ib, pb): do they intersect beyond what
// their shared vertices allow?  Degenerate faces are not tested.
template <class P>
__device__ bool x_faces_intersect(P& pr, const int* ia, const XP3* pa, const int* ib, const XP3* pb) {
    if (x_proj_axis(pr, pa[0], pa[1], pa[2]) < 0 || x_proj_axis(pr, pb[0], pb[1], pb[2]) < 0) return false;
    int share = 0, sa[3] = {-1, -1, -1}, sb[3] = {-1, -1, -1};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (ia[i] == ib[j] && share < 3) {
                sa[share] = i;
                sb[share] = j;
                ++share;
            }
    if (share == 0) return x_tri_tri(pr, pa, pb);
    if (share >= 3) return true;
    if (share == 1) {
        const int i = sa[0], j = sb[0];
        const XP3 v = pa[i], a = pa[(i + 1) % 3], b = pa[(i + 2) % 3];
        const XP3 c = pb[(j + 1) % 3], d = pb[(j + 2) % 3];
        return x_seg_tri(pr, a, b, pb[0], pb[1], pb[2]) || x_seg_tri(pr, c, d, pa[0], pa[1], pa[2]) ||
               x_in_sector(pr, v, a, b, c) || x_in_sector(pr, v, a, b, d) || x_in_sector(pr, v, c, d, a) ||
               x_in_sector(pr, v, c, d, b);
    }
    const int oa = 3 - sa[0] - sa[1], ob = 3 - sb[0] - sb[1];
    const XP3 u = pa[sa[0]], w = pa[sa[1]], a = pa[oa], b = pb[ob];
    if (pr.o3(u, w, a, b) != 0) return false;
    const int d = x_proj_axis(pr, u, w, a);
    if (d < 0) return false;
    return pr.o2(x_proj(u, d), x_proj(w, d), x_proj(a, d)) == pr.o2(x_proj(u, d), x_proj(w, d), x_proj(b, d));
}

}  // namespace
}  // namespace pamo
