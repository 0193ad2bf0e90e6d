// Exact orientation predicates and exact triangle-intersection tests for
// double coordinates (float32 positions convert exactly).  Test
// infrastructure only: included by the CPU oracles.
//
//   * orient3d / orient2d: floating-point filter, then exact evaluation with
//     floating-point expansions (error-free two-sum / two-product, expansion
//     growth and scaling, after J. R. Shewchuk, "Adaptive Precision
//     Floating-Point Arithmetic and Fast Robust Geometric Predicates",
//     Discrete Comput. Geom. 18, 1997);
//   * closed segment-triangle and triangle-triangle intersection from those
//     signs only (coplanar cases in a projection where the triangle is
//     non-degenerate);
//   * faces_intersect: mesh semantics -- faces with no common vertex index
//     must not meet; faces sharing one vertex may meet only there; faces
//     sharing an edge may meet only along it (a flat fold, both opposite
//     corners on the same side of the edge in a common plane, is an
//     intersection); repeated faces intersect; a degenerate face (collinear
//     corners) is not tested.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace exact {

// ---------------------------------------------------------------- expansions
static inline void two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    const double bv = x - a, av = x - bv, br = b - bv, ar = a - av;
    y = ar + br;
}
static inline void fast_two_sum(double a, double b, double& x, double& y) {  // |a| >= |b|
    x = a + b;
    const double bv = x - a;
    y = b - bv;
}
static inline void two_diff(double a, double b, double& x, double& y) {
    x = a - b;
    const double bv = a - x, av = x + bv, br = bv - b, ar = a - av;
    y = ar + br;
}
static inline void split(double a, double& hi, double& lo) {
    const double c = 134217729.0 * a;  // 2^27 + 1
    const double abig = c - a;
    hi = c - abig;
    lo = a - hi;
}
static inline void two_product(double a, double b, double& x, double& y) {
    x = a * b;
    double ah, al, bh, bl;
    split(a, ah, al);
    split(b, bh, bl);
    const double e1 = x - ah * bh, e2 = e1 - al * bh, e3 = e2 - ah * bl;
    y = al * bl - e3;
}

using Exp = std::vector<double>;  // nonoverlapping, increasing magnitude, no zeros

static Exp from_diff(double a, double b) {
    double x, y;
    two_diff(a, b, x, y);
    Exp e;
    if (y != 0.0) e.push_back(y);
    if (x != 0.0) e.push_back(x);
    return e;
}

static Exp grow(const Exp& e, double b) {
    Exp h;
    h.reserve(e.size() + 1);
    double q = b;
    for (double c : e) {
        double s, hh;
        two_sum(q, c, s, hh);
        q = s;
        if (hh != 0.0) h.push_back(hh);
    }
    if (q != 0.0) h.push_back(q);
    return h;
}

static Exp add(const Exp& e, const Exp& f) {
    Exp h = e;
    for (double c : f) h = grow(h, c);
    return h;
}

static Exp neg(Exp e) {
    for (double& c : e) c = -c;
    return e;
}

static Exp scale(const Exp& e, double b) {
    Exp h;
    if (e.empty() || b == 0.0) return h;
    h.reserve(2 * e.size());
    double q, hh;
    two_product(e[0], b, q, hh);
    if (hh != 0.0) h.push_back(hh);
    for (size_t i = 1; i < e.size(); ++i) {
        double p1, p0, s;
        two_product(e[i], b, p1, p0);
        two_sum(q, p0, s, hh);
        if (hh != 0.0) h.push_back(hh);
        fast_two_sum(p1, s, q, hh);
        if (hh != 0.0) h.push_back(hh);
    }
    if (q != 0.0) h.push_back(q);
    return h;
}

static Exp mul(const Exp& e, const Exp& f) {
    Exp h;
    for (double c : f) h = add(h, scale(e, c));
    return h;
}

static int sign_of(const Exp& e) { return e.empty() ? 0 : (e.back() > 0.0 ? 1 : -1); }

static const double kEps = 1.1102230246251565e-16;  // 2^-53
static const double kO3dBound = (7.0 + 56.0 * kEps) * kEps;
static const double kCcwBound = (3.0 + 16.0 * kEps) * kEps;

struct P3 {
    double x, y, z;
};

static int orient3d_exact(const P3& a, const P3& b, const P3& c, const P3& d) {
    const Exp adx = from_diff(a.x, d.x), ady = from_diff(a.y, d.y), adz = from_diff(a.z, d.z);
    const Exp bdx = from_diff(b.x, d.x), bdy = from_diff(b.y, d.y), bdz = from_diff(b.z, d.z);
    const Exp cdx = from_diff(c.x, d.x), cdy = from_diff(c.y, d.y), cdz = from_diff(c.z, d.z);
    const Exp m1 = add(mul(bdy, cdz), neg(mul(bdz, cdy)));
    const Exp m2 = add(mul(cdy, adz), neg(mul(cdz, ady)));
    const Exp m3 = add(mul(ady, bdz), neg(mul(adz, bdy)));
    return sign_of(add(add(mul(adx, m1), mul(bdx, m2)), mul(cdx, m3)));
}

// Sign of det[a-d; b-d; c-d] (positive when d lies below the plane of a, b, c
// as seen with a, b, c counter-clockwise).
static int orient3d(const P3& a, const P3& b, const P3& c, const P3& d) {
    double adx, ady, adz, bdx, bdy, bdz, cdx, cdy, cdz, t;
    bool exact_diffs = true;
    two_diff(a.x, d.x, adx, t); exact_diffs &= t == 0.0;
    two_diff(a.y, d.y, ady, t); exact_diffs &= t == 0.0;
    two_diff(a.z, d.z, adz, t); exact_diffs &= t == 0.0;
    two_diff(b.x, d.x, bdx, t); exact_diffs &= t == 0.0;
    two_diff(b.y, d.y, bdy, t); exact_diffs &= t == 0.0;
    two_diff(b.z, d.z, bdz, t); exact_diffs &= t == 0.0;
    two_diff(c.x, d.x, cdx, t); exact_diffs &= t == 0.0;
    two_diff(c.y, d.y, cdy, t); exact_diffs &= t == 0.0;
    two_diff(c.z, d.z, cdz, t); exact_diffs &= t == 0.0;
    if (exact_diffs) {
        const double bc = bdx * cdy, cb = cdx * bdy, ca = cdx * ady, ac = adx * cdy, ab = adx * bdy, ba = bdx * ady;
        const double det = adz * (bc - cb) + bdz * (ca - ac) + cdz * (ab - ba);
        const double perm = (std::fabs(bc) + std::fabs(cb)) * std::fabs(adz) +
                            (std::fabs(ca) + std::fabs(ac)) * std::fabs(bdz) +
                            (std::fabs(ab) + std::fabs(ba)) * std::fabs(cdz);
        const double bound = kO3dBound * perm;
        if (det > bound) return 1;
        if (-det > bound) return -1;
        if (perm == 0.0) return 0;
    }
    return orient3d_exact(a, b, c, d);
}

struct P2 {
    double x, y;
};

static int orient2d_exact(const P2& a, const P2& b, const P2& c) {
    const Exp acx = from_diff(a.x, c.x), acy = from_diff(a.y, c.y);
    const Exp bcx = from_diff(b.x, c.x), bcy = from_diff(b.y, c.y);
    return sign_of(add(mul(acx, bcy), neg(mul(acy, bcx))));
}

// Sign of (a - c) x (b - c): positive when a, b, c turn counter-clockwise.
static int orient2d(const P2& a, const P2& b, const P2& c) {
    double acx, acy, bcx, bcy, t;
    bool exact_diffs = true;
    two_diff(a.x, c.x, acx, t); exact_diffs &= t == 0.0;
    two_diff(a.y, c.y, acy, t); exact_diffs &= t == 0.0;
    two_diff(b.x, c.x, bcx, t); exact_diffs &= t == 0.0;
    two_diff(b.y, c.y, bcy, t); exact_diffs &= t == 0.0;
    if (exact_diffs) {
        const double l = acx * bcy, r = acy * bcx;
        const double det = l - r, bound = kCcwBound * (std::fabs(l) + std::fabs(r));
        if (det > bound) return 1;
        if (-det > bound) return -1;
        if (l == 0.0 && r == 0.0) return 0;
    }
    return orient2d_exact(a, b, c);
}

// ---------------------------------------------------------------- 2-D tests
static P2 proj(const P3& p, int drop) {
    if (drop == 0) return P2{p.y, p.z};
    if (drop == 1) return P2{p.z, p.x};
    return P2{p.x, p.y};
}

// An axis whose projection keeps triangle (a, b, c) non-degenerate; -1 when
// the triangle is degenerate (collinear corners).
static int proj_axis(const P3& a, const P3& b, const P3& c) {
    for (int d = 0; d < 3; ++d)
        if (orient2d(proj(a, d), proj(b, d), proj(c, d)) != 0) return d;
    return -1;
}

static bool on_segment(const P2& p, const P2& q, const P2& r) {  // r collinear with p, q
    return std::min(p.x, q.x) <= r.x && r.x <= std::max(p.x, q.x) && std::min(p.y, q.y) <= r.y &&
           r.y <= std::max(p.y, q.y);
}

static bool seg_seg_2d(const P2& p1, const P2& p2, const P2& q1, const P2& q2) {
    const int o1 = orient2d(p1, p2, q1), o2 = orient2d(p1, p2, q2);
    const int o3 = orient2d(q1, q2, p1), o4 = orient2d(q1, q2, p2);
    if (o1 * o2 < 0 && o3 * o4 < 0) return true;
    if (o1 == 0 && on_segment(p1, p2, q1)) return true;
    if (o2 == 0 && on_segment(p1, p2, q2)) return true;
    if (o3 == 0 && on_segment(q1, q2, p1)) return true;
    if (o4 == 0 && on_segment(q1, q2, p2)) return true;
    return false;
}

static bool point_in_tri_2d(const P2& p, const P2& a, const P2& b, const P2& c) {
    const int s1 = orient2d(a, b, p), s2 = orient2d(b, c, p), s3 = orient2d(c, a, p);
    const bool has_neg = s1 < 0 || s2 < 0 || s3 < 0, has_pos = s1 > 0 || s2 > 0 || s3 > 0;
    return !(has_neg && has_pos);
}

static bool seg_tri_2d(const P2& p, const P2& q, const P2& a, const P2& b, const P2& c) {
    return point_in_tri_2d(p, a, b, c) || point_in_tri_2d(q, a, b, c) || seg_seg_2d(p, q, a, b) ||
           seg_seg_2d(p, q, b, c) || seg_seg_2d(p, q, c, a);
}

static bool tri_tri_2d(const P2* t, const P2* u) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (seg_seg_2d(t[i], t[(i + 1) % 3], u[j], u[(j + 1) % 3])) return true;
    return point_in_tri_2d(t[0], u[0], u[1], u[2]) || point_in_tri_2d(u[0], t[0], t[1], t[2]);
}

// ---------------------------------------------------------------- 3-D tests
// Closed segment pq against closed non-degenerate triangle abc.
static bool seg_tri(const P3& p, const P3& q, const P3& a, const P3& b, const P3& c) {
    const int o1 = orient3d(a, b, c, p), o2 = orient3d(a, b, c, q);
    if ((o1 > 0 && o2 > 0) || (o1 < 0 && o2 < 0)) return false;
    if (o1 == 0 && o2 == 0) {
        const int d = proj_axis(a, b, c);
        return seg_tri_2d(proj(p, d), proj(q, d), proj(a, d), proj(b, d), proj(c, d));
    }
    const int s1 = orient3d(p, q, a, b), s2 = orient3d(p, q, b, c), s3 = orient3d(p, q, c, a);
    return (s1 >= 0 && s2 >= 0 && s3 >= 0) || (s1 <= 0 && s2 <= 0 && s3 <= 0);
}

// Closed non-degenerate triangles t and u share at least one point.
static bool tri_tri(const P3* t, const P3* u) {
    const int s0 = orient3d(t[0], t[1], t[2], u[0]), s1 = orient3d(t[0], t[1], t[2], u[1]),
              s2 = orient3d(t[0], t[1], t[2], u[2]);
    if ((s0 > 0 && s1 > 0 && s2 > 0) || (s0 < 0 && s1 < 0 && s2 < 0)) return false;
    if (s0 == 0 && s1 == 0 && s2 == 0) {
        const int d = proj_axis(t[0], t[1], t[2]);
        P2 a[3], b[3];
        for (int i = 0; i < 3; ++i) {
            a[i] = proj(t[i], d);
            b[i] = proj(u[i], d);
        }
        return tri_tri_2d(a, b);
    }
    for (int i = 0; i < 3; ++i)
        if (seg_tri(t[i], t[(i + 1) % 3], u[0], u[1], u[2]) || seg_tri(u[i], u[(i + 1) % 3], t[0], t[1], t[2]))
            return true;
    return false;
}

// x coplanar with non-degenerate triangle (v, a, b) and the segment v -> x
// entering the closed angle of the triangle at v (x != v).
static bool in_sector(const P3& v, const P3& a, const P3& b, const P3& x) {
    if (orient3d(v, a, b, x) != 0) return false;
    if (x.x == v.x && x.y == v.y && x.z == v.z) return false;
    const int d = proj_axis(v, a, b);
    const P2 V = proj(v, d), A = proj(a, d), B = proj(b, d), X = proj(x, d);
    const int o = orient2d(V, A, B);
    const int s1 = orient2d(V, A, X), s2 = orient2d(V, X, B);
    if (s1 * o < 0 || s2 * o < 0) return false;
    // collinear with an edge v -> e: inside only in the same direction, that is
    // when x - v and e - v have equal signs coordinate by coordinate
    auto sg = [](double p, double q) { return (p > q) - (p < q); };
    auto same_dir = [&](const P2& E) { return sg(E.x, V.x) == sg(X.x, V.x) && sg(E.y, V.y) == sg(X.y, V.y); };
    if (s1 == 0 && !same_dir(A)) return false;
    if (s2 == 0 && !same_dir(B)) return false;
    return true;
}

// Faces fa, fb (vertex indices ia, ib; corners pa, pb): do they intersect
// beyond what their shared vertices allow?
static bool faces_intersect(const int* ia, const P3* pa, const int* ib, const P3* pb) {
    if (proj_axis(pa[0], pa[1], pa[2]) < 0 || proj_axis(pb[0], pb[1], pb[2]) < 0) return false;  // degenerate
    int share = 0, sa[3] = {-1, -1, -1}, sb[3] = {-1, -1, -1};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (ia[i] == ib[j] && share < 3) {
                sa[share] = i;
                sb[share] = j;
                ++share;
            }
    if (share == 0) return tri_tri(pa, pb);
    if (share >= 3) return true;
    if (share == 1) {
        const int i = sa[0], j = sb[0];
        const P3 &v = pa[i], &a = pa[(i + 1) % 3], &b = pa[(i + 2) % 3];
        const P3 &c = pb[(j + 1) % 3], &d = pb[(j + 2) % 3];
        return seg_tri(a, b, pb[0], pb[1], pb[2]) || seg_tri(c, d, pa[0], pa[1], pa[2]) || in_sector(v, a, b, c) ||
               in_sector(v, a, b, d) || in_sector(v, c, d, a) || in_sector(v, c, d, b);
    }
    // shared edge (u, v): opposite corners a (of fa) and b (of fb)
    const int oa = 3 - sa[0] - sa[1], ob = 3 - sb[0] - sb[1];
    const P3 &u = pa[sa[0]], &w = pa[sa[1]], &a = pa[oa], &b = pb[ob];
    if (orient3d(u, w, a, b) != 0) return false;
    const int d = proj_axis(u, w, a);
    return orient2d(proj(u, d), proj(w, d), proj(a, d)) == orient2d(proj(u, d), proj(w, d), proj(b, d));
}

}  // namespace exact
