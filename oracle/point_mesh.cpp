// CPU oracle of the surface metrics: point-to-mesh squared distance and the
// Hausdorff distance with libigl 2.5.1's semantics, which the reference calls
// in simp_cuda/safe_project/src/pamo_safe_project/metrics.py:41-49
// (igl.point_mesh_squared_distance, igl.hausdorff).
//
// Test infrastructure only (loaded by tests/ through oracle/metrics.py): it
// checks the GPU metrics, it is never part of the product path.
//
// pm_point_tri_sqdist: Ericson's closest point on a triangle (Real-Time
// Collision Detection 5.1.5, the Voronoi-region test libigl's
// point_simplex_squared_distance follows), in double.
// pm_point_tri_sqdist_alt: an independent construction used to pin it: the
// projection onto the triangle's plane when it falls inside (sign tests of
// the sub-triangle normals), otherwise the nearest of the three edges
// (clamped segment projection).
// pm_point_mesh_sqdist: brute force over every face; the nearest face is the
// first one (in face order) reaching the minimum.
// pm_hausdorff: max over the vertices of each mesh of the squared distance to
// the other surface, square-rooted (libigl hausdorff.cpp: vertex-to-surface
// both ways).
#include <algorithm>
#include <cmath>

#include <omp.h>

namespace {

struct P3 {
    double x, y, z;
};
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 mul(double s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline P3 at(const double* p, long long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

P3 closest_ericson(P3 p, P3 a, P3 b, P3 c) {
    const P3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;  // vertex region a
    const P3 bp = sub(p, b);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;  // vertex region b
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return add(a, mul(d1 / (d1 - d3), ab));  // edge ab
    const P3 cp = sub(p, c);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;  // vertex region c
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return add(a, mul(d2 / (d2 - d6), ac));  // edge ac
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
        return add(b, mul((d4 - d3) / ((d4 - d3) + (d5 - d6)), sub(c, b)));  // edge bc
    const double den = 1.0 / (va + vb + vc);  // interior
    const double v = vb * den, w = vc * den;
    return add(add(a, mul(v, ab)), mul(w, ac));  // (a + v ab) + w ac, Eigen's left-to-right sum
}

double sqdist_ericson(P3 p, P3 a, P3 b, P3 c) {
    const P3 d = sub(closest_ericson(p, a, b, c), p);
    return dot(d, d);
}

double seg_sqdist(P3 p, P3 a, P3 b) {
    const P3 ab = sub(b, a);
    const double L = dot(ab, ab);
    double t = L > 0.0 ? dot(sub(p, a), ab) / L : 0.0;
    t = std::min(1.0, std::max(0.0, t));
    const P3 d = sub(add(a, mul(t, ab)), p);
    return dot(d, d);
}

double sqdist_alt(P3 p, P3 a, P3 b, P3 c) {
    const P3 n = cross(sub(b, a), sub(c, a));
    const double nn = dot(n, n);
    if (nn > 0.0) {
        const double h = dot(sub(p, a), n) / nn;
        const P3 q = sub(p, mul(h, n));  // projection onto the plane
        const bool in = dot(cross(sub(b, a), sub(q, a)), n) >= 0.0 && dot(cross(sub(c, b), sub(q, b)), n) >= 0.0 &&
                        dot(cross(sub(a, c), sub(q, c)), n) >= 0.0;
        if (in) return h * h * nn;
    }
    return std::min(seg_sqdist(p, a, b), std::min(seg_sqdist(p, b, c), seg_sqdist(p, c, a)));
}

}  // namespace

extern "C" {

double pm_point_tri_sqdist(const double* p, const double* a, const double* b, const double* c) {
    return sqdist_ericson(at(p, 0), at(a, 0), at(b, 0), at(c, 0));
}

double pm_point_tri_sqdist_alt(const double* p, const double* a, const double* b, const double* c) {
    return sqdist_alt(at(p, 0), at(a, 0), at(b, 0), at(c, 0));
}

void pm_point_mesh_sqdist(const double* P, long long nP, const double* V, const int* F, int nF, double* sqd,
                          int* I) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long long i = 0; i < nP; ++i) {
        const P3 p = at(P, i);
        double best = HUGE_VAL;
        int bf = -1;
        for (int f = 0; f < nF; ++f) {
            const double d = sqdist_ericson(p, at(V, F[3 * (long long)f]), at(V, F[3 * (long long)f + 1]),
                                            at(V, F[3 * (long long)f + 2]));
            if (d < best) {
                best = d;
                bf = f;
            }
        }
        sqd[i] = best;
        if (I) I[i] = bf;
    }
}

double pm_hausdorff(const double* VA, int nVA, const int* FA, int nFA, const double* VB, int nVB, const int* FB,
                    int nFB) {
    double m = 0.0;
    double* tmp = new double[std::max(nVA, nVB)];
    pm_point_mesh_sqdist(VA, nVA, VB, FB, nFB, tmp, nullptr);
    for (int i = 0; i < nVA; ++i) m = std::max(m, tmp[i]);
    pm_point_mesh_sqdist(VB, nVB, VA, FA, nFA, tmp, nullptr);
    for (int i = 0; i < nVB; ++i) m = std::max(m, tmp[i]);
    delete[] tmp;
    return std::sqrt(m);
}

}  // extern "C"
