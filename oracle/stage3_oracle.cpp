// CPU oracle for PaMO stage 3 (barrier-energy safe projection).
//
// TEST INFRASTRUCTURE ONLY: this file is a plain restatement of the
// reference's Warp kernels, used by tests/ (and the CPU baseline leg of
// bench.py) to check the GPU library.  The product path never calls it.
//
// Reference files restated (simp_cuda/safe_project/src/pamo_safe_project/):
//   system.py:304-470 (detect / energy / diff / hess_dx / clamp / ccd /
//     line search / step), cg_solver.py:46-158, kernels/solver_kernels.py,
//   kernels/geometry_kernels.py:82-143 (closest point, bounds),
//   kernels/energy_kernels/contact_detection.py:119-216 (brute force here,
//     with the same box test and distance threshold),
//   kernels/distance_kernels/distance_kernels.py:8-494 (classification and
//     distances), kernels/ccd_kernels.py:8-110,
//   kernels/energy_kernels/{distance,elastic,hinge,collision}_energy.py,
//   kernels/utils_kernels.py:21-39 (SPD projection: own eigen-clamp of the
//     9x9 leading block, then translation invariance).
//
// State (positions, p, r, z, v) is held in T (double or float); energies
// and dot products accumulate in double.  Exact derivatives of the distance
// functions and of the hinge energy are taken by forward-mode automatic
// differentiation (Jet: value, 12-gradient, 12x12 Hessian) rather than by
// restating the reference's generated closed forms; the membrane uses the
// reference's frozen-basis Hessian-vector product, the barrier its
// Gauss-Newton Hessian.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace s3o {

// ------------------------------------------------------------ jets
struct Jet {
    Jet() = default;
    Jet(double x) : v(x) {}
    double v = 0.0;
    double g[12] = {};
    double h[12][12] = {};
};
static Jet jconst(double v) { Jet r; r.v = v; return r; }
static Jet jvar(double v, int k) { Jet r; r.v = v; r.g[k] = 1.0; return r; }
static Jet unary(const Jet& a, double f, double f1, double f2) {
    Jet r;
    r.v = f;
    for (int i = 0; i < 12; ++i) r.g[i] = f1 * a.g[i];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) r.h[i][j] = f1 * a.h[i][j] + f2 * a.g[i] * a.g[j];
    return r;
}
static Jet operator+(const Jet& a, const Jet& b) {
    Jet r; r.v = a.v + b.v;
    for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] + b.g[i];
    for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) r.h[i][j] = a.h[i][j] + b.h[i][j];
    return r;
}
static Jet operator-(const Jet& a, const Jet& b) {
    Jet r; r.v = a.v - b.v;
    for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] - b.g[i];
    for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) r.h[i][j] = a.h[i][j] - b.h[i][j];
    return r;
}
static Jet operator-(const Jet& a) { return unary(a, -a.v, -1.0, 0.0); }
static Jet operator*(const Jet& a, const Jet& b) {
    Jet r; r.v = a.v * b.v;
    for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j)
            r.h[i][j] = a.h[i][j] * b.v + a.v * b.h[i][j] + a.g[i] * b.g[j] + b.g[i] * a.g[j];
    return r;
}
static Jet inv(const Jet& a) { return unary(a, 1.0 / a.v, -1.0 / (a.v * a.v), 2.0 / (a.v * a.v * a.v)); }
static Jet operator/(const Jet& a, const Jet& b) { return a * inv(b); }
static Jet jsqrt(const Jet& a) {
    double s = std::sqrt(a.v);
    return unary(a, s, 0.5 / s, -0.25 / (s * a.v));
}
static Jet jabs(const Jet& a) { return a.v < 0 ? -a : a; }
static Jet jatan2(const Jet& y, const Jet& x) {
    const double r2 = x.v * x.v + y.v * y.v;
    const double ty = x.v / r2, tx = -y.v / r2;  // d/dy, d/dx
    const double r4 = r2 * r2;
    const double tyy = -2.0 * x.v * y.v / r4, txx = 2.0 * x.v * y.v / r4, txy = (y.v * y.v - x.v * x.v) / r4;
    Jet r;
    r.v = std::atan2(y.v, x.v);
    for (int i = 0; i < 12; ++i) r.g[i] = ty * y.g[i] + tx * x.g[i];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j)
            r.h[i][j] = ty * y.h[i][j] + tx * x.h[i][j] + tyy * y.g[i] * y.g[j] + txx * x.g[i] * x.g[j] +
                        txy * (x.g[i] * y.g[j] + y.g[i] * x.g[j]);
    return r;
}

// ------------------------------------------------------------ vectors over a scalar
template <typename S>
struct V3 {
    S x, y, z;
};
template <typename S> V3<S> operator+(const V3<S>& a, const V3<S>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename S> V3<S> operator-(const V3<S>& a, const V3<S>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename S> V3<S> operator*(const S& s, const V3<S>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename S> S dot(const V3<S>& a, const V3<S>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename S> V3<S> cross(const V3<S>& a, const V3<S>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static double ssqrt(double a) { return std::sqrt(a); }
static float ssqrt(float a) { return std::sqrt(a); }
static Jet ssqrt(const Jet& a) { return jsqrt(a); }
static double sval(double a) { return a; }
static double sval(float a) { return a; }
static double sval(const Jet& a) { return a.v; }
template <typename S> S length(const V3<S>& a) { return ssqrt(dot(a, a)); }
template <typename S> V3<S> normalize(const V3<S>& a) {
    S l = length(a);
    if (sval(l) > 0.0) {
        S il = S(1) / l;
        return il * a;
    }
    return {S(0), S(0), S(0)};
}
template <typename S> S sabs(const S& a) { return a < S(0) ? -a : a; }
static Jet sabs(const Jet& a) { return jabs(a); }

// ------------------------------------------------------------ classification / distances
enum { PT = 0, PP01 = 1, PP02 = 2, PE012 = 3, PP03 = 4, PE031 = 5, PE023 = 6 };
enum { EPP02 = 0, EPP03 = 1, EPP12 = 2, EPP13 = 3, EPE023 = 4, EPE123 = 5, EPE201 = 6, EPE301 = 7, EEE = 8 };
enum { BLOCK_PT = 3, BLOCK_EE = 4 };

template <typename S> S d_pp(const V3<S>& a, const V3<S>& b) { return length(b - a); }
template <typename S> S d_pe(const V3<S>& p, const V3<S>& a, const V3<S>& b) {
    return length(cross(a - p, b - p)) / length(a - b);
}
template <typename S> S d_pt(const V3<S>& p, const V3<S>& a, const V3<S>& b, const V3<S>& c) {
    return sabs(dot(p - a, normalize(cross(b - a, c - a))));
}
template <typename S> S d_ee(const V3<S>& a, const V3<S>& b, const V3<S>& c, const V3<S>& d) {
    return sabs(dot(a - c, normalize(cross(b - a, d - c))));
}

// distance_kernels.py:75-226
template <typename T>
int pt_classify(V3<T> x0, V3<T> x1, V3<T> x2, V3<T> x3) {
    V3<T> x10 = x0 - x1, x12 = x2 - x1, x13 = x3 - x1;
    T area_sq = dot(cross(x12, x13), cross(x12, x13));
    if (area_sq <= T(1e-6) * (dot(x12, x12) * dot(x13, x13)) || area_sq <= T(1e-16)) {
        T d1 = length(x10), d2 = length(x0 - x2), d3 = length(x0 - x3);
        if (d1 < d2) return d1 < d3 ? PP01 : PP03;
        return d2 < d3 ? PP02 : PP03;
    }
    V3<T> n = normalize(cross(x12, x13));
    V3<T> xp = x0 - dot(n, x10) * n;
    x10 = xp - x1;
    V3<T> x20 = xp - x2, x30 = xp - x3;
    T a00 = dot(x12, x12), a01 = dot(x12, x13), a11 = dot(x13, x13), b0 = dot(x10, x12), b1 = dot(x10, x13);
    T det = a00 * a11 - a01 * a01;
    T beta = b0 * a11 - b1 * a01, gamma = a00 * b1 - b0 * a01, alpha = det - beta - gamma;
    int nneg = (alpha < 0) + (beta < 0) + (gamma < 0);
    if (nneg == 0) return PT;
    if (nneg == 1) {
        if (alpha < 0) {
            V3<T> x23 = x3 - x2;
            T r = dot(x23, x0 - x2);
            return r < 0 ? PP02 : (r > dot(x23, x23) ? PP03 : PE023);
        }
        if (beta < 0) {
            V3<T> x31 = x1 - x3;
            T r = dot(x31, x0 - x3);
            return r < 0 ? PP03 : (r > dot(x31, x31) ? PP01 : PE031);
        }
        T r = dot(x12, x10);
        return r < 0 ? PP01 : (r > dot(x12, x12) ? PP02 : PE012);
    }
    if (alpha >= 0) {
        T p = dot(x12, x10), q = dot(x13, x10);
        if (p > 0 && q <= 0) return dot(x12, x20) > 0 ? PP02 : PE012;
        if (q > 0 && p <= 0) return dot(x13, x30) > 0 ? PP03 : PE031;
        return PP01;
    }
    if (beta >= 0) {
        V3<T> x23 = x3 - x2, x21 = x1 - x2;
        T p = dot(x23, x20), q = dot(x21, x20);
        if (p > 0 && q <= 0) return dot(x23, x30) > 0 ? PP03 : PE023;
        if (q > 0 && p <= 0) return dot(x21, x10) > 0 ? PP01 : PE012;
        return PP02;
    }
    V3<T> x31 = x1 - x3, x32 = x2 - x3;
    T p = dot(x31, x30), q = dot(x32, x30);
    if (p > 0 && q <= 0) return dot(x31, x10) > 0 ? PP01 : PE031;
    if (q > 0 && p <= 0) return dot(x32, x20) > 0 ? PP02 : PE023;
    return PP03;
}

// distance_kernels.py:325-441
template <typename T>
int ee_classify(V3<T> x0, V3<T> x1, V3<T> x2, V3<T> x3, T sin_thres) {
    V3<T> e0 = x0 - x1, e1 = x2 - x3;
    T e0_sq = dot(e0, e0), e1_sq = dot(e1, e1);
    T a00 = e0_sq, a01 = -dot(e0, e1), a10 = dot(e0, e1), a11 = -e1_sq;
    T b0 = dot(x3 - x1, e0), b1 = dot(x3 - x1, e1);
    T mu0 = -(b0 * a11 - b1 * a01), mu1 = -(a00 * b1 - a10 * b0), detA = -(a00 * a11 - a01 * a10);
    T sn = length(cross(normalize(e0), normalize(e1)));
    if (mu0 >= 0 && mu0 <= detA && mu1 >= 0 && mu1 <= detA && sn >= sin_thres && e0_sq > T(1e-15) &&
        e1_sq > T(1e-15))
        return EEE;
    T best = T(1e9);
    int type = -1;
    auto take = [&](T d, int t) {
        if (d < best) { best = d; type = t; }
    };
    T t0 = dot(x0 - x3, e1);
    if (t0 >= e1_sq) take(d_pp(x0, x2), EPP02);
    else if (t0 <= 0) take(d_pp(x0, x3), EPP03);
    else take(d_pe(x0, x2, x3), EPE023);
    T t1 = dot(x1 - x3, e1);
    if (t1 >= e1_sq) take(d_pp(x1, x2), EPP12);
    else if (t1 <= 0) take(d_pp(x1, x3), EPP13);
    else take(d_pe(x1, x2, x3), EPE123);
    T t2 = dot(x2 - x1, e0);
    if (t2 >= e0_sq) take(d_pp(x2, x0), EPP02);
    else if (t2 <= 0) take(d_pp(x2, x1), EPP12);
    else take(d_pe(x2, x0, x1), EPE201);
    T t3 = dot(x3 - x1, e0);
    if (t3 >= e0_sq) take(d_pp(x3, x0), EPP03);
    else if (t3 <= 0) take(d_pp(x3, x1), EPP13);
    else take(d_pe(x3, x0, x1), EPE301);
    return type;
}

template <typename S>
S pt_pair_distance(const V3<S>& x0, const V3<S>& x1, const V3<S>& x2, const V3<S>& x3, int t) {
    switch (t) {
        case PT: return d_pt(x0, x1, x2, x3);
        case PP01: return d_pp(x0, x1);
        case PP02: return d_pp(x0, x2);
        case PE012: return d_pe(x0, x1, x2);
        case PP03: return d_pp(x0, x3);
        case PE031: return d_pe(x0, x3, x1);
        case PE023: return d_pe(x0, x2, x3);
    }
    return S(1e9);
}
template <typename S>
S ee_pair_distance(const V3<S>& x0, const V3<S>& x1, const V3<S>& x2, const V3<S>& x3, int t) {
    switch (t) {
        case EPP02: return d_pp(x0, x2);
        case EPP03: return d_pp(x0, x3);
        case EPP12: return d_pp(x1, x2);
        case EPP13: return d_pp(x1, x3);
        case EPE023: return d_pe(x0, x2, x3);
        case EPE123: return d_pe(x1, x2, x3);
        case EPE201: return d_pe(x2, x0, x1);
        case EPE301: return d_pe(x3, x0, x1);
        case EEE: return d_ee(x0, x1, x2, x3);
    }
    return S(1e9);
}
static Jet pt_pair_distance(const V3<Jet>& x0, const V3<Jet>& x1, const V3<Jet>& x2, const V3<Jet>& x3, int t) {
    switch (t) {
        case PT: return d_pt(x0, x1, x2, x3);
        case PP01: return d_pp(x0, x1);
        case PP02: return d_pp(x0, x2);
        case PE012: return d_pe(x0, x1, x2);
        case PP03: return d_pp(x0, x3);
        case PE031: return d_pe(x0, x3, x1);
        case PE023: return d_pe(x0, x2, x3);
    }
    return jconst(1e9);
}
static Jet ee_pair_distance(const V3<Jet>& x0, const V3<Jet>& x1, const V3<Jet>& x2, const V3<Jet>& x3, int t) {
    switch (t) {
        case EPP02: return d_pp(x0, x2);
        case EPP03: return d_pp(x0, x3);
        case EPP12: return d_pp(x1, x2);
        case EPP13: return d_pp(x1, x3);
        case EPE023: return d_pe(x0, x2, x3);
        case EPE123: return d_pe(x1, x2, x3);
        case EPE201: return d_pe(x2, x0, x1);
        case EPE301: return d_pe(x3, x0, x1);
        case EEE: return d_ee(x0, x1, x2, x3);
    }
    return jconst(1e9);
}

static void jet_inputs(const double* x12, V3<Jet> X[4]) {
    for (int a = 0; a < 4; ++a)
        X[a] = {jvar(x12[3 * a], 3 * a), jvar(x12[3 * a + 1], 3 * a + 1), jvar(x12[3 * a + 2], 3 * a + 2)};
}

// gradient of the classified distance (type fixed), 12 values
static void pair_grad(const double* x12, bool is_pt, int type, double* g12) {
    V3<Jet> X[4];
    jet_inputs(x12, X);
    Jet d = is_pt ? pt_pair_distance(X[0], X[1], X[2], X[3], type) : ee_pair_distance(X[0], X[1], X[2], X[3], type);
    for (int k = 0; k < 12; ++k) g12[k] = d.g[k];
}

// IPC barrier (collision_energy.py:8-28)
static double barrier(double d, double dh) { return d < dh ? -(d - dh) * (d - dh) * std::log(d / dh) : 0.0; }
static double barrier_d1(double d, double dh) {
    return d < dh ? -2.0 * (d - dh) * std::log(d / dh) - (d - dh) * (d - dh) / d : 0.0;
}
static double barrier_d2(double d, double dh) {
    // second derivative of -(d - dh)^2 ln(d / dh)
    const double t = d - dh;
    return d < dh ? -2.0 * std::log(d / dh) - 4.0 * t / d + t * t / (d * d) : 0.0;
}

// compute_hinge_angle (hinge_energy.py:4-18), generic scalar
template <typename S>
S hinge_angle_s(const V3<S>& x0, const V3<S>& x1, const V3<S>& x2, const V3<S>& x3);
template <>
double hinge_angle_s<double>(const V3<double>& x0, const V3<double>& x1, const V3<double>& x2, const V3<double>& x3) {
    V3<double> e = normalize(x2 - x1);
    V3<double> n1 = normalize(cross(e, x0 - x1)), n2 = normalize(cross(x3 - x1, e));
    return std::atan2(dot(cross(n1, n2), e), dot(n1, n2));
}
template <>
Jet hinge_angle_s<Jet>(const V3<Jet>& x0, const V3<Jet>& x1, const V3<Jet>& x2, const V3<Jet>& x3) {
    V3<Jet> e = normalize(x2 - x1);
    V3<Jet> n1 = normalize(cross(e, x0 - x1)), n2 = normalize(cross(x3 - x1, e));
    return jatan2(dot(cross(n1, n2), e), dot(n1, n2));
}

// closest point on a triangle by exhaustive feature test (interior
// projection if inside, else the nearest of the three segments)
template <typename T>
V3<T> closest_on_tri(V3<T> p, V3<T> a, V3<T> b, V3<T> c) {
    auto seg = [&](V3<T> u, V3<T> w) {
        V3<T> e = w - u;
        T ee = dot(e, e);
        T t = ee > 0 ? dot(p - u, e) / ee : T(0);
        t = std::min(std::max(t, T(0)), T(1));
        return u + t * e;
    };
    V3<T> n = cross(b - a, c - a);
    T nn = dot(n, n);
    if (nn > 0) {
        V3<T> q = p - (dot(p - a, n) / nn) * n;
        T w0 = dot(cross(b - q, c - q), n), w1 = dot(cross(c - q, a - q), n), w2 = dot(cross(a - q, b - q), n);
        if (w0 >= 0 && w1 >= 0 && w2 >= 0) return q;
    }
    V3<T> c0 = seg(a, b), c1 = seg(b, c), c2 = seg(c, a);
    T d0 = dot(c0 - p, c0 - p), d1 = dot(c1 - p, c1 - p), d2 = dot(c2 - p, c2 - p);
    if (d0 <= d1 && d0 <= d2) return c0;
    return d1 <= d2 ? c1 : c2;
}

// ------------------------------------------------------------ eigen-clamp SPD projection
// cyclic Jacobi on a symmetric n x n matrix (double), eigenvalues clamped at 0
static void spd_clamp(double* A, int n) {
    std::vector<double> V(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
        if (off < 1e-30) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<double> lam(n);
    for (int k = 0; k < n; ++k) lam[k] = std::max(A[k * n + k], 0.0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += V[r * n + k] * lam[k] * V[c * n + k];
            A[r * n + c] = s;
        }
}

// 12x12 hinge Hessian H (row-major, vertex-major dofs) -> projected blocks
static void hinge_project(double* H) {
    double A[81];
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) A[r * 9 + c] = 0.5 * (H[r * 12 + c] + H[c * 12 + r]);
    spd_clamp(A, 9);
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) H[r * 12 + c] = A[r * 9 + c];
    // block (3, j) = -sum_{a<3} block (a, j); block (j, 3) = block (3, j)^T;
    // block (3, 3) = -sum_{a<3} block (a, 3)
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0.0;
                for (int a = 0; a < 3; ++a) s += H[(3 * a + r) * 12 + 3 * j + c];
                H[(9 + r) * 12 + 3 * j + c] = -s;
            }
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) H[(3 * j + r) * 12 + 9 + c] = H[(9 + c) * 12 + 3 * j + r];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int a = 0; a < 3; ++a) s += H[(3 * a + r) * 12 + 9 + c];
            H[(9 + r) * 12 + 9 + c] = -s;
        }
}

// ------------------------------------------------------------ the system
struct Params {
    int mask, n_newton, n_cg, n_ls, max_blocks, ccd_max_iters;
    double d_hat, kappa, radius0, ee_thres, ccd_slack, ccd_thick, E, nu, k_hinge, k_m2gt;
};
enum { E_M2GT = 1, E_ELASTIC = 2, E_HINGE = 4, E_COLLISION = 8, E_DENSE = 16 };

struct Contact {
    int i[4];
    int type1;
};

template <typename T>
struct System {
    using P3 = V3<T>;
    Params P{};
    double mu = 0, la = 0;
    int nV = 0;
    std::vector<P3> q, qn, qd, target, p;
    std::vector<int> F, E, H;  // 3/face, 2/edge, 4/hinge
    std::vector<double> vor;
    std::vector<T> area;
    std::vector<T> idm;  // 4 per face
    std::vector<T> rest_angle, rest_elen;
    std::vector<P3> gtV;
    std::vector<int> gtF;
    std::vector<double> dA, db;
    double dc = 0;
    bool do_nothing = false;
    T radius = 0;
    std::vector<Contact> C;
    std::vector<T> cd;
    std::vector<int> ctype2;
    std::vector<P3> cdd;           // 4 per contact
    std::vector<double> hblocks;   // 144 per hinge (projected, row-major 12x12)
    std::vector<double> hraw;      // unprojected
    std::vector<P3> grad, diag;
    double energy = 0, energy_prev = 0;
    T ccd_step = 1;
    long long stat_contacts = 0, stat_newton = 0;

    static P3 mk(double x, double y, double z) { return {T(x), T(y), T(z)}; }
    P3 at(const std::vector<P3>& v, int i) const { return v[i]; }

    void set_mesh(const double* V, int n, const int* f, int nf, const int* e, int ne, const int* h, int nh,
                  const double* vr, int dn) {
        nV = n;
        q.resize(n); qn.resize(n); qd.resize(n); target.assign(n, mk(1e9, 1e9, 1e9)); p.assign(n, mk(0, 0, 0));
        grad.assign(n, mk(0, 0, 0)); diag.assign(n, mk(0, 0, 0));
        for (int i = 0; i < n; ++i) q[i] = mk(V[3 * i], V[3 * i + 1], V[3 * i + 2]);
        qd = q;
        F.assign(f, f + 3 * nf);
        E.assign(e, e + 2 * ne);
        H.assign(h, h + 4 * nh);
        vor.assign(vr, vr + n);
        do_nothing = dn != 0;
        radius = T(P.radius0);
        // elastic_preprocess_kernel (elastic_energy.py:8-38)
        area.resize(nf);
        idm.resize(4 * nf);
        for (int t = 0; t < nf; ++t) {
            P3 x0 = q[F[3 * t]], x1 = q[F[3 * t + 1]], x2 = q[F[3 * t + 2]];
            P3 x01 = x1 - x0, x02 = x2 - x0;
            area[t] = T(0.5) * length(cross(x01, x02));
            P3 b1 = normalize(x01);
            P3 b2 = normalize(x02 - dot(x02, b1) * b1);
            T a = dot(b1, x01), b = dot(b1, x02), c = dot(b2, x01), d = dot(b2, x02);
            T det = a * d - b * c;
            idm[4 * t] = d / det; idm[4 * t + 1] = -b / det; idm[4 * t + 2] = -c / det; idm[4 * t + 3] = a / det;
        }
        rest_angle.resize(nh);
        rest_elen.resize(nh);
        for (int k = 0; k < nh; ++k) {
            V3<double> X[4];
            for (int a = 0; a < 4; ++a) {
                P3 v = q[H[4 * k + a]];
                X[a] = {(double)v.x, (double)v.y, (double)v.z};
            }
            rest_angle[k] = T(hinge_angle_s<double>(X[0], X[1], X[2], X[3]));
            rest_elen[k] = length(q[H[4 * k + 2]] - q[H[4 * k + 1]]);
        }
        C.clear();
        stat_contacts = stat_newton = 0;
    }

    // ---- update_closest_point_on_target_bvh_kernel (geometry_kernels.py:82-94), brute force
    void update_target() {
        if (!(P.mask & E_M2GT)) return;
        const int ng = (int)gtF.size() / 3;
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < nV; ++i) {
            double best = 1e300;
            P3 bp = mk(1e9, 1e9, 1e9);
            for (int f = 0; f < ng; ++f) {
                P3 c = closest_on_tri(q[i], gtV[gtF[3 * f]], gtV[gtF[3 * f + 1]], gtV[gtF[3 * f + 2]]);
                P3 d = c - q[i];
                double d2 = (double)dot(d, d);
                if (d2 < best) { best = d2; bp = c; }
            }
            target[i] = bp;
        }
    }

    // ---- contact detection (contact_detection.py:119-216, energy.py:896-948), brute force
    static bool box_hit(P3 lo, P3 hi, P3 blo, P3 bhi) {
        return !(hi.x < blo.x || bhi.x < lo.x || hi.y < blo.y || bhi.y < lo.y || hi.z < blo.z || bhi.z < lo.z);
    }
    static P3 vmin(P3 a, P3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
    static P3 vmax(P3 a, P3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

    void detect() {
        C.clear();
        if (P.mask & E_COLLISION) {
            const int nf = (int)F.size() / 3, ne = (int)E.size() / 2;
            std::vector<P3> tlo(nf), thi(nf), elo(ne), ehi(ne);
            for (int t = 0; t < nf; ++t) {
                P3 a = q[F[3 * t]], b = q[F[3 * t + 1]], c = q[F[3 * t + 2]];
                tlo[t] = vmin(a, vmin(b, c));
                thi[t] = vmax(a, vmax(b, c));
            }
            for (int e = 0; e < ne; ++e) {
                P3 a = q[E[2 * e]], b = q[E[2 * e + 1]];
                elo[e] = vmin(a, b);
                ehi[e] = vmax(a, b);
            }
            while (true) {
                C.clear();
                const T thr = T(2) * radius + T(P.d_hat);
                const P3 r3 = {thr, thr, thr};
                for (int i0 = 0; i0 < nV; ++i0) {
                    P3 lo = q[i0] - r3, hi = q[i0] + r3;
                    for (int t = 0; t < nf; ++t) {
                        if (!box_hit(lo, hi, tlo[t], thi[t])) continue;
                        int a = F[3 * t], b = F[3 * t + 1], c = F[3 * t + 2];
                        if (i0 == a || i0 == b || i0 == c) continue;
                        int ty = pt_classify(q[i0], q[a], q[b], q[c]);
                        T d = pt_pair_distance(q[i0], q[a], q[b], q[c], ty);
                        if (d <= thr) C.push_back({{i0, a, b, c}, BLOCK_PT});
                    }
                }
                for (int e0 = 0; e0 < ne; ++e0) {
                    P3 lo = elo[e0] - r3, hi = ehi[e0] + r3;
                    int a0 = E[2 * e0], a1 = E[2 * e0 + 1];
                    for (int e1 = e0 + 1; e1 < ne; ++e1) {
                        if (!box_hit(lo, hi, elo[e1], ehi[e1])) continue;
                        int b0 = E[2 * e1], b1 = E[2 * e1 + 1];
                        if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) continue;
                        int ty = ee_classify(q[a0], q[a1], q[b0], q[b1], T(P.ee_thres));
                        T d = ee_pair_distance(q[a0], q[a1], q[b0], q[b1], ty);
                        if (d <= thr) C.push_back({{a0, a1, b0, b1}, BLOCK_EE});
                    }
                }
                if ((long long)C.size() > P.max_blocks) {
                    radius = radius / T(2);
                    continue;
                }
                break;
            }
        }
        qd = q;
        stat_contacts += (long long)C.size();
    }

    // ---- energies
    double e_m2gt(const std::vector<P3>& x) const {
        double s = 0;
        for (int i = 0; i < nV; ++i) {
            P3 d = x[i] - target[i];
            s += 0.5 * (double)dot(d, d) * vor[i] * P.k_m2gt;
        }
        return s;
    }
    // membrane of one face in FP64 from T positions
    double e_elastic_face(const std::vector<P3>& x, int t) const {
        V3<double> x0 = cvt(x[F[3 * t]]), x1 = cvt(x[F[3 * t + 1]]), x2 = cvt(x[F[3 * t + 2]]);
        V3<double> x01 = x1 - x0, x02 = x2 - x0;
        V3<double> b1 = normalize(x01), b2 = normalize(x02 - dot(x02, b1) * b1);
        double Ds[4] = {dot(b1, x01), dot(b1, x02), dot(b2, x01), dot(b2, x02)};
        double Dm[4] = {(double)idm[4 * t], (double)idm[4 * t + 1], (double)idm[4 * t + 2], (double)idm[4 * t + 3]};
        double Fm[4];
        mul22(Ds, Dm, Fm);
        double Em[4];
        green(Fm, Em);
        double tr = Em[0] + Em[3];
        return (double)area[t] * (mu * (Em[0] * Em[0] + Em[1] * Em[1] + Em[2] * Em[2] + Em[3] * Em[3]) +
                                  0.5 * la * tr * tr);
    }
    double e_elastic(const std::vector<P3>& x) const {
        double s = 0;
        for (int t = 0; t < (int)F.size() / 3; ++t) s += e_elastic_face(x, t);
        return s;
    }
    double e_hinge(const std::vector<P3>& x) const {
        double s = 0;
        for (int k = 0; k < (int)H.size() / 4; ++k) {
            double th = hinge_angle_s<double>(cvt(x[H[4 * k]]), cvt(x[H[4 * k + 1]]), cvt(x[H[4 * k + 2]]),
                                              cvt(x[H[4 * k + 3]]));
            double dt = th - (double)rest_angle[k];
            s += 0.5 * P.k_hinge * dt * dt * (double)rest_elen[k];
        }
        return s;
    }
    // collision energy; also refreshes per-contact distance and type
    double e_collision(const std::vector<P3>& x) {
        double s = 0;
        cd.resize(C.size());
        ctype2.resize(C.size());
        for (size_t b = 0; b < C.size(); ++b) {
            const Contact& c = C[b];
            P3 x0 = x[c.i[0]], x1 = x[c.i[1]], x2 = x[c.i[2]], x3 = x[c.i[3]];
            int t2;
            T d;
            if (c.type1 == BLOCK_PT) {
                t2 = pt_classify(x0, x1, x2, x3);
                d = pt_pair_distance(x0, x1, x2, x3, t2);
            } else {
                t2 = ee_classify(x0, x1, x2, x3, T(P.ee_thres));
                d = ee_pair_distance(x0, x1, x2, x3, t2);
            }
            ctype2[b] = t2;
            cd[b] = d;
            if ((double)d < P.d_hat) s += barrier((double)d, P.d_hat) * P.kappa;
        }
        return s;
    }
    double e_dense(const std::vector<P3>& x) const {
        const int n3 = 3 * nV;
        double s = dc;
        for (int r = 0; r < n3; ++r) {
            double ax = 0;
            for (int k = 0; k < n3; ++k) ax += dA[(size_t)r * n3 + k] * comp(x, k);
            s += 0.5 * comp(x, r) * ax + db[r] * comp(x, r);
        }
        return s;
    }
    double energy_terms(const std::vector<P3>& x, int mask, double out[5]) {
        double e[5] = {0, 0, 0, 0, 0};
        if (mask & E_M2GT) e[0] = e_m2gt(x);
        if (mask & E_ELASTIC) e[1] = e_elastic(x);
        if (mask & E_HINGE) e[2] = e_hinge(x);
        if (mask & E_COLLISION) e[3] = e_collision(x);
        if ((mask & E_DENSE) && !dA.empty()) e[4] = e_dense(x);
        if (out) std::memcpy(out, e, sizeof(e));
        return e[0] + e[1] + e[2] + e[3] + e[4];
    }

    static V3<double> cvt(P3 a) { return {(double)a.x, (double)a.y, (double)a.z}; }
    static double comp(const std::vector<P3>& x, int k) {
        const P3& v = x[k / 3];
        return (double)(k % 3 == 0 ? v.x : (k % 3 == 1 ? v.y : v.z));
    }
    static void mul22(const double* A, const double* B, double* C) {
        C[0] = A[0] * B[0] + A[1] * B[2]; C[1] = A[0] * B[1] + A[1] * B[3];
        C[2] = A[2] * B[0] + A[3] * B[2]; C[3] = A[2] * B[1] + A[3] * B[3];
    }
    static void green(const double* Fm, double* Em) {
        Em[0] = 0.5 * (Fm[0] * Fm[0] + Fm[2] * Fm[2] - 1.0);
        Em[1] = 0.5 * (Fm[0] * Fm[1] + Fm[2] * Fm[3]);
        Em[2] = Em[1];
        Em[3] = 0.5 * (Fm[1] * Fm[1] + Fm[3] * Fm[3] - 1.0);
    }

    // membrane gradient with the local basis computed at basis_x (the true
    // gradient when basis_x == x; elastic_diff_kernel, elastic_energy.py:89-146)
    void elastic_grad_basis(const std::vector<P3>& x, const std::vector<P3>& basis_x, std::vector<V3<double>>& g) const {
        for (int t = 0; t < (int)F.size() / 3; ++t) {
            const int i0 = F[3 * t], i1 = F[3 * t + 1], i2 = F[3 * t + 2];
            V3<double> x01 = cvt(x[i1]) - cvt(x[i0]), x02 = cvt(x[i2]) - cvt(x[i0]);
            V3<double> y01 = cvt(basis_x[i1]) - cvt(basis_x[i0]), y02 = cvt(basis_x[i2]) - cvt(basis_x[i0]);
            V3<double> b1 = normalize(y01), b2 = normalize(y02 - dot(y02, b1) * b1);
            double Ds[4] = {dot(b1, x01), dot(b1, x02), dot(b2, x01), dot(b2, x02)};
            double Dm[4] = {(double)idm[4 * t], (double)idm[4 * t + 1], (double)idm[4 * t + 2], (double)idm[4 * t + 3]};
            double Fm[4], Em[4];
            mul22(Ds, Dm, Fm);
            green(Fm, Em);
            double tr = Em[0] + Em[3];
            double S[4] = {2 * mu * Em[0] + la * tr, 2 * mu * Em[1], 2 * mu * Em[2], 2 * mu * Em[3] + la * tr};
            double Pm[4], Hm[4];
            mul22(Fm, S, Pm);
            double DmT[4] = {Dm[0], Dm[2], Dm[1], Dm[3]};
            mul22(Pm, DmT, Hm);
            const double W = (double)area[t];
            V3<double> f1 = (W * Hm[0]) * b1 + (W * Hm[2]) * b2;
            V3<double> f2 = (W * Hm[1]) * b1 + (W * Hm[3]) * b2;
            g[i0] = g[i0] - f1 - f2;
            g[i1] = g[i1] + f1;
            g[i2] = g[i2] + f2;
        }
    }

    // ---- _compute_diff: grad (coefficient -1), diag, hinge blocks, contact gradients
    void diff() {
        const int mask = P.mask;
        std::vector<V3<double>> g(nV, {0, 0, 0}), dg(nV, {0, 0, 0});
        if (mask & E_M2GT)
            for (int i = 0; i < nV; ++i) {
                double w = vor[i] * P.k_m2gt;
                g[i] = g[i] + w * (cvt(q[i]) - cvt(target[i]));
                dg[i] = dg[i] + V3<double>{w, w, w};
            }
        if (mask & E_ELASTIC) elastic_grad_basis(q, q, g);
        const int nh = (int)H.size() / 4;
        hblocks.assign((size_t)144 * nh, 0.0);
        hraw.assign((size_t)144 * nh, 0.0);
        if (mask & E_HINGE)
            for (int k = 0; k < nh; ++k) {
                double x12[12];
                for (int a = 0; a < 4; ++a) {
                    V3<double> v = cvt(q[H[4 * k + a]]);
                    x12[3 * a] = v.x; x12[3 * a + 1] = v.y; x12[3 * a + 2] = v.z;
                }
                V3<Jet> X[4];
                jet_inputs(x12, X);
                Jet th = hinge_angle_s<Jet>(X[0], X[1], X[2], X[3]);
                Jet dt = th - jconst((double)rest_angle[k]);
                Jet e = jconst(0.5 * P.k_hinge * (double)rest_elen[k]) * dt * dt;
                double* Hk = &hraw[(size_t)144 * k];
                for (int r = 0; r < 12; ++r)
                    for (int c = 0; c < 12; ++c) Hk[r * 12 + c] = e.h[r][c];
                for (int a = 0; a < 4; ++a) {
                    int vi = H[4 * k + a];
                    g[vi] = g[vi] + V3<double>{e.g[3 * a], e.g[3 * a + 1], e.g[3 * a + 2]};
                    dg[vi] = dg[vi] + V3<double>{Hk[(3 * a) * 12 + 3 * a], Hk[(3 * a + 1) * 12 + 3 * a + 1],
                                                 Hk[(3 * a + 2) * 12 + 3 * a + 2]};
                }
                double* Bk = &hblocks[(size_t)144 * k];
                std::memcpy(Bk, Hk, sizeof(double) * 144);
                hinge_project(Bk);
            }
        cdd.assign(4 * C.size(), mk(0, 0, 0));
        if (mask & E_COLLISION)
            for (size_t b = 0; b < C.size(); ++b) {
                double d = (double)cd[b];
                if (d >= P.d_hat) continue;
                double x12[12], g12[12];
                for (int a = 0; a < 4; ++a) {
                    V3<double> v = cvt(q[C[b].i[a]]);
                    x12[3 * a] = v.x; x12[3 * a + 1] = v.y; x12[3 * a + 2] = v.z;
                }
                pair_grad(x12, C[b].type1 == BLOCK_PT, ctype2[b], g12);
                const double c1 = barrier_d1(d, P.d_hat) * P.kappa, c2 = barrier_d2(d, P.d_hat) * P.kappa;
                for (int a = 0; a < 4; ++a) {
                    V3<double> ga = {g12[3 * a], g12[3 * a + 1], g12[3 * a + 2]};
                    cdd[4 * b + a] = mk(ga.x, ga.y, ga.z);
                    int vi = C[b].i[a];
                    g[vi] = g[vi] + c1 * ga;
                    dg[vi] = dg[vi] + V3<double>{c2 * ga.x * ga.x, c2 * ga.y * ga.y, c2 * ga.z * ga.z};
                }
            }
        if ((mask & E_DENSE) && !dA.empty()) {
            const int n3 = 3 * nV;
            for (int r = 0; r < n3; ++r) {
                double ax = db[r];
                for (int k = 0; k < n3; ++k) ax += dA[(size_t)r * n3 + k] * comp(q, k);
                double* gr = r % 3 == 0 ? &g[r / 3].x : (r % 3 == 1 ? &g[r / 3].y : &g[r / 3].z);
                double* dr = r % 3 == 0 ? &dg[r / 3].x : (r % 3 == 1 ? &dg[r / 3].y : &dg[r / 3].z);
                *gr += ax;
                *dr += dA[(size_t)r * n3 + r];
            }
        }
        for (int i = 0; i < nV; ++i) {
            grad[i] = mk(-g[i].x, -g[i].y, -g[i].z);
            diag[i] = mk(dg[i].x, dg[i].y, dg[i].z);
        }
    }

    // ---- _compute_hess_dx
    void hvp(const std::vector<P3>& dx, std::vector<P3>& out) const {
        const int mask = P.mask;
        std::vector<V3<double>> o(nV, {0, 0, 0});
        if (mask & E_M2GT)
            for (int i = 0; i < nV; ++i) o[i] = o[i] + (vor[i] * P.k_m2gt) * cvt(dx[i]);
        if (mask & E_ELASTIC)
            for (int t = 0; t < (int)F.size() / 3; ++t) {  // elastic_hess_dx_kernel (elastic_energy.py:149-220)
                const int i0 = F[3 * t], i1 = F[3 * t + 1], i2 = F[3 * t + 2];
                V3<double> x01 = cvt(q[i1]) - cvt(q[i0]), x02 = cvt(q[i2]) - cvt(q[i0]);
                V3<double> d01 = cvt(dx[i1]) - cvt(dx[i0]), d02 = cvt(dx[i2]) - cvt(dx[i0]);
                V3<double> b1 = normalize(x01), b2 = normalize(x02 - dot(x02, b1) * b1);
                double Ds[4] = {dot(b1, x01), dot(b1, x02), dot(b2, x01), dot(b2, x02)};
                double dDs[4] = {dot(b1, d01), dot(b1, d02), dot(b2, d01), dot(b2, d02)};
                double Dm[4] = {(double)idm[4 * t], (double)idm[4 * t + 1], (double)idm[4 * t + 2],
                                (double)idm[4 * t + 3]};
                double Fm[4], dF[4], Em[4];
                mul22(Ds, Dm, Fm);
                mul22(dDs, Dm, dF);
                green(Fm, Em);
                double tr = Em[0] + Em[3];
                double FtdF[4] = {Fm[0] * dF[0] + Fm[2] * dF[2], Fm[0] * dF[1] + Fm[2] * dF[3],
                                  Fm[1] * dF[0] + Fm[3] * dF[2], Fm[1] * dF[1] + Fm[3] * dF[3]};
                double dE[4] = {FtdF[0], 0.5 * (FtdF[1] + FtdF[2]), 0.5 * (FtdF[1] + FtdF[2]), FtdF[3]};
                double trd = dE[0] + dE[3];
                double S[4] = {2 * mu * Em[0] + la * tr, 2 * mu * Em[1], 2 * mu * Em[2], 2 * mu * Em[3] + la * tr};
                double dS[4] = {2 * mu * dE[0] + la * trd, 2 * mu * dE[1], 2 * mu * dE[2], 2 * mu * dE[3] + la * trd};
                double a1[4], a2[4], dP[4], dH[4];
                mul22(dF, S, a1);
                mul22(Fm, dS, a2);
                for (int k = 0; k < 4; ++k) dP[k] = a1[k] + a2[k];
                double DmT[4] = {Dm[0], Dm[2], Dm[1], Dm[3]};
                mul22(dP, DmT, dH);
                const double W = (double)area[t];
                V3<double> f1 = (W * dH[0]) * b1 + (W * dH[2]) * b2;
                V3<double> f2 = (W * dH[1]) * b1 + (W * dH[3]) * b2;
                o[i0] = o[i0] - f1 - f2;
                o[i1] = o[i1] + f1;
                o[i2] = o[i2] + f2;
            }
        if (mask & E_HINGE)
            for (int k = 0; k < (int)H.size() / 4; ++k) {
                const double* B = &hblocks[(size_t)144 * k];
                double v[12];
                for (int a = 0; a < 4; ++a) {
                    V3<double> d = cvt(dx[H[4 * k + a]]);
                    v[3 * a] = d.x; v[3 * a + 1] = d.y; v[3 * a + 2] = d.z;
                }
                for (int a = 0; a < 4; ++a) {
                    double r[3] = {0, 0, 0};
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 12; ++j) r[i] += B[(3 * a + i) * 12 + j] * v[j];
                    int vi = H[4 * k + a];
                    o[vi] = o[vi] + V3<double>{r[0], r[1], r[2]};
                }
            }
        if (mask & E_COLLISION)
            for (size_t b = 0; b < C.size(); ++b) {
                double d = (double)cd[b];
                if (d >= P.d_hat) continue;
                double c2 = barrier_d2(d, P.d_hat) * P.kappa, s = 0;
                for (int a = 0; a < 4; ++a) s += dot(cvt(cdd[4 * b + a]), cvt(dx[C[b].i[a]]));
                for (int a = 0; a < 4; ++a) o[C[b].i[a]] = o[C[b].i[a]] + (s * c2) * cvt(cdd[4 * b + a]);
            }
        if ((mask & E_DENSE) && !dA.empty()) {
            const int n3 = 3 * nV;
            for (int r = 0; r < n3; ++r) {
                double ax = 0;
                for (int k = 0; k < n3; ++k) ax += dA[(size_t)r * n3 + k] * comp(dx, k);
                double* orr = r % 3 == 0 ? &o[r / 3].x : (r % 3 == 1 ? &o[r / 3].y : &o[r / 3].z);
                *orr += ax;
            }
        }
        out.resize(nV);
        for (int i = 0; i < nV; ++i) out[i] = mk(o[i].x, o[i].y, o[i].z);
    }

    static double vdot(const std::vector<P3>& a, const std::vector<P3>& b) {
        double s = 0;
        for (size_t i = 0; i < a.size(); ++i)
            s += (double)a[i].x * b[i].x + (double)a[i].y * b[i].y + (double)a[i].z * b[i].z;
        return s;
    }

    // ---- CGSolver.solve (cg_solver.py:96-158, solver_kernels.py:4-58)
    void cg() {
        std::vector<P3> r = grad, z(nV), v(nV), Av;
        p.assign(nV, mk(0, 0, 0));
        for (int i = 0; i < nV; ++i)
            z[i] = {r[i].x / (diag[i].x + T(1e-6)), r[i].y / (diag[i].y + T(1e-6)), r[i].z / (diag[i].z + T(1e-6))};
        v = z;
        double zr = vdot(z, r);
        for (int it = 0; it < P.n_cg; ++it) {
            hvp(v, Av);
            const double vAv = vdot(v, Av);
            const double zr_last = zr;
            if (vAv > 1e-16) {
                const T alpha = T(zr_last / vAv);
                for (int i = 0; i < nV; ++i) {
                    p[i] = p[i] + alpha * v[i];
                    r[i] = r[i] - alpha * Av[i];
                }
            }
            for (int i = 0; i < nV; ++i) z[i] = {r[i].x / diag[i].x, r[i].y / diag[i].y, r[i].z / diag[i].z};
            zr = vdot(z, r);
            const T s = T(zr / zr_last);
            for (int i = 0; i < nV; ++i) v[i] = z[i] + s * v[i];
        }
    }

    // ---- clamp_p_kernel (solver_kernels.py:77-94)
    void clamp() {
        for (int i = 0; i < nV; ++i) {
            P3 dq = qn[i] + p[i] - qd[i];
            if (length(dq) > T(P.radius0)) p[i] = T(P.radius0) * normalize(dq) + qd[i] - qn[i];
        }
    }

    // ---- accd_kernel (ccd_kernels.py:8-110)
    void ccd() {
        ccd_step = 1;
        for (size_t b = 0; b < C.size(); ++b) {
            const Contact& c = C[b];
            P3 x[4], v[4];
            for (int a = 0; a < 4; ++a) {
                x[a] = q[c.i[a]];
                v[a] = p[c.i[a]];
            }
            const bool is_pt = c.type1 == BLOCK_PT;
            P3 va = T(0.25) * (((v[0] + v[1]) + v[2]) + v[3]);
            for (int a = 0; a < 4; ++a) v[a] = v[a] - va;
            T lv = is_pt ? length(v[0]) + std::max(length(v[1]), std::max(length(v[2]), length(v[3])))
                         : std::max(length(v[0]), length(v[1])) + std::max(length(v[2]), length(v[3]));
            if (lv < T(1e-12)) continue;
            auto dist = [&]() {
                if (is_pt) return pt_pair_distance(x[0], x[1], x[2], x[3], pt_classify(x[0], x[1], x[2], x[3]));
                return ee_pair_distance(x[0], x[1], x[2], x[3], ee_classify(x[0], x[1], x[2], x[3], T(P.ee_thres)));
            };
            const T xi = T(P.ccd_thick);
            T d = dist();
            T dmt = std::max((d * d - xi * xi) / (d + xi), T(0));
            const T g = (T(1) - T(P.ccd_slack)) * dmt;
            T t = 0, tl = T(P.ccd_slack) * dmt / lv;
            for (int it = 0;; ++it) {
                for (int a = 0; a < 4; ++a) x[a] = x[a] + tl * v[a];
                d = dist();
                dmt = std::max((d * d - xi * xi) / (d + xi), T(0));
                if ((it > 0 && dmt <= g) || it >= P.ccd_max_iters) {
                    ccd_step = std::min(ccd_step, t);
                    break;
                }
                t = t + tl;
                if (t > T(1)) break;
                tl = T(0.8) * dmt / lv;
            }
        }
    }

    // ---- step (system.py:411-470)
    void step(bool keep_target = false) {
        if (do_nothing) return;
        if (!keep_target) update_target();
        for (int it = 0; it < P.n_newton; ++it) {
            ++stat_newton;
            qn = q;
            detect();
            energy = energy_terms(q, P.mask, nullptr);
            diff();
            cg();
            clamp();
            ccd();
            energy_prev = energy;
            for (int h = 0; h < P.n_ls; ++h) {
                if (!(h > 0 && energy < energy_prev)) {
                    const T f = ccd_step * T(std::pow(0.5, h));
                    for (int i = 0; i < nV; ++i) q[i] = qn[i] + f * p[i];
                }
                energy = energy_terms(q, P.mask, nullptr);
            }
        }
    }
};

template <typename S> V3<S> operator-(const V3<S>& a) { return {-a.x, -a.y, -a.z}; }

struct Handle {
    int fp64;
    System<double> d;
    System<float> f;
};

}  // namespace s3o

using s3o::Handle;

#define DISPATCH(h, ...)                   \
    do {                                   \
        if ((h)->fp64) {                   \
            auto& S = (h)->d;              \
            __VA_ARGS__;                   \
        } else {                           \
            auto& S = (h)->f;              \
            __VA_ARGS__;                   \
        }                                  \
    } while (0)

template <typename Sys>
static void put_vec(const Sys& S, const decltype(Sys::q)& v, double* out) {
    for (size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = (double)v[i].x;
        out[3 * i + 1] = (double)v[i].y;
        out[3 * i + 2] = (double)v[i].z;
    }
    (void)S;
}
template <typename Sys>
static void get_vec(Sys& S, const double* in, decltype(Sys::q)& v) {
    v.resize(S.nV);
    for (int i = 0; i < S.nV; ++i) v[i] = Sys::mk(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
}

extern "C" {

// params: mask, n_newton, n_cg, n_ls, max_blocks, ccd_max_iters, d_hat, kappa,
// radius, ee_thres, ccd_slack, ccd_thick, young, poisson, k_hinge, k_m2gt
void* s3o_create(int fp64, const double* prm) {
    Handle* h = new Handle();
    h->fp64 = fp64;
    s3o::Params P;
    P.mask = (int)prm[0]; P.n_newton = (int)prm[1]; P.n_cg = (int)prm[2]; P.n_ls = (int)prm[3];
    P.max_blocks = (int)prm[4]; P.ccd_max_iters = (int)prm[5];
    P.d_hat = prm[6]; P.kappa = prm[7]; P.radius0 = prm[8]; P.ee_thres = prm[9]; P.ccd_slack = prm[10];
    P.ccd_thick = prm[11]; P.E = prm[12]; P.nu = prm[13]; P.k_hinge = prm[14]; P.k_m2gt = prm[15];
    double mu = P.E / (2.0 * (1.0 + P.nu)), la = P.E * P.nu / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
    // the reference stores mu and lambda as float32
    mu = (double)(float)mu;
    la = (double)(float)la;
    h->d.P = h->f.P = P;
    h->d.mu = h->f.mu = mu;
    h->d.la = h->f.la = la;
    return h;
}
void s3o_destroy(void* p) { delete (Handle*)p; }

void s3o_set_target(void* p, const double* V, int nV, const int* F, int nF) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        S.gtV.resize(nV);
        for (int i = 0; i < nV; ++i) S.gtV[i] = S.mk(V[3 * i], V[3 * i + 1], V[3 * i + 2]);
        S.gtF.assign(F, F + 3 * nF);
    });
}
void s3o_set_mesh(void* p, const double* V, int nV, const int* F, int nF, const int* E, int nE, const int* Hh, int nH,
                  const double* vor, int do_nothing) {
    Handle* h = (Handle*)p;
    DISPATCH(h, S.set_mesh(V, nV, F, nF, E, nE, Hh, nH, vor, do_nothing));
}
void s3o_set_dense(void* p, const double* A, const double* b, double c) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        const size_t n3 = 3 * (size_t)S.nV;
        S.dA.assign(A, A + n3 * n3);
        S.db.assign(b, b + n3);
        S.dc = c;
    });
}
void s3o_set_q(void* p, const double* q) { Handle* h = (Handle*)p; DISPATCH(h, get_vec(S, q, S.q)); }
void s3o_get_q(void* p, double* q) { Handle* h = (Handle*)p; DISPATCH(h, put_vec(S, S.q, q)); }
void s3o_target(void* p, double* out) {
    Handle* h = (Handle*)p;
    DISPATCH(h, { S.update_target(); put_vec(S, S.target, out); });
}
// overrides the per-vertex targets (e.g. with the device's, where two
// input faces are equally close and the nearest one is a tie)
void s3o_set_targets(void* p, const double* t) { Handle* h = (Handle*)p; DISPATCH(h, get_vec(S, t, S.target)); }
long long s3o_contacts(void* p, int* out, long long cap) {
    Handle* h = (Handle*)p;
    long long n = 0;
    DISPATCH(h, {
        S.detect();
        n = (long long)S.C.size();
        for (long long b = 0; b < std::min(n, cap); ++b) {
            for (int a = 0; a < 4; ++a) out[5 * b + a] = S.C[b].i[a];
            out[5 * b + 4] = S.C[b].type1;
        }
    });
    return n;
}
void s3o_energy(void* p, double* out6) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        S.detect();
        out6[5] = S.energy_terms(S.q, S.P.mask, out6);
    });
}
void s3o_diff(void* p, double* grad, double* diag) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        S.detect();
        S.energy_terms(S.q, S.P.mask, nullptr);
        S.diff();
        put_vec(S, S.grad, grad);
        put_vec(S, S.diag, diag);
    });
}
void s3o_hvp(void* p, const double* dx, double* out) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        S.detect();
        S.energy_terms(S.q, S.P.mask, nullptr);
        S.diff();
        decltype(S.q) d, o;
        get_vec(S, dx, d);
        S.hvp(d, o);
        put_vec(S, o, out);
    });
}
void s3o_solve(void* p, double* out) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        S.detect();
        S.energy_terms(S.q, S.P.mask, nullptr);
        S.diff();
        S.cg();
        put_vec(S, S.p, out);
    });
}
void s3o_ccd(void* p, const double* pin, double* pout, double* step) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        S.qn = S.q;
        S.detect();
        S.energy_terms(S.q, S.P.mask, nullptr);
        get_vec(S, pin, S.p);
        S.clamp();
        S.ccd();
        put_vec(S, S.p, pout);
        *step = (double)S.ccd_step;
    });
}
// projected (projected = 1) or raw hinge Hessians of the last diff, (n_h, 144)
void s3o_hinge_blocks(void* p, double* out, int projected) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        const auto& B = projected ? S.hblocks : S.hraw;
        std::copy(B.begin(), B.end(), out);
    });
}
// n steps; keep_target = 1 leaves the first step's targets as set
void s3o_step(void* p, int n, int keep_target) {
    Handle* h = (Handle*)p;
    DISPATCH(h, { for (int k = 0; k < n; ++k) S.step(keep_target && k == 0); });
}
void s3o_stats(void* p, double* out4) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        out4[0] = S.energy;
        out4[1] = (double)S.stat_contacts;
        out4[2] = (double)S.stat_newton;
        out4[3] = (double)S.radius;
    });
}

// ---- single functions, for pinning against closed forms and finite differences
int s3o_pt_classify(const double* x12) {
    using V = s3o::V3<double>;
    V x[4];
    for (int a = 0; a < 4; ++a) x[a] = {x12[3 * a], x12[3 * a + 1], x12[3 * a + 2]};
    return s3o::pt_classify(x[0], x[1], x[2], x[3]);
}
int s3o_ee_classify(const double* x12, double thres) {
    using V = s3o::V3<double>;
    V x[4];
    for (int a = 0; a < 4; ++a) x[a] = {x12[3 * a], x12[3 * a + 1], x12[3 * a + 2]};
    return s3o::ee_classify(x[0], x[1], x[2], x[3], thres);
}
double s3o_pair_distance(const double* x12, int is_pt, int type) {
    using V = s3o::V3<double>;
    V x[4];
    for (int a = 0; a < 4; ++a) x[a] = {x12[3 * a], x12[3 * a + 1], x12[3 * a + 2]};
    return is_pt ? s3o::pt_pair_distance(x[0], x[1], x[2], x[3], type)
                 : s3o::ee_pair_distance(x[0], x[1], x[2], x[3], type);
}
void s3o_pair_grad(const double* x12, int is_pt, int type, double* g12) { s3o::pair_grad(x12, is_pt != 0, type, g12); }
double s3o_hinge_angle(const double* x12) {
    using V = s3o::V3<double>;
    V x[4];
    for (int a = 0; a < 4; ++a) x[a] = {x12[3 * a], x12[3 * a + 1], x12[3 * a + 2]};
    return s3o::hinge_angle_s<double>(x[0], x[1], x[2], x[3]);
}
void s3o_closest_on_tri(const double* p, const double* tri9, double* out3) {
    using V = s3o::V3<double>;
    V c = s3o::closest_on_tri(V{p[0], p[1], p[2]}, V{tri9[0], tri9[1], tri9[2]}, V{tri9[3], tri9[4], tri9[5]},
                              V{tri9[6], tri9[7], tri9[8]});
    out3[0] = c.x; out3[1] = c.y; out3[2] = c.z;
}
double s3o_barrier(double d, double dh, int order) {
    return order == 0 ? s3o::barrier(d, dh) : (order == 1 ? s3o::barrier_d1(d, dh) : s3o::barrier_d2(d, dh));
}
void s3o_spd_project_hinge(const double* H144, double* out) {
    std::copy(H144, H144 + 144, out);
    s3o::hinge_project(out);
}
// membrane gradient at x with the local basis taken from basis_x (n_v,3)
void s3o_elastic_grad_basis(void* p, const double* x, const double* basis_x, double* g) {
    Handle* h = (Handle*)p;
    DISPATCH(h, {
        decltype(S.q) X, B;
        get_vec(S, x, X);
        get_vec(S, basis_x, B);
        std::vector<s3o::V3<double>> G(S.nV, {0, 0, 0});
        S.elastic_grad_basis(X, B, G);
        for (int i = 0; i < S.nV; ++i) { g[3 * i] = G[i].x; g[3 * i + 1] = G[i].y; g[3 * i + 2] = G[i].z; }
    });
}
double s3o_energy_at(void* p, const double* x, int mask) {
    Handle* h = (Handle*)p;
    double e = 0;
    DISPATCH(h, {
        decltype(S.q) X;
        get_vec(S, x, X);
        e = S.energy_terms(X, mask, nullptr);
    });
    return e;
}

}  // extern "C"
