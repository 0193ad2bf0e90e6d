// Stage-3 element math (FP32 state, as the reference's Warp kernels), shared
// by the kernels of stage3.hip.  Distances and their classification follow
// simp_cuda/safe_project/src/pamo_safe_project/kernels/distance_kernels/
// distance_kernels.py:8-465; gradients are the closed forms of the
// unsigned distances (derived here, not the generated code of grad_funcs.py;
// both are the gradient of the same function).
#pragma once
#include <hip/hip_runtime.h>

namespace pamo {
namespace s3 {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 v3(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ float4 f4(V3 a) { return make_float4(a.x, a.y, a.z, 0.0f); }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ float len_sq(V3 a) { return dot(a, a); }
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
// wp.normalize: zero vector stays zero
__device__ __forceinline__ V3 normalize(V3 a) {
    float l = len(a);
    return l > 0.0f ? a / l : mk(0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }

// PTContactTypes / EEContactTypes (defs.py)
enum { PT = 0, PP01 = 1, PP02 = 2, PE012 = 3, PP03 = 4, PE031 = 5, PE023 = 6 };
enum { EPP02 = 0, EPP03 = 1, EPP12 = 2, EPP13 = 3, EPE023 = 4, EPE123 = 5, EPE201 = 6, EPE301 = 7, EEE = 8 };
enum { BLOCK_PT = 3, BLOCK_EE = 4 };

__device__ __forceinline__ float pp_dist(V3 a, V3 b) { return len(b - a); }
__device__ __forceinline__ float pe_dist(V3 p, V3 a, V3 b) { return len(cross(a - p, b - p)) / len(a - b); }
__device__ __forceinline__ float pt_dist(V3 p, V3 a, V3 b, V3 c) { return fabsf(dot(p - a, normalize(cross(b - a, c - a)))); }
__device__ __forceinline__ float ee_dist(V3 a, V3 b, V3 c, V3 d) { return fabsf(dot(a - c, normalize(cross(b - a, d - c)))); }

// pt_pair_classify (distance_kernels.py:75-226)
__device__ inline int pt_classify(V3 x0, V3 x1, V3 x2, V3 x3) {
    V3 x10 = x0 - x1, x12 = x2 - x1, x13 = x3 - x1;
    float area_sq = len_sq(cross(x12, x13));
    if (area_sq <= 1e-6f * (len_sq(x12) * len_sq(x13)) || area_sq <= 1e-16f) {
        float d1 = len(x10), d2 = len(x0 - x2), d3 = len(x0 - x3);
        if (d1 < d2) return d1 < d3 ? PP01 : PP03;
        return d2 < d3 ? PP02 : PP03;
    }
    V3 n = normalize(cross(x12, x13));
    V3 xp = x0 - dot(n, x10) * n;
    x10 = xp - x1;
    V3 x20 = xp - x2, x30 = xp - x3;
    float a00 = dot(x12, x12), a01 = dot(x12, x13), a11 = dot(x13, x13);
    float b0 = dot(x10, x12), b1 = dot(x10, x13);
    float det = a00 * a11 - a01 * a01;
    float beta = b0 * a11 - b1 * a01;
    float gamma = a00 * b1 - b0 * a01;
    float alpha = det - beta - gamma;
    // wp.step(x) = 1 if x < 0 else 0
    float nneg = (alpha < 0.0f ? 1.0f : 0.0f) + (beta < 0.0f ? 1.0f : 0.0f) + (gamma < 0.0f ? 1.0f : 0.0f);
    if (nneg == 0.0f) return PT;
    if (nneg == 1.0f) {
        if (alpha < 0.0f) {
            V3 x23 = x3 - x2, x20o = x0 - x2;
            float dr = dot(x23, x20o);
            if (dr < 0.0f) return PP02;
            if (dr > dot(x23, x23)) return PP03;
            return PE023;
        } else if (beta < 0.0f) {
            V3 x31 = x1 - x3, x30o = x0 - x3;
            float dr = dot(x31, x30o);
            if (dr < 0.0f) return PP03;
            if (dr > dot(x31, x31)) return PP01;
            return PE031;
        } else if (gamma < 0.0f) {
            float dr = dot(x12, x10);
            if (dr < 0.0f) return PP01;
            if (dr > dot(x12, x12)) return PP02;
            return PE012;
        }
        return -1;
    }
    if (alpha >= 0.0f) {
        float d12 = dot(x12, x10), d13 = dot(x13, x10);
        if (d12 > 0.0f && d13 <= 0.0f) return dot(x12, x20) > 0.0f ? PP02 : PE012;
        if (d13 > 0.0f && d12 <= 0.0f) return dot(x13, x30) > 0.0f ? PP03 : PE031;
        return PP01;
    } else if (beta >= 0.0f) {
        V3 x23 = x3 - x2, x21 = x1 - x2;
        float d23 = dot(x23, x20), d21 = dot(x21, x20);
        if (d23 > 0.0f && d21 <= 0.0f) return dot(x23, x30) > 0.0f ? PP03 : PE023;
        if (d21 > 0.0f && d23 <= 0.0f) return dot(x21, x10) > 0.0f ? PP01 : PE012;
        return PP02;
    } else {
        V3 x31 = x1 - x3, x32 = x2 - x3;
        float d31 = dot(x31, x30), d32 = dot(x32, x30);
        if (d31 > 0.0f && d32 <= 0.0f) return dot(x31, x10) > 0.0f ? PP01 : PE031;
        if (d32 > 0.0f && d31 <= 0.0f) return dot(x32, x20) > 0.0f ? PP02 : PE023;
        return PP03;
    }
}

__device__ inline float pt_distance(V3 x0, V3 x1, V3 x2, V3 x3, int t) {
    switch (t) {
        case PT: return pt_dist(x0, x1, x2, x3);
        case PP01: return pp_dist(x0, x1);
        case PP02: return pp_dist(x0, x2);
        case PE012: return pe_dist(x0, x1, x2);
        case PP03: return pp_dist(x0, x3);
        case PE031: return pe_dist(x0, x3, x1);
        case PE023: return pe_dist(x0, x2, x3);
    }
    return 1e9f;
}

// ee_pair_classify (distance_kernels.py:325-441)
__device__ inline int ee_classify(V3 x0, V3 x1, V3 x2, V3 x3, float sin_thres) {
    V3 e0 = x0 - x1, e1 = x2 - x3;
    float e0_sq = dot(e0, e0), e1_sq = dot(e1, e1);
    float a00 = e0_sq, a01 = -dot(e0, e1), a10 = dot(e0, e1), a11 = -e1_sq;
    float b0 = dot(x3 - x1, e0), b1 = dot(x3 - x1, e1);
    float mu0 = -(b0 * a11 - b1 * a01);
    float mu1 = -(a00 * b1 - a10 * b0);
    float detA = -(a00 * a11 - a01 * a10);
    float edge_sin = len(cross(normalize(e0), normalize(e1)));
    if (mu0 >= 0.0f && mu0 <= detA && mu1 >= 0.0f && mu1 <= detA && edge_sin >= sin_thres && e0_sq > 1e-15f &&
        e1_sq > 1e-15f)
        return EEE;
    float best = 1e9f;
    int type = -1;
    float d;
    float t0 = dot(x0 - x3, e1);
    if (t0 >= e1_sq) { d = pp_dist(x0, x2); if (d < best) { best = d; type = EPP02; } }
    else if (t0 <= 0.0f) { d = pp_dist(x0, x3); if (d < best) { best = d; type = EPP03; } }
    else { d = pe_dist(x0, x2, x3); if (d < best) { best = d; type = EPE023; } }
    float t1 = dot(x1 - x3, e1);
    if (t1 >= e1_sq) { d = pp_dist(x1, x2); if (d < best) { best = d; type = EPP12; } }
    else if (t1 <= 0.0f) { d = pp_dist(x1, x3); if (d < best) { best = d; type = EPP13; } }
    else { d = pe_dist(x1, x2, x3); if (d < best) { best = d; type = EPE123; } }
    float t2 = dot(x2 - x1, e0);
    if (t2 >= e0_sq) { d = pp_dist(x2, x0); if (d < best) { best = d; type = EPP02; } }
    else if (t2 <= 0.0f) { d = pp_dist(x2, x1); if (d < best) { best = d; type = EPP12; } }
    else { d = pe_dist(x2, x0, x1); if (d < best) { best = d; type = EPE201; } }
    float t3 = dot(x3 - x1, e0);
    if (t3 >= e0_sq) { d = pp_dist(x3, x0); if (d < best) { best = d; type = EPP03; } }
    else if (t3 <= 0.0f) { d = pp_dist(x3, x1); if (d < best) { best = d; type = EPP13; } }
    else { d = pe_dist(x3, x0, x1); if (d < best) { best = d; type = EPE301; } }
    return type;
}

__device__ inline float ee_distance(V3 x0, V3 x1, V3 x2, V3 x3, int t) {
    switch (t) {
        case EPP02: return pp_dist(x0, x2);
        case EPP03: return pp_dist(x0, x3);
        case EPP12: return pp_dist(x1, x2);
        case EPP13: return pp_dist(x1, x3);
        case EPE023: return pe_dist(x0, x2, x3);
        case EPE123: return pe_dist(x1, x2, x3);
        case EPE201: return pe_dist(x2, x0, x1);
        case EPE301: return pe_dist(x3, x0, x1);
        case EEE: return ee_dist(x0, x1, x2, x3);
    }
    return 1e9f;
}

// ------------------------------------------------------------ gradients of unsigned distances
// d = |x0 - x1|
__device__ __forceinline__ void pp_grad(V3 a, V3 b, V3& ga, V3& gb) {
    ga = normalize(a - b);
    gb = normalize(b - a);
}

// d = |(a - p) x (b - p)| / |a - b|: with u = a - p, w = b - p, c = u x w,
// dA = c^ . dc, dc = du x w + u x dw; L = |b - a|.
__device__ inline void pe_grad(V3 p, V3 a, V3 b, V3& gp, V3& ga, V3& gb) {
    V3 u = a - p, w = b - p;
    V3 c = cross(u, w);
    float A = len(c);
    V3 ch = A > 0.0f ? c / A : mk(0.0f, 0.0f, 0.0f);
    V3 e = b - a;
    float L = len(e);
    V3 eh = e / L;
    V3 dAa = cross(w, ch), dAb = cross(ch, u);
    V3 dAp = -(dAa + dAb);
    float iL = 1.0f / L, r = A / (L * L);
    gp = iL * dAp;
    ga = iL * dAa + r * eh;
    gb = iL * dAb - r * eh;
}

// d = |s|, s = (p - a) . n^, n = (b - a) x (c - a); u = (w - s n^) / |n|.
__device__ inline void pt_grad(V3 p, V3 a, V3 b, V3 c, V3& gp, V3& ga, V3& gb, V3& gc) {
    V3 e1 = b - a, e2 = c - a;
    V3 n = cross(e1, e2);
    float nl = len(n);
    V3 nh = n / nl;
    V3 w = p - a;
    float s = dot(w, nh);
    V3 u = (w - s * nh) / nl;
    V3 db = cross(e2, u), dc = cross(u, e1);
    float sg = s < 0.0f ? -1.0f : 1.0f;
    gp = sg * nh;
    gb = sg * db;
    gc = sg * dc;
    ga = sg * (-nh - db - dc);
}

// d = |s|, s = (x0 - x2) . n^, n = (x1 - x0) x (x3 - x2).
__device__ inline void ee_grad(V3 x0, V3 x1, V3 x2, V3 x3, V3& g0, V3& g1, V3& g2, V3& g3) {
    V3 e0 = x1 - x0, e1 = x3 - x2;
    V3 n = cross(e0, e1);
    float nl = len(n);
    V3 nh = n / nl;
    V3 w = x0 - x2;
    float s = dot(w, nh);
    V3 u = (w - s * nh) / nl;
    V3 a = cross(e1, u), b = cross(u, e0);
    float sg = s < 0.0f ? -1.0f : 1.0f;
    g0 = sg * (nh - a);
    g1 = sg * a;
    g2 = sg * (-nh - b);
    g3 = sg * b;
}

// gradient of the classified distance into g[4]
__device__ inline void pt_pair_grad(V3 x0, V3 x1, V3 x2, V3 x3, int t, V3 g[4]) {
    V3 z = mk(0.0f, 0.0f, 0.0f);
    g[0] = g[1] = g[2] = g[3] = z;
    switch (t) {
        case PT: pt_grad(x0, x1, x2, x3, g[0], g[1], g[2], g[3]); break;
        case PP01: pp_grad(x0, x1, g[0], g[1]); break;
        case PP02: pp_grad(x0, x2, g[0], g[2]); break;
        case PE012: pe_grad(x0, x1, x2, g[0], g[1], g[2]); break;
        case PP03: pp_grad(x0, x3, g[0], g[3]); break;
        case PE031: pe_grad(x0, x3, x1, g[0], g[3], g[1]); break;
        case PE023: pe_grad(x0, x2, x3, g[0], g[2], g[3]); break;
    }
}

__device__ inline void ee_pair_grad(V3 x0, V3 x1, V3 x2, V3 x3, int t, V3 g[4]) {
    V3 z = mk(0.0f, 0.0f, 0.0f);
    g[0] = g[1] = g[2] = g[3] = z;
    switch (t) {
        case EPP02: pp_grad(x0, x2, g[0], g[2]); break;
        case EPP03: pp_grad(x0, x3, g[0], g[3]); break;
        case EPP12: pp_grad(x1, x2, g[1], g[2]); break;
        case EPP13: pp_grad(x1, x3, g[1], g[3]); break;
        case EPE023: pe_grad(x0, x2, x3, g[0], g[2], g[3]); break;
        case EPE123: pe_grad(x1, x2, x3, g[1], g[2], g[3]); break;
        case EPE201: pe_grad(x2, x0, x1, g[2], g[0], g[1]); break;
        case EPE301: pe_grad(x3, x0, x1, g[3], g[0], g[1]); break;
        case EEE: ee_grad(x0, x1, x2, x3, g[0], g[1], g[2], g[3]); break;
    }
}

// IPC log barrier b(d) = -(d - dh)^2 ln(d / dh) for d < dh (collision_energy.py:8-28)
__device__ __forceinline__ float barrier(float d, float dh) {
    return d < dh ? -(d - dh) * (d - dh) * logf(d / dh) : 0.0f;
}
__device__ __forceinline__ float barrier_d1(float d, float dh) {
    float t = d - dh;
    return d < dh ? -2.0f * t * logf(d / dh) - t * t / d : 0.0f;
}
__device__ __forceinline__ float barrier_d2(float d, float dh) {
    return d < dh ? -2.0f * logf(d / dh) - 2.0f * (d - dh) / d - 1.0f + (dh * dh) / (d * d) : 0.0f;
}

// compute_hinge_angle (hinge_energy.py:4-18)
__device__ __forceinline__ float hinge_angle(V3 x0, V3 x1, V3 x2, V3 x3) {
    V3 e12 = normalize(x2 - x1);
    V3 n1 = normalize(cross(e12, x0 - x1));
    V3 n2 = normalize(cross(x3 - x1, e12));
    float c = dot(n1, n2);
    float s = dot(cross(n1, n2), e12);
    return atan2f(s, c);
}

// closest point on a triangle (Ericson, Real-Time Collision Detection 5.1.5)
__device__ inline V3 closest_on_tri(V3 p, V3 a, V3 b, V3 c) {
    V3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    V3 bp = p - b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + (d1 / (d1 - d3)) * ab;
    V3 cp = p - c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + (d2 / (d2 - d6)) * ac;
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    float den = 1.0f / (va + vb + vc);
    float v = vb * den, w = vc * den;
    return a + v * ab + w * ac;
}

}  // namespace s3
}  // namespace pamo
