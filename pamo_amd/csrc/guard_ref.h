// Triangle/triangle guard with the reference's decision rules, as device
// code.  Follows simp_cuda/src/bvh/self_intersect.cuh:129-149,229-378,
// tri_tri_3d.cuh:99-276 and math.cuh:43-52: shift by the floored mean,
// coplanar skip (|side| <= max|coord| * 1e-6), skip on two shared indices or
// two coordinate-coincident vertex pairs, FP64 planes + FP32 sides
// segment-overlap test, and the 1e-3 shared-vertex tolerance.  Arithmetic
// order is kept so results match the reference bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace pamo {

struct Tri9 {
    float v[3][3];
};

__device__ __forceinline__ double plane_side_d(const double p[4], const float* x) {
    return (p[0] * (double)x[0] + p[1] * (double)x[1] + p[2] * (double)x[2]) + p[3];
}

__device__ __forceinline__ void plane_of(const float T[3][3], double p[4]) {
    double ba0 = (double)T[0][0] - (double)T[1][0];
    double ba1 = (double)T[0][1] - (double)T[1][1];
    double ba2 = (double)T[0][2] - (double)T[1][2];
    double bc0 = (double)T[2][0] - (double)T[1][0];
    double bc1 = (double)T[2][1] - (double)T[1][1];
    double bc2 = (double)T[2][2] - (double)T[1][2];
    p[0] = ba1 * bc2 - ba2 * bc1;
    p[1] = ba2 * bc0 - ba0 * bc2;
    p[2] = ba0 * bc1 - ba1 * bc0;
    p[3] = -(p[0] * (double)T[1][0] + p[1] * (double)T[1][1] + p[2] * (double)T[1][2]);
}

__device__ inline bool ref_coplanar_d(const float A[3][3], const float B[3][3]) {
    float sc = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) sc = fmaxf(sc, fabsf(A[i][j]));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) sc = fmaxf(sc, fabsf(B[i][j]));
    double p[4];
    plane_of(A, p);
    const double eps = (double)sc * 1e-6;
#pragma unroll
    for (int v = 0; v < 3; ++v)
        if (!(fabs(plane_side_d(p, B[v])) <= eps)) return false;
    return true;
}

struct V3f {
    float x, y, z;
};

// One triangle's contribution to the overlap test along the plane-plane
// line: the lone vertex (on its own side of the other plane) is V1, the two
// others V0 / V2.  Register-only: the three vertex rotations are selected
// with conditional moves, never by indexing.
__device__ __forceinline__ void isect_range(V3f a, V3f b, V3f c, float s0, float s1, float s2, const double dir[3],
                                            float& rmin, float& rmax, V3f& l0, V3f& l1) {
    int rot;  // (o0, o1, o2) = (rot, rot + 1, rot + 2) mod 3
    if ((s0 && s1) && (s0 < 0.0f) == (s1 < 0.0f)) rot = 1;
    else if ((s1 && s2) && (s1 < 0.0f) == (s2 < 0.0f)) rot = 2;
    else if (!s1 && (s2 < 0.0f) != (s0 < 0.0f)) rot = 1;
    else rot = 0;
    V3f P0 = rot == 0 ? a : (rot == 1 ? b : c);
    V3f P1 = rot == 0 ? b : (rot == 1 ? c : a);
    V3f P2 = rot == 0 ? c : (rot == 1 ? a : b);
    float q0 = rot == 0 ? s0 : (rot == 1 ? s1 : s2);
    float q1 = rot == 0 ? s1 : (rot == 1 ? s2 : s0);
    float q2 = rot == 0 ? s2 : (rot == 1 ? s0 : s1);
    const double db = dir[0] * (double)P1.x + dir[1] * (double)P1.y + dir[2] * (double)P1.z;
    if (q1) {
        const double da = dir[0] * (double)P0.x + dir[1] * (double)P0.y + dir[2] * (double)P0.z;
        const double dc = dir[0] * (double)P2.x + dir[1] * (double)P2.y + dir[2] * (double)P2.z;
        float f0 = q1 / (q1 - q0);
        float f1 = q1 / (q1 - q2);
        double off0 = f0 * (da - db);
        double off1 = f1 * (dc - db);
        if (off0 > off1) {
            double td = off0; off0 = off1; off1 = td;
            float tf = f0; f0 = f1; f1 = tf;
            V3f tv = P0; P0 = P2; P2 = tv;
        }
        rmin = (float)(db + off0);
        rmax = (float)(db + off1);
        const float g0 = 1.0f - f0, g1 = 1.0f - f1;
        l0 = {g0 * P1.x + f0 * P0.x, g0 * P1.y + f0 * P0.y, g0 * P1.z + f0 * P0.z};
        l1 = {g1 * P1.x + f1 * P2.x, g1 * P1.y + f1 * P2.y, g1 * P1.z + f1 * P2.z};
    } else {
        rmin = rmax = (float)db;
        l0 = P1;
        l1 = P1;
    }
}

__device__ inline bool ref_isect_d(const float A[3][3], const float B[3][3], float r1[3], float r2[3]) {
    double pa[4], pb[4];
    plane_of(A, pa);
    const float b0 = (float)plane_side_d(pa, B[0]), b1 = (float)plane_side_d(pa, B[1]),
                b2 = (float)plane_side_d(pa, B[2]);
    if (!b0 && !b1 && !b2) return false;
    if ((b0 && b1 && b2) && (b0 < 0.0f) == (b1 < 0.0f) && (b0 < 0.0f) == (b2 < 0.0f)) return false;
    plane_of(B, pb);
    const float a0 = (float)plane_side_d(pb, A[0]), a1 = (float)plane_side_d(pb, A[1]),
                a2 = (float)plane_side_d(pb, A[2]);
    if ((a0 && a1 && a2) && (a0 < 0.0f) == (a1 < 0.0f) && (a0 < 0.0f) == (a2 < 0.0f)) return false;
    double dir[3];
    dir[0] = pa[1] * pb[2] - pa[2] * pb[1];
    dir[1] = pa[2] * pb[0] - pa[0] * pb[2];
    dir[2] = pa[0] * pb[1] - pa[1] * pb[0];
    float min0, max0, min1, max1;
    V3f l00, l01, l10, l11;
    isect_range({A[0][0], A[0][1], A[0][2]}, {A[1][0], A[1][1], A[1][2]}, {A[2][0], A[2][1], A[2][2]}, a0, a1, a2,
                dir, min0, max0, l00, l01);
    isect_range({B[0][0], B[0][1], B[0][2]}, {B[1][0], B[1][1], B[1][2]}, {B[2][0], B[2][1], B[2][2]}, b0, b1, b2,
                dir, min1, max1, l10, l11);
    if ((max0 > min1) && (min0 < max1)) {
        V3f x, y;
        if (min0 >= min1) {
            x = l00;
            y = (max0 <= max1) ? l01 : l11;
        } else if (max0 <= max1) {
            x = l01;
            y = l10;
        } else {
            x = l10;
            y = l11;
        }
        r1[0] = x.x; r1[1] = x.y; r1[2] = x.z;
        r2[0] = y.x; r2[1] = y.y; r2[2] = y.z;
        return true;
    }
    return false;
}

// One ordered pair: cur (the candidate found by traversal) and qry (the
// face that issued the query).
__device__ inline bool ref_guard_d(const float cur[3][3], const int ci[3], const float qry[3][3], const int qi[3],
                                   float epsilon) {
    float A[3][3], B[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float m = (cur[0][d] + cur[1][d] + cur[2][d] + qry[0][d] + qry[1][d] + qry[2][d]) / 6.0f;
        float s = (float)(int)floorf(m);
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            A[v][d] = cur[v][d] - s;
            B[v][d] = qry[v][d] - s;
        }
    }
    int shared = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) shared += (ci[a] == qi[b]) ? 1 : 0;
    if (ref_coplanar_d(A, B)) return false;
    if (shared == 2) return false;
    int same = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bool any = false;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            any |= (fabsf(A[a][0] - B[b][0]) < 1e-6f) && (fabsf(A[a][1] - B[b][1]) < 1e-6f) &&
                   (fabsf(A[a][2] - B[b][2]) < 1e-6f);
        same += any ? 1 : 0;
    }
    if (same >= 2) return false;
    float r1[3], r2[3];
    if (!ref_isect_d(A, B, r1, r2)) return false;
    const float dx = fabsf(r2[0] - r1[0]), dy = fabsf(r2[1] - r1[1]), dz = fabsf(r2[2] - r1[2]);
    float big = -__builtin_huge_valf();
    if (dx > big) big = dx;
    if (dy > big) big = dy;
    if (dz > big) big = dz;
    if (big < epsilon && shared == 1) return false;
    return true;
}

}  // namespace pamo
