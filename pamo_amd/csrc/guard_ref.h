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

__device__ inline bool ref_isect_d(const float A[3][3], const float B[3][3], float r1[3], float r2[3]) {
    float side[2][3];
    double pa[4], pb[4];
    plane_of(A, pa);
#pragma unroll
    for (int v = 0; v < 3; ++v) side[1][v] = (float)plane_side_d(pa, B[v]);
    if (!side[1][0] && !side[1][1] && !side[1][2]) return false;
    if ((side[1][0] && side[1][1] && side[1][2]) && (side[1][0] < 0.0f) == (side[1][1] < 0.0f) &&
        (side[1][0] < 0.0f) == (side[1][2] < 0.0f))
        return false;
    plane_of(B, pb);
#pragma unroll
    for (int v = 0; v < 3; ++v) side[0][v] = (float)plane_side_d(pb, A[v]);
    if ((side[0][0] && side[0][1] && side[0][2]) && (side[0][0] < 0.0f) == (side[0][1] < 0.0f) &&
        (side[0][0] < 0.0f) == (side[0][2] < 0.0f))
        return false;
    double dir[3];
    dir[0] = pa[1] * pb[2] - pa[2] * pb[1];
    dir[1] = pa[2] * pb[0] - pa[0] * pb[2];
    dir[2] = pa[0] * pb[1] - pa[1] * pb[0];
    float rmin[2], rmax[2], loc[2][2][3];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const float(*T)[3] = t == 0 ? A : B;
        const float* sd = side[t];
        int o0, o1, o2;
        if ((sd[0] && sd[1]) && (sd[0] < 0.0f) == (sd[1] < 0.0f)) {
            o0 = 1; o1 = 2; o2 = 0;
        } else if ((sd[1] && sd[2]) && (sd[1] < 0.0f) == (sd[2] < 0.0f)) {
            o0 = 2; o1 = 0; o2 = 1;
        } else if (!sd[1] && (sd[2] < 0.0f) != (sd[0] < 0.0f)) {
            o0 = 1; o1 = 2; o2 = 0;
        } else {
            o0 = 0; o1 = 1; o2 = 2;
        }
        const double db = dir[0] * (double)T[o1][0] + dir[1] * (double)T[o1][1] + dir[2] * (double)T[o1][2];
        const float sc = sd[o1];
        if (sc) {
            const double da = dir[0] * (double)T[o0][0] + dir[1] * (double)T[o0][1] + dir[2] * (double)T[o0][2];
            const double dc = dir[0] * (double)T[o2][0] + dir[1] * (double)T[o2][1] + dir[2] * (double)T[o2][2];
            float f0 = sc / (sc - sd[o0]);
            float f1 = sc / (sc - sd[o2]);
            double off0 = f0 * (da - db);
            double off1 = f1 * (dc - db);
            if (off0 > off1) {
                double td = off0; off0 = off1; off1 = td;
                float tf = f0; f0 = f1; f1 = tf;
                int ti = o0; o0 = o2; o2 = ti;
            }
            rmin[t] = (float)(db + off0);
            rmax[t] = (float)(db + off1);
            const float s0 = 1.0f - f0, s1 = 1.0f - f1;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                loc[t][0][d] = s0 * T[o1][d] + f0 * T[o0][d];
                loc[t][1][d] = s1 * T[o1][d] + f1 * T[o2][d];
            }
        } else {
            rmin[t] = rmax[t] = (float)db;
#pragma unroll
            for (int d = 0; d < 3; ++d) loc[t][0][d] = loc[t][1][d] = T[o1][d];
        }
    }
    if ((rmax[0] > rmin[1]) && (rmin[0] < rmax[1])) {
        const float *a, *b;
        if (rmin[0] >= rmin[1]) {
            a = loc[0][0];
            b = (rmax[0] <= rmax[1]) ? loc[0][1] : loc[1][1];
        } else if (rmax[0] <= rmax[1]) {
            a = loc[0][1];
            b = loc[1][0];
        } else {
            a = loc[1][0];
            b = loc[1][1];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) { r1[d] = a[d]; r2[d] = b[d]; }
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
