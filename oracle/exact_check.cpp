// Exhaustive exact self-intersection check of a triangle mesh.
//
// Test infrastructure only (loaded by tests/ through oracle/exact.py): it
// checks outputs, it is never part of the product path.
//
// Predicates are exact for double inputs (float32 positions convert to
// double exactly):
//   * orient3d / orient2d evaluated with a floating-point filter and, when the
//     filter cannot decide, exactly with floating-point expansions (error-free
//     transformations two-sum / two-product and expansion growth and scaling,
//     after J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and
//     Fast Robust Geometric Predicates", Discrete Comput. Geom. 18, 1997);
//   * closed segment-triangle and triangle-triangle intersection built only
//     from those signs (coplanar cases in a projection where the triangle is
//     non-degenerate).
// Mesh semantics: two faces with no common vertex index must not meet at
// all; faces sharing one vertex may meet only at that vertex; faces sharing
// an edge may meet only along it (a flat fold, the two opposite vertices on
// the same side of the edge in a common plane, is an intersection); repeated
// faces are intersections.  Degenerate faces (collinear corners) are counted
// separately and excluded from the pair tests.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "exact_pred.h"


extern "C" {

int ex_orient3d(const double* a, const double* b, const double* c, const double* d) {
    using exact::P3;
    return exact::orient3d(P3{a[0], a[1], a[2]}, P3{b[0], b[1], b[2]}, P3{c[0], c[1], c[2]}, P3{d[0], d[1], d[2]});
}

int ex_orient2d(const double* a, const double* b, const double* c) {
    using exact::P2;
    return exact::orient2d(P2{a[0], a[1]}, P2{b[0], b[1]}, P2{c[0], c[1]});
}

// Two faces given by 9 coordinates and 3 vertex ids each.
int ex_faces_intersect(const double* a9, const int* ia, const double* b9, const int* ib) {
    exact::P3 pa[3], pb[3];
    for (int i = 0; i < 3; ++i) {
        pa[i] = exact::P3{a9[3 * i], a9[3 * i + 1], a9[3 * i + 2]};
        pb[i] = exact::P3{b9[3 * i], b9[3 * i + 1], b9[3 * i + 2]};
    }
    return exact::faces_intersect(ia, pa, ib, pb) ? 1 : 0;
}

// All intersecting face pairs (a < b) of the mesh: closed-box sort-and-sweep
// broad phase, exact narrow phase.  Faces with index -1 are skipped.  Writes
// up to cap pairs, the number of box-overlap candidates to *n_candidates and
// the number of degenerate faces to *n_degenerate; returns the number of
// intersecting pairs.
long long ex_self_intersections(const float* pts, int nP, const int* tris, int nT, int* pairs, long long cap,
                                long long* n_candidates, long long* n_degenerate) {
    using exact::P3;
    std::vector<int> live;
    std::vector<char> degen(nT, 0);
    long long ndeg = 0;
    for (int f = 0; f < nT; ++f) {
        const int* t = tris + 3 * f;
        if (t[0] < 0 || t[1] < 0 || t[2] < 0 || t[0] >= nP || t[1] >= nP || t[2] >= nP) continue;
        const P3 a{pts[3 * t[0]], pts[3 * t[0] + 1], pts[3 * t[0] + 2]};
        const P3 b{pts[3 * t[1]], pts[3 * t[1] + 1], pts[3 * t[1] + 2]};
        const P3 c{pts[3 * t[2]], pts[3 * t[2] + 1], pts[3 * t[2] + 2]};
        if (exact::proj_axis(a, b, c) < 0) {
            degen[f] = 1;
            ++ndeg;
            continue;
        }
        live.push_back(f);
    }
    const int n = (int)live.size();
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const int* t = tris + 3 * live[k];
        for (int d = 0; d < 3; ++d) {
            float l = pts[3 * t[0] + d], h = l;
            for (int c = 1; c < 3; ++c) {
                l = std::min(l, pts[3 * t[c] + d]);
                h = std::max(h, pts[3 * t[c] + d]);
            }
            lo[3 * (size_t)k + d] = l;
            hi[3 * (size_t)k + d] = h;
        }
    }
    std::vector<int> ord(n);
    for (int k = 0; k < n; ++k) ord[k] = k;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
        return lo[3 * (size_t)a] < lo[3 * (size_t)b] || (lo[3 * (size_t)a] == lo[3 * (size_t)b] && a < b);
    });
    long long total = 0, cand = 0;
    std::vector<std::vector<int>> found(omp_get_max_threads());
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total, cand)
    for (int r = 0; r < n; ++r) {
        const int k = ord[r];
        auto& out = found[omp_get_thread_num()];
        for (int s = r + 1; s < n && lo[3 * (size_t)ord[s]] <= hi[3 * (size_t)k]; ++s) {
            const int m = ord[s];
            if (hi[3 * (size_t)m + 1] < lo[3 * (size_t)k + 1] || hi[3 * (size_t)k + 1] < lo[3 * (size_t)m + 1] ||
                hi[3 * (size_t)m + 2] < lo[3 * (size_t)k + 2] || hi[3 * (size_t)k + 2] < lo[3 * (size_t)m + 2])
                continue;
            ++cand;
            const int fa = live[k], fb = live[m];
            P3 pa[3], pb[3];
            for (int i = 0; i < 3; ++i) {
                const int va = tris[3 * fa + i], vb = tris[3 * fb + i];
                pa[i] = P3{pts[3 * va], pts[3 * va + 1], pts[3 * va + 2]};
                pb[i] = P3{pts[3 * vb], pts[3 * vb + 1], pts[3 * vb + 2]};
            }
            if (exact::faces_intersect(tris + 3 * fa, pa, tris + 3 * fb, pb)) {
                ++total;
                out.push_back(std::min(fa, fb));
                out.push_back(std::max(fa, fb));
            }
        }
    }
    std::vector<std::pair<int, int>> all;
    for (auto& v : found)
        for (size_t i = 0; i < v.size(); i += 2) all.emplace_back(v[i], v[i + 1]);
    std::sort(all.begin(), all.end());
    for (long long i = 0; i < (long long)all.size() && i < cap; ++i) {
        pairs[2 * i] = all[i].first;
        pairs[2 * i + 1] = all[i].second;
    }
    if (n_candidates) *n_candidates = cand;
    if (n_degenerate) *n_degenerate = ndeg;
    return total;
}

}  // extern "C"
