// Synthetic mesh workloads (host C++, OpenMP).
//
// Implicit union of rotated ellipsoids and axis-aligned boxes, contoured
// with marching tetrahedra on a regular grid.  Every cube is split into the
// six Kuhn tetrahedra along its (0,0,0)-(1,1,1) diagonal, so neighbouring
// cubes agree on every face diagonal; with no grid value exactly zero the
// zero set of the piecewise-linear interpolant is a closed 2-manifold, and
// sharing one output vertex per crossed grid edge makes the mesh watertight.
// Triangles are oriented so their normal points toward increasing f
// (outward, since f < 0 inside), decided combinatorially from grid
// coordinates, so the orientation is consistent whatever the rounding of the
// float32 positions.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Shape {
    int kind;        // 0 ellipsoid, 1 box
    double c[3];     // centre
    double a[3];     // semi-axes / half sizes
    double R[9];     // rotation (row-major), ellipsoids only
};

inline double eval_shape(const Shape& s, double x, double y, double z) {
    double d[3] = {x - s.c[0], y - s.c[1], z - s.c[2]};
    if (s.kind == 1) {
        double m = -1e30;
        for (int k = 0; k < 3; ++k) m = std::max(m, std::fabs(d[k]) - s.a[k]);
        return m;
    }
    // local = R^T d, scaled by semi-axes; value ~ distance near the surface
    double l[3];
    for (int k = 0; k < 3; ++k) l[k] = (s.R[0 * 3 + k] * d[0] + s.R[1 * 3 + k] * d[1] + s.R[2 * 3 + k] * d[2]) / s.a[k];
    double r = std::sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
    double amin = std::min(s.a[0], std::min(s.a[1], s.a[2]));
    return (r - 1.0) * amin;
}

// edge types from a base grid vertex: +x +y +z +xy +yz +xz +xyz
const int ET[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};

inline int edge_type(int dx, int dy, int dz) {
    for (int t = 0; t < 7; ++t)
        if (ET[t][0] == dx && ET[t][1] == dy && ET[t][2] == dz) return t;
    return -1;
}

}  // namespace

extern "C" {

// shapes: n_shapes records of 16 doubles: kind, c[3], a[3], R[9].
// Grid: (n+1)^3 vertices on [lo, hi]^3; the surface is the level set
// {min over shapes = level}.  Grid values with |f| < min_abs are moved to
// +-min_abs (0 goes to +), which keeps every crossing point at least about
// min_abs / |grad f| away from the grid nodes, so that no two crossing points
// round to the same float32 position.  Outputs are allocated by the library;
// free with synth_free.  Returns 0 on success.
int synth_mt(const double* shapes, int n_shapes, int n, double lo, double hi, double level, double min_abs,
             float** out_v, long long* out_nv, int** out_f, long long* out_nf) {
    std::vector<Shape> S(n_shapes);
    for (int s = 0; s < n_shapes; ++s) {
        const double* r = shapes + 16 * s;
        S[s].kind = (int)r[0];
        for (int k = 0; k < 3; ++k) { S[s].c[k] = r[1 + k]; S[s].a[k] = r[4 + k]; }
        for (int k = 0; k < 9; ++k) S[s].R[k] = r[7 + k];
    }
    const long long N1 = n + 1;
    const double h = (hi - lo) / n;
    std::vector<float> f((size_t)(N1 * N1 * N1));
#pragma omp parallel for schedule(dynamic, 1)
    for (long long k = 0; k < N1; ++k) {
        double z = lo + h * k;
        for (long long j = 0; j < N1; ++j) {
            double y = lo + h * j;
            for (long long i = 0; i < N1; ++i) {
                double x = lo + h * i;
                double v = 1e30;
                for (const Shape& s : S) v = std::min(v, eval_shape(s, x, y, z));
                double w = v - level;
                if (std::fabs(w) < min_abs) w = w < 0.0 ? -min_abs : min_abs;
                float fv = (float)w;
                if (fv == 0.0f) fv = 1e-12f;  // keep every vertex off the zero set
                f[(size_t)((k * N1 + j) * N1 + i)] = fv;
            }
        }
    }
    auto F = [&](long long i, long long j, long long k) -> float { return f[(size_t)((k * N1 + j) * N1 + i)]; };
    // pass 1: crossing edges per base layer, in (j, i, type) order
    std::vector<long long> lay_cnt(N1 + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (long long k = 0; k < N1; ++k) {
        long long c = 0;
        for (long long j = 0; j < N1; ++j)
            for (long long i = 0; i < N1; ++i) {
                bool s0 = F(i, j, k) < 0;
                for (int t = 0; t < 7; ++t) {
                    long long i2 = i + ET[t][0], j2 = j + ET[t][1], k2 = k + ET[t][2];
                    if (i2 >= N1 || j2 >= N1 || k2 >= N1) continue;
                    if ((F(i2, j2, k2) < 0) != s0) ++c;
                }
            }
        lay_cnt[k] = c;
    }
    std::vector<long long> lay_off(N1 + 1, 0);
    for (long long k = 0; k < N1; ++k) lay_off[k + 1] = lay_off[k] + lay_cnt[k];
    const long long nv = lay_off[N1];
    float* V = (float*)std::malloc(sizeof(float) * 3 * (size_t)std::max(nv, 1LL));
    // dense per-layer id table builder
    auto build_layer = [&](long long k, std::vector<int>& tab, bool write_pos) {
        tab.assign((size_t)(N1 * N1 * 7), -1);
        if (k >= N1) return;
        long long id = lay_off[k];
        for (long long j = 0; j < N1; ++j)
            for (long long i = 0; i < N1; ++i) {
                float f0 = F(i, j, k);
                bool s0 = f0 < 0;
                for (int t = 0; t < 7; ++t) {
                    long long i2 = i + ET[t][0], j2 = j + ET[t][1], k2 = k + ET[t][2];
                    if (i2 >= N1 || j2 >= N1 || k2 >= N1) continue;
                    float f1 = F(i2, j2, k2);
                    if ((f1 < 0) == s0) continue;
                    tab[(size_t)((j * N1 + i) * 7 + t)] = (int)id;
                    if (write_pos) {
                        double tt = (double)f0 / ((double)f0 - (double)f1);
                        double p0[3] = {lo + h * i, lo + h * j, lo + h * k};
                        double p1[3] = {lo + h * i2, lo + h * j2, lo + h * k2};
                        for (int d = 0; d < 3; ++d) V[3 * id + d] = (float)(p0[d] + tt * (p1[d] - p0[d]));
                    }
                    ++id;
                }
            }
    };
    // pass 2: vertex positions, one vertex layer per task
#pragma omp parallel for schedule(dynamic, 1)
    for (long long k = 0; k < N1; ++k) {
        std::vector<int> tmp;
        build_layer(k, tmp, true);
    }
    // pass 3: triangles per cell layer
    std::vector<std::vector<int>> tris((size_t)n);
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    long long chunk = (n + nth - 1) / nth;
#pragma omp parallel num_threads(nth)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        long long k0 = tid * chunk, k1 = std::min<long long>(n, k0 + chunk);
        std::vector<int> tabA, tabB;
        if (k0 < k1) build_layer(k0, tabA, false);
        for (long long k = k0; k < k1; ++k) {
            build_layer(k + 1, tabB, false);
            std::vector<int>& out = tris[(size_t)k];
            auto eid = [&](const int* a, const int* b) -> int {
                // a, b: absolute grid coords (i,j,k); base = componentwise min
                int base[3] = {std::min(a[0], b[0]), std::min(a[1], b[1]), std::min(a[2], b[2])};
                int t = edge_type(std::abs(a[0] - b[0]), std::abs(a[1] - b[1]), std::abs(a[2] - b[2]));
                const std::vector<int>& tab = (base[2] == k) ? tabA : tabB;
                return tab[(size_t)(((long long)base[1] * N1 + base[0]) * 7 + t)];
            };
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
                    for (int p = 0; p < 6; ++p) {
                        int q[4][3];
                        q[0][0] = i; q[0][1] = j; q[0][2] = (int)k;
                        for (int s = 1; s < 4; ++s) {
                            std::memcpy(q[s], q[s - 1], sizeof(q[s]));
                            if (s < 4) q[s][perms[p][s - 1]] += 1;
                        }
                        float fv[4];
                        int neg = 0;
                        for (int s = 0; s < 4; ++s) {
                            fv[s] = F(q[s][0], q[s][1], q[s][2]);
                            neg += fv[s] < 0;
                        }
                        if (neg == 0 || neg == 4) continue;
                        int in[4], ni = 0, ou[4], no = 0;
                        for (int s = 0; s < 4; ++s) {
                            if (fv[s] < 0) in[ni++] = s;
                            else ou[no++] = s;
                        }
                        // surface vertices with the tet edge each lies on
                        int poly[4], pe[4][2], np = 0;
                        if (ni == 1 || no == 1) {
                            int lone = ni == 1 ? in[0] : ou[0];
                            for (int s = 0; s < 4; ++s)
                                if (s != lone) {
                                    pe[np][0] = lone;
                                    pe[np][1] = s;
                                    poly[np++] = eid(q[lone], q[s]);
                                }
                        } else {
                            // quad: a0-b0, a0-b1, a1-b1, a1-b0 (cyclic)
                            const int ends[4][2] = {{in[0], ou[0]}, {in[0], ou[1]}, {in[1], ou[1]}, {in[1], ou[0]}};
                            for (int e = 0; e < 4; ++e) {
                                pe[e][0] = ends[e][0];
                                pe[e][1] = ends[e][1];
                                poly[e] = eid(q[ends[e][0]], q[ends[e][1]]);
                            }
                            np = 4;
                        }
                        // Orientation toward a positive tet corner, decided exactly.  A
                        // surface triangle's corners sit on three tet edges; for any
                        // interpolation parameters in (0,1) the triangle neither
                        // degenerates nor passes through a tet corner, so the sign
                        // taken at the edge midpoints (integer grid coordinates,
                        // doubled) is the sign at the actual crossing points.
                        const int pos = ou[0];
                        auto emit = [&](int a, int b, int c) {
                            long long m[3][3];
                            const int sel[3] = {a, b, c};
                            for (int r = 0; r < 3; ++r)
                                for (int d = 0; d < 3; ++d) m[r][d] = (long long)q[pe[sel[r]][0]][d] + q[pe[sel[r]][1]][d];
                            long long u[3], w[3], z[3];
                            for (int d = 0; d < 3; ++d) {
                                u[d] = m[1][d] - m[0][d];
                                w[d] = m[2][d] - m[0][d];
                                z[d] = 2LL * q[pos][d] - m[0][d];
                            }
                            const long long det = z[0] * (u[1] * w[2] - u[2] * w[1]) + z[1] * (u[2] * w[0] - u[0] * w[2]) +
                                                  z[2] * (u[0] * w[1] - u[1] * w[0]);
                            if (det >= 0) { out.push_back(poly[a]); out.push_back(poly[b]); out.push_back(poly[c]); }
                            else { out.push_back(poly[a]); out.push_back(poly[c]); out.push_back(poly[b]); }
                        };
                        if (np == 3) emit(0, 1, 2);
                        else { emit(0, 1, 2); emit(0, 2, 3); }
                    }
                }
            tabA.swap(tabB);
        }
    }
    long long nf = 0;
    for (auto& t : tris) nf += (long long)t.size() / 3;
    int* Fo = (int*)std::malloc(sizeof(int) * 3 * (size_t)std::max(nf, 1LL));
    long long o = 0;
    for (auto& t : tris) {
        std::memcpy(Fo + o, t.data(), sizeof(int) * t.size());
        o += (long long)t.size();
    }
    *out_v = V;
    *out_nv = nv;
    *out_f = Fo;
    *out_nf = nf;
    return 0;
}

void synth_free(void* p) { std::free(p); }

}  // extern "C"
