// Stage 3 of PaMO on gfx950: barrier-energy safe projection of the
// simplified mesh back onto the input surface, behind the C ABI of
// include/pamo_hip.h (pamo_s3_*).
//
// Reference behaviour followed (simp_cuda/safe_project/src/pamo_safe_project/):
//   process / normalisation ... processing.py:15-108
//   register_mesh ............. system.py:167-272
//   step (10 Newton) .......... system.py:411-470
//   closest point on GT ....... kernels/geometry_kernels.py:82-94
//   contact detection ......... energy.py:896-948, energy_kernels/contact_detection.py:119-216
//   energies .................. energy_kernels/{distance,elastic,hinge,collision}_energy.py
//   SPD projection ............ kernels/utils_kernels.py:21-39
//   Jacobi PCG ................ cg_solver.py:96-158, kernels/solver_kernels.py:4-58
//   clamp / ACCD / line search  solver_kernels.py:61-94, ccd_kernels.py:8-110, system.py:354-409
//
// State is FP32 (as the reference); the total energy and every CG dot
// product are reduced in FP64 in a fixed two-level order (per-block partial
// sums, then one block).  Contacts are listed in an order fixed by the
// trees and positions, and the gradient, Jacobi diagonal and Hessian-vector
// products are summed per vertex over per-element slots in a fixed order
// (the reference scatters them with FP32 atomics), so a step's result does
// not depend on scheduling.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pamo_hip.h"
#include "prim.h"
#include "stage2_kernels.h"
#include "stage3_math.h"
#include "lbvh_tree.h"

namespace pamo {
namespace s3 {

__device__ __forceinline__ float box_dist2(V3 p, float4 lo, float4 hi) {
    float dx = fmaxf(fmaxf(lo.x - p.x, 0.0f), p.x - hi.x);
    float dy = fmaxf(fmaxf(lo.y - p.y, 0.0f), p.y - hi.y);
    float dz = fmaxf(fmaxf(lo.z - p.z, 0.0f), p.z - hi.z);
    return dx * dx + dy * dy + dz * dz;
}

// ------------------------------------------------------------ parameters / context
struct Params {
    int n_newton, n_cg, n_ls, max_blocks, ccd_max_iters, mask;
    float d_hat, kappa, radius0, ee_thres, ccd_slack, ccd_thick, mu, la, k_hinge, k_m2gt;
};

struct Ctx {
    hipStream_t stream = nullptr;
    int device = 0;
    Params P{};
    int nV = 0, nF = 0, nE = 0, nH = 0, nGV = 0, nGF = 0;
    int capV = 0, capF = 0, capE = 0, capH = 0, capG = 0;
    bool do_nothing = false;
    double min_angle_thres = 0.0;  // degrees (Stage3Config.min_angle_thres)
    float radius = 0.0f;
    // vertices (float4)
    float4 *q = nullptr, *qn = nullptr, *qd = nullptr, *qrest = nullptr, *p = nullptr, *target = nullptr;
    float4 *grad = nullptr, *diag = nullptr, *r = nullptr, *v = nullptr, *Av = nullptr, *z = nullptr;
    float* vor = nullptr;
    // topology
    int4* tri = nullptr;
    int2* edge = nullptr;
    int4* hinge = nullptr;
    float *area = nullptr, *rest_angle = nullptr, *rest_elen = nullptr;
    float4* inv_dm = nullptr;
    float* hblock = nullptr;  // 144 per hinge (4x4 blocks of 3x3, row-major)
    float* hsym = nullptr;    // 45 per hinge: projected 9x9, packed upper triangle
    // incidence lists (sorted by element) and per-slot HVP outputs
    int *vf_off = nullptr, *vf_ent = nullptr, *vh_off = nullptr, *vh_ent = nullptr;
    int *vc_off = nullptr, *vc_ent = nullptr;
    // inverse incidence: position in the vertex lists of face slot 3 f + k /
    // hinge slot 4 h + a, so element contributions land in list order
    int *vf_dst = nullptr, *vh_dst = nullptr;
    // gradient-diagonal terms per hinge corner and per active-contact corner
    // (vertex-list order) for the fixed-order per-vertex sums of compute_diff
    float4 *hdg = nullptr, *cdg = nullptr;
    float4 *fc = nullptr, *hc = nullptr;
    float* dn = nullptr;
    // active contacts (d < d_hat) of the current Newton iteration
    int *act = nullptr, *act_flag = nullptr, *act_pos = nullptr, *n_act = nullptr;
    float4* cslot = nullptr;  // per active-contact corner H dx terms, vertex-list order
    int4* cvert = nullptr;    // per active contact: vertices
    // per active-contact corner in vertex-list order: its distance gradient
    // and its contact; per active contact the CG scalar c2 (g . dx)
    float4* cdv = nullptr;
    int* cj = nullptr;
    float* sA = nullptr;
    float* c2j = nullptr;     // per active contact: barrier'' kappa
    int* c_dst = nullptr;     // list position of active-contact corner slot 4 j + a
    int nA = 0;
    uint32_t *skey = nullptr, *sval = nullptr, *skey2 = nullptr, *sval2 = nullptr;
    long long cap_sort = 0;
    // CG partial sums: z.r double-buffered, v.Av
    double *zr_part = nullptr, *vav_part = nullptr;
    // GT
    float4* gt_v = nullptr;
    int4* gt_f = nullptr;
    Tree gt_tree, tri_tree, edge_tree;
    bool contact_trees_built = false;  // tri_tree / edge_tree hold this mesh's hierarchy
    // contacts
    int capC = 0;
    int4* cidx = nullptr;
    int *ctype1 = nullptr, *ctype2 = nullptr;
    // detection order: second contact arrays (swapped with cidx / ctype1
    // after ordering), per-contact (wave, rank) tags, per-wave counts and bases
    int4* cidx2 = nullptr;
    int* ctype1b = nullptr;
    int2* ctag = nullptr;
    int *wave_count = nullptr, *wave_base = nullptr;
    int cap_wave = 0;
    float* cd = nullptr;
    float4* cdd = nullptr;  // 4 per contact
    int* ccount = nullptr;
    int* hcount = nullptr;
    // scalars (device): 0 energy, 1 energy_prev, 2 vAv, 3 zr, 4 zr_last
    double* sc = nullptr;
    double* hsc = nullptr;
    float* ccd_step = nullptr;
    double* part = nullptr;
    int cap_part = 0;
    // dense quadratic test energy
    double *dense_A = nullptr, *dense_b = nullptr;
    double dense_c = 0.0;
    int* ws = nullptr;
    long long cap_ws = 0;
    // statistics
    long long stat_contacts = 0, stat_newton = 0, stat_cg = 0, stat_ls = 0, stat_active = 0;
    // the CG loop replayed from a captured graph (updated each solve); off by
    // default: on the 20k-face headline output five steps took 0.169 s with
    // the graph and 0.163 s launching the kernels one by one
    int cg_graph = 0;
    hipGraphExec_t cg_exec = nullptr;
    double last_energy = 0.0;
    // vertex merge at the stage-2 -> stage-3 handoff: referenced flags, group
    // first vertex, kept flags and their prefix sum (n + 1 each), hash table
    int *mg_ref = nullptr, *mg_first = nullptr, *mg_keep = nullptr, *mg_pos = nullptr, *mg_tab = nullptr;
    int* mg_err = nullptr;
    long long cap_mg = 0, cap_mg_tab = 0;
};

enum { SC_E = 0, SC_EPREV = 1, SC_VAV = 2, SC_ZR = 3, SC_ZRLAST = 4, SC_N = 8 };

static void ctx_free(Ctx& c) {
    void* ptrs[] = {c.q, c.qn, c.qd, c.qrest, c.p, c.target, c.grad, c.diag, c.r, c.v, c.Av, c.z, c.vor, c.tri,
                    c.edge, c.hinge, c.area, c.rest_angle, c.rest_elen, c.inv_dm, c.hblock, c.gt_v, c.gt_f,
                    c.cidx, c.ctype1, c.ctype2, c.cd, c.cdd, c.ccount, c.sc, c.ccd_step, c.part, c.ws,
                    c.dense_A, c.dense_b, c.hsym, c.vf_off, c.vf_ent, c.vh_off, c.vh_ent, c.vc_off, c.vc_ent,
                    c.fc, c.hc, c.dn, c.act, c.act_flag, c.act_pos, c.n_act, c.cslot, c.c_dst, c.cvert, c.c2j, c.cdv, c.cj, c.sA, c.skey, c.sval, c.skey2,
                    c.vf_dst, c.vh_dst, c.hdg, c.cdg,
                    c.sval2, c.zr_part, c.vav_part,
                    c.mg_ref, c.mg_first, c.mg_keep, c.mg_pos, c.mg_tab, c.mg_err,
                    c.cidx2, c.ctype1b, c.ctag, c.wave_count, c.wave_base};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (c.hcount) (void)hipHostFree(c.hcount);
    if (c.hsc) (void)hipHostFree(c.hsc);
    if (c.cg_exec) (void)hipGraphExecDestroy(c.cg_exec);
    c.cg_exec = nullptr;
    tree_free(c.gt_tree);
    tree_free(c.tri_tree);
    tree_free(c.edge_tree);
}

static int ensure_ws(Ctx& c, long long n) {
    long long need = std::max(radix_workspace(n), scan_workspace(n + 1)) + 256;
    if (need > c.cap_ws) {
        c.cap_ws = need + need / 4;
        PAMO_TRY(dalloc(&c.ws, (size_t)c.cap_ws, c.stream));
    }
    return 0;
}

static int ensure_part(Ctx& c, long long items) {
    int need = (int)(items / kB + 8);
    if (need > c.cap_part) {
        c.cap_part = need + need / 4;
        PAMO_TRY(dalloc(&c.part, (size_t)c.cap_part, c.stream));
    }
    return 0;
}

// ------------------------------------------------------------ deterministic FP64 reductions
// Sum over the 64 lanes of a wave by an xor butterfly: each step adds a
// lane's value and its partner's, which both lanes form alike, so every lane
// ends with the same value whatever the scheduling.
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int m = kWave / 2; m > 0; m >>= 1) x += __shfl_xor(x, m);
    return x;
}

// Sum over the block: wave sums, then the waves' sums in wave order; every
// thread returns it.
__device__ __forceinline__ double block_sum(double x, double* red) {
    constexpr int kW = kB / kWave;
    x = wave_sum(x);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = x;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < kW; ++w) r += red[w];
    __syncthreads();
    return r;
}

// out[slot] = sum of part[0..n) in a fixed order (one block).
__global__ void __launch_bounds__(kB) finish_sum_k(const double* __restrict__ part, int n, double* out, int slot) {
    __shared__ double red[kB];
    double acc = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * kB < n; i += 4 * kB) {  // four loads in flight, added in order
        const double a = part[i], b = part[i + kB], c = part[i + 2 * kB], d = part[i + 3 * kB];
        acc += a;
        acc += b;
        acc += c;
        acc += d;
    }
    for (; i < n; i += kB) acc += part[i];
    double s = block_sum(acc, red);
    if (threadIdx.x == 0) out[slot] = s;
}


// ------------------------------------------------------------ preprocessing kernels
// elastic_preprocess_kernel (elastic_energy.py:8-38)
__global__ void elastic_pre_k(const float4* __restrict__ x, const int4* __restrict__ tri, int nF, float* area,
                              float4* inv_dm) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nF) return;
    int4 f = tri[t];
    V3 x0 = v3(x[f.x]), x1 = v3(x[f.y]), x2 = v3(x[f.z]);
    V3 x01 = x1 - x0, x02 = x2 - x0;
    float W = 0.5f * len(cross(x01, x02));
    V3 b1 = normalize(x01);
    V3 b2 = normalize(x02 - dot(x02, b1) * b1);
    float a = dot(b1, x01), b = dot(b1, x02), cc = dot(b2, x01), d = dot(b2, x02);
    float det = a * d - b * cc;
    area[t] = W;
    inv_dm[t] = make_float4(d / det, -b / det, -cc / det, a / det);
}

__global__ void hinge_rest_k(const float4* __restrict__ x, const int4* __restrict__ h, int nH, float* angle,
                             float* elen) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nH) return;
    int4 k = h[i];
    V3 x0 = v3(x[k.x]), x1 = v3(x[k.y]), x2 = v3(x[k.z]), x3 = v3(x[k.w]);
    angle[i] = hinge_angle(x0, x1, x2, x3);
    elen[i] = len(x1 - x2);
}

__global__ void normalize_points_k(const float* __restrict__ src, int n, double scale, double t0, double t1,
                                   double t2, float4* __restrict__ dst, float4* __restrict__ dst2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 o = make_float4((float)((double)src[3 * (long long)i] * scale + t0),
                           (float)((double)src[3 * (long long)i + 1] * scale + t1),
                           (float)((double)src[3 * (long long)i + 2] * scale + t2), 0.0f);
    dst[i] = o;
    if (dst2) dst2[i] = o;
}

__global__ void load_faces4_k(const int* __restrict__ src, int n, int4* __restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = make_int4(src[3 * (long long)i], src[3 * (long long)i + 1], src[3 * (long long)i + 2], 0);
}

__global__ void tri_items_k(const float4* __restrict__ x, const int4* __restrict__ f, int n, float inflate,
                            float4* __restrict__ item) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int4 t = f[i];
    V3 a = v3(x[t.x]), b = v3(x[t.y]), c = v3(x[t.z]);
    V3 lo = vmin(a, vmin(b, c)), hi = vmax(a, vmax(b, c));
    item[2 * (long long)i] = make_float4(lo.x - inflate, lo.y - inflate, lo.z - inflate, 0.0f);
    item[2 * (long long)i + 1] = make_float4(hi.x + inflate, hi.y + inflate, hi.z + inflate, 0.0f);
}

__global__ void edge_items_k(const float4* __restrict__ x, const int2* __restrict__ e, int n,
                             float4* __restrict__ item) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int2 t = e[i];
    V3 a = v3(x[t.x]), b = v3(x[t.y]);
    V3 lo = vmin(a, b), hi = vmax(a, b);
    item[2 * (long long)i] = f4(lo);
    item[2 * (long long)i + 1] = f4(hi);
}

// ------------------------------------------------------------ closest point on the input surface (B4)
// Nearest triangle by branch and bound over the GT tree: nearer child first,
// subtrees whose box is no closer than the best distance so far skipped.
__global__ void __launch_bounds__(kB) closest_point_k(const float4* __restrict__ x, int nV, TreeView T,
                                                     const float4* __restrict__ gv, const int4* __restrict__ gf,
                                                     float4* __restrict__ target, int* err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nV) return;
    const V3 p = v3(x[i]);
    float best = 1e18f;  // squared; query radius 1e9
    V3 bestp = mk(1e9f, 1e9f, 1e9f);
    const int n = T.n;
    auto leaf = [&](int f) {
        int4 t = gf[f];
        V3 c = closest_on_tri(p, v3(gv[t.x]), v3(gv[t.y]), v3(gv[t.z]));
        float d2 = len_sq(c - p);
        if (d2 < best) {
            best = d2;
            bestp = c;
        }
    };
    if (n == 1) {
        leaf((int)T.val[0]);
    } else if (n > 1) {
        int cur = 0, prev = -1;
        const long long max_steps = 3LL * (2LL * n - 1) + 1;
        for (long long steps = 0; cur >= 0; ++steps) {
            if (steps >= max_steps || cur > 2 * n - 2) {
                walk_overrun(err);
                break;
            }
            const int par = cur == 0 ? -1 : T.parent[cur];
            int next;
            if (prev == par) {
                if (box_dist2(p, T.box[2 * (long long)cur], T.box[2 * (long long)cur + 1]) >= best) next = par;
                else if (cur >= n - 1) {
                    leaf((int)T.val[cur - (n - 1)]);
                    next = par;
                } else {
                    int l = T.left[cur], r = T.right[cur];
                    float dl = box_dist2(p, T.box[2 * (long long)l], T.box[2 * (long long)l + 1]);
                    float dr = box_dist2(p, T.box[2 * (long long)r], T.box[2 * (long long)r + 1]);
                    next = dl <= dr ? l : r;
                }
            } else {
                int l = T.left[cur], r = T.right[cur];
                float dl = box_dist2(p, T.box[2 * (long long)l], T.box[2 * (long long)l + 1]);
                float dr = box_dist2(p, T.box[2 * (long long)r], T.box[2 * (long long)r + 1]);
                int first = dl <= dr ? l : r, second = dl <= dr ? r : l;
                next = prev == first ? second : par;
            }
            prev = cur;
            cur = next;
        }
    }
    target[i] = f4(bestp);
}

// ------------------------------------------------------------ contact detection (B5)
// detect_pt_contact_bvh_kernel / detect_ee_contact_bvh_kernel
// (contact_detection.py:119-216): PT = vertex i0 against every triangle not
// holding it whose box overlaps x0 +- (2r + d_hat); EE = edge e0 against every
// edge e1 > e0 sharing no vertex whose box overlaps e0's box grown by the same
// amount; a pair is kept when its classified distance is <= 2r + d_hat.
constexpr int kDetQ = 128;     // per-wave pair queue (< 64 pending + one append of <= 64)
constexpr int kDetW = 4;       // queries per wave
constexpr int kDetStack = 512; // frontier entries per wave

// One candidate (query item qa, tree item qb): the classified distance within
// thr keeps it as a contact block (contact_detection.py:120-214).
template <bool kPT>
__device__ __forceinline__ bool det_pair(const float4* __restrict__ x, const int4* __restrict__ tri,
                                         const int2* __restrict__ edge, int qa, int qb, float thr, float ee_thres,
                                         int4& ids) {
    if (kPT) {
        const int4 t = tri[qb];
        if (qa == t.x || qa == t.y || qa == t.z) return false;
        const V3 x0 = v3(x[qa]), x1 = v3(x[t.x]), x2 = v3(x[t.y]), x3 = v3(x[t.z]);
        const int ty = pt_classify(x0, x1, x2, x3);
        ids = make_int4(qa, t.x, t.y, t.z);
        return pt_distance(x0, x1, x2, x3, ty) <= thr;
    }
    const int2 e0 = edge[qa], e1 = edge[qb];
    if (e0.x == e1.x || e0.x == e1.y || e0.y == e1.x || e0.y == e1.y) return false;
    const V3 x0 = v3(x[e0.x]), x1 = v3(x[e0.y]), x2 = v3(x[e1.x]), x3 = v3(x[e1.y]);
    const int ty = ee_classify(x0, x1, x2, x3, ee_thres);
    ids = make_int4(e0.x, e0.y, e1.x, e1.y);
    return ee_distance(x0, x1, x2, x3, ty) <= thr;
}

// Output of the detection: contacts go to raw slots reserved with one atomic
// per emission step, each tagged with (wave, rank in the wave's emission
// order); the wave's count goes to wave_count.  A wave's emission order
// depends only on the trees and positions, so placing each contact at
// (exclusive scan of the wave counts)[wave] + rank (contact_order_k) lists
// the contacts in an order that does not depend on scheduling.
struct DetOut {
    int* counter;
    int cap;
    int4* cidx;
    int* ctype1;
    int2* ctag;
    int* wave_count;
    int wave0;  // global index of this launch's first wave
};

// The keeping lanes among the lanes active here emit in lane order; wcnt is
// the wave's running count in LDS.
__device__ __forceinline__ void det_emit(bool keep, int4 ids, int type, const DetOut& o, int wid, int* wcnt) {
    const uint64_t km = __ballot(keep);
    if (km == 0) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int leader = __ffsll((unsigned long long)km) - 1;
    const int k = (int)__popcll(km);
    int base = 0, before = 0;
    if (lane == leader) {
        base = atomicAdd(o.counter, k);
        before = *wcnt;
        *wcnt = before + k;
    }
    base = __shfl(base, leader);
    before = __shfl(before, leader);
    if (keep) {
        const int r = lanes_below(km);
        const int slot = base + r;
        if (slot < o.cap) {
            o.cidx[slot] = ids;
            o.ctype1[slot] = type;
            o.ctag[slot] = make_int2(wid, before + r);
        }
    }
}

// Stackless walk below internal node x (from its left child until x's
// escape), classifying and appending in this lane.
template <bool kPT>
__device__ __forceinline__ void det_walk_subtree(const float4* __restrict__ x, const int4* __restrict__ tri,
                                              const int2* __restrict__ edge, const float4* __restrict__ node,
                                              int n_leaves, int cur, int stop, V3 lo, V3 hi, int q, float thr,
                                              float ee_thres, const DetOut& o, int wid, int* wcnt, int* err) {
    for (long long visits = 0; cur != stop; ++visits) {
        if (visits >= 2LL * n_leaves || cur < 0) {
            walk_overrun(err);
            return;
        }
        const float4 blo = node[2 * (long long)cur], bhi = node[2 * (long long)cur + 1];
        const int lw = __float_as_int(blo.w), esc = __float_as_int(bhi.w);
        if (box_hit(lo, hi, blo, bhi)) {
            if (cur >= n_leaves - 1) {
                int4 ids = make_int4(0, 0, 0, 0);
                const bool keep = (kPT || lw > q) && det_pair<kPT>(x, tri, edge, q, lw, thr, ee_thres, ids);
                det_emit(keep, ids, kPT ? BLOCK_PT : BLOCK_EE, o, wid, wcnt);
                cur = esc;
            } else {
                cur = lw;
            }
        } else {
            cur = esc;
        }
    }
}

// PT (query = vertex, tree over triangles) or EE (query = edge, tree over
// edges, pairs taken from the larger edge index) candidates within
// thr = 2 r + d_hat of the query's box; classified contacts are appended at
// counter (contact_detection.py:120-214).
//
// A wave serves kDetW queries with one shared LIFO frontier of
// (query, node) entries in LDS: each step pops up to 64 entries, one per
// lane, pushes both children of overlapping internal nodes and queues
// overlapping leaves; the queue is classified 64 pairs at a time, and output
// slots are reserved with one atomic per wave.  When the frontier is full, a
// lane walks the rest of its subtree by escape links.
template <bool kPT>
__global__ void __launch_bounds__(kB) detect_k(const float4* __restrict__ x, int nq, const int4* __restrict__ tri,
                                               const int2* __restrict__ edge, const float4* __restrict__ node,
                                               const int* __restrict__ right, int n_leaves, float d_hat,
                                               float radius, float ee_thres, DetOut o, int* err) {
    constexpr int kW = kB / kWave;
    __shared__ int2 stk[kW][kDetStack];
    __shared__ int2 pq[kW][kDetQ];
    __shared__ float4 qbox[kW][kDetW][2];
    __shared__ int wcnt[kW];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    const long long wl = (long long)blockIdx.x * kW + w;
    const int wid = o.wave0 + (int)wl;
    const long long q0 = wl * kDetW;
    if (q0 >= nq || n_leaves <= 0) return;  // uniform per wave (such waves count 0: see contact_order)
    if (lane == 0) wcnt[w] = 0;
    const int nl = nq - q0 < kDetW ? (int)(nq - q0) : kDetW;
    const float thr = 2.0f * radius + d_hat;
    if (lane < nl) {
        const int q = (int)q0 + lane;
        V3 lo, hi;
        if (kPT) {
            const V3 x0 = v3(x[q]);
            lo = x0 - mk(thr, thr, thr);
            hi = x0 + mk(thr, thr, thr);
        } else {
            const int2 e = edge[q];
            const V3 a = v3(x[e.x]), b = v3(x[e.y]);
            lo = vmin(a, b) - mk(thr, thr, thr);
            hi = vmax(a, b) + mk(thr, thr, thr);
        }
        qbox[w][lane][0] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        qbox[w][lane][1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        stk[w][lane] = make_int2(lane, 0);  // root
    }
    wave_sync();
    // emits the classified pairs of pq[w][qn - take, qn) with one atomic per wave
    auto classify = [&](int qn, int take) {
        bool keep = false;
        int4 ids = make_int4(0, 0, 0, 0);
        if (lane < take) {
            const int2 e = pq[w][qn - take + lane];
            keep = det_pair<kPT>(x, tri, edge, (int)q0 + e.x, e.y, thr, ee_thres, ids);
        }
        det_emit(keep, ids, kPT ? BLOCK_PT : BLOCK_EE, o, wid, &wcnt[w]);
    };
    int size = nl, qn = 0;  // wave-uniform
    const long long max_steps = (long long)nl * (2LL * n_leaves - 1);
    for (long long steps = 0; size > 0; ++steps) {
        if (steps >= max_steps) {  // wave-uniform
            if (lane == 0) walk_overrun(err);
            if (lane == 0) o.wave_count[wid] = 0;
            return;
        }
        const int take = size < kWave ? size : kWave;
        int l = 0, cur = -1;
        if (lane < take) {
            const int2 e = stk[w][size - take + lane];
            l = e.x;
            cur = e.y;
        }
        size -= take;
        int c0 = -1, c1 = -1, esc = -1, leaf = -1;
        V3 lo = mk(0.f, 0.f, 0.f), hi = lo;
        if (cur >= 0) {
            const float4 blo = node[2 * (long long)cur], bhi = node[2 * (long long)cur + 1];
            const int r = cur < n_leaves - 1 ? right[cur] : -1;
            const float4 ql = qbox[w][l][0], qh = qbox[w][l][1];
            lo = mk(ql.x, ql.y, ql.z);
            hi = mk(qh.x, qh.y, qh.z);
            if (box_hit(lo, hi, blo, bhi)) {
                const int lw = __float_as_int(blo.w);
                if (cur >= n_leaves - 1) {
                    if (kPT || lw > (int)q0 + l) leaf = lw;
                } else {
                    c0 = lw;
                    c1 = r;
                    esc = __float_as_int(bhi.w);
                }
            }
        }
        wave_sync();  // every pop has read its entry before pushes reuse the slots
        const uint64_t lm = __ballot(leaf >= 0);
        if (leaf >= 0) pq[w][qn + lanes_below(lm)] = make_int2(l, leaf);
        qn += (int)__popcll(lm);
        const uint64_t pm = __ballot(c0 >= 0);
        const int below = lanes_below(pm);
        const bool fits = size + 2 * (below + 1) <= kDetStack;  // pushes that fit form a prefix
        if (c0 >= 0) {
            if (fits) {
                stk[w][size + 2 * below] = make_int2(l, c0);
                stk[w][size + 2 * below + 1] = make_int2(l, c1);
            } else {
                det_walk_subtree<kPT>(x, tri, edge, node, n_leaves, c0, esc, lo, hi, (int)q0 + l, thr, ee_thres,
                                      o, wid, &wcnt[w], err);
            }
        }
        size += 2 * (int)__popcll(__ballot(c0 >= 0 && fits));
        wave_sync();
        if (qn >= kWave) {
            classify(qn, kWave);
            qn -= kWave;
            wave_sync();
        }
    }
    if (qn > 0) classify(qn, qn);
    wave_sync();
    if (lane == 0) o.wave_count[wid] = wcnt[w];
}

// Contact s moves to (wave base)[ctag.x] + ctag.y.
__global__ void contact_order_k(int n, const int2* __restrict__ ctag, const int* __restrict__ wave_base,
                                const int4* __restrict__ cidx, const int* __restrict__ ctype1,
                                int4* __restrict__ cidx2, int* __restrict__ ctype1b) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int2 t = ctag[s];
    const int d = wave_base[t.x] + t.y;
    cidx2[d] = cidx[s];
    ctype1b[d] = ctype1[s];
}

// ------------------------------------------------------------ energies (B7)
struct Topo {
    const int4* tri;
    const float* area;
    const float4* inv_dm;
    const int4* hinge;
    const float *rest_angle, *rest_elen;
    const float* vor;
    const float4* target;
    const int4* cidx;
    const int* ctype1;
};

__device__ __forceinline__ void mat22_mul(const float A[4], const float B[4], float C[4]) {
    C[0] = A[0] * B[0] + A[1] * B[2];
    C[1] = A[0] * B[1] + A[1] * B[3];
    C[2] = A[2] * B[0] + A[3] * B[2];
    C[3] = A[2] * B[1] + A[3] * B[3];
}

struct Membrane {
    V3 b1, b2;
    float F[4], E[4], trE;
};

// Local 2D frame, deformation gradient and Green strain (elastic_energy.py:64-82)
__device__ __forceinline__ Membrane membrane(V3 x0, V3 x1, V3 x2, float4 idm) {
    Membrane m;
    V3 x01 = x1 - x0, x02 = x2 - x0;
    m.b1 = normalize(x01);
    m.b2 = normalize(x02 - dot(x02, m.b1) * m.b1);
    float Ds[4] = {dot(m.b1, x01), dot(m.b1, x02), dot(m.b2, x01), dot(m.b2, x02)};
    float Dm[4] = {idm.x, idm.y, idm.z, idm.w};
    mat22_mul(Ds, Dm, m.F);
    // E = 0.5 (F^T F - I)
    float FtF[4] = {m.F[0] * m.F[0] + m.F[2] * m.F[2], m.F[0] * m.F[1] + m.F[2] * m.F[3],
                    m.F[1] * m.F[0] + m.F[3] * m.F[2], m.F[1] * m.F[1] + m.F[3] * m.F[3]};
    m.E[0] = 0.5f * (FtF[0] - 1.0f);
    m.E[1] = 0.5f * FtF[1];
    m.E[2] = 0.5f * FtF[2];
    m.E[3] = 0.5f * (FtF[3] - 1.0f);
    m.trE = m.E[0] + m.E[3];
    return m;
}

// Per-item energy: items [0, nV) M2GT, [nV, +nF) elastic, [+nH) hinge,
// [+nC) collision (also stores the distance and its classification).
__global__ void __launch_bounds__(kB) energy_k(const float4* __restrict__ x, Topo tp, Params P, int nV, int nF,
                                               int nH, int nC, float* __restrict__ cd, int* __restrict__ ctype2,
                                               double* __restrict__ part, const float* __restrict__ reach) {
    __shared__ double red[kB];
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < nV) {
        V3 d = v3(x[i]) - v3(tp.target[i]);
        float w = tp.vor[i] * P.k_m2gt;
        e = (double)(0.5f * len_sq(d) * w);
    } else if ((i -= nV) < nF) {
        int4 f = tp.tri[i];
        Membrane m = membrane(v3(x[f.x]), v3(x[f.y]), v3(x[f.z]), tp.inv_dm[i]);
        // tr(E E^T) = sum of squares
        float tEEt = m.E[0] * m.E[0] + m.E[1] * m.E[1] + m.E[2] * m.E[2] + m.E[3] * m.E[3];
        e = (double)(tp.area[i] * (P.mu * tEEt + 0.5f * P.la * m.trE * m.trE));
    } else if ((i -= nF) < nH) {
        int4 k = tp.hinge[i];
        float th = hinge_angle(v3(x[k.x]), v3(x[k.y]), v3(x[k.z]), v3(x[k.w]));
        float dt = th - tp.rest_angle[i];
        e = (double)(0.5f * P.k_hinge * dt * dt * tp.rest_elen[i]);
    } else if ((i -= nH) < nC) {
        // Line-search points (reach given): x = q_n + f p with f <= the CCD
        // step s and |p_v| <= pmax, so a pair's distance differs from its
        // distance cd at q_n by at most 2 s pmax; a pair farther than d_hat
        // by that much (plus a tenth of d_hat for rounding) has energy 0 at
        // every such point and is not evaluated.  cd / ctype2 keep the q_n
        // values there (no later pass reads the line-search ones).
        // (a NaN bound or distance compares false: the pair is evaluated)
        const bool far = reach && cd[i] > P.d_hat + 2.0f * reach[0] * reach[1] + 0.1f * P.d_hat;
        if (!far) {
            int4 k = tp.cidx[i];
            V3 x0 = v3(x[k.x]), x1 = v3(x[k.y]), x2 = v3(x[k.z]), x3 = v3(x[k.w]);
            float d;
            int t2;
            if (tp.ctype1[i] == BLOCK_PT) {
                t2 = pt_classify(x0, x1, x2, x3);
                d = pt_distance(x0, x1, x2, x3, t2);
            } else {
                t2 = ee_classify(x0, x1, x2, x3, P.ee_thres);
                d = ee_distance(x0, x1, x2, x3, t2);
            }
            if (!reach) {
                ctype2[i] = t2;
                cd[i] = d;
            }
            if (d < P.d_hat) e = (double)(barrier(d, P.d_hat) * P.kappa);
        }
    }
    double s = block_sum(e, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}


// Symmetric eigen-clamp of the 9x9 leading block (cyclic Jacobi), then rows /
// columns of vertex 3 rebuilt by translation invariance (sum over a block
// row / column vanishes), as block_spd_project_kernel (utils_kernels.py:21-39).
// Also writes the projected 9x9 as a packed upper triangle (45 floats,
// row-major, entry k at S45[k * stride]: the hinges' k-th entries are
// contiguous), which is all the Hessian-vector product needs.
__device__ void spd_project_hinge(float* B /* 144: block (a,b) at 9*(4*a+b), row-major 3x3 */,
                                  float* __restrict__ S45, long long stride) {
    float A[9][9], Vm[9][9];
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            A[r][c] = B[9 * (4 * (r / 3) + (c / 3)) + 3 * (r % 3) + (c % 3)];
            Vm[r][c] = r == c ? 1.0f : 0.0f;
        }
    // symmetrise
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int c = r + 1; c < 9; ++c) {
            float s = 0.5f * (A[r][c] + A[c][r]);
            A[r][c] = s;
            A[c][r] = s;
        }
    for (int sweep = 0; sweep < 8; ++sweep) {
#pragma unroll
        for (int pp = 0; pp < 8; ++pp)
#pragma unroll
            for (int qq = pp + 1; qq < 9; ++qq) {
                float apq = A[pp][qq];
                if (fabsf(apq) > 1e-30f) {
                    float theta = (A[qq][qq] - A[pp][pp]) / (2.0f * apq);
                    float t = (theta >= 0.0f ? 1.0f : -1.0f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
                    float cs = 1.0f / sqrtf(t * t + 1.0f), sn = t * cs;
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        float akp = A[k][pp], akq = A[k][qq];
                        A[k][pp] = cs * akp - sn * akq;
                        A[k][qq] = sn * akp + cs * akq;
                    }
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        float apk = A[pp][k], aqk = A[qq][k];
                        A[pp][k] = cs * apk - sn * aqk;
                        A[qq][k] = sn * apk + cs * aqk;
                    }
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        float vkp = Vm[k][pp], vkq = Vm[k][qq];
                        Vm[k][pp] = cs * vkp - sn * vkq;
                        Vm[k][qq] = sn * vkp + cs * vkq;
                    }
                }
            }
    }
    float lam[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) lam[k] = fmaxf(A[k][k], 0.0f);
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < 9; ++k) s += Vm[r][k] * lam[k] * Vm[c][k];
            B[9 * (4 * (r / 3) + (c / 3)) + 3 * (r % 3) + (c % 3)] = s;
            if (c >= r) S45[(9 * r - (r * (r - 1)) / 2 + (c - r)) * stride] = s;
        }
    // row 3: block[3][j] = -(block[0][j] + block[1][j] + block[2][j]) for j < 3
    for (int j = 0; j < 3; ++j)
        for (int e = 0; e < 9; ++e)
            B[9 * (4 * 3 + j) + e] = -(B[9 * (4 * 0 + j) + e] + B[9 * (4 * 1 + j) + e] + B[9 * (4 * 2 + j) + e]);
    // column 3: block[j][3] = block[3][j]^T
    for (int j = 0; j < 3; ++j)
        for (int rr = 0; rr < 3; ++rr)
            for (int cc = 0; cc < 3; ++cc) B[9 * (4 * j + 3) + 3 * rr + cc] = B[9 * (4 * 3 + j) + 3 * cc + rr];
    // block[3][3] = -(block[0][3] + block[1][3] + block[2][3])
    for (int e = 0; e < 9; ++e)
        B[9 * (4 * 3 + 3) + e] = -(B[9 * (4 * 0 + 3) + e] + B[9 * (4 * 1 + 3) + e] + B[9 * (4 * 2 + 3) + e]);
}

struct M3 {
    float m[9];
};
__device__ __forceinline__ M3 outer(V3 a, V3 b) {
    return {{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z}};
}
__device__ __forceinline__ M3 tr(M3 a) { return {{a.m[0], a.m[3], a.m[6], a.m[1], a.m[4], a.m[7], a.m[2], a.m[5], a.m[8]}}; }
__device__ __forceinline__ M3 add(M3 a, M3 b) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] + b.m[k]; return r; }
__device__ __forceinline__ M3 sub(M3 a, M3 b) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] - b.m[k]; return r; }
__device__ __forceinline__ M3 scl(M3 a, float s) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] * s; return r; }
__device__ __forceinline__ M3 sym(M3 a) { return add(a, tr(a)); }
__device__ __forceinline__ void put(float* B, int a, int b, M3 x) { for (int k = 0; k < 9; ++k) B[9 * (4 * a + b) + k] = x.m[k]; }

// Gradient (coefficient `coeff`) and Jacobi diagonal terms and, for hinges,
// the projected 12x12 Hessian blocks; for contacts the distance gradient.
// Items as in energy_k.  The M2GT terms are stored to grad / diag; the
// terms of faces, hinges and active contacts go to their slots in the
// vertex lists (f_dst, h_dst, c_dst), summed per vertex by diff_gather_k.
struct DiffSlots {
    const int *f_dst, *h_dst, *c_dst;
    float4 *fc, *hc, *hdg, *cslot, *cdg;
    long long hstride;  // hsym: entry k of hinge i at k * hstride + i
    // per active contact j for the Hessian-vector products: vertices, barrier''
    // kappa, distance gradient of corner a at cdd[a * cstride + j]
    int4* cvert;
    float* c2j;
    long long cstride;
};
__global__ void __launch_bounds__(kB) diff_k(const float4* __restrict__ x, Topo tp, Params P, int nV, int nF, int nH,
                                             int nC, float coeff, const float* __restrict__ cd,
                                             const int* __restrict__ ctype2, float4* __restrict__ grad,
                                             float4* __restrict__ diag, float* __restrict__ hblock,
                                             float* __restrict__ hsym, const int* __restrict__ act,
                                             const int* __restrict__ n_act, float4* __restrict__ cdd, DiffSlots ds) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nV) {
        float w = tp.vor[i] * P.k_m2gt;
        V3 g = (v3(x[i]) - v3(tp.target[i])) * w;
        grad[i] = f4(g * coeff);
        diag[i] = f4(mk(w, w, w));
    } else if ((i -= nV) < nF) {
        int4 f = tp.tri[i];
        Membrane m = membrane(v3(x[f.x]), v3(x[f.y]), v3(x[f.z]), tp.inv_dm[i]);
        // P = F (2 mu E + la tr(E) I); H = coeff W P inv_Dm^T
        float S[4] = {2.0f * P.mu * m.E[0] + P.la * m.trE, 2.0f * P.mu * m.E[1], 2.0f * P.mu * m.E[2],
                      2.0f * P.mu * m.E[3] + P.la * m.trE};
        float Pm[4];
        mat22_mul(m.F, S, Pm);
        float4 idm = tp.inv_dm[i];
        float IdT[4] = {idm.x, idm.z, idm.y, idm.w};
        float H[4];
        mat22_mul(Pm, IdT, H);
        float cw = coeff * tp.area[i];
        for (int k = 0; k < 4; ++k) H[k] *= cw;
        V3 f1 = H[0] * m.b1 + H[2] * m.b2;
        V3 f2 = H[1] * m.b1 + H[3] * m.b2;
        V3 f0 = -f1 - f2;
        ds.fc[ds.f_dst[3 * i]] = f4(f0);
        ds.fc[ds.f_dst[3 * i + 1]] = f4(f1);
        ds.fc[ds.f_dst[3 * i + 2]] = f4(f2);
    } else if ((i -= nF) < nH) {
        int4 k = tp.hinge[i];
        V3 x0 = v3(x[k.x]), x1 = v3(x[k.y]), x2 = v3(x[k.z]), x3 = v3(x[k.w]);
        V3 e0 = x2 - x1, e1 = x0 - x2, e2 = x0 - x1, e1_ = x3 - x2, e2_ = x3 - x1;
        float e0_lsq = len_sq(e0);
        float e0_len = sqrtf(e0_lsq), e1_len = sqrtf(len_sq(e1)), e2_len = sqrtf(len_sq(e2));
        float e1_len_ = sqrtf(len_sq(e1_)), e2_len_ = sqrtf(len_sq(e2_));
        V3 e0u = normalize(e0), e1u = normalize(e1), e2u = normalize(e2), e1u_ = normalize(e1_), e2u_ = normalize(e2_);
        float ca1 = dot(e2u, e0u), ca2 = -dot(e1u, e0u), ca1_ = dot(e2u_, e0u), ca2_ = -dot(e1u_, e0u);
        float h0 = len(cross(e1, e2)) / e0_len, h1 = len(cross(e0, e2)) / e1_len, h2 = len(cross(e0, e1)) / e2_len;
        float h0_ = len(cross(e1_, e2_)) / e0_len, h1_ = len(cross(e0, e2_)) / e1_len_,
              h2_ = len(cross(e0, e1_)) / e2_len_;
        V3 n = normalize(cross(e0, e2)), n_ = normalize(cross(e2_, e0));
        float cth = dot(n, n_), sth = dot(cross(n, n_), e0) / e0_len;
        float th = atan2f(sth, cth);
        float dth = th - tp.rest_angle[i];
        float dE = P.k_hinge * dth * tp.rest_elen[i];
        float d2E = P.k_hinge * tp.rest_elen[i];
        V3 g1 = n * (ca2 / h1) + n_ * (ca2_ / h1_);
        V3 g2 = n * (ca1 / h2) + n_ * (ca1_ / h2_);
        V3 g0 = n * (-1.0f / h0);
        V3 g3 = n_ * (-1.0f / h0_);
        ds.hc[ds.h_dst[4 * i]] = f4(dE * g0 * coeff);
        ds.hc[ds.h_dst[4 * i + 1]] = f4(dE * g1 * coeff);
        ds.hc[ds.h_dst[4 * i + 2]] = f4(dE * g2 * coeff);
        ds.hc[ds.h_dst[4 * i + 3]] = f4(dE * g3 * coeff);
        V3 m0 = normalize(cross(e0, n)), m1 = normalize(cross(e1, n)), m2 = normalize(cross(n, e2));
        V3 m0_ = normalize(cross(n_, e0)), m1_ = normalize(cross(n_, e1_)), m2_ = normalize(cross(e2_, n_));
        M3 M0 = outer(n, m0), M1 = outer(n, m1), M2 = outer(n, m2);
        M3 M0_ = outer(n_, m0_), M1_ = outer(n_, m1_), M2_ = outer(n_, m2_);
        M3 N0 = scl(M0, 1.0f / e0_lsq), N0_ = scl(M0_, 1.0f / e0_lsq);
        M3 P10 = scl(tr(M0), ca1 / (h1 * h0)), P11 = scl(tr(M1), ca1 / (h1 * h1)), P12 = scl(tr(M2), ca1 / (h1 * h2));
        M3 P20 = scl(tr(M0), ca2 / (h2 * h0)), P21 = scl(tr(M1), ca2 / (h2 * h1)), P22 = scl(tr(M2), ca2 / (h2 * h2));
        M3 P10_ = scl(tr(M0_), ca1_ / (h1_ * h0_)), P11_ = scl(tr(M1_), ca1_ / (h1_ * h1_)),
           P12_ = scl(tr(M2_), ca1_ / (h1_ * h2_));
        M3 P20_ = scl(tr(M0_), ca2_ / (h2_ * h0_)), P21_ = scl(tr(M1_), ca2_ / (h2_ * h1_)),
           P22_ = scl(tr(M2_), ca2_ / (h2_ * h2_));
        M3 Q0 = scl(M0, 1.0f / (h0 * h0)), Q1 = scl(M1, 1.0f / (h0 * h1)), Q2 = scl(M2, 1.0f / (h0 * h2));
        M3 Q0_ = scl(M0_, 1.0f / (h0_ * h0_)), Q1_ = scl(M1_, 1.0f / (h0_ * h1_)), Q2_ = scl(M2_, 1.0f / (h0_ * h2_));
        float* B = hblock + 144 * (long long)i;
        put(B, 0, 0, add(scl(sym(Q0), -dE), scl(outer(g0, g0), d2E)));
        put(B, 3, 3, add(scl(sym(Q0_), -dE), scl(outer(g3, g3), d2E)));
        put(B, 1, 1, add(scl(sub(add(sub(sym(P11), N0), sym(P11_)), N0_), dE), scl(outer(g1, g1), d2E)));
        put(B, 2, 2, add(scl(sub(add(sub(sym(P22), N0), sym(P22_)), N0_), dE), scl(outer(g2, g2), d2E)));
        M3 b10 = add(scl(sub(P10, Q1), dE), scl(outer(g1, g0), d2E));
        M3 b20 = add(scl(sub(P20, Q2), dE), scl(outer(g2, g0), d2E));
        M3 b13 = add(scl(sub(P10_, Q1_), dE), scl(outer(g1, g3), d2E));
        M3 b23 = add(scl(sub(P20_, Q2_), dE), scl(outer(g2, g3), d2E));
        M3 b12 = add(scl(add(add(add(add(add(P12, tr(P21)), N0), P12_), tr(P21_)), N0_), dE), scl(outer(g1, g2), d2E));
        M3 b03 = scl(outer(g0, g3), d2E);
        put(B, 1, 0, b10); put(B, 0, 1, tr(b10));
        put(B, 2, 0, b20); put(B, 0, 2, tr(b20));
        put(B, 1, 3, b13); put(B, 3, 1, tr(b13));
        put(B, 2, 3, b23); put(B, 3, 2, tr(b23));
        put(B, 1, 2, b12); put(B, 2, 1, tr(b12));
        put(B, 0, 3, b03); put(B, 3, 0, tr(b03));
        // the Jacobi diagonal takes the blocks before projection (hinge_energy.py:255-256)
        for (int a = 0; a < 4; ++a) {
            const float* D = B + 9 * (4 * a + a);
            ds.hdg[ds.h_dst[4 * i + a]] = f4(mk(D[0], D[4], D[8]));
        }
        spd_project_hinge(B, hsym + i, ds.hstride);
    } else if ((i -= nH) < nC && i < n_act[0]) {
        // active contacts only (d < d_hat); their distance gradients are kept
        // in list order for the Hessian-vector products
        const long long j = i;
        i = act[j];
        float d = cd[i];
        int4 k = tp.cidx[i];
        V3 x0 = v3(x[k.x]), x1 = v3(x[k.y]), x2 = v3(x[k.z]), x3 = v3(x[k.w]);
        V3 g[4];
        if (tp.ctype1[i] == BLOCK_PT) pt_pair_grad(x0, x1, x2, x3, ctype2[i], g);
        else ee_pair_grad(x0, x1, x2, x3, ctype2[i], g);
        // barrier' kappa coeff g and barrier'' kappa g*g per corner
        const float c1 = barrier_d1(d, P.d_hat) * P.kappa * coeff;
        const float c2 = barrier_d2(d, P.d_hat) * P.kappa;
        ds.cvert[j] = k;
        ds.c2j[j] = c2;
        for (int a = 0; a < 4; ++a) {
            cdd[a * ds.cstride + j] = f4(g[a]);
            const int e = ds.c_dst[4 * j + a];
            ds.cslot[e] = f4(g[a] * c1);
            ds.cdg[e] = f4(c2 * mk(g[a].x * g[a].x, g[a].y * g[a].y, g[a].z * g[a].z));
        }
    }
}

// ------------------------------------------------------------ Hessian-vector products and PCG (B7, B8)
// The product y = H dx is formed without atomics in two passes: per element
// (membrane face, hinge, active contact, dense row) the contribution to each
// of its vertices is written to its own slot; then per vertex the slots of
// its incident elements (incidence lists sorted by element) are summed in a
// fixed order, with the M2GT term w dx added.  The CG direction is formed on
// the fly, dx = z + beta v, beta = zr_k / zr_{k-1} re-reduced inside every
// block from the FP64 partial sums (fixed order), so one CG iteration is
// three launches: elements, gather (+ v . Av partials), update (+ z . r
// partials).

// sum of part[0..n) in a fixed order; every thread returns it.  Up to 4
// partials per lane every wave forms it on its own (lane-strided sums, then
// the butterfly), with no barrier; more partials are spread over the whole
// block (all its threads must call).
__device__ __forceinline__ double reduce_parts(const double* __restrict__ part, int n, double* red) {
    if (n > 4 * kWave) {
        double acc = 0.0;
        int i = threadIdx.x;
        for (; i + 3 * kB < n; i += 4 * kB) {
            const double a = part[i], b = part[i + kB], c = part[i + 2 * kB], d = part[i + 3 * kB];
            acc += a;
            acc += b;
            acc += c;
            acc += d;
        }
        for (; i < n; i += kB) acc += part[i];
        return block_sum(acc, red);
    }
    double acc = 0.0;
    int i = threadIdx.x & (kWave - 1);
    for (; i + 3 * kWave < n; i += 4 * kWave) {  // four loads in flight, added in order
        const double a = part[i], b = part[i + kWave], c = part[i + 2 * kWave], d = part[i + 3 * kWave];
        acc += a;
        acc += b;
        acc += c;
        acc += d;
    }
    for (; i < n; i += kWave) acc += part[i];
    return wave_sum(acc);
}

struct CgVec {
    const float4* z;      // current z
    const float4* v_old;  // previous direction (unused when first)
    const double* zr_new; // partials of z_k . r_k
    const double* zr_old; // partials of z_{k-1} . r_{k-1}
    int nparts;
    int first;            // dx = z (k = 0, or a plain product)
};

__device__ __forceinline__ float cg_beta(const CgVec& cv, double* red) {
    if (cv.first) return 0.0f;
    const double a = reduce_parts(cv.zr_new, cv.nparts, red);
    const double b = reduce_parts(cv.zr_old, cv.nparts, red);
    return (float)(a / b);
}
__device__ __forceinline__ V3 cg_dx(const CgVec& cv, float beta, int v) {
    return cv.first ? v3(cv.z[v]) : v3(cv.z[v]) + beta * v3(cv.v_old[v]);
}

// items: [0, nF) faces -> fc[f_dst[3f + k]]; [nF, +nH) hinges -> hc[h_dst[4h + a]]
// (vertex-list order, so the gather reads them contiguously);
// [.., +nA) active contacts -> sA[j] = c2 (g . dx);
// [.., +3 nV_dense) dense rows
__global__ void __launch_bounds__(kB) hvp_elem_k(const float4* __restrict__ x, Topo tp, Params P, int nF, int nH,
                                                 int nA_max, const int* __restrict__ n_act, int n_dense,
                                                 CgVec cv, const int4* __restrict__ cvert, const float* __restrict__ c2j,
                                                 const float* __restrict__ hsym, long long hstride,
                                                 const float4* __restrict__ cdd, long long cstride,
                                                 const double* __restrict__ dA, const int* __restrict__ f_dst,
                                                 const int* __restrict__ h_dst, float4* __restrict__ fc,
                                                 float4* __restrict__ hc, float* __restrict__ sA,
                                                 float* __restrict__ dn) {
    __shared__ double red[kB];
    const float beta = cg_beta(cv, red);
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nF) {
        // elastic_hess_dx_kernel (elastic_energy.py:149-220), local basis fixed
        int4 f = tp.tri[i];
        V3 x0 = v3(x[f.x]), x1 = v3(x[f.y]), x2 = v3(x[f.z]);
        V3 d0 = cg_dx(cv, beta, f.x), d1 = cg_dx(cv, beta, f.y), d2 = cg_dx(cv, beta, f.z);
        float4 idm = tp.inv_dm[i];
        Membrane m = membrane(x0, x1, x2, idm);
        V3 dx01 = d1 - d0, dx02 = d2 - d0;
        float dDs[4] = {dot(m.b1, dx01), dot(m.b1, dx02), dot(m.b2, dx01), dot(m.b2, dx02)};
        float Dm[4] = {idm.x, idm.y, idm.z, idm.w};
        float dF[4];
        mat22_mul(dDs, Dm, dF);
        float a[4] = {dF[0] * m.F[0] + dF[2] * m.F[2], dF[0] * m.F[1] + dF[2] * m.F[3],
                      dF[1] * m.F[0] + dF[3] * m.F[2], dF[1] * m.F[1] + dF[3] * m.F[3]};
        float b[4] = {m.F[0] * dF[0] + m.F[2] * dF[2], m.F[0] * dF[1] + m.F[2] * dF[3],
                      m.F[1] * dF[0] + m.F[3] * dF[2], m.F[1] * dF[1] + m.F[3] * dF[3]};
        float dE[4];
        for (int k = 0; k < 4; ++k) dE[k] = 0.5f * (a[k] + b[k]);
        float trdE = dE[0] + dE[3];
        float S[4] = {2.0f * P.mu * m.E[0] + P.la * m.trE, 2.0f * P.mu * m.E[1], 2.0f * P.mu * m.E[2],
                      2.0f * P.mu * m.E[3] + P.la * m.trE};
        float dS[4] = {2.0f * P.mu * dE[0] + P.la * trdE, 2.0f * P.mu * dE[1], 2.0f * P.mu * dE[2],
                       2.0f * P.mu * dE[3] + P.la * trdE};
        float t1[4], t2[4];
        mat22_mul(dF, S, t1);
        mat22_mul(m.F, dS, t2);
        float dP[4] = {t1[0] + t2[0], t1[1] + t2[1], t1[2] + t2[2], t1[3] + t2[3]};
        float IdT[4] = {idm.x, idm.z, idm.y, idm.w};
        float dH[4];
        mat22_mul(dP, IdT, dH);
        float W = tp.area[i];
        for (int k = 0; k < 4; ++k) dH[k] *= W;
        V3 df1 = dH[0] * m.b1 + dH[2] * m.b2;
        V3 df2 = dH[1] * m.b1 + dH[3] * m.b2;
        fc[f_dst[3 * i]] = f4(-df1 - df2);
        fc[f_dst[3 * i + 1]] = f4(df1);
        fc[f_dst[3 * i + 2]] = f4(df2);
    } else if ((i -= nF) < nH) {
        // projected hinge Hessian: rows / columns of vertex 3 follow from the
        // 9x9 block by translation invariance, so y = A (d_a - d_3) and the
        // fourth vertex receives -(y_0 + y_1 + y_2)
        int4 k = tp.hinge[i];
        V3 d3 = cg_dx(cv, beta, k.w);
        V3 e[3] = {cg_dx(cv, beta, k.x) - d3, cg_dx(cv, beta, k.y) - d3, cg_dx(cv, beta, k.z) - d3};
        float u[9] = {e[0].x, e[0].y, e[0].z, e[1].x, e[1].y, e[1].z, e[2].x, e[2].y, e[2].z};
        const float* S = hsym + i;
        float y[9];
#pragma unroll
        for (int r = 0; r < 9; ++r) y[r] = 0.0f;
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int c = r; c < 9; ++c) {
                const float a = S[(9 * r - (r * (r - 1)) / 2 + (c - r)) * hstride];
                y[r] += a * u[c];
                if (c != r) y[c] += a * u[r];
            }
        V3 y0 = mk(y[0], y[1], y[2]), y1 = mk(y[3], y[4], y[5]), y2 = mk(y[6], y[7], y[8]);
        hc[h_dst[4 * i]] = f4(y0);
        hc[h_dst[4 * i + 1]] = f4(y1);
        hc[h_dst[4 * i + 2]] = f4(y2);
        hc[h_dst[4 * i + 3]] = f4(-(y0 + y1 + y2));
    } else if ((i -= nH) < nA_max) {
        if (i >= n_act[0]) return;
        const int4 k = cvert[i];
        const float c2 = c2j[i];
        const V3 g0 = v3(cdd[i]), g1 = v3(cdd[cstride + i]), g2 = v3(cdd[2 * cstride + i]),
                 g3 = v3(cdd[3 * cstride + i]);
        const float sdot = dot(g0, cg_dx(cv, beta, k.x)) + dot(g1, cg_dx(cv, beta, k.y)) +
                           dot(g2, cg_dx(cv, beta, k.z)) + dot(g3, cg_dx(cv, beta, k.w));
        // H_c dx = g (c2 g . dx): the scalar per contact, contiguous; the
        // gather multiplies each corner's gradient by it
        sA[i] = sdot * c2;
    } else if ((i -= nA_max) < n_dense) {
        const int n3 = (int)n_dense;
        double acc = 0.0;
        for (int kk = 0; kk < n3; ++kk) {
            const V3 d = cg_dx(cv, beta, kk / 3);
            const float comp = kk % 3 == 0 ? d.x : (kk % 3 == 1 ? d.y : d.z);
            acc += dA[i * n3 + kk] * (double)comp;
        }
        dn[i] = (float)acc;
    }
}

// lanes per vertex in the per-vertex gathers
constexpr int kGather = 32;

// acc + s[e] + s[e + kGather] + ... for e < end, in that order; four loads
// are issued before their adds, so a long list (a vertex with hundreds of
// contact corners) waits for one memory latency per four entries
__device__ __forceinline__ V3 gather_add(V3 acc, const float4* __restrict__ s, int e, int end) {
    for (; e + 3 * kGather < end; e += 4 * kGather) {
        const float4 a = s[e], b = s[e + kGather], c = s[e + 2 * kGather], d = s[e + 3 * kGather];
        acc = acc + v3(a);
        acc = acc + v3(b);
        acc = acc + v3(c);
        acc = acc + v3(d);
    }
    for (; e < end; e += kGather) acc = acc + v3(s[e]);
    return acc;
}

// acc + g[e] sA[cj[e]] + ... for e < end (stride kGather), in that order,
// four entries' loads issued before their adds
__device__ __forceinline__ V3 gather_contacts(V3 acc, const float4* __restrict__ g, const int* __restrict__ cj,
                                              const float* __restrict__ sA, int e, int end) {
    for (; e + 3 * kGather < end; e += 4 * kGather) {
        const int j0 = cj[e], j1 = cj[e + kGather], j2 = cj[e + 2 * kGather], j3 = cj[e + 3 * kGather];
        const float4 g0 = g[e], g1 = g[e + kGather], g2 = g[e + 2 * kGather], g3 = g[e + 3 * kGather];
        const float s0 = sA[j0], s1 = sA[j1], s2 = sA[j2], s3 = sA[j3];
        acc = acc + v3(g0) * s0;
        acc = acc + v3(g1) * s1;
        acc = acc + v3(g2) * s2;
        acc = acc + v3(g3) * s3;
    }
    for (; e < end; e += kGather) acc = acc + v3(g[e]) * sA[cj[e]];
    return acc;
}

struct Incid {
    const int* vf_off;  // vertex -> face corner entries of fc (list order)
    const int* vh_off;  // vertex -> hinge corner entries of hc / hdg (list order)
    const int* vc_off;  // vertex -> active contact corner entries of cslot / cdg (list order)
};

// Av = H dx per vertex; the direction dx is stored to v_out; block partials
// of dx . Av to vav.  A group of kGather lanes serves one vertex (its
// incidence lists are strided over the group and reduced with a fixed
// butterfly), so a vertex with many incident contacts does not hold up a
// whole wave.

// Gradient and Jacobi diagonal per vertex: diff_k leaves the M2GT term in
// grad / diag and every element's term in its slot of the vertex lists
// (faces fc, hinges hc and hdg, active contacts cslot and cdg); kGather lanes
// per vertex sum the slots in list order (faces, then hinges, then contacts,
// strided over the group) and a fixed shuffle tree, so the result does not
// depend on scheduling.
__global__ void __launch_bounds__(kB) diff_gather_k(int nV, int use_elem, int use_hinge, int use_contact, Incid ic,
                                                    const float4* __restrict__ fc, const float4* __restrict__ hc,
                                                    const float4* __restrict__ hdg, const float4* __restrict__ cslot,
                                                    const float4* __restrict__ cdg, float4* __restrict__ grad,
                                                    float4* __restrict__ diag) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int v = (int)(t / kGather), g = (int)(t % kGather);
    V3 ga = mk(0.0f, 0.0f, 0.0f), da = ga;
    if (v < nV) {
        if (use_elem) ga = gather_add(ga, fc, ic.vf_off[v] + g, ic.vf_off[v + 1]);
        if (use_hinge) {
            ga = gather_add(ga, hc, ic.vh_off[v] + g, ic.vh_off[v + 1]);
            da = gather_add(da, hdg, ic.vh_off[v] + g, ic.vh_off[v + 1]);
        }
        if (use_contact) {
            ga = gather_add(ga, cslot, ic.vc_off[v] + g, ic.vc_off[v + 1]);
            da = gather_add(da, cdg, ic.vc_off[v] + g, ic.vc_off[v + 1]);
        }
    }
#pragma unroll
    for (int m = kGather / 2; m > 0; m >>= 1) {
        ga.x += __shfl_xor(ga.x, m, kGather);
        ga.y += __shfl_xor(ga.y, m, kGather);
        ga.z += __shfl_xor(ga.z, m, kGather);
        da.x += __shfl_xor(da.x, m, kGather);
        da.y += __shfl_xor(da.y, m, kGather);
        da.z += __shfl_xor(da.z, m, kGather);
    }
    if (v < nV && g == 0) {
        grad[v] = f4(v3(grad[v]) + ga);
        diag[v] = f4(v3(diag[v]) + da);
    }
}

__global__ void __launch_bounds__(kB) hvp_gather_k(int nV, Params P, int use_m2gt, int use_elem, int use_hinge,
                                                   int use_contact, int use_dense, const float* __restrict__ vor,
                                                   CgVec cv, Incid ic, const float4* __restrict__ fc,
                                                   const float4* __restrict__ hc, const float4* __restrict__ cdv,
                                                   const int* __restrict__ cj, const float* __restrict__ sA,
                                                   const float4* __restrict__ cdd, const float* __restrict__ dn,
                                                   float4* __restrict__ v_out, float4* __restrict__ Av,
                                                   double* __restrict__ vav) {
    __shared__ double red[kB];
    const float beta = cg_beta(cv, red);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int v = (int)(t / kGather), g = (int)(t % kGather);
    V3 acc = mk(0.0f, 0.0f, 0.0f);
    if (v < nV) {
        if (use_elem) acc = gather_add(acc, fc, ic.vf_off[v] + g, ic.vf_off[v + 1]);
        if (use_hinge) acc = gather_add(acc, hc, ic.vh_off[v] + g, ic.vh_off[v + 1]);
        if (use_contact) acc = gather_contacts(acc, cdv, cj, sA, ic.vc_off[v] + g, ic.vc_off[v + 1]);
    }
#pragma unroll
    for (int m = kGather / 2; m > 0; m >>= 1) {
        acc.x += __shfl_xor(acc.x, m, kGather);
        acc.y += __shfl_xor(acc.y, m, kGather);
        acc.z += __shfl_xor(acc.z, m, kGather);
    }
    double dotp = 0.0;
    if (v < nV && g == 0) {
        const V3 d = cg_dx(cv, beta, v);
        if (use_m2gt) acc = acc + d * (vor[v] * P.k_m2gt);
        if (use_dense) acc = acc + mk(dn[3 * v], dn[3 * v + 1], dn[3 * v + 2]);
        Av[v] = f4(acc);
        v_out[v] = f4(d);
        dotp = (double)d.x * acc.x + (double)d.y * acc.y + (double)d.z * acc.z;
    }
    const double sum = block_sum(dotp, red);
    if (threadIdx.x == 0) vav[blockIdx.x] = sum;
}

// r = grad; z = r / (diag + 1e-6); p = 0; partials of z . r (cg_solver.py:120-142)
__global__ void __launch_bounds__(kB) cg_init_k(const float4* __restrict__ grad, const float4* __restrict__ diag,
                                                int n, float4* __restrict__ r, float4* __restrict__ z,
                                                float4* __restrict__ p, double* __restrict__ part) {
    __shared__ double red[kB];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double x = 0.0;
    if (i < n) {
        V3 rr = v3(grad[i]), dg = v3(diag[i]);
        V3 zz = mk(rr.x / (dg.x + 1e-6f), rr.y / (dg.y + 1e-6f), rr.z / (dg.z + 1e-6f));
        r[i] = f4(rr);
        z[i] = f4(zz);
        p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        x = (double)zz.x * rr.x + (double)zz.y * rr.y + (double)zz.z * rr.z;
    }
    double s = block_sum(x, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// alpha = zr_k / vAv (if vAv > 1e-16); p += alpha v; r -= alpha Av; z = r / diag;
// partials of z . r (update_p_r_z_compute_zr_kernel, solver_kernels.py:4-23)
__global__ void __launch_bounds__(kB) cg_update_k(const float4* __restrict__ v, const float4* __restrict__ Av,
                                                  const float4* __restrict__ diag, const double* __restrict__ vav,
                                                  int n_vav, const double* __restrict__ zr_cur, int nparts, int n,
                                                  float4* __restrict__ p, float4* __restrict__ r,
                                                  float4* __restrict__ z, double* __restrict__ zr_next) {
    __shared__ double red[kB];
    const double vAv = reduce_parts(vav, n_vav, red);
    const double zr = reduce_parts(zr_cur, nparts, red);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double x = 0.0;
    if (i < n) {
        V3 rr = v3(r[i]);
        if (vAv > 1e-16) {
            float alpha = (float)(zr / vAv);
            p[i] = f4(v3(p[i]) + alpha * v3(v[i]));
            rr = rr - alpha * v3(Av[i]);
            r[i] = f4(rr);
        }
        V3 dg = v3(diag[i]);
        V3 zz = mk(rr.x / dg.x, rr.y / dg.y, rr.z / dg.z);
        z[i] = f4(zz);
        x = (double)zz.x * rr.x + (double)zz.y * rr.y + (double)zz.z * rr.z;
    }
    double s = block_sum(x, red);
    if (threadIdx.x == 0) zr_next[blockIdx.x] = s;
}

// ---- incidence lists
// dst[ent[e]] = e: the list position of every slot
__global__ void invert_k(const int* __restrict__ ent, int n, int* __restrict__ dst) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[ent[e]] = e;
}
// the active-contact corner of list entry e (slot s = 4 j + a): its distance
// gradient (corner-major cdd) and its contact j
__global__ void contact_lists_k(const int* __restrict__ ent, int n, const float4* __restrict__ cdd, long long cstride,
                                float4* __restrict__ cdv, int* __restrict__ cj) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s = ent[e], j = s >> 2, a = s & 3;
    cdv[e] = cdd[a * cstride + j];
    cj[e] = j;
}
// keys[s] = vertex of slot s of element s / width, vals[s] = s
__global__ void slot_keys_k(const int* __restrict__ elem, int width, int stride, int n_slots,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    const int e = s / width, k = s % width;
    keys[s] = (uint32_t)elem[(long long)e * stride + k];
    vals[s] = (uint32_t)s;
}
// active contacts: slot s = 4 j + a of active contact act[j]
__global__ void contact_keys_k(const int4* __restrict__ cidx, const int* __restrict__ act, const int* __restrict__ n_act,
                               uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= 4 * n_act[0]) return;
    const int4 k = cidx[act[s >> 2]];
    const int a = s & 3;
    keys[s] = (uint32_t)(a == 0 ? k.x : (a == 1 ? k.y : (a == 2 ? k.z : k.w)));
    vals[s] = (uint32_t)s;
}
// off[v] = first position whose sorted key is >= v, for v in [0, nV]
__global__ void slot_offsets_k(const uint32_t* __restrict__ keys, int n, int nV, int* __restrict__ off) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int lo = i == 0 ? 0 : (int)keys[i - 1] + 1;
    const int hi = i == n ? nV : (int)keys[i];
    for (int v = lo; v <= hi; ++v) off[v] = i;
}
// active contacts: cd < d_hat
__global__ void act_flag_k(const float* __restrict__ cd, int nC, float d_hat, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nC) return;
    flag[i] = (i < nC && cd[i] < d_hat) ? 1 : 0;
}
__global__ void act_list_k(const int* __restrict__ flag, const int* __restrict__ pos, int nC, int* __restrict__ act,
                           int* __restrict__ n_act) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nC) return;
    if (i == nC) n_act[0] = pos[nC];
    else if (flag[i]) act[pos[i]] = i;
}

__global__ void copy_slot_k(double* sc, int from, int to) { sc[to] = sc[from]; }

// ------------------------------------------------------------ clamp, CCD, line search (B9, B10)
// also the largest |p| after the clamp into ccd[1] (integer max on the
// non-negative float's bits, one atomic per wave)
__global__ void clamp_k(float4* __restrict__ p, const float4* __restrict__ qn, const float4* __restrict__ qd, int n,
                        float radius, float* __restrict__ ccd) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (i < n) {
        V3 pi = v3(p[i]), qni = v3(qn[i]), qdi = v3(qd[i]);
        V3 dq = qni + pi - qdi;
        if (len(dq) > radius) {
            pi = normalize(dq) * radius + qdi - qni;
            p[i] = f4(pi);
        }
        m = len(pi);
        if (!(m >= 0.0f)) m = __builtin_huge_valf();  // a NaN step: no pair is skipped
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax((int*)ccd + 1, __float_as_int(m));
}

// Additive CCD (Li et al. 2021), ccd_kernels.py:8-110; min step by an
// integer atomic on the non-negative float's bits.
__global__ void __launch_bounds__(kB) accd_k(const float4* __restrict__ x, const float4* __restrict__ vel, Topo tp,
                                             Params P, int nC, float* __restrict__ ccd_step) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nC) return;
    // the result is the minimum over pairs and a pair's t only grows, so a
    // pair stops as soon as t reaches the minimum found so far
    const volatile float* vstep = ccd_step;
    if (*vstep <= 0.0f) return;
    int4 k = tp.cidx[b];
    V3 x0 = v3(x[k.x]), x1 = v3(x[k.y]), x2 = v3(x[k.z]), x3 = v3(x[k.w]);
    V3 v0 = v3(vel[k.x]), v1 = v3(vel[k.y]), v2 = v3(vel[k.z]), v3_ = v3(vel[k.w]);
    const bool is_pt = tp.ctype1[b] == BLOCK_PT;
    V3 va = (v0 + v1 + v2 + v3_) / 4.0f;
    v0 = v0 - va; v1 = v1 - va; v2 = v2 - va; v3_ = v3_ - va;
    float lv = is_pt ? len(v0) + fmaxf(len(v1), fmaxf(len(v2), len(v3_)))
                     : fmaxf(len(v0), len(v1)) + fmaxf(len(v2), len(v3_));
    if (lv < 1e-12f) return;
    auto dist = [&]() {
        if (is_pt) {
            int t = pt_classify(x0, x1, x2, x3);
            return pt_distance(x0, x1, x2, x3, t);
        }
        int t = ee_classify(x0, x1, x2, x3, P.ee_thres);
        return ee_distance(x0, x1, x2, x3, t);
    };
    float d = dist();
    float dsq = d * d;
    float dmt = fmaxf((dsq - P.ccd_thick * P.ccd_thick) / (sqrtf(dsq) + P.ccd_thick), 0.0f);
    const float g = (1.0f - P.ccd_slack) * dmt;
    float t = 0.0f;
    float tl = P.ccd_slack * dmt / lv;
    int it = 0;
    while (true) {
        x0 = x0 + tl * v0; x1 = x1 + tl * v1; x2 = x2 + tl * v2; x3 = x3 + tl * v3_;
        d = dist();
        dsq = d * d;
        dmt = fmaxf((dsq - P.ccd_thick * P.ccd_thick) / (sqrtf(dsq) + P.ccd_thick), 0.0f);
        if ((it > 0 && dmt <= g) || it >= P.ccd_max_iters) {
            atomicMin((int*)ccd_step, __float_as_int(t));
            return;
        }
        t = t + tl;
        if (t > 1.0f) return;
        if ((it & 7) == 7 && t >= *vstep) return;
        tl = 0.8f * dmt / lv;
        ++it;
    }
}

// q = q_n + alpha 0.5^n p unless n > 0 and the energy already decreased
// (solver_kernels.py:61-74)
__global__ void line_search_k(const float4* __restrict__ qn, const float4* __restrict__ p,
                              const float* __restrict__ alpha, int nh, const double* __restrict__ sc, int n,
                              float4* __restrict__ q) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (nh > 0 && sc[SC_E] < sc[SC_EPREV]) return;
    float f = alpha[0] * powf(0.5f, (float)nh);
    q[i] = f4(v3(qn[i]) + f * v3(p[i]));
}

// ------------------------------------------------------------ dense quadratic test energy
// E = 0.5 x^T A x + b^T x + c over the first n3 = 3 * nV coordinates
// (DummyEnergyCalculator, energy.py:62-111); A row-major FP64, n3 <= 3072.
__global__ void __launch_bounds__(kB) dense_energy_k(const float4* __restrict__ x, int nV, const double* __restrict__ A,
                                                     const double* __restrict__ b, double c0, double* sc, int slot) {
    __shared__ double red[kB];
    const int n3 = 3 * nV;
    double acc = 0.0;
    for (int r = threadIdx.x; r < n3; r += kB) {
        double ax = 0.0;
        for (int k = 0; k < n3; ++k) ax += A[(long long)r * n3 + k] * (double)((const float*)&x[k / 3])[k % 3];
        const double xr = (double)((const float*)&x[r / 3])[r % 3];
        acc += 0.5 * xr * ax + b[r] * xr;
    }
    double s = block_sum(acc, red);
    if (threadIdx.x == 0) sc[slot] += s + c0;
}

__global__ void __launch_bounds__(kB) dense_diff_k(const float4* __restrict__ x, int nV, const double* __restrict__ A,
                                                   const double* __restrict__ b, float coeff, float4* __restrict__ grad,
                                                   float4* __restrict__ diag) {
    const int n3 = 3 * nV;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n3) return;
    double ax = b[r];
    for (int k = 0; k < n3; ++k) ax += A[(long long)r * n3 + k] * (double)((const float*)&x[k / 3])[k % 3];
    atomicAdd(&((float*)&grad[r / 3])[r % 3], (float)ax * coeff);
    atomicAdd(&((float*)&diag[r / 3])[r % 3], (float)A[(long long)r * n3 + r]);
}


// ------------------------------------------------------------ host orchestration
#define L3(kern, n, ...)                                                                                  \
    do {                                                                                                  \
        hipLaunchKernelGGL(kern, dim3(grid_for((long long)(n), kB)), dim3(kB), 0, c.stream, __VA_ARGS__); \
        PAMO_HIP(hipGetLastError());                                                                      \
    } while (0)

static Topo topo(Ctx& c) {
    return Topo{c.tri, c.area, c.inv_dm, c.hinge, c.rest_angle, c.rest_elen, c.vor, c.target, c.cidx, c.ctype1};
}

// Item counts of the enabled energy terms (items of a disabled term are
// simply not launched).
struct Counts {
    int v, f, h, c;
    long long total() const { return (long long)v + f + h + c; }
};
static Counts counts(const Ctx& c, int nC, int mask) {
    return Counts{(mask & PAMO_S3_E_M2GT) ? c.nV : 0, (mask & PAMO_S3_E_ELASTIC) ? c.nF : 0,
                  (mask & PAMO_S3_E_HINGE) ? c.nH : 0, (mask & PAMO_S3_E_COLLISION) ? nC : 0};
}

// reach: null, or (CCD step, max |p|) at a line-search point (see energy_k)
static int compute_energy(Ctx& c, int nC, int slot, int mask, const float* reach = nullptr) {
    const Counts k = counts(c, nC, mask);
    PAMO_TRY(ensure_part(c, k.total()));
    const unsigned nb = grid_for(k.total(), kB);
    hipLaunchKernelGGL(energy_k, dim3(nb), dim3(kB), 0, c.stream, (const float4*)c.q, topo(c), c.P, k.v, k.f, k.h, k.c,
                       c.cd, c.ctype2, c.part, reach);
    hipLaunchKernelGGL(finish_sum_k, dim3(1), dim3(kB), 0, c.stream, (const double*)c.part, (int)nb, c.sc, slot);
    if ((mask & PAMO_S3_E_DENSE) && c.dense_A)
        hipLaunchKernelGGL(dense_energy_k, dim3(1), dim3(kB), 0, c.stream, (const float4*)c.q, c.nV,
                           (const double*)c.dense_A, (const double*)c.dense_b, c.dense_c, c.sc, slot);
    PAMO_HIP(hipGetLastError());
    return 0;
}

// grad (coefficient -1) and Jacobi diagonal; hinge blocks and contact
// distance gradients are stored for the Hessian-vector products.
static int ensure_sort(Ctx& c, long long n) {
    if (n > c.cap_sort) {
        c.cap_sort = n + n / 4 + 1024;
        PAMO_TRY(dalloc(&c.skey, (size_t)c.cap_sort, c.stream));
        PAMO_TRY(dalloc(&c.sval, (size_t)c.cap_sort, c.stream));
        PAMO_TRY(dalloc(&c.skey2, (size_t)c.cap_sort, c.stream));
        PAMO_TRY(dalloc(&c.sval2, (size_t)c.cap_sort, c.stream));
    }
    return ensure_ws(c, n + 64);
}

static int key_bits(int nV) {
    int b = 8;
    while (b < 32 && (nV - 1) >> b) b += 8;
    return b;
}

// off (nV + 1) / ent: the slots of each vertex in ascending slot order, from
// keys[s] = vertex of slot s (already in skey / sval).
static int finish_incidence(Ctx& c, long long n, int* off, int* ent) {
    PAMO_TRY(radix_sort_pairs(c.skey, c.sval, c.skey2, c.sval2, (int)n, key_bits(c.nV), c.ws, c.stream));
    L3(slot_offsets_k, n + 1, (const uint32_t*)c.skey, (int)n, c.nV, off);
    PAMO_HIP(hipMemcpyAsync(ent, c.sval, sizeof(int) * n, hipMemcpyDeviceToDevice, c.stream));
    return 0;
}

static int build_incidence(Ctx& c, const int* elem, int n_elem, int width, int stride, int* off, int* ent) {
    const long long n = (long long)n_elem * width;
    if (n == 0) return fill(off, 0, c.nV + 1, c.stream);
    PAMO_TRY(ensure_sort(c, n));
    L3(slot_keys_k, n, elem, width, stride, (int)n, c.skey, c.sval);
    return finish_incidence(c, n, off, ent);
}

// _compute_diff (system.py:334-339): grad (coefficient -1) and Jacobi
// diagonal; projected hinge Hessians; the active contacts (d < d_hat) are
// listed and their distance gradients kept, with a vertex -> contact slot
// incidence list for the Hessian-vector products.
static int compute_diff(Ctx& c, int nC, int mask) {
    PAMO_TRY(fill(c.grad, make_float4(0, 0, 0, 0), c.nV, c.stream));
    PAMO_TRY(fill(c.diag, make_float4(0, 0, 0, 0), c.nV, c.stream));
    c.nA = 0;
    if ((mask & PAMO_S3_E_COLLISION) && nC > 0) {
        PAMO_TRY(ensure_ws(c, (long long)nC + 64));
        L3(act_flag_k, nC + 1, (const float*)c.cd, nC, c.P.d_hat, c.act_flag);
        PAMO_TRY(scan_exclusive(c.act_flag, c.act_pos, nC + 1, c.ws, c.stream));
        L3(act_list_k, nC + 1, (const int*)c.act_flag, (const int*)c.act_pos, nC, c.act, c.n_act);
        PAMO_HIP(hipMemcpyAsync(c.hcount, c.n_act, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        PAMO_HIP(hipStreamSynchronize(c.stream));
        c.nA = c.hcount[0];
        c.stat_active += c.nA;
    } else {
        PAMO_TRY(fill(c.n_act, 0, 1, c.stream));
    }
    if (c.nA > 0) {
        // vertex -> active-contact corner lists (list positions c_dst of the
        // corner slots 4 j + a), before diff_k writes the slots
        const long long n = 4LL * c.nA;
        PAMO_TRY(ensure_sort(c, n));
        L3(contact_keys_k, n, (const int4*)c.cidx, (const int*)c.act, (const int*)c.n_act, c.skey, c.sval);
        PAMO_TRY(finish_incidence(c, n, c.vc_off, c.vc_ent));
        L3(invert_k, n, (const int*)c.vc_ent, (int)n, c.c_dst);
    } else {
        PAMO_TRY(fill(c.vc_off, 0, c.nV + 1, c.stream));
    }
    Counts k = counts(c, c.nA, mask);
    DiffSlots ds{c.vf_dst, c.vh_dst, c.c_dst, c.fc, c.hc, c.hdg, c.cslot, c.cdg, (long long)c.capH,
                 c.cvert, c.c2j, (long long)c.capC};
    if (k.total() > 0)
        L3(diff_k, k.total(), (const float4*)c.q, topo(c), c.P, k.v, k.f, k.h, k.c, -1.0f, (const float*)c.cd,
           (const int*)c.ctype2, c.grad, c.diag, c.hblock, c.hsym, (const int*)c.act, (const int*)c.n_act, c.cdd, ds);
    if (c.nA > 0)
        L3(contact_lists_k, 4LL * c.nA, (const int*)c.vc_ent, 4 * c.nA, (const float4*)c.cdd, (long long)c.capC, c.cdv,
           c.cj);
    if (k.f + k.h + k.c > 0) {
        Incid ic{c.vf_off, c.vh_off, c.vc_off};
        L3(diff_gather_k, (long long)c.nV * kGather, c.nV, (int)(k.f > 0), (int)(k.h > 0), (int)(k.c > 0), ic,
           (const float4*)c.fc, (const float4*)c.hc, (const float4*)c.hdg, (const float4*)c.cslot,
           (const float4*)c.cdg, c.grad, c.diag);
    }
    if ((mask & PAMO_S3_E_DENSE) && c.dense_A)
        L3(dense_diff_k, 3 * c.nV, (const float4*)c.q, c.nV, (const double*)c.dense_A, (const double*)c.dense_b, -1.0f,
           c.grad, c.diag);
    return 0;
}

// Av = H dx with dx from cv (see hvp_elem_k); dx is also stored to v_out
// and block partials of dx . Av to vav.
static int compute_hvp(Ctx& c, int mask, const CgVec& cv, float4* v_out, float4* Av, double* vav) {
    const int nF = (mask & PAMO_S3_E_ELASTIC) ? c.nF : 0;
    const int nH = (mask & PAMO_S3_E_HINGE) ? c.nH : 0;
    const int nA = (mask & PAMO_S3_E_COLLISION) ? c.nA : 0;
    const int nd = ((mask & PAMO_S3_E_DENSE) && c.dense_A) ? 3 * c.nV : 0;
    const long long total = (long long)nF + nH + nA + nd;
    if (total > 0)
        L3(hvp_elem_k, total, (const float4*)c.q, topo(c), c.P, nF, nH, nA, (const int*)c.n_act, nd, cv,
           (const int4*)c.cvert, (const float*)c.c2j, (const float*)c.hsym, (long long)c.capH, (const float4*)c.cdd,
           (long long)c.capC,
           (const double*)c.dense_A, (const int*)c.vf_dst, (const int*)c.vh_dst, c.fc, c.hc, c.sA, c.dn);
    Incid ic{c.vf_off, c.vh_off, c.vc_off};
    L3(hvp_gather_k, (long long)c.nV * kGather, c.nV, c.P, (int)((mask & PAMO_S3_E_M2GT) != 0), (int)(nF > 0), (int)(nH > 0),
       (int)(nA > 0), (int)(nd > 0), (const float*)c.vor, cv, ic, (const float4*)c.fc, (const float4*)c.hc,
       (const float4*)c.cdv, (const int*)c.cj, (const float*)c.sA, (const float4*)c.cdd, (const float*)c.dn, v_out,
       Av, vav);
    return 0;
}

static int ensure_contacts(Ctx& c, int n) {
    if (n <= c.capC) return 0;
    int a = n + n / 4 + 1024;
    c.capC = a;
    PAMO_TRY(dalloc(&c.cidx, a, c.stream));
    PAMO_TRY(dalloc(&c.ctype1, a, c.stream));
    PAMO_TRY(dalloc(&c.cidx2, a, c.stream));
    PAMO_TRY(dalloc(&c.ctype1b, a, c.stream));
    PAMO_TRY(dalloc(&c.ctag, a, c.stream));
    PAMO_TRY(dalloc(&c.ctype2, a, c.stream));
    PAMO_TRY(dalloc(&c.cd, a, c.stream));
    PAMO_TRY(dalloc(&c.cdd, 4 * (size_t)a, c.stream));
    PAMO_TRY(dalloc(&c.act, a, c.stream));
    PAMO_TRY(dalloc(&c.act_flag, (size_t)a + 1, c.stream));
    PAMO_TRY(dalloc(&c.act_pos, (size_t)a + 1, c.stream));
    PAMO_TRY(dalloc(&c.cslot, 4 * (size_t)a, c.stream));
    PAMO_TRY(dalloc(&c.c_dst, 4 * (size_t)a, c.stream));
    PAMO_TRY(dalloc(&c.vc_ent, 4 * (size_t)a, c.stream));
    PAMO_TRY(dalloc(&c.cdg, 4 * (size_t)a, c.stream));
    PAMO_TRY(dalloc(&c.cvert, a, c.stream));
    PAMO_TRY(dalloc(&c.cdv, 4 * (size_t)a, c.stream));
    PAMO_TRY(dalloc(&c.cj, 4 * (size_t)a, c.stream));
    PAMO_TRY(dalloc(&c.sA, a, c.stream));
    PAMO_TRY(dalloc(&c.c2j, a, c.stream));
    return 0;
}

// _detect_contact (system.py:304-318, energy.py:896-948): refit the edge and
// triangle trees, gather PT and EE pairs within 2r + d_hat, halve r and
// retry while the count exceeds max_blocks; q_prev_detection <- q.
// Reports a tree walk that exceeded its bound (hcount[1] read from ccount[1]
// at the last synchronisation) and clears the flag.
static int check_walks(Ctx& c) {
    if (c.hcount[1] == 0) return 0;
    c.hcount[1] = 0;
    PAMO_HIP(hipMemsetAsync(c.ccount + 1, 0, sizeof(int), c.stream));
    PAMO_CHECK(false, "stage 3: a tree walk exceeded its bound (corrupted LBVH)");
}

static int detect_contacts(Ctx& c, int& nC) {
    nC = 0;
    if (c.P.mask & PAMO_S3_E_COLLISION) {
        PAMO_TRY(tree_alloc(c.tri_tree, c.nF, c.stream));
        PAMO_TRY(tree_alloc(c.edge_tree, c.nE, c.stream));
        // the hierarchies are built at the first detection on a mesh and refit
        // after that: Newton steps move vertices by at most the clamp radius,
        // and the contact set found does not depend on the tree's shape
        L3(tri_items_k, c.nF, (const float4*)c.q, (const int4*)c.tri, c.nF, 0.0f, c.tri_tree.item);
        L3(edge_items_k, c.nE, (const float4*)c.q, (const int2*)c.edge, c.nE, c.edge_tree.item);
        if (c.contact_trees_built) {
            PAMO_TRY(tree_refit(c.tri_tree, c.nF, c.stream));
            PAMO_TRY(tree_refit(c.edge_tree, c.nE, c.stream));
        } else {
            PAMO_TRY(tree_build(c.tri_tree, c.nF, c.ws, c.stream));
            PAMO_TRY(tree_build(c.edge_tree, c.nE, c.ws, c.stream));
            c.contact_trees_built = true;
        }
        while (true) {
            PAMO_TRY(fill(c.ccount, 0, 1, c.stream));
            const int cap = std::min(c.capC, c.P.max_blocks);
            // launches of whole blocks: waves past the queries count 0
            const long long pt_blocks = grid_for(((long long)c.nV + kDetW - 1) / kDetW * kWave, kB);
            const long long ee_blocks = grid_for(((long long)c.nE + kDetW - 1) / kDetW * kWave, kB);
            const int pt_waves = (int)(pt_blocks * (kB / kWave)), n_waves = (int)((pt_blocks + ee_blocks) * (kB / kWave));
            if (n_waves + 1 > c.cap_wave) {
                c.cap_wave = n_waves + 1 + n_waves / 4;
                PAMO_TRY(dalloc(&c.wave_count, (size_t)c.cap_wave, c.stream));
                PAMO_TRY(dalloc(&c.wave_base, (size_t)c.cap_wave, c.stream));
            }
            PAMO_TRY(fill(c.wave_count, 0, n_waves + 1, c.stream));
            DetOut o{c.ccount, cap, c.cidx, c.ctype1, c.ctag, c.wave_count, 0};
            L3(detect_k<true>, pt_blocks * kB, (const float4*)c.q, c.nV, (const int4*)c.tri, (const int2*)c.edge,
               (const float4*)c.tri_tree.box, (const int*)c.tri_tree.right, c.nF, c.P.d_hat, c.radius, c.P.ee_thres,
               o, c.ccount + 1);
            o.wave0 = pt_waves;
            L3(detect_k<false>, ee_blocks * kB, (const float4*)c.q, c.nE, (const int4*)c.tri, (const int2*)c.edge,
               (const float4*)c.edge_tree.box, (const int*)c.edge_tree.right, c.nE, c.P.d_hat, c.radius,
               c.P.ee_thres, o, c.ccount + 1);
            PAMO_HIP(hipMemcpyAsync(c.hcount, c.ccount, 2 * sizeof(int), hipMemcpyDeviceToHost, c.stream));
            PAMO_HIP(hipStreamSynchronize(c.stream));
            PAMO_TRY(check_walks(c));
            const int n = c.hcount[0];
            if (n > c.P.max_blocks) {
                PAMO_CHECK(c.radius > 1e-30f, "contact detection: more than max_blocks pairs within d_hat");
                c.radius /= 2.0f;
                continue;
            }
            if (n > c.capC) {
                PAMO_TRY(ensure_contacts(c, n));
                continue;
            }
            nC = n;
            // contacts in (wave, emission rank) order
            if (nC > 0) {
                PAMO_TRY(ensure_ws(c, (long long)n_waves + 64));
                PAMO_TRY(scan_exclusive(c.wave_count, c.wave_base, n_waves + 1LL, c.ws, c.stream));
                L3(contact_order_k, nC, nC, (const int2*)c.ctag, (const int*)c.wave_base, (const int4*)c.cidx,
                   (const int*)c.ctype1, c.cidx2, c.ctype1b);
                std::swap(c.cidx, c.cidx2);
                std::swap(c.ctype1, c.ctype1b);
            }
            break;
        }
    }
    PAMO_HIP(hipMemcpyAsync(c.qd, c.q, sizeof(float4) * c.nV, hipMemcpyDeviceToDevice, c.stream));
    c.stat_contacts += nC;
    return 0;
}

static int update_target(Ctx& c) {
    if (c.P.mask & PAMO_S3_E_M2GT)
        L3(closest_point_k, c.nV, (const float4*)c.q, c.nV, view(c.gt_tree), (const float4*)c.gt_v,
           (const int4*)c.gt_f, c.target, c.ccount + 1);
    return 0;
}

// CGSolver.solve (cg_solver.py:96-158): Jacobi PCG with n_cg fixed
// iterations, p = H^-1 grad; three launches per iteration.
static int cg_record(Ctx& c) {
    const int nV = c.nV, mask = c.P.mask;
    const int np = (int)grid_for(nV, kB);
    L3(cg_init_k, nV, (const float4*)c.grad, (const float4*)c.diag, nV, c.r, c.z, c.p, c.zr_part);
    for (int k = 0; k < c.P.n_cg; ++k) {
        ++c.stat_cg;
        const int cur = k & 1;
        double* zr_cur = c.zr_part + (size_t)cur * np;
        double* zr_oth = c.zr_part + (size_t)(cur ^ 1) * np;
        CgVec cv{c.z, c.v, zr_cur, zr_oth, np, k == 0 ? 1 : 0};
        PAMO_TRY(compute_hvp(c, mask, cv, c.v, c.Av, c.vav_part));
        L3(cg_update_k, nV, (const float4*)c.v, (const float4*)c.Av, (const float4*)c.diag, (const double*)c.vav_part,
           (int)grid_for((long long)nV * kGather, kB), (const double*)zr_cur, np, nV, c.p, c.r, c.z, zr_oth);
    }
    return 0;
}

// The reference captures its PCG loop in a CUDA graph (cg_solver.py:150-156,
// config.py:21).  Here the loop's 3 n_cg + 1 kernels are captured from the
// stream each solve and the executable graph is updated in place (the
// active-contact count, hence one grid size, changes between Newton
// iterations), then launched as one unit.  No memset or copy is captured.
static int cg_solve(Ctx& c) {
    if (!c.cg_graph) return cg_record(c);
    hipGraph_t g = nullptr;
    PAMO_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    const int rc = cg_record(c);
    const hipError_t ec = hipStreamEndCapture(c.stream, &g);
    if (rc != 0) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    PAMO_HIP(ec);
    if (c.cg_exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
        if (hipGraphExecUpdate(c.cg_exec, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();  // not updatable (shape changed): instantiate anew
            (void)hipGraphExecDestroy(c.cg_exec);
            c.cg_exec = nullptr;
        }
    }
    hipError_t ei = hipSuccess;
    if (!c.cg_exec) ei = hipGraphInstantiate(&c.cg_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PAMO_HIP(ei);
    PAMO_HIP(hipGraphLaunch(c.cg_exec, c.stream));
    return 0;
}

// _clamp_p + _ccd (system.py:354-391)
static int clamp_and_ccd(Ctx& c, int nC) {
    PAMO_TRY(fill(c.ccd_step + 1, 0.0f, 1, c.stream));
    L3(clamp_k, c.nV, c.p, (const float4*)c.qn, (const float4*)c.qd, c.nV, c.P.radius0, c.ccd_step);
    PAMO_TRY(fill(c.ccd_step, 1.0f, 1, c.stream));
    if (nC > 0) L3(accd_k, nC, (const float4*)c.q, (const float4*)c.p, topo(c), c.P, nC, c.ccd_step);
    return 0;
}

static int check_ready(const Ctx& c, const char* who) {
    PAMO_CHECK(c.nV > 0, std::string(who) + ": no mesh registered");
    PAMO_CHECK(!(c.P.mask & PAMO_S3_E_M2GT) || c.gt_tree.n > 0, std::string(who) + ": no target surface registered");
    PAMO_CHECK(!(c.P.mask & (PAMO_S3_E_ELASTIC | PAMO_S3_E_COLLISION)) || c.nF > 0,
               std::string(who) + ": the enabled energies need faces");
    PAMO_CHECK(!(c.P.mask & PAMO_S3_E_COLLISION) || c.nE > 0, std::string(who) + ": contacts need edges");
    PAMO_CHECK(!(c.P.mask & PAMO_S3_E_DENSE) || c.dense_A, std::string(who) + ": dense energy not set");
    return 0;
}

// One step(): target update, then n_newton Newton iterations (system.py:411-470).
// on_newton (may be null) is called after each Newton iteration, once its
// device work has finished.
static int step_once(Ctx& c, pamo_s3_newton_fn on_newton = nullptr, void* user = nullptr) {
    if (c.do_nothing) return 0;
    const int nV = c.nV, mask = c.P.mask;
    PAMO_TRY(update_target(c));
    for (int it = 0; it < c.P.n_newton; ++it) {
        ++c.stat_newton;
        PAMO_HIP(hipMemcpyAsync(c.qn, c.q, sizeof(float4) * nV, hipMemcpyDeviceToDevice, c.stream));
        int nC = 0;
        PAMO_TRY(detect_contacts(c, nC));
        PAMO_TRY(compute_energy(c, nC, SC_E, mask));
        PAMO_TRY(compute_diff(c, nC, mask));
        PAMO_TRY(cg_solve(c));
        PAMO_TRY(clamp_and_ccd(c, nC));
        // _line_search (system.py:393-409)
        hipLaunchKernelGGL(copy_slot_k, dim3(1), dim3(1), 0, c.stream, c.sc, (int)SC_E, (int)SC_EPREV);
        for (int h = 0; h < c.P.n_ls; ++h) {
            ++c.stat_ls;
            L3(line_search_k, nV, (const float4*)c.qn, (const float4*)c.p, (const float*)c.ccd_step, h,
               (const double*)c.sc, nV, c.q);
            PAMO_TRY(compute_energy(c, nC, SC_E, mask, c.ccd_step));
        }
        if (on_newton) {
            PAMO_HIP(hipStreamSynchronize(c.stream));
            on_newton(it, user);
        }
    }
    PAMO_HIP(hipMemcpyAsync(c.hsc, c.sc, sizeof(double) * SC_N, hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.hcount, c.ccount, 2 * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    PAMO_TRY(check_walks(c));
    c.last_energy = c.hsc[SC_E];
    return 0;
}

// ------------------------------------------------------------ host-side mesh preparation
// The connectivity work of register_mesh (system.py:175-228) that the
// reference does in trimesh / igl: unique edges in trimesh's order (by
// (max, min)), hinges in (face, corner) order, Voronoi areas (Meyer et al.
// 2003 mixed areas, igl MASSMATRIX_TYPE_VORONOI) and the smallest interior
// angle, all in FP64 on the normalised positions.
struct HostPrep {
    std::vector<int> edges, hinges;
    std::vector<float> vor;
    double min_angle_deg = 180.0;
};

static void host_prepare(const float* V, int nV, const int* F, int nF, double scale, const double* t, HostPrep& hp) {
    std::vector<double> X(3 * (size_t)nV);
    for (size_t i = 0; i < 3 * (size_t)nV; ++i) X[i] = (double)V[i] * scale + t[i % 3];
    // unique edges
    std::vector<uint64_t> keys(3 * (size_t)nF);
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) {
            uint32_t a = (uint32_t)F[3 * f + k], b = (uint32_t)F[3 * f + (k + 1) % 3];
            uint32_t lo = std::min(a, b), hi = std::max(a, b);
            keys[3 * (size_t)f + k] = ((uint64_t)hi << 32) | lo;
        }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    hp.edges.resize(2 * keys.size());
    for (size_t e = 0; e < keys.size(); ++e) {
        hp.edges[2 * e] = (int)(keys[e] & 0xFFFFFFFFu);
        hp.edges[2 * e + 1] = (int)(keys[e] >> 32);
    }
    // hinges: directed edge (a, b) of slot s = corner * nF + face; the partner
    // is the smallest slot holding (b, a)
    const size_t nS = 3 * (size_t)nF;
    std::vector<std::pair<uint64_t, uint32_t>> dir(nS);
    for (int k = 0; k < 3; ++k)
        for (int f = 0; f < nF; ++f) {
            const uint32_t a = (uint32_t)F[3 * f + k], b = (uint32_t)F[3 * f + (k + 1) % 3];
            const size_t sl = (size_t)k * nF + f;
            dir[sl] = {((uint64_t)a << 32) | b, (uint32_t)sl};
        }
    std::vector<std::pair<uint64_t, uint32_t>> sorted = dir;
    std::sort(sorted.begin(), sorted.end());
    hp.hinges.clear();
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) {
            const int a = F[3 * f + k], b = F[3 * f + (k + 1) % 3], o = F[3 * f + (k + 2) % 3];
            if (!(a < b)) continue;
            const uint64_t rk = ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
            auto it = std::lower_bound(sorted.begin(), sorted.end(), std::make_pair(rk, (uint32_t)0));
            if (it == sorted.end() || it->first != rk) continue;
            const uint32_t sl = it->second;
            const int f2 = (int)(sl % (uint32_t)nF), k2 = (int)(sl / (uint32_t)nF);
            hp.hinges.push_back(o);
            hp.hinges.push_back(a);
            hp.hinges.push_back(b);
            hp.hinges.push_back(F[3 * f2 + (k2 + 2) % 3]);
        }
    // Voronoi areas and the smallest interior angle
    std::vector<double> vor(nV, 0.0);
    double min_ang = 1e300;
    bool any_nan = false;  // numpy's min propagates NaN (a zero-length edge), and NaN <= thres is false
    for (int f = 0; f < nF; ++f) {
        const double* P[3] = {&X[3 * (size_t)F[3 * f]], &X[3 * (size_t)F[3 * f + 1]], &X[3 * (size_t)F[3 * f + 2]]};
        auto d3 = [](const double* a, const double* b, double* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; };
        auto nrm = [](const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); };
        double e12[3], e20[3], e01[3];
        d3(P[1], P[2], e12);
        d3(P[2], P[0], e20);
        d3(P[0], P[1], e01);
        const double l[3] = {nrm(e12), nrm(e20), nrm(e01)};  // opposite corner k
        const double l2[3] = {l[0] * l[0], l[1] * l[1], l[2] * l[2]};
        const double cs[3] = {(l2[1] + l2[2] - l2[0]) / (2 * l[1] * l[2]), (l2[2] + l2[0] - l2[1]) / (2 * l[2] * l[0]),
                              (l2[0] + l2[1] - l2[2]) / (2 * l[0] * l[1])};
        double u[3], w[3], c[3];
        d3(P[1], P[0], u);
        d3(P[2], P[0], w);
        c[0] = u[1] * w[2] - u[2] * w[1];
        c[1] = u[2] * w[0] - u[0] * w[2];
        c[2] = u[0] * w[1] - u[1] * w[0];
        const double area = 0.5 * nrm(c);
        double bary[3] = {cs[0] * l[0], cs[1] * l[1], cs[2] * l[2]};
        const double bs = bary[0] + bary[1] + bary[2];
        for (int k = 0; k < 3; ++k) bary[k] /= bs;
        const double part[3] = {bary[0] * area, bary[1] * area, bary[2] * area};
        double quad[3] = {(part[1] + part[2]) * 0.5, (part[2] + part[0]) * 0.5, (part[0] + part[1]) * 0.5};
        for (int k = 0; k < 3; ++k)
            if (cs[k] < 0) {
                quad[k] = 0.5 * area;
                quad[(k + 1) % 3] = 0.25 * area;
                quad[(k + 2) % 3] = 0.25 * area;
            }
        for (int k = 0; k < 3; ++k) vor[F[3 * f + k]] += quad[k];
        // interior angles (igl.internal_angles)
        for (int k = 0; k < 3; ++k) {
            double a[3], b[3];
            d3(P[(k + 1) % 3], P[k], a);
            d3(P[(k + 2) % 3], P[k], b);
            const double cc = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / (nrm(a) * nrm(b));
            if (!(cc == cc)) {
                any_nan = true;
                continue;
            }
            const double ang = std::acos(std::min(1.0, std::max(-1.0, cc)));
            min_ang = std::min(min_ang, ang);
        }
    }
    hp.vor.resize(nV);
    for (int i = 0; i < nV; ++i) hp.vor[i] = (float)vor[i];
    hp.min_angle_deg = nF == 0 ? 180.0 : (any_nan ? NAN : min_ang * 180.0 / M_PI);
}

// numpy's V.mean(axis=0) of a float64 (n, 3) array adds the rows one after
// another (from the first row) in FP64; the same sequential sum of the
// float32 vertices widened to FP64, in one workgroup: waves 1..15 copy the
// next chunk of coordinates to LDS while lanes 0..2 of wave 0 add up the
// current chunk, one coordinate per lane, 64 values per LDS batch.
constexpr int kSeqThreads = 1024;
constexpr int kSeqChunk = 2048;            // multiple of 32
constexpr int kSeqPitch = kSeqChunk + 4;

__device__ inline void seq_fill(const float* __restrict__ v, long long base, int cnt, float* __restrict__ dst, int t,
                                int nt) {
    for (int j = t; j < 3 * cnt; j += nt) {
        const int r = j / 3, d = j - 3 * r;  // coalesced reads of the interleaved rows
        dst[d * kSeqPitch + r] = v[3 * base + j];
    }
}

// acc + row[0] + ... + row[n - 1] in FP64, left to right (row 16-byte
// aligned), 64 values per LDS batch as in centroid_chain
__device__ __forceinline__ double seq_chain(double acc, const float* __restrict__ row, int n) {
    const float4* r4 = reinterpret_cast<const float4*>(row);
    const int full = n >> 6;
    for (int b = 0; b < full; ++b) {
        float4 v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = r4[16 * b + j];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            acc = acc + (double)v[j].x;
            acc = acc + (double)v[j].y;
            acc = acc + (double)v[j].z;
            acc = acc + (double)v[j].w;
        }
    }
    for (int i = full << 6; i < n; ++i) acc = acc + (double)row[i];
    return acc;
}

__global__ void __launch_bounds__(kSeqThreads) seq_sum3_k(const float* __restrict__ v, int n,
                                                         double* __restrict__ out3) {
    __shared__ __align__(16) float buf[2][3 * kSeqPitch];
    const int t = threadIdx.x;
    const int nchunks = (n + kSeqChunk - 1) / kSeqChunk;
    seq_fill(v, 0, n < kSeqChunk ? n : kSeqChunk, buf[0], t, kSeqThreads);
    __syncthreads();
    double acc = 0.0;
    for (int k = 0; k < nchunks; ++k) {
        const long long base = (long long)k * kSeqChunk;
        const int cnt = (int)((n - base) < kSeqChunk ? (n - base) : kSeqChunk);
        if (t >= kWave) {
            const long long nb = base + kSeqChunk;
            if (nb < n)
                seq_fill(v, nb, (int)((n - nb) < kSeqChunk ? (n - nb) : kSeqChunk), buf[(k + 1) & 1], t - kWave,
                         kSeqThreads - kWave);
        } else if (t < 3) {
            const float* row = buf[k & 1] + t * kSeqPitch;
            if (k == 0) {
                // from the first row, then one at a time up to 32 (aligned), then in batches
                acc = (double)row[0];
                const int head = cnt < 32 ? cnt : 32;
                for (int i = 1; i < head; ++i) acc = acc + (double)row[i];
                acc = seq_chain(acc, row + head, cnt - head);
            } else {
                acc = seq_chain(acc, row, cnt);
            }
        }
        __syncthreads();
    }
    if (t < 3) out3[t] = acc;
}

// normalization transform (processing.py:15-20): bounds reduced on the device
__global__ void __launch_bounds__(kB) gt_stats_k(const float* __restrict__ v, int n, double* __restrict__ part) {
    __shared__ double red[kB];
    __shared__ float rlo[6][kBlock];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sx = 0, sy = 0, sz = 0;
    const float inf = __builtin_huge_valf();
    float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    if (i < n) {
        const float x = v[3 * (long long)i], y = v[3 * (long long)i + 1], z = v[3 * (long long)i + 2];
        sx = x; sy = y; sz = z;
        lo[0] = hi[0] = x; lo[1] = hi[1] = y; lo[2] = hi[2] = z;
    }
    const double a = block_sum(sx, red), b = block_sum(sy, red), c = block_sum(sz, red);
    block_minmax6(lo, hi, rlo);
    if (threadIdx.x == 0) {
        double* o = part + 9 * (long long)blockIdx.x;
        o[0] = a; o[1] = b; o[2] = c;
        for (int k = 0; k < 6; ++k) o[3 + k] = rlo[k][0];
    }
}

// ------------------------------------------------------------ handoff vertex merge
// trimesh 4.4.0 builds the stage-3 input with Trimesh(vertices, faces) and its
// default process=True (pamo/__init__.py:194), which merges vertices
// (trimesh.grouping.merge_vertices): each vertex is keyed by its float64
// coordinates times 10^8 (tol.merge = 1e-8) rounded half to even to int64;
// among the vertices a face references (all of them when there are no
// faces), vertices with equal keys become the first of them in index order
// (unique_rows(..., keep_order=True)); the kept vertices stay in index
// order, with their own coordinates, and faces are re-indexed.
//
// The groups come from an open-addressing table of vertex indices: a vertex
// claims the first slot on its probe path that is empty or holds its key and
// lowers it to its own index; slots never change key once claimed, so all
// vertices of one key meet in one slot whatever the order of the inserts, and
// the slot ends with the smallest index of the group.

__device__ __forceinline__ bool merge_key(const float* __restrict__ v, int i, long long k[3]) {
    bool finite = true;
    for (int c = 0; c < 3; ++c) {
        const float x = v[3 * (long long)i + c];
        finite = finite && isfinite(x);
        k[c] = finite ? (long long)rint((double)x * 1e8) : 0;
    }
    return finite;
}

__device__ __forceinline__ uint32_t merge_hash(const long long k[3]) {
    uint64_t h = 0x243F6A8885A308D3ull;
    for (int c = 0; c < 3; ++c) {
        h = (h ^ (uint64_t)k[c]) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
    }
    h *= 0xBF58476D1CE4E5B9ull;
    return (uint32_t)(h >> 32);
}

// err bit 0: a face index outside [0, n_v); bit 1: a referenced vertex is not
// finite; bit 2: a probe path exceeded the table (cannot happen at load 1/2)
__global__ void merge_mark_k(const int* __restrict__ f, long long n3, int n_v, int* __restrict__ ref,
                             int* __restrict__ err) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const int x = f[i];
    if (x < 0 || x >= n_v)
        atomicOr(err, 1);
    else
        ref[x] = 1;
}

__global__ void merge_insert_k(const float* __restrict__ v, int n_v, const int* __restrict__ ref, int* tab,
                               uint32_t mask, int* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_v || !ref[i]) return;
    long long k[3];
    if (!merge_key(v, i, k)) {
        atomicOr(err, 2);
        return;
    }
    uint32_t h = merge_hash(k) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const int cur = atomicCAS(&tab[h], -1, i);
        if (cur == -1) return;
        long long q[3];
        merge_key(v, cur, q);
        if (q[0] == k[0] && q[1] == k[1] && q[2] == k[2]) {
            atomicMin(&tab[h], i);
            return;
        }
        h = (h + 1) & mask;
    }
    atomicOr(err, 4);
}

// first[i]: the group's first vertex (-1 if i is not kept in any group);
// keep[i] = 1 for the first vertex of each group; keep[n_v] = 0 for the scan
__global__ void merge_first_k(const float* __restrict__ v, int n_v, const int* __restrict__ ref,
                              const int* __restrict__ tab, uint32_t mask, int* __restrict__ first,
                              int* __restrict__ keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n_v) return;
    if (i == n_v) {
        keep[i] = 0;
        return;
    }
    first[i] = -1;
    keep[i] = 0;
    long long k[3];
    if (!ref[i] || !merge_key(v, i, k)) return;
    uint32_t h = merge_hash(k) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const int cur = tab[h];
        if (cur < 0) return;
        long long q[3];
        merge_key(v, cur, q);
        if (q[0] == k[0] && q[1] == k[1] && q[2] == k[2]) {
            first[i] = cur;
            keep[i] = cur == i;
            return;
        }
        h = (h + 1) & mask;
    }
}

__global__ void merge_out_k(const float* __restrict__ v, int n_v, const int* __restrict__ keep,
                            const int* __restrict__ pos, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_v || !keep[i]) return;
    const long long o = 3LL * pos[i], s = 3LL * i;
    out[o] = v[s];
    out[o + 1] = v[s + 1];
    out[o + 2] = v[s + 2];
}

// (indices the mark pass flagged, and vertices without a group, which the
// error bits report, are left as they are)
__global__ void merge_faces_k(int* __restrict__ f, long long n3, int n_v, const int* __restrict__ first,
                              const int* __restrict__ pos) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const int x = f[i];
    if (x < 0 || x >= n_v) return;
    const int g = first[x];
    if (g >= 0) f[i] = pos[g];
}

}  // namespace s3
}  // namespace pamo

// ================================================================ C ABI

using pamo::s3::Ctx;

extern "C" {

int pamo_s3_create(pamo_s3** out, int device, const pamo_s3_params* prm) {
    pamo::clear_error();
    PAMO_CHECK(out && prm, "pamo_s3_create: null argument");
    Ctx* c = new Ctx();
    c->device = device;
    pamo::s3::Params& P = c->P;
    P.n_newton = prm->n_newton_iters;
    P.n_cg = prm->n_cg_iters;
    P.n_ls = prm->n_line_search_iters;
    P.max_blocks = prm->max_blocks;
    P.ccd_max_iters = prm->ccd_max_iters;
    P.mask = prm->energy_mask;
    P.d_hat = (float)prm->d_hat;
    P.kappa = (float)prm->coll_stiffness;
    P.radius0 = (float)prm->contact_detection_radius;
    P.ee_thres = (float)prm->ee_classify_thres;
    P.ccd_slack = (float)prm->ccd_slackness;
    P.ccd_thick = (float)prm->ccd_thickness;
    const double E = prm->elas_young_modulus, nu = prm->elas_poisson_ratio;
    P.mu = (float)(E / (2.0 * (1.0 + nu)));
    P.la = (float)(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
    P.k_hinge = (float)prm->hinge_stiffness;
    P.k_m2gt = (float)prm->mesh2gt_dist_stiffness;
    c->min_angle_thres = prm->min_angle_thres;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // ccount[0]: contacts found; ccount[1]: a tree walk exceeded its bound
    if (e == hipSuccess) e = hipMalloc((void**)&c->ccount, 2 * sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(c->ccount, 0, 2 * sizeof(int), c->stream);
    if (e == hipSuccess) e = hipMalloc((void**)&c->n_act, sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(c->n_act, 0, sizeof(int), c->stream);
    if (e == hipSuccess) e = hipMalloc((void**)&c->sc, sizeof(double) * pamo::s3::SC_N);
    if (e == hipSuccess) e = hipMemsetAsync(c->sc, 0, sizeof(double) * pamo::s3::SC_N, c->stream);
    // ccd_step[0]: the CCD step; [1]: the largest |p| after the clamp
    if (e == hipSuccess) e = hipMalloc((void**)&c->ccd_step, 2 * sizeof(float));
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->hcount, 2 * sizeof(int), 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->hsc, sizeof(double) * pamo::s3::SC_N, 0);
    if (e != hipSuccess) {
        pamo::set_error(std::string("pamo_s3_create: ") + hipGetErrorString(e));
        pamo::s3::ctx_free(*c);
        delete c;
        return 1;
    }
    *out = reinterpret_cast<pamo_s3*>(c);
    return 0;
}

int pamo_s3_destroy(pamo_s3* s) {
    if (!s) return 0;
    Ctx* c = reinterpret_cast<Ctx*>(s);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pamo::s3::ctx_free(*c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int pamo_s3_set_target(pamo_s3* s, const float* d_gt_v, int n_gt_v, const int* d_gt_f, int n_gt_f, double scale,
                       const double* t3) {
    pamo::clear_error();
    PAMO_CHECK(s && d_gt_v && d_gt_f && t3, "pamo_s3_set_target: null argument");
    PAMO_CHECK(n_gt_v > 0 && n_gt_f > 0, "pamo_s3_set_target: empty mesh");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    if (n_gt_v > c.capG || n_gt_f > c.capG) {
        c.capG = std::max(n_gt_v, n_gt_f) + 1024;
        PAMO_TRY(pamo::s3::dalloc(&c.gt_v, c.capG, c.stream));
        PAMO_TRY(pamo::s3::dalloc(&c.gt_f, c.capG, c.stream));
    }
    c.nGV = n_gt_v;
    c.nGF = n_gt_f;
    PAMO_TRY(pamo::s3::ensure_ws(c, n_gt_f + 64));
    hipLaunchKernelGGL(pamo::s3::normalize_points_k, dim3(pamo::grid_for(n_gt_v, 256)), dim3(256), 0, c.stream,
                       d_gt_v, n_gt_v, scale, t3[0], t3[1], t3[2], c.gt_v, (float4*)nullptr);
    hipLaunchKernelGGL(pamo::s3::load_faces4_k, dim3(pamo::grid_for(n_gt_f, 256)), dim3(256), 0, c.stream, d_gt_f,
                       n_gt_f, c.gt_f);
    PAMO_HIP(hipGetLastError());
    PAMO_TRY(pamo::s3::tree_alloc(c.gt_tree, n_gt_f, c.stream));
    hipLaunchKernelGGL(pamo::s3::tri_items_k, dim3(pamo::grid_for(n_gt_f, 256)), dim3(256), 0, c.stream,
                       (const float4*)c.gt_v, (const int4*)c.gt_f, n_gt_f, 0.0f, c.gt_tree.item);
    PAMO_HIP(hipGetLastError());
    PAMO_TRY(pamo::s3::tree_build(c.gt_tree, n_gt_f, c.ws, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

int pamo_s3_set_mesh(pamo_s3* s, const float* d_v, int n_v, const int* d_f, int n_f, double scale, const double* t3,
                     const int* h_edges, int n_e, const int* h_hinges, int n_h, const float* h_voronoi,
                     int do_nothing) {
    pamo::clear_error();
    PAMO_CHECK(s && d_v && t3 && h_voronoi, "pamo_s3_set_mesh: null argument");
    PAMO_CHECK(n_v > 0 && n_f >= 0 && n_e >= 0 && n_h >= 0, "pamo_s3_set_mesh: bad sizes");
    PAMO_CHECK((n_f == 0 || d_f) && (n_e == 0 || h_edges) && (n_h == 0 || h_hinges),
               "pamo_s3_set_mesh: null topology array");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    using pamo::s3::dalloc;
    if (n_v > c.capV) {
        c.capV = n_v + n_v / 4 + 64;
        float4** vv[] = {&c.q, &c.qn, &c.qd, &c.qrest, &c.p, &c.target, &c.grad, &c.diag, &c.r, &c.v, &c.Av, &c.z};
        for (float4** q : vv) PAMO_TRY(dalloc(q, c.capV, c.stream));
        PAMO_TRY(dalloc(&c.vor, c.capV, c.stream));
        PAMO_TRY(dalloc(&c.vf_off, (size_t)c.capV + 1, c.stream));
        PAMO_TRY(dalloc(&c.vh_off, (size_t)c.capV + 1, c.stream));
        PAMO_TRY(dalloc(&c.vc_off, (size_t)c.capV + 1, c.stream));
        PAMO_TRY(dalloc(&c.dn, 3 * (size_t)c.capV, c.stream));
        const size_t np = pamo::grid_for(c.capV, pamo::s3::kB) + 2;
        PAMO_TRY(dalloc(&c.zr_part, 2 * np, c.stream));
        PAMO_TRY(dalloc(&c.vav_part, pamo::grid_for((long long)c.capV * pamo::s3::kGather, pamo::s3::kB) + 2, c.stream));
    }
    if (n_f > c.capF) {
        c.capF = n_f + n_f / 4 + 64;
        PAMO_TRY(dalloc(&c.tri, c.capF, c.stream));
        PAMO_TRY(dalloc(&c.area, c.capF, c.stream));
        PAMO_TRY(dalloc(&c.inv_dm, c.capF, c.stream));
        PAMO_TRY(dalloc(&c.vf_ent, 3 * (size_t)c.capF, c.stream));
        PAMO_TRY(dalloc(&c.vf_dst, 3 * (size_t)c.capF, c.stream));
        PAMO_TRY(dalloc(&c.fc, 3 * (size_t)c.capF, c.stream));
    }
    if (n_e > c.capE) {
        c.capE = n_e + n_e / 4 + 64;
        PAMO_TRY(dalloc(&c.edge, c.capE, c.stream));
    }
    if (n_h > c.capH) {
        c.capH = n_h + n_h / 4 + 64;
        PAMO_TRY(dalloc(&c.hinge, c.capH, c.stream));
        PAMO_TRY(dalloc(&c.rest_angle, c.capH, c.stream));
        PAMO_TRY(dalloc(&c.rest_elen, c.capH, c.stream));
        PAMO_TRY(dalloc(&c.hblock, 144 * (size_t)c.capH, c.stream));
        PAMO_TRY(dalloc(&c.hsym, 45 * (size_t)c.capH, c.stream));
        PAMO_TRY(dalloc(&c.vh_ent, 4 * (size_t)c.capH, c.stream));
        PAMO_TRY(dalloc(&c.vh_dst, 4 * (size_t)c.capH, c.stream));
        PAMO_TRY(dalloc(&c.hc, 4 * (size_t)c.capH, c.stream));
        PAMO_TRY(dalloc(&c.hdg, 4 * (size_t)c.capH, c.stream));
    }
    c.nV = n_v;
    c.nF = n_f;
    c.nE = n_e;
    c.nH = n_h;
    c.contact_trees_built = false;
    c.do_nothing = do_nothing != 0;
    c.radius = c.P.radius0;
    c.stat_contacts = c.stat_newton = c.stat_cg = c.stat_ls = c.stat_active = 0;
    PAMO_TRY(pamo::s3::ensure_ws(c, std::max(n_f, n_e) + 64));
    PAMO_TRY(pamo::s3::ensure_part(c, (long long)n_v + n_f + n_h + 1024));
    PAMO_TRY(pamo::s3::ensure_contacts(c, 4 * n_v + 1024));
    hipLaunchKernelGGL(pamo::s3::normalize_points_k, dim3(pamo::grid_for(n_v, 256)), dim3(256), 0, c.stream, d_v,
                       n_v, scale, t3[0], t3[1], t3[2], c.q, c.qrest);
    hipLaunchKernelGGL(pamo::s3::load_faces4_k, dim3(pamo::grid_for(n_f, 256)), dim3(256), 0, c.stream, d_f, n_f,
                       c.tri);
    PAMO_HIP(hipGetLastError());
    if (n_e > 0)
        PAMO_HIP(hipMemcpyAsync(c.edge, h_edges, sizeof(int) * 2 * (size_t)n_e, hipMemcpyHostToDevice, c.stream));
    if (n_h > 0)
        PAMO_HIP(hipMemcpyAsync(c.hinge, h_hinges, sizeof(int) * 4 * (size_t)n_h, hipMemcpyHostToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.vor, h_voronoi, sizeof(float) * (size_t)n_v, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(pamo::s3::elastic_pre_k, dim3(pamo::grid_for(n_f, 256)), dim3(256), 0, c.stream,
                       (const float4*)c.qrest, (const int4*)c.tri, n_f, c.area, c.inv_dm);
    if (n_h > 0)
        hipLaunchKernelGGL(pamo::s3::hinge_rest_k, dim3(pamo::grid_for(n_h, 256)), dim3(256), 0, c.stream,
                           (const float4*)c.qrest, (const int4*)c.hinge, n_h, c.rest_angle, c.rest_elen);
    PAMO_HIP(hipGetLastError());
    // vertex -> face corner / hinge slot incidence, fixed for this mesh
    PAMO_TRY(pamo::s3::build_incidence(c, (const int*)c.tri, n_f, 3, 4, c.vf_off, c.vf_ent));
    PAMO_TRY(pamo::s3::build_incidence(c, (const int*)c.hinge, n_h, 4, 4, c.vh_off, c.vh_ent));
    if (n_f > 0)
        hipLaunchKernelGGL(pamo::s3::invert_k, dim3(pamo::grid_for(3LL * n_f, 256)), dim3(256), 0, c.stream,
                           (const int*)c.vf_ent, 3 * n_f, c.vf_dst);
    if (n_h > 0)
        hipLaunchKernelGGL(pamo::s3::invert_k, dim3(pamo::grid_for(4LL * n_h, 256)), dim3(256), 0, c.stream,
                           (const int*)c.vh_ent, 4 * n_h, c.vh_dst);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

int pamo_s3_step(pamo_s3* s, int n_steps) {
    pamo::clear_error();
    PAMO_CHECK(s != nullptr, "pamo_s3_step: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_TRY(pamo::s3::check_ready(c, "pamo_s3_step"));
    PAMO_HIP(hipSetDevice(c.device));
    for (int k = 0; k < n_steps; ++k) PAMO_TRY(pamo::s3::step_once(c));
    return 0;
}

int pamo_s3_step_eval(pamo_s3* s, pamo_s3_newton_fn on_newton, void* user) {
    pamo::clear_error();
    PAMO_CHECK(s != nullptr && on_newton != nullptr, "pamo_s3_step_eval: null argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_TRY(pamo::s3::check_ready(c, "pamo_s3_step_eval"));
    PAMO_HIP(hipSetDevice(c.device));
    return pamo::s3::step_once(c, on_newton, user);
}

int pamo_s3_get_vertices(pamo_s3* s, float* d_out, double scale, const double* t3) {
    pamo::clear_error();
    PAMO_CHECK(s && d_out, "pamo_s3_get_vertices: null argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    // normalised -> input frame in FP64: (q - t) / scale
    std::vector<float4> h(c.nV);
    PAMO_HIP(hipMemcpyAsync(h.data(), c.q, sizeof(float4) * c.nV, hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    std::vector<float> o(3 * (size_t)c.nV);
    for (int i = 0; i < c.nV; ++i) {
        o[3 * (size_t)i] = (float)(((double)h[i].x - t3[0]) / scale);
        o[3 * (size_t)i + 1] = (float)(((double)h[i].y - t3[1]) / scale);
        o[3 * (size_t)i + 2] = (float)(((double)h[i].z - t3[2]) / scale);
    }
    PAMO_HIP(hipMemcpyAsync(d_out, o.data(), sizeof(float) * o.size(), hipMemcpyHostToDevice, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

int pamo_s3_get_normalized(pamo_s3* s, double* h_out) {
    pamo::clear_error();
    PAMO_CHECK(s && h_out, "pamo_s3_get_normalized: null argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    std::vector<float4> h(c.nV);
    PAMO_HIP(hipMemcpyAsync(h.data(), c.q, sizeof(float4) * c.nV, hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    for (int i = 0; i < c.nV; ++i) {
        h_out[3 * (size_t)i] = h[i].x;
        h_out[3 * (size_t)i + 1] = h[i].y;
        h_out[3 * (size_t)i + 2] = h[i].z;
    }
    return 0;
}

int pamo_s3_set_dense(pamo_s3* s, const double* h_A, const double* h_b, double c0, int n_v) {
    pamo::clear_error();
    PAMO_CHECK(s && h_A && h_b, "pamo_s3_set_dense: null argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_CHECK(n_v == c.nV && n_v > 0 && n_v <= 1024, "pamo_s3_set_dense: n_v must equal the mesh's (<= 1024)");
    PAMO_HIP(hipSetDevice(c.device));
    const size_t n3 = 3 * (size_t)n_v;
    PAMO_TRY(pamo::s3::dalloc(&c.dense_A, n3 * n3, c.stream));
    PAMO_TRY(pamo::s3::dalloc(&c.dense_b, n3, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.dense_A, h_A, sizeof(double) * n3 * n3, hipMemcpyHostToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.dense_b, h_b, sizeof(double) * n3, hipMemcpyHostToDevice, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    c.dense_c = c0;
    return 0;
}

namespace {
int read_vec(Ctx& c, const float4* d, double* out) {
    std::vector<float4> h(c.nV);
    PAMO_HIP(hipMemcpyAsync(h.data(), d, sizeof(float4) * c.nV, hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    for (int i = 0; i < c.nV; ++i) {
        out[3 * (size_t)i] = h[i].x;
        out[3 * (size_t)i + 1] = h[i].y;
        out[3 * (size_t)i + 2] = h[i].z;
    }
    return 0;
}
int write_vec(Ctx& c, const double* in, float4* d) {
    std::vector<float4> h(c.nV);
    for (int i = 0; i < c.nV; ++i)
        h[i] = make_float4((float)in[3 * (size_t)i], (float)in[3 * (size_t)i + 1], (float)in[3 * (size_t)i + 2], 0.0f);
    PAMO_HIP(hipMemcpyAsync(d, h.data(), sizeof(float4) * c.nV, hipMemcpyHostToDevice, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    return 0;
}
double read_slot(Ctx& c, int slot) {
    double v = 0.0;
    if (hipMemcpyAsync(&v, c.sc + slot, sizeof(double), hipMemcpyDeviceToHost, c.stream) != hipSuccess) return NAN;
    if (hipStreamSynchronize(c.stream) != hipSuccess) return NAN;
    return v;
}
}  // namespace

int pamo_s3_debug(pamo_s3* s, int what, const double* h_in, long long n_in, double* h_out, long long cap,
                  long long* n_out) {
    pamo::clear_error();
    PAMO_CHECK(s && n_out, "pamo_s3_debug: null argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_CHECK(c.nV > 0, "pamo_s3_debug: no mesh registered");
    PAMO_HIP(hipSetDevice(c.device));
    using namespace pamo::s3;
    const long long n3 = 3LL * c.nV;
    *n_out = 0;
    auto need_in = [&](long long n) { return h_in != nullptr && n_in >= n; };
    auto need_out = [&](long long n) { return h_out != nullptr && cap >= n; };
    int nC = 0;
    switch (what) {
    case PAMO_S3_DBG_SET_Q:
        PAMO_CHECK(need_in(n3), "pamo_s3_debug: SET_Q needs 3 n_v inputs");
        PAMO_TRY(write_vec(c, h_in, c.q));
        return 0;
    case PAMO_S3_DBG_TARGET:
        PAMO_TRY(check_ready(c, "pamo_s3_debug"));
        PAMO_CHECK(need_out(n3), "pamo_s3_debug: output too small");
        PAMO_TRY(update_target(c));
        PAMO_TRY(read_vec(c, c.target, h_out));
        *n_out = n3;
        return 0;
    case PAMO_S3_DBG_CONTACTS: {
        PAMO_TRY(check_ready(c, "pamo_s3_debug"));
        PAMO_TRY(detect_contacts(c, nC));
        PAMO_CHECK(need_out(5LL * nC), "pamo_s3_debug: output too small");
        std::vector<int4> idx(nC);
        std::vector<int> ty(nC);
        if (nC > 0) {
            PAMO_HIP(hipMemcpyAsync(idx.data(), c.cidx, sizeof(int4) * nC, hipMemcpyDeviceToHost, c.stream));
            PAMO_HIP(hipMemcpyAsync(ty.data(), c.ctype1, sizeof(int) * nC, hipMemcpyDeviceToHost, c.stream));
        }
        PAMO_HIP(hipStreamSynchronize(c.stream));
        for (int i = 0; i < nC; ++i) {
            double* o = h_out + 5LL * i;
            o[0] = idx[i].x; o[1] = idx[i].y; o[2] = idx[i].z; o[3] = idx[i].w; o[4] = ty[i];
        }
        *n_out = nC;
        return 0;
    }
    case PAMO_S3_DBG_ENERGY: {
        PAMO_TRY(check_ready(c, "pamo_s3_debug"));
        PAMO_CHECK(need_out(6), "pamo_s3_debug: output too small");
        PAMO_TRY(detect_contacts(c, nC));
        const int bits[5] = {PAMO_S3_E_M2GT, PAMO_S3_E_ELASTIC, PAMO_S3_E_HINGE, PAMO_S3_E_COLLISION, PAMO_S3_E_DENSE};
        for (int k = 0; k < 5; ++k) {
            if (c.P.mask & bits[k]) {
                PAMO_TRY(compute_energy(c, nC, SC_E, bits[k]));
                h_out[k] = read_slot(c, SC_E);
            } else {
                h_out[k] = 0.0;
            }
        }
        PAMO_TRY(compute_energy(c, nC, SC_E, c.P.mask));
        h_out[5] = read_slot(c, SC_E);
        *n_out = 6;
        return 0;
    }
    case PAMO_S3_DBG_RESOLVE: {
        PAMO_TRY(check_ready(c, "pamo_s3_debug"));
        PAMO_CHECK(need_out(n3), "pamo_s3_debug: output too small");
        PAMO_TRY(cg_solve(c));
        PAMO_TRY(read_vec(c, c.p, h_out));
        *n_out = n3;
        return 0;
    }
    case PAMO_S3_DBG_DIFF:
    case PAMO_S3_DBG_HVP:
    case PAMO_S3_DBG_SOLVE:
    case PAMO_S3_DBG_HINGE: {
        PAMO_TRY(check_ready(c, "pamo_s3_debug"));
        const long long want = what == PAMO_S3_DBG_DIFF ? 2 * n3 : what == PAMO_S3_DBG_HINGE ? 144LL * c.nH : n3;
        PAMO_CHECK(need_out(want), "pamo_s3_debug: output too small");
        PAMO_CHECK(what != PAMO_S3_DBG_HVP || need_in(n3), "pamo_s3_debug: HVP needs 3 n_v inputs");
        PAMO_TRY(detect_contacts(c, nC));
        PAMO_TRY(compute_energy(c, nC, SC_E, c.P.mask));
        PAMO_TRY(compute_diff(c, nC, c.P.mask));
        if (what == PAMO_S3_DBG_DIFF) {
            PAMO_TRY(read_vec(c, c.grad, h_out));
            PAMO_TRY(read_vec(c, c.diag, h_out + n3));
        } else if (what == PAMO_S3_DBG_HVP) {
            PAMO_TRY(write_vec(c, h_in, c.v));
            CgVec cv{c.v, c.v, c.zr_part, c.zr_part, 1, 1};  // dx = v as given
            PAMO_TRY(compute_hvp(c, c.P.mask, cv, c.v, c.Av, c.vav_part));
            PAMO_TRY(read_vec(c, c.Av, h_out));
        } else if (what == PAMO_S3_DBG_SOLVE) {
            PAMO_TRY(cg_solve(c));
            PAMO_TRY(read_vec(c, c.p, h_out));
        } else {
            std::vector<float> hb(144 * (size_t)c.nH);
            if (c.nH > 0)
                PAMO_HIP(hipMemcpyAsync(hb.data(), c.hblock, sizeof(float) * hb.size(), hipMemcpyDeviceToHost,
                                        c.stream));
            PAMO_HIP(hipStreamSynchronize(c.stream));
            for (size_t k = 0; k < hb.size(); ++k) h_out[k] = hb[k];
        }
        *n_out = want;
        return 0;
    }
    case PAMO_S3_DBG_CCD: {
        PAMO_TRY(check_ready(c, "pamo_s3_debug"));
        PAMO_CHECK(need_in(n3) && need_out(n3 + 1), "pamo_s3_debug: CCD needs 3 n_v inputs and 3 n_v + 1 outputs");
        PAMO_HIP(hipMemcpyAsync(c.qn, c.q, sizeof(float4) * c.nV, hipMemcpyDeviceToDevice, c.stream));
        PAMO_TRY(detect_contacts(c, nC));
        PAMO_TRY(compute_energy(c, nC, SC_E, c.P.mask));
        PAMO_TRY(write_vec(c, h_in, c.p));
        PAMO_TRY(clamp_and_ccd(c, nC));
        PAMO_TRY(read_vec(c, c.p, h_out));
        float st = 0.0f;
        PAMO_HIP(hipMemcpyAsync(&st, c.ccd_step, sizeof(float), hipMemcpyDeviceToHost, c.stream));
        PAMO_HIP(hipStreamSynchronize(c.stream));
        h_out[n3] = st;
        *n_out = n3 + 1;
        return 0;
    }
    default:
        pamo::set_error("pamo_s3_debug: unknown item");
        return 1;
    }
}

int pamo_s3_prepare_host(const float* h_v, int n_v, const int* h_f, int n_f, double scale, const double* t3,
                         int* h_edges, int* n_e, int* h_hinges, int* n_h, float* h_voronoi, double* min_angle_deg) {
    pamo::clear_error();
    PAMO_CHECK(h_v && h_f && t3 && h_edges && n_e && h_hinges && n_h && h_voronoi && min_angle_deg,
               "pamo_s3_prepare_host: null argument");
    pamo::s3::HostPrep hp;
    pamo::s3::host_prepare(h_v, n_v, h_f, n_f, scale, t3, hp);
    std::copy(hp.edges.begin(), hp.edges.end(), h_edges);
    std::copy(hp.hinges.begin(), hp.hinges.end(), h_hinges);
    std::copy(hp.vor.begin(), hp.vor.end(), h_voronoi);
    *n_e = (int)(hp.edges.size() / 2);
    *n_h = (int)(hp.hinges.size() / 4);
    *min_angle_deg = hp.min_angle_deg;
    return 0;
}

int pamo_s3_normalization(pamo_s3* s, const float* d_v, int n_v, double* scale, double* t3) {
    pamo::clear_error();
    PAMO_CHECK(s && d_v && scale && t3 && n_v > 0, "pamo_s3_normalization: bad argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    const int nb = (int)pamo::grid_for(n_v, pamo::s3::kB);
    PAMO_TRY(pamo::s3::ensure_part(c, 9LL * nb * pamo::s3::kB));
    hipLaunchKernelGGL(pamo::s3::gt_stats_k, dim3(nb), dim3(pamo::s3::kB), 0, c.stream, d_v, n_v, c.part);
    PAMO_HIP(hipGetLastError());
    // the mean as numpy's sequential FP64 sum (one workgroup), after the
    // bounds' partials
    double* d_sum = c.part + 9LL * nb;
    hipLaunchKernelGGL(pamo::s3::seq_sum3_k, dim3(1), dim3(pamo::s3::kSeqThreads), 0, c.stream, d_v, n_v, d_sum);
    PAMO_HIP(hipGetLastError());
    std::vector<double> h(9 * (size_t)nb + 3);
    PAMO_HIP(hipMemcpyAsync(h.data(), c.part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], h[9 * (size_t)b + 3 + k]);
            hi[k] = std::max(hi[k], h[9 * (size_t)b + 6 + k]);
        }
    const double* sum = h.data() + 9 * (size_t)nb;
    const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    *scale = 1.0 / ext;
    for (int k = 0; k < 3; ++k) t3[k] = -(sum[k] / n_v) * *scale;
    return 0;
}

int pamo_s3_set_mesh_device(pamo_s3* s, const float* d_v, int n_v, const int* d_f, int n_f, double scale,
                            const double* t3) {
    pamo::clear_error();
    PAMO_CHECK(s && d_v && d_f && t3 && n_v > 0 && n_f > 0, "pamo_s3_set_mesh_device: bad argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    PAMO_HIP(hipSetDevice(c.device));
    std::vector<float> hv(3 * (size_t)n_v);
    std::vector<int> hf(3 * (size_t)n_f);
    PAMO_HIP(hipMemcpyAsync(hv.data(), d_v, sizeof(float) * hv.size(), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipMemcpyAsync(hf.data(), d_f, sizeof(int) * hf.size(), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    pamo::s3::HostPrep hp;
    pamo::s3::host_prepare(hv.data(), n_v, hf.data(), n_f, scale, t3, hp);
    const int use_hinge = (c.P.mask & PAMO_S3_E_HINGE) != 0;
    const int n_e = (int)(hp.edges.size() / 2), n_h = use_hinge ? (int)(hp.hinges.size() / 4) : 0;
    PAMO_CHECK(!use_hinge || n_h == n_e, "Number of hinges != number of edges (the mesh is not closed and manifold)");
    // system.py:217-222: min angle in degrees against min_angle_thres (a NaN
    // angle compares false, as numpy's does)
    const int do_nothing = hp.min_angle_deg <= c.min_angle_thres ? 1 : 0;
    return pamo_s3_set_mesh(s, d_v, n_v, d_f, n_f, scale, t3, hp.edges.data(), n_e,
                            n_h ? hp.hinges.data() : nullptr, n_h, hp.vor.data(), do_nothing);
}

int pamo_s3_merge_vertices(pamo_s3* s, const float* d_v, int n_v, int* d_f, int n_f, float* d_v_out,
                           int* out_n_v) {
    pamo::clear_error();
    PAMO_CHECK(s && out_n_v && n_v >= 0 && n_f >= 0, "pamo_s3_merge_vertices: bad argument");
    PAMO_CHECK(n_v == 0 || (d_v && d_v_out), "pamo_s3_merge_vertices: null vertex array");
    PAMO_CHECK(n_f == 0 || d_f, "pamo_s3_merge_vertices: null face array");
    PAMO_CHECK(n_v == 0 || d_v_out != d_v, "pamo_s3_merge_vertices: the output must not alias the input");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    *out_n_v = 0;
    if (n_v == 0) return 0;  // an empty mesh is left as it is
    PAMO_HIP(hipSetDevice(c.device));
    using pamo::s3::dalloc;
    if (n_v + 1 > c.cap_mg) {
        c.cap_mg = n_v + 1 + n_v / 4 + 64;
        PAMO_TRY(dalloc(&c.mg_ref, (size_t)c.cap_mg, c.stream));
        PAMO_TRY(dalloc(&c.mg_first, (size_t)c.cap_mg, c.stream));
        PAMO_TRY(dalloc(&c.mg_keep, (size_t)c.cap_mg, c.stream));
        PAMO_TRY(dalloc(&c.mg_pos, (size_t)c.cap_mg, c.stream));
    }
    long long tab = 256;
    while (tab < 2LL * n_v) tab <<= 1;
    if (tab > c.cap_mg_tab) {
        c.cap_mg_tab = tab;
        PAMO_TRY(dalloc(&c.mg_tab, (size_t)tab, c.stream));
    }
    if (!c.mg_err) PAMO_TRY(dalloc(&c.mg_err, 1, c.stream));
    PAMO_TRY(pamo::s3::ensure_ws(c, (long long)n_v + 64));
    const long long n3 = 3LL * n_f;
    // without faces every vertex counts as referenced (merge_vertices'
    // branch for geometry without faces)
    PAMO_TRY(pamo::fill(c.mg_ref, n_f > 0 ? 0 : 1, n_v, c.stream));
    PAMO_TRY(pamo::fill(c.mg_tab, -1, tab, c.stream));
    PAMO_TRY(pamo::fill(c.mg_err, 0, 1, c.stream));
    if (n3 > 0)
        hipLaunchKernelGGL(pamo::s3::merge_mark_k, dim3(pamo::grid_for(n3, 256)), dim3(256), 0, c.stream, d_f, n3,
                           n_v, c.mg_ref, c.mg_err);
    const uint32_t mask = (uint32_t)(tab - 1);
    hipLaunchKernelGGL(pamo::s3::merge_insert_k, dim3(pamo::grid_for(n_v, 256)), dim3(256), 0, c.stream, d_v, n_v,
                       c.mg_ref, c.mg_tab, mask, c.mg_err);
    hipLaunchKernelGGL(pamo::s3::merge_first_k, dim3(pamo::grid_for(n_v + 1LL, 256)), dim3(256), 0, c.stream, d_v,
                       n_v, c.mg_ref, c.mg_tab, mask, c.mg_first, c.mg_keep);
    PAMO_HIP(hipGetLastError());
    PAMO_TRY(pamo::scan_exclusive(c.mg_keep, c.mg_pos, n_v + 1LL, c.ws, c.stream));
    hipLaunchKernelGGL(pamo::s3::merge_out_k, dim3(pamo::grid_for(n_v, 256)), dim3(256), 0, c.stream, d_v, n_v,
                       c.mg_keep, c.mg_pos, d_v_out);
    if (n3 > 0)
        hipLaunchKernelGGL(pamo::s3::merge_faces_k, dim3(pamo::grid_for(n3, 256)), dim3(256), 0, c.stream, d_f, n3,
                           n_v, c.mg_first, c.mg_pos);
    PAMO_HIP(hipGetLastError());
    int h[2] = {0, 0};
    PAMO_HIP(hipMemcpyAsync(&h[0], c.mg_err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipMemcpyAsync(&h[1], c.mg_pos + n_v, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    PAMO_CHECK(!(h[0] & 1), "pamo_s3_merge_vertices: a face index is out of range");
    PAMO_CHECK(!(h[0] & 2), "pamo_s3_merge_vertices: a referenced vertex is not finite");
    PAMO_CHECK(!(h[0] & 4), "pamo_s3_merge_vertices: the vertex table overflowed");
    *out_n_v = h[1];
    return 0;
}

int pamo_s3_set_option(pamo_s3* s, int option, long long value) {
    pamo::clear_error();
    PAMO_CHECK(s, "pamo_s3_set_option: null argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    if (option == PAMO_S3_OPT_CG_GRAPH) {
        PAMO_CHECK(value == 0 || value == 1, "pamo_s3_set_option: value must be 0 or 1");
        c.cg_graph = (int)value;
        return 0;
    }
    PAMO_CHECK(false, "pamo_s3_set_option: unknown option");
}

int pamo_s3_stats(pamo_s3* s, double* out8) {
    pamo::clear_error();
    PAMO_CHECK(s && out8, "pamo_s3_stats: null argument");
    Ctx& c = *reinterpret_cast<Ctx*>(s);
    out8[0] = c.last_energy;
    out8[1] = (double)c.stat_contacts;
    out8[2] = (double)c.stat_newton;
    out8[3] = (double)c.stat_cg;
    out8[4] = (double)c.stat_ls;
    out8[5] = (double)c.radius;
    out8[6] = c.do_nothing ? 1.0 : 0.0;
    out8[7] = (double)c.nH;
    return 0;
}

int pamo_s3_active_contacts(pamo_s3* s, long long* total) {
    pamo::clear_error();
    PAMO_CHECK(s && total, "pamo_s3_active_contacts: null argument");
    *total = reinterpret_cast<Ctx*>(s)->stat_active;
    return 0;
}

}  // extern "C"
