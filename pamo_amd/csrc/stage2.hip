// Stage 2 of PaMO on gfx950: one QEM collapse round with the intersection
// guard and its undo loop, plus the fused multi-round driver, behind the C
// ABI of include/pamo_hip.h.
//
// Reference behaviour followed (file:line in the reference repository):
//   round ............. simp_cuda/src/cusimp_free.cu:891-1048
//   rings / edges ..... :276-335 (ring order made canonical: ascending face)
//   quadrics .......... :337-366
//   edge cost ......... :368-585
//   independent set ... :587-610, :639-681
//   collapse/skinny ... :612-774
//   cleanup ........... :218-274
//   undo .............. :776-889, :1014-1040
//   LBVH + guard ...... simp_cuda/src/bvh/{bvh,morton_code,query,self_intersect}.cuh
//   round loop ........ simp_cuda/pamo/__init__.py:147-185
#include <algorithm>
#include <cstring>
#include <vector>

#include "prim.h"
#include "stage2_kernels.h"

namespace pamo {

static thread_local std::string g_err;
void set_error(const std::string& msg) {
    if (g_err.empty()) g_err = msg;
}
const char* get_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

// ---------------------------------------------------------------- launch profiler
static const char* kKernelNames[K_COUNT] = {
    "ring_count", "ring_fill", "ring_sort", "edge_count", "edge_fill", "quadric", "edge_cost", "propagate",
    "select_collapse", "line_flag", "mark_changed", "face_box", "morton", "radix_hist", "radix_scatter",
    "karras", "bvh_boxes", "guard_query", "undo_cand", "compact_mesh", "scan_tiles", "scan_sums", "other"};

struct Profiler {
    uint64_t mask = 0;
    struct Pending {
        int id;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    long long launches[K_COUNT] = {};
    double bytes[K_COUNT] = {};
    double ms[K_COUNT] = {};
    hipEvent_t get() {
        if (pool_used == pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[pool_used++];
    }
    void drain() {
        for (auto& q : pending) {
            (void)hipEventSynchronize(q.b);
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, q.a, q.b);
            ms[q.id] += t;
            bytes[q.id] += q.bytes;
            launches[q.id] += 1;
        }
        pending.clear();
        pool_used = 0;
    }
    ~Profiler() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

static thread_local Profiler* g_prof = nullptr;

void prof_begin(int id, double b, hipStream_t st) {
    Profiler* p = g_prof;
    if (!p || !(p->mask & (1ull << id))) return;
    hipEvent_t a = p->get(), e = p->get();
    (void)hipEventRecord(a, st);
    p->pending.push_back({id, b, a, e});
}

void prof_end(int id, hipStream_t st) {
    Profiler* p = g_prof;
    if (!p || !(p->mask & (1ull << id))) return;
    (void)hipEventRecord(p->pending.back().b, st);
    if (p->pending.size() > 4096) p->drain();
}

// ---------------------------------------------------------------- context
struct Ctx {
    hipStream_t stream = nullptr;      // the stream every call runs on
    hipStream_t own_stream = nullptr;  // the context's own non-blocking stream
    int device = 0;
    // logical capacities, grown like the reference (n + n/5)
    size_t cap_pts = 0, cap_tris = 0, cap_edges = 0;
    // physical allocations
    size_t alloc_pts = 0, alloc_tris = 0, alloc_edges = 0, alloc_ws = 0, alloc_parts = 0;
    // point sized
    float4 *p = nullptr, *op = nullptr, *np = nullptr;
    float4* Q = nullptr;  // 3 per point
    int *occ = nullptr, *pts_map = nullptr, *ring_off = nullptr, *ring_fill = nullptr;
    int* eord_cnt = nullptr;  // edges per smaller endpoint, then the launch-order cursor
    int *undo_list = nullptr, *tmp_undo = nullptr, *undo_prev = nullptr;  // 4 * alloc_pts
    uint32_t* invalid = nullptr;  // epochs (invalid_set_k)
    uint32_t inv_epoch = 1;       // the invalid table's current epoch
    uint64_t* vmin = nullptr;  // per-vertex minimum edge key
    // face sized
    int *ti = nullptr, *tj = nullptr, *tk = nullptr, *oi = nullptr, *oj = nullptr, *ok = nullptr;
    int *ni = nullptr, *nj = nullptr, *nk = nullptr;
    int* ring = nullptr;  // 3 * alloc_tris
    int *ring_raw = nullptr, *ring_own = nullptr;  // unsorted ring slots and their vertex
    int4* rv = nullptr;   // 3 * alloc_tris
    int* edge_off = nullptr;
    uint8_t *changed = nullptr, *flagged = nullptr;
    int *face_scan = nullptr, *qlist = nullptr;
    // LBVH
    uint32_t *bkey = nullptr, *bval = nullptr, *bkey2 = nullptr, *bval2 = nullptr;
    int *bleft = nullptr, *bright = nullptr, *bparent = nullptr;  // parent: 2n-1 nodes
    int2* brange = nullptr;
    float4 *bbox = nullptr, *fbox = nullptr, *bpre = nullptr, *bsuf = nullptr;  // 2 float4 per box
    float4 *blk = nullptr, *blkpre = nullptr, *blksuf = nullptr, *sblk = nullptr, *blktab = nullptr;
    long long blktab_stride = 0;  // entries per level of blktab
    float* box_part = nullptr;
    // edge sized
    int *eu = nullptr, *ev = nullptr;
    uint32_t* edge_cost = nullptr;
    int *esel = nullptr, *recs = nullptr, *pflag = nullptr, *pushed = nullptr;
    int* eord = nullptr;  // edge ids grouped by smaller endpoint (edge-cost launch order)
    int* erank = nullptr;  // each edge's rank among the edges of its smaller endpoint
    int* shuf = nullptr;
    int* ws = nullptr;
    LookbackScan lb;  // single-launch scans of the round loop
    int* dcnt = nullptr;
    int* hcnt = nullptr;
    // last round, host view
    int n_pts = 0, n_tris = 0, n_edges = 0, n_rec = 0, n_regular = 0, n_vertices_undo = 0;
    int n_cand = 0;  // records scanned for undo candidates
    int undo_iters = 0;
    int guard_small_stack = 0;  // test option: 32-entry guard frontier
    long long walk_cap = 0;     // test option: lowers the guard's walk bounds
    int placement = 0;          // 0 midpoint collapses (the reference), 1 optimal placement
    // guard tree carried across rounds (pamo_simplify's loop): the leaves of
    // the last build, renumbered at each compaction, dead ones with empty
    // boxes; rebuilt once fewer than tree_keep_pct % of them are live
    int tree_keep_pct = 80;     // 100: rebuild every round
    bool tree_valid = false;    // bval / hierarchy describe the current faces
    bool tree_fresh = false;    // the next round may refit instead of building
    int tree_n = 0;             // leaves of the last tree, as last read back
    // guard launch shape: meshes of at most guard_split_faces faces serve
    // each query group with 2^guard_split_log waves, one per subtree
    int guard_split_faces = 1 << 16;
    int guard_split_log = 3;
    int guard_waves = 4;        // 0: the compiler's register allocation for the guard; 4: capped for 4 waves
    int edge_lanes = 4;         // lanes per edge in the edge-cost kernel (1, 4 or 8)
    int ring_lanes = 8;         // lanes per item in the quadric, selection and undo kernels (1, 4 or 8)
    float* cent_out = nullptr;  // pamo_centroid's result (3 floats)
    float* cent_means = nullptr;  // pamo_centroid's face means, 3 rows of cent_pitch
    long long cap_cent = 0;
    int exact_guard = 0;      // 0 reference predicate, 1 exact, 2 exact with every pair in the exact pass
    int2* ulist = nullptr;    // exact guard: pairs left undecided by the filtered predicates
    size_t alloc_ulist = 0;
    Profiler prof;
};

// Activates the context's profiler for the calling thread during an entry point.
struct ProfScope {
    Profiler* saved;
    explicit ProfScope(Ctx& c) : saved(g_prof) { g_prof = c.prof.mask ? &c.prof : nullptr; }
    ~ProfScope() { g_prof = saved; }
};

template <typename T>
static int dalloc(T** q, size_t n, hipStream_t st) {
    if (*q) PAMO_HIP(hipFree(*q));
    *q = nullptr;
    PAMO_HIP(hipMalloc((void**)q, sizeof(T) * std::max<size_t>(n, 1)));
    // zeroed in order on the owning stream: a null-stream hipMemset is not
    // ordered with the context's non-blocking stream and may complete after
    // the kernels that fill the buffer
    PAMO_HIP(hipMemsetAsync(*q, 0, sizeof(T) * std::max<size_t>(n, 1), st));
    return 0;
}

static int ensure_parts(Ctx& c, size_t n_items) {
    size_t need = 7 * (n_items / 256 + 4);  // 6 floats per block, the whole box, a live count per block
    if (need > c.alloc_parts) {
        c.alloc_parts = need;
        PAMO_TRY(dalloc(&c.box_part, need, c.stream));
    }
    return 0;
}

static int ensure_ws(Ctx& c, long long n_items) {
    long long need = std::max(radix_workspace(n_items), scan_workspace(n_items + 1)) + 256;
    if ((size_t)need > c.alloc_ws) {
        c.alloc_ws = (size_t)need + need / 4;
        PAMO_TRY(dalloc(&c.ws, c.alloc_ws, c.stream));
    }
    return lookback_reserve(c.lb, n_items + 1, c.dcnt + C_WALK_ERR, c.stream);
}

// Point-sized buffers.  The logical capacity follows the reference's growth
// rule; persistent buffers (pts_map, undo lists, invalid table) restart at
// zero whenever it grows, as a fresh allocation would.
// Clears n elements of a buffer on the context's stream (a fresh allocation
// holds zeros).
template <typename T>
static int zero_async(Ctx& c, T* q, size_t n) {
    PAMO_HIP(hipMemsetAsync(q, 0, sizeof(T) * std::max<size_t>(n, 1), c.stream));
    return 0;
}

// Point-sized buffers.  The logical capacity grows like the reference's
// (n + n/5) and the buffers restart at zero when it does; an allocation that
// is already large enough (a context reset for a new mesh) is cleared in
// place instead of being freed, which would synchronise the whole device.
static int grow_pts(Ctx& c, size_t n) {
    if (n <= c.cap_pts) return 0;
    c.cap_pts = n + n / 5;
    size_t a = c.cap_pts + 1;
    if (a <= c.alloc_pts) {
        a = c.alloc_pts;
        PAMO_TRY(zero_async(c, c.p, a)); PAMO_TRY(zero_async(c, c.op, a)); PAMO_TRY(zero_async(c, c.np, a));
        PAMO_TRY(zero_async(c, c.Q, 3 * a));
        PAMO_TRY(zero_async(c, c.occ, a + 1)); PAMO_TRY(zero_async(c, c.pts_map, a + 1));
        PAMO_TRY(zero_async(c, c.ring_off, a + 2)); PAMO_TRY(zero_async(c, c.ring_fill, a + 1));
        PAMO_TRY(zero_async(c, c.eord_cnt, a + 1));
        PAMO_TRY(zero_async(c, c.undo_list, 4 * a)); PAMO_TRY(zero_async(c, c.tmp_undo, 4 * a));
        PAMO_TRY(zero_async(c, c.undo_prev, 4 * a));
        PAMO_TRY(zero_async(c, c.invalid, a)); PAMO_TRY(zero_async(c, c.vmin, a));
        return 0;
    }
    c.alloc_pts = a;
    PAMO_TRY(dalloc(&c.p, a, c.stream));
    PAMO_TRY(dalloc(&c.op, a, c.stream));
    PAMO_TRY(dalloc(&c.np, a, c.stream));
    PAMO_TRY(dalloc(&c.Q, 3 * a, c.stream));
    PAMO_TRY(dalloc(&c.occ, a + 1, c.stream));
    PAMO_TRY(dalloc(&c.pts_map, a + 1, c.stream));
    PAMO_TRY(dalloc(&c.ring_off, a + 2, c.stream));
    PAMO_TRY(dalloc(&c.ring_fill, a + 1, c.stream));
    PAMO_TRY(dalloc(&c.eord_cnt, a + 1, c.stream));
    PAMO_TRY(dalloc(&c.undo_list, 4 * a, c.stream));
    PAMO_TRY(dalloc(&c.tmp_undo, 4 * a, c.stream));
    PAMO_TRY(dalloc(&c.undo_prev, 4 * a, c.stream));
    PAMO_TRY(dalloc(&c.invalid, a, c.stream));  // epochs: zero is never current
    PAMO_TRY(dalloc(&c.vmin, a, c.stream));
    PAMO_TRY(ensure_parts(c, a));
    return ensure_ws(c, (long long)a + 1);
}

static int alloc_edges(Ctx& c, size_t ae) {
    c.alloc_edges = ae;
    PAMO_TRY(dalloc(&c.eu, ae, c.stream));
    PAMO_TRY(dalloc(&c.ev, ae, c.stream));
    PAMO_TRY(dalloc(&c.esel, ae + 1, c.stream));
    PAMO_TRY(dalloc(&c.recs, ae + 1, c.stream));
    PAMO_TRY(dalloc(&c.pflag, ae + 1, c.stream));
    PAMO_TRY(dalloc(&c.pushed, ae + 1, c.stream));
    PAMO_TRY(dalloc(&c.edge_cost, ae + 1, c.stream));
    PAMO_TRY(dalloc(&c.eord, ae + 1, c.stream));
    PAMO_TRY(dalloc(&c.erank, ae + 1, c.stream));
    return 0;
}

static int grow_tris(Ctx& c, size_t n) {
    if (n <= c.cap_tris) return 0;
    c.cap_tris = n + n / 5;
    size_t a = c.cap_tris + 1;
    if (a <= c.alloc_tris) {  // reset context: clear in place (see grow_pts)
        a = c.alloc_tris;
        int* ints[] = {c.ti, c.tj, c.tk, c.oi, c.oj, c.ok, c.ni, c.nj, c.nk};
        for (int* q : ints) PAMO_TRY(zero_async(c, q, a));
        PAMO_TRY(zero_async(c, c.ring, 3 * a)); PAMO_TRY(zero_async(c, c.rv, 3 * a));
        PAMO_TRY(zero_async(c, c.ring_raw, 3 * a)); PAMO_TRY(zero_async(c, c.ring_own, 3 * a));
        PAMO_TRY(zero_async(c, c.edge_off, a + 1));
        PAMO_TRY(zero_async(c, c.changed, a)); PAMO_TRY(zero_async(c, c.flagged, a));
        PAMO_TRY(zero_async(c, c.face_scan, a + 1)); PAMO_TRY(zero_async(c, c.qlist, a + 1));
        PAMO_TRY(zero_async(c, c.shuf, 4 * a + 64));
        PAMO_TRY(zero_async(c, c.eu, c.alloc_edges)); PAMO_TRY(zero_async(c, c.ev, c.alloc_edges));
        PAMO_TRY(zero_async(c, c.esel, c.alloc_edges + 1)); PAMO_TRY(zero_async(c, c.recs, c.alloc_edges + 1));
        PAMO_TRY(zero_async(c, c.pflag, c.alloc_edges + 1)); PAMO_TRY(zero_async(c, c.pushed, c.alloc_edges + 1));
        PAMO_TRY(zero_async(c, c.edge_cost, c.alloc_edges + 1));
        return 0;
    }
    c.alloc_tris = a;
    PAMO_TRY(dalloc(&c.ti, a, c.stream)); PAMO_TRY(dalloc(&c.tj, a, c.stream)); PAMO_TRY(dalloc(&c.tk, a, c.stream));
    PAMO_TRY(dalloc(&c.oi, a, c.stream)); PAMO_TRY(dalloc(&c.oj, a, c.stream)); PAMO_TRY(dalloc(&c.ok, a, c.stream));
    PAMO_TRY(dalloc(&c.ni, a, c.stream)); PAMO_TRY(dalloc(&c.nj, a, c.stream)); PAMO_TRY(dalloc(&c.nk, a, c.stream));
    PAMO_TRY(dalloc(&c.ring, 3 * a, c.stream));
    PAMO_TRY(dalloc(&c.ring_raw, 3 * a, c.stream));
    PAMO_TRY(dalloc(&c.ring_own, 3 * a, c.stream));
    PAMO_TRY(dalloc(&c.rv, 3 * a, c.stream));
    PAMO_TRY(dalloc(&c.edge_off, a + 1, c.stream));
    PAMO_TRY(dalloc(&c.changed, a, c.stream));
    PAMO_TRY(dalloc(&c.flagged, a, c.stream));
    PAMO_TRY(dalloc(&c.face_scan, a + 1, c.stream));
    PAMO_TRY(dalloc(&c.qlist, a + 1, c.stream));
    PAMO_TRY(dalloc(&c.bkey, a, c.stream)); PAMO_TRY(dalloc(&c.bval, a, c.stream));
    PAMO_TRY(dalloc(&c.bkey2, a, c.stream)); PAMO_TRY(dalloc(&c.bval2, a, c.stream));
    PAMO_TRY(dalloc(&c.bleft, a, c.stream)); PAMO_TRY(dalloc(&c.bright, a, c.stream));
    PAMO_TRY(dalloc(&c.bparent, 2 * a, c.stream));
    PAMO_TRY(dalloc(&c.brange, a, c.stream));
    PAMO_TRY(dalloc(&c.bbox, 4 * a, c.stream));  // 2n-1 nodes, 2 float4 each
    PAMO_TRY(dalloc(&c.fbox, 2 * a, c.stream));
    PAMO_TRY(dalloc(&c.bpre, 2 * a, c.stream));
    PAMO_TRY(dalloc(&c.bsuf, 2 * a, c.stream));
    const size_t nb = a / kBlock + 2, ns = nb / kBlock + 2;
    PAMO_TRY(dalloc(&c.blk, 2 * nb, c.stream));
    PAMO_TRY(dalloc(&c.blkpre, 2 * nb, c.stream));
    PAMO_TRY(dalloc(&c.blksuf, 2 * nb, c.stream));
    PAMO_TRY(dalloc(&c.sblk, 2 * ns, c.stream));
    PAMO_TRY(dalloc(&c.blktab, 2 * kTabLevels * nb, c.stream));
    c.blktab_stride = (long long)nb;
    PAMO_TRY(ensure_parts(c, a));
    PAMO_TRY(dalloc(&c.shuf, 4 * a + 64, c.stream));
    if (2 * a + 1 > c.alloc_edges) PAMO_TRY(alloc_edges(c, 2 * a + 1));  // at most 2 edges per face
    return ensure_ws(c, (long long)(4 * a + 64));
}

// edge_cost keeps stale values between rounds; it restarts at zero when the
// logical edge capacity grows (reference: cusimp_free.cu:180-194).
static int grow_edges(Ctx& c, size_t n, bool* reallocated = nullptr) {
    if (n <= c.cap_edges) return 0;
    c.cap_edges = n + n / 5;
    if (c.cap_edges + 1 > c.alloc_edges) {
        PAMO_TRY(alloc_edges(c, c.cap_edges + 1));
        if (reallocated) *reallocated = true;
    } else {
        PAMO_HIP(hipMemsetAsync(c.edge_cost, 0, sizeof(uint32_t) * (c.alloc_edges + 1), c.stream));
    }
    return 0;
}

static void free_all(Ctx& c) {
    void* ptrs[] = {c.p, c.op, c.np, c.Q, c.occ, c.pts_map, c.ring_off, c.ring_fill, c.eord_cnt, c.undo_list, c.tmp_undo,
                    c.undo_prev, c.invalid, c.ti, c.tj, c.tk, c.oi, c.oj, c.ok, c.ni, c.nj, c.nk, c.ring, c.ring_raw, c.ring_own, c.rv,
                    c.edge_off, c.vmin, c.changed, c.flagged, c.face_scan, c.qlist, c.bkey, c.bval, c.bkey2,
                    c.bval2, c.bleft, c.bright, c.bparent, c.brange, c.bbox, c.fbox, c.bpre, c.bsuf, c.blk, c.blkpre,
                    c.blksuf, c.sblk, c.blktab, c.box_part, c.eu, c.ev, c.edge_cost, c.esel, c.recs, c.pflag, c.pushed, c.eord, c.erank,
                    c.shuf, c.ws, c.dcnt, c.cent_out, c.ulist, c.cent_means};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    lookback_free(c.lb);
    if (c.hcnt) (void)hipHostFree(c.hcnt);
}

static Feistel make_feistel(uint64_t m) {
    Feistel f;
    uint32_t total = 4;
    if (m > 16) {
        total = 0;
        uint64_t t = m - 1;
        while (t) { ++total; t >>= 1; }
    }
    f.lbits = total / 2;
    f.rbits = total - f.lbits;
    f.lmask = (1u << f.lbits) - 1;
    f.rmask = (1u << f.rbits) - 1;
    uint64_t s = 1;  // minstd_rand, seed 1 (thrust::default_random_engine)
    for (int r = 0; r < 24; ++r) {
        s = (s * 48271ull) % 2147483647ull;
        f.key[r] = (uint32_t)s;
    }
    return f;
}

// ---------------------------------------------------------------- host orchestration
static Mesh cur_mesh(Ctx& c) { return Mesh{c.p, c.ti, c.tj, c.tk}; }
static Orig orig_mesh(Ctx& c) { return Orig{c.op, c.oi, c.oj, c.ok}; }

#define LAUNCH(kern, n, ...) LAUNCHP(K_OTHER, 0.0, kern, n, __VA_ARGS__)

// Launch with the profiler bracket; `bytes` is the kernel's algorithmic
// (compulsory) HBM traffic for this launch.
#define LAUNCHP(id, bytes, kern, n, ...)                                                                  \
    do {                                                                                                  \
        prof_begin(id, (double)(bytes), c.stream);                                                        \
        hipLaunchKernelGGL(kern, dim3(grid_for((long long)(n), kBlock)), dim3(kBlock), 0, c.stream,        \
                           __VA_ARGS__);                                                                  \
        prof_end(id, c.stream);                                                                           \
        PAMO_HIP(hipGetLastError());                                                                      \
    } while (0)

// Reads the device counters; reports (and clears) the error flags that
// kernels set instead of reading out of bounds or walking a corrupted tree.
static int sync_counters(Ctx& c) {
    PAMO_HIP(hipMemcpyAsync(c.hcnt, c.dcnt, sizeof(int) * C_COUNT, hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    if (c.hcnt[C_WALK_ERR] || c.hcnt[C_BAD_UNDO]) {
        const bool walk = (c.hcnt[C_WALK_ERR] & 1) != 0, scan = (c.hcnt[C_WALK_ERR] & 2) != 0;
        c.hcnt[C_WALK_ERR] = c.hcnt[C_BAD_UNDO] = 0;
        PAMO_HIP(hipMemsetAsync(c.dcnt + C_WALK_ERR, 0, 2 * sizeof(int), c.stream));
        PAMO_CHECK(!walk, "guard: a tree walk exceeded its bound (corrupted LBVH)");
        PAMO_CHECK(!scan, "scan: a look-back wait exceeded its bound");
        PAMO_CHECK(false, "undo list: entries outside the points of the previous call");
    }
    return 0;
}

// Flag the faces of intersecting pairs that involve a changed face
// (self_intersect.cuh:151-392, restricted to the faces that can change an
// undo decision), over an LBVH of the faces live at the start of the round.
// The first guard of a round builds the tree (Morton codes, sort, Karras
// hierarchy); a guard after an undo only recomputes the face and node boxes
// over the same hierarchy (refit): the undo moves vertices and restores
// faces, and every face it can restore is already a leaf, with an empty box
// while it is removed.  Which pairs overlap does not depend on the tree's
// shape, so the flags are those of a tree built afresh.
static int run_guard(Ctx& c, int nT, bool refit) {
    Mesh m = cur_mesh(c);
    const double nP = c.n_pts;
    const int nparts = (int)grid_for(nT + 1, kBlock);  // face_box_k also writes the sentinel box at nT
    int* leaf_part = (int*)(c.box_part + 6 * nparts + 6);
    LAUNCHP(K_FACE_BOX, 12.0 * nT + 16.0 * nP + 32.0 * nT + 4.0 * nT, face_box_k, nT + 1, m, nT, c.fbox, c.box_part,
            leaf_part, c.flagged, (const int*)c.oi);
    if (!refit) {
        hipLaunchKernelGGL(whole_box_k, dim3(1), dim3(kBlock), 0, c.stream, c.box_part, nparts,
                           (const int*)leaf_part, c.dcnt);
        PAMO_HIP(hipGetLastError());
        LAUNCHP(K_MORTON, 4.0 * nT + 32.0 * nT + 8.0 * nT, morton_k, nT, m, orig_mesh(c), nT, (const float4*)c.fbox,
                (const float*)(c.box_part + 6 * nparts), c.bkey, c.bval);
        PAMO_TRY(radix_sort_pairs(c.bkey, c.bval, c.bkey2, c.bval2, nT, 32, c.ws, c.stream));
        LAUNCHP(K_KARRAS, 8.0 * nT + 16.0 * nT + 8.0 * nT, karras_k, nT, c.bkey, c.bval, c.dcnt, c.bleft, c.bright,
                c.bparent, c.brange);
    }
    // leaves of the tree: at most nT for a tree built this round, the last
    // build's count for a tree carried across rounds
    const int nL = refit ? std::max(nT, c.tree_n) : nT;
    // node boxes from leaf ranges: no communication between workgroups
    const unsigned nb0 = grid_for(nL, kBlock), nb1 = grid_for(nb0, kBlock);
    prof_begin(K_BVH_BOXES, 4.0 * nT + 32.0 * nT + 32.0 * nT + 64.0 * nT, c.stream);
    hipLaunchKernelGGL(box_tab_k, dim3(nb0), dim3(kBlock), 0, c.stream, (const float4*)c.fbox,
                       (const uint32_t*)c.bval, (const int*)c.dcnt, 0, c.bbox, (const int2*)c.brange, c.bpre, c.bsuf,
                       c.blk, (float4*)nullptr, 0LL);
    hipLaunchKernelGGL(box_tab_k, dim3(nb1), dim3(kBlock), 0, c.stream, (const float4*)c.blk,
                       (const uint32_t*)nullptr, (const int*)c.dcnt, 1, (float4*)nullptr, (const int2*)nullptr,
                       c.blkpre, c.blksuf, c.sblk, c.blktab, c.blktab_stride);
    hipLaunchKernelGGL(node_box_k, dim3(grid_for(nL, kBlock)), dim3(kBlock), 0, c.stream, (const int2*)c.brange,
                       (const int*)c.dcnt, (const float4*)c.bpre, (const float4*)c.bsuf, (const float4*)c.blktab,
                       c.blktab_stride, (const float4*)c.blkpre, (const float4*)c.blksuf, (const float4*)c.sblk,
                       c.bbox);
    // a refit keeps the hierarchy, so the links in the records' .w fields
    // stand (box_tab_k rewrites each leaf's face id, which a carry renumbers)
    if (!refit)
        hipLaunchKernelGGL(rope_k, dim3(grid_for(2LL * nL, kBlock)), dim3(kBlock), 0, c.stream, (const int*)c.bleft,
                           (const int2*)c.brange, (const int*)c.dcnt, (const uint32_t*)c.bval, c.bbox);
    prof_end(K_BVH_BOXES, c.stream);
    PAMO_HIP(hipGetLastError());
    // queries: changed live faces in Morton order
    PAMO_TRY(scan_op(QueryListOp{c.bval, c.changed, c.ti, nT, c.qlist, c.dcnt}, nL, c.lb, c.stream));
    // meshes up to 256k faces split every query group over 8 subtrees: late
    // rounds hold few queries, and one query with a large box would otherwise
    // hold a whole launch in one wave
    const int part_log = nT <= c.guard_split_faces ? c.guard_split_log : 0;
    const long long guard_threads = ((((long long)nT + kGuardQ - 1) / kGuardQ) << part_log) * kWave;
    const double guard_bytes = 4.0 * nT + 12.0 * nT + 16.0 * nP + 32.0 * nT + 64.0 * nT + 8.0 * nT + 4.0 * nT + 1.0 * nT;
    Undecided u{nullptr, c.dcnt + C_UNDECIDED, 0, c.exact_guard == 2 ? 1 : 0};
    if (c.exact_guard) {
        // list capacity: undecided pairs are near-degenerate ones (a few per
        // changed face at most); the test setting lists every candidate pair
        const size_t cap = c.exact_guard == 2 ? 48 * (size_t)c.alloc_tris : std::max<size_t>(1 << 16, 2 * c.alloc_tris);
        if (cap > c.alloc_ulist) {
            PAMO_TRY(dalloc(&c.ulist, cap, c.stream));
            c.alloc_ulist = cap;
        }
        u.list = c.ulist;
        u.cap = (int)std::min<size_t>(c.alloc_ulist, 0x7fffffff);
        PAMO_HIP(hipMemsetAsync(c.dcnt + C_UNDECIDED, 0, sizeof(int), c.stream));
        LAUNCHP(K_GUARD_QUERY, guard_bytes, (guard_query_k<kGuardStack, true>), guard_threads, m, c.qlist, c.dcnt,
                c.changed, c.fbox, (const float4*)c.bbox, c.bright, 1e-3f, c.flagged, u, part_log, c.dcnt + C_WALK_ERR, c.walk_cap);
        hipLaunchKernelGGL(exact_pairs_k, dim3(1024), dim3(kBlock), 0, c.stream, m, (const int2*)c.ulist,
                           (const int*)(c.dcnt + C_UNDECIDED), u.cap, c.flagged);
        PAMO_HIP(hipGetLastError());
    } else if (c.guard_small_stack) {
        LAUNCHP(K_GUARD_QUERY, guard_bytes, (guard_query_k<kGuardStackSmall, false>), guard_threads, m, c.qlist,
                c.dcnt, c.changed, c.fbox, (const float4*)c.bbox, c.bright, 1e-3f, c.flagged, u, part_log, c.dcnt + C_WALK_ERR, c.walk_cap);
    } else if (c.guard_waves == 4) {
        LAUNCHP(K_GUARD_QUERY, guard_bytes, guard_query_w4_k, guard_threads, m, c.qlist, c.dcnt, c.changed, c.fbox,
                (const float4*)c.bbox, c.bright, 1e-3f, c.flagged, u, part_log, c.dcnt + C_WALK_ERR, c.walk_cap);
    } else {
        LAUNCHP(K_GUARD_QUERY, guard_bytes, (guard_query_k<kGuardStack, false>), guard_threads, m, c.qlist, c.dcnt,
                c.changed, c.fbox, (const float4*)c.bbox, c.bright, 1e-3f, c.flagged, u, part_log, c.dcnt + C_WALK_ERR, c.walk_cap);
    }
    return 0;
}

// Undo candidates among the first n_cand records (get_undo_candidate_kernel,
// cusimp_free.cu:776-827, whose grid covers the regular-collapse count
// rounded up to blocks of 512: line-collapse records past it are never
// candidates).
static int undo_candidates(Ctx& c, int& npushed) {
    const int n = c.n_cand;
    PAMO_TRY(scan_op(UndoListOp{c.recs, c.ring_off, c.ring, c.eu, c.ev, c.flagged, c.pushed, c.dcnt}, n, c.lb,
                     c.stream, K_UNDO_CAND, 16.0 * n + 1.0 * c.n_tris));
    PAMO_TRY(sync_counters(c));
    npushed = c.hcnt[C_PUSHED];
    c.tree_n = c.hcnt[C_NLIVE];  // leaves of the tree this guard used
    PAMO_CHECK(!c.exact_guard || c.hcnt[C_UNDECIDED] <= (long long)c.alloc_ulist,
               "exact guard: more pairs left undecided than the list holds");
    return 0;
}

// One round on the mesh held in (p, ti, tj, tk) with nP points and nT faces
// (cusimp_free.cu:891-1048); the undo-list prologue has already run.
// Vertex -> face rings in ascending face order (:276-304), the edge list
// (:306-335), the quadrics (:337-366) and the edge-cost launch order; sets
// c.n_edges.  Shared by the round and the legacy round.
// ring_off, ring_fill and eord_cnt must hold zeros (round_init_k).
static int topology_and_quadrics(Ctx& c, Mesh m, int nP, int nT) {
    // per-vertex ring counts and per-face edge counts in one pass
    LAUNCHP(K_RING_COUNT, 12.0 * nT + 8.0 * nP + 4.0 * nT, topo_count_k, nT + 1, m, nT, c.ring_off, c.edge_off);
    PAMO_TRY(scan_exclusive_lb(c.ring_off, c.ring_off, nP + 1, c.lb, c.stream));
    LAUNCHP(K_RING_FILL, 12.0 * nT + 12.0 * nT + 32.0 * nT, ring_fill_k, nT, m, nT, c.ring_off, c.ring_fill,
            c.ring_raw, c.ring_own);
    LAUNCHP(K_RING_SORT, 4.0 * nP + 24.0 * nT + 12.0 * nT + 60.0 * nT, ring_rank_k, 3LL * nT, m, c.ring_off,
            c.ring_raw, c.ring_own, 3 * nT, c.ring_off + nP, c.ring, c.rv);
    // edges: at most two per face, which the edge buffers always hold, so the
    // fill and the quadrics run before the host reads the count back (the
    // read then waits on them instead of stalling the stream)
    PAMO_TRY(scan_exclusive_lb(c.edge_off, c.edge_off, nT + 1, c.lb, c.stream));
    PAMO_CHECK((size_t)2 * nT <= c.alloc_edges, "edge buffers smaller than two edges per face");
    LAUNCHP(K_EDGE_FILL, 16.0 * nT + 12.0 * 1.5 * nT, edge_fill_k, nT, m, nT, c.edge_off, c.eu, c.ev, c.dcnt,
            c.eord_cnt, c.erank);
    // quadrics
    const double quad_bytes = 4.0 * nP + 48.0 * nT + 16.0 * nP + 48.0 * nP;
    if (c.ring_lanes == 8)
        LAUNCHP(K_QUADRIC, quad_bytes, quadric_group_k<8>, 8LL * nP, m, c.ring_off, c.rv, nP, c.Q);
    else if (c.ring_lanes == 4)
        LAUNCHP(K_QUADRIC, quad_bytes, quadric_group_k<4>, 4LL * nP, m, c.ring_off, c.rv, nP, c.Q);
    else
        LAUNCHP(K_QUADRIC, quad_bytes, quadric_k, nP, m, c.ring_off, c.rv, nP, c.Q);
    PAMO_TRY(sync_counters(c));
    const int nE = c.hcnt[C_NEDGE];
    c.n_edges = nE;
    bool realloc = false;
    PAMO_TRY(grow_edges(c, (size_t)nE, &realloc));
    if (realloc) {  // fresh edge buffers: fill and count them again
        PAMO_HIP(hipMemsetAsync(c.eord_cnt, 0, sizeof(int) * (nP + 1), c.stream));
        LAUNCHP(K_EDGE_FILL, 16.0 * nT + 12.0 * nE, edge_fill_k, nT, m, nT, c.edge_off, c.eu, c.ev, c.dcnt,
                c.eord_cnt, c.erank);
    }
    // edge-cost launch order: edges grouped by smaller endpoint (counted by
    // edge_fill_k)
    PAMO_TRY(scan_exclusive_lb(c.eord_cnt, c.eord_cnt, nP + 1, c.lb, c.stream));
    LAUNCH(edge_order_scatter_k, nE, c.eu, (const int*)c.erank, nE, (const int*)c.eord_cnt, c.eord);
    return 0;
}

// u32(clamp(thr, 0, 10) / 10 * 2^32) with saturation (:923, cusimp.cu:589)
static uint32_t collapse_threshold(float thr) {
    float x = thr > 0.0f ? thr : 0.0f;
    x = x < 10.0f ? x : 10.0f;
    float q = x / 10.0f * 4294967296.0f;
    return !(q < 4294967296.0f) ? 0xFFFFFFFFu : (uint32_t)q;
}

// One round of the legacy simplifier behind PaSP, CUSimp::forward
// (cusimp.cu:587-720), on the mesh loaded in c: rings, edges and quadrics as
// in the round, the legacy cost, the per-vertex minima and the midpoint
// collapse; no guard, no undo, no line collapse.  Leaves occ and pts_map.
static int pasp_round_device(Ctx& c, int nP, int nT, float scale, float thr) {
    const uint32_t collapse_t = collapse_threshold(thr);
    Mesh m = cur_mesh(c);
    c.n_pts = nP;
    c.n_tris = nT;
    LAUNCH(round_init_k, std::max(nP + 1, nT), m, Mesh{nullptr, nullptr, nullptr, nullptr}, nP, nT, c.occ, c.ring_off,
           c.ring_fill, c.eord_cnt, (unsigned long long*)c.vmin, (uint8_t*)nullptr);
    PAMO_TRY(topology_and_quadrics(c, m, nP, nT));
    const int nE = c.n_edges;
    LAUNCHP(K_EDGE_COST, 12.0 * nE + 4.0 * nP + 48.0 * nT + 16.0 * nP + 48.0 * nP + 4.0 * nE, pasp_cost_k, nE, m,
            c.ring_off, c.rv, c.eu, c.ev, c.eord, nE, c.Q, scale, thr, c.edge_cost);
    LAUNCHP(K_PROPAGATE, 12.0 * nE + 16.0 * nP, propagate_k, nE, c.eu, c.ev, nE, c.edge_cost,
            (unsigned long long*)c.vmin);
    LAUNCHP(K_SELECT_COLLAPSE, 16.0 * nE + 4.0 * nP + 48.0 * nT + 8.0 * nP, pasp_collapse_k, nE, m, c.occ,
            c.ring_off, c.ring, (const int4*)c.rv, c.eu, c.ev, nE, c.edge_cost, (const unsigned long long*)c.vmin,
            collapse_t);
    PAMO_TRY(scan_exclusive_lb(c.occ, c.pts_map, nP + 1, c.lb, c.stream));
    return 0;
}

static int round_device(Ctx& c, int nP, int nT, float scale, float thr, bool is_stuck) {
    const float tres = thr, edge_s = scale;
    const uint32_t collapse_t = collapse_threshold(thr);
    Mesh m = cur_mesh(c);
    c.n_pts = nP;
    c.n_tris = nT;
    // snapshot for undo, occupancy, ring counters, minimum keys, changed flags
    LAUNCH(round_init_k, std::max(nP + 1, nT), m, Mesh{c.op, c.oi, c.oj, c.ok}, nP, nT, c.occ, c.ring_off, c.ring_fill,
           c.eord_cnt, (unsigned long long*)c.vmin, c.changed);
    PAMO_TRY(topology_and_quadrics(c, m, nP, nT));
    const int nE = c.n_edges;
    CostParams cp{edge_s, tres, is_stuck ? 1 : 0, c.inv_epoch};
    // (+ 16 B per vertex: the minimum keys the fused propagation lowers)
    const double cost_bytes = 12.0 * nE + 4.0 * nP + 48.0 * nT + 16.0 * nP + 48.0 * nP + 1.0 * nP + 4.0 * nE + 16.0 * nP;
#define PAMO_EDGE_COST(K, lanes)                                                                              \
    LAUNCHP(K_EDGE_COST, cost_bytes, K, (long long)nE * (lanes), m, c.ring_off, c.rv, c.eu, c.ev, c.eord, nE,     \
            c.invalid, c.Q, cp, c.edge_cost, (unsigned long long*)c.vmin)
    if (c.edge_lanes == 8) {
        if (c.placement) PAMO_EDGE_COST((edge_cost_group_k<true, 8>), 8);
        else PAMO_EDGE_COST((edge_cost_group_k<false, 8>), 8);
    } else if (c.edge_lanes == 4) {
        if (c.placement) PAMO_EDGE_COST((edge_cost_group_k<true, 4>), 4);
        else PAMO_EDGE_COST((edge_cost_group_k<false, 4>), 4);
    } else {
        if (c.placement) PAMO_EDGE_COST(edge_cost_k<true>, 1);
        else PAMO_EDGE_COST(edge_cost_k<false>, 1);
    }
#undef PAMO_EDGE_COST
    // independent set + collapse (the cost kernels have lowered the
    // per-vertex minimum keys)
    const double sel_bytes = 16.0 * nE + 4.0 * nP + 48.0 * nT + 8.0 * nP;
#define PAMO_SELECT(K, lanes)                                                                                 \
    LAUNCHP(K_SELECT_COLLAPSE, sel_bytes, K, (long long)(nE + 1) * (lanes), m, orig_mesh(c), c.occ, c.ring_off,   \
            c.ring, (const int4*)c.rv, c.eu, c.ev, nE, c.edge_cost, (const unsigned long long*)c.vmin, collapse_t,  \
            c.esel, (const float4*)c.Q, c.changed)
    if (c.ring_lanes == 8) {
        if (c.placement) PAMO_SELECT((select_collapse_group_k<true, 8>), 8);
        else PAMO_SELECT((select_collapse_group_k<false, 8>), 8);
    } else if (c.ring_lanes == 4) {
        if (c.placement) PAMO_SELECT((select_collapse_group_k<true, 4>), 4);
        else PAMO_SELECT((select_collapse_group_k<false, 4>), 4);
    } else {
        if (c.placement) PAMO_SELECT(select_collapse_k<true>, 1);
        else PAMO_SELECT(select_collapse_k<false>, 1);
    }
#undef PAMO_SELECT
    PAMO_TRY(scan_op(CollapseListOp{c.esel, c.recs, c.dcnt}, nE, c.lb, c.stream));
    // cleanup
    PAMO_TRY(scan_op(CleanupOp{m, c.eu, c.ev, nE, nT, c.recs, c.dcnt}, (long long)std::max(nE, nT), c.lb, c.stream,
                     K_LINE_FLAG, 8.0 * nE + 16.0 * nP + 4.0 * nE + 12.0 * nT));
    hipLaunchKernelGGL(line_apply_k, dim3(1), dim3(1), 0, c.stream, m, c.occ, c.ring_off, c.ring, c.eu, c.ev,
                       (const int*)c.recs, c.dcnt, c.changed);
    PAMO_HIP(hipGetLastError());
    // guard + undo loop (:1014-1040).  The guard is queued before the record
    // counts are read back, so the read waits on the guard's kernels instead
    // of stalling the stream; a round without undo candidates ignores its
    // flags.  A tree carried from the last round is refit, otherwise built.
    PAMO_TRY(run_guard(c, nT, c.tree_valid && c.tree_fresh));
    c.tree_valid = true;
    // the first undo candidates follow the guard without a host read: their
    // count comes from the device counters, over a launch bounded by the
    // edge count and by half the points plus 512 (regular collapses share
    // no vertex)
    const long long cand_bound = std::min<long long>(nE, nP / 2 + 512);
    PAMO_TRY(scan_op(UndoListOp{c.recs, c.ring_off, c.ring, c.eu, c.ev, c.flagged, c.pushed, c.dcnt}, cand_bound,
                     c.lb, c.stream, K_UNDO_CAND, 16.0 * cand_bound + 1.0 * nT));
    PAMO_TRY(sync_counters(c));
    c.n_regular = c.hcnt[C_NREG];
    c.n_rec = c.n_regular + c.hcnt[C_LINE];
    c.n_cand = std::min(c.n_rec, (c.n_regular + 511) / 512 * 512);
    c.tree_n = c.hcnt[C_NLIVE];
    PAMO_CHECK(c.n_cand <= cand_bound, "undo candidates beyond their launch bound");
    PAMO_CHECK(!c.exact_guard || c.hcnt[C_UNDECIDED] <= (long long)c.alloc_ulist,
               "exact guard: more pairs left undecided than the list holds");
    int first = c.n_vertices_undo;  // entries already placed by the prologue
    int npushed = c.hcnt[C_PUSHED];
    c.undo_iters = 0;
    c.n_vertices_undo += 2 * npushed;
    int it = 0;
    while (npushed != 0) {
        ++it;
        ++c.undo_iters;
        if (c.ring_lanes == 8)
            LAUNCH(undo_apply_group_k<8>, 8LL * npushed, m, orig_mesh(c), c.occ, c.ring_off, c.ring, c.eu, c.ev,
                   c.pushed, npushed, (const int*)c.pts_map, c.undo_list, first);
        else if (c.ring_lanes == 4)
            LAUNCH(undo_apply_group_k<4>, 4LL * npushed, m, orig_mesh(c), c.occ, c.ring_off, c.ring, c.eu, c.ev,
                   c.pushed, npushed, (const int*)c.pts_map, c.undo_list, first);
        else
            LAUNCH(undo_apply_k, npushed, m, orig_mesh(c), c.occ, c.ring_off, c.ring, c.eu, c.ev, c.pushed, npushed,
                   (const int*)c.pts_map, c.undo_list, first);
        first += 2 * npushed;
        PAMO_TRY(run_guard(c, nT, true));
        PAMO_TRY(undo_candidates(c, npushed));
        c.n_vertices_undo += 2 * npushed;
        if (it == 5) break;
    }
    // new vertex ids
    PAMO_TRY(scan_exclusive_lb(c.occ, c.pts_map, nP + 1, c.lb, c.stream));
    return 0;
}

// Undo-list prologue (cusimp_free.cu:895-916).
static int prologue(Ctx& c, const int* undo_in, int n_undo_in, bool is_stuck) {
    c.n_vertices_undo = 0;
    if (is_stuck) {
        if (n_undo_in > 0)
            LAUNCH(stuck_prologue_k, n_undo_in, undo_in, n_undo_in, c.tmp_undo, c.undo_list, c.pts_map,
                   (int)c.alloc_pts, c.dcnt + C_BAD_UNDO);
        c.n_vertices_undo += n_undo_in;
    } else if (n_undo_in != 0) {
        // a new epoch clears the table
        if (c.inv_epoch == 0xFFFFFFFFu) {
            PAMO_HIP(hipMemsetAsync(c.invalid, 0, sizeof(uint32_t) * c.alloc_pts, c.stream));
            c.inv_epoch = 0;
        }
        ++c.inv_epoch;
    }
    if (n_undo_in != 0)
        LAUNCH(invalid_set_k, n_undo_in, undo_in, n_undo_in, c.invalid, (int)c.alloc_pts, c.inv_epoch);
    return 0;
}

// Drop removed vertices and faces and renumber (pamo/__init__.py:163-167):
// stable compaction by pts_map and by a scan of live faces.
// With carry, the guard tree's leaves are renumbered to the compacted
// faces (carry_leaves_k) and the next round refits it unless fewer than
// tree_keep_pct % of its leaves are still live.
static int compact_round(Ctx& c, int nP, int nT, bool carry = false) {
    Mesh m = cur_mesh(c);
    PAMO_TRY(scan_op(CompactMeshOp{m, nP, nT, c.occ, c.pts_map, c.face_scan, c.np, c.ni, c.nj, c.nk, c.dcnt},
                     (long long)std::max(nP, nT) + 1, c.lb, c.stream, K_COMPACT_MESH,
                     4.0 * nP + 4.0 * nP + 16.0 * nP + 16.0 * nP + 12.0 * nT + 4.0 * nT + 12.0 * nT));
    carry = carry && c.tree_valid && c.tree_keep_pct < 100;
    if (carry) {
        // over every leaf (CompactMeshOp has zeroed the live-leaf count): a carried tree can have more leaves than faces
        LAUNCH(carry_leaves_k, std::max(nT, c.tree_n), c.bval, (const int*)c.ti, (const int*)c.face_scan, nT,
               c.dcnt);
    }
    PAMO_TRY(sync_counters(c));
    c.tree_n = c.hcnt[C_NLIVE];
    if (carry) {
        const long long leaves = c.hcnt[C_NLIVE], live = c.hcnt[C_LIVE_LEAVES];
        c.tree_fresh = live > 1 && live * 100 >= leaves * (long long)c.tree_keep_pct;
        c.tree_valid = c.tree_fresh;
    } else {
        c.tree_valid = c.tree_fresh = false;
    }
    std::swap(c.p, c.np);
    std::swap(c.ti, c.ni);
    std::swap(c.tj, c.nj);
    std::swap(c.tk, c.nk);
    return 0;
}

// thrust::shuffle of the faces with thrust::default_random_engine (:937-941)
static int shuffle_faces(Ctx& c, int nT) {
    Feistel fs = make_feistel((uint64_t)nT);
    const uint32_t domain = 1u << (fs.lbits + fs.rbits);
    PAMO_CHECK((size_t)(2 * domain + 2) <= 4 * c.alloc_tris + 64, "shuffle scratch too small");
    int* flag = c.shuf;
    uint32_t* img = (uint32_t*)(c.shuf + domain + 1);
    LAUNCH(shuffle_flags_k, domain + 1, fs, domain, (uint32_t)nT, flag, img);
    PAMO_TRY(scan_exclusive_lb(flag, flag, domain + 1, c.lb, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.ni, c.ti, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.nj, c.tj, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(c.nk, c.tk, sizeof(int) * nT, hipMemcpyDeviceToDevice, c.stream));
    LAUNCH(shuffle_scatter_k, domain, domain, (uint32_t)nT, flag, img, c.ni, c.nj, c.nk, c.ti, c.tj, c.tk);
    return 0;
}

static int load_mesh(Ctx& c, const float* d_pts, int n_pts, const int* d_tris, int n_tris) {
    hipLaunchKernelGGL(load_points_k, dim3(grid_for(n_pts + 1, kBlock)), dim3(kBlock), 0, c.stream, d_pts, c.p,
                       c.op, c.occ, n_pts);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(load_tris_k, dim3(grid_for(n_tris, kBlock)), dim3(kBlock), 0, c.stream, d_tris, c.ti, c.tj,
                       c.tk, n_tris);
    PAMO_HIP(hipGetLastError());
    return 0;
}

}  // namespace pamo

// ================================================================ C ABI
#include "pamo_hip.h"

using pamo::Ctx;

extern "C" {

const char* pamo_last_error(void) { return pamo::get_error(); }

int pamo_ctx_create(pamo_ctx** out, int device) {
    pamo::clear_error();
    PAMO_CHECK(out != nullptr, "pamo_ctx_create: out is null");
    Ctx* c = new Ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    c->stream = c->own_stream;
    if (e == hipSuccess) e = hipMalloc((void**)&c->dcnt, sizeof(int) * pamo::C_COUNT);
    if (e == hipSuccess) e = hipMemsetAsync(c->dcnt, 0, sizeof(int) * pamo::C_COUNT, c->stream);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->hcnt, sizeof(int) * pamo::C_COUNT, 0);
    if (e != hipSuccess) {
        pamo::set_error(std::string("pamo_ctx_create: ") + hipGetErrorString(e));
        pamo::free_all(*c);
        if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
        delete c;
        return 1;
    }
    *out = reinterpret_cast<pamo_ctx*>(c);
    return 0;
}

int pamo_ctx_destroy(pamo_ctx* ctx) {
    if (!ctx) return 0;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    c->prof.drain();
    pamo::free_all(*c);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return 0;
}

// Runs the context's calls on a caller's stream of the same device (own = 0;
// stream may be NULL, the device's null stream), or again on its own stream
// (own = 1).  The work queued before the switch on the previous stream is
// finished first, so the context's buffers are never used by two streams at
// once.
int pamo_ctx_set_stream(pamo_ctx* ctx, void* stream, int own) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_ctx_set_stream: null context");
    PAMO_CHECK(own == 0 || own == 1, "pamo_ctx_set_stream: own must be 0 or 1");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    hipStream_t next = own ? c.own_stream : reinterpret_cast<hipStream_t>(stream);
    if (next == c.stream) return 0;
    PAMO_HIP(hipSetDevice(c.device));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    c.stream = next;
    return 0;
}

// Returns a context to the state of a freshly created one (logical
// capacities and undo bookkeeping), keeping its device allocations: the next
// call behaves exactly as on a new context.  No counterpart in the
// reference, where a fresh CUDSP_Free is constructed per PaMO object.
int pamo_ctx_set_option(pamo_ctx* ctx, int option, long long value) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_ctx_set_option: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (option == PAMO_OPT_GUARD_SMALL_STACK) {
        PAMO_CHECK(value == 0 || value == 1, "pamo_ctx_set_option: value must be 0 or 1");
        c.guard_small_stack = (int)value;
        return 0;
    }
    if (option == PAMO_OPT_GUARD_SPLIT_FACES) {
        PAMO_CHECK(value >= 0 && value <= 0x7fffffff, "pamo_ctx_set_option: split face count out of range");
        c.guard_split_faces = (int)value;
        return 0;
    }
    if (option == PAMO_OPT_GUARD_SPLIT_LOG) {
        PAMO_CHECK(value >= 0 && value <= 4, "pamo_ctx_set_option: split depth must be in [0, 4]");
        c.guard_split_log = (int)value;
        return 0;
    }
    if (option == PAMO_OPT_RING_LANES) {
        PAMO_CHECK(value == 1 || value == 4 || value == 8, "pamo_ctx_set_option: ring lanes must be 1, 4 or 8");
        c.ring_lanes = (int)value;
        return 0;
    }
    if (option == PAMO_OPT_EDGE_LANES) {
        PAMO_CHECK(value == 1 || value == 4 || value == 8, "pamo_ctx_set_option: edge lanes must be 1, 4 or 8");
        c.edge_lanes = (int)value;
        return 0;
    }
    if (option == PAMO_OPT_GUARD_WAVES) {
        PAMO_CHECK(value == 0 || value == 4, "pamo_ctx_set_option: guard waves must be 0 (compiler's choice) or 4");
        c.guard_waves = (int)value;
        return 0;
    }
    if (option == PAMO_OPT_TREE_KEEP) {
        PAMO_CHECK(value >= 1 && value <= 100, "pamo_ctx_set_option: tree keep percentage must be in [1, 100]");
        c.tree_keep_pct = (int)value;
        return 0;
    }
    if (option == PAMO_OPT_PLACEMENT) {
        PAMO_CHECK(value == 0 || value == 1, "pamo_ctx_set_option: placement must be 0 (midpoint) or 1 (optimal)");
        c.placement = (int)value;
        return 0;
    }
    if (option == PAMO_OPT_WALK_LIMIT) {
        PAMO_CHECK(value >= 0, "pamo_ctx_set_option: walk limit must be >= 0");
        c.walk_cap = value;
        return 0;
    }
    PAMO_CHECK(option == PAMO_OPT_EXACT_GUARD, "pamo_ctx_set_option: unknown option");
    PAMO_CHECK(value >= 0 && value <= 2, "pamo_ctx_set_option: exact guard value must be 0, 1 or 2");
    c.exact_guard = (int)value;
    return 0;
}

int pamo_ctx_reset(pamo_ctx* ctx) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_ctx_reset: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    c.cap_pts = c.cap_tris = c.cap_edges = 0;
    c.n_pts = c.n_tris = c.n_edges = c.n_rec = c.n_regular = c.n_vertices_undo = 0;
    c.tree_valid = c.tree_fresh = false;
    c.tree_n = 0;
    c.undo_iters = 0;
    return 0;
}

int pamo_round(pamo_ctx* ctx, const float* d_pts, int n_pts, const int* d_tris, int n_tris, const int* d_undo,
               int n_undo, float scale, float threshold, int is_stuck, int init, float* d_out_pts, int* d_out_tris,
               int* d_out_occ, int* d_out_map, int* d_out_undo, int undo_cap, int* n_undo_out) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_round: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    pamo::ProfScope ps(c);
    PAMO_CHECK(n_pts > 0 && n_tris > 0, "pamo_round: empty mesh");
    PAMO_CHECK(d_pts && d_tris && d_out_pts && d_out_tris && d_out_occ && d_out_map, "pamo_round: null buffer");
    PAMO_CHECK(n_undo >= 0 && (n_undo == 0 || d_undo != nullptr), "pamo_round: bad undo list");
    PAMO_HIP(hipSetDevice(c.device));
    // the prologue reads pts_map / the invalid table of the previous call
    if (c.alloc_pts > 0) {
        PAMO_CHECK(n_undo <= (int)(4 * c.alloc_pts), "pamo_round: undo list longer than the state can hold");
        PAMO_HIP(hipMemcpyAsync(c.undo_prev, d_undo, sizeof(int) * (size_t)n_undo, hipMemcpyDeviceToDevice,
                                c.stream));
        PAMO_TRY(pamo::prologue(c, c.undo_prev, n_undo, is_stuck != 0));
    } else {
        PAMO_CHECK(n_undo == 0, "pamo_round: the first call of a context takes no undo list");
        c.n_vertices_undo = 0;
    }
    PAMO_TRY(pamo::grow_pts(c, (size_t)n_pts));
    PAMO_TRY(pamo::grow_tris(c, (size_t)n_tris));
    PAMO_TRY(pamo::load_mesh(c, d_pts, n_pts, d_tris, n_tris));
    c.tree_valid = c.tree_fresh = false;  // the caller compacts: no tree carries across calls
    c.tree_n = 0;
    if (init) PAMO_TRY(pamo::shuffle_faces(c, n_tris));
    PAMO_TRY(pamo::round_device(c, n_pts, n_tris, scale, threshold, is_stuck != 0));
    PAMO_CHECK(c.n_vertices_undo <= undo_cap, "pamo_round: undo_cap too small");
    pamo::Mesh m = pamo::cur_mesh(c);
    hipLaunchKernelGGL(pamo::store_points_k, dim3(pamo::grid_for(n_pts, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, (const float4*)c.p, n_pts, d_out_pts);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(pamo::store_tris_k, dim3(pamo::grid_for(n_tris, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, m, n_tris, d_out_tris);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipMemcpyAsync(d_out_occ, c.occ, sizeof(int) * n_pts, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(d_out_map, c.pts_map, sizeof(int) * n_pts, hipMemcpyDeviceToDevice, c.stream));
    if (c.n_vertices_undo > 0) {
        PAMO_CHECK(d_out_undo != nullptr, "pamo_round: null undo output");
        PAMO_HIP(hipMemcpyAsync(d_out_undo, c.undo_list, sizeof(int) * c.n_vertices_undo, hipMemcpyDeviceToDevice,
                                c.stream));
    }
    PAMO_HIP(hipStreamSynchronize(c.stream));
    if (n_undo_out) *n_undo_out = c.n_vertices_undo;
    return 0;
}

int pamo_centroid(pamo_ctx* ctx, const float* d_pts, int n_pts, const int* d_tris, int n_tris, float* out3) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr && out3 != nullptr, "pamo_centroid: null argument");
    PAMO_CHECK(n_pts > 0 && n_tris > 0, "pamo_centroid: empty mesh");
    PAMO_CHECK(d_pts != nullptr && d_tris != nullptr, "pamo_centroid: null buffer");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    PAMO_HIP(hipSetDevice(c.device));
    if (!c.cent_out) PAMO_TRY(pamo::dalloc(&c.cent_out, 4, c.stream));
    const long long pitch = ((long long)n_tris + 3) & ~3LL;
    if (3 * pitch > c.cap_cent) {
        c.cap_cent = 3 * pitch + 3 * pitch / 4;
        PAMO_TRY(pamo::dalloc(&c.cent_means, (size_t)c.cap_cent, c.stream));
    }
    hipLaunchKernelGGL(pamo::face_mean_k, dim3(pamo::grid_for(n_tris, 256)), dim3(256), 0, c.stream, d_pts, d_tris,
                       n_tris, pitch, c.cent_means);
    hipLaunchKernelGGL(pamo::centroid_seq_k, dim3(1), dim3(pamo::kCentThreads), 0, c.stream, c.cent_means, pitch,
                       n_tris, c.cent_out);
    PAMO_HIP(hipGetLastError());
    float h[3];
    PAMO_HIP(hipMemcpyAsync(h, c.cent_out, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    for (int d = 0; d < 3; ++d) out3[d] = h[d];
    return 0;
}

int pamo_prof_enable(pamo_ctx* ctx, unsigned long long kernel_mask) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_prof_enable: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    c.prof.drain();
    c.prof.mask = kernel_mask;
    return 0;
}

int pamo_prof_reset(pamo_ctx* ctx) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_prof_reset: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    c.prof.drain();
    for (int k = 0; k < pamo::K_COUNT; ++k) {
        c.prof.launches[k] = 0;
        c.prof.bytes[k] = 0.0;
        c.prof.ms[k] = 0.0;
    }
    return 0;
}

int pamo_prof_count(void) { return pamo::K_COUNT; }

int pamo_prof_get(pamo_ctx* ctx, int kernel, const char** name, long long* launches, double* total_ms,
                  double* bytes) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_prof_get: null context");
    PAMO_CHECK(kernel >= 0 && kernel < pamo::K_COUNT, "pamo_prof_get: bad kernel id");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    c.prof.drain();
    if (name) *name = pamo::kKernelNames[kernel];
    if (launches) *launches = c.prof.launches[kernel];
    if (total_ms) *total_ms = c.prof.ms[kernel];
    if (bytes) *bytes = c.prof.bytes[kernel];
    return 0;
}

int pamo_round_debug(pamo_ctx* ctx, int what, void* d_out, long long cap, long long* n_out) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_round_debug: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    long long n = 0;
    const void* src = nullptr;
    switch (what) {
        case PAMO_DBG_EDGES: {
            n = c.n_edges;
            PAMO_CHECK(2 * n <= cap, "pamo_round_debug: cap too small");
            if (n > 0) {
                PAMO_HIP(hipMemcpy2DAsync(d_out, 8, c.eu, 4, 4, (size_t)n, hipMemcpyDeviceToDevice, c.stream));
                PAMO_HIP(hipMemcpy2DAsync((char*)d_out + 4, 8, c.ev, 4, 4, (size_t)n, hipMemcpyDeviceToDevice,
                                          c.stream));
            }
            PAMO_HIP(hipStreamSynchronize(c.stream));
            if (n_out) *n_out = n;
            return 0;
        }
        case PAMO_DBG_EDGE_COST: n = c.n_edges; src = c.edge_cost; break;
        case PAMO_DBG_RECORDS: n = c.n_rec; src = c.recs; break;
        case PAMO_DBG_RING_OFF: n = c.n_pts + 1; src = c.ring_off; break;
        case PAMO_DBG_RING: {
            int tot = 0;
            PAMO_HIP(hipMemcpyAsync(&tot, c.ring_off + c.n_pts, sizeof(int), hipMemcpyDeviceToHost, c.stream));
            PAMO_HIP(hipStreamSynchronize(c.stream));
            n = tot;
            src = c.ring;
            break;
        }
        case PAMO_DBG_QUADRICS: {  // (10, n_pts) component-major
            n = 10LL * c.n_pts;
            PAMO_CHECK(n <= cap, "pamo_round_debug: cap too small");
            for (int q = 0; q < 10; ++q) {
                const float* base = (const float*)c.Q + q;  // component q of the 12-float record
                PAMO_HIP(hipMemcpy2DAsync((float*)d_out + (long long)q * c.n_pts, 4, base, 48, 4, (size_t)c.n_pts,
                                          hipMemcpyDeviceToDevice, c.stream));
            }
            PAMO_HIP(hipStreamSynchronize(c.stream));
            if (n_out) *n_out = n;
            return 0;
        }
        case PAMO_DBG_BVH_N: {
            int h = 0;
            PAMO_HIP(hipMemcpyAsync(&h, c.dcnt + pamo::C_NLIVE, sizeof(int), hipMemcpyDeviceToHost, c.stream));
            PAMO_HIP(hipStreamSynchronize(c.stream));
            if (n_out) *n_out = h;
            return 0;
        }
        case PAMO_DBG_BVH_CODES:
        case PAMO_DBG_BVH_FACES:
        case PAMO_DBG_BVH_LEFT:
        case PAMO_DBG_BVH_RIGHT:
        case PAMO_DBG_BVH_BOXES: {
            int h = 0;
            PAMO_HIP(hipMemcpyAsync(&h, c.dcnt + pamo::C_NLIVE, sizeof(int), hipMemcpyDeviceToHost, c.stream));
            PAMO_HIP(hipStreamSynchronize(c.stream));
            if (what == PAMO_DBG_BVH_CODES) { n = h; src = c.bkey; }
            else if (what == PAMO_DBG_BVH_FACES) { n = h; src = c.bval; }
            else if (what == PAMO_DBG_BVH_LEFT) { n = h > 0 ? h - 1 : 0; src = c.bleft; }
            else if (what == PAMO_DBG_BVH_RIGHT) { n = h > 0 ? h - 1 : 0; src = c.bright; }
            else { n = h > 0 ? 8LL * (2 * h - 1) : 0; src = c.bbox; }  // (2n-1) nodes x (lo4, hi4)
            break;
        }
        case PAMO_DBG_UNDO_ITERS: if (n_out) *n_out = c.undo_iters; return 0;
        case PAMO_DBG_N_REGULAR: if (n_out) *n_out = c.n_regular; return 0;
        default: PAMO_CHECK(false, "pamo_round_debug: unknown selector");
    }
    PAMO_CHECK(n <= cap, "pamo_round_debug: cap too small");
    if (n > 0) PAMO_HIP(hipMemcpyAsync(d_out, src, 4 * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    if (n_out) *n_out = n;
    return 0;
}

}  // extern "C"

namespace pamo {

// Largest bounding-box extent of the loaded points (pamo/__init__.py:153,
// :227 for PaSP).
static int input_scale(Ctx& c, int n_verts, float& scale) {
    const int nparts = (int)pamo::grid_for(n_verts, pamo::kBlock);
    hipLaunchKernelGGL(pamo::vert_box_k, dim3(nparts), dim3(pamo::kBlock), 0, c.stream, (const float4*)c.p, n_verts,
                       c.box_part);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(pamo::whole_box_k, dim3(1), dim3(pamo::kBlock), 0, c.stream, c.box_part, nparts);
    PAMO_HIP(hipGetLastError());
    float hb[6];
    PAMO_HIP(hipMemcpyAsync(hb, c.box_part + 6 * nparts, sizeof(hb), hipMemcpyDeviceToHost, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    const float ex = hb[3] - hb[0], ey = hb[4] - hb[1], ez = hb[5] - hb[2];
    scale = ex;
    if (ey > scale) scale = ey;
    if (ez > scale) scale = ez;
    return 0;
}

// The round loop shared by pamo_simplify (PaMO.run's stage 2) and
// pamo_blender_simplify: stop at the face target, at 10 faces when
// small_stop is set (PaMO.run only), or after `tolerance` rounds without
// progress.
static int simplify_loop(Ctx& c, const float* d_verts, int n_verts, const int* d_faces, int n_faces, long long target,
                         bool small_stop, int tolerance, float threshold, long long iters, float* d_out_verts,
                         int* d_out_faces, int* n_verts_out, int* n_faces_out, int* face_hist, int hist_cap,
                         int* rounds_out) {
    PAMO_TRY(pamo::grow_pts(c, (size_t)n_verts));
    PAMO_TRY(pamo::grow_tris(c, (size_t)n_faces));
    PAMO_TRY(pamo::load_mesh(c, d_verts, n_verts, d_faces, n_faces));
    c.tree_valid = c.tree_fresh = false;
    c.tree_n = 0;
    float scale;
    PAMO_TRY(input_scale(c, n_verts, scale));
    int nV = n_verts, nF = n_faces;
    int n_undo = 0, count = 0, rounds = 0;
    bool is_stuck = false;
    for (long long it = 0; it < iters; ++it) {
        const int prev = nF;
        if (rounds == 0) {
            c.n_vertices_undo = 0;
            PAMO_TRY(pamo::shuffle_faces(c, nF));
        } else {
            PAMO_TRY(pamo::prologue(c, c.undo_prev, n_undo, is_stuck));
        }
        PAMO_TRY(pamo::round_device(c, nV, nF, scale, threshold, is_stuck));
        ++rounds;
        n_undo = c.n_vertices_undo;
        std::swap(c.undo_prev, c.undo_list);  // this round's list feeds the next prologue
        PAMO_TRY(pamo::compact_round(c, nV, nF, true));
        nV = c.hcnt[pamo::C_NV];
        nF = c.hcnt[pamo::C_NF];
        if (face_hist && rounds <= hist_cap) face_hist[rounds - 1] = nF;
        if (nF <= target || (small_stop && nF <= 10)) break;
        if (nF == prev) ++count;
        else { count = 0; is_stuck = false; }
        if (count >= 2) is_stuck = true;
        if (count == tolerance) break;
    }
    pamo::Mesh m = pamo::cur_mesh(c);
    hipLaunchKernelGGL(pamo::store_points_k, dim3(pamo::grid_for(nV, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, (const float4*)c.p, nV, d_out_verts);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(pamo::store_tris_k, dim3(pamo::grid_for(nF, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, m, nF, d_out_faces);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipStreamSynchronize(c.stream));
    if (n_verts_out) *n_verts_out = nV;
    if (n_faces_out) *n_faces_out = nF;
    if (rounds_out) *rounds_out = rounds;
    return 0;
}

}  // namespace pamo

extern "C" {

int pamo_simplify(pamo_ctx* ctx, const float* d_verts, int n_verts, const int* d_faces, int n_faces, double ratio,
                  int tolerance, float threshold, long long iters, long long min_verts, float* d_out_verts,
                  int* d_out_faces, int* n_verts_out, int* n_faces_out, int* face_hist, int hist_cap,
                  int* rounds_out) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_simplify: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    pamo::ProfScope ps(c);
    PAMO_CHECK(n_verts > 0 && n_faces > 0, "pamo_simplify: empty mesh");
    PAMO_CHECK(d_verts && d_faces && d_out_verts && d_out_faces, "pamo_simplify: null buffer");
    PAMO_HIP(hipSetDevice(c.device));
    long long target = (long long)(ratio * (double)n_faces);
    if (min_verts > target) target = min_verts;
    return pamo::simplify_loop(c, d_verts, n_verts, d_faces, n_faces, target, true, tolerance, threshold, iters,
                               d_out_verts, d_out_faces, n_verts_out, n_faces_out, face_hist, hist_cap, rounds_out);
}

int pamo_pasp_round(pamo_ctx* ctx, const float* d_pts, int n_pts, const int* d_tris, int n_tris, float scale,
                    float threshold, int init, float* d_out_pts, int* d_out_tris, int* d_out_occ, int* d_out_map) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_pasp_round: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    pamo::ProfScope ps(c);
    PAMO_CHECK(n_pts > 0 && n_tris > 0, "pamo_pasp_round: empty mesh");
    PAMO_CHECK(d_pts && d_tris && d_out_pts && d_out_tris && d_out_occ && d_out_map, "pamo_pasp_round: null buffer");
    PAMO_HIP(hipSetDevice(c.device));
    PAMO_TRY(pamo::grow_pts(c, (size_t)n_pts));
    PAMO_TRY(pamo::grow_tris(c, (size_t)n_tris));
    PAMO_TRY(pamo::load_mesh(c, d_pts, n_pts, d_tris, n_tris));
    c.tree_valid = c.tree_fresh = false;  // the caller compacts: no tree carries across calls
    c.tree_n = 0;
    if (init) PAMO_TRY(pamo::shuffle_faces(c, n_tris));
    PAMO_TRY(pamo::pasp_round_device(c, n_pts, n_tris, scale, threshold));
    pamo::Mesh m = pamo::cur_mesh(c);
    hipLaunchKernelGGL(pamo::store_points_k, dim3(pamo::grid_for(n_pts, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, (const float4*)c.p, n_pts, d_out_pts);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(pamo::store_tris_k, dim3(pamo::grid_for(n_tris, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, m, n_tris, d_out_tris);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipMemcpyAsync(d_out_occ, c.occ, sizeof(int) * n_pts, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipMemcpyAsync(d_out_map, c.pts_map, sizeof(int) * n_pts, hipMemcpyDeviceToDevice, c.stream));
    PAMO_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

int pamo_pasp_run(pamo_ctx* ctx, const float* d_verts, int n_verts, const int* d_faces, int n_faces, float threshold,
                  long long iters, float* d_out_verts, int* d_out_faces, int* n_verts_out, int* n_faces_out,
                  int* rounds_out) {
    pamo::clear_error();
    PAMO_CHECK(ctx != nullptr, "pamo_pasp_run: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    pamo::ProfScope ps(c);
    PAMO_CHECK(n_verts > 0 && n_faces > 0, "pamo_pasp_run: empty mesh");
    PAMO_CHECK(d_verts && d_faces && d_out_verts && d_out_faces, "pamo_pasp_run: null buffer");
    PAMO_HIP(hipSetDevice(c.device));
    PAMO_TRY(pamo::grow_pts(c, (size_t)n_verts));
    PAMO_TRY(pamo::grow_tris(c, (size_t)n_faces));
    PAMO_TRY(pamo::load_mesh(c, d_verts, n_verts, d_faces, n_faces));
    float scale;
    PAMO_TRY(pamo::input_scale(c, n_verts, scale));
    // PaSP.run (pamo/__init__.py:224-253): shuffle on the first round, compact
    // after each, stop when a round removes no face
    int nV = n_verts, nF = n_faces, rounds = 0;
    for (long long it = 0; it < iters; ++it) {
        const int prev = nF;
        if (rounds == 0) PAMO_TRY(pamo::shuffle_faces(c, nF));
        PAMO_TRY(pamo::pasp_round_device(c, nV, nF, scale, threshold));
        ++rounds;
        PAMO_TRY(pamo::compact_round(c, nV, nF));
        nV = c.hcnt[pamo::C_NV];
        nF = c.hcnt[pamo::C_NF];
        if (nF == prev) break;
    }
    pamo::Mesh m = pamo::cur_mesh(c);
    hipLaunchKernelGGL(pamo::store_points_k, dim3(pamo::grid_for(nV, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, (const float4*)c.p, nV, d_out_verts);
    PAMO_HIP(hipGetLastError());
    hipLaunchKernelGGL(pamo::store_tris_k, dim3(pamo::grid_for(nF, pamo::kBlock)), dim3(pamo::kBlock), 0,
                       c.stream, m, nF, d_out_faces);
    PAMO_HIP(hipGetLastError());
    PAMO_HIP(hipStreamSynchronize(c.stream));
    if (n_verts_out) *n_verts_out = nV;
    if (n_faces_out) *n_faces_out = nF;
    if (rounds_out) *rounds_out = rounds;
    return 0;
}

int pamo_blender_simplify(pamo_ctx* ctx, const float* in_verts, const int* in_tris, float* out_verts, int* out_tris,
                          int* in_out_nv_ni, float threshold) {
    pamo::clear_error();
    PAMO_CHECK(ctx && in_verts && in_tris && out_verts && out_tris && in_out_nv_ni,
               "pamo_blender_simplify: null argument");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    const int nv = in_out_nv_ni[0], nt = in_out_nv_ni[1];
    PAMO_CHECK(nv > 0 && nt > 0, "pamo_blender_simplify: empty mesh");
    PAMO_HIP(hipSetDevice(c.device));
    for (long long i = 0; i < 3LL * nt; ++i)
        PAMO_CHECK(in_tris[i] >= 0 && in_tris[i] < nv, "pamo_blender_simplify: face index outside the vertices");
    // host arrays staged through device buffers on the context's stream
    float *dv = nullptr, *dvo = nullptr;
    int *dt = nullptr, *dto = nullptr;
    auto release = [&]() {
        void* ptrs[] = {dv, dvo, dt, dto};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
    };
    hipError_t e = hipMalloc((void**)&dv, sizeof(float) * 3 * (size_t)nv);
    if (e == hipSuccess) e = hipMalloc((void**)&dvo, sizeof(float) * 3 * (size_t)nv);
    if (e == hipSuccess) e = hipMalloc((void**)&dt, sizeof(int) * 3 * (size_t)nt);
    if (e == hipSuccess) e = hipMalloc((void**)&dto, sizeof(int) * 3 * (size_t)nt);
    if (e == hipSuccess) e = hipMemcpyAsync(dv, in_verts, sizeof(float) * 3 * (size_t)nv, hipMemcpyHostToDevice, c.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, in_tris, sizeof(int) * 3 * (size_t)nt, hipMemcpyHostToDevice, c.stream);
    if (e != hipSuccess) {
        release();
        pamo::set_error(std::string("pamo_blender_simplify: ") + hipGetErrorString(e));
        return 1;
    }
    int onv = 0, onf = 0;
    // cusimpinterface.cu:57-161: no centring, no face target, stop after four
    // rounds without progress (or 1e6 rounds)
    int rc;
    {
        pamo::ProfScope ps(c);
        rc = pamo::simplify_loop(c, dv, nv, dt, nt, 0, false, 4, threshold, 1000000, dvo, dto, &onv, &onf, nullptr,
                                 0, nullptr);
    }
    if (rc == 0) e = hipMemcpyAsync(out_verts, dvo, sizeof(float) * 3 * (size_t)onv, hipMemcpyDeviceToHost, c.stream);
    if (rc == 0 && e == hipSuccess)
        e = hipMemcpyAsync(out_tris, dto, sizeof(int) * 3 * (size_t)onf, hipMemcpyDeviceToHost, c.stream);
    if (rc == 0 && e == hipSuccess) e = hipStreamSynchronize(c.stream);
    release();
    if (rc != 0) return rc;
    if (e != hipSuccess) {
        pamo::set_error(std::string("pamo_blender_simplify: ") + hipGetErrorString(e));
        return 1;
    }
    in_out_nv_ni[0] = onv;
    in_out_nv_ni[1] = onf;
    return 0;
}

}  // extern "C"
