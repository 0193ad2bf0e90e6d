/* C ABI of libpamo_hip.so: PaMO stage-2 simplification on MI355X (gfx950).
 *
 * All mesh pointers are DEVICE pointers (e.g. torch.Tensor.data_ptr() of a
 * contiguous CUDA tensor).  Positions are float32 (n, 3), faces int32 (n, 3).
 * Every call returns 0 on success and a nonzero status otherwise; the message
 * is available from pamo_last_error() on the calling thread.  A context owns
 * persistent device state and one HIP stream; it is not thread safe and must
 * not be shared between threads.
 */
#ifndef PAMO_HIP_H
#define PAMO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pamo_ctx pamo_ctx;

/* Replaces the construction of cusimp_free::CUDSP_Free / CUSimp_Free
 * (reference simp_cuda/src/pybind.cpp:16-49, cusimp_free.h:184-280). */
int pamo_ctx_create(pamo_ctx** out, int device);
int pamo_ctx_destroy(pamo_ctx* ctx);

/* Returns a context to the state of a freshly created one, keeping its
 * device allocations and the options set with pamo_ctx_set_option (no
 * counterpart in the reference, which constructs a new CUDSP_Free per PaMO
 * object, simp_cuda/pamo/__init__.py:21). */
int pamo_ctx_reset(pamo_ctx* ctx);

/* Stream of the context's calls (the reference runs its kernels on the
 * current CUDA stream, cusimp_free.cu; here a context has its own
 * non-blocking stream by default).  own = 0 runs the following calls on
 * `stream`, a hipStream_t of the context's device (NULL: its null stream),
 * so they are ordered after the caller's earlier work without a host
 * synchronisation; own = 1 returns to the context's stream.  Work queued on
 * the previous stream is finished first.  The calls still return when
 * their results are ready. */
int pamo_ctx_set_stream(pamo_ctx* ctx, void* stream, int own);

/* Test option (no counterpart in the reference): with value 1 the guard
 * query runs with a 32-entry frontier per wave, so the exact overflow walk
 * of the frontier traversal is exercised; 0 restores the default.  The
 * result is identical either way. */
#define PAMO_OPT_GUARD_SMALL_STACK 1
/* Guard predicate (no counterpart in the reference, whose guard predicate
 * is not exact): 0 (default) the reference's predicate and skip rules,
 * bit-exact with the reference; 1 exact -- a candidate pair is flagged
 * exactly when the two faces intersect beyond their shared vertices, decided
 * with FP64-filtered orientation predicates and an exact pass for the pairs
 * the filter cannot decide; 2 as 1 with every pair sent through the exact
 * pass (test setting).  A context keeps the setting across calls. */
#define PAMO_OPT_EXACT_GUARD 2
/* Test option (no counterpart in the reference): every tree walk of the
 * guard stops after its bound -- each node once for an escape-link walk,
 * (queries x nodes) frontier steps -- sets an error flag and makes the call
 * fail with "a tree walk exceeded its bound" instead of looping on a
 * corrupted tree.  A value > 0 lowers both bounds to it, so that tests reach
 * that error path on a valid tree; 0 (default) keeps the bounds. */
#define PAMO_OPT_WALK_LIMIT 3
/* Collapse position (no counterpart in the reference, which always collapses
 * an edge to its midpoint, simp_cuda/src/cusimp_free.cu:491,687): 0 (default)
 * the midpoint, bit-exact with the reference; 1 optimal placement -- the
 * minimiser of the summed quadric Q_u + Q_v from its 3x3 system solved in
 * FP64, falling back to the midpoint when the system is singular or the
 * minimiser lies farther from the midpoint than the edge is long.  The cost,
 * flip and skinny tests and the collapse all use the chosen position.  A
 * context keeps the setting across calls. */
#define PAMO_OPT_PLACEMENT 4
/* Guard tree reuse (no counterpart in the reference, which builds an LBVH
 * at every check, simp_cuda/src/bvh/self_intersect.cuh:151-392): the fused
 * loop keeps the guard tree of one round for the next ones -- its leaves are
 * renumbered at each compaction, removed faces keep empty boxes, and each
 * round refits the boxes -- and builds a new one once fewer than this
 * percentage of its leaves are live (default 80; 100 builds every round).
 * The flags, hence every result, do not depend on the tree. */
#define PAMO_OPT_TREE_KEEP 5
/* Guard launch shape (no counterpart in the reference; results do not
 * change): on meshes of at most PAMO_OPT_GUARD_SPLIT_FACES faces (default
 * 65536) each group of queries is served by 2^PAMO_OPT_GUARD_SPLIT_LOG
 * waves (default 3, at most 4), one per subtree under the root, so that one
 * query with a large box does not hold a late-round launch in one wave. */
#define PAMO_OPT_GUARD_SPLIT_FACES 6
#define PAMO_OPT_GUARD_SPLIT_LOG 7
/* Guard register budget (no counterpart in the reference; results do not
 * change): 0 runs the guard kernel as the compiler allocates its registers
 * (129 vector registers, 3 waves per SIMD), 4 (default) a variant capped at
 * 128 so that 4 waves fit per SIMD. */
#define PAMO_OPT_GUARD_WAVES 8
/* Edge-cost launch shape (no counterpart in the reference; results do not
 * change): 1 lane per edge, or 4 (default) / 8 lanes sharing one edge's
 * ring faces and adding their terms in the single-lane order. */
#define PAMO_OPT_EDGE_LANES 9
/* Ring-walk launch shape (no counterpart in the reference; results do not
 * change): 1 lane per vertex / edge / record in the quadric,
 * selection-and-collapse and undo kernels, or 4 / 8 (default) lanes
 * splitting the one-ring faces (float sums kept in the single-lane order). */
#define PAMO_OPT_RING_LANES 10
int pamo_ctx_set_option(pamo_ctx* ctx, int option, long long value);

/* Last error message of the calling thread ("" if none).
 * Replaces the reference's printed CHECK_CUDA messages (cusimp_free.cu:17-27). */
const char* pamo_last_error(void);

/* One collapse round: CUDSP_Free.forward (reference simp_cuda/src/pybind.cpp:52-103,
 * simp_cuda/src/cusimp_free.cu:891-1048).
 *   inputs : d_pts (n_pts,3) f32, d_tris (n_tris,3) i32, d_undo (n_undo,) i32,
 *            scale, threshold, is_stuck, init (shuffle faces first).
 *   outputs: d_out_pts (n_pts,3), d_out_tris (n_tris,3) with removed rows = -1,
 *            d_out_occ (n_pts,), d_out_map (n_pts,), d_out_undo (<= undo_cap),
 *            *n_undo_out entries of d_out_undo. */
int pamo_round(pamo_ctx* ctx, const float* d_pts, int n_pts, const int* d_tris, int n_tris, const int* d_undo,
               int n_undo, float scale, float threshold, int is_stuck, int init, float* d_out_pts, int* d_out_tris,
               int* d_out_occ, int* d_out_map, int* d_out_undo, int undo_cap, int* n_undo_out);

/* Centring offset of PaMO.run, tris_mean = points[triangles].mean(axis=1)
 * .mean(axis=0) on float32 numpy arrays (reference
 * simp_cuda/pamo/__init__.py:60-72,144), bit for bit: per-face
 * ((a + b) + c) / 3, then a sequential float32 running sum over the faces in
 * order, divided by n_tris.  Face indices must lie in [0, n_pts).  out3 is a
 * host float[3]; the call returns when it is written. */
int pamo_centroid(pamo_ctx* ctx, const float* d_pts, int n_pts, const int* d_tris, int n_tris, float* out3);

/* Intermediate arrays of the last pamo_round (for parity tests). */
#define PAMO_DBG_EDGES 1      /* int32 (n_edges, 2)              */
#define PAMO_DBG_EDGE_COST 2  /* uint32 (n_edges,)               */
#define PAMO_DBG_RECORDS 3    /* int32 collapsed-edge records    */
#define PAMO_DBG_RING_OFF 4   /* int32 (n_pts + 1,)              */
#define PAMO_DBG_RING 5       /* int32 (ring_off[n_pts],)        */
#define PAMO_DBG_QUADRICS 6   /* float32 (10, n_pts) component-major */
#define PAMO_DBG_UNDO_ITERS 7 /* count only                      */
#define PAMO_DBG_N_REGULAR 8  /* count only                      */
#define PAMO_DBG_BVH_N 9      /* count only: leaves of the last round's guard LBVH (the
                                 faces live at the start of that round; a face removed
                                 since keeps its leaf, with an empty box) */
#define PAMO_DBG_BVH_CODES 10 /* uint32 Morton codes, sorted      */
#define PAMO_DBG_BVH_FACES 11 /* int32 face per sorted leaf       */
#define PAMO_DBG_BVH_LEFT 12  /* int32 left child of internal node */
#define PAMO_DBG_BVH_RIGHT 13 /* int32 right child                */
#define PAMO_DBG_BVH_BOXES 14 /* float32 (2n-1, 8): lo xyz0, hi xyz0 */
int pamo_round_debug(pamo_ctx* ctx, int what, void* d_out, long long cap, long long* n_out);

/* The whole stage-2 loop of PaMO.run on already-centred vertices
 * (reference simp_cuda/pamo/__init__.py:147-185): rounds run on the device,
 * compaction included, until n_faces <= max(int(ratio * n_faces), min_verts),
 * n_faces <= 10, or `tolerance` rounds without progress.  Outputs must hold
 * the input sizes.  face_hist (host, hist_cap ints, may be NULL) receives the
 * face count after each round. */
int pamo_simplify(pamo_ctx* ctx, const float* d_verts, int n_verts, const int* d_faces, int n_faces, double ratio,
                  int tolerance, float threshold, long long iters, long long min_verts, float* d_out_verts,
                  int* d_out_faces, int* n_verts_out, int* n_faces_out, int* face_hist, int hist_cap,
                  int* rounds_out);

/* Legacy simplifier behind PaSP (reference simp_cuda/src/cusimp.cu,
 * simp_cuda/pamo/__init__.py:208-253): midpoint collapses of edges whose
 * legacy cost (cusimp.cu:283-482) is at most the threshold, chosen as
 * independent sets; no intersection guard and no undo. */
/* One round: cusimp::CUDSP.forward (pybind.cpp:140-183, cusimp.cu:587-720);
 * outputs as pamo_round's (points (n_pts,3) with removed ones zeroed, faces
 * (n_tris,3) with removed rows -1, occ and map (n_pts,)). */
int pamo_pasp_round(pamo_ctx* ctx, const float* d_pts, int n_pts, const int* d_tris, int n_tris, float scale,
                    float threshold, int init, float* d_out_pts, int* d_out_tris, int* d_out_occ, int* d_out_map);
/* PaSP.run (pamo/__init__.py:224-253) on the device: scale = largest
 * bounding-box extent, rounds with compaction until one removes no face or
 * `iters` rounds. */
int pamo_pasp_run(pamo_ctx* ctx, const float* d_verts, int n_verts, const int* d_faces, int n_faces, float threshold,
                  long long iters, float* d_out_verts, int* d_out_faces, int* n_verts_out, int* n_faces_out,
                  int* rounds_out);

/* The loop of the reference's Blender library (pamo_blender/cusimpinterface.cu:57-161)
 * on host arrays: in_out_nv_ni = {n_v, n_t} in, result counts out; no
 * centring, no face target, stop after four rounds without progress.  The
 * context keeps its persistent state across calls as the library's static
 * CUSimp_Free does.  libcusimp.so exports it under the add-on's name
 * pamo_simplify (include/pamo_blender.h). */
int pamo_blender_simplify(pamo_ctx* ctx, const float* in_verts, const int* in_tris, float* out_verts, int* out_tris,
                          int* in_out_nv_ni, float threshold);

/* Launch profiler (no counterpart in the reference, which only prints host
 * wall clock, simp_cuda/pamo/__init__.py:139-148,187-190): HIP events are
 * recorded around launches of the kernels selected by kernel_mask (bit k =
 * kernel id k); per kernel, pamo_prof_get returns the launch count, summed
 * device time and summed algorithmic bytes (compulsory HBM traffic). */
int pamo_prof_enable(pamo_ctx* ctx, unsigned long long kernel_mask);
int pamo_prof_reset(pamo_ctx* ctx);
int pamo_prof_count(void);
int pamo_prof_get(pamo_ctx* ctx, int kernel, const char** name, long long* launches, double* total_ms,
                  double* bytes);

/* ---------------------------------------------------------------- stage 3
 * Barrier-energy safe projection (reference
 * simp_cuda/safe_project/src/pamo_safe_project/system.py:20-470).  All
 * positions live in the normalised frame x' = x * scale + t, with scale and t
 * computed by the caller as processing.py:15-52 does. */
typedef struct pamo_s3 pamo_s3;

/* Energy terms, Stage3Config.energy_calcs (reference config.py:29-41). */
#define PAMO_S3_E_M2GT 1       /* Mesh2GTDistanceBvhEnergyCalculator (energy.py:346-468) */
#define PAMO_S3_E_ELASTIC 2    /* ElasticEnergyCalculator (energy.py:470-588)            */
#define PAMO_S3_E_HINGE 4      /* HingeEnergyCalculator (energy.py:590-706)              */
#define PAMO_S3_E_COLLISION 8  /* CollisionBvhEnergyCalculator (energy.py:708-948)       */
#define PAMO_S3_E_DENSE 16     /* DummyEnergyCalculator (energy.py:62-111), tests only   */
#define PAMO_S3_E_DEFAULT 15

/* Fields of Stage3Config (reference config.py:6-80) that the path reads. */
typedef struct pamo_s3_params {
    int energy_mask;         /* PAMO_S3_E_* bits; PAMO_S3_E_DEFAULT */
    int n_newton_iters;      /* 10  */
    int n_cg_iters;          /* 40  */
    int n_line_search_iters; /* 10  */
    int max_blocks;          /* 1 << 25: contact-pair cap before the radius halves */
    int ccd_max_iters;       /* 100 */
    double d_hat;            /* 1e-3 */
    double coll_stiffness;   /* 1e2  */
    double contact_detection_radius; /* 1e-2 */
    double ee_classify_thres;        /* 1e-3 */
    double ccd_slackness;            /* 0.7  */
    double ccd_thickness;            /* 1e-6 */
    double elas_young_modulus;       /* 1e-1 */
    double elas_poisson_ratio;       /* 0.0  */
    double hinge_stiffness;          /* 1e-2 */
    double mesh2gt_dist_stiffness;   /* 1e3  */
    double min_angle_thres;          /* 0.0: a mesh whose smallest angle (degrees) is at most this
                                        is not moved (system.py:217-222) */
} pamo_s3_params;

/* Replaces Stage3System.__init__ (system.py:20-70). */
int pamo_s3_create(pamo_s3** out, int device, const pamo_s3_params* params);
int pamo_s3_destroy(pamo_s3* s3);

/* Input ("ground truth") surface: device float32 (n_gt_v,3), int32
 * (n_gt_f,3), normalised in FP64 with scale / t3; builds its LBVH.
 * Replaces the input-surface half of Stage3System.register_mesh
 * (system.py:175-205,266-269: GT arrays and the wp.Mesh over them). */
int pamo_s3_set_target(pamo_s3* s3, const float* d_gt_v, int n_gt_v, const int* d_gt_f, int n_gt_f, double scale,
                       const double* t3);

/* Simplified mesh to project: device float32 (n_v,3) / int32 (n_f,3); host
 * int32 unique edges (n_e,2), hinges (n_h,4: opposite, edge a, edge b,
 * opposite) and per-vertex Voronoi areas of the rest shape (n_v floats).
 * Rest-shape terms (membrane Dm^-1 and areas, hinge rest angles and edge
 * lengths) are computed on the device.  Counts of terms that the energy mask
 * leaves out may be 0.  Replaces the rest of Stage3System.register_mesh
 * (system.py:167-272). */
int pamo_s3_set_mesh(pamo_s3* s3, const float* d_v, int n_v, const int* d_f, int n_f, double scale, const double* t3,
                     const int* h_edges, int n_e, const int* h_hinges, int n_h, const float* h_voronoi,
                     int do_nothing);

/* The connectivity part of Stage3System.register_mesh done on the host
 * (system.py:175-228: trimesh edges_unique, the hinge search of
 * hinge_energy.py:21-57, igl Voronoi mass matrix, igl internal angles):
 * host float32 (n_v,3) / int32 (n_f,3) in, the normalisation (scale, t3)
 * applied in FP64.  Outputs sized 2*(3*n_f) edges ints, 4*(3*n_f) hinge ints
 * and n_v floats; n_e / n_h receive the counts. */
int pamo_s3_prepare_host(const float* h_v, int n_v, const int* h_f, int n_f, double scale, const double* t3,
                         int* h_edges, int* n_e, int* h_hinges, int* n_h, float* h_voronoi, double* min_angle_deg);

/* processing.py:15-20 get_normalization_transform for device float32
 * (n_v,3) vertices, reduced in FP64: scale = 1 / max extent, t = -mean * scale. */
int pamo_s3_normalization(pamo_s3* s3, const float* d_v, int n_v, double* scale, double* t3);

/* pamo_s3_set_mesh with the connectivity prepared by pamo_s3_prepare_host
 * from the device mesh (the whole of register_mesh for the mesh to project,
 * including the min-angle do_nothing rule, system.py:220-224). */
int pamo_s3_set_mesh_device(pamo_s3* s3, const float* d_v, int n_v, const int* d_f, int n_f, double scale,
                            const double* t3);

/* The stage-2 -> stage-3 handoff trimesh.Trimesh(vertices, faces) with its
 * default process=True (pamo/__init__.py:194; trimesh 4.4.0
 * Trimesh.process -> grouping.merge_vertices): vertices referenced by a face
 * whose coordinates times 10^8, rounded half to even, are equal merge into
 * the first of them; the kept vertices, in index order, go to d_v_out
 * (float32 (n_v,3), not aliasing d_v) and *out_n_v receives their count;
 * the int32 (n_f,3) faces d_f are re-indexed in place.  Unreferenced
 * vertices are dropped.  A non-finite referenced vertex is an error (trimesh
 * would drop it and point its faces at vertex 0).  Runs on the stage-3
 * context's stream and returns when done. */
int pamo_s3_merge_vertices(pamo_s3* s3, const float* d_v, int n_v, int* d_f, int n_f, float* d_v_out,
                           int* out_n_v);

/* n_steps calls of Stage3System.step (system.py:411-470): closest-point
 * targets, then n_newton_iters of {contacts, gradient, Jacobi PCG, clamp,
 * ACCD, halving line search}. */
int pamo_s3_step(pamo_s3* s3, int n_steps);
/* Replaces Stage3System.step(eval=True, ...) (system.py:411-470): one step,
 * calling on_newton(newton_iter, user) after each Newton iteration once its
 * device work has finished, so the caller can read the vertices and
 * evaluate them (system.py:466-470). */
typedef void (*pamo_s3_newton_fn)(int newton_iter, void* user);
int pamo_s3_step_eval(pamo_s3* s3, pamo_s3_newton_fn on_newton, void* user);

/* Current positions mapped back to the input frame, (x' - t) / scale, into a
 * device float32 (n_v,3) buffer (processing.py:96-104). */
int pamo_s3_get_vertices(pamo_s3* s3, float* d_out, double scale, const double* t3);

/* Current normalised positions as host float64 (n_v,3) (tests). */
int pamo_s3_get_normalized(pamo_s3* s3, double* h_out);

/* Dense quadratic energy 0.5 x^T A x + b^T x + c over the 3 * n_v
 * coordinates (A row-major FP64, 3n_v x 3n_v, n_v <= 1024): the reference's
 * DummyEnergyCalculator (energy.py:62-111), used by its CG test
 * (simp_cuda/safe_project/tests/test_dummy_energy.py). */
int pamo_s3_set_dense(pamo_s3* s3, const double* h_A, const double* h_b, double c, int n_v);

/* Test access to single parts of a step (no counterpart in the reference;
 * each mirrors one Stage3System method).  h_in / h_out are host FP64. */
#define PAMO_S3_DBG_SET_Q 1    /* in: q (n_v,3) normalised positions                  */
#define PAMO_S3_DBG_TARGET 2   /* _update_vertex_target; out: target (n_v,3)          */
#define PAMO_S3_DBG_CONTACTS 3 /* _detect_contact; out: (n_c,5) i0 i1 i2 i3 type      */
#define PAMO_S3_DBG_ENERGY 4   /* detect + _compute_energy; out: 6 = per-term (M2GT,
                                  elastic, hinge, collision, dense) and the total      */
#define PAMO_S3_DBG_DIFF 5     /* detect + _compute_diff; out: grad (n_v,3), diag (n_v,3) */
#define PAMO_S3_DBG_HVP 6      /* detect + diff + _compute_hess_dx; in: dx (n_v,3);
                                  out: H dx (n_v,3)                                    */
#define PAMO_S3_DBG_SOLVE 7    /* detect + diff + CGSolver.solve; out: p (n_v,3)     */
#define PAMO_S3_DBG_CCD 8      /* detect + _clamp_p + _ccd on in: p (n_v,3);
                                  out: clamped p (n_v,3) then ccd_step                 */
#define PAMO_S3_DBG_HINGE 9    /* after DIFF: projected hinge blocks (n_h,144)        */
#define PAMO_S3_DBG_RESOLVE 10 /* CGSolver.solve again on the gradient and diagonal of
                                  the last DIFF / SOLVE; out: p (n_v,3)              */
int pamo_s3_debug(pamo_s3* s3, int what, const double* h_in, long long n_in, double* h_out, long long cap,
                  long long* n_out);

/* Statistics (no counterpart in the reference): [0] last total energy,
 * [1] contact pairs summed over Newton iterations, [2] Newton iterations,
 * [3] CG iterations, [4] line-search evaluations, [5] current detection
 * radius, [6] do_nothing flag, [7] number of hinges. */
int pamo_s3_stats(pamo_s3* s3, double* out8);
/* Active contacts (distance below d_hat at the start of a Newton iteration)
 * summed over the Newton iterations since the mesh was set: the pairs the
 * gradient pass and the CG products touch (measurement; no counterpart in
 * the reference). */
int pamo_s3_active_contacts(pamo_s3* s3, long long* total);

/* Run-time options (no counterpart in the reference's interface, which
 * fixes its CUDA-graph capture of the PCG loop in config.py:21).
 * PAMO_S3_OPT_CG_GRAPH: 1 captures each solve's PCG kernels from the
 * stream into a HIP graph, updated in place between solves, and launches it
 * as one unit; 0 (default) launches them one by one.  Results do not
 * change. */
#define PAMO_S3_OPT_CG_GRAPH 1
int pamo_s3_set_option(pamo_s3* s3, int option, long long value);

/* ---------------------------------------------------------------- stage 1 (remeshing)
 * Replaces PaMO.remesh and its inputs (simp_cuda/pamo/__init__.py:60-120,
 * 130-141): the triangles normalised into the unit cube as
 * preprocess_mesh does, a narrow-band distance field on a resolution^3
 * grid (band 3 / resolution, in place of torchcumesh2sdf.get_sdf), signed
 * by a flood fill of the outside from the grid border, and the level
 * 0.9 / resolution extracted as a closed manifold triangle mesh by
 * marching cubes with one face rule for every case (in place of pdmc.DMC
 * and the quad split).  The reference picks resolution 256, 128 when the
 * target is at most 1000 faces and 64 at most 50. */
typedef struct pamo_s1 pamo_s1;
int pamo_s1_create(pamo_s1** out, int device);
int pamo_s1_destroy(pamo_s1* s1);
/* Device float32 points (n_points,3) and int32 faces (n_faces,3) with
 * indices already checked; mean3 (host) is the centring offset of PaMO.run
 * (pamo_centroid).  Returns the output sizes; the mesh stays in the context
 * until pamo_s1_get. */
int pamo_s1_remesh(pamo_s1* s1, const float* d_points, int n_points, const int* d_faces, int n_faces,
                   const float* mean3, int resolution, int* n_verts_out, int* n_faces_out);
/* The remeshed vertices, in the centred frame of PaMO.run (as remesh
 * returns them, pamo/__init__.py:87), and faces, into device buffers of the
 * sizes returned. */
int pamo_s1_get(pamo_s1* s1, float* d_verts, int* d_faces);
/* Test access (no counterpart in the reference). */
#define PAMO_S1_DBG_FRAME 1       /* 8 floats: mean xyz, min xyz, max, level           */
#define PAMO_S1_DBG_CASES 2       /* 256 x 18 int8: crossing edges in patch cycle order (12, -1 padded),
                                     edges per patch (4), patches, triangles       */
#define PAMO_S1_DBG_FIELD 3       /* resolution^3 uint32: distance bits (+inf beyond the band) */
#define PAMO_S1_DBG_OUTSIDE 4     /* resolution^3 uint8: 1 for the outside              */
#define PAMO_S1_DBG_NORMALIZED 5  /* n_faces x 3 x 4 float: normalised corners (w = 0)   */
#define PAMO_S1_DBG_VERTS 6       /* n_verts x 4 float: vertices in sample index units   */
int pamo_s1_debug(pamo_s1* s1, int what, void* h_out, long long cap, long long* n_out);

/* ---------------------------------------------------------------- batched and served forms
 * No counterpart in the reference, which runs one mesh per PaMO.run call
 * (simp_cuda/pamo/__init__.py:122-205).  Each job is that whole call --
 * centring, the stage-2 round loop, and stage 3 against the job's own mesh
 * when requested -- and gives the same result as running it alone.  Jobs run
 * concurrently on a pool of n_slots slots (each: stage-2 and stage-3
 * contexts with their own HIP streams, and a worker thread). */

/* A batch in CSR mesh-of-meshes layout: device d_verts (sum V,3) f32 and
 * d_faces (sum F,3) i32 with face indices local to each mesh; host
 * vert_off / face_off (n_meshes+1) int64 offsets (in vertices / faces);
 * host per-mesh ratio and use_stage3.  Outputs go to d_out_verts /
 * d_out_faces at the input offsets (capacity = input sizes), counts to host
 * out_nv / out_nf; *seconds receives the wall time.  s3 may be NULL when no
 * mesh uses stage 3.  exact_guard is the PAMO_OPT_EXACT_GUARD value of every
 * slot's stage-2 context.  Meshes are started longest first.  on_done (may be
 * NULL) is called from a worker thread as each mesh finishes, with the
 * mesh's index, its status (0 = ok) and user. */
typedef void (*pamo_batch_done_fn)(int mesh, int status, void* user);
int pamo_batch_run(int device, int n_slots, const pamo_s3_params* s3, int n_meshes, const float* d_verts,
                   const int* d_faces, const long long* vert_off, const long long* face_off, const double* ratio,
                   const int* use_stage3, int tolerance, float threshold, long long iters, long long min_verts,
                   int exact_guard, float* d_out_verts, int* d_out_faces, int* out_nv, int* out_nf,
                   double* seconds, pamo_batch_done_fn on_done, void* user);

/* Request loop: jobs are admitted in arrival order when a slot is free and
 * the estimated HBM footprint of the running jobs plus the new one fits
 * hbm_budget_bytes (<= 0: 200 GiB); a job larger than the budget runs alone.
 * exact_guard is the PAMO_OPT_EXACT_GUARD value of every slot.  The caller
 * keeps a job's inputs alive until it is done. */
typedef struct pamo_server pamo_server;
int pamo_serve_create(pamo_server** out, int device, int n_slots, long long hbm_budget_bytes,
                      const pamo_s3_params* s3, int exact_guard);
int pamo_serve_destroy(pamo_server* srv);
int pamo_serve_submit(pamo_server* srv, const float* d_verts, int n_verts, const int* d_faces, int n_faces,
                      double ratio, int use_stage3, int tolerance, float threshold, long long iters,
                      long long min_verts, long long* job_id);
/* state: 0 queued, 1 running, 2 done, 3 failed; wait != 0 blocks until done.
 * latency_s = completion - submission; queue_s = start - submission. */
int pamo_serve_poll(pamo_server* srv, long long job_id, int wait, int* state, int* n_verts_out, int* n_faces_out,
                    double* latency_s, double* queue_s);
/* Copies a finished job's output (n_verts_out x 3 f32, n_faces_out x 3 i32)
 * to device buffers and releases the job. */
int pamo_serve_fetch(pamo_server* srv, long long job_id, float* d_out_verts, int* d_out_faces);

/* ---------------------------------------------------------------- surface metrics
 * Point-to-mesh distances and the Hausdorff distance with the semantics of
 * the libigl calls in simp_cuda/safe_project/src/pamo_safe_project/metrics.py
 * (libigl 2.5.1), in FP64 on the device.  All arrays are device pointers:
 * points and vertices (n, 3) f64, faces (n, 3) i32.  A metrics context owns
 * one HIP stream and its tree buffers; calls return when the results are
 * written.  Faces indexing outside the vertices are rejected. */
typedef struct pamo_metrics pamo_metrics;
int pamo_metrics_create(pamo_metrics** out, int device);
int pamo_metrics_destroy(pamo_metrics* m);

/* Replaces igl.point_mesh_squared_distance(P, V, F) -> (sqrD, I, C)
 * (metrics.py:41-42): squared distance of each point to the nearest face,
 * that face's index (ties: smaller index) and the closest point.
 * face_index and closest may be NULL. */
int pamo_point_mesh_sqdist(pamo_metrics* m, const double* points, long long n_points, const double* verts, int n_v,
                           const int* faces, int n_f, double* sqr_d, int* face_index, double* closest);

/* Replaces igl.hausdorff(VA, FA, VB, FB) (metrics.py:49): the larger of the
 * largest distance of a vertex of A to the surface of B and of a vertex of B
 * to the surface of A, written to host *d_out. */
int pamo_hausdorff(pamo_metrics* m, const double* va, int n_va, const int* fa, int n_fa, const double* vb, int n_vb,
                   const int* fb, int n_fb, double* d_out);

/* No counterpart in the reference.  pamo_point_mesh_sqdist for callers that
 * cannot allocate device memory (code between captures, serving loops, hosts
 * that forbid allocation): the same results, with the tree and workspace in
 * a static device array of 16 MiB per GPU instead of a context's buffers, so
 * the call makes no hipMalloc or hipFree.  Runs on the current device;
 * stream is a hipStream_t of that device or NULL.  At most 65536 faces
 * (n_f above that fails with an error status).  Calls on one device share its
 * array: they are serialised by a process-wide lock and each returns after
 * its kernels finish, so calls never overlap on a GPU; work the caller puts
 * on the same array by other means would race.  The first call on a device
 * looks up that device's array address. */
int pamo_point_mesh_sqdist_noalloc(const double* points, long long n_points, const double* verts, int n_v,
                                   const int* faces, int n_f, double* sqr_d, int* face_index, double* closest,
                                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PAMO_HIP_H */
