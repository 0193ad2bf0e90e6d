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

/* Centring offset of PaMO.run: mean over faces of the face centroids
 * (reference simp_cuda/pamo/__init__.py:60-72,144), FP64 accumulation in a
 * fixed order; out3 is a host float[3]. */
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
#define PAMO_DBG_BVH_N 9      /* count only: leaves of the last LBVH */
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

#ifdef __cplusplus
}
#endif

#endif /* PAMO_HIP_H */
