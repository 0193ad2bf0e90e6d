/* C ABI of libcusimp.so, the drop-in for the library the reference's Blender
 * add-on loads (pamo_blender/__init__.py:75-79).  Host pointers. */
#ifndef PAMO_BLENDER_H
#define PAMO_BLENDER_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces pamo_simplify of pamo_blender/cusimpinterface.cu:57-161:
 * in_verts (n_v, 3) f32 and in_tris (n_t, 3) i32 with in_out_nv_ni = {n_v, n_t};
 * the mesh is simplified without centring or face target until four rounds
 * make no progress, with the cost threshold `threshold`.  out_verts / out_tris
 * (capacity n_v / n_t) receive the result and in_out_nv_ni its counts.  On an
 * error a message goes to stderr and the counts are set to 0. */
void pamo_simplify(float* in_verts, int* in_tris, float* out_verts, int* out_tris, int* in_out_nv_ni,
                   float threshold);

#ifdef __cplusplus
}
#endif

#endif /* PAMO_BLENDER_H */
