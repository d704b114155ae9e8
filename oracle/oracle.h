// oracle.h — TEST INFRASTRUCTURE ONLY.  The CPU parity oracle for the HIP
// path tracer: a C++17/OpenMP restatement of the reference CPU renderer
// (li-plus/tinypt src/tinypt/cpu/*), consuming the same tpt_scene_desc the
// HIP backend consumes.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker.
#pragma once

#include "../include/tinypt_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

const char *oracle_last_error(void);

/* RNG known-answer helper: run `n` ops on RandEngine(seed); op 0 = random()
 * (float stored as double), op 1 = rand_int(). */
int oracle_rng_run(uint64_t seed, const int32_t *ops, int32_t n, double *out);
/* raw state after RandEngine(seed): 3 uint16 */
void oracle_rng_state(uint64_t seed, uint16_t *state3);
uint64_t oracle_stream_seed(uint64_t base, uint64_t pixel, uint64_t sample);

oracle_scene *oracle_scene_create(const tpt_scene_desc *desc);
void oracle_scene_destroy(oracle_scene *s);

/* Object-light list in reference order (cpu/scene.cpp:6-36): kind 0 = primitive
 * (id = global primitive id), kind 1 = environment light, kind 2 = ObjectLight of a
 * whole mesh (id = its index in tpt_scene_desc.objects).  Returns count. */
int oracle_object_lights(const oracle_scene *s, int32_t *kinds, int32_t *ids, int32_t cap);

/* Closest hit, geometric only.  mode 0: reference structure (linear top level,
 * one median-split BVHTree(.,4) per mesh shape, recursive near-first traversal);
 * mode 1: brute force over all primitives with a (t, id) tie-break. */
int oracle_trace_closest(const oracle_scene *s, const float *rays, uint64_t n, int32_t *prim, float *t, int mode,
                         int threads);
int oracle_trace_occluded(const oracle_scene *s, const float *rays, uint64_t n, uint8_t *occ, int threads);

/* Full surface hit as Scene::intersect returns it (alpha draws use path_key/depth).
 * out: t, normal[3], shade_normal[3], into, uv[2]  -> 10 floats; returns prim id or -1. */
int32_t oracle_intersect_surface(const oracle_scene *s, const float *ray8, uint64_t path_key, int32_t depth,
                                 float *out10);

int oracle_intersect_surface_batch(const oracle_scene *s, const float *rays, uint64_t n, int32_t *prim, float *out10,
                                   int threads);

/* Render rows [y0, y1) into out (H*W*3, only those rows written).  Returns 0. */
int oracle_render(const oracle_scene *s, const tpt_render_params *p, float *out, int32_t y0, int32_t y1,
                  int threads);

/* Single-sample radiance of one (pixel, sample) stream; returns radiance in out3. */
int oracle_radiance_sample(const oracle_scene *s, const tpt_render_params *p, int32_t x, int32_t y, int32_t sample,
                           float *out3);

/* Unit helpers pinned by the reference gtests. */
int oracle_bvh_octant_golden(int leaf_size, float *boxes /*15*6*/, int32_t *leaf_obj /*15*/, int32_t *is_leaf);
int oracle_distribution1d(const float *w, int32_t n, float u, float *x, float *pdf_at_x);
int oracle_distribution2d(const float *w, int32_t rows, int32_t cols, float u, float v, float *xy, float *pdf_at_xy);
void oracle_env_uv2dir(float u, float v, float *dir3);
void oracle_env_dir2uv(const float *dir3, float *uv2);
void oracle_sobel(const float *height, int32_t w, int32_t h, float scale, float *grad /* h*w*2 */);
void oracle_rotation_from_direction(const float *w3, float *m9);
int oracle_rotation_from_euler(const int32_t *axes3, const float *angles3, int32_t degrees, float *m9);
int oracle_delta_light_sample(const oracle_scene *s, int32_t light, const float *pos3, uint64_t seed, float *dir3);
int oracle_light_sample(const oracle_scene *s, int32_t light_index, const float *pos3, uint64_t seed, float *dir3,
                        float *pdf);
/* BxDF evaluation on a synthetic hit: f (rgb) and pdf for a given out_dir, and one sample. */
int oracle_bxdf_eval(const oracle_scene *s, int32_t material, const float *ray_dir3, const float *normal3,
                     int32_t into, const float *out_dir3, float *f3, float *pdf);
int oracle_bxdf_sample(const oracle_scene *s, int32_t material, const float *ray_dir3, const float *normal3,
                       int32_t into, uint64_t seed, float *out_dir3);

/* Microfacet terms of a Glossy lobe (cpu/material.cpp:35-111): type TPT_BXDF_GLOSSY_BLINN
 * (BlinnPhongModel(shininess)) or TPT_BXDF_GLOSSY_GGX (TrowbridgeReitzModel(2 / (shininess + 2))):
 * D(n, h) (also pdf_h) and G(l, n, v). */
int oracle_micro_terms(int32_t type, float shininess, const float *n3, const float *l3, const float *v3,
                       const float *h3, float *D, float *G);
/* `count` microfacet normals from sample_h (material.cpp:58-66, 88-99), one RandEngine(seed) stream. */
int oracle_micro_sample_h(int32_t type, float shininess, const float *n3, uint64_t seed, int32_t count, float *h3);
/* Camera::shoot_ray(x, y) (cpu/camera.cpp:23-34) with RandEngine(seed) for the lens. */
int oracle_camera_ray(const oracle_scene *s, float x, float y, uint64_t seed, float *org3, float *dir3);

#ifdef __cplusplus
}
#endif
