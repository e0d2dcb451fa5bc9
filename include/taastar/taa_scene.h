/*
 * taastar synthetic scene generator (host, C ABI).
 *
 * No counterpart in the reference's hot path: it stands in for the renderer that produced the
 * resolve's inputs (colour, depth and velocity written by shaders/fwd_geometry.frag:288-295 with
 * the jittered projection of source/taa.hpp:243-253). It draws seeded convex polygons with
 * constant depth, log-normal HDR colour, 1% fireflies, a camera pan and per-polygon motion, and
 * writes the velocity convention of the reference: unjittered motion in UV units, so that
 * historyUv = uv - velocity (shaders/taa.comp:421).
 */
#ifndef TAASTAR_TAA_SCENE_H
#define TAASTAR_TAA_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct taa_scene_config {
    int32_t  width, height;
    int32_t  num_polygons;
    uint64_t seed;
    float    lum_mu, lum_sigma;          /* polygon luminance ~ LogNormal(mu, sigma)          */
    float    firefly_fraction;           /* share of pixels multiplied by U[min, max]         */
    float    firefly_min, firefly_max;
    int32_t  pan_mode;                   /* 0: pan per axis ~ U[-pan, pan]; 1: |pan| ~ U[0,pan],
                                            direction uniform                                  */
    float    pan;                        /* px / frame                                        */
    float    object_speed;               /* per-polygon velocity per axis ~ U[-v, v] px/frame */
    int32_t  jitter_distribution;        /* taa_jitter_config_t::sample_distribution          */
    int32_t  jitter_phase;               /* frame offset into the jitter sequence             */
    float    color_max;                  /* > 0: clamp colour channels (fp16 storage)         */
    float    depth_min, depth_max;       /* polygon depth ~ U[min, max); background 1.0       */
} taa_scene_config_t;

typedef struct taa_scene taa_scene_t;

/* Fills cfg with the defaults of the 640x360 case: 200 polygons, LogNormal(-1, 1.5), 1% fireflies
 * at 10-100x, pan +-8 px/frame, objects +-16 px/frame, Halton(2,3)x8, depth U[0.1, 1). */
int32_t taa_scene_defaults(taa_scene_config_t* cfg);
int32_t taa_scene_create(const taa_scene_config_t* cfg, taa_scene_t** out);
int32_t taa_scene_destroy(taa_scene_t* s);

/* Render frame `frame` (>= 0) into host buffers: colour RGBA float (alpha 1), depth float,
 * velocity RGBA float (x, y in UV units; z = 0; w = 1 on polygons, 0 on the background).
 * jitter_ndc receives the frame's jitter in NDC (may be NULL). threads <= 0 uses all cores.
 * The output does not depend on the thread count. */
int32_t taa_scene_render(taa_scene_t* s, int32_t frame, float* color_rgba, float* depth,
                         float* velocity_rgba, float jitter_ndc[2], int32_t threads);

#ifdef __cplusplus
}
#endif
#endif /* TAASTAR_TAA_SCENE_H */
