/*
 * taastar velocity producer -- the motion-vector image the resolve reads, made from a depth
 * image and the camera matrices, for clients that render no velocity target.
 *
 * In the reference the velocity image is written by the geometry pass (shaders/fwd_geometry.frag:
 * 288-295, also blinnphong_and_normal_mapping.frag:140-146) from the interpolated clip positions
 * of the vertex stage (transform_and_pass_on.vert:87-101, animatedObject.vert:86-102) and the
 * matrices of source/main.cpp:408-424. This entry point computes the same quantity per texel from
 * depth instead of from rasterised vertices.
 *
 * Velocity image contract (IMAGE_FORMAT_VELOCITY, shader_cpu_common.h:77; read by taa.comp:
 * 392-422 as historyUv = currentUv - v.xy, historyDepth = currentDepth - v.z):
 *   xy  = (NDC_current - jitter_current) - (NDC_previous - jitter_previous), times 0.5: the
 *         texel's motion in uv units, current minus previous, without the subpixel jitter;
 *   z   = NDC depth current minus previous;
 *   w   = the object id (0 for static geometry), as fwd_geometry.frag:295 stores movingObjectId.
 * RG formats keep xy only. Half formats round to nearest even and saturate to +-65504 (NaN stays
 * NaN), as the colour storage of the resolve does.
 */
#ifndef TAASTAR_TAA_VELOCITY_H
#define TAASTAR_TAA_VELOCITY_H

#include <stdint.h>

#include "taa_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reprojection of one frame. Matrices are 4x4 column-major (glm layout), applied as M * v. */
typedef struct taa_reproject {
    /* inverse(P_jittered_current * V_current): texel centre + depth -> homogeneous world */
    float inverse_view_proj[16];
    /* P_jittered_previous * V_previous: mPrevFrameProjViewMatrix (main.cpp:423) */
    float prev_view_proj[16];
    /* mJitterCurrentPrev (main.cpp:424): current jitter xy, previous jitter zw, NDC units, as
     * translated into the projection by get_jittered_projection_matrix (taa.hpp:242-252) */
    float jitter_current_prev[4];
    /* device array of n_objects column-major 4x4 matrices: object k's (id k+1) previous model
     * matrix times the inverse of its current one (mMover_additionalModelMatrix_prev and
     * _current, main.cpp:428-429); NULL when n_objects is 0 */
    const float* object_prev_from_current;
    int32_t n_objects;
    int32_t reserved[3]; /* must be 0 */
} taa_reproject_t;

/* Velocity image of one frame from its depth image (fwd_geometry.frag:288-295 per texel).
 *   depth      device, D32F, the jittered depth the resolve also reads (NDC z in [0, 1])
 *   object_id  device, one uint32 per texel (0 = static) or NULL (every texel static); ids above
 *              n_objects move with the camera only and still land in w
 *   velocity   device, TAA_VEL_* format; written for every texel of width x height, row padding
 *              untouched
 * Texel centre of (x, y): uv = ((x + 0.5) / width, (y + 0.5) / height), NDC = uv * 2 - 1, as
 * taa.comp forms currentUv and its clip position (taa.comp:426). FP32 arithmetic without
 * contraction; the perspective division is one correctly rounded reciprocal of w and three
 * products. Stream-ordered, asynchronous, allocates nothing. */
int32_t taa_velocity_from_depth(const taa_reproject_t* reproject, const float* depth,
                                const uint32_t* object_id, void* velocity, int32_t velocity_format,
                                int32_t width, int32_t height, int32_t pitch, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TAASTAR_TAA_VELOCITY_H */
