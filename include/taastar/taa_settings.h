/*
 * taastar settings -- the TAA sections of the reference's settings file, read and written in its
 * on-disk format, and the bridge from those settings to the per-frame structs of taa_abi.h.
 *
 * Format (what source/taa.hpp:1198-1339 writes through mINI's generate(), IniUtil.cpp:50-60):
 * "[section]" lines and "key=value" lines, section and key names lower-cased, no spaces around
 * '=', '\n' line ends; bools as 0/1, ints and uint32 flags in decimal, floats as C's "%f"
 * (std::to_string), vectors as one key per component (name.x, name.y, ...). Reading follows
 * mINI's parser and IniUtil.cpp:101-110: names compare case-insensitively, lines starting with
 * ';' are comments, whitespace around keys and values is ignored, a later key overrides an
 * earlier one, an absent or empty value keeps the current setting; numbers are read as strtof /
 * strtol / strtoul read them (a leading number, the rest of the value ignored), and a bool is
 * true for "1" or "true".
 */
#ifndef TAASTAR_TAA_SETTINGS_H
#define TAASTAR_TAA_SETTINGS_H

#include <stddef.h>
#include <stdint.h>

#include "taa_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TAA_SETTINGS_MAX_SAMPLE_OFFSETS 256

/* The settings of the TAA effect that the reference stores (taa.hpp:1198-1262). */
typedef struct taa_settings {
    taa_params_t param[2];               /* [TAA_Param_0], [TAA_Param_1] (mParameters)           */
    /* [TAA_Primary] */
    taa_bool32 taa_enabled;              /* mTaaEnabled            (default 1)                   */
    int32_t    sample_distribution;      /* mSampleDistribution    (1)                           */
    int32_t    sharpener;                /* mSharpener 0 off, 1 simple, 2 FidelityFX CAS (0)     */
    float      sharpen_factor;           /* mSharpenFactor         (0.5)                         */
    taa_bool32 split_screen;             /* mSplitScreen           (0)                           */
    int32_t    split_x;                  /* mSplitX                (0)                           */
    int32_t    fixed_jitter_index;       /* mFixedJitterIndex      (-1)                          */
    float      jitter_extra_scale;       /* mJitterExtraScale      (1)                           */
    int32_t    jitter_slow_motion;       /* mJitterSlowMotion      (1)                           */
    float      jitter_rotate_degrees;    /* mJitterRotateDegrees   (0)                           */
    taa_bool32 reset_history_on_change;  /* mResetHistoryOnChange  (1)                           */
    int32_t    num_sample_offsets;       /* mDebugSampleOffsets.size (1; at least 1 after a read) */
    float      sample_offsets[TAA_SETTINGS_MAX_SAMPLE_OFFSETS][2]; /* mDebugSampleOffsets_i, pixels */
    /* [TAA_Postprocess]: the display zoom of post_process.comp, carried through unchanged */
    taa_bool32 postprocess_enabled;      /* mPostProcessEnabled    (1)                           */
    uint32_t   zoom;                     /* (0)                                                  */
    uint32_t   show_zoom_box;            /* (1)                                                  */
    int32_t    zoom_src_ltwh[4];         /* (950, 530, 20, 20)                                   */
    int32_t    zoom_dst_ltwh[4];         /* (1710, 10, 200, 200)                                 */
} taa_settings_t;

/* The reference's defaults (taa.hpp:31-76, 101-111, 1346-1424). */
int32_t taa_settings_defaults(taa_settings_t* out);

/* Apply the [TAA_Param_0], [TAA_Param_1], [TAA_Primary] and [TAA_Postprocess] keys found in an
 * INI text (len bytes, need not be NUL-terminated) to *settings, as readSettingsFromIni does to
 * the effect's members (taa.hpp:1265-1333); other sections are ignored. A value that holds no
 * number where one is needed is TAA_ERR_INVALID_ARG (the reference's std::stof / stol throw),
 * with *settings unchanged; a sample-offset count above TAA_SETTINGS_MAX_SAMPLE_OFFSETS is
 * TAA_ERR_CAPACITY. */
int32_t taa_settings_read_ini(const char* text, size_t len, taa_settings_t* settings);

/* The TAA sections as writeSettingsToIni + mINI generate() print them (taa.hpp:1198-1262):
 * at most `capacity` bytes into `out` (NUL-terminated when it fits), *length = bytes needed
 * without the NUL. TAA_ERR_CAPACITY when `capacity` is too small (*length is still set). */
int32_t taa_settings_write_ini(const taa_settings_t* settings, char* out, size_t capacity, size_t* length);

/* The per-frame structs the settings feed (taa.hpp:935-951): the two parameter blocks, split
 * screen and split position into *uniforms (other fields untouched), and the jitter selection
 * into *jitter (user offsets point into settings->sample_offsets). Either output may be NULL. */
int32_t taa_settings_apply(const taa_settings_t* settings, taa_uniforms_t* uniforms, taa_jitter_config_t* jitter);

#ifdef __cplusplus
}
#endif
#endif /* TAASTAR_TAA_SETTINGS_H */
