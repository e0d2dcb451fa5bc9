/*
 * taastar C ABI -- MI355X (gfx950) temporal anti-aliasing resolve + fused CAS sharpen.
 *
 * This is the drop-in boundary for the two GPU passes of cg-tuwien/TAA-STAR that sit on the
 * hot path: the temporal resolve `shaders/taa.comp` and the FidelityFX CAS pass
 * `shaders/sharpen_cas.comp`, both driven by `source/taa.hpp`.
 *
 * Conventions
 *   - Plain C types only; every pointer named "device" is HIP device memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *   - Every entry point returns an int status (TAA_OK or a negative TAA_ERR_*); the message of
 *     the last failure on the calling thread is available from taa_last_error().
 *   - No entry point throws. Only the context/session/plan constructors allocate device memory,
 *     and taa_resolve_frame for a split screen with CAS (its intermediate screen, allocated on
 *     first use per device and kept; taa_resolve_frame_noalloc allocates nothing).
 *   - Images are row-major, tightly packed rows of `pitch` pixels (pitch >= width). Every
 *     image must fit in 4 GiB: pitch * height * 16 bytes (64 for the inspect image) < 2^32.
 *
 * The struct layouts below are byte-identical to the reference's host structs (std140 in the
 * shaders); offsets are pinned with static asserts in the C++ translation units.
 */
#ifndef TAASTAR_TAA_ABI_H
#define TAASTAR_TAA_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TAASTAR_ABI_VERSION 3

typedef int32_t taa_bool32; /* VkBool32: 0 = false, anything else = true */

/* ---- status codes ------------------------------------------------------------------------ */
enum {
    TAA_OK = 0,
    TAA_ERR_INVALID_ARG = -1,  /* null pointer, bad size, bad enum, aliasing outputs */
    TAA_ERR_UNSUPPORTED = -2,  /* a feature the build does not implement (see taa_resolve_frame) */
    TAA_ERR_OUT_OF_MEMORY = -3,/* device allocation failed */
    TAA_ERR_HIP = -4,          /* a HIP runtime call failed */
    TAA_ERR_BUDGET = -5,       /* session admission refused: HBM budget exhausted */
    TAA_ERR_NO_DEVICE = -6,    /* no usable gfx950 device */
    TAA_ERR_CAPACITY = -7      /* a fixed-capacity table (sessions, batch slots) is full */
};

/* ---- storage formats --------------------------------------------------------------------- */
/* colour input, history in/out and result screen share one format (TAA_IMAGE_FORMAT_RGB,
 * shaders/shader_cpu_common.h:55-61) */
enum { TAA_FMT_RGBA32F = 0, TAA_FMT_RGBA16F = 1 };
/* velocity (IMAGE_FORMAT_VELOCITY, shader_cpu_common.h:77). RG formats read back as
 * (r, g, 0, 1), the Vulkan rule for components a format does not store. */
enum { TAA_VEL_RG32F = 0, TAA_VEL_RG16F = 1, TAA_VEL_RGBA32F = 2, TAA_VEL_RGBA16F = 3 };
/* arithmetic of the resolve.
 *   FP32: IEEE fp32 in the operation order of taa.comp (no contraction, correctly rounded
 *         division and square root) -- the parity path.
 *   FP16: fp16 texels entering the history and CAS filters through mixed fp16 x fp32 FMAs (the
 *         CAS minimum / maximum in packed fp16), moments, rectification and the blend in fp32,
 *         hardware reciprocal / square root; texel indices and rejection decisions as in FP32 --
 *         the throughput path. It covers the variance or plain min/max box, Catmull-Rom or
 *         bilinear history, every velocity use (matrix reprojection, movers only, all) and sample
 *         mode, clamp or clip, reject-outside, dynamic anti-ghosting, raw or linear depth test,
 *         velocity-based alpha, Lottes weighting, the near-clamp reduction, the debug image and
 *         fused CAS, with RG or RGBA velocity of the colour storage's precision -- the reference's
 *         default block with its RGBA16F velocity included. Shaped or chroma-shrunk boxes, slow
 *         clip, B-spline history, Karis tone mapping, unjitter, noise, pass-through and bypass
 *         are resolved by the FP32 path (taa_resolve_kernel tells which). Upsampling runs the
 *         upsampling resolve in the FP16 path's arithmetic (TAA_KERNEL_UPSAMPLE_FP16), with every
 *         sample position, texel index and decision exact. */
enum { TAA_ARITH_FP32 = 0, TAA_ARITH_FP16 = 1 };
/* the kernel a resolve runs (taa_resolve_kernel) */
enum {
    TAA_KERNEL_EXACT = 0,    /* FP32 arithmetic in taa.comp's operation order (k_resolve)      */
    TAA_KERNEL_FP16 = 1,     /* the FP16-arithmetic throughput kernel (k_resolve_fast)         */
    TAA_KERNEL_UPSAMPLE = 2, /* the exact FP32 upsampling resolve (k_resolve_up)               */
    TAA_KERNEL_UPSAMPLE_FP16 = 3 /* the upsampling resolve with the FP16 path's arithmetic      */
};

/* ---- reference structs ------------------------------------------------------------------- */

/* `taa<CF>::Parameters` (source/taa.hpp:30-77; GLSL mirror shaders/taa.comp:50-95). 176 bytes.
 * Field names are kept so that a copy of the reference struct can be memcpy'd in. */
typedef struct taa_params {
    float      mAlpha;                  /*   0 */
    int32_t    mColorClampingOrClipping;/*   4  0 = none, 1 = clamp, 2 = clip, 3 = slow clip */
    taa_bool32 mDepthCulling;           /*   8 */
    taa_bool32 mUnjitterNeighbourhood;  /*  12 */
    taa_bool32 mUnjitterCurrentSample;  /*  16 */
    float      mUnjitterFactor;         /*  20 */
    taa_bool32 mPassThrough;            /*  24 */
    taa_bool32 mUseYCoCg;               /*  28 */
    taa_bool32 mShrinkChromaAxis;       /*  32 */
    taa_bool32 mVarianceClipping;       /*  36 */
    taa_bool32 mShapedNeighbourhood;    /*  40 */
    taa_bool32 mLumaWeightingLottes;    /*  44 */
    float      mVarClipGamma;           /*  48 */
    float      mMinAlpha;               /*  52 */
    float      mMaxAlpha;               /*  56 */
    float      mRejectionAlpha;         /*  60 */
    taa_bool32 mRejectOutside;          /*  64 */
    int32_t    mUseVelocityVectors;     /*  68  0 = off, 1 = movers only, 2 = everything */
    int32_t    mVelocitySampleMode;     /*  72  0 = simple, 1 = 3x3 longest, 2 = 3x3 closest */
    int32_t    mInterpolationMode;      /*  76  0 = bilinear, 1 = b-spline, 2 = catmull-rom */
    taa_bool32 mToneMapLumaKaris;       /*  80 */
    taa_bool32 mAddNoise;               /*  84 */
    float      mNoiseFactor;            /*  88 */
    taa_bool32 mReduceBlendNearClamp;   /*  92 */
    taa_bool32 mDynamicAntiGhosting;    /*  96 */
    taa_bool32 mVelBasedAlpha;          /* 100 */
    float      mVelBasedAlphaMax;       /* 104 */
    float      mVelBasedAlphaFactor;    /* 108 */
    taa_bool32 mRayTraceAugment;        /* 112  segmentation mask (taa_resolve_frame_segmented) */
    uint32_t   mRayTraceAugmentFlags;   /* 116 */
    float      mRayTraceAugment_WNrm;   /* 120 */
    float      mRayTraceAugment_WDpt;   /* 124 */
    float      mRayTraceAugment_WMId;   /* 128 */
    float      mRayTraceAugment_WLum;   /* 132 */
    float      mRayTraceAugment_Thresh; /* 136 */
    int32_t    mRayTraceHistoryCount;   /* 140 */
    float      mDebugMask[4];           /* 144 */
    int32_t    mDebugMode;              /* 160 */
    float      mDebugScale;             /* 164 */
    taa_bool32 mDebugCenter;            /* 168 */
    taa_bool32 mDebugToScreenOutput;    /* 172 */
} taa_params_t;                         /* 176 */

/* `taa<CF>::uniforms_for_taa` (source/taa.hpp:113-129; GLSL UBO shaders/taa.comp:102-118).
 * 544 bytes. Matrices are glm::mat4, column-major: m[col*4 + row]. */
typedef struct taa_uniforms {
    float        mHistoryViewProjMatrix[16]; /*   0 */
    float        mInverseViewProjMatrix[16]; /*  64 */
    taa_params_t param[2];                   /* 128  [1] applies right of splitX */
    float        mJitterNdc[4];              /* 480  only .xy used */
    float        mSinTime[4];                /* 496 */
    taa_bool32   splitScreen;                /* 512 */
    int32_t      splitX;                     /* 516 */
    taa_bool32   mUpsampling;                /* 520  TAAU sample selection (taa_resolve_frame only) */
    taa_bool32   mBypassHistoryUpdate;       /* 524 */
    taa_bool32   mResetHistory;              /* 528 */
    float        mCamNearPlane;              /* 532 */
    float        mCamFarPlane;               /* 536 */
    float        pad1;                       /* 540  extension: > 0 selects depth rejection on
                                                linear depth with this absolute threshold;
                                                0 keeps taa.comp:818-822 (raw depth) */
} taa_uniforms_t;                            /* 544 */

/* `taa<CF>::push_constants_for_cas` (source/taa.hpp:88-91; sharpen_cas.comp:13-16). 32 bytes. */
typedef struct taa_cas_consts {
    uint32_t const0[4];
    uint32_t const1[4];
} taa_cas_consts_t;

/* The images one resolve reads and writes: the descriptor set of taa.comp
 * (shaders/taa.comp:24-38, bound at source/taa.hpp:1009-1025). Inputs are borrowed; the
 * caller owns every buffer. No output (result_screen, result_history, debug, inspect) may be the
 * same buffer as an input or as another output: TAA_ERR_INVALID_ARG. */
typedef struct taa_frame_images {
    const void*  color;          /* set0 b1 uCurrentFrame    (color_format)            */
    const float* depth;          /* set0 b2 uCurrentDepth    D32F                       */
    const void*  history;        /* set0 b3 uHistoryFrame    (color_format)            */
    const float* history_depth;  /* set0 b4 uHistoryDepth    D32F (previous frame)      */
    void*        result_screen;  /* set0 b5 uResultScreen    (color_format); CAS output
                                    when a CAS constant block is given                  */
    void*        debug;          /* set0 b6 uDebug           RGBA16F, or NULL (skipped) */
    const void*  velocity;       /* set0 b7 uCurrentVelocity (velocity_format)          */
    void*        result_history; /* set0 b8 uResultHistory   (color_format)            */
    void*        inspect;        /* NULL, or 16 x 32-bit words per pixel of intermediate
                                    values (TAA_INSPECT_* word indices) for parity tests */
    int32_t      width;
    int32_t      height;
    int32_t      pitch;          /* pixels per row of every image; 0 = width            */
    int32_t      color_format;   /* TAA_FMT_*                                           */
    int32_t      velocity_format;/* TAA_VEL_*                                           */
    /* Input (render) size of color, depth, velocity and history_depth; 0 = width / height (and
     * input_pitch 0 = pitch when input_width is 0, else input_width). A smaller input is upsampled to width x height (TAAU, taa.comp:222-257, 724, 750-755):
     * the neighbourhood comes from the input around int(uv * input size), depth and velocity are
     * sampled at the output texel's uv, and with uniforms.mUpsampling the current sample is the
     * input sample whose jittered position falls into the output texel. */
    int32_t      input_width;
    int32_t      input_height;
    int32_t      input_pitch;
    int32_t      reserved[2];    /* must be 0                                           */
} taa_frame_images_t;

/* word indices of the optional inspect output (int32 unless marked float) */
enum {
    TAA_INSPECT_HIST_TC_X = 0,   /* history filter base texel: floor(uv*size - 0.5)    */
    TAA_INSPECT_HIST_TC_Y = 1,
    TAA_INSPECT_CLOSEST_TAP = 2, /* 3x3 closest-depth tap 0..8 (row-major), -1 if unused */
    TAA_INSPECT_FLAGS = 3,       /* TAA_FLAG_* */
    TAA_INSPECT_COLMIN = 4,      /* float[3] neighbourhood box min */
    TAA_INSPECT_COLMAX = 7,      /* float[3] neighbourhood box max */
    TAA_INSPECT_BLEND = 10,      /* float alpha * beta used by the blend */
    TAA_INSPECT_DEPTH_TC_X = 11, /* texel of the history-depth test: int(historyUv*size) */
    TAA_INSPECT_DEPTH_TC_Y = 12,
    TAA_INSPECT_HIST_UV_X = 13,  /* float historyUv */
    TAA_INSPECT_HIST_UV_Y = 14,
    TAA_INSPECT_SEG_TOTAL = 15,  /* float weighted edge indicator of the segmentation mask
                                    (0 when the mask is off or decided before it) */
    TAA_INSPECT_WORDS = 16
};
enum {
    TAA_FLAG_REJECTED = 1, TAA_FLAG_RECTIFIED = 2, TAA_FLAG_OUTSIDE = 4,
    TAA_FLAG_DEPTH_REJECT = 8, TAA_FLAG_USED_VELOCITY = 16, TAA_FLAG_GHOST_REJECT = 32
};

/* ---- host helpers (no GPU needed) -------------------------------------------------------- */

/* ABI version (TAASTAR_ABI_VERSION). */
int32_t taa_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* taa_last_error(void);

/* Reference defaults of Parameters (source/taa.hpp:31-76). */
int32_t taa_params_defaults(taa_params_t* out);

/* CasSetup (shaders/ffx_cas.h:375-394), as called at source/taa.hpp:965. */
int32_t taa_cas_setup(uint32_t const0[4], uint32_t const1[4], float sharpness,
                      float input_width, float input_height, float output_width, float output_height);

/* helpers::halton (source/helper_functions.hpp:10-17): radical inverse of index i >= 1 in base b. */
float taa_halton(int32_t i, int32_t b);

/* Jitter selection of get_jitter_offset_for_frame (source/taa.hpp:150-233). */
typedef struct taa_jitter_config {
    int32_t      sample_distribution; /* 0 circular quad, 1 uniform-4 helix, 2 Halton(2,3)x8,
                                         3 Halton(2,3)x16, 4 regular 16, 5 user offsets      */
    int32_t      slow_motion;         /* mJitterSlowMotion: > 1 divides the frame id          */
    int32_t      fixed_index;         /* mFixedJitterIndex: >= 0 pins the sample index        */
    float        rotate_degrees;      /* mJitterRotateDegrees                                 */
    float        extra_scale;         /* mJitterExtraScale (reference default 1)              */
    int32_t      num_user_offsets;    /* distribution 5: count of (x,y) pairs in pixel units  */
    const float* user_offsets;
} taa_jitter_config_t;

/* Jitter for `frame_id` in NDC units at this width x height. Unlike the reference, whose pixel
 * size is a function-local static fixed by the first call (taa.hpp:154-155), the pixel size is
 * taken from width/height on every call, so sessions of different resolutions are exact. */
int32_t taa_jitter_offset(const taa_jitter_config_t* cfg, uint64_t frame_id, uint32_t width,
                          uint32_t height, float out_ndc[2], uint32_t* out_num_samples);

/* ---- segmentation mask (mRayTraceAugment) ------------------------------------------------ */

/* mRayTraceAugmentFlags bits (shaders/shader_cpu_common.h:31-40) */
enum {
    TAA_RTFLAG_OUT = 0x0001, /* history outside the screen -> edge-AA fallback            */
    TAA_RTFLAG_DIS = 0x0002, /* disocclusion by a moving object -> ray trace             */
    TAA_RTFLAG_NRM = 0x0004, /* normal discontinuity indicator                           */
    TAA_RTFLAG_DPT = 0x0008, /* linear-depth Sobel indicator                              */
    TAA_RTFLAG_MID = 0x0010, /* material-id edge indicator                                */
    TAA_RTFLAG_LUM = 0x0020, /* luminance Sobel indicator                                 */
    TAA_RTFLAG_CNT = 0x0040, /* keep marking for mRayTraceHistoryCount frames            */
    TAA_RTFLAG_ALL = 0x0080, /* mark every texel for ray tracing (debug)                 */
    TAA_RTFLAG_FXD = 0x0100, /* a fixed 100-texel border for the edge-AA fallback (debug) */
    TAA_RTFLAG_FXA = 0x0200  /* run the edge-AA fallback pass on category-1 texels        */
};
/* Category of a mask texel (bits 0-1); bits 16-31 hold the remaining history count. */
enum { TAA_SEG_TAA = 0, TAA_SEG_EDGE_AA = 1, TAA_SEG_RAY_TRACE = 2 };

/* The extra images of the segmentation mask (taa.comp:34-37, 589-702): R32UI material ids of this
 * and the previous frame (bit 31 marks a moving object, fwd_geometry.frag:286), the RGBA32F
 * uv + spherical-normal G-buffer (IMAGE_FORMAT_NORMAL; z azimuth, w elevation), the previous
 * frame's mask and the mask written by this frame (TAA_IMAGE_FORMAT_SEGMASK, R32UI).
 * The renderer's images (material, prev_material, uv_normal) have the input (render) size, the
 * masks the output size (taa.hpp:336-338). `pitch`: row pitch in texels of the masks, 0 = the
 * frame's pitch; at one size also that of the renderer's images. `input_pitch`: row pitch of the
 * renderer's images when the input size differs from the output size, 0 = the frame's input pitch.
 * segmask must not alias an input. */
typedef struct taa_segmentation_images {
    const uint32_t* material;      /* set0 b10 uCurrentMaterial  */
    const uint32_t* prev_material; /* set0 b11 uPreviousMaterial */
    const float*    uv_normal;     /* set0 b13 uCurrentUvNrm     */
    const uint32_t* prev_segmask;  /* set0 b12 uPreviousSegMask  */
    uint32_t*       segmask;       /* set0 b9  uSegMask (written) */
    int32_t         pitch;
    int32_t         input_pitch;
    int32_t         reserved[2];   /* must be 0 */
} taa_segmentation_images_t;

/* Push constants of the edge-AA fallback pass: push_constants_for_fxaa (source/taa.hpp:93-98),
 * byte-identical (20 bytes); defaults subpix 0.75, edge_threshold 0.116, edge_threshold_min 0.0833. */
typedef struct taa_edge_aa_consts {
    float rcp_frame[2];       /* 1 / output size (the pass steps in whole texels; kept for layout) */
    float subpix;             /* amount of sub-texel smoothing, [0, 1]                             */
    float edge_threshold;     /* local luma contrast, relative to the local maximum, below which a
                                 texel is kept                                                      */
    float edge_threshold_min; /* absolute contrast below which a texel is kept (dark regions)     */
} taa_edge_aa_consts_t;

/* ---- per-frame GPU entry points ---------------------------------------------------------- */

/* One taa.comp dispatch over the whole frame (source/taa.hpp:1008-1027) and, when `cas` is not
 * NULL, the CAS pass of source/taa.hpp:1127-1134 fused into the same kernel: result_screen then
 * receives the sharpened screen, exactly what the reference's separate pass would write.
 * Stream-ordered and asynchronous.
 * An input size different from the output size, or mUpsampling, takes the upsampling path (in the
 * arithmetic `arith` asks). TAA_ERR_UNSUPPORTED: mRayTraceAugment (it needs the
 * extra images of taa_resolve_frame_segmented).
 * Split screen (uniforms->splitScreen): columns x > splitX use param[1], the others param[0]
 * (taa.comp:717). With CAS the frame is sharpened after both parts are resolved, in the
 * reference's order (taa.hpp:1008-1027, 1111-1139): the parts write an intermediate screen, one
 * CAS pass reads it into result_screen. That intermediate is the only device memory this call
 * allocates: one buffer per device, on first use (grown when a larger frame needs it) and kept
 * until the process ends; calls on different streams are ordered around it. */
int32_t taa_resolve_frame(const taa_uniforms_t* uniforms, const taa_frame_images_t* images,
                          const taa_cas_consts_t* cas, int32_t arith, void* stream);

/* Which kernel taa_resolve_frame runs for these arguments, without launching anything (the
 * arguments are checked as the resolve checks them). kernels[0] receives the TAA_KERNEL_* of the
 * columns of param[0], kernels[1] that of the columns of param[1] (the same as kernels[0] without
 * split screen). An FP16 request the FP16 kernel does not cover gets TAA_KERNEL_EXACT, and
 * taa_last_error() then names the first reason (e.g. "a debug image is bound"); otherwise it is
 * "". No counterpart in the reference (its one shader has no fallback). */
int32_t taa_resolve_kernel(const taa_uniforms_t* uniforms, const taa_frame_images_t* images,
                           const taa_cas_consts_t* cas, int32_t arith, int32_t kernels[2]);

/* Bytes of the intermediate screen taa_resolve_frame_noalloc holds per device: 64 MiB, enough for
 * pitch x height texels of 3840x2160 RGBA16F or 2560x1440 RGBA32F. */
#define TAA_NOALLOC_SCRATCH_BYTES (64u << 20)

/* taa_resolve_frame for callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops and hosts that forbid hipMalloc. No counterpart in the reference (its effect owns
 * its temp images). Identical results; the one difference is where a split screen with CAS keeps
 * its intermediate screen: in a file-scope device array of TAA_NOALLOC_SCRATCH_BYTES that the
 * library holds on every device it runs on until the process ends, never in allocated memory. A
 * frame whose intermediate (pitch x height texels of color_format) exceeds it is refused with
 * TAA_ERR_CAPACITY before anything is launched. Every call on one device shares that array and
 * rewrites all of it that it reads, so nothing carries over between calls; calls that overlap on
 * one device (other streams, other threads) race on it and must be serialised by the caller.
 * Allocates nothing and calls no allocating routine. */
int32_t taa_resolve_frame_noalloc(const taa_uniforms_t* uniforms, const taa_frame_images_t* images,
                                  const taa_cas_consts_t* cas, int32_t arith, void* stream);

/* taa_resolve_frame with the segmentation mask of mRayTraceAugment (taa.comp:589-702, 772-783,
 * 959): where a Parameters block has mRayTraceAugment set, each texel's mask value is computed
 * from the extra images, written to seg->segmask, and a non-zero category rejects the history
 * (alpha = mRejectionAlpha) as in the reference; debug mode 8 shows the category. Texels of
 * mPassThrough / mBypassHistoryUpdate leave the mask untouched, as the reference's early return
 * does. Category-2 texels are for the caller's ray tracer; category-1 texels for
 * taa_edge_aa_apply. TAA_ARITH_FP16 takes the FP16 path's arithmetic where no selection depends on
 * it (the mask, indices and decisions stay exact), TAA_ARITH_FP32 the exact kernel.
 * With an input size different from the output size, or mUpsampling, the mask is evaluated at the
 * output texel as the reference does (taa.comp:775 passes the output coordinates): the fixed
 * border of TAA_RTFLAG_FXD is measured against the input size, the neighbours are clamped to the
 * input image, and the centre's material id and normal are read at the output coordinates, which
 * lie outside the input image for most texels -- a read Vulkan leaves undefined unless the device
 * bounds-checks images; it reads zero here (material 0, uv/normal (0, 0, 0, 0)), the value
 * robustImageAccess2 defines. */
int32_t taa_resolve_frame_segmented(const taa_uniforms_t* uniforms, const taa_frame_images_t* images,
                                    const taa_segmentation_images_t* seg, const taa_cas_consts_t* cas,
                                    int32_t arith, void* stream);

/* The sharpen_cas.comp pass alone (source/taa.hpp:1127-1134): src, dst in `format`, must not
 * alias. Border texels are read clamped to the edge. */
int32_t taa_cas_apply(const void* src, void* dst, int32_t width, int32_t height, int32_t pitch,
                      int32_t format, const taa_cas_consts_t* cas, void* stream);

/* The edge-AA fallback on the texels the segmentation mask puts in category 1 (the prepare and
 * FXAA passes of source/taa.hpp:1061-1107, antialias_fxaa_prepare.comp, antialias_fxaa.comp).
 * Every texel of dst gets alpha = luma = dot(rgb, (0.299, 0.587, 0.114)) of src, as the prepare
 * pass writes; texels whose mask category is not 1 keep src's rgb (antialias_fxaa.comp:55-77).
 * A category-1 texel whose 4-neighbour luma range reaches max(edge_threshold_min,
 * edge_threshold * local max) is blended with one 4-neighbour across the edge: vertically when the
 * vertical luma Sobel gradient is at least the horizontal one, else horizontally, towards the side
 * with the larger luma step, with weight subpix / 2 * clamp(|weighted 3x3 mean luma - luma| /
 * range, 0, 1) (mean weights: 2 for the 4-neighbours, 1 for the diagonals).
 * This is the library's own filter, not NVIDIA's FXAA: results differ from the reference's
 * pass on category-1 texels. src and dst in `format` with rows of `pitch` texels (0 = width),
 * segmask R32UI with rows of `seg_pitch` (0 = pitch); src and dst must not alias. */
int32_t taa_edge_aa_apply(const void* src, const uint32_t* segmask, void* dst, int32_t width, int32_t height,
                          int32_t pitch, int32_t seg_pitch, int32_t format, const taa_edge_aa_consts_t* consts,
                          void* stream);
/* The reference's defaults of push_constants_for_fxaa for an output of width x height. */
int32_t taa_edge_aa_defaults(int32_t width, int32_t height, taa_edge_aa_consts_t* out);

/* ---- context, sessions, batches ---------------------------------------------------------- */

typedef struct taa_ctx taa_ctx_t;
typedef struct taa_session taa_session_t;

/* A context owns the device tables of one GPU and the HBM admission budget for sessions.
 * hbm_budget_bytes = 0 means 90% of the device's free memory at creation. */
int32_t taa_ctx_create(int32_t device, uint64_t hbm_budget_bytes, taa_ctx_t** out);
int32_t taa_ctx_destroy(taa_ctx_t* ctx);
/* bytes admitted / budget */
int32_t taa_ctx_usage(const taa_ctx_t* ctx, uint64_t* used_bytes, uint64_t* budget_bytes);

/* Device bytes a session of this shape keeps resident (history ping-pong, depth ring,
 * colour and velocity input slots). */
uint64_t taa_session_bytes(int32_t width, int32_t height, int32_t color_format, int32_t velocity_format);
/* ... of a session whose inputs are rendered at input_width x input_height (TAAU). */
uint64_t taa_session_bytes_scaled(int32_t width, int32_t height, int32_t input_width, int32_t input_height,
                                  int32_t color_format, int32_t velocity_format);

/* A session is one stream of frames: it owns what taa<CF> owns per frame in flight
 * (source/taa.hpp:1364-1369) -- the history ping-pong -- plus resident input slots and the
 * previous depth. TAA_ERR_BUDGET when the context's HBM budget cannot hold it. */
int32_t taa_session_open(taa_ctx_t* ctx, int32_t width, int32_t height, int32_t color_format,
                         int32_t velocity_format, taa_session_t** out);
/* A session whose frames are rendered at input_width x input_height and resolved into a
 * width x height history and result (TAAU: taa.hpp:292 mInputResolution / mOutputResolution):
 * the colour, depth and velocity slots have the input size; every resolve takes the upsampling
 * path of taa_resolve_frame. */
int32_t taa_session_open_scaled(taa_ctx_t* ctx, int32_t width, int32_t height, int32_t input_width,
                                int32_t input_height, int32_t color_format, int32_t velocity_format,
                                taa_session_t** out);
int32_t taa_session_close(taa_session_t* s);

/* Device pointers of the slots the next frame's inputs are written into, and of the history
 * the last resolve produced. width / height / pitch describe the input slots (the input size of
 * a scaled session); the history has the session's output size. */
typedef struct taa_session_io {
    void*       color;     /* write: next frame colour (color_format)        */
    float*      depth;     /* write: next frame depth  (D32F)                */
    void*       velocity;  /* write: next frame velocity (velocity_format)   */
    const void* history;   /* read: history written by the last resolve      */
    int32_t     width, height, pitch, color_format, velocity_format;
    uint64_t    frames_resolved;
} taa_session_io_t;
int32_t taa_session_io(taa_session_t* s, taa_session_io_t* out);

/* Resolve the frame in the session's input slots. The first frame of a session, and the first
 * after taa_session_reset, runs with mResetHistory forced on (the reference instead skips TAA
 * on its very first frame, source/taa.hpp:989-990). */
/* A scaled session, or mUpsampling, takes the upsampling path (as taa_resolve_frame does). */
int32_t taa_session_resolve(taa_session_t* s, const taa_uniforms_t* uniforms,
                            const taa_cas_consts_t* cas, void* result_screen, int32_t arith,
                            void* stream);
int32_t taa_session_reset(taa_session_t* s);

#define TAA_BATCH_MAX_ITEMS 1024

/* One frame of one session in a batch. */
typedef struct taa_batch_item {
    taa_session_t*          session;
    const taa_uniforms_t*   uniforms;      /* host memory, read during the call          */
    const taa_cas_consts_t* cas;           /* NULL = no sharpen                          */
    void*                   result_screen; /* device                                      */
} taa_batch_item_t;

/* Resolve frames of many sessions (any mix of resolutions, one colour format, one arith) in
 * ONE kernel launch over the tiles of all frames. At most TAA_BATCH_MAX_ITEMS items
 * (TAA_ERR_CAPACITY beyond). Sessions must be distinct, and so must the
 * result_screen pointers; no result_screen may lie inside a batched session's buffers
 * (TAA_ERR_INVALID_ARG). Stream-ordered.
 * Upsampling frames (scaled sessions, mUpsampling) go through a second batched launch of the
 * upsampling resolve, in the call's arithmetic, in the same call (and in the same captured graph of
 * a plan). */
int32_t taa_batch_submit(taa_ctx_t* ctx, int32_t n, const taa_batch_item_t* items,
                         int32_t arith, void* stream);

/* A batch plan captures descriptor upload + the batched kernel in a hipGraph for a fixed set
 * of sessions; each launch re-reads the items' uniforms and advances every session by one frame.
 * Tile shape per plan comes from the context's autotuner table: a single one-size frame takes its
 * class's shape, several take the shape tuned for the largest class of at most their total
 * pixel count (a launch over many frames behaves like one frame of their total size), or the
 * default shape when it resolves 2% fewer texels over the batch's frames. */
typedef struct taa_batch_plan taa_batch_plan_t;
int32_t taa_batch_plan_create(taa_ctx_t* ctx, int32_t n, const taa_batch_item_t* items,
                              int32_t arith, taa_batch_plan_t** out);
int32_t taa_batch_plan_launch(taa_batch_plan_t* plan, void* stream);
int32_t taa_batch_plan_destroy(taa_batch_plan_t* plan);

/* ---- autotuner --------------------------------------------------------------------------- */

/* Tile shapes the kernels are built for (output tile width x height, 256 threads, LDS row
 * padding). Separate tables for the plain resolve and for the resolve with CAS fused, whose
 * resolve region is the tile plus a 1-px ring. */
typedef struct taa_tile_shape {
    int32_t tile_w, tile_h;
    int32_t lds_pad;
} taa_tile_shape_t;

int32_t taa_tile_shape_count(void);
int32_t taa_tile_shape_get(int32_t index, taa_bool32 cas, taa_tile_shape_t* out);

/* Time every built tile shape for this frame class (width, height, colour format, arith, CAS)
 * on the context's device with a synthetic frame, and record the fastest in the process-wide
 * table that taa_resolve_frame, taa_batch_submit and batch plans consult (a batch of several
 * frames uses the entry of the largest class of at most its total pixel count). */
int32_t taa_autotune(taa_ctx_t* ctx, int32_t width, int32_t height, int32_t color_format,
                     int32_t arith, taa_bool32 cas, int32_t iters, int32_t* best_index,
                     float* best_us);
/* Force a tile shape for a frame class; index -1 restores the built-in default (with CAS: 30x30,
 * or for frames of fewer 30x30 tiles than the device holds workgroups at once, 30x14 and below
 * 512 such tiles 14x14). */
int32_t taa_tile_shape_select(int32_t width, int32_t height, int32_t color_format, int32_t arith,
                              taa_bool32 cas, int32_t index);
/* The shape index the table holds for a frame class. */
int32_t taa_tile_shape_selected(int32_t width, int32_t height, int32_t color_format, int32_t arith,
                                taa_bool32 cas);

#ifdef __cplusplus
}
#endif
#endif /* TAASTAR_TAA_ABI_H */
