"""ctypes mirror of include/taastar/taa_abi.h, taa_velocity.h, taa_settings.h and taa_scene.h (struct layouts are asserted)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtaastar.so")

TAA_OK = 0
TAA_ERR_INVALID_ARG = -1
TAA_ERR_UNSUPPORTED = -2
TAA_ERR_OUT_OF_MEMORY = -3
TAA_ERR_HIP = -4
TAA_ERR_BUDGET = -5
TAA_ERR_NO_DEVICE = -6
TAA_ERR_CAPACITY = -7
NOALLOC_SCRATCH_BYTES = 64 << 20

FMT_RGBA32F, FMT_RGBA16F = 0, 1
VEL_RG32F, VEL_RG16F, VEL_RGBA32F, VEL_RGBA16F = 0, 1, 2, 3
ARITH_FP32, ARITH_FP16 = 0, 1
KERNEL_EXACT, KERNEL_FP16, KERNEL_UPSAMPLE, KERNEL_UPSAMPLE_FP16 = 0, 1, 2, 3

INSPECT_WORDS = 16
INSPECT = dict(HIST_TC_X=0, HIST_TC_Y=1, CLOSEST_TAP=2, FLAGS=3, COLMIN=4, COLMAX=7, BLEND=10,
               DEPTH_TC_X=11, DEPTH_TC_Y=12, HIST_UV_X=13, HIST_UV_Y=14, SEG_TOTAL=15)
RTFLAG = dict(OUT=0x1, DIS=0x2, NRM=0x4, DPT=0x8, MID=0x10, LUM=0x20, CNT=0x40, ALL=0x80, FXD=0x100, FXA=0x200)
SEG_TAA, SEG_EDGE_AA, SEG_RAY_TRACE = 0, 1, 2
FLAG_REJECTED, FLAG_RECTIFIED, FLAG_OUTSIDE, FLAG_DEPTH_REJECT, FLAG_USED_VELOCITY, FLAG_GHOST_REJECT = 1, 2, 4, 8, 16, 32


class Params(C.Structure):
    """taa<CF>::Parameters (source/taa.hpp:30-77), 176 bytes."""
    _fields_ = [
        ("mAlpha", C.c_float), ("mColorClampingOrClipping", C.c_int32), ("mDepthCulling", C.c_int32),
        ("mUnjitterNeighbourhood", C.c_int32), ("mUnjitterCurrentSample", C.c_int32), ("mUnjitterFactor", C.c_float),
        ("mPassThrough", C.c_int32), ("mUseYCoCg", C.c_int32), ("mShrinkChromaAxis", C.c_int32),
        ("mVarianceClipping", C.c_int32), ("mShapedNeighbourhood", C.c_int32), ("mLumaWeightingLottes", C.c_int32),
        ("mVarClipGamma", C.c_float), ("mMinAlpha", C.c_float), ("mMaxAlpha", C.c_float), ("mRejectionAlpha", C.c_float),
        ("mRejectOutside", C.c_int32), ("mUseVelocityVectors", C.c_int32), ("mVelocitySampleMode", C.c_int32),
        ("mInterpolationMode", C.c_int32), ("mToneMapLumaKaris", C.c_int32), ("mAddNoise", C.c_int32),
        ("mNoiseFactor", C.c_float), ("mReduceBlendNearClamp", C.c_int32), ("mDynamicAntiGhosting", C.c_int32),
        ("mVelBasedAlpha", C.c_int32), ("mVelBasedAlphaMax", C.c_float), ("mVelBasedAlphaFactor", C.c_float),
        ("mRayTraceAugment", C.c_int32), ("mRayTraceAugmentFlags", C.c_uint32), ("mRayTraceAugment_WNrm", C.c_float),
        ("mRayTraceAugment_WDpt", C.c_float), ("mRayTraceAugment_WMId", C.c_float), ("mRayTraceAugment_WLum", C.c_float),
        ("mRayTraceAugment_Thresh", C.c_float), ("mRayTraceHistoryCount", C.c_int32),
        ("mDebugMask", C.c_float * 4), ("mDebugMode", C.c_int32), ("mDebugScale", C.c_float),
        ("mDebugCenter", C.c_int32), ("mDebugToScreenOutput", C.c_int32),
    ]


class Uniforms(C.Structure):
    """taa<CF>::uniforms_for_taa (source/taa.hpp:113-129), 544 bytes."""
    _fields_ = [
        ("mHistoryViewProjMatrix", C.c_float * 16), ("mInverseViewProjMatrix", C.c_float * 16),
        ("param", Params * 2), ("mJitterNdc", C.c_float * 4), ("mSinTime", C.c_float * 4),
        ("splitScreen", C.c_int32), ("splitX", C.c_int32), ("mUpsampling", C.c_int32),
        ("mBypassHistoryUpdate", C.c_int32), ("mResetHistory", C.c_int32), ("mCamNearPlane", C.c_float),
        ("mCamFarPlane", C.c_float), ("pad1", C.c_float),
    ]


class CasConsts(C.Structure):
    """push_constants_for_cas (source/taa.hpp:88-91), 32 bytes."""
    _fields_ = [("const0", C.c_uint32 * 4), ("const1", C.c_uint32 * 4)]


class FrameImages(C.Structure):
    _fields_ = [
        ("color", C.c_void_p), ("depth", C.c_void_p), ("history", C.c_void_p), ("history_depth", C.c_void_p),
        ("result_screen", C.c_void_p), ("debug", C.c_void_p), ("velocity", C.c_void_p),
        ("result_history", C.c_void_p), ("inspect", C.c_void_p),
        ("width", C.c_int32), ("height", C.c_int32), ("pitch", C.c_int32), ("color_format", C.c_int32),
        ("velocity_format", C.c_int32), ("input_width", C.c_int32), ("input_height", C.c_int32),
        ("input_pitch", C.c_int32), ("reserved", C.c_int32 * 2),
    ]


class JitterConfig(C.Structure):
    _fields_ = [
        ("sample_distribution", C.c_int32), ("slow_motion", C.c_int32), ("fixed_index", C.c_int32),
        ("rotate_degrees", C.c_float), ("extra_scale", C.c_float), ("num_user_offsets", C.c_int32),
        ("user_offsets", C.POINTER(C.c_float)),
    ]


class SessionIO(C.Structure):
    _fields_ = [
        ("color", C.c_void_p), ("depth", C.c_void_p), ("velocity", C.c_void_p), ("history", C.c_void_p),
        ("width", C.c_int32), ("height", C.c_int32), ("pitch", C.c_int32), ("color_format", C.c_int32),
        ("velocity_format", C.c_int32), ("frames_resolved", C.c_uint64),
    ]


class BatchItem(C.Structure):
    _fields_ = [("session", C.c_void_p), ("uniforms", C.POINTER(Uniforms)), ("cas", C.POINTER(CasConsts)),
                ("result_screen", C.c_void_p)]


class TileShape(C.Structure):
    _fields_ = [("tile_w", C.c_int32), ("tile_h", C.c_int32), ("lds_pad", C.c_int32)]


class SceneConfig(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("num_polygons", C.c_int32), ("seed", C.c_uint64),
        ("lum_mu", C.c_float), ("lum_sigma", C.c_float), ("firefly_fraction", C.c_float),
        ("firefly_min", C.c_float), ("firefly_max", C.c_float), ("pan_mode", C.c_int32), ("pan", C.c_float),
        ("object_speed", C.c_float), ("jitter_distribution", C.c_int32), ("jitter_phase", C.c_int32),
        ("color_max", C.c_float), ("depth_min", C.c_float), ("depth_max", C.c_float),
    ]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("hbm_budget_bytes", C.c_uint64), ("color_format", C.c_int32),
                ("arith", C.c_int32), ("max_batch", C.c_int32), ("max_inflight", C.c_int32),
                ("cas_sharpness", C.c_float)]


class Completion(C.Structure):
    _fields_ = [("session_id", C.c_int32), ("status", C.c_int32), ("tag", C.c_uint64), ("t_submit_us", C.c_double),
                ("t_done_us", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("frames", C.c_uint64), ("pixels", C.c_uint64),
                ("frames_failed", C.c_uint64), ("sessions_open", C.c_uint64), ("admissions_refused", C.c_uint64), ("hbm_used", C.c_uint64),
                ("hbm_budget", C.c_uint64), ("busy_us", C.c_double)]


class ServeSimConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("seed", C.c_uint64), ("mode", C.c_int32), ("sessions", C.c_int32),
                ("n_classes", C.c_int32), ("class_w", C.c_int32 * 8), ("class_h", C.c_int32 * 8),
                ("class_p", C.c_float * 8), ("arrival_rate", C.c_double), ("churn_per_s", C.c_double),
                ("duration_s", C.c_double), ("hbm_budget_bytes", C.c_uint64), ("color_format", C.c_int32),
                ("arith", C.c_int32), ("cas_sharpness", C.c_float), ("max_batch", C.c_int32),
                ("max_inflight", C.c_int32), ("warmup_frames", C.c_int32), ("exact_mix", C.c_int32),
                ("per_session_inputs", C.c_int32), ("pan_max", C.c_float)]


class ServeReport(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("pixels", C.c_uint64), ("batches", C.c_uint64), ("admissions", C.c_uint64),
                ("admissions_refused", C.c_uint64), ("sessions_end", C.c_int32), ("sessions_peak", C.c_int32),
                ("wall_s", C.c_double), ("gpix_per_s", C.c_double), ("offered_fps", C.c_double),
                ("achieved_fps", C.c_double), ("lat_p50_us", C.c_double), ("lat_p90_us", C.c_double),
                ("lat_p99_us", C.c_double), ("lat_max_us", C.c_double), ("lat_mean_us", C.c_double),
                ("mean_batch", C.c_double), ("gpu_busy_frac", C.c_double), ("hbm_peak_bytes", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


assert C.sizeof(Params) == 176 and Params.mDebugMask.offset == 144
assert C.sizeof(Uniforms) == 544 and Uniforms.param.offset == 128 and Uniforms.mJitterNdc.offset == 480
assert Uniforms.mResetHistory.offset == 528 and Uniforms.pad1.offset == 540
assert C.sizeof(CasConsts) == 32



class EdgeAaConsts(C.Structure):
    """push_constants_for_fxaa (source/taa.hpp:93-98), 20 bytes."""
    _fields_ = [("rcp_frame", C.c_float * 2), ("subpix", C.c_float), ("edge_threshold", C.c_float),
                ("edge_threshold_min", C.c_float)]


assert C.sizeof(EdgeAaConsts) == 20


TAA_SETTINGS_MAX_SAMPLE_OFFSETS = 256


class Settings(C.Structure):
    """taa_settings_t (include/taastar/taa_settings.h): the TAA sections of the settings file."""
    _fields_ = [
        ("param", Params * 2), ("taa_enabled", C.c_int32), ("sample_distribution", C.c_int32),
        ("sharpener", C.c_int32), ("sharpen_factor", C.c_float), ("split_screen", C.c_int32), ("split_x", C.c_int32),
        ("fixed_jitter_index", C.c_int32), ("jitter_extra_scale", C.c_float), ("jitter_slow_motion", C.c_int32),
        ("jitter_rotate_degrees", C.c_float), ("reset_history_on_change", C.c_int32),
        ("num_sample_offsets", C.c_int32), ("sample_offsets", (C.c_float * 2) * TAA_SETTINGS_MAX_SAMPLE_OFFSETS),
        ("postprocess_enabled", C.c_int32), ("zoom", C.c_uint32), ("show_zoom_box", C.c_uint32),
        ("zoom_src_ltwh", C.c_int32 * 4), ("zoom_dst_ltwh", C.c_int32 * 4),
    ]


class SegmentationImages(C.Structure):
    """taa_segmentation_images_t: the segmentation mask's extra images (taa.comp:34-37)."""
    _fields_ = [
        ("material", C.c_void_p), ("prev_material", C.c_void_p), ("uv_normal", C.c_void_p),
        ("prev_segmask", C.c_void_p), ("segmask", C.c_void_p), ("pitch", C.c_int32), ("input_pitch", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


assert C.sizeof(SegmentationImages) == 56


class Reproject(C.Structure):
    """taa_reproject_t (include/taastar/taa_velocity.h): the velocity producer's camera and objects."""
    _fields_ = [
        ("inverse_view_proj", C.c_float * 16), ("prev_view_proj", C.c_float * 16), ("jitter_current_prev", C.c_float * 4),
        ("object_prev_from_current", C.c_void_p), ("n_objects", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


assert C.sizeof(Reproject) == 168 and Reproject.n_objects.offset == 152

P = C.POINTER
I32, U32, U64, F32, VP = C.c_int32, C.c_uint32, C.c_uint64, C.c_float, C.c_void_p

# name -> (restype, argtypes); the exported C ABI of libtaastar.so
PROTOTYPES = {
    "taa_abi_version": (I32, []),
    "taa_last_error": (C.c_char_p, []),
    "taa_params_defaults": (I32, [P(Params)]),
    "taa_cas_setup": (I32, [P(U32), P(U32), F32, F32, F32, F32, F32]),
    "taa_halton": (F32, [I32, I32]),
    "taa_jitter_offset": (I32, [P(JitterConfig), U64, U32, U32, P(F32), P(U32)]),
    "taa_resolve_frame": (I32, [P(Uniforms), P(FrameImages), P(CasConsts), I32, VP]),
    "taa_resolve_frame_segmented": (I32, [P(Uniforms), P(FrameImages), P(SegmentationImages), P(CasConsts), I32, VP]),
    "taa_resolve_kernel": (I32, [P(Uniforms), P(FrameImages), P(CasConsts), I32, P(I32)]),
    "taa_resolve_frame_noalloc": (I32, [P(Uniforms), P(FrameImages), P(CasConsts), I32, VP]),
    "taa_cas_apply": (I32, [VP, VP, I32, I32, I32, I32, P(CasConsts), VP]),
    "taa_settings_defaults": (I32, [P(Settings)]),
    "taa_settings_read_ini": (I32, [C.c_char_p, C.c_size_t, P(Settings)]),
    "taa_settings_write_ini": (I32, [P(Settings), C.c_char_p, C.c_size_t, P(C.c_size_t)]),
    "taa_settings_apply": (I32, [P(Settings), P(Uniforms), P(JitterConfig)]),
    "taa_edge_aa_apply": (I32, [VP, VP, VP, I32, I32, I32, I32, I32, P(EdgeAaConsts), VP]),
    "taa_edge_aa_defaults": (I32, [I32, I32, P(EdgeAaConsts)]),
    "taa_velocity_from_depth": (I32, [P(Reproject), VP, VP, VP, I32, I32, I32, I32, VP]),
    "taa_ctx_create": (I32, [I32, U64, P(VP)]),
    "taa_ctx_destroy": (I32, [VP]),
    "taa_ctx_usage": (I32, [VP, P(U64), P(U64)]),
    "taa_session_bytes": (U64, [I32, I32, I32, I32]),
    "taa_session_open": (I32, [VP, I32, I32, I32, I32, P(VP)]),
    "taa_session_open_scaled": (I32, [VP, I32, I32, I32, I32, I32, I32, P(VP)]),
    "taa_session_bytes_scaled": (U64, [I32, I32, I32, I32, I32, I32]),
    "taa_session_close": (I32, [VP]),
    "taa_session_io": (I32, [VP, P(SessionIO)]),
    "taa_session_resolve": (I32, [VP, P(Uniforms), P(CasConsts), VP, I32, VP]),
    "taa_session_reset": (I32, [VP]),
    "taa_batch_submit": (I32, [VP, I32, P(BatchItem), I32, VP]),
    "taa_batch_plan_create": (I32, [VP, I32, P(BatchItem), I32, P(VP)]),
    "taa_batch_plan_launch": (I32, [VP, VP]),
    "taa_batch_plan_destroy": (I32, [VP]),
    "taa_tile_shape_count": (I32, []),
    "taa_tile_shape_get": (I32, [I32, I32, P(TileShape)]),
    "taa_autotune": (I32, [VP, I32, I32, I32, I32, I32, I32, P(I32), P(F32)]),
    "taa_tile_shape_select": (I32, [I32, I32, I32, I32, I32, I32]),
    "taa_tile_shape_selected": (I32, [I32, I32, I32, I32, I32]),
    "taa_scene_defaults": (I32, [P(SceneConfig)]),
    "taa_scene_create": (I32, [P(SceneConfig), P(VP)]),
    "taa_scene_destroy": (I32, [VP]),
    "taa_scene_render": (I32, [VP, I32, P(F32), P(F32), P(F32), P(F32), I32]),
    "taa_server_create": (I32, [P(ServerConfig), P(VP)]),
    "taa_server_destroy": (I32, [VP]),
    "taa_server_open": (I32, [VP, I32, I32, P(I32)]),
    "taa_server_close": (I32, [VP, I32]),
    "taa_server_io": (I32, [VP, I32, P(SessionIO), P(VP)]),
    "taa_server_submit": (I32, [VP, I32, P(Uniforms), U64, C.c_double]),
    "taa_server_pump": (I32, [VP]),
    "taa_server_poll": (I32, [VP, P(Completion), I32, P(I32)]),
    "taa_server_drain": (I32, [VP]),
    "taa_server_stats": (I32, [VP, P(ServerStats)]),
    "taa_server_now_us": (C.c_double, []),
    "taa_serve_simulate": (I32, [P(ServeSimConfig), P(ServeReport)]),
}

_lib = None


def lib():
    """The loaded libtaastar.so. Raises if it has not been built: there is no fallback path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python taa-star_amd/build.py` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class TaaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"taastar error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != TAA_OK:
        raise TaaError(rc, lib().taa_last_error().decode(errors="replace"))
    return rc
