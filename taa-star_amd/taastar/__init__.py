"""taastar: MI355X-native TAA resolve + fused CAS (cg-tuwien/TAA-STAR taa.comp + sharpen_cas.comp).

Python face of the C ABI in include/taastar/ (taa_abi.h, taa_velocity.h, taa_serve.h, taa_scene.h). The compute runs in libtaastar.so
(hand-written gfx950 kernels); torch is used only to own device buffers and streams.
"""
import ctypes as C

import numpy as np

from ._abi import (ARITH_FP16, ARITH_FP32, KERNEL_EXACT, KERNEL_FP16, KERNEL_UPSAMPLE, KERNEL_UPSAMPLE_FP16, FLAG_DEPTH_REJECT, FLAG_GHOST_REJECT, FLAG_OUTSIDE, FLAG_RECTIFIED,
                   FLAG_REJECTED, FLAG_USED_VELOCITY, FMT_RGBA16F, FMT_RGBA32F, INSPECT, INSPECT_WORDS, LIB_PATH,
                   PROTOTYPES, VEL_RG16F, VEL_RG32F, VEL_RGBA16F, VEL_RGBA32F, BatchItem, CasConsts, FrameImages,
                   JitterConfig, Params, SceneConfig, ServeReport, ServerConfig, ServerStats, ServeSimConfig,
                   RTFLAG, SEG_EDGE_AA, SEG_RAY_TRACE, SEG_TAA, Completion, EdgeAaConsts, Reproject, SegmentationImages,
                   SessionIO, Settings,
                   TaaError, TileShape, Uniforms, check, lib)

__all__ = [
    "lib", "check", "TaaError", "Params", "Uniforms", "CasConsts", "FrameImages", "JitterConfig", "SessionIO",
    "BatchItem", "Settings", "read_settings", "write_settings", "EdgeAaConsts", "Reproject", "SegmentationImages", "RTFLAG", "SEG_TAA", "SEG_EDGE_AA", "SEG_RAY_TRACE", "TileShape", "SceneConfig", "ServerConfig", "ServerStats", "ServeSimConfig", "ServeReport",
    "Completion", "default_params", "cas_setup", "halton", "jitter_offset", "Scene",
    "FMT_RGBA32F", "FMT_RGBA16F", "VEL_RG32F", "VEL_RG16F", "VEL_RGBA32F", "VEL_RGBA16F", "ARITH_FP32",
    "ARITH_FP16", "KERNEL_EXACT", "KERNEL_FP16", "KERNEL_UPSAMPLE", "KERNEL_UPSAMPLE_FP16", "INSPECT", "INSPECT_WORDS", "LIB_PATH", "PROTOTYPES",
]


def default_params():
    p = Params()
    check(lib().taa_params_defaults(C.byref(p)))
    return p


def cas_setup(sharpness, in_w, in_h, out_w=None, out_h=None):
    """CasSetup (shaders/ffx_cas.h:375-394) as called at source/taa.hpp:965."""
    k = CasConsts()
    out_w = in_w if out_w is None else out_w
    out_h = in_h if out_h is None else out_h
    check(lib().taa_cas_setup(k.const0, k.const1, sharpness, float(in_w), float(in_h), float(out_w), float(out_h)))
    return k


def halton(i, b):
    return lib().taa_halton(i, b)


def jitter_offset(frame, width, height, distribution=2, slow_motion=0, fixed_index=-1, rotate_degrees=0.0,
                  extra_scale=1.0, user_offsets=None):
    cfg = JitterConfig()
    cfg.sample_distribution = distribution
    cfg.slow_motion = slow_motion
    cfg.fixed_index = fixed_index
    cfg.rotate_degrees = rotate_degrees
    cfg.extra_scale = extra_scale
    keep = None
    if user_offsets is not None:
        keep = np.ascontiguousarray(np.asarray(user_offsets, np.float32).reshape(-1))
        cfg.num_user_offsets = keep.size // 2
        cfg.user_offsets = keep.ctypes.data_as(C.POINTER(C.c_float))
    out = (C.c_float * 2)()
    n = C.c_uint32()
    check(lib().taa_jitter_offset(C.byref(cfg), frame, width, height, out, C.byref(n)))
    return (out[0], out[1]), n.value


class Scene:
    """Seeded synthetic scene (include/taastar/taa_scene.h) rendered on the host."""

    def __init__(self, **kw):
        cfg = SceneConfig()
        check(lib().taa_scene_defaults(C.byref(cfg)))
        for k, v in kw.items():
            if not hasattr(cfg, k):
                raise KeyError(k)
            setattr(cfg, k, v)
        self.cfg = cfg
        self._h = C.c_void_p()
        check(lib().taa_scene_create(C.byref(cfg), C.byref(self._h)))

    @property
    def width(self):
        return self.cfg.width

    @property
    def height(self):
        return self.cfg.height

    def render(self, frame, threads=0):
        """-> (colour RGBA f32 [H,W,4], depth f32 [H,W], velocity RGBA f32 [H,W,4], jitter_ndc (x, y))"""
        H, W = self.cfg.height, self.cfg.width
        col = np.empty((H, W, 4), np.float32)
        dep = np.empty((H, W), np.float32)
        vel = np.empty((H, W, 4), np.float32)
        jit = (C.c_float * 2)()
        fp = C.POINTER(C.c_float)
        check(lib().taa_scene_render(self._h, frame, col.ctypes.data_as(fp), dep.ctypes.data_as(fp),
                                     vel.ctypes.data_as(fp), jit, threads))
        return col, dep, vel, (jit[0], jit[1])

    def close(self):
        if self._h:
            lib().taa_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_settings(text, base=None):
    """Apply the TAA sections of an INI text (the reference's settings file) to `base` (default:
    the reference's defaults); returns the Settings."""
    s = Settings()
    if base is None:
        check(lib().taa_settings_defaults(C.byref(s)))
    else:
        C.memmove(C.byref(s), C.byref(base), C.sizeof(Settings))
    raw = text.encode() if isinstance(text, str) else bytes(text)
    check(lib().taa_settings_read_ini(raw, len(raw), C.byref(s)))
    return s


def write_settings(settings):
    """The TAA sections of a settings file, as the reference writes them."""
    n = C.c_size_t()
    lib().taa_settings_write_ini(C.byref(settings), None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    check(lib().taa_settings_write_ini(C.byref(settings), buf, n.value + 1, C.byref(n)))
    return buf.value.decode()
