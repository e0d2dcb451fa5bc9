"""The C ABI library: loads without a GPU, exports every declared symbol, host helpers are exact."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import taastar
from taastar import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "taastar", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(taa_\w+)\s*\(", text, flags=re.M):
            if not m.group(0).lstrip().startswith("typedef"):
                names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(taastar_lib):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(taastar_lib, n)]
    assert not missing, missing
    # and the Python mirror binds exactly the declared set
    assert names == set(_abi.PROTOTYPES), names ^ set(_abi.PROTOTYPES)


def test_struct_layouts_match_reference():
    # source/taa.hpp:30-77, 113-129, 88-91 (the std140 offsets of the GLSL mirror, taa.comp:50-118)
    assert C.sizeof(taastar.Params) == 176
    assert C.sizeof(taastar.Uniforms) == 544
    assert C.sizeof(taastar.CasConsts) == 32
    U = taastar.Uniforms
    assert (U.param.offset, U.mJitterNdc.offset, U.mSinTime.offset, U.splitScreen.offset, U.mResetHistory.offset,
            U.mCamNearPlane.offset, U.mCamFarPlane.offset, U.pad1.offset) == (128, 480, 496, 512, 528, 532, 536, 540)
    P = taastar.Params
    assert (P.mVarClipGamma.offset, P.mRejectOutside.offset, P.mInterpolationMode.offset,
            P.mRayTraceHistoryCount.offset, P.mDebugMask.offset, P.mDebugToScreenOutput.offset) == (48, 64, 76, 140, 144, 172)


def test_abi_version(taastar_lib):
    assert taastar_lib.taa_abi_version() == 3


def test_param_defaults_match_reference():
    p = taastar.default_params()
    f32 = lambda v: float(np.float32(v))
    # source/taa.hpp:31-76
    assert p.mAlpha == f32(0.05) and p.mColorClampingOrClipping == 1 and p.mVarClipGamma == 1.0
    assert p.mMinAlpha == f32(np.float32(1.0) - np.float32(0.97)) and p.mMaxAlpha == f32(np.float32(1.0) - np.float32(0.88))
    assert p.mRejectionAlpha == 1.0 and p.mUseVelocityVectors == 1 and p.mVelocitySampleMode == 0
    assert p.mInterpolationMode == 0 and p.mNoiseFactor == f32(np.float32(1) / np.float32(510))
    assert p.mVelBasedAlphaMax == f32(0.2) and p.mVelBasedAlphaFactor == f32(np.float32(1) / np.float32(40))
    assert p.mRayTraceAugmentFlags == 0xFFFFFE7F and p.mRayTraceHistoryCount == -1
    assert list(p.mDebugMask) == [1.0, 1.0, 1.0, 0.0] and p.mDebugScale == 1.0 and p.mUnjitterFactor == 1.0
    assert not any([p.mDepthCulling, p.mPassThrough, p.mUseYCoCg, p.mVarianceClipping, p.mLumaWeightingLottes])


# CasSetup known answers at 1920x1080 from the reference's own ffx_cas.h built for the CPU
CAS_KAT = {
    0.0: ([0x3F800000, 0x3F800000, 0, 0], [0xBE000000, 0x0000B000, 0x41000000, 0]),
    0.5: ([0x3F800000, 0x3F800000, 0, 0], [0xBE1D89D9, 0x0000B0EC, 0x41000000, 0]),
    1.0: ([0x3F800000, 0x3F800000, 0, 0], [0xBE4CCCCD, 0x0000B266, 0x41000000, 0]),
}


def test_cas_setup_known_answers():
    for s, (c0, c1) in CAS_KAT.items():
        k = taastar.cas_setup(s, 1920, 1080)
        assert list(k.const0) == c0 and list(k.const1) == c1


def test_cas_setup_matches_oracle_and_reference(oracle):
    ref = oracle.ref_lib()
    for s in np.linspace(-0.25, 1.25, 61):
        for (w, h) in [(640, 360), (1920, 1080), (3840, 2160), (1000, 3)]:
            k = taastar.cas_setup(float(s), w, h)
            c0, c1 = oracle.cas_setup(float(s), w, h)
            assert list(k.const0) == c0 and list(k.const1) == c1
            if ref is not None:
                r0, r1 = (C.c_uint32 * 4)(), (C.c_uint32 * 4)()
                ref.ref_cas_setup(r0, r1, float(s), w, h, w, h)
                assert list(r0) == c0 and list(r1) == c1


def test_halton_and_jitter_match_oracle(oracle):
    for b in (2, 3, 5):
        for i in range(1, 70):
            assert taastar.halton(i, b) == oracle.lib().oracle_halton(i, b)
    for dist in range(5):
        for (w, h) in [(640, 360), (1920, 1080), (1281, 719)]:
            for f in range(40):
                for slow, fixed, rot, sc in [(0, -1, 0.0, 1.0), (3, -1, 0.0, 1.0), (0, 5, 0.0, 1.0), (0, -1, 30.0, 2.0)]:
                    got, n = taastar.jitter_offset(f, w, h, dist, slow, fixed, rot, sc)
                    exp, n2 = oracle.jitter(dist, f, w, h, slow, fixed, rot, sc)
                    assert n == n2
                    assert got == exp or np.allclose(got, exp, rtol=0, atol=1e-7), (dist, w, h, f, got, exp)
    user = [[0.1, -0.2], [0.3, 0.4], [-0.45, 0.05]]
    for f in range(6):
        assert taastar.jitter_offset(f, 800, 600, 5, user_offsets=user) == oracle.jitter(5, f, 800, 600, user=user)


def test_session_bytes():
    lib = taastar.lib()
    w, h = 1920, 1080
    px = w * h
    a = lambda v: (v + 255) // 256 * 256
    assert lib.taa_session_bytes(w, h, 0, 0) == 3 * a(px * 16) + 2 * a(px * 4) + a(px * 8)
    assert lib.taa_session_bytes(w, h, 1, 1) == 3 * a(px * 8) + 2 * a(px * 4) + a(px * 4)
    assert lib.taa_session_bytes(0, h, 1, 1) == 0 and lib.taa_session_bytes(w, h, 7, 1) == 0
    # a scaled (TAAU) session: history ping-pong at the output size, input slots and depths at the input size
    wl, hl = 1280, 720
    pl = wl * hl
    assert lib.taa_session_bytes_scaled(w, h, wl, hl, 1, 1) == 2 * a(px * 8) + a(pl * 8) + 2 * a(pl * 4) + a(pl * 4)
    assert lib.taa_session_bytes_scaled(w, h, w, h, 0, 2) == lib.taa_session_bytes(w, h, 0, 2)
    assert lib.taa_session_bytes_scaled(w, h, 0, hl, 1, 1) == 0


def test_argument_errors_need_no_gpu():
    lib = taastar.lib()
    U = taastar.Uniforms()
    img = taastar.FrameImages()
    # NULLs, bad sizes and unsupported features are refused before anything touches a device
    assert lib.taa_resolve_frame(None, C.byref(img), None, 0, None) == _abi.TAA_ERR_INVALID_ARG
    assert lib.taa_resolve_frame(C.byref(U), C.byref(img), None, 0, None) == _abi.TAA_ERR_INVALID_ARG
    U.param[0].mRayTraceAugment = 1
    assert lib.taa_resolve_frame(C.byref(U), C.byref(img), None, 0, None) == _abi.TAA_ERR_UNSUPPORTED
    assert b"mRayTraceAugment" in lib.taa_last_error()
    assert lib.taa_cas_apply(None, None, 4, 4, 4, 0, None, None) == _abi.TAA_ERR_INVALID_ARG
    # images address with 32-bit byte offsets: a frame whose 16-byte image would reach 4 GiB is refused
    U.param[0].mRayTraceAugment = 0
    big = taastar.FrameImages()
    for f in ("color", "depth", "history", "history_depth", "result_screen", "velocity", "result_history"):
        setattr(big, f, 4096 * (1 + len(f)))
    big.width, big.height, big.pitch = 16384, 16384, 16384
    assert lib.taa_resolve_frame(C.byref(U), C.byref(big), None, 0, None) == _abi.TAA_ERR_INVALID_ARG
    assert b"4 GiB" in lib.taa_last_error()
    # a bad input (render) size and non-zero reserved words are refused as well
    big.width, big.height, big.pitch = 64, 32, 64
    big.input_width, big.input_height = -3, 16
    assert lib.taa_resolve_frame(C.byref(U), C.byref(big), None, 0, None) == _abi.TAA_ERR_INVALID_ARG
    big.input_width, big.input_height, big.input_pitch = 32, 16, 8
    assert lib.taa_resolve_frame(C.byref(U), C.byref(big), None, 0, None) == _abi.TAA_ERR_INVALID_ARG
    big.input_pitch = 0
    big.reserved[1] = 5
    assert lib.taa_resolve_frame(C.byref(U), C.byref(big), None, 0, None) == _abi.TAA_ERR_INVALID_ARG
    assert b"reserved" in lib.taa_last_error()
    assert lib.taa_session_close(None) == _abi.TAA_ERR_INVALID_ARG
    # outputs may share a buffer with neither an input nor another output (other workgroups read
    # their aprons and history footprints while a tile is written)
    good = taastar.FrameImages()
    names = ("color", "depth", "history", "history_depth", "result_screen", "velocity", "result_history", "debug", "inspect")
    for i, f in enumerate(names):
        setattr(good, f, 0x10000 * (i + 1))
    good.width, good.height, good.pitch = 64, 32, 64
    for out_field, in_field in (("result_screen", "color"), ("result_screen", "history_depth"), ("result_history", "velocity"),
                                ("debug", "depth"), ("inspect", "history"), ("debug", "result_screen")):
        bad = taastar.FrameImages.from_buffer_copy(good)
        setattr(bad, out_field, getattr(good, in_field))
        assert lib.taa_resolve_frame(C.byref(U), C.byref(bad), None, 0, None) == _abi.TAA_ERR_INVALID_ARG, out_field
        assert b"alias" in lib.taa_last_error(), (out_field, in_field, lib.taa_last_error())


def test_resolve_kernel_query_needs_no_gpu():
    """taa_resolve_kernel names the kernel a resolve runs, and why an FP16 request falls back."""
    from taastar.workloads import params_c2
    lib = taastar.lib()
    img = taastar.FrameImages()
    for i, f in enumerate(("color", "depth", "history", "history_depth", "result_screen", "velocity", "result_history")):
        setattr(img, f, 0x10000 * (i + 1))
    img.width, img.height, img.pitch = 64, 32, 64
    img.color_format, img.velocity_format = taastar.FMT_RGBA16F, taastar.VEL_RG16F
    U = taastar.Uniforms()
    U.param[0] = params_c2()
    U.param[1] = params_c2()
    k = (C.c_int32 * 2)()

    def query(arith=taastar.ARITH_FP16, images=img):
        taastar.check(lib.taa_resolve_kernel(C.byref(U), C.byref(images), None, arith, k))
        return (k[0], k[1]), lib.taa_last_error().decode()

    assert query() == ((taastar.KERNEL_FP16, taastar.KERNEL_FP16), "")
    assert query(taastar.ARITH_FP32) == ((taastar.KERNEL_EXACT, taastar.KERNEL_EXACT), "")
    # the reference's own defaults: RGBA16F velocity, movers only, a debug image bound
    img.velocity_format = taastar.VEL_RGBA16F
    img.debug = 0x90000
    U.param[0] = taastar.default_params()
    U.param[1] = taastar.default_params()
    assert query() == ((taastar.KERNEL_FP16, taastar.KERNEL_FP16), "")
    img.velocity_format = taastar.VEL_RGBA32F  # another precision than the RGBA16F colour
    (a, b), why = query()
    assert (a, b) == (taastar.KERNEL_EXACT, taastar.KERNEL_EXACT) and "velocity format" in why, why
    img.velocity_format = taastar.VEL_RG16F
    U.param[0].mToneMapLumaKaris = 1
    (a, b), why = query()
    assert a == taastar.KERNEL_EXACT and "mToneMapLumaKaris" in why, why
    U.param[0].mToneMapLumaKaris = 0
    img.debug = None
    U.splitScreen, U.splitX = 1, 20
    U.param[1].mShrinkChromaAxis = 1
    (a, b), why = query()
    assert (a, b) == (taastar.KERNEL_FP16, taastar.KERNEL_EXACT) and "mShrinkChromaAxis (param[1])" in why, why
    U.splitScreen = 0
    img.input_width, img.input_height = 32, 16  # upsampling: its own kernel, in the FP16 path's arithmetic
    assert query() == ((taastar.KERNEL_UPSAMPLE_FP16, taastar.KERNEL_UPSAMPLE_FP16), "")
    assert query(taastar.ARITH_FP32)[0] == (taastar.KERNEL_UPSAMPLE, taastar.KERNEL_UPSAMPLE)
    img.input_width = img.input_height = 0
    assert query()[0] == (taastar.KERNEL_FP16, taastar.KERNEL_FP16)
    # arguments are checked as the resolve checks them
    assert lib.taa_resolve_kernel(C.byref(U), C.byref(img), None, 7, k) == _abi.TAA_ERR_INVALID_ARG
    assert lib.taa_resolve_kernel(C.byref(U), C.byref(img), None, 0, None) == _abi.TAA_ERR_INVALID_ARG


def test_noalloc_refuses_an_oversize_intermediate_without_a_gpu():
    """taa_resolve_frame_noalloc refuses a split screen with CAS whose intermediate screen exceeds
    its fixed array before it touches a device."""
    lib = taastar.lib()
    img = taastar.FrameImages()
    for i, f in enumerate(("color", "depth", "history", "history_depth", "result_screen", "velocity", "result_history")):
        setattr(img, f, 0x10000 * (i + 1))
    img.width, img.height, img.pitch = 3840, 2160, 3840
    img.color_format, img.velocity_format = taastar.FMT_RGBA32F, taastar.VEL_RG32F  # 16 B/texel: 133 MB
    U = taastar.Uniforms()
    U.splitScreen, U.splitX = 1, 100
    K = taastar.cas_setup(0.5, 3840, 2160)
    assert lib.taa_resolve_frame_noalloc(C.byref(U), C.byref(img), C.byref(K), 0, None) == _abi.TAA_ERR_CAPACITY
    msg = lib.taa_last_error().decode()
    assert "132710400" in msg and str(_abi.NOALLOC_SCRATCH_BYTES) in msg, msg


def test_velocity_argument_errors_need_no_gpu():
    lib = taastar.lib()
    R = taastar.Reproject()
    dep, vel, ids = 4096, 1 << 20, 1 << 21   # never dereferenced: every call below is refused first
    call = lambda r, d, i, v, fmt=_abi.VEL_RG16F, w=64, h=32, p=0: lib.taa_velocity_from_depth(
        C.byref(r) if r is not None else None, d, i, v, fmt, w, h, p, None)
    assert call(None, dep, None, vel) == _abi.TAA_ERR_INVALID_ARG
    assert call(R, None, None, vel) == _abi.TAA_ERR_INVALID_ARG
    assert call(R, dep, None, None) == _abi.TAA_ERR_INVALID_ARG
    assert call(R, dep, None, vel, fmt=7) == _abi.TAA_ERR_INVALID_ARG
    assert call(R, dep, None, vel, w=0) == _abi.TAA_ERR_INVALID_ARG
    assert call(R, dep, None, vel, p=10) == _abi.TAA_ERR_INVALID_ARG
    assert call(R, dep, None, vel, w=16384, h=16384) == _abi.TAA_ERR_INVALID_ARG
    assert b"4 GiB" in lib.taa_last_error()
    assert call(R, dep, None, dep) == _abi.TAA_ERR_INVALID_ARG
    assert call(R, dep, vel, vel) == _abi.TAA_ERR_INVALID_ARG
    R.n_objects = 2
    assert call(R, dep, ids, vel) == _abi.TAA_ERR_INVALID_ARG
    assert b"object_prev_from_current" in lib.taa_last_error()
    R.object_prev_from_current = 4096 + 8
    assert call(R, dep, ids, vel) == _abi.TAA_ERR_INVALID_ARG
    assert b"aligned" in lib.taa_last_error()
    R.n_objects, R.object_prev_from_current = -1, None
    assert call(R, dep, ids, vel) == _abi.TAA_ERR_INVALID_ARG
    R.n_objects = 0
    R.reserved[2] = 1
    assert call(R, dep, None, vel) == _abi.TAA_ERR_INVALID_ARG
    assert b"reserved" in lib.taa_last_error()


def test_tile_shapes_table():
    lib = taastar.lib()
    n = lib.taa_tile_shape_count()
    assert n >= 3
    for cas in (0, 1):
        for i in range(n):
            s = taastar.TileShape()
            assert lib.taa_tile_shape_get(i, cas, C.byref(s)) == 0
            assert s.tile_w > 0 and s.tile_h > 0
            region = (s.tile_w + 2 * cas) * (s.tile_h + 2 * cas)
            assert region >= 196
    assert lib.taa_tile_shape_select(640, 360, 0, 0, 0, 2) == 0
    assert lib.taa_tile_shape_selected(640, 360, 0, 0, 0) == 2
    assert lib.taa_tile_shape_select(640, 360, 0, 0, 0, -1) == 0
    assert lib.taa_tile_shape_select(640, 360, 0, 0, 0, 99) == _abi.TAA_ERR_INVALID_ARG


def test_default_cas_tile_shape_follows_the_frame_size():
    """Untuned classes with CAS: 30x30 once a frame holds 1,280 such tiles (one per workgroup the
    device holds at once), 30x14 below that, 14x14 below 512; an entry of the table overrides it and
    -1 restores the built-in choice."""
    lib = taastar.lib()

    def shape(i):
        s = taastar.TileShape()
        assert lib.taa_tile_shape_get(i, 1, C.byref(s)) == 0
        return s.tile_w, s.tile_h

    fmt, arith = taastar.FMT_RGBA16F, taastar.ARITH_FP16
    for (w, h), want in (((640, 360), (14, 14)), ((203, 81), (14, 14)), ((1280, 720), (30, 14)),
                         ((1920, 1080), (30, 30)), ((3840, 2160), (30, 30)), ((7680, 4320), (30, 30))):
        n30 = -(-w // 30) * -(-h // 30)
        assert (n30 < 512) == (want == (14, 14)) and (512 <= n30 < 1280) == (want == (30, 14))
        assert shape(lib.taa_tile_shape_selected(w, h, fmt, arith, 1)) == want, (w, h)
    assert lib.taa_tile_shape_selected(640, 360, fmt, arith, 0) == lib.taa_tile_shape_selected(3840, 2160, fmt, arith, 0)
    assert lib.taa_tile_shape_select(1280, 720, fmt, arith, 1, 2) == 0
    assert shape(lib.taa_tile_shape_selected(1280, 720, fmt, arith, 1)) == (30, 30)
    assert lib.taa_tile_shape_select(1280, 720, fmt, arith, 1, -1) == 0
    assert shape(lib.taa_tile_shape_selected(1280, 720, fmt, arith, 1)) == (30, 14)


def test_scene_is_deterministic_and_well_formed():
    sc = taastar.Scene(width=160, height=96, num_polygons=40, seed=123)
    c1, d1, v1, j1 = sc.render(3, threads=1)
    c2, d2, v2, j2 = sc.render(3, threads=7)
    assert np.array_equal(c1, c2) and np.array_equal(d1, d2) and np.array_equal(v1, v2) and j1 == j2
    assert np.all(np.isfinite(c1)) and np.all(c1[..., :3] >= 0) and np.all(c1[..., 3] == 1)
    assert np.all((d1 >= 0.1) & (d1 <= 1.0)) and (d1 < 1.0).mean() > 0.5
    # jitter of frame 3 is Halton(2,3) phase 3 scaled to NDC
    (jx, jy), _ = taastar.jitter_offset(3, 160, 96, 2)
    assert j1 == (jx, jy)
    # ~1% fireflies over a larger frame
    sc2 = taastar.Scene(width=400, height=300, num_polygons=0, seed=5, firefly_fraction=0.01)
    c, _, v, _ = sc2.render(0)
    frac = (c[..., 1] > 0.5).mean()
    assert 0.006 < frac < 0.014
    # background moves with the pan: velocity is the pan in UV units, mover flag 0
    assert np.allclose(v[..., 2], 0) and np.all(v[..., 3] == 0)


@pytest.mark.parametrize("fmt", [0, 1])
def test_no_gpu_context_is_refused(fmt):
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert taastar.lib().taa_ctx_create(0, 0, C.byref(h)) == _abi.TAA_ERR_NO_DEVICE
