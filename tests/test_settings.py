"""The settings file's TAA sections (include/taastar/taa_settings.h): the writer's text is what the
reference's own INI library (mINI, compiled from the reference tree into oracle/_ref) generates
from it, values read back by that library are the struct's, reading follows its parser, and the
settings feed the uniforms and the jitter selection. Runs on the CPU."""
import ctypes as C
import os

import numpy as np
import pytest

import taastar
from taastar import Settings, read_settings, write_settings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INI_REF = os.path.join(ROOT, "oracle", "_ref", "libini_ref.so")


@pytest.fixture(scope="module")
def mini(tmp_path_factory):
    """The reference's mINI (oracle/_ref/libini_ref.so), or skip when it has not been built."""
    if not os.path.exists(INI_REF):
        pytest.skip("oracle/_ref/libini_ref.so not built (reference tree absent)")
    L = C.CDLL(INI_REF)
    L.ref_ini_regenerate.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]
    L.ref_ini_get.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]
    d = tmp_path_factory.mktemp("ini")
    a, b = str(d / "in.ini").encode(), str(d / "out.ini").encode()

    class M:
        def regenerate(self, text):
            buf = C.create_string_buffer(1 << 20)
            n = L.ref_ini_regenerate(text.encode(), a, b, buf, len(buf))
            assert n >= 0
            return buf.value.decode()

        def get(self, text, section, key):
            buf = C.create_string_buffer(4096)
            assert L.ref_ini_get(text.encode(), a, section.encode(), key.encode(), buf, len(buf)) >= 0
            return buf.value.decode()
    return M()


def randomized(seed):
    rng = np.random.default_rng(seed)
    s = read_settings("")
    for p in s.param:
        for name, ty in p._fields_:
            if name == "mDebugMask":
                for i in range(4):
                    p.mDebugMask[i] = float(rng.uniform(-2, 2))
            elif ty is C.c_float:
                setattr(p, name, float(np.float32(rng.uniform(-3, 3))))
            elif name in ("mColorClampingOrClipping", "mUseVelocityVectors", "mVelocitySampleMode", "mInterpolationMode",
                          "mDebugMode"):
                setattr(p, name, int(rng.integers(-2, 9)))
            else:
                setattr(p, name, int(rng.integers(0, 2)))
    s.sample_distribution, s.sharpener, s.split_x = 5, 2, int(rng.integers(-5, 4000))
    s.sharpen_factor, s.jitter_extra_scale, s.jitter_rotate_degrees = 0.3125, 1.75, -12.5
    s.split_screen, s.taa_enabled, s.reset_history_on_change = 1, 0, 0
    s.fixed_jitter_index, s.jitter_slow_motion = 3, 4
    s.num_sample_offsets = 5
    for i in range(5):
        s.sample_offsets[i][0], s.sample_offsets[i][1] = float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1))
    s.zoom, s.show_zoom_box, s.postprocess_enabled = 1, 0, 0
    for i in range(4):
        s.zoom_src_ltwh[i], s.zoom_dst_ltwh[i] = int(rng.integers(-100, 3000)), int(rng.integers(0, 3000))
    return s


def test_defaults_are_the_reference_members():
    s = read_settings("")
    p = taastar.default_params()
    assert bytes(s.param[0]) == bytes(p) and bytes(s.param[1]) == bytes(p)
    assert (s.taa_enabled, s.sample_distribution, s.sharpener, s.split_screen, s.split_x) == (1, 1, 0, 0, 0)
    assert (s.fixed_jitter_index, s.jitter_slow_motion, s.reset_history_on_change) == (-1, 1, 1)
    assert (s.sharpen_factor, s.jitter_extra_scale, s.jitter_rotate_degrees) == (0.5, 1.0, 0.0)
    assert s.num_sample_offsets == 1 and list(s.sample_offsets[0]) == [0.0, 0.0]
    assert list(s.zoom_src_ltwh) == [950, 530, 20, 20] and list(s.zoom_dst_ltwh) == [1710, 10, 200, 200]
    assert (s.postprocess_enabled, s.zoom, s.show_zoom_box) == (1, 0, 1)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_written_text_is_what_mini_generates(mini, seed):
    """The reference's parser + generator reproduce the writer's text byte for byte (lower-cased
    names, "key=value", section order, no final line end), and every value it reads back is the
    struct's value printed as std::to_string prints it."""
    s = randomized(seed) if seed else read_settings("")
    t = write_settings(s)
    assert mini.regenerate(t) == t
    p0 = s.param[0]
    assert mini.get(t, "TAA_Param_0", "mAlpha") == "%f" % p0.mAlpha
    assert mini.get(t, "TAA_Param_0", "mDebugMask.z") == "%f" % p0.mDebugMask[2]
    assert mini.get(t, "TAA_Param_1", "mInterpolationMode") == str(s.param[1].mInterpolationMode)
    assert mini.get(t, "taa_primary", "mSplitX") == str(s.split_x)
    assert mini.get(t, "TAA_Primary", "mResetHistoryOnChange") == ("1" if s.reset_history_on_change else "0")
    assert mini.get(t, "TAA_Primary", "mDebugSampleOffsets.size") == str(s.num_sample_offsets)
    assert mini.get(t, "TAA_Primary", "mDebugSampleOffsets_0.y") == "%f" % s.sample_offsets[0][1]
    assert mini.get(t, "TAA_Postprocess", "zoomDstLTWH.w") == str(s.zoom_dst_ltwh[3])
    # keys the reference does not store stay out of the file
    assert mini.get(t, "TAA_Param_0", "mShrinkChromaAxis") == ""
    assert mini.get(t, "TAA_Param_0", "mRayTraceAugmentFlags") == ""


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_round_trip(seed):
    """write -> read -> write is the identity on the text; every stored field survives up to the
    six decimals of "%f"."""
    s = randomized(seed)
    t = write_settings(s)
    r = read_settings(t)
    assert write_settings(r) == t
    assert r.param[1].mVarianceClipping == s.param[1].mVarianceClipping
    assert r.param[0].mDebugMode == s.param[0].mDebugMode
    np.testing.assert_allclose(r.param[0].mVarClipGamma, s.param[0].mVarClipGamma, atol=5e-7)
    assert r.num_sample_offsets == 5 and r.split_x == s.split_x and list(r.zoom_src_ltwh) == list(s.zoom_src_ltwh)


def test_reading_follows_the_reference_parser(mini):
    """Case-insensitive names, whitespace, ';' comments, other sections ignored, later duplicates win,
    empty values keep the current setting, bools are "1" or "true", numbers are their leading
    digits (strtol / strtof), a missing sample-offset count keeps one offset."""
    text = "\n".join([
        "; a settings file with more than the TAA sections",
        "[Camera]", "mAlpha=0.9",
        "[TAA_PARAM_0]  ; trailing comment",
        "  MALPHA =  0.25  ", "mVarClipGamma=1.5", "mVarClipGamma=2.5", "mMinAlpha=", "mDebugMode=7xyz",
        "mUseYCoCg=4294967295",
        "[taa_primary]", "mTaaEnabled=true", "mSplitScreen=yes", "mJitterExtraScale=2.5e0junk",
        "mDebugSampleOffsets.size=2", "mDebugSampleOffsets_1.x=0.75",
        "[TAA_Postprocess]", "zoom=7",
    ])
    base = read_settings("")
    s = read_settings(text)
    p = s.param[0]
    assert p.mAlpha == np.float32(0.25) and p.mVarClipGamma == 2.5 and p.mMinAlpha == base.param[0].mMinAlpha
    assert p.mDebugMode == 7 and (p.mUseYCoCg & 0xFFFFFFFF) == 4294967295  # stoul of a uint32 flag
    assert bytes(s.param[1]) == bytes(base.param[1])
    assert s.taa_enabled == 1 and s.split_screen == 0 and s.jitter_extra_scale == 2.5
    assert s.num_sample_offsets == 2 and list(s.sample_offsets[1]) == [0.75, 0.0]
    assert s.zoom == 7
    # what the reference's own parser reads from the same text
    assert mini.get(text, "taa_param_0", "malpha") == "0.25"
    assert mini.get(text, "TAA_Param_0", "mVarClipGamma") == "2.5"
    assert mini.get(text, "TAA_Param_0", "mMinAlpha") == ""


def test_bad_values_are_refused_whole():
    base = read_settings("")
    for bad in ("[TAA_Param_1]\nmAlpha=abc", "[TAA_Primary]\nmSplitX=--3", "[TAA_Postprocess]\nzoom=x1"):
        s = Settings()
        C.memmove(C.byref(s), C.byref(base), C.sizeof(Settings))
        raw = bad.encode()
        assert taastar.lib().taa_settings_read_ini(raw, len(raw), C.byref(s)) == taastar._abi.TAA_ERR_INVALID_ARG
        assert bytes(s) == bytes(base)
    raw = b"[TAA_Primary]\nmDebugSampleOffsets.size=100000"
    s = read_settings("")
    assert taastar.lib().taa_settings_read_ini(raw, len(raw), C.byref(s)) == taastar._abi.TAA_ERR_CAPACITY


def test_settings_feed_uniforms_and_jitter():
    """taa_settings_apply: both parameter blocks and the split into the uniforms; the jitter
    selection (distribution 5 = the user's sample offsets in pixels) into the jitter config."""
    s = randomized(4)
    U = taastar.Uniforms()
    U.mCamNearPlane = 0.25
    J = taastar.JitterConfig()
    assert taastar.lib().taa_settings_apply(C.byref(s), C.byref(U), C.byref(J)) == 0
    assert bytes(U.param[0]) == bytes(s.param[0]) and bytes(U.param[1]) == bytes(s.param[1])
    assert (U.splitScreen, U.splitX, U.mCamNearPlane) == (1, s.split_x, 0.25)
    assert (J.sample_distribution, J.fixed_index, J.slow_motion, J.num_user_offsets) == (5, 3, 4, 5)
    out, n = (C.c_float * 2)(), C.c_uint32()
    assert taastar.lib().taa_jitter_offset(C.byref(J), 0, 1920, 1080, out, C.byref(n)) == 0
    assert n.value == 5
