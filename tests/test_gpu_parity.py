"""GPU FP32 path vs the FP32 transcription of taa.comp + CasFilter, through the C ABI.

Bar (FP32 arithmetic): reprojected texel indices, closest-depth taps, rejection flags and the
neighbourhood box bit-exact; history and screen within 1e-5 relative per channel
(|d| <= 1e-5 * max(|ref|, 1e-3)).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import Sequence, assert_parity, rel_err  # noqa: E402
from taastar import FMT_RGBA16F, FMT_RGBA32F, VEL_RG16F, VEL_RG32F, VEL_RGBA16F, VEL_RGBA32F  # noqa: E402
from taastar.workloads import params_c1, params_c2, workload  # noqa: E402


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_c1_full_sequence_bit_exact(oracle):
    """640x360, 16 frames, FP32 storage: the first workload at its real size, every frame."""
    print("case c1 640x360 x16 fp32", flush=True)
    seq = Sequence(workload("c1"), cas=False)
    for f in range(16):
        out = seq.step(oracle)
        assert_parity(out)
        assert np.array_equal(out["gpu_history"], out["oracle_history"]), f"frame {f}: history not bit-identical"
        assert np.array_equal(out["gpu_screen"], out["oracle_screen"]), f"frame {f}: screen not bit-identical"


@pytest.mark.parametrize("fmt,vfmt", [(FMT_RGBA32F, VEL_RG32F), (FMT_RGBA16F, VEL_RG16F)])
def test_c2_features_fused_cas(oracle, fmt, vfmt):
    """c2 feature set (closest-depth velocity, linear depth rejection, Lottes, reject outside,
    fused CAS 0.5) at 480x272, 8 frames, FP32 and FP16 storage; without the inspect image the same
    frames take the exact kernel's production form, which must write the same bits."""
    print(f"case c2 480x272 x8 fmt={fmt}", flush=True)
    seq = Sequence(workload("c2", fmt=fmt, vel=vfmt), width=480, height=272, cas=True, debug=False, shipped=True)
    for f in range(8):
        out = seq.step(oracle)
        assert_parity(out)


FEATURES = {
    "minmax_rgb": dict(mVarianceClipping=0, mUseYCoCg=0, mColorClampingOrClipping=1),
    "shaped_shrink": dict(mVarianceClipping=0, mShapedNeighbourhood=1, mShrinkChromaAxis=1),
    "slow_clip": dict(mColorClampingOrClipping=3),
    "bilinear_hist": dict(mInterpolationMode=0),
    "bspline_hist": dict(mInterpolationMode=1),
    "longest_vel": dict(mVelocitySampleMode=1),
    "movers_only": dict(mUseVelocityVectors=1),
    "matrix_reproj": dict(mUseVelocityVectors=0),
    "raw_depth_cull": dict(mDepthCulling=1),
    "reduce_near_clamp": dict(mReduceBlendNearClamp=1),
    "vel_alpha": dict(mVelBasedAlpha=1),
    "karis": dict(mToneMapLumaKaris=1, mVarClipGamma=1.25),
    "antighost": dict(mDynamicAntiGhosting=1),
    "unjitter": dict(mUnjitterNeighbourhood=1, mUnjitterCurrentSample=1),
    # the neighbourhood unjittered, the current sample not: getCurrentColor fetches the texel
    # itself instead of reusing the neighbourhood's centre (taa.comp:216-220, 750-753)
    "unjitter_neighbourhood_only": dict(mUnjitterNeighbourhood=1),
    "passthrough": dict(mPassThrough=1),
}


@pytest.mark.parametrize("name", sorted(FEATURES))
@pytest.mark.parametrize("vfmt", [VEL_RG32F, VEL_RGBA32F])
def test_feature_matrix(oracle, name, vfmt):
    """Every option of Parameters the kernel implements, on 160x96 frames, 4 frames each."""
    print(f"case feature {name} vfmt={vfmt}", flush=True)
    p = params_c2()
    for k, v in FEATURES[name].items():
        setattr(p, k, v)
    w = workload("c2", vel=vfmt, params=lambda p=p: p)
    seq = Sequence(w, width=160, height=96, cas=True)
    for f in range(4):
        out = seq.step(oracle)
        assert_parity(out)


def test_debug_modes_and_bypass(oracle):
    """uDebug output (taa.comp:912-942) for modes 0..8, and mBypassHistoryUpdate."""
    from taastar.workloads import uniforms_for
    w = workload("c2")
    seq = Sequence(w, width=96, height=64, cas=False)
    seq.step(oracle)
    for mode in range(9):
        p = params_c2()
        p.mDebugMode = mode
        p.mDebugScale = 2.0
        p.mDebugCenter = mode % 2
        _, _, _, jit = seq.scene.render(seq.frame)
        U = uniforms_for(w, seq.frame, jit, params=p)
        U.mBypassHistoryUpdate = int(mode == 8)
        out = seq.step(oracle, uniforms=U)
        assert_parity(out)
        g, o = out["gpu_debug"], out["oracle_debug"]
        assert np.array_equal(np.isfinite(g), np.isfinite(o))
        fin = np.isfinite(o)
        assert rel_err(g[fin], o[fin], floor=1e-2).max() <= 2e-3, f"debug mode {mode}"


def test_split_screen(oracle):
    """splitScreen: columns x > splitX use param[1] (taa.comp:717), two launches."""
    from taastar.workloads import uniforms_for
    w = workload("c2")
    seq = Sequence(w, width=200, height=80, cas=False)
    for f in range(3):
        col, dep, vel4, jit = seq.scene.render(seq.frame)
        U = uniforms_for(w, seq.frame, jit)
        U.splitScreen = 1
        U.splitX = 77
        U.param[1].mVarianceClipping = 0
        U.param[1].mInterpolationMode = 0
        U.param[1].mUseYCoCg = 0
        out = seq.step(oracle, uniforms=U)
        assert_parity(out)


@pytest.mark.parametrize("fmt,arith", [(FMT_RGBA32F, 0), (FMT_RGBA16F, 0), (FMT_RGBA16F, 1)])
@pytest.mark.parametrize("split_x", [77, -1, 202])
def test_split_screen_with_cas(oracle, fmt, arith, split_x):
    """splitScreen with CAS in one call: both parts resolved, then one CAS pass over the whole
    frame (taa.hpp:1008-1027, 1111-1139), against the oracle's resolve + CasFilter of the same
    frame; splitX -1 puts every column on param[1], the last column every column on param[0]. The
    same frames without the inspect image (param[0]'s columns then take the production form of the
    exact kernel) write the same bits."""
    import ctypes as C
    import taastar
    from gpu_helpers import FP16_PSNR_MIN_DB, INDEX_WORDS, psnr
    from taastar.workloads import uniforms_for
    print(f"case split+cas fmt={fmt} arith={arith} splitX={split_x}", flush=True)
    w = workload("c2", fmt=fmt, vel=VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F)
    seq = Sequence(w, width=203, height=81, cas=True, arith=arith, debug=False, shipped=True)
    for f in range(3):
        col, dep, vel4, jit = seq.scene.render(seq.frame)
        U = uniforms_for(w, seq.frame, jit)
        U.splitScreen = 1
        U.splitX = split_x
        U.param[1].mVarianceClipping = 0
        U.param[1].mInterpolationMode = 0
        U.param[1].mColorClampingOrClipping = 1
        k = (C.c_int32 * 2)()
        taastar.check(taastar.lib().taa_resolve_kernel(C.byref(U), C.byref(seq.bufs.images()), C.byref(seq.cas), arith, k))
        assert tuple(k) == ((1, 1) if arith else (0, 0)), tuple(k)
        out = seq.step(oracle, uniforms=U)
        if arith == 0:
            assert_parity(out)
        else:
            gi, oi = out["gpu_inspect"], out["oracle_inspect"]
            for wd in INDEX_WORDS:
                g, o = gi[..., wd], oi[..., wd]
                if wd == 3:  # flags: "rectified" is a threshold on arithmetic
                    g, o = g & ~2, o & ~2
                assert np.count_nonzero(g != o) == 0, f"frame {f}: inspect word {wd}"
            assert psnr(out["gpu_screen"][..., :3], out["oracle_screen"][..., :3]) > FP16_PSNR_MIN_DB
            seq.hist_o = out["gpu_history"]


def test_fused_cas_equals_separate_pass(oracle):
    """The fused CAS writes exactly what resolve + sharpen_cas.comp would."""
    from taastar.device import FrameBuffers, cas_apply
    from taastar.workloads import cas_for, encode_color, encode_velocity, scene_for, uniforms_for
    for fmt, vfmt in [(FMT_RGBA32F, VEL_RG32F), (FMT_RGBA16F, VEL_RG16F)]:
        w = workload("c2", fmt=fmt, vel=vfmt)
        W, H = 333, 201  # ragged: partial tiles on both axes
        sc = scene_for(w, width=W, height=H)
        a = FrameBuffers(W, H, fmt, vfmt)
        b = FrameBuffers(W, H, fmt, vfmt)
        tmp = torch.zeros_like(b.screen)
        K = cas_for(w, W, H)
        for f in range(4):
            col, dep, vel4, jit = sc.render(f)
            for bb in (a, b):
                bb.upload(encode_color(col, fmt), dep, encode_velocity(vel4, vfmt))
            U = uniforms_for(w, f, jit)
            a.resolve(U, K)
            b.resolve(U, None, screen=tmp)
            cas_apply(tmp, b.screen, fmt, K)
            torch.cuda.synchronize()
            assert torch.equal(a.screen, b.screen), f"fmt {fmt} frame {f}"
            assert torch.equal(a.hist_in, b.hist_in)


def test_deterministic_rerun(oracle):
    seqs = [Sequence(workload("c2"), width=257, height=130, cas=True) for _ in range(2)]
    for f in range(3):
        o = [s.step(oracle, run_oracle=False) for s in seqs]
        assert np.array_equal(o[0]["gpu_screen"], o[1]["gpu_screen"])
        assert np.array_equal(o[0]["gpu_history"], o[1]["gpu_history"])


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_full_size_against_oracle(oracle, name):
    """The real 1080p / 4K frames: 3 frames through the oracle, index/box bit-exact; the
    production form (no inspect image) writes the same bits."""
    print(f"case full-size {name}", flush=True)
    seq = Sequence(workload(name), cas=True, debug=False, shipped=True)
    for f in range(3):
        out = seq.step(oracle)
        assert_parity(out)
