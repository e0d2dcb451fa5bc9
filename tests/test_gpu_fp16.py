"""The FP16-arithmetic path (TAA_ARITH_FP16) on FP16 storage against the FP64 oracle.

Bar: reprojected texel indices, closest-depth taps, depth-test texels and rejection flags
bit-exact against the FP32 transcription (they are computed in fp32 on both sides); the
sharpened screen (display-referred, CAS output in [0,1], peak 1.0) >= 55 dB PSNR against the FP64
oracle on every frame, over the pixels outside the reference's own divergence footprint.

Divergence footprint: taa.comp's Lottes weight w = 1 - |lc - lh| / max(lc, lh, 0.2) is not clamped
(taa.comp:865-869); once Catmull-Rom undershoot makes a luma negative, w*w > 1 makes alpha negative
and the blend extrapolates. On the sigma = 2 HDR scene a few pixels then grow without bound in the
FP64 oracle itself (1e9 and more), and at others the reference's result is chaotic: rounding the
stored history to FP16 alone, with every operation of the FP32 transcription unchanged, moves the
sharpened result by up to 1.0 and takes that transcription to 54 dB against FP64 by frame 15.
Excluded from the 55 dB bar are therefore (a) the pixels whose FP64 history exceeded 4x the frame's
input maximum in the last 8 frames, dilated by the Catmull-Rom support, and (b) the pixels where
the FP32 transcription on FP16 storage is itself more than 0.01 from the FP64 result, dilated by
the CAS neighbourhood. Both sets come from the oracle alone, their size is reported and bounded
(at most 0.13% of the pixels at 960x540 and 0.19% at 4K on these sequences), and the full-frame PSNR must stay within 3 dB of
that of the FP32 transcription on FP16 storage.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import INDEX_WORDS, psnr, to_f16  # noqa: E402
from taastar import FLAG_RECTIFIED  # noqa: E402
from taastar import ARITH_FP16, FMT_RGBA16F, FMT_RGBA32F, INSPECT  # noqa: E402
from taastar.workloads import (cas_for, decode, encode_color, encode_velocity, scene_for, uniforms_for,  # noqa: E402
                               velocity_as_read, workload)
from scipy.ndimage import binary_dilation  # noqa: E402

PSNR_MIN_DB = 55.0
DEPART_ABS = 1e-2          # FP32 transcription on FP16 storage vs FP64, sharpened screen
MASK_MAX_FRACTION = 3e-3   # excluded pixels per frame


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def run_fp16_sequence(oracle, name, W, H, frames, report):
    from taastar.device import FrameBuffers
    w = workload(name)
    sc = scene_for(w, width=W, height=H)
    k = cas_for(w, W, H)
    bufs = FrameBuffers(W, H, w["fmt"], w["vel"], inspect=True)
    h64 = np.zeros((H, W, 4))
    h32 = np.zeros((H, W, 4), np.float32)
    prev = np.zeros((H, W), np.float32)
    recent = []
    worst = 1e9
    for f in range(frames):
        col, dep, vel4, jit = sc.render(f)
        col_s, vel_s = encode_color(col, w["fmt"]), encode_velocity(vel4, w["vel"])
        bufs.upload(col_s, dep, vel_s)
        U = uniforms_for(w, f, jit)
        bufs.resolve(U, k, ARITH_FP16)
        torch.cuda.synchronize()
        g_scr = decode(bufs.screen.cpu().numpy())
        g_his = decode(bufs.hist_in.cpu().numpy())
        g_ins = bufs.inspect.cpu().numpy()
        args = (decode(col_s), dep, velocity_as_read(vel_s))
        o64 = oracle.resolve(U, *args, h64, prev, dtype=np.float64, debug=False, inspect=False)
        s64 = oracle.cas(o64["screen"], k.const1[0], dtype=np.float64)
        # selections are made in fp32 arithmetic: check them against the FP32 transcription
        o32 = oracle.resolve(U, *args, h32, prev, dtype=np.float32, debug=False)
        for wd in INDEX_WORDS:
            g, o = g_ins[..., wd], o32["inspect"][..., wd]
            if wd == INSPECT["FLAGS"]:  # "rectified" is a threshold on arithmetic, not a selection
                g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
            bad = np.count_nonzero(g != o)
            assert bad == 0, f"frame {f}: inspect word {wd} differs at {bad} texels"
        # divergence footprint of the FP64 oracle (see module docstring)
        cmax = float(decode(col_s)[..., :3].max())
        recent.append(np.abs(o64["history"][..., :3]).max(-1) > 4.0 * cmax)
        del recent[:-8]
        s32 = oracle.cas(to_f16(o32["screen"]).astype(np.float32), k.const1[0])
        departs = np.abs(s32[..., :3] - s64[..., :3]).max(-1) > DEPART_ABS
        unstable = binary_dilation(np.logical_or.reduce(recent), iterations=3) | binary_dilation(departs, iterations=1)
        stable = ~unstable
        p_all = psnr(g_scr[..., :3], s64[..., :3], 1.0)
        p_base = psnr(s32[..., :3], s64[..., :3], 1.0)
        p_scr = psnr(g_scr[stable][:, :3], s64[stable][:, :3], 1.0)
        err = float(np.abs(g_scr[stable][:, :3] - s64[stable][:, :3]).max())
        report.append(dict(frame=f, screen_psnr_db=round(p_scr, 2), screen_max_abs=err, full_frame_psnr_db=round(p_all, 2),
                           fp32_fp16storage_psnr_db=round(p_base, 2), masked_pixels=int(unstable.sum())))
        worst = min(worst, p_scr)
        assert p_scr >= PSNR_MIN_DB, f"frame {f}: screen PSNR {p_scr:.2f} dB < {PSNR_MIN_DB}: {report[-1]}"
        assert p_all >= min(p_base - 3.0, 60.0), f"frame {f}: full-frame {p_all:.2f} dB vs FP32 path {p_base:.2f} dB"
        assert unstable.sum() <= MASK_MAX_FRACTION * W * H, report[-1]
        h64 = o64["history"]
        h32 = to_f16(o32["history"]).astype(np.float32)
        prev = dep
    print(f"{name} {W}x{H} x{frames}: worst screen PSNR {worst:.2f} dB", report, flush=True)
    return worst


def test_fp16_path_960x540_16_frames(oracle):
    print("case fp16 c3-scene 960x540 x16", flush=True)
    rep = []
    run_fp16_sequence(oracle, "c3", 960, 540, 16, rep)


def test_fp16_path_full_4k(oracle):
    print("case fp16 c3 3840x2160 x6", flush=True)
    rep = []
    run_fp16_sequence(oracle, "c3", 3840, 2160, 6, rep)


def test_fp16_path_fp32_storage_indices(oracle):
    """ARITH_FP16 on FP32 storage (c2 workload): same exactness of the selection logic."""
    rep = []
    from taastar.device import FrameBuffers
    w = workload("c2")
    W, H = 320, 180
    sc = scene_for(w, width=W, height=H)
    k = cas_for(w, W, H)
    bufs = FrameBuffers(W, H, w["fmt"], w["vel"], inspect=True)
    h32 = np.zeros((H, W, 4), np.float32)
    prev = np.zeros((H, W), np.float32)
    for f in range(6):
        col, dep, vel4, jit = sc.render(f)
        col_s, vel_s = encode_color(col, w["fmt"]), encode_velocity(vel4, w["vel"])
        bufs.upload(col_s, dep, vel_s)
        U = uniforms_for(w, f, jit)
        bufs.resolve(U, k, ARITH_FP16)
        torch.cuda.synchronize()
        g_ins = bufs.inspect.cpu().numpy()
        o = oracle.resolve(U, decode(col_s), dep, velocity_as_read(vel_s), h32, prev, dtype=np.float32, debug=False)
        for wd in INDEX_WORDS:
            g, oo = g_ins[..., wd], o["inspect"][..., wd]
            if wd == INSPECT["FLAGS"]:
                g, oo = g & ~FLAG_RECTIFIED, oo & ~FLAG_RECTIFIED
            assert np.count_nonzero(g != oo) == 0, (f, wd)
        g_scr = decode(bufs.screen.cpu().numpy())
        s = oracle.cas(o["screen"], k.const1[0])
        assert psnr(g_scr[..., :3], s[..., :3]) > 60.0
        h32, prev = o["history"], dep


@pytest.mark.parametrize("fmt", [FMT_RGBA16F, FMT_RGBA32F])
def test_reference_default_block(oracle, fmt):
    """The reference's default parameter block (taa.hpp:31-76: plain min/max box, clamp, bilinear
    history) on the FP16 kernel's general form: texel indices, taps and decisions as the FP32
    transcription's, the sharpened screen above 55 dB against it."""
    from gpu_helpers import Sequence
    from taastar import VEL_RG16F, VEL_RG32F, default_params
    print(f"case fp16 reference defaults fmt={fmt}", flush=True)
    p = default_params()
    p.mUseVelocityVectors = 2  # RG velocity: every texel moves (w reads as 1)
    w = workload("c2", fmt=fmt, vel=VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F, params=lambda: p)
    seq = Sequence(w, width=480, height=272, cas=True, arith=ARITH_FP16, debug=False)
    for f in range(4):
        out = seq.step(oracle)
        gi, oi = out["gpu_inspect"], out["oracle_inspect"]
        for wd in INDEX_WORDS:
            g, o = gi[..., wd], oi[..., wd]
            if wd == INSPECT["FLAGS"]:
                g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
            assert np.count_nonzero(g != o) == 0, f"frame {f}: inspect word {wd}"
        assert psnr(out["gpu_screen"][..., :3], out["oracle_screen"][..., :3]) > 55.0, f"frame {f}"
        seq.hist_o = out["gpu_history"]
