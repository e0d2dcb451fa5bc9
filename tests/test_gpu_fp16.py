"""The FP16-arithmetic path (TAA_ARITH_FP16) on FP16 storage against the FP64 oracle.

Bar (gpu_helpers.FP16OracleChain): reprojected texel indices, closest-depth taps, depth-test
texels and rejection flags bit-exact against the FP32 transcription (they are computed in fp32 on
both sides); the sharpened screen (display-referred, CAS output in [0,1], peak 1.0) >= 55 dB PSNR
against the FP64 oracle on every frame, over the pixels outside the reference's own divergence
footprint (taa.comp:865-869's unclamped Lottes weight; at most 1% of a frame, its size and the
full-frame PSNR printed per frame); the full-frame PSNR within 3 dB of that of the FP32
transcription on FP16 storage. The instantiation without the inspect image (the one the throughput
path launches) writes the same bits as the inspected one.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import INDEX_WORDS, FP16OracleChain, assert_same_bits, psnr  # noqa: E402
from taastar import FLAG_RECTIFIED  # noqa: E402
from taastar import ARITH_FP16, FMT_RGBA16F, FMT_RGBA32F, INSPECT  # noqa: E402
from taastar.workloads import (cas_for, decode, encode_color, encode_velocity, scene_for, uniforms_for,  # noqa: E402
                               velocity_as_read, workload)


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def run_fp16_sequence(oracle, name, W, H, frames, report, shipped=True):
    """Frames of the workload's scene through the FP16 kernel with the inspect image bound (the
    selections are read from it) and, with `shipped`, through the instantiation without it (the
    one the throughput path launches) on the same inputs: both must write the same bits. Every
    frame is held to FP16OracleChain's bar."""
    from taastar.device import FrameBuffers
    w = workload(name)
    sc = scene_for(w, width=W, height=H)
    k = cas_for(w, W, H)
    bufs = FrameBuffers(W, H, w["fmt"], w["vel"], inspect=True)
    ship = FrameBuffers(W, H, w["fmt"], w["vel"]) if shipped else None
    chain = FP16OracleChain(oracle, W, H, k, name)
    for f in range(frames):
        col, dep, vel4, jit = sc.render(f)
        col_s, vel_s = encode_color(col, w["fmt"]), encode_velocity(vel4, w["vel"])
        bufs.upload(col_s, dep, vel_s)
        U = uniforms_for(w, f, jit)
        bufs.resolve(U, k, ARITH_FP16)
        if ship is not None:
            ship.upload(col_s, dep, vel_s)
            ship.resolve(U, k, ARITH_FP16)
        torch.cuda.synchronize()
        if ship is not None:
            assert_same_bits(bufs.screen, ship.screen, f"frame {f}: screen without the inspect image")
            assert_same_bits(bufs.hist_in, ship.hist_in, f"frame {f}: history without the inspect image")
        chain.check(U, col_s, dep, vel_s, decode(bufs.screen.cpu().numpy()), bufs.inspect.cpu().numpy())
    report.extend(chain.reports)
    print(f"{name} {W}x{H} x{frames}: worst screen PSNR {chain.worst:.2f} dB", flush=True)
    return chain.worst


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
    seq = Sequence(w, width=480, height=272, cas=True, arith=ARITH_FP16, debug=False, shipped=True)
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


@pytest.mark.parametrize("mode", [0, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("fmt", [FMT_RGBA16F, FMT_RGBA32F])
def test_reference_defaults_rgba_velocity_and_debug_image(oracle, fmt, mode):
    """The reference's defaults as they are shipped (taa.hpp:31-76 with IMAGE_FORMAT_VELOCITY =
    RGBA16F, shader_cpu_common.h:77): movers-only velocity (.w), the background reprojected through
    the matrices, .z in the expected history depth, a bound debug image -- on the FP16 kernel, not
    the exact one. Indices, taps and decisions bit-exact against the FP32 transcription, the
    sharpened screen > 55 dB against it, the debug image within 2e-3 relative (modes of the
    neighbourhood, the decisions, the velocity and the speed) or 1e-2 on 99.9% of the texels (modes
    6 and 7, the outputs themselves)."""
    import ctypes as C
    from gpu_helpers import Sequence, rel_err
    from taastar import KERNEL_FP16, VEL_RGBA16F, VEL_RGBA32F, default_params
    print(f"case fp16 reference defaults rgba velocity fmt={fmt} debug mode {mode}", flush=True)
    p = default_params()
    p.mDebugMode = mode
    p.mDebugScale = 2.0
    p.mDebugCenter = mode % 2
    w = workload("c2", fmt=fmt, vel=VEL_RGBA16F if fmt == FMT_RGBA16F else VEL_RGBA32F, params=lambda: p)
    seq = Sequence(w, width=240, height=136, cas=True, arith=ARITH_FP16, debug=True, shipped=True)
    for f in range(4):
        assert seq.kernel() == (KERNEL_FP16, KERNEL_FP16)
        out = seq.step(oracle)
        gi, oi = out["gpu_inspect"], out["oracle_inspect"]
        for wd in INDEX_WORDS:
            g, o = gi[..., wd], oi[..., wd]
            if wd == INSPECT["FLAGS"]:
                g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
            assert np.count_nonzero(g != o) == 0, f"frame {f}: inspect word {wd}"
        assert psnr(out["gpu_screen"][..., :3], out["oracle_screen"][..., :3]) > 55.0, f"frame {f}"
        g, o = out["gpu_debug"], out["oracle_debug"]
        assert np.array_equal(np.isfinite(g), np.isfinite(o)), f"frame {f}: debug finiteness"
        fin = np.isfinite(o)
        e = rel_err(g[fin], o[fin], floor=1e-2)
        if mode in (6, 7):
            assert np.quantile(e, 0.999) <= 1e-2, f"frame {f}: debug mode {mode}"
        else:
            assert e.max() <= 2e-3, f"frame {f}: debug mode {mode}: {e.max()}"
        seq.hist_o = out["gpu_history"]

