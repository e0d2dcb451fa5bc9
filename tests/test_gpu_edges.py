"""Edge cases of the resolve on the GPU: frames smaller than a tile or one texel wide, sizes on
and next to tile boundaries, padded rows (pitch > width), hostile velocities and depths, and the
largest practical frame.

Bar: the FP32 path bit-exact against the FP32 transcription (indices, box, history, screen);
the FP16 path with indices and rejection decisions bit-exact against it (the "rectified" flag is
an arithmetic threshold and is masked) and its sharpened screen close to it.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import INDEX_WORDS, Sequence, assert_parity, psnr, rel_err  # noqa: E402
from taastar import ARITH_FP16, ARITH_FP32, FLAG_RECTIFIED, FMT_RGBA16F, FMT_RGBA32F, INSPECT  # noqa: E402
from taastar import VEL_RG16F, VEL_RG32F  # noqa: E402
from taastar.workloads import (cas_for, decode, encode_color, encode_velocity, scene_for, uniforms_for,  # noqa: E402
                               velocity_as_read, workload)


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def assert_fp16_indices(out):
    gi, oi = out["gpu_inspect"], out["oracle_inspect"]
    for wd in INDEX_WORDS:
        g, o = gi[..., wd], oi[..., wd]
        if wd == INSPECT["FLAGS"]:
            g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
        bad = np.count_nonzero(g != o)
        assert bad == 0, f"inspect word {wd}: {bad} texels differ"


# sizes below one tile, one texel wide or tall, and on / next to the 30- and 16-texel tile edges
SIZES = [(1, 1), (2, 3), (1, 37), (37, 1), (5, 5), (29, 31), (30, 30), (31, 29), (61, 15), (16, 17), (33, 64)]


@pytest.mark.parametrize("cas", [True, False])
@pytest.mark.parametrize("W,H", SIZES)
def test_small_and_boundary_sizes_fp32(oracle, W, H, cas):
    print(f"case fp32 {W}x{H} cas={cas}", flush=True)
    seq = Sequence(workload("c2"), width=W, height=H, cas=cas, debug=False, shipped=True)
    for _ in range(3):
        assert_parity(seq.step(oracle))


@pytest.mark.parametrize("cas", [True, False])
@pytest.mark.parametrize("W,H", SIZES)
def test_small_and_boundary_sizes_fp16(oracle, W, H, cas):
    print(f"case fp16 {W}x{H} cas={cas}", flush=True)
    seq = Sequence(workload("c2", fmt=FMT_RGBA16F, vel=VEL_RG16F), width=W, height=H, cas=cas, arith=ARITH_FP16,
                   debug=False, shipped=True)
    for _ in range(3):
        out = seq.step(oracle)
        assert_fp16_indices(out)
        if cas:  # display-referred: the sharpened screen is in [0, 1]
            assert psnr(out["gpu_screen"][..., :3], out["oracle_screen"][..., :3]) > 55.0
        else:
            e = rel_err(out["gpu_history"][..., :3], out["oracle_history"][..., :3], floor=1e-2)
            assert np.quantile(e, 0.999) < 1e-2


@pytest.mark.parametrize("fmt,vfmt,arith", [(FMT_RGBA32F, VEL_RG32F, ARITH_FP32), (FMT_RGBA16F, VEL_RG16F, ARITH_FP16),
                                            (FMT_RGBA16F, VEL_RG16F, ARITH_FP32)])
def test_padded_rows(fmt, vfmt, arith):
    """pitch > width: results equal the packed layout's, and no texel of the row padding is written."""
    import ctypes as C
    from taastar import FrameImages, check, lib
    from taastar.device import FrameBuffers, color_dtype, vel_shape_dtype
    print(f"case pitch fmt={fmt} arith={arith}", flush=True)
    w = workload("c2", fmt=fmt, vel=vfmt)
    W, H, P = 97, 53, 97 + 13
    sc = scene_for(w, width=W, height=H)
    K = cas_for(w, W, H)
    ref = FrameBuffers(W, H, fmt, vfmt)
    cd = color_dtype(fmt)
    vc, vd = vel_shape_dtype(vfmt)
    nan = float("nan")

    def padded(shape, dtype, fill):
        return torch.full((H, P) + shape, fill, dtype=dtype, device="cuda")

    col, dep, vel = padded((4,), cd, nan), padded((), torch.float32, nan), padded((vc,), vd, nan)
    pdep = padded((), torch.float32, nan)
    hist = [padded((4,), cd, 7.0), padded((4,), cd, 7.0)]
    scr = padded((4,), cd, 7.0)
    hist[1][:, :W] = 0.0
    pdep[:, :W] = 0.0
    for f in range(4):
        c, d, v4, jit = sc.render(f)
        c_s, v_s = encode_color(c, fmt), encode_velocity(v4, vfmt)
        ref.upload(c_s, d, v_s)
        pdep[:, :W] = dep[:, :W] if f else 0.0
        col[:, :W] = torch.from_numpy(c_s).cuda()
        dep[:, :W] = torch.from_numpy(d).cuda()
        vel[:, :W] = torch.from_numpy(np.ascontiguousarray(v_s)).cuda()
        U = uniforms_for(w, f, jit)
        ref.resolve(U, K, arith)
        img = FrameImages()
        img.color, img.depth, img.velocity = col.data_ptr(), dep.data_ptr(), vel.data_ptr()
        img.history, img.result_history = hist[(f + 1) & 1].data_ptr(), hist[f & 1].data_ptr()
        img.history_depth, img.result_screen = pdep.data_ptr(), scr.data_ptr()
        img.width, img.height, img.pitch = W, H, P
        img.color_format, img.velocity_format = fmt, vfmt
        check(lib().taa_resolve_frame(C.byref(U), C.byref(img), C.byref(K), arith, None))
        torch.cuda.synchronize()
        assert torch.equal(scr[:, :W], ref.screen), f"frame {f}: screen"
        assert torch.equal(hist[f & 1][:, :W], ref.hist_in), f"frame {f}: history"
        assert bool((scr[:, W:] == 7.0).all()) and bool((hist[f & 1][:, W:] == 7.0).all()), "row padding written"


def hostile_velocity(v4, rng):
    """Replace scattered texels of a velocity field with NaN, +-Inf, huge and near-+-1 values."""
    v = v4.copy()
    H, W = v.shape[:2]
    n = max(1, H * W // 50)
    ys, xs = rng.integers(0, H, n), rng.integers(0, W, n)
    vals = np.array([np.nan, np.inf, -np.inf, 1e9, -1e9, 0.9999, -1.0001, 3e4], np.float32)
    v[ys, xs, 0] = vals[rng.integers(0, len(vals), n)]
    v[ys, xs, 1] = vals[rng.integers(0, len(vals), n)]
    return v


@pytest.mark.parametrize("mode", [0, 2])
def test_hostile_velocity_and_depth_fp32(oracle, mode):
    """NaN / Inf / huge velocities (and NaN depths) follow the transcription bit for bit:
    int(x) saturates, NaN converts to 0, every fetch is clamped to the edge (mode 2 also through
    the exact kernel's production form, bit for bit as the general form)."""
    print(f"case hostile fp32 vel-mode {mode}", flush=True)
    from taastar.workloads import params_c2
    p = params_c2()
    p.mVelocitySampleMode = mode
    seq = Sequence(workload("c2"), width=123, height=77, cas=True, params=p, debug=False, shipped=True)
    rng = np.random.default_rng(5)
    render = seq.scene.render

    def hostile(frame, threads=0):
        c, d, v4, jit = render(frame, threads)
        d = d.copy()
        d[rng.integers(0, d.shape[0], 20), rng.integers(0, d.shape[1], 20)] = np.nan
        return c, d, hostile_velocity(v4, rng), jit

    seq.scene.render = hostile
    for _ in range(4):
        assert_parity(seq.step(oracle))


def test_hostile_velocity_fp16(oracle):
    """The FP16 path on NaN / Inf / huge velocities and NaN depths: texel indices, closest taps and
    rejection decisions as the transcription's."""
    print("case hostile fp16", flush=True)
    seq = Sequence(workload("c2", fmt=FMT_RGBA16F, vel=VEL_RG16F), width=123, height=77, cas=True, arith=ARITH_FP16,
                   debug=False, shipped=True)
    rng = np.random.default_rng(6)
    render = seq.scene.render

    def hostile(frame, threads=0):
        c, d, v4, jit = render(frame, threads)
        d = d.copy()
        d[rng.integers(0, d.shape[0], 20), rng.integers(0, d.shape[1], 20)] = np.nan
        return c, d, hostile_velocity(v4, rng), jit

    seq.scene.render = hostile
    for _ in range(4):
        assert_fp16_indices(seq.step(oracle))


def test_largest_frame_fp32(oracle):
    """7680x4320 (33 Mpx) with RGBA16F storage on the exact path: one reset frame and one blend."""
    print("case 7680x4320 fp32", flush=True)
    seq = Sequence(workload("c3"), width=7680, height=4320, cas=True, debug=False, shipped=True)
    for _ in range(2):
        assert_parity(seq.step(oracle))
