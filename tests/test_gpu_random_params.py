"""Seeded random combinations of the Parameters options against the FP32 transcription of taa.comp.

test_gpu_parity.py turns one option on at a time; this file draws whole parameter blocks (every
option of taa.hpp:30-77 the resolve implements, with random alpha, gamma and rejection alpha,
either storage format, any velocity format, CAS on or off, raw or linear depth rejection), so
that the options' interactions are exercised: Karis tone mapping with shaped neighbourhoods and
slow clipping, movers-only velocity with anti-ghosting, unjitter with B-spline history, ...

Bars: FP32 arithmetic as in test_gpu_parity.py (indices, taps, flags and the neighbourhood box
bit-exact, history and screen within 1e-5 relative). FP16 arithmetic (the throughput kernel when
the block is one it implements, else the exact kernel): indices, taps and flags bit-exact against
the FP32 transcription, screen PSNR > 55 dB against it (peak 1; CAS output in [0, 1]).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import INDEX_WORDS, Sequence, assert_parity, psnr  # noqa: E402
from taastar import (ARITH_FP16, FMT_RGBA16F, FMT_RGBA32F, INSPECT, KERNEL_FP16, VEL_RG16F, VEL_RG32F,  # noqa: E402
                     VEL_RGBA16F, VEL_RGBA32F)
from taastar import FLAG_RECTIFIED  # noqa: E402
from taastar.workloads import params_c2, workload  # noqa: E402


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def random_params(rng, fast_only=False):
    """A parameter block; fast_only keeps to the options of the FP16 throughput kernel."""
    p = params_c2()
    b = lambda: int(rng.integers(2))
    p.mAlpha = float(rng.uniform(0.02, 0.3))
    p.mVarClipGamma = float(rng.uniform(0.5, 1.5))
    p.mRejectionAlpha = float(rng.uniform(0.5, 1.0))
    p.mUseYCoCg = b()
    p.mDepthCulling = b()
    p.mLumaWeightingLottes = b()
    p.mRejectOutside = b()
    p.mColorClampingOrClipping = int(rng.integers(3))
    p.mVelocitySampleMode = int(rng.choice([0, 2]))
    p.mUseVelocityVectors = int(rng.choice([1, 2]))
    if fast_only:  # everything but shaped / chroma-shrunk boxes, slow clip, B-spline, Karis, unjitter, noise
        p.mUseVelocityVectors = int(rng.integers(3))  # 0: matrix reprojection, 1: movers only
        p.mVelocitySampleMode = int(rng.integers(3))  # 1: the 3x3-longest quirk
        p.mVarianceClipping = b()  # off: the plain min/max box
        p.mShapedNeighbourhood = 0
        p.mInterpolationMode = int(rng.choice([0, 2]))  # bilinear or Catmull-Rom history
        p.mReduceBlendNearClamp = b()
        p.mVelBasedAlpha = b()
        p.mDynamicAntiGhosting = b()
        return p
    p.mVarianceClipping = b()
    p.mShapedNeighbourhood = b()
    p.mShrinkChromaAxis = b()
    p.mColorClampingOrClipping = int(rng.integers(4))
    p.mInterpolationMode = int(rng.integers(3))
    p.mUseVelocityVectors = int(rng.integers(3))
    p.mVelocitySampleMode = int(rng.integers(3))
    p.mReduceBlendNearClamp = b()
    p.mVelBasedAlpha = b()
    p.mToneMapLumaKaris = b()
    p.mDynamicAntiGhosting = b()
    p.mUnjitterNeighbourhood = b()
    p.mUnjitterCurrentSample = b()
    return p


def describe(p):
    keys = ("mVarianceClipping", "mShapedNeighbourhood", "mUseYCoCg", "mShrinkChromaAxis", "mColorClampingOrClipping",
            "mInterpolationMode", "mUseVelocityVectors", "mVelocitySampleMode", "mDepthCulling", "mLumaWeightingLottes",
            "mRejectOutside", "mReduceBlendNearClamp", "mVelBasedAlpha", "mToneMapLumaKaris", "mDynamicAntiGhosting",
            "mUnjitterNeighbourhood", "mUnjitterCurrentSample")
    return " ".join(f"{k[1:]}={getattr(p, k)}" for k in keys)


@pytest.mark.parametrize("seed", range(48))
def test_random_parameter_block_fp32(oracle, seed):
    rng = np.random.default_rng(0x7AA5F000 + seed)
    p = random_params(rng)
    fmt = int(rng.choice([FMT_RGBA32F, FMT_RGBA16F]))
    vfmt = int(rng.choice([VEL_RG32F, VEL_RG16F, VEL_RGBA32F, VEL_RGBA16F]))
    cas = bool(rng.integers(2))
    lin = float(rng.choice([0.0, 0.01]))
    print(f"case random fp32 seed={seed} fmt={fmt} vfmt={vfmt} cas={cas} linear={lin} {describe(p)}", flush=True)
    w = workload("c2", fmt=fmt, vel=vfmt, linear_depth=lin, params=lambda p=p: p)
    seq = Sequence(w, width=96, height=64, cas=cas, seed=0x5EED0000 + seed)
    for f in range(3):
        assert_parity(seq.step(oracle))


@pytest.mark.parametrize("seed", range(16))
def test_random_parameter_block_fp16(oracle, seed):
    rng = np.random.default_rng(0x7AA5F100 + seed)
    p = random_params(rng, fast_only=True)
    fmt = int(rng.choice([FMT_RGBA32F, FMT_RGBA16F]))
    rgba = bool(rng.integers(2))  # RG or RGBA velocity of the colour storage's precision
    vfmt = (VEL_RGBA16F if rgba else VEL_RG16F) if fmt == FMT_RGBA16F else (VEL_RGBA32F if rgba else VEL_RG32F)
    cas = bool(rng.integers(2))
    lin = float(rng.choice([0.0, 0.01]))
    print(f"case random fp16 seed={seed} fmt={fmt} vfmt={vfmt} cas={cas} linear={lin} {describe(p)}", flush=True)
    w = workload("c2", fmt=fmt, vel=vfmt, linear_depth=lin, params=lambda p=p: p)
    seq = Sequence(w, width=128, height=80, cas=cas, arith=ARITH_FP16, seed=0x5EED1000 + seed, debug=False,
                   shipped=True)
    for f in range(3):
        assert seq.kernel() == (KERNEL_FP16, KERNEL_FP16), "every drawn block runs on the FP16 kernel"
        out = seq.step(oracle)
        gi, oi = out["gpu_inspect"], out["oracle_inspect"]
        for wd in INDEX_WORDS:
            g, o = gi[..., wd], oi[..., wd]
            if wd == INSPECT["FLAGS"]:  # "rectified" is a threshold on arithmetic, not a selection
                g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
            bad = np.count_nonzero(g != o)
            assert bad == 0, f"frame {f}: inspect word {wd}: {bad} texels differ"
        scr_g, scr_o = out["gpu_screen"][..., :3], out["oracle_screen"][..., :3]
        if cas:
            assert psnr(scr_g, scr_o) > 55.0, f"frame {f}: screen PSNR {psnr(scr_g, scr_o):.2f} dB"
        else:  # unsharpened HDR screen: PSNR of the display-referred (clamped) value
            assert psnr(np.clip(scr_g, 0, 1), np.clip(scr_o, 0, 1)) > 55.0, f"frame {f}"
        # the FP16 path follows the same history: keep the oracle on the GPU's stored history
        seq.hist_o = out["gpu_history"]


UPSCALES = [((96, 54), (192, 108)), ((128, 72), (192, 108)), ((75, 41), (101, 57)), ((120, 68), (120, 68))]


@pytest.mark.parametrize("seed", range(16))
def test_random_parameter_block_upsampling(oracle, seed):
    """Random blocks on the upsampling path: a random scale, mUpsampling on or off (off: the
    nearest input texel), either storage format, CAS on or off."""
    from test_gpu_taau import Upsampler
    rng = np.random.default_rng(0x7AA5F200 + seed)
    p = random_params(rng)
    lo, hi = UPSCALES[int(rng.integers(len(UPSCALES)))]
    fmt = int(rng.choice([FMT_RGBA32F, FMT_RGBA16F]))
    vfmt = int(rng.choice([VEL_RG32F, VEL_RGBA32F]))
    cas = bool(rng.integers(2))
    upsampling = bool(rng.integers(2)) or lo == hi
    lin = float(rng.choice([0.0, 0.01]))
    print(f"case random up seed={seed} {lo}->{hi} mUpsampling={int(upsampling)} fmt={fmt} vfmt={vfmt} cas={cas} "
          f"linear={lin} {describe(p)}", flush=True)
    w = workload("c2", fmt=fmt, vel=vfmt, linear_depth=lin)
    up = Upsampler(w, lo, hi, cas=cas, upsampling=upsampling, params=p)
    for f in range(3):
        assert_parity(up.step(oracle))


@pytest.mark.parametrize("seed", range(8))
def test_random_split_screen(oracle, seed):
    """Split screen with two random blocks (taa.comp:717-718) at a random split column; odd seeds
    with CAS (both parts resolved, then one CAS pass over the frame)."""
    from taastar.workloads import uniforms_for
    rng = np.random.default_rng(0x7AA5F300 + seed)
    p0, p1 = random_params(rng), random_params(rng)
    split = int(rng.integers(1, 159))
    print(f"case random split seed={seed} splitX={split}\n  left  {describe(p0)}\n  right {describe(p1)}", flush=True)
    w = workload("c2", vel=VEL_RGBA32F)
    seq = Sequence(w, width=160, height=64, cas=seed % 2 == 1, seed=0x5EED2000 + seed)
    for f in range(3):
        col, dep, vel4, jit = seq.scene.render(seq.frame)
        U = uniforms_for(w, seq.frame, jit, params=p0)
        U.param[1] = p1
        U.splitScreen = 1
        U.splitX = split
        assert_parity(seq.step(oracle, uniforms=U))


@pytest.mark.parametrize("seed", range(12))
def test_random_segmentation(oracle, seed):
    """The resolve with the segmentation mask: a random subset of the indicator flags
    (shader_cpu_common.h:31-40), random weights, threshold and history count, on a random block."""
    from taastar import RTFLAG
    from test_gpu_segmentation import SegSequence, assert_seg_parity
    rng = np.random.default_rng(0x7AA5F400 + seed)
    p = random_params(rng)
    names = ["NRM", "DPT", "MID", "LUM", "OUT", "DIS", "CNT", "FXD", "ALL"]
    keep = [n for n in names if rng.uniform() < (0.1 if n in ("ALL", "FXD") else 0.5)]
    p.mRayTraceAugment = 1
    p.mRayTraceAugmentFlags = int(sum(RTFLAG[n] for n in keep))
    p.mRayTraceAugment_WNrm, p.mRayTraceAugment_WDpt, p.mRayTraceAugment_WMId, p.mRayTraceAugment_WLum = (
        float(v) for v in rng.uniform(0.0, 1.0, 4))
    p.mRayTraceAugment_Thresh = float(rng.uniform(0.1, 1.0))
    p.mRayTraceHistoryCount = int(rng.integers(0, 10))
    p.mDebugMode = 8
    fmt = int(rng.choice([FMT_RGBA32F, FMT_RGBA16F]))
    cas = bool(rng.integers(2))
    print(f"case random seg seed={seed} flags={'|'.join(keep) or '0'} fmt={fmt} cas={cas} {describe(p)}", flush=True)
    w = workload("c2", fmt=fmt, vel=VEL_RGBA32F)
    seq = SegSequence(w, 160, 96, cas=cas, params=p)
    for f in range(3):
        assert_seg_parity(seq.step(oracle))


@pytest.mark.parametrize("seed", range(12))
def test_random_velocity_producer(oracle, seed):
    """The velocity producer on random cameras: a random frame size, camera move and turn, up to
    five moved objects (and an id beyond the table), any of the four formats; bit-exact."""
    from taastar.workloads import camera_case
    from test_gpu_velocity import bits, produce, stored
    rng = np.random.default_rng(0x7AA5F500 + seed)
    W, H = int(rng.integers(1, 300)), int(rng.integers(1, 200))
    vfmt = int(rng.choice([VEL_RG32F, VEL_RG16F, VEL_RGBA32F, VEL_RGBA16F]))
    n_obj = int(rng.integers(0, 6))
    move = tuple(float(v) for v in rng.uniform(-3e-3, 3e-3, 3))
    turn = float(rng.uniform(-2.0, 2.0))
    pad = int(rng.integers(0, 9))
    print(f"case random velocity seed={seed} {W}x{H} pad={pad} vfmt={vfmt} objects={n_obj} move={move} turn={turn:.3f}",
          flush=True)
    case = camera_case(W, H, seed=0x5EED3000 + seed, n_objects=n_obj, move=move, turn_deg=turn)
    got = produce(case, vfmt, pad=pad).cpu().numpy()
    ref = stored(oracle.velocity(case["reproject"], case["depth"], case["ids"], case["objects"]), vfmt)
    bad = np.count_nonzero(bits(got[:, :W]) != bits(ref))
    assert bad == 0, f"{bad} words differ"
    assert np.all(got[:, W:] == 7.0), "row padding written"


@pytest.mark.parametrize("seed", range(12))
def test_random_edge_aa(oracle, seed):
    """The edge-AA fallback with random constants (subpix, thresholds), a random mask of all
    categories, random frame size and row padding, either storage; HDR colour above 1 included."""
    from taastar.workloads import encode_color
    from test_gpu_edge_aa import bits, consts, run, stored
    rng = np.random.default_rng(0x7AA5F600 + seed)
    W, H = int(rng.integers(1, 260)), int(rng.integers(1, 140))
    fmt = int(rng.choice([FMT_RGBA32F, FMT_RGBA16F]))
    pad = int(rng.integers(0, 7))
    sub, thr, thr_min = float(rng.uniform(0, 1)), float(rng.uniform(0, 0.4)), float(rng.uniform(0, 0.1))
    hdr = float(rng.choice([1.0, 4.0]))
    print(f"case random edge-aa seed={seed} {W}x{H} pad={pad} fmt={fmt} subpix={sub:.3f} thr={thr:.3f} "
          f"thr_min={thr_min:.3f} peak={hdr}", flush=True)
    col = np.concatenate([rng.uniform(0, hdr, (H, W, 3)), np.ones((H, W, 1))], -1).astype(np.float32)
    # smooth regions with edges between them, so that some texels pass the contrast thresholds and some do not
    col[..., :3] = np.where(rng.uniform(size=(H, W, 1)) < 0.7, np.round(col[..., :3] * 2) / 2, col[..., :3])
    src_s = encode_color(col, fmt)
    mask = rng.integers(0, 3, (H, W)).astype(np.uint32) | (rng.integers(0, 5, (H, W)).astype(np.uint32) << 16)
    k = consts(W, H, subpix=sub, edge_threshold=thr, edge_threshold_min=thr_min)
    got = run(src_s, mask, fmt, k, pad=pad)
    ref = stored(oracle.edge_aa(src_s.astype(np.float32), mask, k.subpix, k.edge_threshold, k.edge_threshold_min), fmt)
    bad = np.count_nonzero(bits(got[:, :W]) != bits(ref))
    assert bad == 0, f"{bad} words differ"
    assert np.all(got[:, W:] == 7.0), "row padding written"


@pytest.mark.parametrize("seed", range(8))
def test_random_batch_equals_alone(seed):
    """A random batch -- 2 to 6 sessions of random sizes (some scaled: input below the output),
    one random parameter block per session, CAS on or off for all -- submitted and through a
    graph-captured plan over three frames: every screen bit-identical to its session resolved
    alone. FP32 arithmetic with any block, or FP16 with blocks of the throughput kernel."""
    import ctypes as C
    from taastar import ARITH_FP32, BatchItem, SessionIO, Uniforms, check, lib
    from taastar.workloads import cas_for, encode_color, encode_velocity, scene_for, uniforms_for
    from test_gpu_serve import h2d
    rng = np.random.default_rng(0x7AA5F700 + seed)
    fmt = int(rng.choice([FMT_RGBA32F, FMT_RGBA16F]))
    vfmt = VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F
    arith = int(rng.choice([ARITH_FP32, ARITH_FP16]))
    cas = bool(rng.integers(2))
    n = int(rng.integers(2, 7))
    specs = []
    for k in range(n):
        W, H = int(rng.integers(8, 200)), int(rng.integers(8, 120))
        lo = (W, H) if rng.uniform() < 0.7 else (max(4, W * 2 // 3), max(4, H * 2 // 3))
        specs.append((lo, (W, H), random_params(rng, fast_only=arith == ARITH_FP16)))
    print(f"case random batch seed={seed} fmt={fmt} arith={arith} cas={cas} sizes={[(s[0], s[1]) for s in specs]}",
          flush=True)
    w = workload("c2", fmt=fmt, vel=vfmt)
    scenes = [scene_for(w, width=lo[0], height=lo[1], seed=0x5EED4000 + 97 * seed + k) for k, (lo, _, _) in enumerate(specs)]
    elt = 8 if fmt == FMT_RGBA16F else 16

    def run(mode):
        ctx = C.c_void_p()
        check(lib().taa_ctx_create(0, 0, C.byref(ctx)))
        sess, scrs, Ks = [], [], []
        for lo, hi, _ in specs:
            s = C.c_void_p()
            check(lib().taa_session_open_scaled(ctx, hi[0], hi[1], lo[0], lo[1], fmt, vfmt, C.byref(s)))
            sess.append(s)
            scrs.append(torch.zeros(hi[1] * hi[0] * elt, dtype=torch.uint8, device="cuda"))
            Ks.append(cas_for(w, hi[0], hi[1]) if cas else None)
        Us = [Uniforms() for _ in specs]
        items = (BatchItem * n)()
        for k in range(n):
            items[k].session, items[k].uniforms = sess[k], C.pointer(Us[k])
            items[k].cas = C.pointer(Ks[k]) if cas else None
            items[k].result_screen = scrs[k].data_ptr()
        plan = C.c_void_p()
        if mode == "plan":
            check(lib().taa_batch_plan_create(ctx, n, items, arith, C.byref(plan)))
        outs = []
        for f in range(3):
            for k, (lo, hi, p) in enumerate(specs):
                io = SessionIO()
                check(lib().taa_session_io(sess[k], C.byref(io)))
                col, dep, vel4, jit = scenes[k].render(f)
                h2d(io.color, encode_color(col, fmt))
                h2d(io.depth, dep)
                h2d(io.velocity, encode_velocity(vel4, vfmt))
                U = uniforms_for(w, f, jit, params=p)
                C.memmove(C.addressof(Us[k]), C.addressof(U), C.sizeof(Uniforms))
            if mode == "batch":
                check(lib().taa_batch_submit(ctx, n, items, arith, None))
            elif mode == "plan":
                check(lib().taa_batch_plan_launch(plan, None))
            else:
                for k in range(n):
                    check(lib().taa_session_resolve(sess[k], C.byref(Us[k]), C.byref(Ks[k]) if cas else None,
                                                    C.c_void_p(scrs[k].data_ptr()), arith, None))
            torch.cuda.synchronize()
            outs.append([t.cpu().numpy().copy() for t in scrs])
        if mode == "plan":
            check(lib().taa_batch_plan_destroy(plan))
        for s in sess:
            check(lib().taa_session_close(s))
        check(lib().taa_ctx_destroy(ctx))
        return outs

    alone = run("alone")
    for mode in ("batch", "plan"):
        got = run(mode)
        for f in range(3):
            for k in range(n):
                assert np.array_equal(alone[f][k], got[f][k]), f"{mode}: frame {f} session {k}"


@pytest.mark.filterwarnings("ignore:overflow encountered in cast:RuntimeWarning")  # huge velocities stored as fp16
@pytest.mark.parametrize("seed", range(12))
def test_random_fp16_sizes_and_hostile_inputs(oracle, seed):
    """The FP16 kernel at random frame sizes (1 to 150 texels a side) with scattered NaN / Inf /
    huge velocities and NaN depths: texel indices, closest taps and rejection decisions as the FP32
    transcription's (int(x) saturates, NaN converts to 0, every fetch clamped to the edge)."""
    from test_gpu_edges import hostile_velocity
    rng = np.random.default_rng(0x7AA5F800 + seed)
    p = random_params(rng, fast_only=True)
    fmt = int(rng.choice([FMT_RGBA32F, FMT_RGBA16F]))
    vfmt = VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F
    W, H = int(rng.integers(1, 151)), int(rng.integers(1, 151))
    cas = bool(rng.integers(2))
    print(f"case random fp16 hostile seed={seed} {W}x{H} fmt={fmt} cas={cas} {describe(p)}", flush=True)
    w = workload("c2", fmt=fmt, vel=vfmt, params=lambda p=p: p)
    seq = Sequence(w, width=W, height=H, cas=cas, arith=ARITH_FP16, seed=0x5EED5000 + seed, debug=False,
                   shipped=True)
    render = seq.scene.render

    def hostile(frame, threads=0):
        c, d, v4, jit = render(frame, threads)
        d = d.copy()
        n = max(1, W * H // 40)
        d[rng.integers(0, H, n), rng.integers(0, W, n)] = np.nan
        return c, d, hostile_velocity(v4, rng), jit

    seq.scene.render = hostile
    for f in range(3):
        out = seq.step(oracle)
        gi, oi = out["gpu_inspect"], out["oracle_inspect"]
        for wd in INDEX_WORDS:
            g, o = gi[..., wd], oi[..., wd]
            if wd == INSPECT["FLAGS"]:
                g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
            bad = np.count_nonzero(g != o)
            assert bad == 0, f"frame {f}: inspect word {wd}: {bad} texels differ"
        seq.hist_o = out["gpu_history"]
