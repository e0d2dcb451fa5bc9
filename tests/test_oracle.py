"""Pins the oracle (the FP64 / FP32 transcription of taa.comp + CasFilter) against checks that do
not restate its formulas: closed forms, numpy brute force, textbook kernels, the reference's own
compiled CasSetup, and algebraic invariants. Runs on the CPU."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import taastar
from taastar import INSPECT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32


# ---------------------------------------------------------------------------------------------
# helpers

def params(**kw):
    p = taastar.default_params()
    # a neutral base: RGB, no box, no rectification, no rejection, alpha from kw
    p.mUseYCoCg = 0
    p.mVarianceClipping = 0
    p.mColorClampingOrClipping = 0
    p.mInterpolationMode = 2
    p.mUseVelocityVectors = 2
    p.mVelocitySampleMode = 0
    p.mAlpha = 0.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def uniforms(p, reset=0, near=0.1, far=100.0, pad1=0.0):
    U = taastar.Uniforms()
    U.param[0] = p
    U.param[1] = p
    U.mResetHistory = reset
    U.mCamNearPlane, U.mCamFarPlane, U.pad1 = near, far, pad1
    for i in (0, 5, 10, 15):
        U.mHistoryViewProjMatrix[i] = 1.0
        U.mInverseViewProjMatrix[i] = 1.0
    return U


def frame(H, W, rng, vel=(0.0, 0.0), hist=None, depth=None, color=None):
    col = color if color is not None else np.concatenate(
        [rng.uniform(0.05, 2.0, (H, W, 3)), np.ones((H, W, 1))], -1).astype(f32)
    dep = depth if depth is not None else rng.uniform(0.1, 1.0, (H, W)).astype(f32)
    v = np.zeros((H, W, 4), f32)
    v[..., 0], v[..., 1], v[..., 3] = vel[0], vel[1], 1.0
    h = hist if hist is not None else np.concatenate([rng.uniform(0.05, 2.0, (H, W, 3)), np.zeros((H, W, 1))], -1)
    return col, dep, v, h


def run(oracle, U, col, dep, vel, hist, hdep=None, dtype=np.float64, cr16=True):
    hdep = dep if hdep is None else hdep
    return oracle.resolve(U, col, dep, vel, np.asarray(hist, dtype), hdep, dtype=dtype, cr16=cr16, threads=4)


def box(o):
    ins = o["inspect"]
    mn = ins[..., INSPECT["COLMIN"]:INSPECT["COLMIN"] + 3].copy().view(f32)
    mx = ins[..., INSPECT["COLMAX"]:INSPECT["COLMAX"] + 3].copy().view(f32)
    return mn, mx


def windows(a):
    """3x3 edge-clamped windows of [H, W, ...] in taa.comp's tap order: centre first, then
    (-1,-1) (0,-1) (1,-1) (-1,0) (1,0) (-1,1) (0,1) (1,1)."""
    pad = np.pad(a, [(1, 1), (1, 1)] + [(0, 0)] * (a.ndim - 2), mode="edge")
    H, W = a.shape[:2]
    offs = [(0, 0), (-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]
    return np.stack([pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dx, dy in offs])


def ycocg64(c):
    r, g, b = (c[..., i].astype(np.float64) for i in range(3))
    return np.stack([0.25 * r + 0.5 * g + 0.25 * b, 0.5 * r - 0.5 * b, -0.25 * r + 0.5 * g - 0.25 * b], -1)


# ---------------------------------------------------------------------------------------------
# jitter, CAS constants

def test_halton_known_answers(oracle):
    h2 = [1 / 2, 1 / 4, 3 / 4, 1 / 8, 5 / 8, 3 / 8, 7 / 8, 1 / 16]
    h3 = [1 / 3, 2 / 3, 1 / 9, 4 / 9, 7 / 9, 2 / 9, 5 / 9, 8 / 9]
    for i in range(8):
        assert oracle.lib().oracle_halton(i + 1, 2) == f32(h2[i])  # dyadic: exact
        assert abs(oracle.lib().oracle_halton(i + 1, 3) - h3[i]) <= 2 * np.spacing(f32(h3[i]))


def test_jitter_selection_closed_forms(oracle):
    W, H = 640, 360
    px, py = f32(2) / f32(W), f32(2) / f32(H)
    h2 = [1 / 2, 1 / 4, 3 / 4, 1 / 8, 5 / 8, 3 / 8, 7 / 8, 1 / 16]
    h3 = [1 / 3, 2 / 3, 1 / 9, 4 / 9, 7 / 9, 2 / 9, 5 / 9, 8 / 9]
    for f in range(20):
        (x, y), n = oracle.jitter(2, f, W, H)
        assert n == 8
        k = f % 8
        assert abs(x - px * (h2[k] - 0.5)) < 1e-8 and abs(y - py * (h3[k] - 0.5)) < 1e-8
        (x, y), n = oracle.jitter(0, f, W, H)
        q = [(-.25, -.25), (.25, -.25), (.25, .25), (-.25, .25)][f % 4]
        assert n == 4 and (x, y) == (px * f32(q[0]), py * f32(q[1]))
        (x, y), _ = oracle.jitter(4, f, W, H)
        assert abs(x - px * (2 * (f % 4) - 3) / 8) < 1e-9 and abs(y - py * (2 * ((f % 16) // 4) - 3) / 8) < 1e-9
    # slow motion divides the frame id, a fixed index pins it, rotation by 90 degrees swaps axes
    assert oracle.jitter(2, 13, W, H, slow=4)[0] == oracle.jitter(2, 3, W, H)[0]
    assert oracle.jitter(2, 13, W, H, fixed=5)[0] == oracle.jitter(2, 5, W, H)[0]
    (x0, y0), _ = oracle.jitter(2, 1, W, H)
    (xr, yr), _ = oracle.jitter(2, 1, W, H, rot=90.0)
    assert abs(xr + y0) < 1e-7 and abs(yr - x0) < 1e-7


def test_cas_setup_closed_form(oracle):
    for s in np.linspace(0, 1, 41):
        c0, c1 = oracle.cas_setup(float(s), 1920, 1080)
        peak = np.array(c1[0], np.uint32).view(f32)
        exact = -1.0 / (8.0 + (5.0 - 8.0) * float(f32(s)))
        assert abs(peak - exact) <= 2 * np.spacing(f32(abs(exact)))
        half = np.array(c1[1] & 0xFFFF, np.uint16).view(np.float16)
        assert abs(float(half) - exact) <= abs(exact) * 2 ** -10  # truncated to fp16
        assert c1[1] >> 16 == 0 and c0 == [0x3F800000, 0x3F800000, 0, 0] and c1[2] == 0x41000000


def test_cas_setup_against_committed_reference_vectors(oracle):
    """tests/golden/cas_setup.json was written by tests/golden/make_cas_setup.py from the
    reference's own ffx_cas.h compiled for the CPU."""
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "cas_setup.json")))
    assert len(g["cases"]) >= 100
    for case in g["cases"]:
        c0, c1 = oracle.cas_setup(case["sharpness"], case["in_w"], case["in_h"])
        assert c0 == case["const0"] and c1 == case["const1"], case
        k = taastar.cas_setup(case["sharpness"], case["in_w"], case["in_h"])
        assert list(k.const0) == case["const0"] and list(k.const1) == case["const1"]


# ---------------------------------------------------------------------------------------------
# neighbourhood box (taa.comp:259-319)

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_minmax_box_is_the_3x3_order_statistic(oracle, dtype):
    rng = np.random.default_rng(1)
    col, dep, vel, hist = frame(17, 23, rng)
    o = run(oracle, uniforms(params()), col, dep, vel, hist, dtype=dtype)
    mn, mx = box(o)
    w = windows(col[..., :3])
    assert np.array_equal(mn, w.min(0)) and np.array_equal(mx, w.max(0))


def test_minmax_box_in_ycocg(oracle):
    rng = np.random.default_rng(2)
    col, dep, vel, hist = frame(11, 19, rng)
    o = run(oracle, uniforms(params(mUseYCoCg=1)), col, dep, vel, hist)
    mn, mx = box(o)
    w = windows(ycocg64(col))
    assert np.array_equal(mn, w.min(0).astype(f32)) and np.array_equal(mx, w.max(0).astype(f32))


@pytest.mark.parametrize("gamma", [1.0, 0.75, 0.0])
def test_variance_box_is_mean_plus_minus_gamma_sigma(oracle, gamma):
    rng = np.random.default_rng(3)
    col, dep, vel, hist = frame(13, 21, rng)
    p = params(mVarianceClipping=1, mVarClipGamma=gamma)
    w = windows(col[..., :3].astype(np.float64))
    mean, std = w.mean(0), w.std(0)  # population sigma over the 9 taps
    for dtype, tol in [(np.float64, 1e-6), (np.float32, 2e-5)]:
        mn, mx = box(run(oracle, uniforms(p), col, dep, vel, hist, dtype=dtype))
        assert np.abs(mn - (mean - gamma * std)).max() < tol and np.abs(mx - (mean + gamma * std)).max() < tol
    if gamma == 0.0:  # the box collapses to the mean
        assert np.array_equal(mn, mx)


def test_shaped_box_and_chroma_shrink(oracle):
    rng = np.random.default_rng(4)
    col, dep, vel, hist = frame(9, 14, rng)
    w = windows(ycocg64(col))
    cross = w[[0, 2, 4, 5, 7]]
    mn = (w.min(0) + cross.min(0)) * 0.5
    mx = (w.max(0) + cross.max(0)) * 0.5
    o = run(oracle, uniforms(params(mUseYCoCg=1, mShapedNeighbourhood=1)), col, dep, vel, hist)
    gmn, gmx = box(o)
    assert np.abs(gmn - mn).max() < 1e-6 and np.abs(gmx - mx).max() < 1e-6
    # chroma shrink: same centre, luma extent kept, chroma extents halved
    o = run(oracle, uniforms(params(mUseYCoCg=1, mShapedNeighbourhood=1, mShrinkChromaAxis=1)), col, dep, vel, hist)
    smn, smx = box(o)
    assert np.abs((smn + smx) - (mn + mx)).max() < 1e-6
    ext = (mx - mn) * np.array([1.0, 0.5, 0.5])
    assert np.abs((smx - smn) - ext).max() < 1e-6


# ---------------------------------------------------------------------------------------------
# reprojection (taa.comp:371-438)

def test_closest_depth_tap_first_minimum_in_scan_order(oracle):
    rng = np.random.default_rng(5)
    H, W = 19, 27
    dep = rng.integers(0, 3, (H, W)).astype(f32) * f32(0.25) + f32(0.25)  # heavy ties
    dep[:4, :5] = 0.5  # an all-equal block -> tap 0 (-1,-1)
    col, _, vel, hist = frame(H, W, rng, depth=dep)
    o = run(oracle, uniforms(params(mVelocitySampleMode=2)), col, dep, vel, hist)
    pad = np.pad(dep, 1, mode="edge")
    scan = np.stack([pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    expect = np.argmin(scan, axis=0)  # numpy returns the first minimum
    assert np.array_equal(o["inspect"][..., INSPECT["CLOSEST_TAP"]], expect)
    assert np.all(o["inspect"][1:3, 1:4, INSPECT["CLOSEST_TAP"]] == 0)


def test_history_uv_follows_velocity(oracle):
    rng = np.random.default_rng(6)
    H, W = 12, 20
    vx, vy = f32(0.0137), f32(-0.0291)
    col, dep, vel, hist = frame(H, W, rng, vel=(vx, vy))
    for dtype in (np.float64, np.float32):
        o = run(oracle, uniforms(params()), col, dep, vel, hist, dtype=dtype)
        hu = o["inspect"][..., INSPECT["HIST_UV_X"]].copy().view(f32)
        hv = o["inspect"][..., INSPECT["HIST_UV_Y"]].copy().view(f32)
        x = np.arange(W)[None, :]
        y = np.arange(H)[:, None]
        if dtype == np.float64:
            ex, ey = (x + 0.5) / W - float(vx), (y + 0.5) / H - float(vy)
            assert np.abs(hu - ex).max() < 1e-7 and np.abs(hv - ey).max() < 1e-7
        else:  # float32 IEEE: (x + 0.5) / W - v, correctly rounded at each step
            ex = (x.astype(f32) + f32(0.5)) / f32(W) - vx
            ey = (y.astype(f32) + f32(0.5)) / f32(H) - vy
            assert np.array_equal(hu, np.broadcast_to(ex, (H, W))) and np.array_equal(hv, np.broadcast_to(ey, (H, W)))
        base_x = o["inspect"][..., INSPECT["HIST_TC_X"]]
        assert np.array_equal(base_x, np.floor(hu.astype(np.float64) * W - 0.5).astype(np.int32)) or dtype == np.float32


def test_velocity_sample_modes(oracle):
    rng = np.random.default_rng(7)
    H, W = 10, 12
    col, dep, _, hist = frame(H, W, rng)
    vel = np.zeros((H, W, 4), f32)
    vel[..., :2] = rng.uniform(-0.02, 0.02, (H, W, 2))
    vel[..., 3] = 1.0
    x = (np.arange(W)[None, :] + 0.5) / W
    y = (np.arange(H)[:, None] + 0.5) / H
    hu = lambda o: o["inspect"][..., INSPECT["HIST_UV_X"]].copy().view(f32)
    # mode 1: the reference assigns the last tap of its loop, (+1,+1), clamped (taa.comp:412)
    o = run(oracle, uniforms(params(mVelocitySampleMode=1)), col, dep, vel, hist)
    vpp = np.pad(vel, [(0, 1), (0, 1), (0, 0)], mode="edge")[1:, 1:]
    assert np.abs(hu(o) - (x - vpp[..., 0])).max() < 1e-7
    # mode 2: velocity of the closest-depth tap
    o = run(oracle, uniforms(params(mVelocitySampleMode=2)), col, dep, vel, hist)
    k = o["inspect"][..., INSPECT["CLOSEST_TAP"]]
    vpad = np.pad(vel, [(1, 1), (1, 1), (0, 0)], mode="edge")
    yy, xx = np.mgrid[0:H, 0:W]
    vsel = vpad[yy + k // 3, xx + k % 3]
    assert np.abs(hu(o) - (x - vsel[..., 0])).max() < 1e-7
    # movers only: .w < 0.5 means no velocity -> matrix reprojection with identity = static
    vel[..., 3] = 0.0
    o = run(oracle, uniforms(params(mUseVelocityVectors=1)), col, dep, vel, hist)
    assert np.abs(hu(o) - x).max() < 1e-7


def test_matrix_reprojection_translation(oracle):
    rng = np.random.default_rng(8)
    H, W = 8, 10
    col, dep, vel, hist = frame(H, W, rng)
    U = uniforms(params(mUseVelocityVectors=0))
    U.mHistoryViewProjMatrix[12] = 0.1   # clip-space translation x (column 3, row 0)
    U.mHistoryViewProjMatrix[13] = -0.06
    o = run(oracle, U, col, dep, vel, hist)
    hu = o["inspect"][..., INSPECT["HIST_UV_X"]].copy().view(f32)
    hv = o["inspect"][..., INSPECT["HIST_UV_Y"]].copy().view(f32)
    x = (np.arange(W)[None, :] + 0.5) / W
    y = (np.arange(H)[:, None] + 0.5) / H
    assert np.abs(hu - (x + 0.05)).max() < 1e-7 and np.abs(hv - (y - 0.03)).max() < 1e-7


# ---------------------------------------------------------------------------------------------
# history filters (taa.comp:441-549)

def keys_weight(d):
    d = np.abs(d)
    return np.where(d <= 1, 1.5 * d ** 3 - 2.5 * d ** 2 + 1, np.where(d < 2, -0.5 * d ** 3 + 2.5 * d ** 2 - 4 * d + 2, 0.0))


def filtered_history(oracle, hist, vel_px, mode, dtype=np.float64, cr16=True):
    """Output history with alpha = 0 and no rectification = the filtered history at uv - v."""
    H, W = hist.shape[:2]
    rng = np.random.default_rng(0)
    col, dep, vel, _ = frame(H, W, rng, vel=(vel_px[0] / W, vel_px[1] / H))
    o = run(oracle, uniforms(params(mInterpolationMode=mode)), col, dep, vel, hist, dtype=dtype, cr16=cr16)
    return o["history"], vel


def test_catmull_rom_is_the_keys_cubic(oracle):
    rng = np.random.default_rng(9)
    H, W = 16, 20
    hist = np.concatenate([rng.uniform(0, 1, (H, W, 3)), rng.uniform(0, 1, (H, W, 1))], -1)
    got, vel = filtered_history(oracle, hist, (0.37, -0.81), 2)
    # textbook separable Keys (a = -1/2) convolution with edge clamping, in numpy
    for y in range(H):
        for x in range(W):
            px = ((x + 0.5) / W - float(vel[y, x, 0])) * W - 0.5  # float(): keep numpy out of float32
            py = ((y + 0.5) / H - float(vel[y, x, 1])) * H - 0.5
            bx, by = int(np.floor(px)), int(np.floor(py))
            acc = np.zeros(3)
            for j in range(by - 1, by + 3):
                for i in range(bx - 1, bx + 3):
                    acc += keys_weight(px - i) * keys_weight(py - j) * hist[min(max(j, 0), H - 1), min(max(i, 0), W - 1), :3]
            assert np.allclose(got[y, x, :3], acc, rtol=0, atol=1e-12), (x, y)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_history_filters_reproduce_polynomials(oracle, mode):
    """Catmull-Rom reproduces quadratics, the B-spline and bilinear filters linear ramps."""
    H, W = 24, 24
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    lin = 0.3 + 0.02 * xx - 0.015 * yy
    quad = 0.2 + 0.01 * xx + 0.003 * xx * xx - 0.002 * yy * yy + 0.001 * xx * yy
    fields = [lin, quad] if mode == 2 else [lin]
    for fld in fields:
        hist = np.stack([fld, 2 * fld, fld + 1, np.zeros_like(fld)], -1)
        got, vel = filtered_history(oracle, hist, (0.43, 0.71), mode)
        inner = (slice(4, -4), slice(4, -4))
        # sample position in texel units: x - v (texel centres at integer coordinates), with the
        # velocity as stored (float32 UV units)
        sx, sy = xx - vel[..., 0].astype(np.float64) * W, yy - vel[..., 1].astype(np.float64) * H
        expect = (0.3 + 0.02 * sx - 0.015 * sy) if fld is lin else \
            (0.2 + 0.01 * sx + 0.003 * sx * sx - 0.002 * sy * sy + 0.001 * sx * sy)
        assert np.abs(got[..., 0][inner] - expect[inner]).max() < 1e-9


def test_catmull_rom_9tap_equals_16tap(oracle):
    rng = np.random.default_rng(10)
    H, W = 21, 17
    hist = np.concatenate([rng.uniform(0, 3, (H, W, 3)), rng.uniform(0, 1, (H, W, 1))], -1)
    for v in [(0.3, 0.6), (-2.7, 1.2), (4.4, -3.9), (0.0, 0.0)]:  # includes footprints over the edges
        a, _ = filtered_history(oracle, hist, v, 2, cr16=True)
        b, _ = filtered_history(oracle, hist, v, 2, cr16=False)
        assert np.abs(a - b).max() < 1e-12
        a32, _ = filtered_history(oracle, hist, v, 2, dtype=np.float32, cr16=True)
        b32, _ = filtered_history(oracle, hist, v, 2, dtype=np.float32, cr16=False)
        assert np.abs(a32 - b32).max() < 1e-5


def test_integer_shift_is_exact(oracle):
    """At texel centres (f = 0) Catmull-Rom returns the texel itself, in FP32 too."""
    rng = np.random.default_rng(11)
    H, W = 8, 16  # powers of two: uv and the velocity are exact in fp32
    hist = np.concatenate([rng.uniform(0, 1, (H, W, 3)), np.zeros((H, W, 1))], -1).astype(f32)
    for dtype in (np.float64, np.float32):
        got, _ = filtered_history(oracle, hist, (2.0, -1.0), 2, dtype=dtype)
        src = np.pad(hist, [(1, 1), (2, 2), (0, 0)], mode="edge")
        shifted = src[2:H + 2, 0:W]  # value at (x - 2, y + 1), clamped
        assert np.array_equal(got[..., :3].astype(f32), shifted[..., :3])


# ---------------------------------------------------------------------------------------------
# rejection, rectification, blend (taa.comp:787-900)

def test_clip_lands_on_the_box(oracle):
    rng = np.random.default_rng(12)
    H, W = 14, 18
    col, dep, vel, _ = frame(H, W, rng)
    hist = np.concatenate([rng.uniform(-3, 5, (H, W, 3)), np.zeros((H, W, 1))], -1)
    o = run(oracle, uniforms(params(mColorClampingOrClipping=2, mInterpolationMode=0)), col, dep, vel, hist)
    mn, mx = (b.astype(np.float64) for b in box(o))
    h = o["history"][..., :3]
    c, e = 0.5 * (mx + mn), 0.5 * (mx - mn) + 1e-7
    m = np.abs((h - c) / e).max(-1)
    outside = np.abs((hist[..., :3] - c) / e).max(-1) > 1
    assert outside.mean() > 0.5
    assert np.abs(m[outside] - 1.0).max() < 1e-9                      # clipped onto the surface
    assert np.array_equal(h[~outside], hist[..., :3][~outside])        # inside: untouched
    # the clipped point lies on the segment from the box centre to the history sample
    d0, d1 = hist[..., :3] - c, h - c
    cross = np.cross(d0[outside], d1[outside])
    assert np.abs(cross).max() < 1e-9
    o = run(oracle, uniforms(params(mColorClampingOrClipping=1, mInterpolationMode=0)), col, dep, vel, hist)
    assert np.array_equal(o["history"][..., :3], np.clip(hist[..., :3], mn, mx))


def test_current_sample_unjitter(oracle):
    """getCurrentColor (taa.comp:216-220) samples the texel itself unless mUnjitterCurrentSample is
    set, whatever mUnjitterNeighbourhood says; with it, a bilinear clamp-to-edge sample at
    (p + 0.5) / size + jitter * 0.5 * mUnjitterFactor. On a reset frame the screen is that sample."""
    rng = np.random.default_rng(29)
    H, W = 9, 12
    col, dep, vel, hist = frame(H, W, rng)
    jit = (0.6 / W, -0.4 / H)
    for unj_nb, unj_cur in ((1, 0), (0, 0), (1, 1), (0, 1)):
        U = uniforms(params(mUnjitterNeighbourhood=unj_nb, mUnjitterCurrentSample=unj_cur, mUnjitterFactor=1.0), reset=1)
        U.mJitterNdc[0], U.mJitterNdc[1] = jit
        o = run(oracle, U, col, dep, vel, hist)
        if not unj_cur:
            assert np.array_equal(o["screen"][..., :3], col[..., :3].astype(np.float64)), (unj_nb, unj_cur)
            continue
        # numpy bilinear, clamp to edge, at the unjittered texel centres
        c = col[..., :3].astype(np.float64)
        ju, jv = np.float64(f32(jit[0])) * 0.5, np.float64(f32(jit[1])) * 0.5
        px = (np.arange(W) + 0.5) / W + ju
        py = (np.arange(H) + 0.5) / H + jv
        fx, fy = px * W - 0.5, py * H - 0.5
        x0, y0 = np.floor(fx).astype(int), np.floor(fy).astype(int)
        tx, ty = fx - x0, fy - y0
        xa, xb = np.clip(x0, 0, W - 1), np.clip(x0 + 1, 0, W - 1)
        ya, yb = np.clip(y0, 0, H - 1), np.clip(y0 + 1, 0, H - 1)
        top = c[ya][:, xa] * (1 - tx)[None, :, None] + c[ya][:, xb] * tx[None, :, None]
        bot = c[yb][:, xa] * (1 - tx)[None, :, None] + c[yb][:, xb] * tx[None, :, None]
        ref = top * (1 - ty)[:, None, None] + bot * ty[:, None, None]
        assert np.abs(o["screen"][..., :3] - ref).max() < 1e-9, (unj_nb, unj_cur)
        assert not np.allclose(ref, c)  # the jitter really moved the sample


def test_blend_closed_forms(oracle):
    rng = np.random.default_rng(13)
    H, W = 10, 13
    col, dep, vel, hist = frame(H, W, rng)
    # reset: output = current colour, also through the YCoCg round trip
    for ycocg in (0, 1):
        o = run(oracle, uniforms(params(mUseYCoCg=ycocg, mAlpha=0.3), reset=1), col, dep, vel, hist)
        assert np.array_equal(o["screen"][..., :3], col[..., :3].astype(np.float64))
        assert np.all(o["screen"][..., 3] == 1)
    # alpha: out = hist*(1-a) + cur*a
    a = float(f32(0.3))  # mAlpha is a float field
    o = run(oracle, uniforms(params(mAlpha=a, mInterpolationMode=0)), col, dep, np.zeros_like(vel) + [0, 0, 0, 1], hist)
    assert np.abs(o["history"][..., :3] - (hist[..., :3] * (1 - a) + col[..., :3].astype(np.float64) * a)).max() < 1e-12
    # Lottes with history == current: diff = 0 -> alpha = mMinAlpha
    p = params(mLumaWeightingLottes=1, mUseYCoCg=1, mInterpolationMode=0)
    o = run(oracle, uniforms(p), col, dep, np.zeros_like(vel) + [0, 0, 0, 1], col.astype(np.float64))
    blend = o["inspect"][..., INSPECT["BLEND"]].copy().view(f32)
    assert np.all(blend == f32(p.mMinAlpha))
    # Lottes with |lc - lh| >= max(lc, lh): alpha = mMaxAlpha
    o = run(oracle, uniforms(p), col, dep, np.zeros_like(vel) + [0, 0, 0, 1], np.zeros_like(hist))
    assert np.all(o["inspect"][..., INSPECT["BLEND"]].copy().view(f32) == f32(p.mMaxAlpha))


def test_rejections(oracle):
    rng = np.random.default_rng(14)
    H, W = 12, 16
    col, dep, vel, hist = frame(H, W, rng, vel=(0.2, 0.0))  # shift 0.2 of the width: the left part leaves the screen
    p = params(mRejectOutside=1, mRejectionAlpha=1.0, mAlpha=0.05)
    o = run(oracle, uniforms(p), col, dep, vel, hist)
    hu = (np.arange(W) + 0.5) / W - 0.2
    out = np.broadcast_to(hu < 0, (H, W))
    flags = o["inspect"][..., INSPECT["FLAGS"]]
    assert np.array_equal((flags & taastar.FLAG_OUTSIDE) != 0, out)
    assert np.array_equal(o["history"][out][:, :3], col[out][:, :3].astype(np.float64))  # alpha 1
    # depth rejection, raw depth (taa.comp:818-822) and the linear-depth extension
    _, _, vel0, _ = frame(H, W, rng)
    hdep = rng.uniform(0.1, 1.0, (H, W)).astype(f32)
    o = run(oracle, uniforms(params(mDepthCulling=1)), col, dep, vel0, hist, hdep=hdep)
    hd, d = hdep.astype(np.float64), dep.astype(np.float64)
    rej_raw = np.abs(hd - d) > 0.1 * (1.0 - hd)
    assert np.array_equal((o["inspect"][..., INSPECT["FLAGS"]] & taastar.FLAG_DEPTH_REJECT) != 0, rej_raw)
    n, fr = 0.1, 100.0
    lin = lambda z: n * fr / (fr + z * (n - fr))
    o = run(oracle, uniforms(params(mDepthCulling=1), near=n, far=fr, pad1=0.05), col, dep, vel0, hist, hdep=hdep)
    rej_lin = np.abs(lin(hd) - lin(d)) > 0.05
    assert np.array_equal((o["inspect"][..., INSPECT["FLAGS"]] & taastar.FLAG_DEPTH_REJECT) != 0, rej_lin)


def test_passthrough_bypass_and_karis(oracle):
    rng = np.random.default_rng(15)
    H, W = 9, 11
    col, dep, vel, hist = frame(H, W, rng)
    o = run(oracle, uniforms(params(mPassThrough=1)), col, dep, vel, hist)
    assert np.array_equal(o["screen"][..., :3], col[..., :3].astype(np.float64))
    assert np.array_equal(o["history"][..., :3], hist[..., :3]) and np.all(o["history"][..., 3] == 1)
    U = uniforms(params())
    U.mBypassHistoryUpdate = 1
    o = run(oracle, U, col, dep, vel, hist)
    assert np.array_equal(o["screen"][..., :3], hist[..., :3]) and np.array_equal(o["history"][..., :3], hist[..., :3])
    # Karis: history holds tone-mapped colour c/(1+max c); the screen un-maps it back
    o = run(oracle, uniforms(params(mToneMapLumaKaris=1), reset=1), col, dep, vel, hist)
    c = col[..., :3].astype(np.float64)
    assert np.abs(o["history"][..., :3] - c / (1 + c.max(-1, keepdims=True))).max() < 1e-15
    assert np.abs(o["screen"][..., :3] - c).max() < 1e-12


# ---------------------------------------------------------------------------------------------
# CAS (ffx_cas.h:424-537)

def u2f(u):
    return np.array(u, np.uint32).view(f32)


def test_cas_integer_approximations_approximate_the_right_functions():
    # APrxLoRcpF1 / APrxLoSqrtF1 / APrxMedRcpF1 (ffx_a.h:1455-1457) on fp32 bit patterns
    x = np.geomspace(1e-3, 1e3, 2001).astype(f32)
    b = x.view(np.uint32)
    lo_rcp = (np.uint32(0x7EF07EBB) - b).view(f32)
    lo_sqrt = ((b >> 1) + np.uint32(0x1FBC4639)).view(f32)
    m = (np.uint32(0x7EF19FFF) - b).view(f32)
    med_rcp = m * (-m * x + f32(2))
    assert np.abs(lo_rcp * x - 1).max() < 0.07
    assert np.abs(lo_sqrt / np.sqrt(x) - 1).max() < 0.045
    assert np.abs(med_rcp * x - 1).max() < 4e-3


def test_cas_flat_and_range(oracle):
    c0, c1 = oracle.cas_setup(0.5, 64, 48)
    img = np.zeros((48, 64, 4), np.float64)
    for v in (0.02, 0.3, 0.77):
        img[...] = v
        out = oracle.cas(img, c1[0], dtype=np.float64)
        assert np.abs(out[..., :3] - v).max() < 3e-3 * v  # flat input: only the rcp approximation error
    rng = np.random.default_rng(16)
    hdr = rng.lognormal(-1, 2, (48, 64, 4))
    out = oracle.cas(hdr, c1[0], dtype=np.float64)
    assert out[..., :3].min() >= 0 and out[..., :3].max() <= 1 and np.all(out[..., 3] == 1)
    # amp = 0 when the local max exceeds 1: sharpness has no effect, output = sat(e)
    img = np.full((48, 64, 4), 1.5)
    img[::2, ::2] = 0.4
    outs = [oracle.cas(img, oracle.cas_setup(s, 64, 48)[1][0], dtype=np.float64) for s in (0.0, 1.0)]
    assert np.array_equal(outs[0], outs[1])


def test_cas_sharpens_an_edge(oracle):
    c0, c1 = oracle.cas_setup(1.0, 32, 8)
    img = np.full((8, 32, 4), 0.2)
    img[:, 16:] = 0.6
    out = oracle.cas(img, c1[0], dtype=np.float64)
    assert out[4, 15, 0] < 0.2 and out[4, 16, 0] > 0.6  # overshoot on both sides of the step
    assert abs(out[4, 3, 0] - 0.2) < 1e-3 and abs(out[4, 28, 0] - 0.6) < 2e-3
    a = oracle.cas(img.astype(f32), c1[0], dtype=np.float32)
    assert np.abs(a - out).max() < 1e-6


def test_fp32_transcription_tracks_fp64(oracle):
    """The two instantiations on a real scene, 4 frames of the full feature set."""
    from taastar.workloads import (cas_for, decode, encode_color, encode_velocity, scene_for, uniforms_for,
                                   velocity_as_read, workload)
    w = workload("c2")
    W, H = 128, 72
    sc = scene_for(w, width=W, height=H)
    k = cas_for(w, W, H)
    h32 = np.zeros((H, W, 4), f32)
    h64 = np.zeros((H, W, 4))
    prev = np.zeros((H, W), f32)
    for f in range(4):
        col, dep, vel4, jit = sc.render(f)
        U = uniforms_for(w, f, jit)
        args = (decode(encode_color(col, w["fmt"])), dep, velocity_as_read(encode_velocity(vel4, w["vel"])))
        o32 = oracle.resolve(U, *args, h32, prev, dtype=np.float32, threads=4)
        o64 = oracle.resolve(U, *args, h64, prev, dtype=np.float64, threads=4)
        s32 = oracle.cas(o32["screen"], k.const1[0], dtype=np.float32)
        s64 = oracle.cas(o64["screen"], k.const1[0], dtype=np.float64)
        d = np.abs(s32 - s64)
        assert d.max() < 1e-3 and d.mean() < 1e-6
        h32, h64, prev = o32["history"], o64["history"], dep


# ---------------------------------------------------------------------------------------------
# TAAU: input (render) size differs from the output size, or mUpsampling is set

def test_taau_scale1_zero_jitter_equals_plain(oracle):
    """At equal sizes with power-of-two dimensions every uv lands exactly on a texel centre, the
    bilinear depth / velocity samples reduce to fetches and the upsampled current sample is the
    texel itself: mUpsampling changes nothing (all features that sample the current frame on)."""
    rng = np.random.default_rng(41)
    H, W = 32, 64
    col, dep, vel, hist = frame(H, W, rng, vel=(0.015625, -0.03125))
    vel[..., 0] += rng.integers(-4, 5, (H, W)).astype(f32) / 256.0  # dyadic: exact uv arithmetic
    p = params(mUseYCoCg=1, mVarianceClipping=1, mColorClampingOrClipping=2, mVelocitySampleMode=2,
               mDepthCulling=1, mRejectOutside=1, mLumaWeightingLottes=1, mAlpha=0.1)
    a = run(oracle, uniforms(p), col, dep, vel, hist)
    U = uniforms(p)
    U.mUpsampling = 1
    b = run(oracle, U, col, dep, vel, hist)
    for k in ("screen", "history"):
        np.testing.assert_array_equal(a[k], b[k])
    np.testing.assert_array_equal(a["inspect"], b["inspect"])


def test_taau_nearest_input_without_upsampling(oracle):
    """2x output without mUpsampling: iuv_lores = int(uv * lo) = output texel // 2, so a reset frame
    is the nearest-neighbour enlargement of the input (taa.comp:724, 753)."""
    rng = np.random.default_rng(42)
    Hl, Wl = 12, 20
    col, dep, vel, _ = frame(Hl, Wl, rng)
    hist = np.zeros((2 * Hl, 2 * Wl, 4))
    o = run(oracle, uniforms(params(), reset=1), col, dep, vel, hist)
    want = np.repeat(np.repeat(col[..., :3].astype(np.float64), 2, 0), 2, 1)
    np.testing.assert_array_equal(o["screen"][..., :3], want)


def test_taau_upsampled_sample_selection(oracle):
    """2x output with mUpsampling and no jitter: input texel i (centre i + 0.5) lands on output
    texel 2i + 1, so only odd/odd output texels receive a current sample (beta = 1); the others get
    beta = 0. On a reset frame (alpha = beta = 1) the former are the input texel, the latter the
    zero colour getCurrentUpsampledColor returns when nothing was found (taa.comp:222-257, 897)."""
    rng = np.random.default_rng(43)
    Hl, Wl = 10, 14
    col, dep, vel, _ = frame(Hl, Wl, rng)
    H, W = 2 * Hl, 2 * Wl
    U = uniforms(params(), reset=1)
    U.mUpsampling = 1
    o = run(oracle, U, col, dep, vel, np.zeros((H, W, 4)))
    scr = o["screen"][..., :3]
    odd = np.zeros((H, W), bool)
    odd[1::2, 1::2] = True
    np.testing.assert_array_equal(scr[odd].reshape(Hl, Wl, 3), col[..., :3].astype(np.float64))
    assert not scr[~odd].any()
    # a blended frame (alpha 0.5): texels without a sample keep the history (alpha * beta = 0)
    hist = rng.uniform(0.1, 1.0, (H, W, 4))
    U.mResetHistory = 0
    p = params(mAlpha=0.5)
    U.param[0] = p
    U.param[1] = p
    o = run(oracle, U, col, dep, vel, hist)
    np.testing.assert_allclose(o["history"][~odd][:, :3], hist[~odd][:, :3], rtol=1e-12)
    np.testing.assert_allclose(o["history"][odd][:, :3].reshape(Hl, Wl, 3),
                               0.5 * hist[odd][:, :3].reshape(Hl, Wl, 3) + 0.5 * col[..., :3], rtol=1e-12)


def test_taau_jittered_sample_moves_with_the_jitter(oracle):
    """texelJitter = -JITTER_UV * lo. +1/4 input texel moves input texel i's sample to i + 0.75,
    output 2i + 1.5 -> texel 2i + 1 (odd columns); -1/4 moves it to i + 0.25, output 2i + 0.5 ->
    texel 2i (even columns) (taa.comp:238-245)."""
    rng = np.random.default_rng(44)
    Hl, Wl = 6, 8
    col, dep, vel, _ = frame(Hl, Wl, rng)
    H, W = 2 * Hl, 2 * Wl
    for tj, first in ((0.25, 1), (-0.25, 0)):
        U = uniforms(params(mUnjitterFactor=1.0), reset=1)
        U.mUpsampling = 1
        # JITTER_UV = mJitterNdc * 0.5 * mUnjitterFactor; texelJitter = -JITTER_UV * lo
        U.mJitterNdc[0] = -2.0 * tj / Wl
        U.mJitterNdc[1] = 0.0
        o = run(oracle, U, col, dep, vel, np.zeros((H, W, 4)))
        got = o["screen"][1::2, :, :3]  # odd rows carry samples (no vertical jitter)
        lit = np.flatnonzero(got.any(-1).any(0))
        assert lit[0] == first and np.all(np.diff(lit) == 2), (tj, lit)


def test_taau_depth_test_and_velocity_at_input_size(oracle):
    """Velocity is sampled bilinearly at the output uv (a linear ramp is reproduced), and the depth
    test fetches uHistoryDepth at int(historyUv * input size) (taa.comp:393, 818)."""
    rng = np.random.default_rng(45)
    Hl, Wl, H, W = 16, 24, 40, 60
    col, dep, vel, _ = frame(Hl, Wl, rng)
    xs = (np.arange(Wl) + 0.5) / Wl
    vel[..., 0] = (0.01 + 0.02 * xs)[None, :]
    vel[..., 1] = -0.005
    p = params(mDepthCulling=1)
    o = run(oracle, uniforms(p), col, dep, vel, np.zeros((H, W, 4)))
    u = (np.arange(W) + 0.5) / W
    v = (np.arange(H) + 0.5) / H
    # inside the outermost half input texel the clamp makes the ramp flat
    inner = (u > 0.5 / Wl) & (u < 1 - 0.5 / Wl)
    hu = o["inspect"][..., INSPECT["HIST_UV_X"]].copy().view(f32)
    hv = o["inspect"][..., INSPECT["HIST_UV_Y"]].copy().view(f32)
    np.testing.assert_allclose(hu[:, inner], (u - (0.01 + 0.02 * u))[None, inner].repeat(H, 0), atol=2e-7)
    np.testing.assert_allclose(hv, (v + 0.005)[:, None].repeat(W, 1), atol=1e-7)
    dtx = o["inspect"][..., INSPECT["DEPTH_TC_X"]]
    dty = o["inspect"][..., INSPECT["DEPTH_TC_Y"]]
    np.testing.assert_array_equal(dtx, (hu.astype(np.float64) * Wl).astype(np.int64))
    np.testing.assert_array_equal(dty, (hv.astype(np.float64) * Hl).astype(np.int64))


def test_kernels_division_by_nine_is_the_ieee_division(oracle):
    """The FP32 kernel divides the variance moments by 9 with a multiply and two FMAs; for every
    single-precision input that equals the correctly rounded IEEE division m / 9 the transcription
    performs (exhaustive over all 2^32 bit patterns, OpenMP)."""
    L = oracle.lib()
    L.oracle_div9_mismatches.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
    L.oracle_div9_mismatches.restype = C.c_uint64
    assert L.oracle_div9_mismatches(0, 1 << 32, os.cpu_count() or 1) == 0


# ---------------------------------------------------------------------------------------------
# velocity producer (fwd_geometry.frag:288-295 from depth)

def view_distances(H, W, rng):
    """Piecewise-constant view distances (0.5 .. 40 units) in vertical bands plus a slanted plane."""
    d = np.repeat(rng.uniform(0.5, 40.0, W // 8 + 1), 8)[:W][None, :].repeat(H, 0)
    d[H // 2:] = 2.0 + 0.05 * np.arange(W)[None, :]
    return d


def depth_and_distance(dist, P):
    """NDC depth of each distance, stored as float32, and the distance that float32 depth encodes."""
    from taastar.workloads import depth_from_view_distance
    d32 = depth_from_view_distance(dist, P).astype(f32)
    return d32, P[2, 3] / (d32.astype(np.float64) + P[2, 2])


def test_velocity_still_camera_jitter_cancels(oracle):
    """A still camera with different current and previous jitters: the depth is rendered through the
    jittered projection and the jitters are removed in NDC (fwd_geometry.frag:292-293), so every
    texel's motion is zero to the float32 rounding of the matrices."""
    from taastar.workloads import perspective, reproject_for, translate
    rng = np.random.default_rng(70)
    H, W = 24, 40
    P = perspective(aspect=W / H)
    V = translate(0.3, -1.0, -5.0)
    d32, _ = depth_and_distance(view_distances(H, W, rng), P)
    R, _ = reproject_for(P, V, V, (0.7 / W, -0.4 / H), (-0.9 / W, 0.6 / H))
    v = oracle.velocity(R, d32, dtype=np.float64)
    assert np.abs(v[..., :3]).max() < 2e-6          # a dropped or mis-signed jitter term gives > 1e-2
    assert np.all(v[..., 3] == 0.0)


def test_velocity_translated_camera_closed_form(oracle):
    """Camera moving by (dx, 0, dz) between frames, no jitter: a point at view distance r with NDC
    (nx, ny) was at r + dz, x + dx in the previous view, so v.x = (nx - P00 (nx r / P00 + dx) / (r + dz)) / 2,
    v.y = ny (1 - r / (r + dz)) / 2, v.z = depth(r) - depth(r + dz)."""
    from taastar.workloads import depth_from_view_distance, perspective, reproject_for, translate
    rng = np.random.default_rng(71)
    H, W = 30, 48
    P = perspective(fovy_deg=50.0, aspect=W / H)
    dx, dz = 0.02, 0.05
    V_prev = translate(0.0, -1.0, -3.0)
    V_cur = translate(-dx, 0.0, dz) @ V_prev     # camera moved +dx along x and dz towards -z
    d32, r = depth_and_distance(view_distances(H, W, rng), P)
    R, _ = reproject_for(P, V_cur, V_prev, (0.0, 0.0), (0.0, 0.0))
    v = oracle.velocity(R, d32, dtype=np.float64)
    nx = ((np.arange(W) + 0.5) / W * 2 - 1)[None, :]
    ny = ((np.arange(H) + 0.5) / H * 2 - 1)[:, None]
    xc = nx * r / P[0, 0]
    ex = 0.5 * (nx - P[0, 0] * (xc + dx) / (r + dz))
    ey = 0.5 * ny * (1.0 - r / (r + dz))
    ez = d32.astype(np.float64) - depth_from_view_distance(r + dz, P)
    np.testing.assert_allclose(v[..., 0], ex, atol=2e-6)
    np.testing.assert_allclose(v[..., 1], ey, atol=2e-6)
    np.testing.assert_allclose(v[..., 2], ez, atol=2e-6)
    assert np.abs(ex).max() > 1e-3 and np.abs(ez).max() > 1e-4  # the motion is far above the tolerance


def test_velocity_moving_object_closed_form(oracle):
    """Still camera, object 1 moved by ox between frames (previous = current + ox): its texels get
    v.x = -P00 ox / (2 r), v.y = v.z = 0, w = 1; static texels and ids beyond the object table
    get no motion and keep their id in w (fwd_geometry.frag:295)."""
    from taastar.workloads import perspective, reproject_for, translate
    rng = np.random.default_rng(72)
    H, W = 18, 32
    P = perspective(aspect=W / H)
    V = translate(0.0, 0.0, -2.0)
    ox = 0.01
    d32, r = depth_and_distance(view_distances(H, W, rng), P)
    R, obj = reproject_for(P, V, V, (0.0, 0.0), (0.0, 0.0), objects=[translate(ox, 0.0, 0.0)])
    ids = np.zeros((H, W), np.uint32)
    ids[: H // 3] = 1
    ids[H // 3: 2 * H // 3] = 3
    v = oracle.velocity(R, d32, ids=ids, objects=obj, dtype=np.float64)
    moved = ids == 1
    np.testing.assert_allclose(v[..., 0][moved], (-0.5 * P[0, 0] * ox / r)[moved], atol=2e-6)
    assert np.abs(v[..., 1:3][moved]).max() < 2e-6
    assert np.abs(v[..., :3][~moved]).max() < 2e-6
    np.testing.assert_array_equal(v[..., 3], ids.astype(np.float64))


def test_velocity_agrees_with_the_resolves_matrix_path(oracle):
    """With zero jitter the producer and the resolve's matrix reprojection (taa.comp:424-431, the
    other transcription) describe the same motion: uv - v.xy is the matrix path's history uv and
    depth - v.z its expected history depth."""
    from taastar.workloads import perspective, reproject_for, rotate_y, translate
    rng = np.random.default_rng(73)
    H, W = 20, 36
    P = perspective(aspect=W / H)
    V_prev = translate(0.1, -1.0, -4.0)
    V_cur = rotate_y(0.7) @ translate(-0.03, 0.01, 0.04) @ V_prev
    d32, _ = depth_and_distance(view_distances(H, W, rng), P)
    R, _ = reproject_for(P, V_cur, V_prev, (0.0, 0.0), (0.0, 0.0))
    v = oracle.velocity(R, d32, dtype=np.float64)
    p = params(mUseVelocityVectors=0)
    U = uniforms(p)
    U.mInverseViewProjMatrix[:] = list(R.inverse_view_proj)
    U.mHistoryViewProjMatrix[:] = list(R.prev_view_proj)
    col, _, vel, hist = frame(H, W, rng)
    o = run(oracle, U, col, d32, vel, hist)
    hu = o["inspect"][..., INSPECT["HIST_UV_X"]].copy().view(f32)
    hv = o["inspect"][..., INSPECT["HIST_UV_Y"]].copy().view(f32)
    u = ((np.arange(W) + 0.5) / W)[None, :]
    vv = ((np.arange(H) + 0.5) / H)[:, None]
    np.testing.assert_allclose(u - v[..., 0], hu, atol=1e-6)
    np.testing.assert_allclose(vv - v[..., 1], hv, atol=1e-6)
    assert np.abs(v[..., :2]).max() > 5e-3


def test_velocity_fp32_tracks_fp64(oracle):
    from taastar.workloads import camera_case
    c = camera_case(96, 54, seed=3, n_objects=2)
    a = oracle.velocity(c["reproject"], c["depth"], c["ids"], c["objects"], dtype=np.float64)
    b = oracle.velocity(c["reproject"], c["depth"], c["ids"], c["objects"], dtype=np.float32)
    assert np.abs(a - b).max() < 2e-5


# ---------------------------------------------------------------------------------------------
# segmentation mask (taa.comp:566-702, 773-784, 932-938, 959)

def seg_case(H, W, rng, flags, **weights):
    """A frame with polygons of constant depth and colour, their material ids and normals, and a
    previous frame equal to it; only `flags` set, indicator weights from `weights`."""
    from taastar.workloads import segmentation_inputs
    n = 12
    cx, cy = rng.uniform(0, W, n), rng.uniform(0, H, n)
    yy, xx = np.mgrid[0:H, 0:W]
    lab = np.argmin((xx[..., None] - cx) ** 2 + (yy[..., None] - cy) ** 2, axis=-1)  # Voronoi cells
    dep = rng.uniform(0.2, 0.9, n)[lab].astype(f32)
    col = np.concatenate([rng.uniform(0.05, 2.0, (n, 3))[lab], np.ones((H, W, 1))], -1).astype(f32)
    mat, uvn = segmentation_inputs(dep)
    p = params(mRayTraceAugment=1, mRayTraceAugmentFlags=flags, mRayTraceHistoryCount=8,
               mRayTraceAugment_WNrm=weights.get("nrm", 0.0), mRayTraceAugment_WDpt=weights.get("dpt", 0.0),
               mRayTraceAugment_WMId=weights.get("mid", 0.0), mRayTraceAugment_WLum=weights.get("lum", 0.0),
               mRayTraceAugment_Thresh=weights.get("thresh", 1e30), mRejectionAlpha=1.0, mAlpha=0.1)
    v = np.zeros((H, W, 4), f32)
    v[..., 3] = 1.0
    hist = np.concatenate([rng.uniform(0.05, 2.0, (H, W, 3)), np.zeros((H, W, 1))], -1)
    return dict(p=p, col=col, dep=dep, vel=v, hist=hist, mat=mat, prev_mat=mat.copy(), uvn=uvn,
                prev_mask=np.zeros((H, W), np.uint32))


def run_seg(oracle, c, U=None, dtype=np.float64, mask_init=None):
    U = U if U is not None else uniforms(c["p"])
    return oracle.resolve_seg(U, c["col"], c["dep"], c["vel"], np.asarray(c["hist"], dtype), c["dep"], c["mat"],
                              c["prev_mat"], c["uvn"], c["prev_mask"], mask_init=mask_init, dtype=dtype, threads=4)


def seg_total(o):
    return o["inspect"][..., INSPECT["SEG_TOTAL"]].copy().view(f32).astype(np.float64)


def test_seg_sincos_is_within_two_ulp(oracle):
    a = np.concatenate([np.linspace(-10, 10, 400001), np.linspace(-200, 200, 40001)]).astype(f32)
    s, c = oracle.seg_sincos(a)
    a64 = a.astype(np.float64)
    for got, ref in ((s, np.sin(a64)), (c, np.cos(a64))):
        ulp = np.spacing(np.abs(ref).astype(f32)).astype(np.float64)
        assert (np.abs(got - ref) / np.maximum(ulp, 2.0 ** -149)).max() < 2.0


def test_seg_depth_and_luminance_indicators_are_sobel_magnitudes(oracle):
    """taa.comp:651-660, 673-682: |Sobel| of linear depth and of YCoCg luma with clamped edges,
    against scipy's Sobel operator."""
    from scipy import ndimage
    rng = np.random.default_rng(80)
    H, W = 40, 56
    for flag, key in ((taastar.RTFLAG["DPT"], "dpt"), (taastar.RTFLAG["LUM"], "lum")):
        c = seg_case(H, W, rng, flag, **{key: 1.0})
        o = run_seg(oracle, c)
        if key == "dpt":
            d = c["dep"].astype(np.float64)
            img = 0.1 * 100.0 / (100.0 + d * (0.1 - 100.0))
        else:
            rgb = c["col"][..., :3].astype(np.float64)
            img = 0.25 * rgb[..., 0] + 0.5 * rgb[..., 1] + 0.25 * rgb[..., 2]
        ref = np.hypot(ndimage.sobel(img, 0, mode="nearest"), ndimage.sobel(img, 1, mode="nearest"))
        np.testing.assert_allclose(seg_total(o), ref, rtol=1e-6, atol=1e-6)
        assert ref.max() > 0.05
        assert np.all(o["segmask"] == 0)   # threshold not reached


def test_seg_normal_and_material_indicators(oracle):
    """taa.comp:640-648: 1 - max(0, min of the 4-neighbour normal dots); :665-670: 1 where a
    4-neighbour has another material id."""
    rng = np.random.default_rng(81)
    H, W = 36, 50
    c = seg_case(H, W, rng, taastar.RTFLAG["NRM"], nrm=1.0)
    o = run_seg(oracle, c)
    z, w = c["uvn"][..., 2].astype(np.float64), c["uvn"][..., 3].astype(np.float64)
    n = np.stack([np.cos(z) * np.cos(w), np.sin(z) * np.cos(w), np.sin(w)], -1)
    pad = np.pad(n, ((1, 1), (1, 1), (0, 0)), mode="edge")
    dots = [np.sum(n * pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W], -1) for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1))]
    ref = 1.0 - np.maximum(0.0, np.min(dots, 0))
    np.testing.assert_allclose(seg_total(o), ref, atol=1e-6)
    assert ref.max() > 0.1
    c = seg_case(H, W, rng, taastar.RTFLAG["MID"], mid=1.0)
    o = run_seg(oracle, c)
    m = np.pad(c["mat"], 1, mode="edge")
    edge = np.zeros((H, W), bool)
    for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        edge |= m[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] != c["mat"]
    np.testing.assert_array_equal(seg_total(o), edge.astype(np.float64))
    assert edge.any() and not edge.all()


def test_seg_threshold_rejects_and_marks_for_ray_tracing(oracle):
    """total >= threshold -> category 2 with the history count in bits 16-31; such texels take
    mRejectionAlpha (taa.comp:778-781); debug mode 8 shows the category (:932-938)."""
    rng = np.random.default_rng(82)
    c = seg_case(32, 48, rng, taastar.RTFLAG["MID"] | taastar.RTFLAG["CNT"], mid=0.25, thresh=0.25)
    c["p"].mDebugMode, c["p"].mDebugScale = 8, 1.0
    for i in range(4):
        c["p"].mDebugMask[i] = 1.0
    o = run_seg(oracle, c)
    edge = seg_total(o) >= 0.25
    np.testing.assert_array_equal(o["segmask"], np.where(edge, 2 | (8 << 16), 0).astype(np.uint32))
    flags = o["inspect"][..., INSPECT["FLAGS"]]
    assert np.array_equal((flags & taastar.FLAG_REJECTED) != 0, edge)
    blend = o["inspect"][..., INSPECT["BLEND"]].copy().view(f32)
    assert np.all(blend[edge] == 1.0) and np.all(blend[~edge] == f32(0.1))
    np.testing.assert_array_equal(o["debug"][edge], np.broadcast_to([1.0, 1.0, 0.0, 0.0], (edge.sum(), 4)))
    np.testing.assert_array_equal(o["debug"][~edge], np.broadcast_to([0.0, 0.0, 1.0, 0.0], ((~edge).sum(), 4)))


def test_seg_decision_paths(oracle):
    """The early decisions of calc_segmentation_value: the fixed 100-texel border (category 1),
    ray-trace-all (2), history outside the screen (1), disocclusion by a moving object at the
    reprojected texel (2), and the history count kept from the previous mask (2, count - 1)."""
    R = taastar.RTFLAG
    rng = np.random.default_rng(83)
    H, W = 220, 240
    c = seg_case(H, W, rng, R["FXD"])
    o = run_seg(oracle, c)
    yy, xx = np.mgrid[0:H, 0:W]
    border = (xx < 100) | (yy < 100) | (xx >= W - 100) | (yy >= H - 100)
    np.testing.assert_array_equal(o["segmask"], border.astype(np.uint32))
    c = seg_case(24, 40, rng, R["ALL"] | R["OUT"])
    assert np.all(run_seg(oracle, c)["segmask"] == 2)
    # a velocity of 6 texels to the right: the first 6 columns' history lies outside -> 1
    c = seg_case(24, 40, rng, R["OUT"])
    c["vel"][..., 0] = 6.0 / 40
    m = run_seg(oracle, c)["segmask"]
    assert np.all(m[:, :6] == 1) and np.all(m[:, 6:] == 0)
    # disocclusion: the previous frame's texel under the history position belongs to a moving
    # object (bit 31) with another id; static previous ids never trigger
    c = seg_case(24, 40, rng, R["DIS"] | R["CNT"])
    c["vel"][..., 1] = -3.0 / 24                      # history 3 rows below
    prev = c["mat"] & np.uint32(0x7FFFFFFF)            # no moving objects ...
    prev[10:14, :] = np.uint32(0x80000000 | 77)         # ... but these rows
    prev[16:18, :] = np.uint32(78)
    c["prev_mat"] = prev
    m = run_seg(oracle, c)["segmask"]
    hit = np.zeros((24, 40), bool)
    hit[7:11, :] = True                               # y + 3 in rows 10..13 (rows 21..23 map outside: no test)
    np.testing.assert_array_equal(m, np.where(hit, 2 | (8 << 16), 0).astype(np.uint32))
    # history count: the previous mask's count decrements while the texel is not marked again
    c = seg_case(24, 40, rng, R["CNT"])
    c["prev_mask"][:, :10] = 2 | (3 << 16)
    c["prev_mask"][:, 10:20] = 2 | (1 << 16)
    m = run_seg(oracle, c)["segmask"]
    assert np.all(m[:, :10] == 2 | (2 << 16)) and np.all(m[:, 10:20] == 2) and np.all(m[:, 20:] == 0)


def seg_case_up(Hl, Wl, H, W, rng, flags, **weights):
    """seg_case at the input size Hl x Wl with an output-sized (H x W) history and previous mask."""
    c = seg_case(Hl, Wl, rng, flags, **weights)
    c["hist"] = np.concatenate([rng.uniform(0.05, 2.0, (H, W, 3)), np.zeros((H, W, 1))], -1)
    c["prev_mask"] = np.zeros((H, W), np.uint32)
    return c


def test_seg_upsampled_mask_at_output_coordinates(oracle):
    """An input smaller than the output: calc_segmentation_value still runs at the output texel
    (taa.comp:775) -- the fixed border against the input size (textureSize_loRes, :595), the
    neighbours clamped to the input image (CLAMP_TO_TEX, :42), the centre's material id and normal
    read at the output coordinates, zero outside the input image (robustImageAccess2's value for
    that undefined read: material 0, normal of angles (0, 0) = (1, 0, 0)), the previous mask at the
    output texel."""
    R = taastar.RTFLAG
    rng = np.random.default_rng(86)
    Hl, Wl, H, W = 220, 240, 330, 360
    m = run_seg(oracle, seg_case_up(Hl, Wl, H, W, rng, R["FXD"]))["segmask"]
    yy, xx = np.mgrid[0:H, 0:W]
    np.testing.assert_array_equal(m, ((xx < 100) | (yy < 100) | (xx >= Wl - 100) | (yy >= Hl - 100)).astype(np.uint32))

    Hl, Wl, H, W = 24, 40, 48, 80
    yy, xx = np.mgrid[0:H, 0:W]
    inside = (xx < Wl) & (yy < Hl)
    cx = lambda d: np.clip(xx + d, 0, Wl - 1)
    cy = lambda d: np.clip(yy + d, 0, Hl - 1)
    # material edges: own id (0 outside the input) against the clamped 4-neighbours
    c = seg_case_up(Hl, Wl, H, W, rng, R["MID"], mid=1.0)
    mat = c["mat"]
    own = np.where(inside, mat[np.minimum(yy, Hl - 1), np.minimum(xx, Wl - 1)], 0)
    edge = np.zeros((H, W), bool)
    for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        edge |= mat[cy(dy), cx(dx)] != own
    np.testing.assert_array_equal(seg_total(run_seg(oracle, c)), edge.astype(np.float64))
    assert edge[inside].any() and not edge[inside].all() and edge[~inside].all()
    # normal dots: the centre's normal (1, 0, 0) outside the input
    c = seg_case_up(Hl, Wl, H, W, rng, R["NRM"], nrm=1.0)
    z, w = c["uvn"][..., 2].astype(np.float64), c["uvn"][..., 3].astype(np.float64)
    n = np.stack([np.cos(z) * np.cos(w), np.sin(z) * np.cos(w), np.sin(w)], -1)
    nc = np.where(inside[..., None], n[np.minimum(yy, Hl - 1), np.minimum(xx, Wl - 1)], np.array([1.0, 0.0, 0.0]))
    dots = [np.sum(nc * n[cy(dy), cx(dx)], -1) for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1))]
    np.testing.assert_allclose(seg_total(run_seg(oracle, c)), 1.0 - np.maximum(0.0, np.min(dots, 0)), atol=1e-6)
    # history count from the output-sized previous mask
    c = seg_case_up(Hl, Wl, H, W, rng, R["CNT"])
    c["prev_mask"][:, 50:] = 2 | (3 << 16)
    m = run_seg(oracle, c)["segmask"]
    assert np.all(m[:, 50:] == 2 | (2 << 16)) and np.all(m[:, :50] == 0)


def test_seg_passthrough_leaves_the_mask(oracle):
    rng = np.random.default_rng(84)
    c = seg_case(16, 20, rng, taastar.RTFLAG["ALL"])
    c["p"].mPassThrough = 1
    init = np.full((16, 20), 12345, np.uint32)
    assert np.all(run_seg(oracle, c, mask_init=init)["segmask"] == 12345)


def test_seg_fp32_tracks_fp64(oracle):
    rng = np.random.default_rng(85)
    R = taastar.RTFLAG
    c = seg_case(48, 64, rng, R["NRM"] | R["DPT"] | R["MID"] | R["LUM"], nrm=0.5, dpt=0.015, mid=0.25, lum=0.5,
                 thresh=1e30)
    a = seg_total(run_seg(oracle, c, dtype=np.float64))
    b = seg_total(run_seg(oracle, c, dtype=np.float32))
    np.testing.assert_allclose(b, a, rtol=1e-5, atol=1e-5)


# ---------------------------------------------------------------------------------------------
# edge-AA fallback on category-1 texels

def luma64(rgb):
    rgb = rgb.astype(np.float64)
    return 0.299 * rgb[..., 0] + 0.587 * rgb[..., 1] + 0.114 * rgb[..., 2]


def test_edge_aa_copies_outside_category_1_and_flat_regions(oracle):
    """Texels of categories 0 and 2, and category-1 texels without enough contrast, keep their rgb;
    every texel's alpha is its luma (antialias_fxaa_prepare.comp:21-23, antialias_fxaa.comp:55-77)."""
    rng = np.random.default_rng(90)
    H, W = 20, 30
    src = np.concatenate([rng.uniform(0, 1, (H, W, 3)), np.zeros((H, W, 1))], -1)
    mask = rng.choice(np.array([0, 2, 2 | (5 << 16)], np.uint32), (H, W))
    o = oracle.edge_aa(src, mask, dtype=np.float64)
    np.testing.assert_array_equal(o[..., :3], src[..., :3])
    np.testing.assert_allclose(o[..., 3], luma64(src), rtol=1e-15)
    flat = np.concatenate([np.full((H, W, 3), 0.4), np.zeros((H, W, 1))], -1)
    flat[5, 7, :3] = 0.42                       # contrast 0.02 < max(0.0833, 0.116 * 0.42)
    o = oracle.edge_aa(flat, np.ones((H, W), np.uint32), dtype=np.float64)
    np.testing.assert_array_equal(o[..., :3], flat[..., :3])


def test_edge_aa_step_edge_closed_form(oracle):
    """A vertical step edge (luma 0 | 1), all texels in category 1: the two columns at the edge
    move towards each other by subpix / 6 (weighted mean 1/3 or 2/3 of the range), all other
    texels stay; with mirroring and transposition the result mirrors and transposes."""
    H, W, e = 12, 16, 7
    img = np.zeros((H, W, 4))
    img[:, e + 1:, :3] = 1.0
    mask = np.ones((H, W), np.uint32)
    for subpix in (0.75, 1.0, 0.3):
        o = oracle.edge_aa(img, mask, subpix=subpix, dtype=np.float64)
        ref = img[..., :3].copy()
        ref[:, e] = subpix / 6
        ref[:, e + 1] = 1 - subpix / 6
        np.testing.assert_allclose(o[..., :3], ref, atol=1e-12)
        om = oracle.edge_aa(img[:, ::-1], mask, subpix=subpix, dtype=np.float64)
        np.testing.assert_allclose(om, o[:, ::-1], atol=1e-12)
        ot = oracle.edge_aa(np.ascontiguousarray(img.transpose(1, 0, 2)), mask.T.copy(), subpix=subpix, dtype=np.float64)
        np.testing.assert_allclose(ot, o.transpose(1, 0, 2), atol=1e-12)
    # only masked texels change: mask the edge's left column only
    m2 = np.zeros((H, W), np.uint32)
    m2[:, e] = 1
    o = oracle.edge_aa(img, m2, dtype=np.float64)
    assert np.allclose(o[:, e, :3], 0.125) and np.array_equal(o[:, e + 1, :3], img[:, e + 1, :3])


def test_edge_aa_thresholds(oracle):
    """A step of contrast c on a background b: filtered iff c >= max(thr_min, thr * (b + c))."""
    H, W, e = 8, 10, 4
    for b, c, expect in ((0.5, 0.05, False), (0.5, 0.09, True), (1.0, 0.12, False), (1.0, 0.14, True),
                         (0.01, 0.08, False), (0.01, 0.09, True)):
        img = np.full((H, W, 4), b)
        img[:, e + 1:, :3] = b + c
        o = oracle.edge_aa(img, np.ones((H, W), np.uint32), thr=0.116, thr_min=0.0833, dtype=np.float64)
        changed = not np.array_equal(o[..., :3], img[..., :3])
        assert changed == expect, (b, c)


def test_edge_aa_fp32_tracks_fp64(oracle):
    rng = np.random.default_rng(91)
    src = np.concatenate([rng.uniform(0, 1, (40, 50, 3)), np.ones((40, 50, 1))], -1).astype(f32)
    mask = rng.integers(0, 3, (40, 50)).astype(np.uint32)
    a = oracle.edge_aa(src, mask, dtype=np.float64)
    b = oracle.edge_aa(src, mask, dtype=np.float32)
    np.testing.assert_allclose(b, a, atol=2e-6)


# ---------------------------------------------------------------------------------------------
# branches pinned by closed forms: slow clip, velocity-based alpha, reduce-blend-near-clamp,
# dynamic anti-ghosting, debug modes 0-7, and CasFilter's green weight and approximate reciprocal

def np_bilinear(img, hu, hv):
    """Clamp-to-edge bilinear sample of img [H, W, C] at uv positions hu, hv [H, W] (numpy, FP64)."""
    H, W = img.shape[:2]
    fx, fy = hu * W - 0.5, hv * H - 0.5
    x0, y0 = np.floor(fx).astype(int), np.floor(fy).astype(int)
    tx, ty = (fx - x0)[..., None], (fy - y0)[..., None]
    xa, xb = np.clip(x0, 0, W - 1), np.clip(x0 + 1, 0, W - 1)
    ya, yb = np.clip(y0, 0, H - 1), np.clip(y0 + 1, 0, H - 1)
    top = img[ya, xa] * (1 - tx) + img[ya, xb] * tx
    bot = img[yb, xa] * (1 - tx) + img[yb, xb] * tx
    return top * (1 - ty) + bot * ty


def slab_clip(p, q, mn, mx):
    """Where the segment from p (inside the box [mn, mx]) towards q leaves the box: p + t (q - p)
    with t the smallest slab parameter, capped at 1 (the textbook ray / box slab test)."""
    d = q - p
    with np.errstate(divide="ignore", invalid="ignore"):
        t_hi = np.where(d > 0, (mx - p) / d, np.inf)
        t_lo = np.where(d < 0, (mn - p) / d, np.inf)
    t = np.minimum(1.0, np.minimum(t_hi, t_lo).min(-1))
    return p + t[..., None] * d, t


def blend_of(o):
    return o["inspect"][..., INSPECT["BLEND"]].copy().view(f32).astype(np.float64)


@pytest.mark.parametrize("variance", [0, 1])
def test_slow_clip_is_the_slab_intersection(oracle, variance):
    """clipAabbSlow (taa.comp:348-369, mode 3): the history moves along the segment from the
    clip-towards colour (the centre tap for the min/max box, the mean for the variance box, both
    inside the box) to the history sample and stops where that segment leaves the box; a history
    inside the box is unchanged."""
    rng = np.random.default_rng(92)
    H, W = 8, 16  # powers of two: zero velocity samples the history texel exactly
    col, dep, vel, _ = frame(H, W, rng)
    hist = np.concatenate([rng.uniform(-3, 5, (H, W, 3)), np.zeros((H, W, 1))], -1)
    hist[:3, :4, :3] = col[:3, :4, :3] * 0.999 + 0.0005  # a few inside the box (near the centre tap)
    p = params(mColorClampingOrClipping=3, mInterpolationMode=0, mVarianceClipping=variance)
    o = run(oracle, uniforms(p), col, dep, vel, hist)
    w = windows(col[..., :3].astype(np.float64))
    if variance:
        mean, std = w.mean(0), w.std(0)
        towards, mn, mx = mean, mean - std, mean + std
    else:
        towards, mn, mx = col[..., :3].astype(np.float64), w.min(0), w.max(0)
    expect, t = slab_clip(towards, hist[..., :3], mn, mx)
    got = o["history"][..., :3]
    assert np.abs(got - expect).max() < 1e-6
    inside = t >= 1.0
    assert 0 < inside.sum() < inside.size and np.abs(got[inside] - hist[..., :3][inside]).max() < 1e-12
    # outside: on a face of the box, inside every slab
    out = ~inside
    on_face = (np.abs(got - mn) < 1e-6) | (np.abs(got - mx) < 1e-6)
    assert np.all(on_face[out].any(-1))
    assert np.all(got[out] >= mn[out] - 1e-6) and np.all(got[out] <= mx[out] + 1e-6)


def test_velocity_based_alpha(oracle):
    """taa.comp:855-859: no motion leaves alpha unchanged; once pixelSpeed * mVelBasedAlphaFactor
    reaches 1 alpha is max(alpha, mVelBasedAlphaMax); in between it moves linearly towards that
    maximum and never decreases."""
    rng = np.random.default_rng(93)
    H, W = 8, 16
    col, dep, _, hist = frame(H, W, rng)
    vmax, fac = 0.625, 4.0
    for a in (0.125, 0.875):
        for speed, frac in ((0.0, 0.0), (0.0625, 0.25), (0.125, 0.5), (0.25, 1.0), (0.5, 1.0)):
            vel = np.zeros((H, W, 4), f32)
            vel[..., 0], vel[..., 1], vel[..., 3] = speed * 0.6, -speed * 0.8, 1.0  # |v| = speed
            p = params(mAlpha=a, mVelBasedAlpha=1, mVelBasedAlphaMax=vmax, mVelBasedAlphaFactor=fac)
            alpha = blend_of(run(oracle, uniforms(p), col, dep, vel, hist))
            expect = max(a, a + (vmax - a) * frac)
            assert np.abs(alpha - expect).max() < 1e-6, (a, speed)
    # off: alpha is mAlpha whatever the speed
    vel[..., 0] = 0.3
    alpha = blend_of(run(oracle, uniforms(params(mAlpha=0.125, mVelBasedAlphaMax=vmax)), col, dep, vel, hist))
    assert np.all(alpha == 0.125)


def test_reduce_blend_near_clamp(oracle):
    """taa.comp:873-894: alpha scales by clamp(4 d, 0, 1), d = 2 |min(lh - lmin, lmax - lh)| /
    (lmax - lmin) on the luminance of the box and of the unrectified history: a history at the
    box centre keeps alpha, one on a face gives alpha = 0, one an eighth of the box inside (or
    outside) a face halves it; a box under 0.001 in luminance keeps alpha."""
    rng = np.random.default_rng(94)
    H, W = 8, 16
    g = rng.uniform(0.1, 0.9, (H, W)).astype(f32)
    g[1:6, 9:15] = f32(0.4)  # flat: the box of the texels inside has zero extent
    col = np.stack([g, g, g, np.ones_like(g)], -1)  # grey: luminance = the value
    dep = np.full((H, W), 0.5, f32)
    vel = np.zeros((H, W, 4), f32)
    vel[..., 3] = 1
    w = windows(g.astype(np.float64))
    lmin, lmax = w.min(0), w.max(0)
    choices = np.array([0.5, 0.0, 1.0, 1 / 16, 1 / 8, -0.5, 1 + 1 / 16, 0.3])
    factor = np.array([1.0, 0.0, 0.0, 0.5, 1.0, 1.0, 0.5, 1.0])
    k = rng.integers(0, len(choices), (H, W))
    hv = lmin + choices[k] * (lmax - lmin)
    hist = np.stack([hv, hv, hv, np.zeros_like(hv)], -1)
    for ycocg in (0, 1):
        # clamping on: the factor reads the history before rectification (taa.comp:878)
        p = params(mAlpha=0.5, mReduceBlendNearClamp=1, mInterpolationMode=0, mUseYCoCg=ycocg, mColorClampingOrClipping=1)
        alpha = blend_of(run(oracle, uniforms(p), col, dep, vel, hist))
        flat = (lmax - lmin) < 0.001
        assert flat.sum() == 12
        expect = np.where(flat, 0.5, 0.5 * factor[k])
        assert np.abs(alpha - expect).max() < 1e-6, ycocg
    # off: alpha unchanged
    alpha = blend_of(run(oracle, uniforms(params(mAlpha=0.5, mInterpolationMode=0)), col, dep, vel, hist))
    assert np.all(alpha == 0.5)


def test_dynamic_anti_ghosting(oracle):
    """taa.comp:796-811: a texel moves when |v.x| or |v.y| exceeds 1e-5 and v.w >= 0.5. Where none
    of the five cross taps (clamped to the edge) moves but the filtered history's alpha is > 0, the
    history is rejected (alpha = mRejectionAlpha); the new history alpha is 1 exactly where the
    centre moves, 0 elsewhere."""
    rng = np.random.default_rng(95)
    H, W = 16, 16
    col, dep, _, _ = frame(H, W, rng)
    vel = np.zeros((H, W, 4), f32)
    kind = rng.integers(0, 6, (H, W))
    # 0 still; 1 mover shifted by whole texels; 2 mover flag but no motion; 3 motion, w = 0.3;
    # 4 motion at w = 0.5 exactly; 5 mover with a motion of 5e-6 (below the threshold)
    sx = rng.integers(-2, 3, (H, W)).astype(f32) / f32(W)
    vel[..., 0] = np.select([kind == 1, kind == 3, kind == 4, kind == 5], [sx, sx + f32(1 / 16), f32(-1 / 16), f32(5e-6)], 0)
    vel[..., 1] = np.where(kind == 1, f32(1 / 16), 0)
    vel[..., 3] = np.select([kind == 1, kind == 2, kind == 3, kind == 4, kind == 5], [1.0, 1.0, 0.3, 0.5, 1.0], 0.0)
    hist = np.concatenate([rng.uniform(0.1, 1, (H, W, 3)), rng.integers(0, 2, (H, W, 1)).astype(np.float64)], -1)
    p = params(mDynamicAntiGhosting=1, mAlpha=0.125, mRejectionAlpha=0.75, mInterpolationMode=0)
    o = run(oracle, uniforms(p), col, dep, vel, hist)
    v64 = vel.astype(np.float64)
    mov = ((np.abs(v64[..., 0]) > 1e-5) | (np.abs(v64[..., 1]) > 1e-5)) & (v64[..., 3] >= 0.5)
    assert np.array_equal(mov, np.isin(kind, (1, 4)) & ((v64[..., 0] != 0) | (v64[..., 1] != 0)))
    mp = np.pad(mov, 1, mode="edge")
    movement = mov | mp[1:-1, :-2] | mp[1:-1, 2:] | mp[:-2, 1:-1] | mp[2:, 1:-1]
    x = (np.arange(W)[None, :] + 0.5) / W
    y = (np.arange(H)[:, None] + 0.5) / H
    ha = np_bilinear(hist[..., 3:], x - v64[..., 0], y - v64[..., 1])[..., 0]
    rejected = ~movement & (ha > 0)
    assert 0 < rejected.sum() < rejected.size and (~movement & (ha == 0)).any()
    flags = o["inspect"][..., INSPECT["FLAGS"]]
    assert np.array_equal((flags & taastar.FLAG_GHOST_REJECT) != 0, rejected)
    assert np.array_equal(blend_of(o), np.where(rejected, 0.75, 0.125))
    assert np.array_equal(o["history"][..., 3], mov.astype(np.float64))
    # off: nothing rejected, history alpha 0
    o = run(oracle, uniforms(params(mAlpha=0.125, mInterpolationMode=0)), col, dep, vel, hist)
    assert not np.any(o["inspect"][..., INSPECT["FLAGS"]] & taastar.FLAG_GHOST_REJECT)
    assert np.all(o["history"][..., 3] == 0)


def test_debug_modes_closed_forms(oracle):
    """taa.comp:912-942: 0 the box extent per channel, 1 its volume, 2 (rejected, length of the
    rectification), 3 alpha (without beta), 4 the velocity texel, 5 pixelSpeed, 6 the screen, 7 the
    history output; then * mDebugScale * mDebugMask and, with mDebugCenter, * 0.5 + 0.5."""
    rng = np.random.default_rng(96)
    H, W = 8, 16
    col, dep, _, hist = frame(H, W, rng)
    hist[..., :3] = rng.uniform(-0.5, 2.5, (H, W, 3))
    vel = np.zeros((H, W, 4), f32)
    vel[..., 0], vel[..., 2], vel[..., 3] = f32(2 / 16), f32(0.25), f32(1.0)  # two texels left
    c = col[..., :3].astype(np.float64)
    w = windows(c)
    mn, mx = w.min(0), w.max(0)
    xs = np.clip(np.arange(W) - 2, 0, W - 1)
    h = hist[:, xs, :3]  # bilinear at the texel centre two to the left, clamped
    rej = np.broadcast_to((np.arange(W) + 0.5) / W - 2 / 16 < 0, (H, W))
    hc = np.clip(h, mn, mx)
    a = np.where(rej, 0.75, 0.25)[..., None]
    out = hc * (1 - a) + c * a
    ones, zeros = np.ones((H, W)), np.zeros((H, W))
    expect = {
        0: np.concatenate([mx - mn, zeros[..., None]], -1),
        1: np.stack([np.prod(mx - mn, -1)] * 3 + [zeros], -1),
        2: np.stack([rej.astype(np.float64), np.linalg.norm(hc - h, axis=-1), zeros, zeros], -1),
        3: np.stack([a[..., 0]] * 3 + [zeros], -1),
        4: vel.astype(np.float64),
        5: np.stack([ones * 0.125, zeros, zeros, zeros], -1),
        6: np.concatenate([out, ones[..., None]], -1),
        7: np.concatenate([out, zeros[..., None]], -1),
    }
    for mode, e in expect.items():
        for scale, mask, center in ((1.0, (1, 1, 1, 1), 0), (2.0, (1, 0.5, 0, 1), 1)):
            p = params(mColorClampingOrClipping=1, mInterpolationMode=0, mAlpha=0.25, mRejectOutside=1,
                       mRejectionAlpha=0.75, mDebugMode=mode, mDebugScale=scale, mDebugCenter=center)
            for i in range(4):
                p.mDebugMask[i] = mask[i]
            g = run(oracle, uniforms(p), col, dep, vel, hist)["debug"]
            ref = e * scale * np.asarray(mask, np.float64)
            if center:
                ref = ref * 0.5 + 0.5
            assert np.abs(g - ref).max() < 1e-9, (mode, scale, center)


def test_cas_weight_is_green_for_all_channels(oracle):
    """ffx_cas.h:508-522: one weight, from green's soft minimum and maximum, for all channels.
    (a) Green flat above 1 has amplitude 0: a red step is not sharpened, out_r / e_r is the same
    at every texel. (b) Steps of different contrast in all channels: solving out_c * D - w * S_c =
    e_c (S_c the four-neighbour sum) for (D, w) from red and green predicts blue, and w matches
    green's amplitude, not red's."""
    c0, c1 = oracle.cas_setup(1.0, 32, 8)
    img = np.full((8, 32, 4), 1.5)
    img[:, :, 0] = 0.2
    img[:, 16:, 0] = 0.7
    out = oracle.cas(img, c1[0], dtype=np.float64)
    ratio = out[..., 0] / img[..., 0]
    assert np.ptp(ratio) < 1e-12 and abs(ratio[0, 0] - 1) < 4e-3
    # (b) red 0.05 -> 0.15 (soft amplitude sqrt(1/3)), green 0.4 -> 0.6 (sqrt(2/3)), blue 0.3 -> 0.5
    c0, c1 = oracle.cas_setup(0.5, 32, 8)
    img = np.ones((8, 32, 4))
    for ch, (lo, hi) in enumerate(((0.05, 0.15), (0.4, 0.6), (0.3, 0.5))):
        img[:, :16, ch], img[:, 16:, ch] = lo, hi
    out = oracle.cas(img, c1[0], dtype=np.float64)
    y = 4
    peak = float(u2f(c1[0]))
    for x in (15, 16):
        e = img[y, x, :3]
        S = img[y - 1, x, :3] + img[y + 1, x, :3] + img[y, x - 1, :3] + img[y, x + 1, :3]
        A = np.array([[out[y, x, 0], -S[0]], [out[y, x, 1], -S[1]]])
        D, wgt = np.linalg.solve(A, e[:2])
        assert 0 < out[y, x, :3].min() and out[y, x, :3].max() < 1
        assert abs(out[y, x, 2] * D - wgt * S[2] - e[2]) < 1e-9
        amp_g = np.sqrt(min(0.4, 1 - 0.6) / 0.6)
        amp_r = np.sqrt(min(0.05, 1 - 0.15) / 0.15)
        assert abs(wgt - peak * amp_g) < 0.1 * abs(peak * amp_g)   # the integer approximations: within ~8%
        assert abs(wgt - peak * amp_r) > 0.15 * abs(peak * amp_g)


def test_cas_reciprocal_is_the_approximate_newton_step(oracle):
    """The final normalisation uses APrxMedRcpF1 (ffx_a.h:1457): one Newton step from an integer
    estimate, m(x) = b (2 - b x), so x m(x) = 1 - (1 - b x)^2 never exceeds 1. On a flat image of
    value v the filter returns v x m(x) (x = 1 + 4 w): at or below v, within 4e-3, and strictly
    below v for almost every v -- an exact reciprocal would return v itself."""
    c0, c1 = oracle.cas_setup(0.5, 16, 8)
    below = 0
    vs = np.linspace(0.01, 0.99, 99)
    for v in vs:
        out = oracle.cas(np.full((8, 16, 4), v), c1[0], dtype=np.float64)[..., :3]
        assert np.ptp(out) == 0
        r = out[0, 0, 0] / v
        assert 1 - 4e-3 < r <= 1, v
        below += r < 1 - 1e-7
    assert below >= 0.9 * len(vs)
