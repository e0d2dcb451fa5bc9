"""The resolve workloads: parameter sets, scene settings and storage encodings.

c1  640x360,   16 frames, FP32 storage, variance clip + Catmull-Rom, simple velocity
c2  1920x1080, 64 frames, FP32 storage, + closest-depth velocity, linear-depth rejection,
    Lottes luma weighting, reject-outside, fused CAS 0.5
c3  3840x2160, 64 frames, FP16 storage (RGBA16F / RG16F / D32F), c2 feature set, sigma 2.0
c4  48 streams of mixed resolution in one persistent launch (FP32 and FP16 storage)
c5  served load: 256 sessions, Poisson arrivals, FP16 storage, HBM-budget admission
"""
import numpy as np

from . import (FMT_RGBA16F, FMT_RGBA32F, VEL_RG16F, VEL_RG32F, VEL_RGBA16F, VEL_RGBA32F, Reproject, Scene,
               Uniforms, cas_setup, default_params)

NEAR, FAR = 0.1, 100.0          # camera planes for the linear-depth test (main.cpp:3780 uses near 0.1)
LINEAR_DEPTH_THRESHOLD = 0.01   # c2/c3: reject when |linear(hd) - linear(expected)| > 0.01
FP16_COLOR_MAX = 32768.0        # scene colour clamp for FP16 storage (keeps HDR fireflies finite)

RESOLUTIONS = {"360p": (640, 360), "720p": (1280, 720), "1080p": (1920, 1080), "1440p": (2560, 1440),
               "4k": (3840, 2160)}
SERVE_MIX = (("720p", 0.40), ("1080p", 0.35), ("1440p", 0.15), ("4k", 0.10))


def params_c1():
    p = default_params()
    p.mVarianceClipping = 1
    p.mVarClipGamma = 1.0
    p.mColorClampingOrClipping = 2
    p.mInterpolationMode = 2
    p.mAlpha = 0.1
    p.mUseVelocityVectors = 2
    p.mVelocitySampleMode = 0
    p.mUseYCoCg = 1
    return p


def params_c2():
    p = params_c1()
    p.mVelocitySampleMode = 2
    p.mDepthCulling = 1
    p.mLumaWeightingLottes = 1
    p.mRejectOutside = 1
    return p


WORKLOADS = {
    "c1": dict(width=640, height=360, frames=16, polygons=200, seed=0x7AA50001, lum_sigma=1.5,
               fmt=FMT_RGBA32F, vel=VEL_RG32F, params=params_c1, cas=None, linear_depth=0.0),
    "c2": dict(width=1920, height=1080, frames=64, polygons=2000, seed=0x7AA50002, lum_sigma=1.5,
               fmt=FMT_RGBA32F, vel=VEL_RG32F, params=params_c2, cas=0.5, linear_depth=LINEAR_DEPTH_THRESHOLD),
    "c3": dict(width=3840, height=2160, frames=64, polygons=8000, seed=0x7AA50003, lum_sigma=2.0,
               fmt=FMT_RGBA16F, vel=VEL_RG16F, params=params_c2, cas=0.5, linear_depth=LINEAR_DEPTH_THRESHOLD),
}


def workload(name, **over):
    w = dict(WORKLOADS[name])
    w.update(over)
    return w


def scene_for(w, seed=None, width=None, height=None, polygons=None, pan_mode=0, pan=8.0, jitter_phase=0):
    width = width or w["width"]
    height = height or w["height"]
    polygons = w["polygons"] if polygons is None else polygons
    # keep the polygon density of the workload when a test shrinks the frame
    if width != w["width"] or height != w["height"]:
        polygons = max(1, int(round(polygons * width * height / (w["width"] * w["height"]))))
    return Scene(width=width, height=height, num_polygons=polygons, seed=w["seed"] if seed is None else seed,
                 lum_sigma=w["lum_sigma"], pan_mode=pan_mode, pan=pan, jitter_phase=jitter_phase,
                 color_max=FP16_COLOR_MAX if w["fmt"] == FMT_RGBA16F else 0.0)


def uniforms_for(w, frame, jitter_ndc, reset=None, params=None):
    U = Uniforms()
    p = params if params is not None else w["params"]()
    U.param[0] = p
    U.param[1] = p
    U.mJitterNdc[0], U.mJitterNdc[1] = jitter_ndc
    U.mCamNearPlane = NEAR
    U.mCamFarPlane = FAR
    U.pad1 = w.get("linear_depth", 0.0)
    U.mResetHistory = int(frame == 0 if reset is None else reset)
    for i in (0, 5, 10, 15):  # identity matrices (only read by matrix reprojection)
        U.mHistoryViewProjMatrix[i] = 1.0
        U.mInverseViewProjMatrix[i] = 1.0
    return U


def cas_for(w, width=None, height=None):
    if w.get("cas") is None:
        return None
    return cas_setup(w["cas"], width or w["width"], height or w["height"])


# ---- storage encodings --------------------------------------------------------------------

def encode_color(rgba, fmt):
    """float32 [H,W,4] -> storage array ([H,W,4] float32 or float16)."""
    return np.ascontiguousarray(rgba.astype(np.float16 if fmt == FMT_RGBA16F else np.float32))


def decode(arr):
    return np.ascontiguousarray(arr.astype(np.float32))


def encode_velocity(vel4, vfmt):
    if vfmt == VEL_RG32F:
        return np.ascontiguousarray(vel4[..., :2].astype(np.float32))
    if vfmt == VEL_RG16F:
        return np.ascontiguousarray(vel4[..., :2].astype(np.float16))
    if vfmt == VEL_RGBA32F:
        return np.ascontiguousarray(vel4.astype(np.float32))
    return np.ascontiguousarray(vel4.astype(np.float16))


def velocity_as_read(vel_stored):
    """What the resolve reads from a velocity image: RG formats come back as (r, g, 0, 1)."""
    v = vel_stored.astype(np.float32)
    if v.shape[-1] == 2:
        out = np.zeros(v.shape[:-1] + (4,), np.float32)
        out[..., :2] = v
        out[..., 3] = 1.0
        return out
    return np.ascontiguousarray(v)


def bytes_per_pixel(w, cas=True):
    """Algorithmic HBM bytes per output pixel: each input read once, each output written once."""
    cb = 8 if w["fmt"] == FMT_RGBA16F else 16
    vb = {VEL_RG32F: 8, VEL_RG16F: 4, VEL_RGBA32F: 16, VEL_RGBA16F: 8}[w["vel"]]
    p = w["params"]()
    reads = cb + vb + cb  # colour, velocity, history
    # the current depth: closest-depth velocity, the depth test, or matrix reprojection (mode 0, or
    # mode 1 on an RGBA velocity image whose .w can send a texel there)
    rgba = w["vel"] in (VEL_RGBA32F, VEL_RGBA16F)
    if p.mVelocitySampleMode == 2 or p.mDepthCulling or p.mUseVelocityVectors == 0 or (p.mUseVelocityVectors == 1 and rgba):
        reads += 4
    if p.mDepthCulling:
        reads += 4        # previous depth
    writes = cb + cb      # history, screen (sharpened when CAS is fused)
    return reads + writes


# ---- camera matrices for the velocity producer --------------------------------------------
# Row-major float64 math (A @ v); colmajor() gives glm's column-major float32 layout.

def perspective(fovy_deg=60.0, aspect=16.0 / 9.0, near=NEAR, far=FAR):
    """Right-handed perspective with depth 0 at the near plane and 1 at the far plane
    (glm::perspectiveRH_ZO): a point at view-space z = -d projects to NDC z in [0, 1]."""
    f = 1.0 / np.tan(np.radians(fovy_deg) / 2.0)
    P = np.zeros((4, 4))
    P[0, 0], P[1, 1] = f / aspect, f
    P[2, 2], P[2, 3] = far / (near - far), -(far * near) / (far - near)
    P[3, 2] = -1.0
    return P


def translate(x, y, z):
    T = np.eye(4)
    T[:3, 3] = (x, y, z)
    return T


def rotate_y(deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    R = np.eye(4)
    R[0, 0], R[0, 2], R[2, 0], R[2, 2] = c, s, -s, c
    return R


def jittered(P, jx, jy):
    """The jittered projection of get_jittered_projection_matrix (taa.hpp:247): NDC translated by the jitter."""
    return translate(jx, jy, 0.0) @ P


def colmajor(A):
    return np.ascontiguousarray(np.asarray(A, np.float64).T.reshape(-1).astype(np.float32))


def reproject_for(P, V_cur, V_prev, jit_cur, jit_prev, objects=None):
    """Reproject for a camera moving from V_prev to V_cur under projection P, with jitters (NDC).
    objects: [n, 4, 4] float64 previous-from-current object matrices (row-major math), or None.
    Returns (Reproject, column-major float32 object array or None); the caller points
    object_prev_from_current at that array (host memory for the oracle, device for the GPU)."""
    r = Reproject()
    Mi = np.linalg.inv(jittered(P, *jit_cur) @ V_cur)
    Mp = jittered(P, *jit_prev) @ V_prev
    r.inverse_view_proj[:] = colmajor(Mi).tolist()
    r.prev_view_proj[:] = colmajor(Mp).tolist()
    r.jitter_current_prev[:] = [jit_cur[0], jit_cur[1], jit_prev[0], jit_prev[1]]
    obj = None
    if objects is not None and len(objects):
        obj = np.ascontiguousarray(np.stack([colmajor(o) for o in objects]))
        r.n_objects = len(objects)
    return r, obj


def depth_from_view_distance(dist, P):
    """NDC depth of points at view-space distance `dist` (positive) straight through P."""
    return (P[2, 2] * -dist + P[2, 3]) / dist


def camera_case(W, H, seed=0, n_objects=0, move=(5e-4, -3e-4, 1e-3), turn_deg=0.2):
    """A frame for the velocity producer: the seeded scene's depth at W x H, a camera that moves
    by `move` and turns by `turn_deg` between frames, Halton-sized jitters, and n_objects moved
    objects covering horizontal bands of the frame (ids 1..n_objects, id n_objects + 1 beyond the
    table). The scene's depths lie 0.13-0.15 units in front of the camera, so these motions are a
    few texels at 1080p. Returns dict(depth, ids, reproject, objects)."""
    rng = np.random.default_rng(seed)
    d = scene_for(workload("c2"), seed=seed, width=W, height=H).render(1)[1]
    P = perspective(aspect=W / H)
    V_prev = translate(0.0, -1.5, -4.0)
    V_cur = rotate_y(turn_deg) @ translate(-move[0], -move[1], -move[2]) @ V_prev
    jc = (float(rng.uniform(-1, 1)) / W, float(rng.uniform(-1, 1)) / H)
    jp = (float(rng.uniform(-1, 1)) / W, float(rng.uniform(-1, 1)) / H)
    objects = [translate(*rng.uniform(-5e-4, 5e-4, 3)) @ rotate_y(float(rng.uniform(-0.02, 0.02))) for _ in range(n_objects)]
    r, obj = reproject_for(P, V_cur, V_prev, jc, jp, objects)
    ids = None
    if n_objects:
        band = np.minimum((np.arange(H) * (n_objects + 2)) // max(H, 1), n_objects + 1).astype(np.uint32)
        ids = np.ascontiguousarray(np.broadcast_to(band[:, None], (H, W)))
    return dict(depth=np.ascontiguousarray(d, np.float32), ids=ids, reproject=r, objects=obj)


# ---- segmentation-mask inputs ----------------------------------------------------------------

def segmentation_inputs(depth, moving_every=5):
    """G-buffer images of the segmentation mask for a frame of the seeded scene, whose polygons have
    distinct constant depths: a polygon's material id is its depth rank modulo 61, plus 1, with bit
    31 (moving object) on ids divisible by `moving_every`; its flat normal's angles (azimuth z in
    [0, 2 pi), elevation w in [-pi/2, pi/2)) follow from the id. Returns (material [H, W] uint32,
    uv_normal [H, W, 4] float32)."""
    H, W = depth.shape
    _, inv = np.unique(depth, return_inverse=True)
    ids = (inv.reshape(H, W) % 61 + 1).astype(np.uint32)
    mat = np.where(ids % moving_every == 0, ids | np.uint32(0x80000000), ids).astype(np.uint32)
    uvn = np.empty((H, W, 4), np.float32)
    uvn[..., 0] = ((np.arange(W) + 0.5) / W)[None, :]
    uvn[..., 1] = ((np.arange(H) + 0.5) / H)[:, None]
    uvn[..., 2] = (ids * 0.6180339887) % (2 * np.pi)
    uvn[..., 3] = (ids * 0.3819660113) % np.pi - np.pi / 2
    return np.ascontiguousarray(mat), uvn
