"""The resolve with the segmentation mask (taa_resolve_frame_segmented, mRayTraceAugment;
taa.comp:589-702, 773-784, 932-938, 959) on the GPU against the FP32 transcription.

Bar: as the FP32 path — the mask words bit-exact, the weighted indicator sum bit-exact, texel
indices, flags and box bit-exact, history and screen within 1e-5 relative (observed 0).
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import assert_parity, round_storage, to_f16  # noqa: E402
from taastar import (FMT_RGBA16F, FMT_RGBA32F, INSPECT, RTFLAG, VEL_RG16F, VEL_RG32F, SegmentationImages,  # noqa: E402
                     check, lib)
from taastar.device import FrameBuffers  # noqa: E402
from taastar.workloads import (cas_for, decode, encode_color, encode_velocity, params_c2, scene_for,  # noqa: E402
                               segmentation_inputs, uniforms_for, velocity_as_read, workload)


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


class SegSequence:
    """Frames of one seeded scene with material ids and normals, resolved with the segmentation
    mask on the GPU and by the FP32 oracle side by side."""

    def __init__(self, w, width, height, cas=True, params=None, split=None):
        self.w, self.W, self.H = w, width, height
        self.fmt, self.vfmt = w["fmt"], w["vel"]
        self.scene = scene_for(w, width=width, height=height, pan_mode=1, pan=3.0)
        self.cas = cas_for(w, width, height) if cas else None
        self.params, self.split = params, split
        self.bufs = FrameBuffers(width, height, self.fmt, self.vfmt, debug=True, inspect=True)
        z = lambda *s, dtype: torch.zeros(*s, dtype=dtype, device="cuda")
        self.mat = [z(height, width, dtype=torch.int32), z(height, width, dtype=torch.int32)]
        self.uvn = z(height, width, 4, dtype=torch.float32)
        self.mask = [z(height, width, dtype=torch.int32), z(height, width, dtype=torch.int32)]
        self.hist_o = np.zeros((height, width, 4), np.float32)
        self.prev_depth = np.zeros((height, width), np.float32)
        self.prev_mat_o = np.zeros((height, width), np.uint32)
        self.mask_o = np.zeros((height, width), np.uint32)
        self.k = 0

    def step(self, oracle):
        k = self.k
        col, dep, vel4, jit = self.scene.render(k)
        col_s, vel_s = encode_color(col, self.fmt), encode_velocity(vel4, self.vfmt)
        mat, uvn = segmentation_inputs(dep)
        self.bufs.upload(col_s, dep, vel_s)
        self.mat[k & 1].copy_(torch.from_numpy(mat.view(np.int32)))
        self.uvn.copy_(torch.from_numpy(uvn))
        U = uniforms_for(self.w, k, jit, params=self.params)
        if self.split is not None:
            U.splitScreen, U.splitX = 1, self.split
            U.param[1].mRayTraceAugment = 0
        sg = SegmentationImages()
        sg.material, sg.prev_material = self.mat[k & 1].data_ptr(), self.mat[(k + 1) & 1].data_ptr()
        sg.uv_normal, sg.prev_segmask, sg.segmask = self.uvn.data_ptr(), self.mask[(k + 1) & 1].data_ptr(), \
            self.mask[k & 1].data_ptr()
        img = self.bufs.images()
        check(lib().taa_resolve_frame_segmented(C.byref(U), C.byref(img), C.byref(sg),
                                                C.byref(self.cas) if self.cas else None, 0, None))
        self.bufs.k += 1
        torch.cuda.synchronize()
        o = oracle.resolve_seg(U, decode(col_s), dep, velocity_as_read(vel_s), self.hist_o, self.prev_depth, mat,
                               self.prev_mat_o, uvn, self.mask_o)
        o_hist = round_storage(o["history"], self.fmt)
        o_scr = round_storage(o["screen"], self.fmt)
        if self.cas is not None:
            o_scr = round_storage(oracle.cas(o_scr, self.cas.const1[0]), self.fmt)
        out = dict(gpu_history=decode(self.bufs.hist_in.cpu().numpy()), gpu_screen=decode(self.bufs.screen.cpu().numpy()),
                   gpu_inspect=self.bufs.inspect.cpu().numpy(),
                   gpu_debug=self.bufs.debug.cpu().numpy().astype(np.float32),
                   gpu_mask=self.mask[k & 1].cpu().numpy().view(np.uint32),
                   oracle_history=o_hist, oracle_screen=o_scr, oracle_inspect=o["inspect"],
                   oracle_debug=to_f16(o["debug"]).astype(np.float32), oracle_mask=o["segmask"])
        self.hist_o, self.prev_depth = o_hist, dep
        self.prev_mat_o = mat
        self.mask_o = o["segmask"]
        self.k += 1
        return out


def assert_seg_parity(out, split=None):
    assert_parity(out)
    gi, oi = out["gpu_inspect"], out["oracle_inspect"]
    st = INSPECT["SEG_TOTAL"]
    bad = np.count_nonzero(gi[..., st] != oi[..., st])
    assert bad == 0, f"segmentation total: {bad} texels differ"
    gm, om = out["gpu_mask"], out["oracle_mask"]
    if split is not None:  # the right half has no mask: untouched zeros on the GPU, oracle zeros
        gm, om = gm[:, :split + 1], om[:, :split + 1]
    bad = np.count_nonzero(gm != om)
    assert bad == 0, f"segmentation mask: {bad} texels differ"
    assert np.array_equal(out["gpu_debug"], out["oracle_debug"])


def seg_params(flags=None, **kw):
    p = params_c2()
    p.mRayTraceAugment = 1
    p.mRayTraceHistoryCount = 8
    if flags is not None:
        p.mRayTraceAugmentFlags = flags
    p.mDebugMode = 8
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("fmt", [FMT_RGBA32F, FMT_RGBA16F])
def test_segmentation_default_flags(oracle, fmt):
    """The reference's default flags (all indicators, outside, disocclusion, history count) on a
    panning scene: categories 0, 1 and 2 all occur, with the fused CAS."""
    print(f"case seg default fmt={fmt}", flush=True)
    w = workload("c2", fmt=fmt, vel=VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F)
    seq = SegSequence(w, 320, 180, params=seg_params())
    cats = set()
    for _ in range(5):
        out = seq.step(oracle)
        assert_seg_parity(out)
        cats |= set(np.unique(out["gpu_mask"] & 3).tolist())
    assert cats == {0, 1, 2}, cats


VARIANTS = {
    "border_fxd": dict(flags=RTFLAG["FXD"] | RTFLAG["NRM"]),
    "all": dict(flags=RTFLAG["ALL"]),
    "depth_lum_only": dict(flags=RTFLAG["DPT"] | RTFLAG["LUM"], mRayTraceAugment_Thresh=0.3),
    "count_and_material": dict(flags=RTFLAG["MID"] | RTFLAG["CNT"] | RTFLAG["DIS"]),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_segmentation_flag_variants(oracle, name):
    print(f"case seg variant {name}", flush=True)
    v = dict(VARIANTS[name])
    w = workload("c2")
    seq = SegSequence(w, 256, 232, cas=False, params=seg_params(v.pop("flags"), **v))
    for _ in range(4):
        assert_seg_parity(seq.step(oracle))


@pytest.mark.parametrize("cas", [False, True])
def test_segmentation_split_screen(oracle, cas):
    """Split screen: the left parameter block has the mask, the right one not; with CAS the frame
    is sharpened after both parts are resolved."""
    print(f"case seg split cas={cas}", flush=True)
    w = workload("c2")
    seq = SegSequence(w, 200, 120, cas=cas, params=seg_params(), split=97)
    for _ in range(3):
        assert_seg_parity(seq.step(oracle), split=97)


def test_segmentation_refusals():
    from taastar._abi import TAA_ERR_INVALID_ARG as EINVAL, TAA_ERR_UNSUPPORTED as EUNSUP
    w = workload("c2")
    bufs = FrameBuffers(64, 32, w["fmt"], w["vel"])
    U = uniforms_for(w, 0, (0.0, 0.0), params=seg_params())
    img = bufs.images()
    L = lib()
    assert L.taa_resolve_frame(C.byref(U), C.byref(img), None, 0, None) == EUNSUP
    sg = SegmentationImages()
    assert L.taa_resolve_frame_segmented(C.byref(U), C.byref(img), C.byref(sg), None, 0, None) == EINVAL
    m = torch.zeros(32, 64, dtype=torch.int32, device="cuda")
    n = torch.zeros(32, 64, 4, dtype=torch.float32, device="cuda")
    sg.material = sg.prev_material = sg.prev_segmask = m.data_ptr()
    sg.uv_normal, sg.segmask = n.data_ptr(), m.data_ptr()
    assert L.taa_resolve_frame_segmented(C.byref(U), C.byref(img), C.byref(sg), None, 0, None) == EINVAL  # alias
    img.input_width, img.input_height = 32, 16
    m2 = torch.zeros(32, 64, dtype=torch.int32, device="cuda")
    sg.segmask = m2.data_ptr()
    assert L.taa_resolve_frame_segmented(C.byref(U), C.byref(img), C.byref(sg), None, 0, None) == EUNSUP  # TAAU
