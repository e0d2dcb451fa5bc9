"""The resolve with the segmentation mask (taa_resolve_frame_segmented, mRayTraceAugment;
taa.comp:589-702, 773-784, 932-938, 959) on the GPU against the FP32 transcription, at one size
and upsampling (the mask evaluated at the output texel, its inputs at the render size).

Bar: as the FP32 path — the mask words bit-exact, the weighted indicator sum bit-exact, texel
indices, flags and box bit-exact, history and screen within 1e-5 relative (observed 0). In the
FP16 arithmetic: the mask, the indicator sum, the texel indices, taps and decisions bit-exact, the
sharpened screen > 55 dB against the transcription run on the GPU's previous history.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import INDEX_WORDS, assert_parity, assert_same_bits, psnr, round_storage, to_f16  # noqa: E402
from taastar import (ARITH_FP16, FLAG_RECTIFIED, FMT_RGBA16F, FMT_RGBA32F, INSPECT, RTFLAG, VEL_RG16F,  # noqa: E402
                     VEL_RG32F, SegmentationImages, check, lib)
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

    def __init__(self, w, width, height, cas=True, params=None, split=None, input_size=None, upsampling=False,
                 arith=0, shipped=False):
        self.w, self.W, self.H = w, width, height
        Wl, Hl = input_size if input_size is not None else (width, height)
        self.fmt, self.vfmt = w["fmt"], w["vel"]
        self.scene = scene_for(w, width=Wl, height=Hl, pan_mode=1, pan=3.0)
        self.cas = cas_for(w, width, height) if cas else None
        self.params, self.split, self.upsampling, self.arith = params, split, upsampling, arith
        self.bufs = FrameBuffers(width, height, self.fmt, self.vfmt, debug=True, inspect=True, input_size=(Wl, Hl))
        z = lambda *s, dtype: torch.zeros(*s, dtype=dtype, device="cuda")
        self.mat = [z(Hl, Wl, dtype=torch.int32), z(Hl, Wl, dtype=torch.int32)]
        self.uvn = z(Hl, Wl, 4, dtype=torch.float32)
        self.mask = [z(height, width, dtype=torch.int32), z(height, width, dtype=torch.int32)]
        # shipped: the same frames also through buffers without the debug and inspect images (with the
        # production feature set and CAS the one-size masked resolve then runs its production form);
        # step() asserts that both write the same history, screen and mask bits
        self.ship = FrameBuffers(width, height, self.fmt, self.vfmt, input_size=(Wl, Hl)) if shipped else None
        self.mask_s = [z(height, width, dtype=torch.int32), z(height, width, dtype=torch.int32)] if shipped else None
        self.hist_o = np.zeros((height, width, 4), np.float32)
        self.prev_depth = np.zeros((Hl, Wl), np.float32)
        self.prev_mat_o = np.zeros((Hl, Wl), np.uint32)
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
        U.mUpsampling = int(self.upsampling)
        if self.split is not None:
            U.splitScreen, U.splitX = 1, self.split
            U.param[1].mRayTraceAugment = 0
        sg = SegmentationImages()
        sg.material, sg.prev_material = self.mat[k & 1].data_ptr(), self.mat[(k + 1) & 1].data_ptr()
        sg.uv_normal, sg.prev_segmask, sg.segmask = self.uvn.data_ptr(), self.mask[(k + 1) & 1].data_ptr(), \
            self.mask[k & 1].data_ptr()
        img = self.bufs.images()
        check(lib().taa_resolve_frame_segmented(C.byref(U), C.byref(img), C.byref(sg),
                                                C.byref(self.cas) if self.cas else None, self.arith, None))
        self.bufs.k += 1
        if self.ship is not None:
            self.ship.upload(col_s, dep, vel_s)
            s2 = SegmentationImages()
            s2.material, s2.prev_material, s2.uv_normal = sg.material, sg.prev_material, sg.uv_normal
            s2.prev_segmask, s2.segmask = self.mask_s[(k + 1) & 1].data_ptr(), self.mask_s[k & 1].data_ptr()
            check(lib().taa_resolve_frame_segmented(C.byref(U), C.byref(self.ship.images()), C.byref(s2),
                                                    C.byref(self.cas) if self.cas else None, self.arith, None))
            self.ship.k += 1
        torch.cuda.synchronize()
        if self.ship is not None:
            assert_same_bits(self.bufs.screen, self.ship.screen, f"frame {k}: screen without debug and inspect images")
            assert_same_bits(self.bufs.hist_in, self.ship.hist_in, f"frame {k}: history without debug and inspect images")
            assert torch.equal(self.mask[k & 1], self.mask_s[k & 1]), f"frame {k}: mask without debug and inspect images"
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
        # the FP16 arithmetic's history departs from the transcription's in its low bits: the
        # oracle continues from the GPU's history, so that every frame is held to one step's bar
        self.hist_o = out["gpu_history"] if self.arith == ARITH_FP16 else o_hist
        self.prev_depth = dep
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


def assert_seg_fp16(out):
    """The FP16 arithmetic's bar: mask, indicator sum, indices, taps and decisions exact; the
    sharpened screen > 55 dB."""
    gi, oi = out["gpu_inspect"], out["oracle_inspect"]
    for wd in INDEX_WORDS + [INSPECT["SEG_TOTAL"]]:
        g, o = gi[..., wd], oi[..., wd]
        if wd == INSPECT["FLAGS"]:
            g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
        assert np.count_nonzero(g != o) == 0, f"inspect word {wd}"
    assert np.count_nonzero(out["gpu_mask"] != out["oracle_mask"]) == 0, "segmentation mask"
    assert psnr(out["gpu_screen"][..., :3], out["oracle_screen"][..., :3]) > 55.0


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
    panning scene: categories 0, 1 and 2 all occur, with the fused CAS; without the debug and inspect
    images (the production form of the masked resolve) the same bits."""
    print(f"case seg default fmt={fmt}", flush=True)
    w = workload("c2", fmt=fmt, vel=VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F)
    seq = SegSequence(w, 320, 180, params=seg_params(), shipped=True)
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


UP_CASES = {"2x": ((160, 96), (320, 192)), "1.5x": ((200, 120), (300, 180)), "1.33x_ragged": ((75, 41), (101, 57))}


@pytest.mark.parametrize("fmt", [FMT_RGBA32F, FMT_RGBA16F])
@pytest.mark.parametrize("name", sorted(UP_CASES))
def test_segmentation_upsampled(oracle, name, fmt):
    """TAAU (mUpsampling) with the mask: the reference's default flags, every texel's mask, sum and
    resolve against the transcription."""
    print(f"case seg upsampled {name} fmt={fmt}", flush=True)
    lo, hi = UP_CASES[name]
    w = workload("c2", fmt=fmt, vel=VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F)
    seq = SegSequence(w, hi[0], hi[1], params=seg_params(), input_size=lo, upsampling=True)
    cats = set()
    for _ in range(4):
        out = seq.step(oracle)
        assert_seg_parity(out)
        cats |= set(np.unique(out["gpu_mask"] & 3).tolist())
    assert {0, 2} <= cats, cats


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_segmentation_upsampled_flag_variants(oracle, name):
    """The flag variants on the upsampling path; 240x220 input, so that the fixed border (measured
    against the input size) shows inside the 360x330 output."""
    print(f"case seg upsampled variant {name}", flush=True)
    v = dict(VARIANTS[name])
    seq = SegSequence(workload("c2"), 360, 330, cas=False, params=seg_params(v.pop("flags"), **v), input_size=(240, 220),
                      upsampling=True)
    for _ in range(3):
        assert_seg_parity(seq.step(oracle))


@pytest.mark.parametrize("up", [False, True])
def test_segmentation_fp16_arithmetic(oracle, up):
    """TAA_ARITH_FP16 with the mask, at one size and upsampling: the mask and every selection exact;
    without the debug and inspect images the same bits."""
    print(f"case seg fp16 up={up}", flush=True)
    w = workload("c3")
    seq = SegSequence(w, 320, 192, params=seg_params(), input_size=(160, 96) if up else None, upsampling=up,
                      arith=ARITH_FP16, shipped=True)
    for _ in range(4):
        assert_seg_fp16(seq.step(oracle))


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
    m2 = torch.zeros(32, 64, dtype=torch.int32, device="cuda")
    sg.segmask = m2.data_ptr()
    sg.input_pitch = -1
    assert L.taa_resolve_frame_segmented(C.byref(U), C.byref(img), C.byref(sg), None, 0, None) == EINVAL
    sg.input_pitch, sg.reserved[0] = 0, 1
    assert L.taa_resolve_frame_segmented(C.byref(U), C.byref(img), C.byref(sg), None, 0, None) == EINVAL
    sg.reserved[0] = 0
    img.input_width, img.input_height = 32, 16
    sg.input_pitch = 31  # narrower than the input width
    assert L.taa_resolve_frame_segmented(C.byref(U), C.byref(img), C.byref(sg), None, 0, None) == EINVAL
