"""The upsampling (TAAU) path on the GPU against the FP32 transcription: input (render) size
different from the output size, and mUpsampling (taa.comp:222-257, 724, 750-755, 818).

Bar: FP32 arithmetic bit-exact like the FP32 path at one size -- texel indices, depth-test texels,
closest taps, flags and the neighbourhood box identical; history and screen within 1e-5 relative
(observed 0). FP16 arithmetic: indices, taps and decisions identical, the screen > 55 dB.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import assert_parity, assert_same_bits, round_storage, to_f16  # noqa: E402
from taastar import (ARITH_FP16, ARITH_FP32, FMT_RGBA16F, FMT_RGBA32F, INSPECT_WORDS, VEL_RG16F, VEL_RG32F,  # noqa: E402
                     VEL_RGBA32F, FrameImages, check, lib)
from taastar.device import color_dtype, vel_shape_dtype  # noqa: E402
from taastar.workloads import (cas_for, decode, encode_color, encode_velocity, params_c2, scene_for,  # noqa: E402
                               uniforms_for, velocity_as_read, workload)


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


class Upsampler:
    """A render-resolution scene resolved into an output-resolution history, GPU and oracle. Every
    frame is also resolved into a second history chain without the debug and inspect images (the
    launch that takes the upsampling kernel's production form when the parameters are the
    production set); both chains must hold the same bits."""

    def __init__(self, w, lo, hi, cas=True, upsampling=True, params=None, arith=ARITH_FP32):
        self.w, self.fmt, self.vfmt = w, w["fmt"], w["vel"]
        (self.Wl, self.Hl), (self.W, self.H) = lo, hi
        self.scene = scene_for(w, width=self.Wl, height=self.Hl)
        self.cas = cas_for(w, self.W, self.H) if cas else None
        self.upsampling, self.params, self.arith = upsampling, params, arith
        cd = color_dtype(self.fmt)
        vc, vd = vel_shape_dtype(self.vfmt)
        z = lambda *s, dtype: torch.zeros(*s, dtype=dtype, device="cuda")
        self.color, self.depth = z(self.Hl, self.Wl, 4, dtype=cd), z(self.Hl, self.Wl, dtype=torch.float32)
        self.vel, self.prev_depth = z(self.Hl, self.Wl, vc, dtype=vd), z(self.Hl, self.Wl, dtype=torch.float32)
        self.hist = [z(self.H, self.W, 4, dtype=cd), z(self.H, self.W, 4, dtype=cd)]
        self.screen = z(self.H, self.W, 4, dtype=cd)
        self.hist_s = [z(self.H, self.W, 4, dtype=cd), z(self.H, self.W, 4, dtype=cd)]
        self.screen_s = z(self.H, self.W, 4, dtype=cd)
        self.debug = z(self.H, self.W, 4, dtype=torch.float16)
        self.inspect = z(self.H, self.W, INSPECT_WORDS, dtype=torch.int32)
        self.hist_o = np.zeros((self.H, self.W, 4), np.float32)
        self.prev_o = np.zeros((self.Hl, self.Wl), np.float32)
        self.k = 0

    def step(self, oracle):
        col, dep, vel4, jit = self.scene.render(self.k)
        col_s, vel_s = encode_color(col, self.fmt), encode_velocity(vel4, self.vfmt)
        self.prev_depth.copy_(self.depth)
        self.color.copy_(torch.from_numpy(col_s))
        self.depth.copy_(torch.from_numpy(dep))
        self.vel.copy_(torch.from_numpy(np.ascontiguousarray(vel_s)))
        U = uniforms_for(self.w, self.k, jit, params=self.params)
        U.mUpsampling = int(self.upsampling)
        img = FrameImages()
        img.color, img.depth, img.velocity = self.color.data_ptr(), self.depth.data_ptr(), self.vel.data_ptr()
        img.history, img.result_history = self.hist[(self.k + 1) & 1].data_ptr(), self.hist[self.k & 1].data_ptr()
        img.history_depth, img.result_screen = self.prev_depth.data_ptr(), self.screen.data_ptr()
        img.debug, img.inspect = self.debug.data_ptr(), self.inspect.data_ptr()
        img.width, img.height, img.pitch = self.W, self.H, self.W
        img.input_width, img.input_height, img.input_pitch = self.Wl, self.Hl, self.Wl
        img.color_format, img.velocity_format = self.fmt, self.vfmt
        k = (C.c_int32 * 2)()
        check(lib().taa_resolve_kernel(C.byref(U), C.byref(img), C.byref(self.cas) if self.cas else None, self.arith, k))
        self.last_kernel = (k[0], k[1])
        check(lib().taa_resolve_frame(C.byref(U), C.byref(img), C.byref(self.cas) if self.cas else None, self.arith,
                                      None))
        img.history, img.result_history = self.hist_s[(self.k + 1) & 1].data_ptr(), self.hist_s[self.k & 1].data_ptr()
        img.result_screen, img.debug, img.inspect = self.screen_s.data_ptr(), None, None
        check(lib().taa_resolve_frame(C.byref(U), C.byref(img), C.byref(self.cas) if self.cas else None, self.arith,
                                      None))
        torch.cuda.synchronize()
        assert_same_bits(self.screen, self.screen_s, f"frame {self.k}: screen without debug and inspect images")
        assert_same_bits(self.hist[self.k & 1], self.hist_s[self.k & 1],
                         f"frame {self.k}: history without debug and inspect images")
        o = oracle.resolve(U, decode(col_s), dep, velocity_as_read(vel_s), self.hist_o, self.prev_o)
        o_hist = round_storage(o["history"], self.fmt)
        o_scr = round_storage(o["screen"], self.fmt)
        if self.cas is not None:
            o_scr = round_storage(oracle.cas(o_scr, self.cas.const1[0]), self.fmt)
        out = dict(gpu_history=decode(self.hist[self.k & 1].cpu().numpy()), gpu_screen=decode(self.screen.cpu().numpy()),
                   gpu_inspect=self.inspect.cpu().numpy(), gpu_debug=self.debug.cpu().numpy().astype(np.float32),
                   oracle_history=o_hist, oracle_screen=o_scr, oracle_inspect=o["inspect"],
                   oracle_debug=to_f16(o["debug"]).astype(np.float32))
        self.hist_o, self.prev_o = o_hist, dep
        self.k += 1
        return out


CASES = {
    "2x_upsampled": dict(lo=(96, 54), hi=(192, 108), upsampling=True),
    "1.5x_upsampled": dict(lo=(128, 72), hi=(192, 108), upsampling=True),
    "2x_nearest": dict(lo=(96, 54), hi=(192, 108), upsampling=False),
    "1.33x_ragged_nocas": dict(lo=(75, 41), hi=(101, 57), upsampling=True, cas=False),
    "1x_upsampling_flag": dict(lo=(160, 90), hi=(160, 90), upsampling=True),
    # an input larger than the output: the tile's input texels do not fit the staged square and
    # every fetch reads global memory
    "0.5x_supersampled": dict(lo=(192, 108), hi=(96, 54), upsampling=True),
}


@pytest.mark.parametrize("fmt", [FMT_RGBA32F, FMT_RGBA16F])
@pytest.mark.parametrize("name", sorted(CASES))
def test_taau_sequence(oracle, name, fmt):
    c = dict(CASES[name])
    print(f"case taau {name} fmt={fmt}", flush=True)
    w = workload("c2", fmt=fmt, vel=VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F)
    up = Upsampler(w, c.pop("lo"), c.pop("hi"), **c)
    for _ in range(5):
        assert_parity(up.step(oracle))


VARIANTS = {
    "simple_velocity": dict(mVelocitySampleMode=0),
    "longest_velocity": dict(mVelocitySampleMode=1),
    "antighost_unjitter": dict(mDynamicAntiGhosting=1, mUnjitterNeighbourhood=1, mUnjitterCurrentSample=1),
    "unjitter_neighbourhood_nearest": dict(mUnjitterNeighbourhood=1),
    "matrix_reproj_raw_depth": dict(mUseVelocityVectors=0),
    "debug_velocity": dict(mDebugMode=4),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_taau_feature_variants(oracle, name):
    print(f"case taau variant {name}", flush=True)
    p = params_c2()
    for k, v in VARIANTS[name].items():
        setattr(p, k, v)
    w = workload("c2", vel=VEL_RGBA32F)
    up = Upsampler(w, (80, 48), (160, 96), params=p, upsampling=not name.endswith("_nearest"))
    if name == "matrix_reproj_raw_depth":
        up.w = dict(up.w, linear_depth=0.0)
    for _ in range(4):
        out = up.step(oracle)
        assert_parity(out)
        if name == "debug_velocity":
            assert np.array_equal(out["gpu_debug"], out["oracle_debug"])


@pytest.mark.parametrize("fmt", [FMT_RGBA32F, FMT_RGBA16F])
@pytest.mark.parametrize("name", sorted(CASES))
def test_taau_fp16_arithmetic(oracle, name, fmt):
    """ARITH_FP16 on the upsampling path: the upsampling resolve with the FP16 path's arithmetic
    (TAA_KERNEL_UPSAMPLE_FP16). Texel indices, depth-test texels, closest taps and decisions
    bit-exact against the FP32 transcription; the screen > 55 dB against it (sharpened: peak 1;
    unsharpened: its display-referred, clamped value); the oracle follows the GPU's history."""
    from gpu_helpers import INDEX_WORDS, psnr
    from taastar import FLAG_RECTIFIED, INSPECT, KERNEL_UPSAMPLE_FP16
    c = dict(CASES[name])
    print(f"case taau fp16 {name} fmt={fmt}", flush=True)
    w = workload("c2", fmt=fmt, vel=VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F)
    up = Upsampler(w, c.pop("lo"), c.pop("hi"), arith=ARITH_FP16, **c)
    for f in range(5):
        out = up.step(oracle)
        assert up.last_kernel == (KERNEL_UPSAMPLE_FP16, KERNEL_UPSAMPLE_FP16)
        gi, oi = out["gpu_inspect"], out["oracle_inspect"]
        for wd in INDEX_WORDS:
            g, o = gi[..., wd], oi[..., wd]
            if wd == INSPECT["FLAGS"]:
                g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
            assert np.count_nonzero(g != o) == 0, f"frame {f}: inspect word {wd}"
        g, o = out["gpu_screen"][..., :3], out["oracle_screen"][..., :3]
        if up.cas is None:
            g, o = np.clip(g, 0, 1), np.clip(o, 0, 1)
        assert psnr(g, o) > 55.0, f"frame {f}: screen PSNR {psnr(g, o):.2f} dB"
        up.hist_o = out["gpu_history"]


def test_scaled_session_equals_resolve_frame(oracle):
    """A session opened at render size 96x54 -> output 192x108 (taa_session_open_scaled) gives the
    bits of taa_resolve_frame with the same input size, frame after frame; its input slots have the
    render size."""
    from taastar import SessionIO
    print("case taau scaled session", flush=True)
    w = workload("c2", fmt=FMT_RGBA16F, vel=VEL_RG16F)
    up = Upsampler(w, (96, 54), (192, 108), arith=ARITH_FP16)  # the session below resolves in FP16 too
    ctx, s = C.c_void_p(), C.c_void_p()
    check(lib().taa_ctx_create(0, 0, C.byref(ctx)))
    check(lib().taa_session_open_scaled(ctx, 192, 108, 96, 54, w["fmt"], w["vel"], C.byref(s)))
    scr = torch.zeros(108, 192, 4, dtype=torch.float16, device="cuda")
    for f in range(4):
        ref = up.step(oracle)  # resolve_frame on the GPU (and the oracle, already checked)
        io = SessionIO()
        check(lib().taa_session_io(s, C.byref(io)))
        assert (io.width, io.height, io.pitch) == (96, 54, 96)
        hmemcpy = lib().hipMemcpy
        hmemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        torch.cuda.synchronize()
        for ptr, src in ((io.color, up.color), (io.depth, up.depth), (io.velocity, up.vel)):
            nbytes = src.numel() * src.element_size()
            assert hmemcpy(C.c_void_p(ptr), C.c_void_p(src.data_ptr()), nbytes, 3) == 0  # device to device
        U = uniforms_for(w, f, up.scene.render(f)[3])
        U.mUpsampling = 1
        check(lib().taa_session_resolve(s, C.byref(U), C.byref(up.cas), C.c_void_p(scr.data_ptr()), ARITH_FP16, None))
        torch.cuda.synchronize()
        assert torch.equal(scr, up.screen), f"frame {f}: screen"
    check(lib().taa_session_close(s))
    check(lib().taa_ctx_destroy(ctx))


def test_batch_with_scaled_sessions_equals_alone():
    """A batch holding two one-size sessions and a scaled one, submitted and through a graph-captured
    plan: every result equals the session resolved alone. The upsampling frames go through the
    batched upsampling kernel; the third session sets mUpsampling on odd frames, which moves it
    between the two launches (and makes the plan recapture)."""
    from taastar import BatchItem, SessionIO, Uniforms
    print("case taau batch mix", flush=True)
    w = workload("c2", fmt=FMT_RGBA16F, vel=VEL_RG16F)
    specs = [((160, 90), (160, 90)), ((120, 68), (240, 136)), ((96, 96), (96, 96))]  # (input, output)
    scenes = [scene_for(w, width=i[0], height=i[1], seed=500 + k) for k, (i, _) in enumerate(specs)]
    hmemcpy = lib().hipMemcpy
    hmemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def upload(s, sc, f, Wl, Hl):
        io = SessionIO()
        check(lib().taa_session_io(s, C.byref(io)))
        col, dep, vel4, jit = sc.render(f)
        for ptr, arr in ((io.color, encode_color(col, w["fmt"])), (io.depth, dep), (io.velocity, encode_velocity(vel4, w["vel"]))):
            a = np.ascontiguousarray(arr)
            torch.cuda.synchronize()
            assert hmemcpy(C.c_void_p(ptr), C.c_void_p(a.ctypes.data), a.nbytes, 1) == 0
        return uniforms_for(w, f, jit)

    def run(mode):
        ctx = C.c_void_p()
        check(lib().taa_ctx_create(0, 0, C.byref(ctx)))
        sess, scrs, Ks = [], [], []
        for (i, o) in specs:
            s = C.c_void_p()
            check(lib().taa_session_open_scaled(ctx, o[0], o[1], i[0], i[1], w["fmt"], w["vel"], C.byref(s)))
            sess.append(s)
            scrs.append(torch.zeros(o[1], o[0], 4, dtype=torch.float16, device="cuda"))
            Ks.append(cas_for(w, o[0], o[1]))
        outs = []
        Us = [Uniforms() for _ in specs]
        items = (BatchItem * len(specs))()
        for k in range(len(specs)):
            items[k].session, items[k].uniforms = sess[k], C.pointer(Us[k])
            items[k].cas, items[k].result_screen = C.pointer(Ks[k]), scrs[k].data_ptr()
        plan = C.c_void_p()
        if mode == "plan":
            check(lib().taa_batch_plan_create(ctx, len(specs), items, ARITH_FP16, C.byref(plan)))
        for f in range(6):
            for k, (i, o) in enumerate(specs):
                U = upload(sess[k], scenes[k], f, *i)
                U.mUpsampling = int(i != o or (k == 2 and f % 2 == 1))
                C.memmove(C.addressof(Us[k]), C.addressof(U), C.sizeof(Uniforms))
            if mode == "batch":
                check(lib().taa_batch_submit(ctx, len(specs), items, ARITH_FP16, None))
            elif mode == "plan":
                check(lib().taa_batch_plan_launch(plan, None))
            else:
                for k in range(len(specs)):
                    check(lib().taa_session_resolve(sess[k], C.byref(Us[k]), C.byref(Ks[k]),
                                                    C.c_void_p(scrs[k].data_ptr()), ARITH_FP16, None))
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
        for f in range(6):
            for k in range(len(specs)):
                assert np.array_equal(alone[f][k], got[f][k]), f"{mode}: frame {f} session {k}"
