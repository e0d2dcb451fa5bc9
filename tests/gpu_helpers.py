"""Shared helpers of the GPU parity tests: run a frame sequence on the GPU and through the oracle."""
import numpy as np

from taastar import FMT_RGBA16F, INSPECT
from taastar.workloads import (cas_for, decode, encode_color, encode_velocity, scene_for, uniforms_for,
                               velocity_as_read)

REL_TOL = 1e-5      # FP32 path vs FP32 transcription, relative per channel
ABS_FLOOR = 1e-3    # ... of max(|ref|, ABS_FLOOR): YCoCg / Catmull-Rom intermediates cross zero


def to_f16(a):
    """fp32 -> fp16 as the kernels store it: nearest even, saturating to +-65504, NaN kept."""
    a = np.asarray(a, np.float32)
    with np.errstate(invalid="ignore"):
        return np.where(np.isnan(a), a, np.clip(a, -65504.0, 65504.0)).astype(np.float16)


def round_storage(a, fmt):
    return to_f16(a).astype(np.float32) if fmt == FMT_RGBA16F else a.astype(np.float32)


def rel_err(got, ref, floor=ABS_FLOOR):
    """|got - ref| / max(|ref|, floor); identical values (also Inf, and NaN vs NaN) count as 0,
    a finite vs non-finite mismatch as Inf."""
    got = got.astype(np.float64)
    ref = ref.astype(np.float64)
    same = (got == ref) | (np.isnan(got) & np.isnan(ref))
    with np.errstate(invalid="ignore", over="ignore"):
        e = np.abs(got - ref) / np.maximum(np.abs(ref), floor)
    e = np.where(same, 0.0, e)
    return np.where(np.isnan(e), np.inf, e)


def psnr(got, ref, peak=1.0):
    mse = np.mean((got.astype(np.float64) - ref.astype(np.float64)) ** 2)
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


INDEX_WORDS = [INSPECT["HIST_TC_X"], INSPECT["HIST_TC_Y"], INSPECT["CLOSEST_TAP"], INSPECT["FLAGS"],
               INSPECT["DEPTH_TC_X"], INSPECT["DEPTH_TC_Y"]]
BOX_WORDS = list(range(INSPECT["COLMIN"], INSPECT["COLMIN"] + 3)) + list(range(INSPECT["COLMAX"], INSPECT["COLMAX"] + 3))


class Sequence:
    """Frames of one seeded scene, resolved on the GPU and by the FP32 oracle side by side."""

    def __init__(self, w, width=None, height=None, cas=True, arith=0, params=None, seed=None, pan_mode=0, pan=8.0,
                 debug=True, shipped=False):
        import torch
        from taastar.device import FrameBuffers
        self.w = w
        self.W = width or w["width"]
        self.H = height or w["height"]
        self.fmt, self.vfmt = w["fmt"], w["vel"]
        self.scene = scene_for(w, seed=seed, width=self.W, height=self.H, pan_mode=pan_mode, pan=pan)
        self.cas = cas_for(w, self.W, self.H) if cas else None
        self.arith = arith
        self.params = params
        # a bound debug image sends an FP16 request to the exact kernel (the throughput kernel
        # writes no debug image): debug=False keeps the FP16 kernel on the FP16 path
        self.bufs = FrameBuffers(self.W, self.H, self.fmt, self.vfmt, debug=debug, inspect=True)
        # shipped: the same frames also through buffers without the inspect image (the FP16 kernel
        # has a separate instantiation for it, and without inspect and debug images the exact kernel
        # runs its production form for the production feature set); step() asserts both write the
        # same bits
        self.ship = FrameBuffers(self.W, self.H, self.fmt, self.vfmt, debug=debug) if shipped else None
        self.hist_o = np.zeros((self.H, self.W, 4), np.float32)
        self.prev_depth = np.zeros((self.H, self.W), np.float32)
        self.frame = 0
        self.torch = torch

    def kernel(self, uniforms=None):
        """(kernel of param[0]'s columns, of param[1]'s) the next step's resolve runs (taa_resolve_kernel)."""
        import ctypes as C
        from taastar import check, lib
        U = uniforms if uniforms is not None else uniforms_for(self.w, self.frame, (0.0, 0.0), params=self.params)
        k = (C.c_int32 * 2)()
        check(lib().taa_resolve_kernel(C.byref(U), C.byref(self.bufs.images()), C.byref(self.cas) if self.cas else None,
                                       self.arith, k))
        return k[0], k[1]

    def step(self, oracle, run_oracle=True, uniforms=None):
        col, dep, vel4, jit = self.scene.render(self.frame)
        col_s = encode_color(col, self.fmt)
        vel_s = encode_velocity(vel4, self.vfmt)
        self.bufs.upload(col_s, dep, vel_s)
        U = uniforms if uniforms is not None else uniforms_for(self.w, self.frame, jit, params=self.params)
        self.bufs.resolve(U, self.cas, self.arith)
        if self.ship is not None:
            self.ship.upload(col_s, dep, vel_s)
            self.ship.resolve(U, self.cas, self.arith)
        self.torch.cuda.synchronize()
        if self.ship is not None:
            assert_same_bits(self.bufs.screen, self.ship.screen, f"frame {self.frame}: screen without the inspect image")
            assert_same_bits(self.bufs.hist_in, self.ship.hist_in, f"frame {self.frame}: history without the inspect image")
        g_hist = decode(self.bufs.hist_in.cpu().numpy())  # hist_in now points at the frame just written
        g_scr = decode(self.bufs.screen.cpu().numpy())
        g_ins = self.bufs.inspect.cpu().numpy()
        g_dbg = self.bufs.debug.cpu().numpy().astype(np.float32) if self.bufs.debug is not None else None
        out = dict(gpu_history=g_hist, gpu_screen=g_scr, gpu_inspect=g_ins, gpu_debug=g_dbg, uniforms=U)
        if run_oracle:
            o = oracle.resolve(U, decode(col_s), dep, velocity_as_read(vel_s), self.hist_o, self.prev_depth)
            o_hist = round_storage(o["history"], self.fmt)
            o_scr = round_storage(o["screen"], self.fmt)
            if self.cas is not None:
                o_scr = round_storage(oracle.cas(o_scr, self.cas.const1[0]), self.fmt)
            o_dbg = to_f16(o["debug"]).astype(np.float32)  # uDebug is RGBA16F
            out.update(oracle_history=o_hist, oracle_screen=o_scr, oracle_inspect=o["inspect"], oracle_debug=o_dbg)
            self.hist_o = o_hist
        self.prev_depth = dep
        self.frame += 1
        return out


def assert_same_bits(a, b, what):
    """Two device tensors of one dtype (fp16 or fp32 words) hold identical bit patterns, NaNs aside:
    a NaN's sign and payload depend on which operand carried it into each operation, which the
    contract leaves open (the oracle checks count NaN against NaN as equal too), so any NaN matches
    any NaN; every other word must match bit for bit."""
    import torch
    it, ft = (torch.int16, torch.float16) if a.element_size() == 2 else (torch.int32, torch.float32)
    fa, fb = a.view(ft), b.view(ft)
    diff = (a.view(it) != b.view(it)) & ~(torch.isnan(fa) & torch.isnan(fb))
    bad = int(diff.sum().item())
    if bad:
        idx = torch.nonzero(diff.reshape(-1))[:4, 0]
        pairs = [(float(fa.reshape(-1)[i]), float(fb.reshape(-1)[i])) for i in idx.tolist()]
        raise AssertionError(f"{what}: {bad} words differ, e.g. {pairs}")


def assert_parity(out, rel_tol=REL_TOL, exact_box=True):
    gi, oi = out["gpu_inspect"], out["oracle_inspect"]
    for wd in INDEX_WORDS:
        bad = np.count_nonzero(gi[..., wd] != oi[..., wd])
        assert bad == 0, f"inspect word {wd}: {bad} texels differ"
    if exact_box:
        bad = np.count_nonzero(gi[..., BOX_WORDS] != oi[..., BOX_WORDS])
        assert bad == 0, f"neighbourhood box: {bad} words differ"
    for k in ("history", "screen"):
        e = rel_err(out["gpu_" + k], out["oracle_" + k]).max()
        assert e <= rel_tol, f"{k}: max rel err {e:.3g} > {rel_tol}"


# ---- the FP16 path against the FP64 oracle ------------------------------------------------------

FP16_PSNR_MIN_DB = 55.0
FP16_DEPART_ABS = 1e-2          # FP32 transcription on FP16 storage vs FP64, sharpened screen
# cap on the excluded pixels per frame: the footprint comes from the oracle alone and grows with
# motion (at most 0.26% of a frame over the 64 4K frames of the 8 px/frame pan, 0.31% by frame 8 of
# a 4K session panning 24 px/frame); the cap only guards against a mask that would hide the frame
FP16_MASK_MAX_FRACTION = 1e-2


class FP16OracleChain:
    """One stream's history under the FP64 oracle (the accuracy reference) and under the FP32
    transcription on FP16 storage (the selections, and what the reference's own arithmetic gives on
    FP16 storage), fed the inputs the GPU resolved, frame by frame.

    check() holds the FP16 path to its bar: every texel index, closest-depth tap, depth-test texel
    and rejection decision equal to the FP32 transcription's (when the inspect words are given);
    the sharpened screen (display-referred, peak 1.0) >= 55 dB PSNR against the FP64 oracle outside
    the oracle's own divergence footprint, which must stay under 1% of the frame; the full-frame
    PSNR within 3 dB of the FP32 transcription's on FP16 storage (or above 60 dB).

    Divergence footprint: taa.comp's Lottes weight w = 1 - |lc - lh| / max(lc, lh, 0.2) is not
    clamped (taa.comp:865-869); once Catmull-Rom undershoot makes a luma negative, w*w > 1 makes
    alpha negative and the blend extrapolates, so a few pixels grow without bound in the FP64 oracle
    itself and others are chaotic (rounding the history to FP16 alone moves them by up to 1.0).
    Excluded: (a) pixels whose FP64 history exceeded 4x the frame's input maximum in the last 8
    frames, dilated by the Catmull-Rom support; (b) pixels where the FP32 transcription on FP16
    storage is itself more than 0.01 from the FP64 result, dilated by the CAS neighbourhood."""

    def __init__(self, oracle, W, H, cas, name=""):
        self.oracle, self.W, self.H, self.cas, self.name = oracle, W, H, cas, name
        self.h64 = np.zeros((H, W, 4))
        self.h32 = np.zeros((H, W, 4), np.float32)
        self.prev = np.zeros((H, W), np.float32)
        self.recent = []
        self.frame = 0
        self.worst = float("inf")
        self.reports = []

    def check(self, U, col_s, dep, vel_s, g_scr, g_ins=None):
        rep = self.measure(U, col_s, dep, vel_s, g_scr, g_ins)
        f, p_scr, p_all, p_base = rep["frame"], rep["screen_psnr_db"], rep["full_frame_psnr_db"], rep["fp32_fp16storage_psnr_db"]
        assert p_scr >= FP16_PSNR_MIN_DB, f"{self.name} frame {f}: screen PSNR {p_scr:.2f} dB < {FP16_PSNR_MIN_DB}: {rep}"
        assert p_all >= min(p_base - 3.0, 60.0), f"{self.name} frame {f}: full-frame {p_all:.2f} dB vs FP32 path {p_base:.2f} dB"
        assert rep["masked_pixels"] <= FP16_MASK_MAX_FRACTION * self.W * self.H, rep
        return rep

    def measure(self, U, col_s, dep, vel_s, g_scr, g_ins=None):
        """One frame's figures (the report check() holds to the bar); advances the chains."""
        from scipy.ndimage import binary_dilation
        from taastar import FLAG_RECTIFIED
        oracle, k, f = self.oracle, self.cas, self.frame
        args = (decode(col_s), dep, velocity_as_read(vel_s))
        o64 = oracle.resolve(U, *args, self.h64, self.prev, dtype=np.float64, debug=False, inspect=False)
        o32 = oracle.resolve(U, *args, self.h32, self.prev, dtype=np.float32, debug=False, inspect=g_ins is not None)
        if g_ins is not None:  # selections are made in fp32 arithmetic on both sides
            for wd in INDEX_WORDS:
                g, o = g_ins[..., wd], o32["inspect"][..., wd]
                if wd == INSPECT["FLAGS"]:  # "rectified" is a threshold on arithmetic, not a selection
                    g, o = g & ~FLAG_RECTIFIED, o & ~FLAG_RECTIFIED
                bad = np.count_nonzero(g != o)
                assert bad == 0, f"{self.name} frame {f}: inspect word {wd} differs at {bad} texels"
        if k is not None:
            s64 = oracle.cas(o64["screen"], k.const1[0], dtype=np.float64)
            s32 = oracle.cas(to_f16(o32["screen"]).astype(np.float32), k.const1[0])
        else:  # unsharpened HDR screen: compare the display-referred (clamped) values
            s64 = np.clip(o64["screen"], 0.0, 1.0)
            s32 = np.clip(to_f16(o32["screen"]).astype(np.float32), 0.0, 1.0)
            g_scr = np.clip(g_scr, 0.0, 1.0)
        cmax = float(decode(col_s)[..., :3].max())
        self.recent.append(np.abs(o64["history"][..., :3]).max(-1) > 4.0 * cmax)
        del self.recent[:-8]
        departs = np.abs(s32[..., :3] - s64[..., :3]).max(-1) > FP16_DEPART_ABS
        unstable = binary_dilation(np.logical_or.reduce(self.recent), iterations=3) | binary_dilation(departs, iterations=1)
        stable = ~unstable
        p_all = psnr(g_scr[..., :3], s64[..., :3], 1.0)
        p_base = psnr(s32[..., :3], s64[..., :3], 1.0)
        p_scr = psnr(g_scr[stable][:, :3], s64[stable][:, :3], 1.0) if stable.any() else float("inf")
        err = float(np.abs(g_scr[stable][:, :3] - s64[stable][:, :3]).max()) if stable.any() else 0.0
        rep = dict(frame=f, screen_psnr_db=p_scr, screen_max_abs=err, full_frame_psnr_db=p_all,
                   fp32_fp16storage_psnr_db=p_base, masked_pixels=int(unstable.sum()))
        self.reports.append(rep)
        print(f"{self.name} {self.W}x{self.H} frame {f}: PSNR {p_scr:.2f} dB outside {int(unstable.sum())} masked "
              f"pixels, full frame {p_all:.2f} dB (FP32 transcription on FP16 storage: {p_base:.2f} dB)", flush=True)
        self.worst = min(self.worst, p_scr)
        self.h64 = o64["history"]
        self.h32 = to_f16(o32["history"]).astype(np.float32)
        self.prev = dep
        self.frame += 1
        return rep


class FP32OracleChain:
    """One stream under the FP32 transcription (the exact kernel's bar: history and screen within
    1e-5 relative per channel, the inspect words bit-exact when given), fed the GPU's inputs."""

    def __init__(self, oracle, W, H, fmt, cas, name=""):
        self.oracle, self.W, self.H, self.fmt, self.cas, self.name = oracle, W, H, fmt, cas, name
        self.hist = np.zeros((H, W, 4), np.float32)
        self.prev = np.zeros((H, W), np.float32)
        self.frame = 0

    def check(self, U, col_s, dep, vel_s, g_scr, g_his, g_ins=None, rel_tol=REL_TOL):
        oracle = self.oracle
        o = oracle.resolve(U, decode(col_s), dep, velocity_as_read(vel_s), self.hist, self.prev, debug=False,
                           inspect=g_ins is not None)
        o_his = round_storage(o["history"], self.fmt)
        o_scr = round_storage(o["screen"], self.fmt)
        if self.cas is not None:
            o_scr = round_storage(oracle.cas(o_scr, self.cas.const1[0]), self.fmt)
        if g_ins is not None:
            for wd in INDEX_WORDS + BOX_WORDS:
                bad = np.count_nonzero(g_ins[..., wd] != o["inspect"][..., wd])
                assert bad == 0, f"{self.name} frame {self.frame}: inspect word {wd}: {bad} texels differ"
        for what, g, r in (("history", g_his, o_his), ("screen", g_scr, o_scr)):
            e = rel_err(g, r).max()
            assert e <= rel_tol, f"{self.name} frame {self.frame}: {what} max rel err {e:.3g} > {rel_tol}"
        self.hist = o_his
        self.prev = dep
        self.frame += 1

