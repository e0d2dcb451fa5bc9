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
                 debug=True):
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
        self.hist_o = np.zeros((self.H, self.W, 4), np.float32)
        self.prev_depth = np.zeros((self.H, self.W), np.float32)
        self.frame = 0
        self.torch = torch

    def step(self, oracle, run_oracle=True, uniforms=None):
        col, dep, vel4, jit = self.scene.render(self.frame)
        col_s = encode_color(col, self.fmt)
        vel_s = encode_velocity(vel4, self.vfmt)
        self.bufs.upload(col_s, dep, vel_s)
        U = uniforms if uniforms is not None else uniforms_for(self.w, self.frame, jit, params=self.params)
        self.bufs.resolve(U, self.cas, self.arith)
        self.torch.cuda.synchronize()
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
