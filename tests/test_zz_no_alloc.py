"""taa_resolve_frame_noalloc against taa_resolve_frame (same bits), through the C ABI.

The no-allocation entry point keeps the intermediate screen of a split screen with CAS in a
file-scope device array shared by every call on the GPU; its results must equal those of
taa_resolve_frame, whose intermediate the library allocates. Two calls in a row, the second on a
smaller frame, for both storage formats and arithmetics; also the cases without the intermediate.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import taastar  # noqa: E402
from taastar import ARITH_FP16, ARITH_FP32, FMT_RGBA16F, FMT_RGBA32F, VEL_RG16F, VEL_RG32F, check, lib  # noqa: E402
from taastar.device import FrameBuffers  # noqa: E402
from taastar.workloads import cas_for, encode_color, encode_velocity, scene_for, uniforms_for, workload  # noqa: E402


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _same(a, b):
    it = torch.int16 if a.element_size() == 2 else torch.int32
    return int((a.view(it) != b.view(it)).sum().item()) == 0


@pytest.mark.parametrize("fmt,arith", [(FMT_RGBA16F, ARITH_FP16), (FMT_RGBA32F, ARITH_FP32)])
def test_noalloc_equals_resolve_frame(fmt, arith):
    w = workload("c2", fmt=fmt, vel=VEL_RG16F if fmt == FMT_RGBA16F else VEL_RG32F)
    L = lib()
    for call, (W, H) in enumerate([(640, 360), (331, 187)]):  # the second call on a smaller frame
        print(f"case noalloc call {call} {W}x{H} fmt={fmt} arith={arith}", flush=True)
        sc = scene_for(w, width=W, height=H, seed=0x0A110C + call)
        K = cas_for(w, W, H)
        a = FrameBuffers(W, H, fmt, w["vel"])
        b = FrameBuffers(W, H, fmt, w["vel"])
        for f in range(3):
            col, dep, vel4, jit = sc.render(f)
            for bb in (a, b):
                bb.upload(encode_color(col, fmt), dep, encode_velocity(vel4, w["vel"]))
            U = uniforms_for(w, f, jit)
            U.splitScreen = int(f != 1)  # frame 1: no split screen, no intermediate
            U.splitX = W // 3 + f
            U.param[1].mVarianceClipping = 0
            U.param[1].mInterpolationMode = 0
            check(L.taa_resolve_frame(C.byref(U), C.byref(a.images()), C.byref(K), arith, None))
            check(L.taa_resolve_frame_noalloc(C.byref(U), C.byref(b.images()), C.byref(K), arith, None))
            a.k += 1
            b.k += 1
            torch.cuda.synchronize()
            assert _same(a.screen, b.screen), f"call {call} frame {f}: screen"
            assert _same(a.hist_in, b.hist_in), f"call {call} frame {f}: history"
        # without CAS there is no intermediate either
        U.splitScreen = 1
        check(L.taa_resolve_frame(C.byref(U), C.byref(a.images()), None, arith, None))
        check(L.taa_resolve_frame_noalloc(C.byref(U), C.byref(b.images()), None, arith, None))
        torch.cuda.synchronize()
        assert _same(a.screen, b.screen) and _same(a.history[0], b.history[0]) and _same(a.history[1], b.history[1])
        del a, b
