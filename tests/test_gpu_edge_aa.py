"""The edge-AA fallback pass (taa_edge_aa_apply) on the GPU against the FP32 oracle.

Bar: bit-exact in the storage format (float32 bits, or the fp16 rounding of the oracle's float32
result); row padding untouched.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import to_f16  # noqa: E402
from taastar import FMT_RGBA16F, FMT_RGBA32F, EdgeAaConsts, check, lib  # noqa: E402
from taastar.device import color_dtype  # noqa: E402
from taastar.workloads import encode_color, scene_for, workload  # noqa: E402


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def consts(W, H, **kw):
    k = EdgeAaConsts()
    check(lib().taa_edge_aa_defaults(W, H, C.byref(k)))
    for a, v in kw.items():
        setattr(k, a, v)
    return k


def run(src_s, mask, fmt, k, pad=0):
    H, W = mask.shape
    P = W + pad
    cd = color_dtype(fmt)
    src = torch.full((H, P, 4), float("nan"), dtype=cd, device="cuda")
    src[:, :W] = torch.from_numpy(src_s).cuda()
    m = torch.zeros((H, P), dtype=torch.int32, device="cuda")
    m[:, :W] = torch.from_numpy(mask.view(np.int32)).cuda()
    dst = torch.full((H, P, 4), 7.0, dtype=cd, device="cuda")
    check(lib().taa_edge_aa_apply(C.c_void_p(src.data_ptr()), C.c_void_p(m.data_ptr()), C.c_void_p(dst.data_ptr()), W, H,
                                  P, P, fmt, C.byref(k), None))
    torch.cuda.synchronize()
    return dst.cpu().numpy()


def stored(a, fmt):
    return to_f16(a) if fmt == FMT_RGBA16F else np.asarray(a, np.float32)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint16 if a.dtype == np.float16 else np.uint32)


@pytest.mark.parametrize("fmt", [FMT_RGBA32F, FMT_RGBA16F])
def test_edge_aa_matches_the_oracle(oracle, fmt):
    """The seeded scene's colour (clamped to the display range, as after tone mapping) with a mask of
    all three categories in blocks and scattered texels, 1920x1080."""
    print(f"case edge-aa 1920x1080 fmt={fmt}", flush=True)
    W, H = 1920, 1080
    col = np.clip(scene_for(workload("c2"), width=W, height=H).render(3)[0], 0.0, 1.0)
    src_s = encode_color(col, fmt)
    rng = np.random.default_rng(31)
    mask = np.repeat(np.repeat(rng.integers(0, 3, (H // 8 + 1, W // 8 + 1)), 8, 0), 8, 1)[:H, :W].astype(np.uint32)
    mask[rng.integers(0, H, 5000), rng.integers(0, W, 5000)] = 1
    mask |= np.uint32(3 << 16)
    k = consts(W, H)
    got = run(src_s, mask, fmt, k)
    ref = stored(oracle.edge_aa(src_s.astype(np.float32), mask, k.subpix, k.edge_threshold, k.edge_threshold_min), fmt)
    bad = np.count_nonzero(bits(got) != bits(ref))
    assert bad == 0, f"{bad} words differ"
    changed = np.any(got[..., :3].astype(np.float32) != src_s[..., :3].astype(np.float32), -1)
    assert changed.sum() > 1000 and np.all((mask[changed] & 3) == 1)


@pytest.mark.parametrize("W,H", [(1, 1), (3, 2), (65, 3), (130, 9)])
def test_edge_aa_small_sizes_and_padding(oracle, W, H):
    print(f"case edge-aa {W}x{H}", flush=True)
    rng = np.random.default_rng(W * 100 + H)
    src = np.concatenate([rng.uniform(0, 1, (H, W, 3)), np.ones((H, W, 1))], -1).astype(np.float32)
    mask = np.ones((H, W), np.uint32)
    k = consts(W, H, subpix=1.0, edge_threshold=0.0, edge_threshold_min=0.0)
    got = run(src, mask, FMT_RGBA32F, k, pad=5)
    ref = oracle.edge_aa(src, mask, 1.0, 0.0, 0.0)
    assert np.count_nonzero(bits(got[:, :W]) != bits(ref)) == 0
    assert np.all(got[:, W:] == 7.0), "row padding written"


def test_edge_aa_argument_errors():
    L = lib()
    k = consts(8, 8)
    t = torch.zeros(8, 8, 4, dtype=torch.float32, device="cuda")
    m = torch.zeros(8, 8, dtype=torch.int32, device="cuda")
    from taastar._abi import TAA_ERR_INVALID_ARG as EINVAL
    p = lambda x: C.c_void_p(x.data_ptr())
    assert L.taa_edge_aa_apply(p(t), p(m), p(t), 8, 8, 8, 8, FMT_RGBA32F, C.byref(k), None) == EINVAL
    assert L.taa_edge_aa_apply(p(t), p(m), None, 8, 8, 8, 8, FMT_RGBA32F, C.byref(k), None) == EINVAL
    assert L.taa_edge_aa_apply(p(t), p(m), p(m), 8, 8, 4, 8, FMT_RGBA32F, C.byref(k), None) == EINVAL
