"""The velocity producer (taa_velocity_from_depth, fwd_geometry.frag:288-295 from depth) on the GPU.

Bar: bit-exact against the FP32 oracle in the stored format (float32 bits, or the fp16 rounding of
the oracle's float32 result), for every velocity format with and without object ids; row padding
untouched; the produced image drives the resolve exactly as the oracle's does.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from gpu_helpers import Sequence, assert_parity, to_f16  # noqa: E402
from taastar import VEL_RG16F, VEL_RG32F, VEL_RGBA16F, VEL_RGBA32F  # noqa: E402
from taastar.device import vel_shape_dtype, velocity_from_depth  # noqa: E402
from taastar.workloads import camera_case, params_c2, workload  # noqa: E402

FORMATS = {"rg32f": VEL_RG32F, "rg16f": VEL_RG16F, "rgba32f": VEL_RGBA32F, "rgba16f": VEL_RGBA16F}


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def stored(v4, vfmt):
    """The oracle's float32 velocity as the format stores it."""
    ch, _ = vel_shape_dtype(vfmt)
    v = np.ascontiguousarray(v4[..., :ch], np.float32)
    return to_f16(v) if vfmt in (VEL_RG16F, VEL_RGBA16F) else v


def produce(case, vfmt, pad=0, sentinel=7.0):
    """Run the producer on a case; returns the velocity tensor (with `pad` extra texels per row)."""
    d = case["depth"]
    H, W = d.shape
    P = W + pad
    ch, dt = vel_shape_dtype(vfmt)
    dep = torch.full((H, P), float("nan"), dtype=torch.float32, device="cuda")
    dep[:, :W] = torch.from_numpy(d).cuda()
    ids = None
    if case["ids"] is not None:
        ids = torch.zeros((H, P), dtype=torch.int32, device="cuda")
        ids[:, :W] = torch.from_numpy(case["ids"].astype(np.int32)).cuda()
    objs = None if case["objects"] is None else torch.from_numpy(case["objects"]).cuda()
    vel = torch.full((H, P, ch), sentinel, dtype=dt, device="cuda")
    velocity_from_depth(case["reproject"], dep, vel, vfmt, ids=ids, objects=objs, width=W)
    torch.cuda.synchronize()
    return vel


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint16 if a.dtype == np.float16 else np.uint32)


@pytest.mark.parametrize("objects", [0, 3])
@pytest.mark.parametrize("fmt", sorted(FORMATS))
def test_velocity_matches_the_oracle(oracle, fmt, objects):
    vfmt = FORMATS[fmt]
    print(f"case velocity 1920x1080 {fmt} objects={objects}", flush=True)
    case = camera_case(1920, 1080, seed=11, n_objects=objects)
    got = produce(case, vfmt).cpu().numpy()
    ref = stored(oracle.velocity(case["reproject"], case["depth"], case["ids"], case["objects"]), vfmt)
    bad = np.count_nonzero(bits(got) != bits(ref))
    assert bad == 0, f"{bad} words differ"
    assert np.abs(ref[..., :2].astype(np.float32)).max() > 1e-3   # the camera really moved


@pytest.mark.parametrize("W,H", [(1, 1), (37, 1), (1, 37), (65, 5), (130, 7)])
def test_velocity_small_sizes_and_padded_rows(oracle, W, H):
    print(f"case velocity {W}x{H} padded", flush=True)
    case = camera_case(W, H, seed=12, n_objects=2)
    got = produce(case, VEL_RGBA16F, pad=11).cpu().numpy()
    ref = stored(oracle.velocity(case["reproject"], case["depth"], case["ids"], case["objects"]), VEL_RGBA16F)
    assert np.count_nonzero(bits(got[:, :W]) != bits(ref)) == 0
    assert np.all(got[:, W:] == 7.0), "row padding written"


def test_produced_velocity_drives_the_resolve(oracle):
    """Moving-object velocity (mUseVelocityVectors = 1 uses v only where w >= 0.5, taa.comp:395) and
    closest-depth sampling with the producer's RGBA16F image: the resolve matches the oracle fed
    with the same image, frame after frame."""
    print("case velocity -> resolve", flush=True)
    p = params_c2()
    p.mUseVelocityVectors = 1
    w = workload("c2", vel=VEL_RGBA16F)
    seq = Sequence(w, width=240, height=136, cas=True, params=p)
    render = seq.scene.render

    def with_produced_velocity(frame, threads=0):
        c, d, _, jit = render(frame, threads)
        case = camera_case(240, 136, seed=20 + frame, n_objects=2)
        case["depth"] = np.ascontiguousarray(d, np.float32)
        v = produce(case, VEL_RGBA16F).cpu().numpy().astype(np.float32)
        ref = stored(oracle.velocity(case["reproject"], case["depth"], case["ids"], case["objects"]), VEL_RGBA16F)
        assert np.count_nonzero(bits(v.astype(np.float16)) != bits(ref)) == 0
        return c, d, v, jit

    seq.scene.render = with_produced_velocity
    for _ in range(4):
        assert_parity(seq.step(oracle))
