"""The workloads at their full sizes, every frame against the oracle.

c2  1920x1080, FP32 storage and arithmetic, all 64 frames: texel indices, closest taps, decisions
    and the neighbourhood box bit-exact, history and sharpened screen within 1e-5 relative; the
    exact kernel's production form (the launch without the inspect image) writes the same bits.
c3  3840x2160, FP16 storage and arithmetic, all 64 frames, through the shipped FP16 kernel (no
    inspect image) as well as the inspected one: gpu_helpers.FP16OracleChain's bar.
c4  48 streams of mixed resolution (16x720p, 16x1080p, 8x1440p, 8x4K; per-stream seed, jitter phase
    and a pan of 0-24 px/frame), resolved by one graph-captured batched launch per round
    (taa_batch_plan_launch) for two rounds, FP32 and FP16 storage: every frame of every stream against
    the oracle (FP32 within 1e-5 relative; FP16 to the FP16 bar).
c5  a shortened served load: sessions drawn from the served resolution mix (720p 40%, 1080p 35%,
    1440p 15%, 4K 10%), frames arriving in Poisson-sized bursts that the server batches across
    sessions, sessions leaving and arriving; the sampled sessions' every frame against the oracle.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import taastar  # noqa: E402
from gpu_helpers import FP16OracleChain, FP32OracleChain, Sequence, assert_parity  # noqa: E402
from taastar import (ARITH_FP16, ARITH_FP32, FMT_RGBA16F, FMT_RGBA32F, BatchItem, Completion, ServerConfig,  # noqa: E402
                     SessionIO, check, lib)
from taastar.workloads import (RESOLUTIONS, SERVE_MIX, cas_for, decode, encode_color, encode_velocity,  # noqa: E402
                               scene_for, uniforms_for, workload)
from test_gpu_fp16 import run_fp16_sequence  # noqa: E402
from test_gpu_serve import d2h, h2d, io_of, new_ctx, open_session  # noqa: E402

C4_MIX = (("720p", 16), ("1080p", 16), ("1440p", 8), ("4k", 8))
C4_SEED = 0x7AA50004
C5_SEED = 0x7AA50005


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.timeout(900)
def test_c2_1080p_all_64_frames(oracle):
    print("case c2 1920x1080 fp32 x64 (inspected general form and shipped production form)", flush=True)
    seq = Sequence(workload("c2"), cas=True, debug=False, shipped=True)
    for f in range(64):
        assert_parity(seq.step(oracle))
        if f % 8 == 7:
            print(f"c2 frame {f} ok", flush=True)


@pytest.mark.timeout(1800)
def test_c3_4k_all_64_frames_fp16(oracle):
    print("case c3 3840x2160 fp16 x64 (inspected and shipped instantiations)", flush=True)
    rep = []
    run_fp16_sequence(oracle, "c3", 3840, 2160, 64, rep, shipped=True)
    print("c3 masked pixels per frame, max:", max(r["masked_pixels"] for r in rep),
          "full-frame PSNR min:", min(r["full_frame_psnr_db"] for r in rep), flush=True)


def c4_streams(fmt):
    """The 48 streams: resolution, scene (seed 0x7AA50004 ^ i, jitter phase i mod 8, pan of
    magnitude U[0, 24] px/frame in a uniform direction)."""
    w = workload("c3" if fmt == FMT_RGBA16F else "c2")
    out = []
    i = 0
    for cls, n in C4_MIX:
        W, H = RESOLUTIONS[cls]
        for _ in range(n):
            sc = scene_for(w, seed=C4_SEED ^ i, width=W, height=H, pan_mode=1, pan=24.0, jitter_phase=i % 8)
            out.append((W, H, sc))
            i += 1
    return w, out


@pytest.mark.timeout(1800)
@pytest.mark.parametrize("fmt,arith", [(FMT_RGBA32F, ARITH_FP32), (FMT_RGBA16F, ARITH_FP16)], ids=["fp32", "fp16"])
def test_c4_48_streams_batched_plan(oracle, fmt, arith):
    print(f"case c4 48 streams fmt={fmt} arith={arith}", flush=True)
    w, streams = c4_streams(fmt)
    elt = 8 if fmt == FMT_RGBA16F else 16
    npdt = np.float16 if fmt == FMT_RGBA16F else np.float32
    ctx = new_ctx()
    sess = [open_session(ctx, w, W, H) for W, H, _ in streams]
    scrs = [torch.zeros(H * W * elt, dtype=torch.uint8, device="cuda") for W, H, _ in streams]
    Ks = [cas_for(w, W, H) for W, H, _ in streams]
    Us = [taastar.Uniforms() for _ in streams]
    items = (BatchItem * len(streams))()
    for i in range(len(streams)):
        items[i].session = sess[i]
        items[i].uniforms = C.pointer(Us[i])
        items[i].cas = C.pointer(Ks[i])
        items[i].result_screen = scrs[i].data_ptr()
    if fmt == FMT_RGBA16F:
        chains = [FP16OracleChain(oracle, W, H, Ks[i], f"c4 stream {i}") for i, (W, H, _) in enumerate(streams)]
    else:
        chains = [FP32OracleChain(oracle, W, H, fmt, Ks[i], f"c4 stream {i}") for i, (W, H, _) in enumerate(streams)]
    plan = C.c_void_p()
    check(lib().taa_batch_plan_create(ctx, len(streams), items, arith, C.byref(plan)))
    for rnd in range(2):
        inputs = []
        for i, (W, H, sc) in enumerate(streams):
            col, dep, vel4, jit = sc.render(rnd)
            col_s, vel_s = encode_color(col, w["fmt"]), encode_velocity(vel4, w["vel"])
            io = io_of(sess[i])
            h2d(io.color, col_s)
            h2d(io.depth, dep)
            h2d(io.velocity, vel_s)
            U = uniforms_for(w, rnd, jit)
            C.memmove(C.addressof(Us[i]), C.addressof(U), C.sizeof(taastar.Uniforms))
            inputs.append((col_s, dep, vel_s))
        check(lib().taa_batch_plan_launch(plan, None))
        torch.cuda.synchronize()
        for i, (W, H, _) in enumerate(streams):
            col_s, dep, vel_s = inputs[i]
            g_scr = decode(scrs[i].cpu().numpy().view(npdt).reshape(H, W, 4))
            if fmt == FMT_RGBA16F:
                chains[i].check(Us[i], col_s, dep, vel_s, g_scr)
            else:
                g_his = decode(d2h(io_of(sess[i]).history, (H, W, 4), npdt))
                chains[i].check(Us[i], col_s, dep, vel_s, g_scr, g_his)
            inputs[i] = None
        print(f"c4 round {rnd}: 48 streams match the oracle", flush=True)
    check(lib().taa_batch_plan_destroy(plan))
    for s in sess:
        check(lib().taa_session_close(s))
    check(lib().taa_ctx_destroy(ctx))


class ServedSession:
    def __init__(self, srv, w, rng, j, sampled=False):
        cls = rng.choice(len(SERVE_MIX), p=[p for _, p in SERVE_MIX])
        self.W, self.H = RESOLUTIONS[SERVE_MIX[cls][0]]
        self.scene = scene_for(w, seed=C5_SEED ^ j, width=self.W, height=self.H, pan_mode=1, pan=24.0,
                               jitter_phase=j % 8)
        sid = C.c_int32()
        check(lib().taa_server_open(srv, self.W, self.H, C.byref(sid)))
        self.sid = sid.value
        self.frame = 0
        self.sampled = sampled
        self.pending = None  # inputs of the last submitted frame, checked once its result is read
        self.chain = None


@pytest.mark.timeout(1800)
def test_c5_served_load_sampled_sessions(oracle):
    """Shortened c5: 24 sessions (then churn), 40 bursts of Poisson(6) frames of distinct sessions
    each, pumped into batches of the server's continuous batching; one session of each resolution
    class sampled and checked frame by frame against the oracle."""
    print("case c5 served load, sampled sessions against the oracle", flush=True)
    w = workload("c3")
    rng = np.random.default_rng(C5_SEED)
    cfg = ServerConfig(device=0, hbm_budget_bytes=0, color_format=FMT_RGBA16F, arith=ARITH_FP16, max_batch=64,
                       max_inflight=3, cas_sharpness=0.5)
    srv = C.c_void_p()
    check(lib().taa_server_create(C.byref(cfg), C.byref(srv)))
    sessions = [ServedSession(srv, w, rng, j) for j in range(24)]
    next_j = 24
    seen = set()
    for s in sessions:  # sample the first session of each resolution class
        if (s.W, s.H) not in seen:
            seen.add((s.W, s.H))
            s.sampled = True
            s.chain = FP16OracleChain(oracle, s.W, s.H, cas_for(w, s.W, s.H), f"c5 session {s.sid} {s.W}x{s.H}")
    assert len(seen) >= 3, "the mix should draw at least three resolution classes"
    submitted = 0
    checked = 0
    batches_before = taastar.ServerStats()
    check(lib().taa_server_stats(srv, C.byref(batches_before)))

    def read_and_check(s):
        nonlocal checked
        if not s.pending:
            return
        scr = C.c_void_p()
        check(lib().taa_server_io(srv, s.sid, None, C.byref(scr)))
        g = decode(d2h(scr.value, (s.H, s.W, 4), np.float16))
        U, col_s, dep, vel_s = s.pending
        s.chain.check(U, col_s, dep, vel_s, g)
        s.pending = None
        checked += 1

    for burst in range(40):
        if burst and burst % 8 == 0:  # churn: an unsampled session leaves, a new one arrives
            leave = [s for s in sessions if not s.sampled]
            gone = leave[int(rng.integers(len(leave)))]
            check(lib().taa_server_close(srv, gone.sid))
            sessions.remove(gone)
            sessions.append(ServedSession(srv, w, rng, next_j))
            next_j += 1
        k = min(len(sessions), max(1, int(rng.poisson(6))))
        for s in [sessions[i] for i in rng.choice(len(sessions), size=k, replace=False)]:
            if s.sampled:
                read_and_check(s)  # waits for its previous frame, then holds it to the bar
            io = SessionIO()
            check(lib().taa_server_io(srv, s.sid, C.byref(io), None))
            col, dep, vel4, jit = s.scene.render(s.frame)
            col_s, vel_s = encode_color(col, w["fmt"]), encode_velocity(vel4, w["vel"])
            h2d(io.color, col_s)
            h2d(io.depth, dep)
            h2d(io.velocity, vel_s)
            U = uniforms_for(w, s.frame, jit)
            check(lib().taa_server_submit(srv, s.sid, C.byref(U), (s.sid << 16) | s.frame, -1.0))
            if s.sampled:
                s.pending = (U, col_s, dep, vel_s)
            s.frame += 1
            submitted += 1
        check(lib().taa_server_pump(srv))
    check(lib().taa_server_drain(srv))
    for s in sessions:
        if s.sampled:
            read_and_check(s)
    comp = (Completion * 512)()
    n = C.c_int32()
    check(lib().taa_server_poll(srv, comp, 512, C.byref(n)))
    assert n.value == submitted and all(comp[i].status == 0 for i in range(n.value))
    st = taastar.ServerStats()
    check(lib().taa_server_stats(srv, C.byref(st)))
    assert st.frames == submitted and st.frames_failed == 0
    mean_batch = submitted / max(1, st.batches - batches_before.batches)
    print(f"c5: {submitted} frames in {st.batches} batches (mean {mean_batch:.2f}), {checked} sampled frames "
          f"checked against the oracle", flush=True)
    assert mean_batch > 1.5, "frames of different sessions should share batches"
    assert checked >= 12
    check(lib().taa_server_destroy(srv))
