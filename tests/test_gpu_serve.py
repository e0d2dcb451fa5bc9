"""Sessions, batches, graph plans and the continuous-batching server on the GPU.

Bar: a session's outputs are bit-identical whether it is resolved alone (taa_session_resolve),
in a batch of mixed resolutions (taa_batch_submit, taa_batch_plan_launch) or by the server in a
mix of sessions; the per-pixel arithmetic does not depend on what a frame is batched with.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import taastar  # noqa: E402
from taastar import (ARITH_FP16, ARITH_FP32, FMT_RGBA16F, FMT_RGBA32F, BatchItem, ServerConfig, ServeReport,  # noqa: E402
                     ServeSimConfig, SessionIO, Completion, check, lib)
from taastar._abi import TAA_ERR_BUDGET, TAA_ERR_INVALID_ARG  # noqa: E402
from taastar.workloads import (cas_for, encode_color, encode_velocity, scene_for, uniforms_for, workload)  # noqa: E402


@pytest.fixture(autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _hip_memcpy():
    """hipMemcpy of the HIP runtime libtaastar.so is linked against (dlsym through its handle)."""
    fn = lib().hipMemcpy
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    fn.restype = C.c_int
    return fn


def h2d(ptr, arr):
    a = np.ascontiguousarray(arr)
    torch.cuda.synchronize()
    assert _hip_memcpy()(C.c_void_p(ptr), C.c_void_p(a.ctypes.data), a.nbytes, 1) == 0  # host to device


def d2h(ptr, shape, dtype):
    out = np.empty(shape, dtype)
    torch.cuda.synchronize()
    assert _hip_memcpy()(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), out.nbytes, 2) == 0  # device to host
    return out


class Stream:
    """One seeded scene feeding a session's input slots."""

    def __init__(self, w, W, H, seed):
        self.w, self.W, self.H = w, W, H
        self.scene = scene_for(w, width=W, height=H, seed=seed)
        self.frames = [self.scene.render(f) for f in range(6)]

    def upload(self, io, f):
        col, dep, vel4, jit = self.frames[f]
        h2d(io.color, encode_color(col, self.w["fmt"]))
        h2d(io.depth, dep)
        h2d(io.velocity, encode_velocity(vel4, self.w["vel"]))
        return uniforms_for(self.w, f, jit)


def new_ctx():
    ctx = C.c_void_p()
    check(lib().taa_ctx_create(0, 0, C.byref(ctx)))
    return ctx


def open_session(ctx, w, W, H):
    s = C.c_void_p()
    check(lib().taa_session_open(ctx, W, H, w["fmt"], w["vel"], C.byref(s)))
    return s


def io_of(s):
    io = SessionIO()
    check(lib().taa_session_io(s, C.byref(io)))
    return io


SIZES = [(160, 90), (96, 200), (333, 121)]


def test_batch_refuses_aliasing_screens():
    """Two items writing one screen, or a screen inside a batched session's buffers, are refused
    before anything is launched."""
    w = workload("c2", fmt=FMT_RGBA16F, vel=1)
    ctx = new_ctx()
    sess = [open_session(ctx, w, 64, 32), open_session(ctx, w, 48, 40)]
    scr = [torch.zeros(64 * 40 * 8, dtype=torch.uint8, device="cuda") for _ in sess]
    U = taastar.Uniforms()
    items = (BatchItem * 2)()
    for i in range(2):
        items[i].session = sess[i]
        items[i].uniforms = C.pointer(U)
        items[i].cas = None
        items[i].result_screen = scr[i].data_ptr()
    for screens in ((scr[0].data_ptr(), scr[0].data_ptr()), (scr[0].data_ptr(), io_of(sess[0]).color),
                    (io_of(sess[1]).velocity, scr[1].data_ptr())):
        items[0].result_screen, items[1].result_screen = screens
        assert lib().taa_batch_submit(ctx, 2, items, ARITH_FP16, None) == TAA_ERR_INVALID_ARG
        assert b"result_screen" in lib().taa_last_error()
        plan = C.c_void_p()
        assert lib().taa_batch_plan_create(ctx, 2, items, ARITH_FP16, C.byref(plan)) == TAA_ERR_INVALID_ARG
    for s in sess:
        check(lib().taa_session_close(s))
    check(lib().taa_ctx_destroy(ctx))


@pytest.mark.parametrize("fmt,arith", [(FMT_RGBA32F, ARITH_FP32), (FMT_RGBA16F, ARITH_FP16), (FMT_RGBA16F, ARITH_FP32)])
def test_alone_equals_batched_and_graph(fmt, arith):
    w = workload("c2", fmt=fmt, vel=1 if fmt == FMT_RGBA16F else 0)
    ctx = new_ctx()
    streams = [Stream(w, W, H, 1000 + i) for i, (W, H) in enumerate(SIZES)]
    elt = 8 if fmt == FMT_RGBA16F else 16
    npdt = np.float16 if fmt == FMT_RGBA16F else np.float32
    # A: every stream alone, per-frame launches
    alone = []
    for st in streams:
        s = open_session(ctx, w, st.W, st.H)
        scr = torch.zeros(st.H * st.W * elt, dtype=torch.uint8, device="cuda")
        K = cas_for(w, st.W, st.H)
        outs = []
        for f in range(5):
            U = st.upload(io_of(s), f)
            check(lib().taa_session_resolve(s, C.byref(U), C.byref(K), C.c_void_p(scr.data_ptr()), arith, None))
            torch.cuda.synchronize()
            outs.append(scr.cpu().numpy().view(npdt).copy())
        alone.append(outs)
        check(lib().taa_session_close(s))
    # B: all streams in one batch per frame; C: the same through a graph-captured plan
    for use_plan in (False, True):
        sess = [open_session(ctx, w, st.W, st.H) for st in streams]
        scrs = [torch.zeros(st.H * st.W * elt, dtype=torch.uint8, device="cuda") for st in streams]
        Ks = [cas_for(w, st.W, st.H) for st in streams]
        Us = [taastar.Uniforms() for _ in streams]
        items = (BatchItem * len(streams))()
        for i in range(len(streams)):
            items[i].session = sess[i]
            items[i].uniforms = C.pointer(Us[i])
            items[i].cas = C.pointer(Ks[i])
            items[i].result_screen = scrs[i].data_ptr()
        plan = C.c_void_p()
        if use_plan:
            check(lib().taa_batch_plan_create(ctx, len(streams), items, arith, C.byref(plan)))
        for f in range(5):
            for i, st in enumerate(streams):
                C.memmove(C.addressof(Us[i]), C.addressof(st.upload(io_of(sess[i]), f)), C.sizeof(taastar.Uniforms))
            if use_plan:
                check(lib().taa_batch_plan_launch(plan, None))
            else:
                check(lib().taa_batch_submit(ctx, len(streams), items, arith, None))
            torch.cuda.synchronize()
            for i in range(len(streams)):
                got = scrs[i].cpu().numpy().view(npdt)
                assert np.array_equal(got, alone[i][f]), f"plan={use_plan} stream {i} frame {f}"
        if use_plan:
            check(lib().taa_batch_plan_destroy(plan))
        for s in sess:
            check(lib().taa_session_close(s))
    check(lib().taa_ctx_destroy(ctx))


def test_server_mix_equals_alone():
    """Continuous batching: frames of one session give the same bits alone and in a served mix."""
    w = workload("c3", width=256, height=144)
    cfg = ServerConfig(device=0, hbm_budget_bytes=0, color_format=FMT_RGBA16F, arith=ARITH_FP16, max_batch=64,
                       max_inflight=3, cas_sharpness=0.5)
    results = {}
    for mix in (False, True):
        srv = C.c_void_p()
        check(lib().taa_server_create(C.byref(cfg), C.byref(srv)))
        probe = Stream(w, 200, 120, 77)
        others = [Stream(w, W, H, 900 + i) for i, (W, H) in enumerate([(320, 180), (128, 128), (256, 144)])] if mix else []
        sids = []
        for st in [probe] + others:
            sid = C.c_int32()
            check(lib().taa_server_open(srv, st.W, st.H, C.byref(sid)))
            sids.append(sid.value)
        outs = []
        for f in range(5):
            for st, sid in zip([probe] + others, sids):
                io = SessionIO()
                check(lib().taa_server_io(srv, sid, C.byref(io), None))
                U = st.upload(io, f)
                check(lib().taa_server_submit(srv, sid, C.byref(U), f, -1.0))
            check(lib().taa_server_drain(srv))
            scr = C.c_void_p()
            check(lib().taa_server_io(srv, sids[0], None, C.byref(scr)))
            outs.append(d2h(scr.value, (120, 200, 4), np.float16))
            comp = (Completion * 16)()
            n = C.c_int32()
            check(lib().taa_server_poll(srv, comp, 16, C.byref(n)))
            assert n.value == len(sids)
            assert all(comp[i].t_done_us >= comp[i].t_submit_us for i in range(n.value))
        results[mix] = outs
        check(lib().taa_server_destroy(srv))
    for f in range(5):
        assert np.array_equal(results[False][f], results[True][f]), f"frame {f}"


def test_server_admission_budget():
    cfg = ServerConfig(device=0, hbm_budget_bytes=3 * lib().taa_session_bytes(1280, 720, 1, 1) + (1 << 20) * 60,
                       color_format=FMT_RGBA16F, arith=ARITH_FP16, max_batch=8, max_inflight=2, cas_sharpness=-1.0)
    srv = C.c_void_p()
    check(lib().taa_server_create(C.byref(cfg), C.byref(srv)))
    sid = C.c_int32()
    opened = 0
    for _ in range(3):
        if lib().taa_server_open(srv, 1280, 720, C.byref(sid)) == 0:
            opened += 1
    assert opened == 3
    assert lib().taa_server_open(srv, 3840, 2160, C.byref(sid)) == TAA_ERR_BUDGET
    # closing one frees room for another of the same size (recycled without reallocation)
    check(lib().taa_server_close(srv, 1))
    assert lib().taa_server_open(srv, 1280, 720, C.byref(sid)) == 0
    st = taastar.ServerStats()
    check(lib().taa_server_stats(srv, C.byref(st)))
    assert st.admissions_refused == 1 and st.hbm_used <= st.hbm_budget
    check(lib().taa_server_destroy(srv))


def test_simulation_smoke():
    cfg = ServeSimConfig()
    cfg.device, cfg.seed, cfg.mode, cfg.sessions = 0, 7, 0, 12
    cfg.n_classes = 2
    cfg.class_w[0], cfg.class_h[0], cfg.class_p[0] = 640, 360, 0.7
    cfg.class_w[1], cfg.class_h[1], cfg.class_p[1] = 1280, 720, 0.3
    cfg.arrival_rate, cfg.churn_per_s, cfg.duration_s = 1500.0, 0.5, 0.5
    cfg.color_format, cfg.arith, cfg.cas_sharpness = FMT_RGBA16F, ARITH_FP16, 0.5
    cfg.max_batch, cfg.max_inflight, cfg.warmup_frames = 64, 3, 1
    rep = ServeReport()
    check(lib().taa_serve_simulate(C.byref(cfg), C.byref(rep)))
    d = rep.as_dict()
    print(d, flush=True)
    assert rep.frames > 200 and 0 < rep.lat_p50_us <= rep.lat_p99_us
    assert rep.offered_fps > 1000 and abs(rep.achieved_fps - rep.offered_fps) < 0.1 * rep.offered_fps


@pytest.mark.parametrize("mode", [2, 3, 4])
def test_simulation_per_frame_and_batched_modes(mode):
    """Per-frame (async / waited) and graph-batched rounds over an exact 3:1 class mix."""
    cfg = ServeSimConfig()
    cfg.device, cfg.seed, cfg.mode, cfg.sessions, cfg.exact_mix = 0, 5, mode, 8, 1
    cfg.n_classes = 2
    cfg.class_w[0], cfg.class_h[0], cfg.class_p[0] = 320, 180, 0.75
    cfg.class_w[1], cfg.class_h[1], cfg.class_p[1] = 640, 360, 0.25
    cfg.duration_s = 0.3
    cfg.color_format, cfg.arith, cfg.cas_sharpness = FMT_RGBA16F, ARITH_FP16, 0.5
    cfg.max_batch, cfg.max_inflight, cfg.warmup_frames = 16, 2, 1
    rep = ServeReport()
    check(lib().taa_serve_simulate(C.byref(cfg), C.byref(rep)))
    print(mode, rep.as_dict(), flush=True)
    assert rep.sessions_end == 8 and rep.frames > 0 and rep.frames % 8 == 0
    # 6 sessions of 320x180 and 2 of 640x360 per round of every session
    assert rep.pixels == (rep.frames // 8) * (6 * 320 * 180 + 2 * 640 * 360)
    if mode == 3:
        assert 0 < rep.lat_p50_us <= rep.lat_p99_us
    if mode == 4:
        assert rep.mean_batch == 8 and rep.lat_mean_us > 0


def test_server_refuses_bad_frames_and_keeps_serving():
    """A frame whose uniforms a batched launch refuses (split screen, mRayTraceAugment without the
    mask images, an unknown history filter) is refused by taa_server_submit with that status and
    never joins a batch: the other sessions' frames complete with status 0 and closing returns."""
    from taastar._abi import TAA_ERR_UNSUPPORTED
    w = workload("c3", width=160, height=96)
    cfg = ServerConfig(device=0, hbm_budget_bytes=0, color_format=FMT_RGBA16F, arith=ARITH_FP16, max_batch=8,
                       max_inflight=2, cas_sharpness=0.5)
    srv = C.c_void_p()
    check(lib().taa_server_create(C.byref(cfg), C.byref(srv)))
    streams = [Stream(w, 160, 96, 300 + i) for i in range(3)]
    sids = []
    for st in streams:
        sid = C.c_int32()
        check(lib().taa_server_open(srv, st.W, st.H, C.byref(sid)))
        sids.append(sid.value)
    for f in range(3):
        for k, (st, sid) in enumerate(zip(streams, sids)):
            io = SessionIO()
            check(lib().taa_server_io(srv, sid, C.byref(io), None))
            U = st.upload(io, f)
            if k == 1:
                bad = [("splitScreen", 1, TAA_ERR_UNSUPPORTED), ("rta", 1, TAA_ERR_UNSUPPORTED),
                       ("interp", 5, TAA_ERR_INVALID_ARG)][f]
                Ub = taastar.Uniforms.from_buffer_copy(U)
                if bad[0] == "splitScreen":
                    Ub.splitScreen = 1
                elif bad[0] == "rta":
                    Ub.param[0].mRayTraceAugment = 1
                else:
                    Ub.param[0].mInterpolationMode = bad[1]
                assert lib().taa_server_submit(srv, sid, C.byref(Ub), 100 + f, -1.0) == bad[2], bad
            check(lib().taa_server_submit(srv, sid, C.byref(U), f, -1.0))
        check(lib().taa_server_drain(srv))
        comp = (Completion * 16)()
        n = C.c_int32()
        check(lib().taa_server_poll(srv, comp, 16, C.byref(n)))
        assert n.value == 3 and all(comp[i].status == 0 and comp[i].tag == f for i in range(3))
    st = taastar.ServerStats()
    check(lib().taa_server_stats(srv, C.byref(st)))
    assert st.frames == 9 and st.frames_failed == 0
    for sid in sids:
        check(lib().taa_server_close(srv, sid))
    check(lib().taa_server_destroy(srv))


def test_server_io_waits_for_the_sessions_frames():
    """taa_server_io hands out a session's slots only once its queued and running frames are done
    (so the next frame's inputs cannot overwrite what a running resolve reads)."""
    w = workload("c3", width=1280, height=720)
    cfg = ServerConfig(device=0, hbm_budget_bytes=0, color_format=FMT_RGBA16F, arith=ARITH_FP16, max_batch=8,
                       max_inflight=2, cas_sharpness=0.5)
    srv = C.c_void_p()
    check(lib().taa_server_create(C.byref(cfg), C.byref(srv)))
    st = Stream(w, 1280, 720, 41)
    sid = C.c_int32()
    check(lib().taa_server_open(srv, 1280, 720, C.byref(sid)))
    io = SessionIO()
    check(lib().taa_server_io(srv, sid.value, C.byref(io), None))
    U = st.upload(io, 0)
    for f in range(4):  # four queued frames of one session
        check(lib().taa_server_submit(srv, sid.value, C.byref(U), f, -1.0))
    check(lib().taa_server_pump(srv))
    check(lib().taa_server_io(srv, sid.value, C.byref(io), None))
    stats = taastar.ServerStats()
    check(lib().taa_server_stats(srv, C.byref(stats)))
    assert stats.frames == 4 and io.frames_resolved == 4
    check(lib().taa_server_destroy(srv))
