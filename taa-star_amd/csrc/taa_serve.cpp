// Continuous-batching server and load simulator (include/taastar/taa_serve.h).
//
// One host thread drives the GPU: taa_server_pump() harvests finished batches (event queries) and
// launches every queued frame whose session is idle into a free in-flight slot (its own stream),
// as one batched resolve (one workgroup per tile of all its frames). Completion times are GPU timestamps (event elapsed time from
// a reference event whose host time is known), so latency does not include host polling delays.
// Closed sessions go to a per-resolution recycle list instead of hipFree, so churn never stalls the
// device; their bytes stay counted against the budget until reused or released.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <random>
#include <vector>

#include "taa_internal.hpp"
#include "taastar/taa_abi.h"
#include "taastar/taa_scene.h"
#include "taastar/taa_serve.h"

namespace {

double host_us() {
    using namespace std::chrono;
    return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

struct Req {
    uint64_t tag;
    double t_submit;
    taa_uniforms_t U;
};

struct Sess {
    taa_session_t* s = nullptr;
    void* screen = nullptr;
    int w = 0, h = 0;
    uint64_t bytes = 0;  // session + screen
    taa_cas_consts_t cas;
    std::deque<Req> q;
    bool busy = false;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev_start = nullptr, ev_done = nullptr;
    bool busy = false;
    std::vector<int> sids;
    std::vector<Req> reqs;
    uint64_t pixels = 0;
};

struct Recycled {
    taa_session_t* s;
    void* screen;
    uint64_t bytes;
};

size_t color_bpp(int fmt) { return fmt == TAA_FMT_RGBA16F ? 8 : 16; }

}  // namespace

struct taa_server {
    taa_server_config_t cfg;
    taa_ctx_t* ctx = nullptr;
    bool cas_on = false;
    std::map<int, Sess> sess;
    std::map<std::pair<int, int>, std::vector<Recycled>> recycle;
    uint64_t budget = 0, used = 0;  // sessions + screens, open or recycled
    int next_id = 1;
    std::vector<Slot> slots;
    std::deque<int> ready;
    std::vector<taa_completion_t> done;
    hipEvent_t ev0 = nullptr;
    double t0_host = 0;
    taa_server_stats_t st;
    std::vector<std::pair<double, double>> spans;  // GPU [start, done] of batches (host clock)
};

namespace {

uint64_t server_session_bytes(const taa_server* s, int w, int h) {
    const int vf = s->cfg.color_format == TAA_FMT_RGBA16F ? TAA_VEL_RG16F : TAA_VEL_RG32F;
    return taa_session_bytes(w, h, s->cfg.color_format, vf) + (((uint64_t)w * h * color_bpp(s->cfg.color_format) + 255) & ~255ull);
}

double gpu_time_us(taa_server* s, hipEvent_t e) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, s->ev0, e);
    return s->t0_host + 1000.0 * (double)ms;
}

void harvest(taa_server* s) {
    for (Slot& sl : s->slots) {
        if (!sl.busy || hipEventQuery(sl.ev_done) != hipSuccess) continue;
        const double t_start = gpu_time_us(s, sl.ev_start), t_done = gpu_time_us(s, sl.ev_done);
        s->spans.emplace_back(t_start, t_done);
        s->st.busy_us += t_done - t_start;
        for (size_t i = 0; i < sl.sids.size(); ++i) {
            taa_completion_t c;
            std::memset(&c, 0, sizeof(c));
            c.session_id = sl.sids[i];
            c.tag = sl.reqs[i].tag;
            c.t_submit_us = sl.reqs[i].t_submit;
            c.t_done_us = t_done;
            s->done.push_back(c);
            auto it = s->sess.find(sl.sids[i]);
            if (it != s->sess.end()) {
                it->second.busy = false;
                if (!it->second.q.empty()) s->ready.push_back(sl.sids[i]);
            }
        }
        s->st.frames += sl.sids.size();
        s->st.pixels += sl.pixels;
        sl.busy = false;
        sl.sids.clear();
        sl.reqs.clear();
    }
}

int launch_ready(taa_server* s) {
    for (Slot& sl : s->slots) {
        if (s->ready.empty()) break;
        if (sl.busy) continue;
        std::vector<taa_batch_item_t> items;
        sl.pixels = 0;
        while (!s->ready.empty() && (int)items.size() < s->cfg.max_batch) {
            const int sid = s->ready.front();
            s->ready.pop_front();
            auto found = s->sess.find(sid);
            if (found == s->sess.end()) continue;
            Sess& se = found->second;
            if (se.busy || se.q.empty()) continue;
            sl.reqs.push_back(se.q.front());
            se.q.pop_front();
            se.busy = true;
            sl.sids.push_back(sid);
            taa_batch_item_t it;
            it.session = se.s;
            it.uniforms = &sl.reqs.back().U;
            it.cas = s->cas_on ? &se.cas : nullptr;
            it.result_screen = se.screen;
            items.push_back(it);
            sl.pixels += (uint64_t)se.w * se.h;
        }
        if (items.empty()) continue;
        // sl.reqs may have reallocated while filling: point the items at their final storage
        for (size_t i = 0; i < items.size(); ++i) items[i].uniforms = &sl.reqs[i].U;
        hipEventRecord(sl.ev_start, sl.st);
        int rc = taa_batch_submit(s->ctx, (int)items.size(), items.data(), s->cfg.arith, sl.st);
        if (rc) {
            // nothing of this batch runs: its frames complete with the error, their sessions are
            // free again (and queued behind the failure if more frames wait), the slot is empty
            const double now = host_us();
            for (size_t i = 0; i < sl.sids.size(); ++i) {
                taa_completion_t c;
                std::memset(&c, 0, sizeof(c));
                c.session_id = sl.sids[i];
                c.status = rc;
                c.tag = sl.reqs[i].tag;
                c.t_submit_us = sl.reqs[i].t_submit;
                c.t_done_us = now;
                s->done.push_back(c);
                auto it = s->sess.find(sl.sids[i]);
                if (it != s->sess.end()) {
                    it->second.busy = false;
                    if (!it->second.q.empty()) s->ready.push_back(sl.sids[i]);
                }
            }
            s->st.frames_failed += sl.sids.size();
            sl.sids.clear();
            sl.reqs.clear();
            sl.pixels = 0;
            return rc;
        }
        hipEventRecord(sl.ev_done, sl.st);
        sl.busy = true;
        s->st.batches++;
    }
    return TAA_OK;
}

}  // namespace

extern "C" {

double taa_server_now_us(void) { return host_us(); }

int32_t taa_server_create(const taa_server_config_t* cfg, taa_server_t** out) {
    if (!cfg || !out) return TAA_ERR_INVALID_ARG;
    *out = nullptr;
    if (cfg->max_batch < 1 || cfg->max_batch > 1024 || cfg->max_inflight < 1 || cfg->max_inflight > 8)
        return TAA_ERR_INVALID_ARG;
    if (cfg->color_format != TAA_FMT_RGBA32F && cfg->color_format != TAA_FMT_RGBA16F) return TAA_ERR_INVALID_ARG;
    taa_server* s = new (std::nothrow) taa_server();
    if (!s) return TAA_ERR_OUT_OF_MEMORY;
    s->cfg = *cfg;
    int rc = taa_ctx_create(cfg->device, 0, &s->ctx);
    if (rc) { delete s; return rc; }
    uint64_t ctx_budget = 0;
    taa_ctx_usage(s->ctx, nullptr, &ctx_budget);
    s->budget = cfg->hbm_budget_bytes ? std::min<uint64_t>(cfg->hbm_budget_bytes, ctx_budget) : ctx_budget;
    s->cas_on = cfg->cas_sharpness >= 0.0f;
    s->slots.resize(cfg->max_inflight);
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&sl.ev_start) != hipSuccess ||
            hipEventCreate(&sl.ev_done) != hipSuccess) {
            taa_server_destroy(s);
            return TAA_ERR_HIP;
        }
    }
    hipEventCreate(&s->ev0);
    hipEventRecord(s->ev0, nullptr);
    hipEventSynchronize(s->ev0);
    s->t0_host = host_us();
    std::memset(&s->st, 0, sizeof(s->st));
    *out = s;
    return TAA_OK;
}

int32_t taa_server_destroy(taa_server_t* s) {
    if (!s) return TAA_ERR_INVALID_ARG;
    taa_server_drain(s);
    for (auto& kv : s->sess) {
        taa_session_close(kv.second.s);
        hipFree(kv.second.screen);
    }
    for (auto& kv : s->recycle)
        for (Recycled& r : kv.second) {
            taa_session_close(r.s);
            hipFree(r.screen);
        }
    for (Slot& sl : s->slots) {
        if (sl.ev_start) hipEventDestroy(sl.ev_start);
        if (sl.ev_done) hipEventDestroy(sl.ev_done);
        if (sl.st) hipStreamDestroy(sl.st);
    }
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ctx) taa_ctx_destroy(s->ctx);
    delete s;
    return TAA_OK;
}

int32_t taa_server_open(taa_server_t* s, int32_t w, int32_t h, int32_t* sid) {
    if (!s || !sid || w <= 0 || h <= 0) return TAA_ERR_INVALID_ARG;
    Sess se;
    se.w = w;
    se.h = h;
    auto& pool = s->recycle[{w, h}];
    if (!pool.empty()) {  // reuse a closed session of this resolution: already counted in `used`
        Recycled r = pool.back();
        pool.pop_back();
        se.s = r.s;
        se.screen = r.screen;
        se.bytes = r.bytes;
        taa_session_reset(se.s);
    } else {
        const uint64_t need = server_session_bytes(s, w, h);
        if (s->used + need > s->budget) {
            // release recycled sessions of other resolutions before refusing
            for (auto& kv : s->recycle) {
                while (!kv.second.empty() && s->used + need > s->budget) {
                    Recycled r = kv.second.back();
                    kv.second.pop_back();
                    taa_session_close(r.s);
                    hipFree(r.screen);
                    s->used -= r.bytes;
                }
            }
            if (s->used + need > s->budget) {
                s->st.admissions_refused++;
                return TAA_ERR_BUDGET;
            }
        }
        const int vf = s->cfg.color_format == TAA_FMT_RGBA16F ? TAA_VEL_RG16F : TAA_VEL_RG32F;
        int rc = taa_session_open(s->ctx, w, h, s->cfg.color_format, vf, &se.s);
        if (rc) {
            if (rc == TAA_ERR_BUDGET) s->st.admissions_refused++;
            return rc;
        }
        if (hipMalloc(&se.screen, (size_t)w * h * color_bpp(s->cfg.color_format)) != hipSuccess) {
            taa_session_close(se.s);
            return TAA_ERR_OUT_OF_MEMORY;
        }
        se.bytes = need;
        s->used += need;
    }
    if (s->cas_on) taa_cas_setup(se.cas.const0, se.cas.const1, s->cfg.cas_sharpness, (float)w, (float)h, (float)w, (float)h);
    const int id = s->next_id++;
    s->sess[id] = se;
    s->st.sessions_open = s->sess.size();
    *sid = id;
    return TAA_OK;
}

int32_t taa_server_close(taa_server_t* s, int32_t sid) {
    if (!s) return TAA_ERR_INVALID_ARG;
    auto it = s->sess.find(sid);
    if (it == s->sess.end()) return TAA_ERR_INVALID_ARG;
    while (it->second.busy || !it->second.q.empty()) {
        int rc = taa_server_pump(s);
        if (rc) return rc;
    }
    s->recycle[{it->second.w, it->second.h}].push_back(Recycled{it->second.s, it->second.screen, it->second.bytes});
    s->sess.erase(it);
    s->st.sessions_open = s->sess.size();
    return TAA_OK;
}

int32_t taa_server_io(taa_server_t* s, int32_t sid, taa_session_io_t* io, void** screen) {
    if (!s) return TAA_ERR_INVALID_ARG;
    auto it = s->sess.find(sid);
    if (it == s->sess.end()) return TAA_ERR_INVALID_ARG;
    // the slots are the ones the session's queued and running frames read: hand them out only
    // once every such frame has completed, so writes through them cannot race a resolve
    while (it->second.busy || !it->second.q.empty()) {
        int rc = taa_server_pump(s);
        if (rc) return rc;
    }
    if (screen) *screen = it->second.screen;
    return io ? taa_session_io(it->second.s, io) : TAA_OK;
}

int32_t taa_server_submit(taa_server_t* s, int32_t sid, const taa_uniforms_t* U, uint64_t tag, double arrival_us) {
    if (!s || !U) return TAA_ERR_INVALID_ARG;
    auto it = s->sess.find(sid);
    if (it == s->sess.end()) return TAA_ERR_INVALID_ARG;
    // refuse here what the batched launch would refuse, so one bad frame cannot fail a batch of
    // other sessions' frames
    if (int rc = taastar::check_batch_uniforms(U)) return rc;
    Sess& se = it->second;
    Req r;
    r.tag = tag;
    r.t_submit = arrival_us >= 0 ? arrival_us : host_us();
    r.U = *U;
    const bool was_idle = !se.busy && se.q.empty();
    se.q.push_back(r);
    if (was_idle) s->ready.push_back(sid);
    return TAA_OK;
}

int32_t taa_server_pump(taa_server_t* s) {
    if (!s) return TAA_ERR_INVALID_ARG;
    harvest(s);
    return launch_ready(s);
}

int32_t taa_server_poll(taa_server_t* s, taa_completion_t* out, int32_t max, int32_t* n_out) {
    if (!s || !n_out || (max > 0 && !out)) return TAA_ERR_INVALID_ARG;
    const int n = std::min<int>(max, (int)s->done.size());
    for (int i = 0; i < n; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + n);
    *n_out = n;
    return TAA_OK;
}

int32_t taa_server_drain(taa_server_t* s) {
    if (!s) return TAA_ERR_INVALID_ARG;
    for (;;) {
        int rc = taa_server_pump(s);
        if (rc) return rc;
        bool busy = !s->ready.empty();
        for (const Slot& sl : s->slots) busy = busy || sl.busy;
        if (!busy) return TAA_OK;
    }
}

int32_t taa_server_stats(const taa_server_t* s, taa_server_stats_t* out) {
    if (!s || !out) return TAA_ERR_INVALID_ARG;
    *out = s->st;
    out->hbm_used = s->used;
    out->hbm_budget = s->budget;
    return TAA_OK;
}

}  // extern "C"

// ---- load simulation ------------------------------------------------------------------------

namespace {

uint16_t f32_to_f16(float f) {  // round to nearest even, overflow to Inf
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const int32_t exp = (int32_t)((u >> 23) & 0xff);
    uint32_t man = u & 0x7fffffu;
    if (exp == 0xff) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
    int32_t e = exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        const int shift = 14 - e;
        uint32_t half = man >> shift;
        const uint32_t rem = man & ((1u << shift) - 1), mid = 1u << (shift - 1);
        if (rem > mid || (rem == mid && (half & 1u))) half++;
        return (uint16_t)(sign | half);
    }
    uint32_t half = ((uint32_t)e << 10) | (man >> 13);
    const uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) half++;
    return (uint16_t)(sign | half);
}

struct PoolFrame {
    int w, h;
    void* color = nullptr;
    float* depth = nullptr;
    void* vel = nullptr;
    float jitter[8][2];
};

int make_pool_frame(PoolFrame& pf, int w, int h, int fmt, uint64_t seed, int pan_mode = 0, float pan = -1.0f,
                    int jitter_phase = 0) {
    taa_scene_config_t sc;
    taa_scene_defaults(&sc);
    sc.width = w;
    sc.height = h;
    if (pan >= 0.0f) {
        sc.pan_mode = pan_mode;
        sc.pan = pan;
    }
    sc.jitter_phase = jitter_phase;
    sc.num_polygons = std::max(1, (int)std::lround(2000.0 * w * h / (1920.0 * 1080.0)));
    sc.seed = seed;
    sc.color_max = fmt == TAA_FMT_RGBA16F ? 32768.0f : 0.0f;
    taa_scene_t* scene = nullptr;
    if (taa_scene_create(&sc, &scene)) return TAA_ERR_INVALID_ARG;
    const size_t px = (size_t)w * h;
    std::vector<float> col(px * 4), dep(px), vel(px * 4);
    float jit[2];
    taa_scene_render(scene, 1, col.data(), dep.data(), vel.data(), jit, 0);
    taa_jitter_config_t jc;
    std::memset(&jc, 0, sizeof(jc));
    jc.sample_distribution = 2;
    jc.fixed_index = -1;
    jc.extra_scale = 1.0f;
    for (int k = 0; k < 8; ++k)
        taa_jitter_offset(&jc, (uint64_t)(k + jitter_phase), (uint32_t)w, (uint32_t)h, pf.jitter[k], nullptr);
    taa_scene_destroy(scene);
    pf.w = w;
    pf.h = h;
    const bool s16 = fmt == TAA_FMT_RGBA16F;
    std::vector<uint16_t> col16, vel16;
    std::vector<float> vel32;
    const void* hc = col.data();
    const void* hv = nullptr;
    if (s16) {
        col16.resize(px * 4);
        for (size_t i = 0; i < px * 4; ++i) col16[i] = f32_to_f16(col[i]);
        vel16.resize(px * 2);
        for (size_t i = 0; i < px; ++i) { vel16[2 * i] = f32_to_f16(vel[4 * i]); vel16[2 * i + 1] = f32_to_f16(vel[4 * i + 1]); }
        hc = col16.data();
        hv = vel16.data();
    } else {
        vel32.resize(px * 2);
        for (size_t i = 0; i < px; ++i) { vel32[2 * i] = vel[4 * i]; vel32[2 * i + 1] = vel[4 * i + 1]; }
        hv = vel32.data();
    }
    const size_t cb = px * color_bpp(fmt), vb = px * (s16 ? 4 : 8);
    if (hipMalloc(&pf.color, cb) != hipSuccess || hipMalloc((void**)&pf.depth, px * 4) != hipSuccess || hipMalloc(&pf.vel, vb) != hipSuccess)
        return TAA_ERR_OUT_OF_MEMORY;
    hipMemcpy(pf.color, hc, cb, hipMemcpyHostToDevice);
    hipMemcpy(pf.depth, dep.data(), px * 4, hipMemcpyHostToDevice);
    hipMemcpy(pf.vel, hv, vb, hipMemcpyHostToDevice);
    return TAA_OK;
}

taa_params_t sim_params() {
    taa_params_t p;
    taa_params_defaults(&p);
    p.mVarianceClipping = 1;
    p.mVarClipGamma = 1.0f;
    p.mColorClampingOrClipping = 2;
    p.mInterpolationMode = 2;
    p.mAlpha = 0.1f;
    p.mUseVelocityVectors = 2;
    p.mVelocitySampleMode = 2;
    p.mUseYCoCg = 1;
    p.mDepthCulling = 1;
    p.mLumaWeightingLottes = 1;
    p.mRejectOutside = 1;
    return p;
}

void sim_uniforms(taa_uniforms_t& U, const float (&jitter)[8][2], uint64_t frame) {
    std::memset(&U, 0, sizeof(U));
    U.param[0] = sim_params();
    U.param[1] = U.param[0];
    for (int i : {0, 5, 10, 15}) U.mHistoryViewProjMatrix[i] = U.mInverseViewProjMatrix[i] = 1.0f;
    U.mJitterNdc[0] = jitter[frame % 8][0];
    U.mJitterNdc[1] = jitter[frame % 8][1];
    U.mCamNearPlane = 0.1f;
    U.mCamFarPlane = 100.0f;
    U.pad1 = 0.01f;
}

double percentile(std::vector<double>& v, double q) {
    if (v.empty()) return 0.0;
    const size_t k = std::min(v.size() - 1, (size_t)std::floor(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return v[k];
}

double union_length(std::vector<std::pair<double, double>> iv, double lo, double hi) {
    std::sort(iv.begin(), iv.end());
    double total = 0, cs = 0, ce = -1e300;
    for (auto& p : iv) {
        double a = std::max(p.first, lo), b = std::min(p.second, hi);
        if (b <= a) continue;
        if (a > ce) {
            if (ce > cs) total += ce - cs;
            cs = a;
            ce = b;
        } else {
            ce = std::max(ce, b);
        }
    }
    if (ce > cs) total += ce - cs;
    return total;
}

struct SimSession {
    int sid;
    int cls;
    int own = -1;  // index of the session's own input frame (per-session inputs), else -1
    uint64_t frame = 0;
    bool departing = false;
    float jitter[8][2];
};

}  // namespace

extern "C" int32_t taa_serve_simulate(const taa_serve_sim_config_t* cfg, taa_serve_report_t* out) {
    if (!cfg || !out || cfg->n_classes < 1 || cfg->n_classes > 8 || cfg->sessions < 1) return TAA_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    hipSetDevice(cfg->device);
    std::mt19937_64 rng(cfg->seed);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    // class distribution
    std::vector<double> cdf(cfg->n_classes);
    double acc = 0;
    for (int i = 0; i < cfg->n_classes; ++i) cdf[i] = (acc += cfg->class_p[i]);
    auto draw_class = [&]() {
        const double r = U01(rng) * acc;
        for (int i = 0; i < cfg->n_classes; ++i)
            if (r < cdf[i]) return i;
        return cfg->n_classes - 1;
    };
    std::vector<PoolFrame> pool(cfg->n_classes);
    for (int i = 0; i < cfg->n_classes; ++i) {
        int rc = make_pool_frame(pool[i], cfg->class_w[i], cfg->class_h[i], cfg->color_format, cfg->seed ^ (0x7AA50005ull + i));
        if (rc) return rc;
    }
    // per-session inputs: every admitted session gets an input frame of its own (own scene, pan and
    // jitter phase), rendered before the measured run -- one per initial session plus a reserve per
    // class for the sessions churn admits; a departing session's frame goes back to its class's
    // free list, so the active sessions' inputs stay distinct and nothing is rendered while timing
    std::vector<PoolFrame> own;
    std::vector<std::vector<int>> own_free(cfg->n_classes);
    uint64_t own_rendered = 0;
    auto render_own = [&](int cls) -> int {
        PoolFrame pf;
        const float pan = (float)(U01(rng) * (double)std::max(0.0f, cfg->pan_max));
        const uint64_t k = own_rendered++;
        int r = make_pool_frame(pf, cfg->class_w[cls], cfg->class_h[cls], cfg->color_format,
                                cfg->seed ^ (0x7AA50004ull + k), 1, pan, (int)(k % 8));
        if (r) return r;
        own.push_back(pf);
        own_free[cls].push_back((int)own.size() - 1);
        return 0;
    };
    hipStream_t copy_st;
    hipStreamCreateWithFlags(&copy_st, hipStreamNonBlocking);

    taa_server_config_t sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.device = cfg->device;
    sc.hbm_budget_bytes = cfg->hbm_budget_bytes;
    sc.color_format = cfg->color_format;
    sc.arith = cfg->arith;
    sc.max_batch = cfg->max_batch > 0 ? cfg->max_batch : 256;
    sc.max_inflight = cfg->max_inflight > 0 ? cfg->max_inflight : 3;
    sc.cas_sharpness = cfg->cas_sharpness;
    taa_server_t* srv = nullptr;
    int rc = taa_server_create(&sc, &srv);
    if (rc) return rc;

    std::vector<SimSession> ss;
    auto open_one = [&](int cls) -> int {
        int sid = 0;
        out->admissions++;
        int r = taa_server_open(srv, cfg->class_w[cls], cfg->class_h[cls], &sid);
        if (r == TAA_ERR_BUDGET) { out->admissions_refused++; return 0; }
        if (r) return r;
        taa_session_io_t io;
        taa_server_io(srv, sid, &io, nullptr);
        int fi = -1;
        if (cfg->per_session_inputs) {  // a frame of its own (rendered now only if the reserve ran out)
            if (own_free[cls].empty()) {
                if (int r = render_own(cls)) return r;
            }
            fi = own_free[cls].back();
            own_free[cls].pop_back();
        }
        const PoolFrame& pf = fi >= 0 ? own[fi] : pool[cls];
        const size_t px = (size_t)pf.w * pf.h;
        hipMemcpyAsync(io.color, pf.color, px * color_bpp(cfg->color_format), hipMemcpyDeviceToDevice, copy_st);
        hipMemcpyAsync(io.depth, pf.depth, px * 4, hipMemcpyDeviceToDevice, copy_st);
        // the depth ring's other slot holds the previous depth: same content (a static first frame)
        hipMemcpyAsync(io.velocity, pf.vel, px * (cfg->color_format == TAA_FMT_RGBA16F ? 4 : 8), hipMemcpyDeviceToDevice, copy_st);
        hipStreamSynchronize(copy_st);
        SimSession x;
        x.sid = sid;
        x.cls = cls;
        x.own = fi;
        std::memcpy(x.jitter, pf.jitter, sizeof(x.jitter));
        ss.push_back(x);
        out->sessions_peak = std::max<int32_t>(out->sessions_peak, (int32_t)ss.size());
        return 0;
    };
    if (cfg->exact_mix) {  // largest-remainder apportionment of `sessions` over the classes
        std::vector<int> cnt(cfg->n_classes);
        std::vector<std::pair<double, int>> rem;
        int given = 0;
        for (int i = 0; i < cfg->n_classes; ++i) {
            const double q = cfg->sessions * cfg->class_p[i] / acc;
            cnt[i] = (int)std::floor(q);
            given += cnt[i];
            rem.emplace_back(q - cnt[i], i);
        }
        std::sort(rem.begin(), rem.end(), [](auto& a, auto& b) { return a.first > b.first; });
        for (int k = 0; given < cfg->sessions; ++k, ++given) cnt[rem[k % rem.size()].second]++;
        for (int i = 0; i < cfg->n_classes && rc == 0; ++i)
            for (int j = 0; j < cnt[i] && rc == 0; ++j) rc = open_one(i);
    } else {
        for (int i = 0; i < cfg->sessions; ++i) {
            rc = open_one(draw_class());
            if (rc) break;
        }
    }
    if (rc == 0 && cfg->per_session_inputs) {
        // the initial sessions rendered their frames as they were admitted; now the reserve for the
        // sessions churn admits: the expected number of churn events over the run, times 1.5,
        // split by the class mix, so that no frame is rendered while the run is measured
        const double events = cfg->churn_per_s * (double)cfg->sessions * cfg->duration_s * 1.5 + 4.0;
        for (int i = 0; i < cfg->n_classes && rc == 0; ++i)
            for (int j = 0; j < (int)std::ceil(events * cfg->class_p[i] / acc) && rc == 0; ++j) rc = render_own(i);
    }
    std::vector<double> lat;
    std::vector<taa_completion_t> comp(4096);
    uint64_t frames = 0, pixels = 0;
    auto find_sess = [&](int sid) -> SimSession* {
        for (auto& x : ss)
            if (x.sid == sid) return &x;
        return nullptr;
    };
    auto submit = [&](SimSession& x, double t) {
        taa_uniforms_t Uf;
        sim_uniforms(Uf, x.jitter, x.frame);
        taa_server_submit(srv, x.sid, &Uf, x.frame, t);
        x.frame++;
    };
    auto collect = [&](double t_lo, double t_hi, bool resubmit) {
        int n = 0;
        do {
            taa_server_poll(srv, comp.data(), (int)comp.size(), &n);
            for (int i = 0; i < n; ++i) {
                const taa_completion_t& c = comp[i];
                SimSession* x = find_sess(c.session_id);
                if (c.t_submit_us >= t_lo && c.t_submit_us <= t_hi) {
                    lat.push_back(c.t_done_us - c.t_submit_us);
                    frames++;
                    if (x) pixels += (uint64_t)cfg->class_w[x->cls] * cfg->class_h[x->cls];
                }
                if (resubmit && x && !x->departing) submit(*x, -1.0);
            }
        } while (n == (int)comp.size());
    };

    double t_begin = 0, t_end = 0;
    if (rc == 0 && (cfg->mode == 0 || cfg->mode == 1)) {
        // warm-up: every session resolves `warmup_frames` frames (includes its reset frame)
        for (int k = 0; k < std::max(1, cfg->warmup_frames); ++k) {
            for (auto& x : ss) submit(x, -1.0);
            taa_server_drain(srv);
        }
        int dummy;
        do taa_server_poll(srv, comp.data(), (int)comp.size(), &dummy); while (dummy);
        taa_server_stats_t st0;
        taa_server_stats(srv, &st0);
        t_begin = host_us();
        t_end = t_begin + cfg->duration_s * 1e6;
        if (cfg->mode == 1) {
            for (auto& x : ss) submit(x, -1.0);
            while (host_us() < t_end) {
                taa_server_pump(srv);
                collect(t_begin, t_end, true);
            }
        } else {
            std::exponential_distribution<double> gap(cfg->arrival_rate / 1e6);
            const double churn_rate = cfg->churn_per_s * (double)cfg->sessions / 1e6;  // events per us
            std::exponential_distribution<double> cgap(churn_rate > 0 ? churn_rate : 1.0);
            double next_arr = t_begin + gap(rng);
            double next_churn = churn_rate > 0 ? t_begin + cgap(rng) : 1e300;
            uint64_t offered = 0;
            while (host_us() < t_end) {
                const double now = host_us();
                while (next_arr <= now && next_arr < t_end) {
                    // a uniformly chosen active session receives the frame
                    std::vector<int> act;
                    for (size_t i = 0; i < ss.size(); ++i)
                        if (!ss[i].departing) act.push_back((int)i);
                    if (!act.empty()) {
                        SimSession& x = ss[act[(size_t)(U01(rng) * act.size()) % act.size()]];
                        submit(x, next_arr);
                        offered++;
                    }
                    next_arr += gap(rng);
                }
                while (next_churn <= now && next_churn < t_end) {
                    if (!ss.empty()) {  // one session departs, a new one (class from the mix) asks for admission
                        const size_t k = (size_t)(U01(rng) * ss.size()) % ss.size();
                        const int sid = ss[k].sid;
                        if (ss[k].own >= 0) own_free[ss[k].cls].push_back(ss[k].own);
                        ss.erase(ss.begin() + k);
                        taa_server_close(srv, sid);
                    }
                    open_one(draw_class());
                    next_churn += cgap(rng);
                }
                taa_server_pump(srv);
                collect(t_begin, t_end, false);
            }
            out->offered_fps = (double)offered / cfg->duration_s;
        }
        taa_server_drain(srv);
        collect(t_begin, t_end, false);
        taa_server_stats_t st1;
        taa_server_stats(srv, &st1);
        out->batches = st1.batches - st0.batches;
        out->mean_batch = out->batches ? (double)(st1.frames - st0.frames) / (double)out->batches : 0.0;
        out->gpu_busy_frac = union_length(srv->spans, t_begin, t_end) / (t_end - t_begin);
        out->hbm_peak_bytes = st1.hbm_used;
    } else if (rc == 0 && cfg->mode == 4) {
        // every session advances one frame per graph launch (descriptor upload + batched kernel)
        hipStream_t st;
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        std::vector<taa_batch_item_t> items(ss.size());
        std::vector<taa_uniforms_t> us(ss.size());
        std::vector<taa_cas_consts_t> cs(ss.size());
        uint64_t round_px = 0;
        for (size_t i = 0; i < ss.size(); ++i) {
            auto it = srv->sess.find(ss[i].sid);
            cs[i] = it->second.cas;
            items[i].session = it->second.s;
            items[i].uniforms = &us[i];
            items[i].cas = srv->cas_on ? &cs[i] : nullptr;
            items[i].result_screen = it->second.screen;
            round_px += (uint64_t)cfg->class_w[ss[i].cls] * cfg->class_h[ss[i].cls];
        }
        taa_batch_plan_t* plan = nullptr;
        rc = taa_batch_plan_create(srv->ctx, (int)items.size(), items.data(), cfg->arith, &plan);
        auto refresh = [&]() {
            for (size_t i = 0; i < ss.size(); ++i) sim_uniforms(us[i], ss[i].jitter, ss[i].frame++);
        };
        for (int k = 0; rc == 0 && k < std::max(1, cfg->warmup_frames); ++k) {
            refresh();
            rc = taa_batch_plan_launch(plan, st);
        }
        hipStreamSynchronize(st);
        hipEvent_t e0, e1;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        t_begin = host_us();
        t_end = t_begin + cfg->duration_s * 1e6;
        hipEventRecord(e0, st);
        uint64_t rounds = 0;
        while (rc == 0 && host_us() < t_end) {
            refresh();
            rc = taa_batch_plan_launch(plan, st);
            rounds++;
        }
        hipEventRecord(e1, st);
        hipStreamSynchronize(st);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        frames = rounds * ss.size();
        pixels = rounds * round_px;
        out->batches = rounds;
        out->mean_batch = (double)ss.size();
        out->lat_mean_us = rounds ? 1000.0 * ms / (double)rounds : 0.0;  // GPU time per round
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        if (plan) taa_batch_plan_destroy(plan);
        hipStreamDestroy(st);
    } else if (rc == 0) {
        // per-frame launches: sessions round-robin, one taa_session_resolve each
        hipStream_t st;
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        std::vector<taa_session_t*> hs;
        std::vector<void*> scr;
        std::vector<taa_cas_consts_t> cs(ss.size());
        for (size_t i = 0; i < ss.size(); ++i) {
            auto it = srv->sess.find(ss[i].sid);
            hs.push_back(it->second.s);
            scr.push_back(it->second.screen);
            cs[i] = it->second.cas;
        }
        auto one = [&](size_t i) {
            taa_uniforms_t Uf;
            sim_uniforms(Uf, ss[i].jitter, ss[i].frame++);
            return taa_session_resolve(hs[i], &Uf, srv->cas_on ? &cs[i] : nullptr, scr[i], cfg->arith, st);
        };
        for (int k = 0; k < std::max(1, cfg->warmup_frames); ++k)
            for (size_t i = 0; i < ss.size() && rc == 0; ++i) rc = one(i);
        hipStreamSynchronize(st);
        t_begin = host_us();
        t_end = t_begin + cfg->duration_s * 1e6;
        while (rc == 0 && host_us() < t_end) {
            for (size_t i = 0; i < ss.size() && rc == 0; ++i) {
                const double t0 = host_us();
                rc = one(i);
                if (cfg->mode == 3) {
                    hipStreamSynchronize(st);
                    lat.push_back(host_us() - t0);
                }
                frames++;
                pixels += (uint64_t)cfg->class_w[ss[i].cls] * cfg->class_h[ss[i].cls];
            }
        }
        hipStreamSynchronize(st);
        hipStreamDestroy(st);
        out->batches = frames;
        out->mean_batch = 1.0;
    }
    const double t_stop = host_us();
    out->wall_s = (t_stop - t_begin) * 1e-6;
    if (cfg->mode == 0 || cfg->mode == 1) out->wall_s = (t_end - t_begin) * 1e-6;
    out->frames = frames;
    out->pixels = pixels;
    out->gpix_per_s = out->wall_s > 0 ? (double)pixels / out->wall_s / 1e9 : 0.0;
    out->achieved_fps = out->wall_s > 0 ? (double)frames / out->wall_s : 0.0;
    if (!lat.empty()) {  // (mode 4 reports the GPU time per round in lat_mean_us instead)
        double sum = 0;
        for (double v : lat) sum += v;
        out->lat_mean_us = sum / (double)lat.size();
        out->lat_max_us = *std::max_element(lat.begin(), lat.end());
        out->lat_p50_us = percentile(lat, 0.50);
        out->lat_p90_us = percentile(lat, 0.90);
        out->lat_p99_us = percentile(lat, 0.99);
    }
    out->sessions_end = (int32_t)ss.size();
    taa_server_destroy(srv);
    hipStreamDestroy(copy_st);
    for (auto* v : {&pool, &own})
        for (auto& pf : *v) {
            hipFree(pf.color);
            hipFree(pf.depth);
            hipFree(pf.vel);
        }
    return rc;
}
