// Seeded synthetic scene: the input producer for tests and benchmarks (include/taastar/taa_scene.h).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "taastar/taa_abi.h"
#include "taastar/taa_scene.h"

namespace {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uni(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
double normal(uint64_t& s) {
    double u1 = uni(s), u2 = uni(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
// stateless hash of (seed, x, y, frame) -> [0, 1)
double hash01(uint64_t seed, uint32_t x, uint32_t y, uint32_t f, uint32_t salt) {
    uint64_t s = seed ^ ((uint64_t)x * 0x9E3779B1ull) ^ ((uint64_t)y << 21) ^ ((uint64_t)f << 42) ^ ((uint64_t)salt * 0xD6E8FEB86659FD93ull);
    splitmix64(s);
    return uni(s);
}

struct Poly {
    double cx, cy;       // centre at frame 0 (scene px)
    double vx, vy;       // own velocity px/frame
    int nv;
    double ox[8], oy[8]; // vertex offsets from the centre, counter-clockwise
    double r;
    float depth;
    float rgb[3];
    double sdx, sdy;     // stripe direction (unit), period 8 px
};

}  // namespace

struct taa_scene {
    taa_scene_config_t cfg;
    std::vector<Poly> polys;
    double panx = 0, pany = 0;
    double margin = 0;
};

extern "C" {

int32_t taa_scene_defaults(taa_scene_config_t* c) {
    if (!c) return TAA_ERR_INVALID_ARG;
    std::memset(c, 0, sizeof(*c));
    c->width = 640;
    c->height = 360;
    c->num_polygons = 200;
    c->seed = 0x7AA50001ull;
    c->lum_mu = -1.0f;
    c->lum_sigma = 1.5f;
    c->firefly_fraction = 0.01f;
    c->firefly_min = 10.0f;
    c->firefly_max = 100.0f;
    c->pan_mode = 0;
    c->pan = 8.0f;
    c->object_speed = 16.0f;
    c->jitter_distribution = 2;
    c->jitter_phase = 0;
    c->color_max = 0.0f;
    c->depth_min = 0.1f;
    c->depth_max = 1.0f;
    return TAA_OK;
}

int32_t taa_scene_create(const taa_scene_config_t* cfg, taa_scene_t** out) {
    if (!cfg || !out || cfg->width <= 0 || cfg->height <= 0 || cfg->num_polygons < 0) return TAA_ERR_INVALID_ARG;
    taa_scene* s = new (std::nothrow) taa_scene();
    if (!s) return TAA_ERR_OUT_OF_MEMORY;
    s->cfg = *cfg;
    const double W = cfg->width, H = cfg->height;
    const int N = cfg->num_polygons;
    uint64_t g = cfg->seed;
    if (cfg->pan_mode == 1) {
        const double mag = uni(g) * cfg->pan, ang = uni(g) * 6.283185307179586;
        s->panx = mag * std::cos(ang);
        s->pany = mag * std::sin(ang);
    } else {
        s->panx = (2.0 * uni(g) - 1.0) * cfg->pan;
        s->pany = (2.0 * uni(g) - 1.0) * cfg->pan;
    }
    // radius scale: about two thirds of the frame covered by polygons
    const double rs = N > 0 ? 1.2 * std::sqrt(W * H / (double)N) : 1.0;
    s->margin = 1.5 * rs + 2.0;
    s->polys.resize(N);
    for (int i = 0; i < N; ++i) {
        uint64_t ps = cfg->seed ^ (0xA24BAED4963EE407ull * (uint64_t)(i + 1));
        Poly& p = s->polys[i];
        p.cx = -s->margin + uni(ps) * (W + 2 * s->margin);
        p.cy = -s->margin + uni(ps) * (H + 2 * s->margin);
        p.vx = (2.0 * uni(ps) - 1.0) * cfg->object_speed;
        p.vy = (2.0 * uni(ps) - 1.0) * cfg->object_speed;
        p.r = rs * (0.3 + 1.2 * uni(ps));
        p.nv = 3 + (int)(uni(ps) * 6.0);
        if (p.nv > 8) p.nv = 8;
        double ang[8];
        for (int k = 0; k < p.nv; ++k) ang[k] = uni(ps) * 6.283185307179586;
        std::sort(ang, ang + p.nv);
        for (int k = 0; k < p.nv; ++k) {
            p.ox[k] = p.r * std::cos(ang[k]);
            p.oy[k] = p.r * std::sin(ang[k]);
        }
        p.depth = (float)(cfg->depth_min + uni(ps) * (cfg->depth_max - cfg->depth_min));
        const double L = std::exp(cfg->lum_mu + cfg->lum_sigma * normal(ps));
        double a[3] = {0.2 + 0.8 * uni(ps), 0.2 + 0.8 * uni(ps), 0.2 + 0.8 * uni(ps)};
        const double y = 0.25 * a[0] + 0.5 * a[1] + 0.25 * a[2];
        for (int c = 0; c < 3; ++c) p.rgb[c] = (float)(a[c] * L / y);
        const double sd = uni(ps) * 6.283185307179586;
        p.sdx = std::cos(sd);
        p.sdy = std::sin(sd);
    }
    *out = s;
    return TAA_OK;
}

int32_t taa_scene_destroy(taa_scene_t* s) {
    if (!s) return TAA_ERR_INVALID_ARG;
    delete s;
    return TAA_OK;
}

int32_t taa_scene_render(taa_scene_t* s, int32_t frame, float* color, float* depth, float* vel, float jitter_ndc[2],
                         int32_t threads) {
    if (!s || !color || !depth || !vel || frame < 0) return TAA_ERR_INVALID_ARG;
    const taa_scene_config_t& C = s->cfg;
    const int W = C.width, H = C.height;
    // jitter of this frame (the projection translate of taa.hpp:248): geometry moves by +j px,
    // i.e. pixel (x, y) samples the scene at (x + 0.5 - jx, y + 0.5 - jy)
    taa_jitter_config_t jc;
    std::memset(&jc, 0, sizeof(jc));
    jc.sample_distribution = C.jitter_distribution;
    jc.fixed_index = -1;
    jc.extra_scale = 1.0f;
    float jn[2] = {0.f, 0.f};
    if (C.jitter_distribution >= 0 && C.jitter_distribution <= 4)
        taa_jitter_offset(&jc, (uint64_t)frame + (uint64_t)C.jitter_phase, (uint32_t)W, (uint32_t)H, jn, nullptr);
    if (jitter_ndc) { jitter_ndc[0] = jn[0]; jitter_ndc[1] = jn[1]; }
    const double jx = (double)jn[0] * 0.5 * W, jy = (double)jn[1] * 0.5 * H;
    const double f = frame;
    const double DW = W + 2 * s->margin, DH = H + 2 * s->margin;
    const int N = (int)s->polys.size();
    // polygon positions this frame (wrapped into the margin-extended domain)
    std::vector<double> px(N), py(N);
    for (int i = 0; i < N; ++i) {
        const Poly& p = s->polys[i];
        double x = p.cx + (s->panx + p.vx) * f + s->margin, y = p.cy + (s->pany + p.vy) * f + s->margin;
        x = std::fmod(x, DW); if (x < 0) x += DW;
        y = std::fmod(y, DH); if (y < 0) y += DH;
        px[i] = x - s->margin + jx;
        py[i] = y - s->margin + jy;
    }
    const float bgvx = (float)(s->panx / W), bgvy = (float)(s->pany / H);
    if (threads <= 0) threads = omp_get_max_threads();
    const int bands = std::max(1, std::min(H, threads * 4));
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int b = 0; b < bands; ++b) {
        const int y0 = (int)((long long)H * b / bands), y1 = (int)((long long)H * (b + 1) / bands);
        // background: checker that moves with the camera pan
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t i = (size_t)y * W + x;
                const double sx = x + 0.5 - jx - s->panx * f, sy = y + 0.5 - jy - s->pany * f;
                const int cell = ((int)std::floor(sx / 24.0) + (int)std::floor(sy / 24.0)) & 1;
                const float l = cell ? 0.12f : 0.05f;
                color[4 * i + 0] = l * 0.9f;
                color[4 * i + 1] = l;
                color[4 * i + 2] = l * 1.2f;
                color[4 * i + 3] = 1.0f;
                depth[i] = 1.0f;
                vel[4 * i + 0] = bgvx;
                vel[4 * i + 1] = bgvy;
                vel[4 * i + 2] = 0.0f;
                vel[4 * i + 3] = 0.0f;
            }
        for (int k = 0; k < N; ++k) {
            const Poly& p = s->polys[k];
            const double cx = px[k], cy = py[k];
            int ya = std::max(y0, (int)std::floor(cy - p.r - 0.5)), yb = std::min(y1 - 1, (int)std::ceil(cy + p.r - 0.5));
            if (ya > yb) continue;
            const int xa = std::max(0, (int)std::floor(cx - p.r - 0.5)), xb = std::min(W - 1, (int)std::ceil(cx + p.r - 0.5));
            if (xa > xb) continue;
            const float pvx = (float)((s->panx + p.vx) / W), pvy = (float)((s->pany + p.vy) / H);
            for (int y = ya; y <= yb; ++y)
                for (int x = xa; x <= xb; ++x) {
                    const size_t i = (size_t)y * W + x;
                    if (!(p.depth < depth[i])) continue;
                    const double qx = x + 0.5 - cx, qy = y + 0.5 - cy;  // sample relative to the centre
                    bool inside = true;
                    for (int e = 0; e < p.nv && inside; ++e) {
                        const int e1 = (e + 1) % p.nv;
                        const double ex = p.ox[e1] - p.ox[e], ey = p.oy[e1] - p.oy[e];
                        inside = ex * (qy - p.oy[e]) - ey * (qx - p.ox[e]) >= 0.0;
                    }
                    if (!inside) continue;
                    const double t = (qx * p.sdx + qy * p.sdy) / 8.0;
                    const float m = ((int)std::floor(t) & 1) ? 1.15f : 0.85f;
                    color[4 * i + 0] = p.rgb[0] * m;
                    color[4 * i + 1] = p.rgb[1] * m;
                    color[4 * i + 2] = p.rgb[2] * m;
                    depth[i] = p.depth;
                    vel[4 * i + 0] = pvx;
                    vel[4 * i + 1] = pvy;
                    vel[4 * i + 3] = 1.0f;
                }
        }
        // fireflies and the storage clamp
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t i = (size_t)y * W + x;
                if (C.firefly_fraction > 0.f && hash01(C.seed, (uint32_t)x, (uint32_t)y, (uint32_t)frame, 1) < C.firefly_fraction) {
                    const float k = (float)(C.firefly_min + (C.firefly_max - C.firefly_min) * hash01(C.seed, (uint32_t)x, (uint32_t)y, (uint32_t)frame, 2));
                    for (int c = 0; c < 3; ++c) color[4 * i + c] *= k;
                }
                if (C.color_max > 0.f)
                    for (int c = 0; c < 3; ++c) color[4 * i + c] = std::min(color[4 * i + c], C.color_max);
            }
    }
    return TAA_OK;
}

}  // extern "C"
