// TEST INFRASTRUCTURE -- the oracle's C entry points (loaded by tests/ through ctypes and by
// bench.py's cpu_baseline leg). Nothing in the product links or loads this library.
//
// Parity status of each oracle function (how it is pinned, see tests/test_oracle.py):
//   halton / jitter selection  closed-form Halton KATs; jitter tables re-derived from them
//   cas_setup                  oracle/_ref (the reference's own ffx_a.h + ffx_cas.h compiled
//                              for the CPU) at many sharpness values; closed form -1/lerp(8,5,s)
//   rgb<->YCoCg, tonemap       round trip on dyadic values; closed forms
//   neighbourhood box          brute force on 3x3 patches (min/max order statistics; variance
//                              against numpy mean/std); invariants min<=centre<=max
//   closest-depth tap          brute force incl. all-equal ties -> tap 0
//   velocity / reprojection    constant fields give historyUv = uv - v; identity matrices
//   segmentation mask          each indicator against an independent numpy form (scipy-style
//                              Sobel, unit normals from angles, 4-neighbour id compare), the
//                              decision paths (border, all, outside, moving-object disocclusion,
//                              history count) on constructed frames; FP32 sin/cos within a few
//                              ulp of the library functions
//   edge-AA fallback           closed forms: texels outside category 1 and flat or low-contrast
//                              texels copied with alpha = luma; a straight step edge blends each
//                              side by subpix/6 towards the other; transposing or mirroring the
//                              image transposes or mirrors the result
//   velocity producer          jitter cancels on a still camera; closed-form shifts of a
//                              translated perspective camera and of a moved object; with zero
//                              jitter uv - v.xy and depth - v.z equal the resolve's matrix path
//   Catmull-Rom                9-tap (bilinear) form == 16-tap form; exact on linear ramps
//                              and at texel centres; textbook Keys(a=-1/2) kernel in numpy
//   clip / blend / rejection   closed forms: inside box -> identity, alpha=1 -> current, ...
//   CasFilter                  numpy restatement with the same integer steps; amp = 0 case;
//                              output in [0,1]
// Parity unpinned (no reference run is possible here; documented definitions are used):
//   sampler weight quantisation, out-of-bounds loads, GLSL sin() noise, FMA contraction.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "taa_oracle.hpp"

namespace {

// Segmentation inputs of one frame (material, prev_material, uv_normal at the input size,
// prev_segmask and mask_out at the output size); mask_out receives uSegMask where the reference
// stores it and keeps its content elsewhere.
struct SegIO {
    const uint32_t* material;
    const uint32_t* prev_material;
    const float* uv_normal;
    const uint32_t* prev_segmask;
    uint32_t* mask_out;
};

template <class T>
int resolve_frame(const taa_uniforms_t* U, int W, int H, int Wl, int Hl, const float* color, const float* depth,
                  const float* vel4, const T* hist, const float* hdepth, T* screen, T* hist_out,
                  T* debug, int32_t* inspect, int cr16, int threads, const SegIO* seg = nullptr) {
    if (!U || W <= 0 || H <= 0 || Wl <= 0 || Hl <= 0 || !color || !depth || !vel4 || !hist || !hdepth || !screen || !hist_out)
        return -1;
    if (seg && (!seg->material || !seg->prev_material || !seg->uv_normal || !seg->prev_segmask || !seg->mask_out))
        return -1;
    oracle::Frame<T> F;
    F.W = W; F.H = H; F.Wl = Wl; F.Hl = Hl;
    F.color = color; F.depth = depth; F.vel = vel4; F.hist = hist; F.hdepth = hdepth;
    if (seg) {
        F.material = seg->material;
        F.prev_material = seg->prev_material;
        F.uv_normal = seg->uv_normal;
        F.prev_segmask = seg->prev_segmask;
    }
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            oracle::PixelOut<T> o = oracle::resolve_pixel<T>(*U, F, oracle::iv2{x, y}, cr16 != 0);
            size_t i = (size_t)y * W + x;
            T* s = screen + 4 * i;
            s[0] = o.screen.x; s[1] = o.screen.y; s[2] = o.screen.z; s[3] = o.screen.w;
            T* h = hist_out + 4 * i;
            h[0] = o.history.x; h[1] = o.history.y; h[2] = o.history.z; h[3] = o.history.w;
            if (debug) {
                T* d = debug + 4 * i;
                d[0] = o.debug.x; d[1] = o.debug.y; d[2] = o.debug.z; d[3] = o.debug.w;
            }
            if (inspect) std::memcpy(inspect + (size_t)TAA_INSPECT_WORDS * i, o.inspect, sizeof(o.inspect));
            if (seg && o.seg_written) seg->mask_out[i] = o.segmask;
        }
    }
    return 0;
}

template <class T>
int cas_frame(const T* src, T* dst, int W, int H, uint32_t const1_x, int threads) {
    if (!src || !dst || W <= 0 || H <= 0) return -1;
    if (threads <= 0) threads = 1;
    auto load = [&](int x, int y) {
        x = oracle::iclamp(x, 0, W - 1);
        y = oracle::iclamp(y, 0, H - 1);
        const T* p = src + 4 * ((size_t)y * W + x);
        return oracle::v3<T>{p[0], p[1], p[2]};
    };
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            oracle::v3<T> c = oracle::cas_filter<T>(load, x, y, const1_x);
            T* d = dst + 4 * ((size_t)y * W + x);
            d[0] = c.x; d[1] = c.y; d[2] = c.z; d[3] = T(1);
        }
    return 0;
}

template <class T>
int edge_aa_frame(const T* src, const uint32_t* mask, T* dst, int W, int H, T subpix, T thr, T thr_min, int threads) {
    if (!src || !mask || !dst || W <= 0 || H <= 0) return -1;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            auto at = [&](int dx, int dy) {
                const int xx = oracle::iclamp(x + dx, 0, W - 1), yy = oracle::iclamp(y + dy, 0, H - 1);
                const T* p = src + 4 * ((size_t)yy * W + xx);
                return oracle::v3<T>{p[0], p[1], p[2]};
            };
            const size_t i = (size_t)y * W + x;
            oracle::v4<T> o = oracle::edge_aa_texel<T>(at, (mask[i] & 3u) == 1u, subpix, thr, thr_min);
            dst[4 * i + 0] = o.x;
            dst[4 * i + 1] = o.y;
            dst[4 * i + 2] = o.z;
            dst[4 * i + 3] = o.w;
        }
    return 0;
}

template <class T>
int velocity_frame(const taa_reproject_t& R, const float* depth, const uint32_t* ids, int W, int H, T* out,
                   int threads) {
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            T u = (T(x) + T(0.5)) / T(W), v = (T(y) + T(0.5)) / T(H);
            oracle::v4<T> m = oracle::velocity_texel<T>(R, u, v, T(depth[i]), ids ? ids[i] : 0u);
            out[4 * i + 0] = m.x;
            out[4 * i + 1] = m.y;
            out[4 * i + 2] = m.z;
            out[4 * i + 3] = m.w;
        }
    return 0;
}

}  // namespace

extern "C" {

// W x H: output (history, screen) size; Wl x Hl: input (colour, depth, velocity, history depth)
int oracle_resolve_f32(const taa_uniforms_t* U, int W, int H, int Wl, int Hl, const float* color, const float* depth,
                       const float* vel4, const float* hist, const float* hdepth, float* screen,
                       float* hist_out, float* debug, int32_t* inspect, int cr16, int threads) {
    return resolve_frame<float>(U, W, H, Wl, Hl, color, depth, vel4, hist, hdepth, screen, hist_out, debug, inspect, cr16,
                                threads);
}

int oracle_resolve_f64(const taa_uniforms_t* U, int W, int H, int Wl, int Hl, const float* color, const float* depth,
                       const float* vel4, const double* hist, const float* hdepth, double* screen,
                       double* hist_out, double* debug, int32_t* inspect, int cr16, int threads) {
    return resolve_frame<double>(U, W, H, Wl, Hl, color, depth, vel4, hist, hdepth, screen, hist_out, debug, inspect, cr16,
                                 threads);
}

// The resolve with the segmentation mask: the extra images and the mask written. W x H: output
// size; Wl x Hl: input size (colour, depth, velocity, history depth, material ids, G-buffer).
int oracle_resolve_seg_f32(const taa_uniforms_t* U, int W, int H, int Wl, int Hl, const float* color, const float* depth,
                           const float* vel4, const float* hist, const float* hdepth, float* screen, float* hist_out,
                           float* debug, int32_t* inspect, const uint32_t* material, const uint32_t* prev_material,
                           const float* uv_normal, const uint32_t* prev_segmask, uint32_t* mask_out, int threads) {
    SegIO sg{material, prev_material, uv_normal, prev_segmask, mask_out};
    return resolve_frame<float>(U, W, H, Wl, Hl, color, depth, vel4, hist, hdepth, screen, hist_out, debug, inspect, 1,
                                threads, &sg);
}

int oracle_resolve_seg_f64(const taa_uniforms_t* U, int W, int H, int Wl, int Hl, const float* color, const float* depth,
                           const float* vel4, const double* hist, const float* hdepth, double* screen, double* hist_out,
                           double* debug, int32_t* inspect, const uint32_t* material, const uint32_t* prev_material,
                           const float* uv_normal, const uint32_t* prev_segmask, uint32_t* mask_out, int threads) {
    SegIO sg{material, prev_material, uv_normal, prev_segmask, mask_out};
    return resolve_frame<double>(U, W, H, Wl, Hl, color, depth, vel4, hist, hdepth, screen, hist_out, debug, inspect, 1,
                                 threads, &sg);
}

// sin / cos of the segmentation normals in FP32 (the evaluation the kernels repeat), elementwise.
void oracle_seg_sincos_f32(const float* a, float* s, float* c, int n) {
    for (int i = 0; i < n; ++i) oracle::seg_sincos(a[i], s[i], c[i]);
}

// Edge-AA fallback over a W x H RGBA image with its R32UI mask (category = mask & 3).
int oracle_edge_aa_f32(const float* src, const uint32_t* mask, float* dst, int W, int H, float subpix, float thr,
                       float thr_min, int threads) {
    return edge_aa_frame<float>(src, mask, dst, W, H, subpix, thr, thr_min, threads);
}

int oracle_edge_aa_f64(const double* src, const uint32_t* mask, double* dst, int W, int H, double subpix, double thr,
                       double thr_min, int threads) {
    return edge_aa_frame<double>(src, mask, dst, W, H, subpix, thr, thr_min, threads);
}

// Velocity image (xyzw per texel) of a W x H depth image; ids may be NULL (all static). The object
// matrices are host memory here.
int oracle_velocity_f32(const taa_reproject_t* R, const float* depth, const uint32_t* ids, int W, int H, float* out,
                        int threads) {
    return velocity_frame<float>(*R, depth, ids, W, H, out, threads);
}

int oracle_velocity_f64(const taa_reproject_t* R, const float* depth, const uint32_t* ids, int W, int H, double* out,
                        int threads) {
    return velocity_frame<double>(*R, depth, ids, W, H, out, threads);
}

// Check of a numerical identity the GPU kernels use (not part of the restatement): the number of
// single-precision bit patterns x in [begin, end) for which q0 = x * RN(1/9),
// r = fma(-q0, 9, x), q = (q0 == 0 or inf) ? q0 : fma(r, RN(1/9), q0) differs from the IEEE
// quotient x / 9 (two NaNs compare equal).
uint64_t oracle_div9_mismatches(uint64_t begin, uint64_t end, int threads) {
    if (threads <= 0) threads = 1;
    const float y = 1.0f / 9.0f;
    uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static) num_threads(threads)
    for (long long u = (long long)begin; u < (long long)end; ++u) {
        const uint32_t ui = (uint32_t)u;
        float x;
        std::memcpy(&x, &ui, 4);
        volatile float exact = x / 9.0f;
        const float q0 = x * y;
        const float r = std::fma(-q0, 9.0f, x);
        const float q = (q0 == 0.0f || std::isinf(q0)) ? q0 : std::fma(r, y, q0);
        const float e = exact;
        uint32_t a, b;
        std::memcpy(&a, &e, 4);
        std::memcpy(&b, &q, 4);
        if (a != b && !(std::isnan(e) && std::isnan(q))) bad++;
    }
    return bad;
}

int oracle_cas_f32(const float* src, float* dst, int W, int H, uint32_t const1_x, int threads) {
    return cas_frame<float>(src, dst, W, H, const1_x, threads);
}

int oracle_cas_f64(const double* src, double* dst, int W, int H, uint32_t const1_x, int threads) {
    return cas_frame<double>(src, dst, W, H, const1_x, threads);
}

void oracle_cas_setup(uint32_t c0[4], uint32_t c1[4], float s, float inX, float inY, float outX, float outY) {
    oracle::cas_setup(c0, c1, s, inX, inY, outX, outY);
}

float oracle_halton(int i, int b) { return oracle::halton(i, b); }

uint32_t oracle_half_bits_trunc(float f) { return oracle::half_bits_trunc(f); }

// get_jitter_offset_for_frame (source/taa.hpp:150-233) with the pixel size of (W, H).
int oracle_jitter(int dist, uint64_t frame, int slow_motion, int fixed_index, float rotate_deg,
                  float extra_scale, int n_user, const float* user, uint32_t W, uint32_t H,
                  float out[2], uint32_t* n_out) {
    float px = 2.0f / float(W), py = 2.0f / float(H);
    std::vector<float> ox, oy;
    auto add = [&](float x, float y) { ox.push_back(px * x); oy.push_back(py * y); };
    const float e = 1.f / 8.f;
    float sx = 1.f, sy = 1.f;
    switch (dist) {
        case 0: add(-0.25f, -0.25f); add(0.25f, -0.25f); add(0.25f, 0.25f); add(-0.25f, 0.25f); break;
        case 1: add(-0.25f, -0.25f); add(0.25f, 0.25f); add(0.25f, -0.25f); add(-0.25f, 0.25f); break;
        case 2:
        case 3: {
            int n = dist == 2 ? 8 : 16;
            for (int i = 0; i < n; ++i) add(oracle::halton(i + 1, 2) - 0.5f, oracle::halton(i + 1, 3) - 0.5f);
            break;
        }
        case 4:
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) add(float(2 * c - 3) * e, float(2 * r - 3) * e);
            break;
        case 5:
            if (n_user <= 0 || !user) return -1;
            for (int i = 0; i < n_user; ++i) { ox.push_back(user[2 * i]); oy.push_back(user[2 * i + 1]); }
            sx = px; sy = py;
            break;
        default: return -1;
    }
    if (slow_motion > 1) frame /= (uint64_t)slow_motion;
    if (fixed_index >= 0) frame = (uint64_t)fixed_index;
    size_t k = (size_t)(frame % ox.size());
    float x = ox[k] * sx, y = oy[k] * sy;
    if (rotate_deg != 0.f) {
        float rad = rotate_deg * 0.01745329251994329576923690768489f;
        float s = std::sin(rad), c = std::cos(rad);
        float nx = x * c - y * s, ny = x * s + y * c;
        x = nx; y = ny;
    }
    out[0] = x * extra_scale;
    out[1] = y * extra_scale;
    if (n_out) *n_out = (uint32_t)ox.size();
    return 0;
}

}  // extern "C"
