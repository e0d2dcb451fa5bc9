// Arithmetic of the FP16 path shared by its kernels (taa_fast.hip, and the upsampling resolve of
// taa_kernels.hip in its FP16 form): fp16 texels through mixed fp16 x fp32 FMAs, fp32 moments and
// blend, the hardware reciprocal and square root; the decisions that select texels stay those of
// the FP32 path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "taa_device.hpp"

namespace taastar {
namespace {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

// Linear-depth rejection |L(hd) - L(d)| > thr with L(x) = n*f / (f + x*(n - f)), decided exactly as
// with the IEEE divisions of the FP32 path. The quotients through the hardware reciprocal are
// within 2^-22 (relative) of the correctly rounded ones, so their difference is within
// 2^-21 * (|a| + |b|) of the FP32 path's; every case farther than twice that from the threshold is
// decided by them, the rest (and any non-finite case) by the IEEE divisions themselves.
__device__ __forceinline__ bool linear_depth_reject(float hd, float d, float n, float f, float thr) {
    const float nf = n * f, q1 = f + hd * (n - f), q2 = f + d * (n - f);  // linearize_depth's operations
    const float a = nf * rcp_fast(q1), b = nf * rcp_fast(q2);
    const float diff = fabsf(a - b), m = 0x1p-20f * (fabsf(a) + fabsf(b));
    if (diff > thr + m) return true;
    if (diff < thr - m) return false;
    return fabsf(nf / q1 - nf / q2) > thr;
}

// Catmull-Rom weights (taa.comp:464-467) for one axis, fp32
struct W4 { float w[4]; };
__device__ __forceinline__ W4 cr_weights(float f) {
    const float f2 = f * f;
    W4 r;
    r.w[0] = f * __builtin_fmaf(f, __builtin_fmaf(f, -0.5f, 1.0f), -0.5f);
    r.w[1] = __builtin_fmaf(f2, __builtin_fmaf(f, 1.5f, -2.5f), 1.0f);
    r.w[2] = f * __builtin_fmaf(f, __builtin_fmaf(f, -1.5f, 2.0f), 0.5f);
    r.w[3] = f2 * __builtin_fmaf(f, 0.5f, -0.5f);
    return r;
}

// one texel of the history as three (ALPHA: four) fp32 channels (fp16 storage feeds mixed FMAs)
template <bool S16>
struct Texel {
    float r, g, b, a;
};
template <bool S16>
__device__ __forceinline__ Texel<S16> load_texel(const void* base, uint32_t row_bytes, int i) {
    if constexpr (S16) {
        uint2 v = gload<uint2>(base, row_bytes, 8 * i);
        h2 rg = __builtin_bit_cast(h2, v.x);
        h2 ba = __builtin_bit_cast(h2, v.y);
        return {(float)rg.x, (float)rg.y, (float)ba.x, (float)ba.y};
    } else {
        float4 v = gload<float4>(base, row_bytes, 16 * i);
        return {v.x, v.y, v.z, v.w};
    }
}

// history sample: rgb and, when the dynamic anti-ghosting reads it, alpha (taa.comp:807)
struct Hist {
    f3 c;
    float a;
};

// Catmull-Rom 4x4 history filter (separable, fp32 accumulation), rgb (ALPHA: and alpha).
template <bool S16, bool ALPHA = false>
__device__ __forceinline__ Hist history_cr_fast(const void* hist, int W, int H, int pitch, int kx, int ky, float fx, float fy) {
    constexpr uint32_t TB = S16 ? 8u : 16u;  // bytes per texel
    const W4 wx = cr_weights(fx), wy = cr_weights(fy);
    f3 acc = {0.f, 0.f, 0.f};
    float acca = 0.f;
    // the whole 4x4 footprint inside the image. kx, ky come unclamped from to_int(floor(...)) and
    // saturate at INT_MIN / INT_MAX for huge or infinite history positions: the raw form of the test
    const bool interior = cr_inside_raw(kx, W) && cr_inside_raw(ky, H);
    if (interior) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            // byte offset of the row's first texel; the four texels follow as immediate offsets
            const uint32_t rb = ((uint32_t)(ky - 1 + j) * (uint32_t)pitch + (uint32_t)(kx - 1)) * TB;
            f3 row = {0.f, 0.f, 0.f};
            float rowa = 0.f;
            Texel<S16> t4[4];
            if constexpr (S16) {  // the row's four fp16 texels as two 16-byte loads (8-byte aligned)
                const u32x4_a8 lo = gload<u32x4_a8>(hist, rb, 0), hi = gload<u32x4_a8>(hist, rb, 16);
                const uint32_t w8[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const h2 rg = __builtin_bit_cast(h2, w8[2 * i]), ba = __builtin_bit_cast(h2, w8[2 * i + 1]);
                    t4[i] = {(float)rg.x, (float)rg.y, (float)ba.x, (float)ba.y};
                }
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) t4[i] = load_texel<S16>(hist, rb, i);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const Texel<S16> t = t4[i];
                row.x = __builtin_fmaf(t.r, wx.w[i], row.x);
                row.y = __builtin_fmaf(t.g, wx.w[i], row.y);
                row.z = __builtin_fmaf(t.b, wx.w[i], row.z);
                if constexpr (ALPHA) rowa = __builtin_fmaf(t.a, wx.w[i], rowa);
            }
            acc.x = __builtin_fmaf(row.x, wy.w[j], acc.x);
            acc.y = __builtin_fmaf(row.y, wy.w[j], acc.y);
            acc.z = __builtin_fmaf(row.z, wy.w[j], acc.z);
            if constexpr (ALPHA) acca = __builtin_fmaf(rowa, wy.w[j], acca);
        }
    } else {
        const int kxc = cr_base(kx, W), kyc = cr_base(ky, H);
        int xs[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) xs[i] = clampi(kxc - 1 + i, 0, W - 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t rb = (uint32_t)clampi(kyc - 1 + j, 0, H - 1) * (uint32_t)pitch;
            f3 row = {0.f, 0.f, 0.f};
            float rowa = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                Texel<S16> t = load_texel<S16>(hist, (rb + (uint32_t)xs[i]) * TB, 0);
                row.x = __builtin_fmaf(t.r, wx.w[i], row.x);
                row.y = __builtin_fmaf(t.g, wx.w[i], row.y);
                row.z = __builtin_fmaf(t.b, wx.w[i], row.z);
                if constexpr (ALPHA) rowa = __builtin_fmaf(t.a, wx.w[i], rowa);
            }
            acc.x = __builtin_fmaf(row.x, wy.w[j], acc.x);
            acc.y = __builtin_fmaf(row.y, wy.w[j], acc.y);
            acc.z = __builtin_fmaf(row.z, wy.w[j], acc.z);
            if constexpr (ALPHA) acca = __builtin_fmaf(rowa, wy.w[j], acca);
        }
    }
    return {acc, acca};
}

// Bilinear history (sample_history_rgba with mInterpolationMode 0, taa.comp:545-549): the 2x2 texels
// at (kx, ky) .. (kx + 1, ky + 1), clamped to the edge, weights (1 - f, f); fp32 accumulation.
template <bool S16>
__device__ __forceinline__ Hist history_bilinear_fast(const void* hist, int W, int H, int pitch, int kx, int ky, float fx, float fy) {
    constexpr uint32_t TB = S16 ? 8u : 16u;
    const uint32_t xa = (uint32_t)clampi(kx, 0, W - 1), xb = (uint32_t)clampi(kx, -1, W - 2) + 1u;
    const uint32_t ya = (uint32_t)clampi(ky, 0, H - 1) * (uint32_t)pitch, yb = ((uint32_t)clampi(ky, -1, H - 2) + 1u) * (uint32_t)pitch;
    const Texel<S16> a = load_texel<S16>(hist, (ya + xa) * TB, 0), b = load_texel<S16>(hist, (ya + xb) * TB, 0);
    const Texel<S16> c = load_texel<S16>(hist, (yb + xa) * TB, 0), d = load_texel<S16>(hist, (yb + xb) * TB, 0);
    const float ux = 1.0f - fx, uy = 1.0f - fy;
    const f3 top = {__builtin_fmaf(b.r, fx, a.r * ux), __builtin_fmaf(b.g, fx, a.g * ux), __builtin_fmaf(b.b, fx, a.b * ux)};
    const f3 bot = {__builtin_fmaf(d.r, fx, c.r * ux), __builtin_fmaf(d.g, fx, c.g * ux), __builtin_fmaf(d.b, fx, c.b * ux)};
    const float ta = __builtin_fmaf(b.a, fx, a.a * ux), ba = __builtin_fmaf(d.a, fx, c.a * ux);
    return {{__builtin_fmaf(bot.x, fy, top.x * uy), __builtin_fmaf(bot.y, fy, top.y * uy), __builtin_fmaf(bot.z, fy, top.z * uy)},
            __builtin_fmaf(ba, fy, ta * uy)};
}

// CAS (ffx_cas.h:424-537) with contraction: the same filter as cas_filter(), the four-neighbour
// sum factored as (b + d + f + h) * w + e.
__device__ __forceinline__ f3 cas_filter_fma(f3 b, f3 d, f3 e, f3 f, f3 h, float peak) {
    const float mxG = fmaxf(fmaxf(d.y, fmaxf(e.y, f.y)), fmaxf(b.y, h.y));
    const float mnG = fminf(fminf(d.y, fminf(e.y, f.y)), fminf(b.y, h.y));
    float amp = sat(fminf(mnG, 1.0f - mxG) * prx_lo_rcp(mxG));
    amp = prx_lo_sqrt(amp);
    const float w = amp * peak;
    const float rw = prx_med_rcp(__builtin_fmaf(4.0f, w, 1.0f));
    f3 o;
    o.x = sat(__builtin_fmaf((b.x + d.x) + (f.x + h.x), w, e.x) * rw);
    o.y = sat(__builtin_fmaf((b.y + d.y) + (f.y + h.y), w, e.y) * rw);
    o.z = sat(__builtin_fmaf((b.z + d.z) + (f.z + h.z), w, e.z) * rw);
    return o;
}

// CAS on fp16 texels (RGBA16F storage): the green minimum and maximum are exact selections on
// (r,g) pairs in packed-half arithmetic; the amplitude, the weight and the weighted cross sum
// are fp32 with the fp16 texels entering through mixed-precision FMAs, so HDR values cannot
// overflow or cancel in half precision (ffx_cas.h:424-537, output alpha 1).
__device__ __forceinline__ h2 hmax(h2 a, h2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ h2 hmin(h2 a, h2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ h2 as_h2(uint32_t v) { return __builtin_bit_cast(h2, v); }
__device__ __forceinline__ float cas_channel(_Float16 b, _Float16 d, _Float16 e, _Float16 f, _Float16 h, float w, float rw) {
    float acc = __builtin_fmaf((float)b, w, (float)e);
    acc = __builtin_fmaf((float)d, w, acc);
    acc = __builtin_fmaf((float)f, w, acc);
    acc = __builtin_fmaf((float)h, w, acc);
    return sat(acc * rw);
}
// two values already in [0, 1] into one dword of fp16 (round to nearest even)
__device__ __forceinline__ uint32_t pack_unit_f2h(float a, float b) {
    const f32x2_t v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, f16x2_t));
}
__device__ __forceinline__ uint2 cas_filter_h2(uint2 b, uint2 d, uint2 e, uint2 f, uint2 h, float peak) {
    const h2 brg = as_h2(b.x), drg = as_h2(d.x), erg = as_h2(e.x), frg = as_h2(f.x), hrg = as_h2(h.x);
    const h2 mx = hmax(hmax(drg, hmax(erg, frg)), hmax(brg, hrg));
    const h2 mn = hmin(hmin(drg, hmin(erg, frg)), hmin(brg, hrg));
    const float mxG = (float)mx.y, mnG = (float)mn.y;
    float amp = sat(fminf(mnG, 1.0f - mxG) * prx_lo_rcp(mxG));
    amp = prx_lo_sqrt(amp);
    const float w = amp * peak;
    const float rw = prx_med_rcp(__builtin_fmaf(4.0f, w, 1.0f));
    const h2 bba = as_h2(b.y), dba = as_h2(d.y), eba = as_h2(e.y), fba = as_h2(f.y), hba = as_h2(h.y);
    const float r = cas_channel(brg.x, drg.x, erg.x, frg.x, hrg.x, w, rw);
    const float g = cas_channel(brg.y, drg.y, erg.y, frg.y, hrg.y, w, rw);
    const float bl = cas_channel(bba.x, dba.x, eba.x, fba.x, hba.x, w, rw);
    return make_uint2(pack_unit_f2h(r, g), pack_unit_f2h(bl, 1.0f));
}

// clipAabb (taa.comp:323-345) with one reciprocal: ma = max_i |v_i| / e_i is M / D for
// M = max_i |v_i| prod_{j != i} e_j and D = e_x e_y e_z, and the clipped point is p + v D / M
__device__ __forceinline__ f3 clip_aabb_fast(f3 mn, f3 mx, f3 q) {
    const f3 p = {0.5f * (mx.x + mn.x), 0.5f * (mx.y + mn.y), 0.5f * (mx.z + mn.z)};
    const float ex = __builtin_fmaf(0.5f, mx.x - mn.x, 1e-7f), ey = __builtin_fmaf(0.5f, mx.y - mn.y, 1e-7f),
                ez = __builtin_fmaf(0.5f, mx.z - mn.z, 1e-7f);
    const f3 v = {q.x - p.x, q.y - p.y, q.z - p.z};
    const float exy = ex * ey, D = exy * ez;
    const float M = fmaxf(fabsf(v.x) * (ey * ez), fmaxf(fabsf(v.y) * (ex * ez), fabsf(v.z) * exy));
    if (M > D) {
        const float s = D * rcp_fast(M);
        return {__builtin_fmaf(v.x, s, p.x), __builtin_fmaf(v.y, s, p.y), __builtin_fmaf(v.z, s, p.z)};
    }
    return q;
}

}  // namespace
}  // namespace taastar
